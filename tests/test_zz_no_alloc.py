"""lc_flat_sum_multi_noalloc (the flux sums on device arrays, with the work
memory in the library's static device scratch array, no allocation) against
lc_flat_sum_star{1,2}_multi (host arrays, the library's own device buffers)
through the C ABI: bit for bit in both precisions, for both stars, with
beaming and lensing, on two calls in a row, the second with a smaller input
(a large grid first, so stale scratch contents would show), and the size
limit reported as an error before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden_faces, load_golden

pytestmark = pytest.mark.gpu


class _DeviceArrays:
    """Device buffers through the HIP runtime the engine library itself
    links (the test allocates them; the entry point under test does not)."""

    def __init__(self):
        import lcurve_re_amd as L
        L._lib()   # loads libamdhip64 with the engine
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipSetDevice.argtypes = [C.c_int]
        self.bufs = []
        assert self.hip.hipSetDevice(0) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, n):
        out = np.empty(n, np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _noalloc(engine, star, flat, ph, ldc, beam, spin, VFAC, XCOFM, glens, rl, prec):
    import lcurve_re_amd as L
    from lcurve_re_amd import _abi
    D = _DeviceArrays()
    try:
        g = _abi.lc_flat_grid()
        g.n, g.n0, g.n1 = flat.n, flat.n0, flat.n1
        for k in ("dx", "dy", "dz", "px", "py", "pz", "flux", "in1", "out1"):
            setattr(g, k, C.cast(D.put(getattr(flat, k)), _abi._DP))
        pb = _abi.lc_phase_batch()
        pb.n = ph.n
        for k in ("ex", "ey", "ez", "phin"):
            setattr(pb, k, C.cast(D.put(getattr(ph, k)), _abi._DP))
        out = D.put(np.full(ph.n, -1.0))
        st = L._lib().lc_flat_sum_multi_noalloc(engine._ctx, int(star), C.byref(g), C.byref(pb), C.byref(ldc),
                                                beam, spin, VFAC, XCOFM, int(glens), rl, int(prec), out)
        return st, D.get(out, ph.n)
    finally:
        D.free()


def _large_case():
    """A detached binary at nlat 100 (both stars eclipsed), continuum fluxes,
    5000 sub-phases."""
    import lcurve_re_amd as L
    from lcurve_re_amd import _abi, workloads as W
    from oracle import pyoracle as P
    m = L.Model.from_json(W.model_json(phys={"beam_factor1": 1.3, "beam_factor2": 1.6},
                                       comp={"nlat1f": 100, "nlat2f": 100, "nlat1c": 100, "nlat2c": 100}))
    f1, f2 = P.star_grid(m.pod, 0, True), P.star_grid(m.pod, 1, True)
    P.star_continuum(m.pod, f1, f2)
    pod = m.pod
    ph = P.phases_for(pod.iangle.value, np.arange(5000) / 5000.0)
    VFAC = np.float32(pod.velocity_scale.value) / (2.99792458e8 / 1.0e3)
    XCOFM = pod.q.value / (1.0 + pod.q.value)
    cases = []
    for star, f, lt, c in ((0, f1, 1, (0.6, -0.2, 0.3, -0.1)), (1, f2, 0, (0.45, 0.12, -0.05, 0.02))):
        ldc = _abi.lc_ldc(*c, 0.0, lt, 0)
        cases.append((star, P.Flat(f), ph, ldc, 1.3 if star == 0 else 1.6, 1.0, VFAC, XCOFM, 0, 0.0))
    return cases


def _golden_cases(name):
    from lcurve_re_amd import _abi
    from oracle import pyoracle as P
    g = load_golden(name)
    VFAC, XCOFM = g["VFAC_XCOFM"]
    ph = P.Phases(g["ph_ex"], g["ph_ey"], g["ph_ez"], g["ph_phin"])
    cases, i = [], 0
    while f"sum{i}" in g:
        star, lt, beam, glens, rl, spin, c1, c2, c3, c4 = g[f"sum{i}_spec"]
        flat = P.Flat(golden_faces(g, f"s{int(star)}f"))
        cases.append((int(star), flat, ph, _abi.lc_ldc(c1, c2, c3, c4, 0.0, int(lt), 0), beam, spin, VFAC, XCOFM,
                      int(glens), rl))
        i += 1
    return cases


def _check(engine, case, prec, tag):
    star, flat, ph, ldc, beam, spin, VFAC, XCOFM, glens, rl = case
    ref = engine.flat_sum(star, flat.view(), ph.view(), ldc, beam, spin, VFAC, XCOFM, glens, rl, precision=prec)
    st, got = _noalloc(engine, star, flat, ph, ldc, beam, spin, VFAC, XCOFM, glens, rl, prec)
    assert st == 0, (tag, st)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, (tag, bad[:5], got[bad[:5]], ref[bad[:5]])


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_host_entry_large_then_small(engine, prec):
    large = _large_case()
    for k, case in enumerate(large):
        print(f"noalloc: large case star {case[0]} faces {case[1].n} sub-phases {case[2].n} precision {prec}",
              flush=True)
        _check(engine, case, prec, f"large {k}")
    for name in ("semidetached", "spheres"):
        for k, case in enumerate(_golden_cases(name)):
            print(f"noalloc: golden {name} sum {k} star {case[0]} faces {case[1].n} precision {prec}", flush=True)
            _check(engine, case, prec, f"{name} {k}")


def test_noalloc_rejects_inputs_beyond_its_scratch(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import _abi
    ldc = _abi.lc_ldc(0.6, 0.0, 0.0, 0.0, 0.0, 0, 0)
    # 2M faces x 390 sub-phases need ~135 MB of work memory: refused on the
    # host before any kernel runs (the arrays are never read)
    D = _DeviceArrays()
    try:
        g = _abi.lc_flat_grid()
        p = C.cast(D.put(np.zeros(8)), _abi._DP)
        g.n, g.n0, g.n1 = 2_000_000, 2_000_000, 2_000_000
        for k in ("dx", "dy", "dz", "px", "py", "pz", "flux"):
            setattr(g, k, p)
        pb = _abi.lc_phase_batch()
        pb.n = 390
        for k in ("ex", "ey", "ez", "phin"):
            setattr(pb, k, p)
        out = D.put(np.zeros(390))
        st = L._lib().lc_flat_sum_multi_noalloc(engine._ctx, 0, C.byref(g), C.byref(pb), C.byref(ldc), 0.0, 1.0,
                                                0.0, 0.0, 0, 0.0, 0, out)
    finally:
        D.free()
    assert st == _abi.LC_E_ARG
    msg = L._lib().lc_last_error(engine._ctx).decode()
    print("noalloc size limit:", msg, flush=True)
    assert "work memory" in msg and "2000000 faces" in msg
