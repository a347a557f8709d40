"""GPU parity: the HIP engine against the reference's golden vectors and
the CPU oracle, called through the C ABI.

Tolerances (the north-star bar is <= 1e-5 relative error vs FP64):
* FP64 flux mode: bit-identical to flat_sum_star{1,2}_multi of the
  reference (golden vectors) on the same grid and sub-phases.
* mixed flux mode (FP32 per face, FP64 accumulation): max |gpu - ref| /
  max |ref| <= 1e-5 per summed curve.
* face grids: the device evaluates sin/cos/acos/atan2 with the GPU math
  library, so a bisection may take the other branch where a comparison is
  within an ulp; positions agree to 1e-7 (the radius accuracy delta/10 is
  1e-8), eclipse phases to 1e-6 (the phase accuracy is 1e-7), and at most
  0.5% of the faces may switch between eclipsed / not at the tolerance edge.
* light curves: max |gpu - oracle| / max |oracle| <= 1e-5.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import CASES, golden_faces, golden_model, load_golden

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _sum_inputs(g, i):
    from oracle import pyoracle as P
    from lcurve_re_amd import _abi
    star, lt, beam, glens, rl, spin, c1, c2, c3, c4 = g[f"sum{i}_spec"]
    star, lt, glens = int(star), int(lt), int(glens)
    faces = golden_faces(g, f"s{star}f")
    ldc = _abi.lc_ldc(c1, c2, c3, c4, 0.0, lt, 0)
    ph = P.Phases(g["ph_ex"], g["ph_ey"], g["ph_ez"], g["ph_phin"])
    return star, P.Flat(faces), ph, ldc, beam, spin, glens, rl


@pytest.mark.parametrize("case", CASES)
def test_flux_fp64_bit_exact_vs_reference(case, engine):
    import lcurve_re_amd as L
    g = load_golden(case)
    VFAC, XCOFM = g["VFAC_XCOFM"]
    i = 0
    while f"sum{i}" in g:
        star, flat, ph, ldc, beam, spin, glens, rl = _sum_inputs(g, i)
        print(f"case={case} sum={i} star={star} fp64", flush=True)
        out = engine.flat_sum(star, flat.view(), ph.view(), ldc, beam, spin, VFAC, XCOFM, glens, rl,
                              precision=L.LC_PREC_FP64)
        ref = g[f"sum{i}"]
        bad = np.nonzero(out != ref)[0]
        assert bad.size == 0, (case, i, bad[:5], out[bad[:5]], ref[bad[:5]])
        i += 1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_flux_ragged_faces_vs_reference(precision, engine):
    """Faces with several eclipse ranges (the FlatGrid's last class, as an
    opaque disc's ranges make them) through lc_flat_sum_star{1,2}_multi:
    FP64 bit-identical to the compiled reference's sums, mixed <= 1e-5; the
    classes straddle the 1024-face blocks, beaming, lensing and mucrit."""
    import lcurve_re_amd as L
    from test_oracle_golden import ragged_inputs
    g = load_golden("flat_ragged")
    VFAC, XCOFM = g["VFAC_XCOFM"]
    prec = L.LC_PREC_FP64 if precision == "fp64" else L.LC_PREC_MIXED
    i = 0
    while f"sum{i}" in g:
        star, flat, ph, ldc, beam, spin, glens, rl = ragged_inputs(g, i)
        print(f"ragged sum={i} star={star} n={flat.n} n1={flat.n1} {precision}", flush=True)
        out = engine.flat_sum(star, flat.view(), ph.view(), ldc, beam, spin, VFAC, XCOFM, glens, rl,
                              precision=prec)
        ref = g[f"sum{i}"]
        if prec == L.LC_PREC_FP64:
            bad = np.nonzero(out != ref)[0]
            assert bad.size == 0, (i, bad[:5], out[bad[:5]], ref[bad[:5]])
        else:
            assert rel_err(out, ref) <= 1e-5, (i, rel_err(out, ref))
        i += 1
    assert i == 6


@pytest.mark.parametrize("case", CASES)
def test_flux_mixed_vs_reference(case, engine):
    import lcurve_re_amd as L
    g = load_golden(case)
    VFAC, XCOFM = g["VFAC_XCOFM"]
    i = 0
    while f"sum{i}" in g:
        star, flat, ph, ldc, beam, spin, glens, rl = _sum_inputs(g, i)
        print(f"case={case} sum={i} star={star} mixed", flush=True)
        out = engine.flat_sum(star, flat.view(), ph.view(), ldc, beam, spin, VFAC, XCOFM, glens, rl,
                              precision=L.LC_PREC_MIXED)
        e = rel_err(out, g[f"sum{i}"])
        assert e <= 1e-5, (case, i, e)
        i += 1


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("key", ["s0f", "s1f", "s0c", "s1c"])
def test_grid_vs_reference(case, key, engine):
    from oracle import pyoracle as P
    g = load_golden(case)
    m = golden_model(g)
    star, fine = int(key[1]), key[2] == "f"
    ref = golden_faces(g, key)
    f = P.Faces(m.grid_size(star, fine))
    assert f.n == ref.n
    print(f"case={case} grid={key} n={f.n}", flush=True)
    engine.set_star_grid(m, star, fine, f.view())
    for k in ("px", "py", "pz", "dx", "dy", "dz"):
        d = np.max(np.abs(getattr(f, k) - getattr(ref, k)))
        assert d <= 1e-7, (case, key, k, d)
    assert np.max(np.abs(f.area / ref.area - 1.0)) <= 1e-6
    assert np.max(np.abs(f.gravity / ref.gravity - 1.0)) <= 1e-6
    flips = np.count_nonzero(f.necl != ref.necl)
    assert flips <= max(1, int(0.005 * f.n)), (case, key, flips)
    both = (f.necl == 1) & (ref.necl == 1)
    if both.any():
        assert np.max(np.abs(f.ingress[both] - ref.ingress[both])) <= 1e-6
        assert np.max(np.abs(f.egress[both] - ref.egress[both])) <= 1e-6


@pytest.mark.parametrize("case", CASES)
def test_continuum_vs_oracle(case, engine):
    from oracle import pyoracle as P
    g = load_golden(case)
    m = golden_model(g)
    f1, f2 = golden_faces(g, "s0f"), golden_faces(g, "s1f")
    o1, o2 = f1.copy(), f2.copy()
    P.star_continuum(m.pod, o1, o2)
    print(f"case={case} continuum", flush=True)
    engine.set_star_continuum(m, f1.view(), f2.view())
    for a, b in ((f1.flux, o1.flux), (f2.flux, o2.flux)):
        assert np.max(np.abs(a.astype(np.float64) / b - 1.0)) <= 2e-7


def _lc_case(name, nphase):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    g = load_golden(name)
    m = golden_model(g)
    t = np.arange(nphase) / nphase
    data = L.make_data(t, expose=0.002, ndiv=3, ferr=0.01, flux=1.0, weight=1.0)
    return m, data


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("prec", [0, 1])
def test_light_curve_vs_oracle(case, prec, engine):
    from oracle import pyoracle as P
    m, data = _lc_case(case, 600)
    print(f"case={case} light curve precision={prec}", flush=True)
    calc, io = engine.light_curve_comp(m, data, scale=True, rdata=True, diagnostics=True, precision=prec)
    st, ref, rio = P.light_curve(m.pod, data, scale=True, rdata=True, diagnostics=True)
    assert st == 0
    assert rel_err(calc, ref) <= 1e-5, rel_err(calc, ref)
    assert abs(io.sfac[0] / rio.sfac[0] - 1) <= 1e-5
    assert abs(io.chisq / rio.chisq - 1) <= 1e-4
    assert io.wnok == rio.wnok
    assert abs(io.wdwarf / rio.wdwarf - 1) <= 1e-5
    for k in ("logg1", "logg2", "rv1", "rv2"):
        assert abs(getattr(io, k) - getattr(rio, k)) <= 1e-6 * max(1.0, abs(getattr(rio, k))), k


@pytest.mark.parametrize("mucrit", [(0.0, 0.0), (0.05, 0.12)])
@pytest.mark.parametrize("limb", ["Poly", "Claret"])
def test_limb_laws_and_mucrit_vs_oracle(engine, mucrit, limb):
    """Both limb-darkening laws with and without a mu cut-off, both GPU
    precisions (the mixed kernels sum faces seen at a whole tile through
    FP64 moments for the polynomial law) against the FP64 oracle."""
    import json

    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    m = L.Model.from_json(json.dumps(W.model_json(
        phys={"q": 0.4, "iangle": 84.0, "r1": 0.12, "r2": 0.25},
        comp={"nlat1f": 32, "nlat2f": 32, "nlat1c": 32, "nlat2c": 32, "limb1": limb, "limb2": limb,
              "mucrit1": mucrit[0], "mucrit2": mucrit[1]})))
    data = L.make_data(np.arange(1500) / 1500.0, ferr=1e-3, flux=1.0)
    st, ref, _ = P.light_curve(m.pod, data, scale=False, rdata=True)
    assert st == 0
    for prec, tol in ((L.LC_PREC_FP64, 1e-9), (L.LC_PREC_MIXED, 1e-5)):
        calc, _ = engine.light_curve_comp(m, data, scale=False, rdata=True, diagnostics=False, precision=prec)
        print(f"limb {limb} mucrit {mucrit} precision {prec}: rel err {rel_err(calc, ref):.3e}", flush=True)
        assert rel_err(calc, ref) <= tol


def test_ring_gap_grid_vs_oracle(engine):
    """nlat 153: the rings lay out fewer faces than numface() allots, and
    the faces in between keep the default face (filled by one kernel for
    all such grids of a call); a batch of such instances and others."""
    import json

    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    models = [L.Model.from_json(json.dumps(W.model_json(
        phys={"q": 0.45, "iangle": 86.0, "r1": 0.1, "r2": 0.3},
        comp={"nlat1f": nl, "nlat2f": nl, "nlat1c": nl, "nlat2c": nl}))) for nl in (153, 60, 153)]
    data = L.make_data(np.arange(900) / 900.0, ferr=1e-3, flux=1.0)
    for prec, tol in ((L.LC_PREC_FP64, 1e-9), (L.LC_PREC_MIXED, 1e-5)):
        calc, _, status = engine.light_curve_comp_batch(models, data, scale=False, rdata=True, precision=prec)
        assert (status == 0).all(), status
        for i, m in enumerate(models):
            st, ref, _ = P.light_curve(m.pod, data, scale=False, rdata=True)
            assert st == 0
            print(f"gap grid model {i} precision {prec}: rel err {rel_err(calc[i], ref):.3e}", flush=True)
            assert rel_err(calc[i], ref) <= tol


def test_detached_small_fp64_vs_oracle(engine):
    """The small detached benchmark (nlat 40, 1000 phases) end to end."""
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    w = W.detached_small()
    m = L.Model.from_json(w["model"])
    data = L.make_data(w["time"])
    for prec in (L.LC_PREC_FP64, L.LC_PREC_MIXED):
        calc, _ = engine.light_curve_comp(m, data, scale=False, rdata=True, diagnostics=False, precision=prec)
        st, ref, _ = P.light_curve(m.pod, data, scale=False, rdata=True)
        print(f"detached_small precision={prec} rel err {rel_err(calc, ref):.3e}", flush=True)
        assert rel_err(calc, ref) <= (1e-9 if prec == L.LC_PREC_FP64 else 1e-5)


def test_batch_matches_single(engine):
    """A batch of instances gives the same curves as one call per instance."""
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    w = W.batch_prior(m=6, nphase=400, nlat_lo=20, nlat_hi=40)
    models = [L.Model.from_json(c) for c in w["models"]]
    data = L.make_data(w["time"])
    calc, io, status = engine.light_curve_comp_batch(models, data, scale=False, rdata=True)
    assert (status == 0).all(), status
    for i, m in enumerate(models):
        c1, _ = engine.light_curve_comp(m, data, scale=False, rdata=True, diagnostics=False)
        assert np.array_equal(c1, calc[i]), i


def test_packed_batch_and_out_array(engine):
    """A ModelBatch and a caller-owned output array give the list path's
    curves bit for bit; rows of failed instances keep their old contents."""
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    w = W.batch_prior(m=4, nphase=300, nlat_lo=20, nlat_hi=30)
    models = [L.Model.from_json(c) for c in w["models"]]
    models.insert(2, L.Model.from_json(W.model_json(phys={"r2": 0.6})))  # Roche overflow
    data = L.make_data(w["time"])
    ref, _, st_ref = engine.light_curve_comp_batch(models, data)
    out = np.full((len(models), len(data)), -7.0)
    got, _, st = engine.light_curve_comp_batch(L.ModelBatch(models), data, out=out)
    assert got is out
    assert list(st) == list(st_ref) and st[2] == 1
    for i in range(len(models)):
        if i == 2:
            assert (out[i] == -7.0).all()
        else:
            assert np.array_equal(out[i], ref[i]), i
    with pytest.raises(ValueError):
        engine.light_curve_comp_batch(models, data, out=np.zeros((len(models), len(data) + 1)))


def test_batch_with_curves_left_on_the_device(engine):
    """lc_light_curve_comp_batch_dev writes the curves of the host-array
    entry point, bit for bit, into device memory; twice in a row (the second
    call smaller, then the first shape again: captured graphs are replayed
    with the caller's buffer), failed rows untouched, scalars identical;
    host pointers and _fast sampling are refused."""
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    for m, nphase in ((5, 700), (3, 250), (5, 700)):
        w = W.batch_prior(m=m, nphase=nphase, nlat_lo=20, nlat_hi=36)
        models = [L.Model.from_json(c) for c in w["models"]]
        models.insert(1, L.Model.from_json(W.model_json(phys={"r2": 0.6})))  # Roche overflow
        data = L.make_data(w["time"], flux=np.linspace(0.5, 1.5, nphase))
        print(f"device output: {len(models)} instances x {nphase} phases", flush=True)
        ref, io_ref, st_ref = engine.light_curve_comp_batch(models, data, scale=True, rdata=True, diagnostics=True)
        dev = L.DeviceBuffer((len(models), nphase)).fill(-7.0)
        io, st = engine.light_curve_comp_batch_dev(models, data, dev.ptr, scale=True, rdata=True, diagnostics=True)
        got = dev.to_host()
        assert list(st) == list(st_ref) and st[1] == 1
        for i in range(len(models)):
            if i == 1:
                assert (got[i] == -7.0).all()
                continue
            assert np.array_equal(got[i], ref[i]), i
            assert io[i].chisq == io_ref[i].chisq and list(io[i].sfac) == list(io_ref[i].sfac)
            assert io[i].wdwarf == io_ref[i].wdwarf and io[i].logg1 == io_ref[i].logg1
    host = np.zeros((len(models), nphase))
    with pytest.raises(L.LcurveError):
        engine.light_curve_comp_batch_dev(models, data, host.ctypes.data)
    with pytest.raises(L.LcurveError):
        engine.light_curve_comp_batch_dev(models, data, dev.ptr, max_model_points=100)
    # the ordinary entry point still works after the refusals
    again, _, _ = engine.light_curve_comp_batch(models, data, scale=True, rdata=True, diagnostics=True)
    assert np.array_equal(again[0], ref[0])


def test_roche_overflow_raises(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    m = L.Model.from_json(W.model_json(phys={"r2": 0.6}))
    data = L.make_data(np.linspace(0, 1, 10))
    with pytest.raises(L.LcurveError) as e:
        engine.light_curve_comp(m, data)
    assert e.value.status == 1
    assert "larger than its Roche lobe" in str(e.value)
    # in a batch only that instance fails
    ok = L.Model.from_json(W.model_json(comp={"nlat1f": 10, "nlat2f": 10, "nlat1c": 10, "nlat2c": 10}))
    calc, io, status = engine.light_curve_comp_batch([ok, m, ok], data)
    assert list(status) == [0, 1, 0]
    assert np.array_equal(calc[0], calc[2])
