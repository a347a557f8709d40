"""The CPU oracle against golden vectors of the compiled reference.

tests/golden/*.npz were written by tests/golden/make_golden.py from
oracle/_ref/libref.so (the reference compiled from its own sources with
-ffp-contract=off, like the oracle): grids from Lcurve::add_faces (which
runs Roche::face_chop_batch, Roche::face, Roche::ingress_egress with
pot_min/linmin/dbrent/fblink, and sphere_eclipse), flat flux sums from
flat_sum_star{1,2}_multi, Point-path sums from comp_star1/2, re_scale,
comp_radius, xl1/xl11/xl12, planck, dlpdlt and numface.  Every comparison
is bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import CASES, golden_faces, golden_model, load_golden


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("key", ["s0f", "s1f", "s0c", "s1c"])
def test_grid_matches_reference(case, key, oracle_lib):
    from oracle import pyoracle as P
    g = load_golden(case)
    m = golden_model(g)
    star, fine = int(key[1]), key[2] == "f"
    f = P.star_grid(m.pod, star, fine)
    ref = golden_faces(g, key)
    assert f.n == ref.n
    for k in ("px", "py", "pz", "dx", "dy", "dz", "area", "gravity", "necl", "ingress", "egress"):
        a, b = getattr(f, k), getattr(ref, k)
        assert np.array_equal(a, b), (case, key, k, np.max(np.abs(a - b)))


@pytest.mark.parametrize("case", CASES)
def test_flat_sums_match_reference(case, oracle_lib):
    from oracle import pyoracle as P
    from lcurve_re_amd import _abi
    g = load_golden(case)
    ph = P.Phases(g["ph_ex"], g["ph_ey"], g["ph_ez"], g["ph_phin"])
    VFAC, XCOFM = g["VFAC_XCOFM"]
    i = 0
    while f"sum{i}" in g:
        star, lt, beam, glens, rl, spin, c1, c2, c3, c4 = g[f"sum{i}_spec"]
        star, lt, glens = int(star), int(lt), int(glens)
        faces = golden_faces(g, f"s{star}f")
        ldc = _abi.lc_ldc(c1, c2, c3, c4, 0.0, lt, 0)
        out = P.flat_sum(star, P.Flat(faces), ph, ldc, beam, spin, VFAC, XCOFM, glens, rl)
        assert np.array_equal(out, g[f"sum{i}"]), (case, i, np.max(np.abs(out - g[f"sum{i}"])))
        i += 1
    assert i >= 6


@pytest.mark.parametrize("case", CASES)
def test_point_sums_match_reference(case, oracle_lib):
    """orc_point_sum is the raw per-phase sum; at ndiv=1, expose=0 and unit
    Ginterp scale the reference comp_star value equals it exactly."""
    from oracle import pyoracle as P
    from lcurve_re_amd import _abi
    g = load_golden(case)
    m = golden_model(g)
    for star, phase, expose, ndiv, beam, val in g["point_sums"]:
        if int(ndiv) != 1:
            continue
        faces = golden_faces(g, f"s{int(star)}f")
        ldc = (_abi.lc_ldc(0.6, -0.2, 0.3, -0.1, 0.0, 1, 0) if star == 0 else _abi.lc_ldc(0.5315, 0, 0, 0, 0.0, 0, 0))
        v = faces.view()
        got = oracle_lib.orc_point_sum(int(star), C.byref(v), C.byref(ldc), m.iangle, phase, m.q, beam,
                                       m.velocity_scale, 0, 0.0)
        # comp_star returns (0 + (1 * scale * s)) / 1 with scale 1 for these phases
        assert got == val, (case, star, phase, got, val)


def test_scalars_match_reference(oracle_lib):
    g = load_golden("scalars")
    err = C.c_int()
    for q, s, x1, x11, x12 in g["xl"]:
        assert oracle_lib.orc_xl1(q, C.byref(err)) == x1
        assert oracle_lib.orc_xl11(q, s, C.byref(err)) == x11
        assert oracle_lib.orc_xl12(q, s, C.byref(err)) == x12
    for w, t, b, dl in g["planck"]:
        assert oracle_lib.orc_planck(w, t) == b
        assert oracle_lib.orc_dlpdlt(w, t) == dl
    for n, nf in g["numface"]:
        assert oracle_lib.orc_numface(int(n), 0, 0.0, 0.0, 0, 0) == nf


def test_re_scale_matches_reference(oracle_lib):
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    g = load_golden("scalars")
    data = L.make_data(g["rs_time"], flux=g["rs_flux"], ferr=g["rs_ferr"], weight=g["rs_weight"])
    fit = g["rs_fit"].copy()
    chisq, wnok = C.c_double(), C.c_double()
    s = oracle_lib.orc_re_scale(data, len(fit), P.dptr(fit), C.byref(chisq), C.byref(wnok))
    assert np.array_equal(fit, g["rs_fit_out"])
    assert (s, chisq.value, wnok.value) == tuple(g["rs_scale"])


def ragged_inputs(g, i):
    """Inputs of flat_ragged.npz sum i: the FlatGrid with the extra ranges,
    sub-phases, LDC and the element parameters."""
    from oracle import pyoracle as P
    from lcurve_re_amd import _abi
    star, lt, beam, glens, rl, spin, c1, c2, c3, c4, muc = g[f"sum{i}_spec"]
    star, lt, glens = int(star), int(lt), int(glens)
    key = f"s{star}"
    faces = golden_faces(g, key)
    extra = {}
    for f, a, b in zip(g[f"{key}_extra_face"], g[f"{key}_extra_in"], g[f"{key}_extra_out"]):
        extra.setdefault(int(f), []).append((float(a), float(b)))
    ldc = _abi.lc_ldc(c1, c2, c3, c4, muc, lt, 0)
    ph = P.Phases(g["ph_ex"], g["ph_ey"], g["ph_ez"], g["ph_phin"])
    return star, P.Flat(faces, extra), ph, ldc, beam, spin, glens, rl


def test_flat_sums_with_ragged_faces_match_reference(oracle_lib):
    """Faces with several eclipse ranges (the FlatGrid's last class): the
    oracle's blocked sums against flat_sum_star{1,2}_multi of the compiled
    reference, bit for bit; the classes straddle the 1024-face blocks."""
    from oracle import pyoracle as P
    g = load_golden("flat_ragged")
    VFAC, XCOFM = g["VFAC_XCOFM"]
    i = 0
    while f"sum{i}" in g:
        star, flat, ph, ldc, beam, spin, glens, rl = ragged_inputs(g, i)
        assert flat.n1 < flat.n and flat.n1 // 1024 < (flat.n - 1) // 1024
        out = P.flat_sum(star, flat, ph, ldc, beam, spin, VFAC, XCOFM, glens, rl)
        assert np.array_equal(out, g[f"sum{i}"]), (i, np.max(np.abs(out - g[f"sum{i}"])))
        i += 1
    assert i == 6
