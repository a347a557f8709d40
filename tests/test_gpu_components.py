"""Disc, disc edge and bright spot (light_curve_comp.cpp:272-333, 483-497)
on the GPU against golden curves of the reference compiled from its sources
(tests/golden/curves.npz, tests/golden/make_golden_curves.py):
cataclysmic-variable-like binaries with a flared disc, its outer edge
irradiated by the donor, and the bright spot, including a spherical white
dwarf, a thick disc, a case without eclipses by star 1 and two opaque discs
that hide parts of both stars (their ranges join the stars' own, and those
grids' flux sums take several ranges per face), each unscaled
with diagnostics, scaled and through light_curve_comp_fast, in both
precisions.  Errors are max |gpu - ref| / max |ref| and the pointwise
max |gpu - ref| / |ref|, both within 1e-5.  The bright spot's position
comes from this engine's own integration of the gas stream (RK4), the
reference's from Bulirsch-Stoer to 1e-8, so agreement is to that accuracy,
not to the bit."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "curves.npz")
CASES = ["cv_disc_spot", "cv_disc", "cv_spot_sphere", "cv_thick_no_eclipse1", "cv_disc_opaque", "cv_disc_spot_opaque"]


def errors(calc, ref):
    d = np.abs(np.asarray(calc) - np.asarray(ref))
    scale = np.max(np.abs(ref))
    nz = np.abs(ref) > 1e-12 * scale
    return float(d.max() / scale), float(np.max(d[nz] / np.abs(ref[nz])))


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD, allow_pickle=False))


@pytest.mark.parametrize("case", CASES)
def test_components_vs_reference(engine, gold, case):
    import lcurve_re_amd as L
    m = L.Model.from_json(json.loads(bytes(gold[f"{case}_model_json"]).decode()))
    data = L.make_data(gold["pts_time"], expose=gold["pts_expose"], ndiv=gold["pts_ndiv"], ferr=gold["pts_ferr"],
                       weight=gold["pts_weight"], flux=gold[f"{case}_dflux"])
    for prec in (L.LC_PREC_FP64, L.LC_PREC_MIXED):
        for mode, kw in (("plain", dict(scale=False, diagnostics=True, sfac=1.7)),
                         ("scaled", dict(scale=True, diagnostics=True)),
                         ("fast", dict(scale=True, diagnostics=False, max_model_points=60))):
            print(f"components {case} {mode} precision {prec}", flush=True)
            mmp = kw.pop("max_model_points", 0)
            if mmp:
                calc, io = engine.light_curve_comp_fast(m, data, max_model_points=mmp, rdata=True, precision=prec,
                                                        **kw)
            else:
                calc, io = engine.light_curve_comp(m, data, rdata=True, precision=prec, **kw)
            ref = gold[f"{case}_{mode}_calc"]
            e_max, e_pt = errors(calc, ref)
            rio = gold[f"{case}_{mode}_io"]
            print(f"   max-normalised {e_max:.3e} pointwise {e_pt:.3e}; chisq {io.chisq:.10g} vs {rio[6]:.10g}",
                  flush=True)
            assert e_max <= 1e-5 and e_pt <= 1e-5, (case, mode, prec, e_max, e_pt)
            assert abs(io.chisq / rio[6] - 1) <= 1e-4
            if mode == "scaled":
                assert abs(io.sfac[0] / rio[0] - 1) <= 1e-5


def test_components_in_a_batch_match_single_calls(engine, gold):
    """Instances with and without components in one batch give each its own
    curve (the component tables are per instance): bit for bit in FP64 per
    face (in FP32 per face the flux tile shape, chosen per call, changes the
    FP32 summation grouping of the stars' sums)."""
    import lcurve_re_amd as L
    ms = [L.Model.from_json(json.loads(bytes(gold[f"{c}_model_json"]).decode())) for c in CASES]
    plain = L.Model.from_json(json.loads(bytes(gold["detached_base_model_json"]).decode()))
    models = [ms[0], plain, ms[1], ms[2], plain, ms[3], ms[4], ms[5], ms[4]]
    data = L.make_data(np.arange(700) / 700.0, ferr=1e-3, weight=1.0, flux=1.0)
    print("components: mixed batch of 9", flush=True)
    calc, _, st = engine.light_curve_comp_batch(models, data, scale=False, rdata=True, precision=L.LC_PREC_FP64)
    assert not st.any()
    for i, m in enumerate(models):
        c1, _ = engine.light_curve_comp(m, data, scale=False, rdata=True, diagnostics=False,
                                        precision=L.LC_PREC_FP64)
        assert np.array_equal(c1, calc[i]), i


@pytest.mark.parametrize("case", CASES)
def test_components_scaled_on_their_own(engine, gold, case):
    """iscale with components (light_curve_comp.cpp:378-416, 511-573): five
    columns (star 1, disc, edge, spot, star 2) fitted by least squares with
    svdfit's singular-value cut, and the caller's factors when scale is
    off.  The curve and chi**2 are compared; factors of nearly collinear
    columns are ill-determined (the reference's own SVD lands ~1e-9 off the
    optimum), so they are printed, and checked where a column is alone."""
    import lcurve_re_amd as L
    cfg = json.loads(bytes(gold[f"{case}_model_json"]).decode())
    cfg["model_parameters"]["iscale"] = "1"
    m = L.Model.from_json(cfg)
    data = L.make_data(gold["pts_time"], expose=gold["pts_expose"], ndiv=gold["pts_ndiv"], ferr=gold["pts_ferr"],
                       weight=gold["pts_weight"], flux=gold[f"{case}_dflux"])
    for prec in (L.LC_PREC_FP64, L.LC_PREC_MIXED):
        for mode, kw in (("iscale_scaled", dict(scale=True)), ("iscale_given", dict(scale=False,
                                                                                     sfac=[1.7, 0.8, 1.3, 0.6, 1.1]))):
            print(f"components iscale {case} {mode} precision {prec}", flush=True)
            calc, io = engine.light_curve_comp(m, data, rdata=True, diagnostics=True, precision=prec, **kw)
            ref = gold[f"{case}_{mode}_calc"]
            rio = gold[f"{case}_{mode}_io"]
            e_max, e_pt = errors(calc, ref)
            print(f"   max-normalised {e_max:.3e} pointwise {e_pt:.3e}; sfac {list(io.sfac)} vs {list(rio[:5])}",
                  flush=True)
            tol = 2e-9 if prec == L.LC_PREC_FP64 else 1e-5   # FP64 measured <= 3.3e-10 pointwise
            assert e_max <= tol and e_pt <= tol, (case, mode, prec, e_max, e_pt)
            assert abs(io.chisq / rio[6] - 1) <= 1e-4
            if mode == "iscale_given":
                assert list(io.sfac) == [1.7, 0.8, 1.3, 0.6, 1.1]
            else:
                for k in range(5):   # a column the model lacks has factor 0 in both
                    if rio[k] == 0.0:
                        assert io.sfac[k] == 0.0, (k, io.sfac[k])
