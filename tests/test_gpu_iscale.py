"""GPU parity of the iscale path (each star scaled on its own:
light_curve_comp.cpp:378-416, 511-573) against the FP64 oracle, whose
factors are pinned to numpy's SVD least squares (test_oracle_iscale.py).

The engine solves the two-star least squares through the 2x2 Gram matrix
with the same singular-value cut (1e-5 of the largest); the oracle runs a
Jacobi SVD on the full design matrix.  Tolerances: factors 1e-9 relative
in FP64 mode and 1e-5 in mixed mode (well conditioned, where both methods
give the least-squares answer); curves 1e-5 of their maximum (1e-9 in FP64
mode)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _model(iscale=1, **phys):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    p = {"q": 0.4, "iangle": 84.0, "r1": 0.12, "r2": 0.25}
    p.update(phys)
    return L.Model.from_json(json.dumps(W.model_json(
        phys=p, comp={"nlat1f": 32, "nlat2f": 32, "nlat1c": 32, "nlat2c": 32, "iscale": iscale})))


def _data(m, t, a0, a4, seed):
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    f1, f2 = P.star_curves(m.pod, L.make_data(t, ferr=1.0, flux=1.0))
    rng = np.random.default_rng(seed)
    n = len(t)
    flux = a0 * f1 + a4 * f2
    ferr = np.abs(flux).max() * (1e-3 + 1e-3 * rng.random(n))
    flux = flux + ferr * rng.standard_normal(n)
    weight = np.where(rng.random(n) < 0.1, 0.0, 1.0 + rng.random(n))
    return L.make_data(t, ferr=ferr, flux=flux, weight=weight)


@pytest.mark.parametrize("case", ["eclipsing", "one_phase_rank_one", "async_beamed"])
def test_iscale_fit_vs_oracle(engine, case):
    """async_beamed: the star components use spin 1 in the beaming term, as
    the reference's comp_star1_flat / comp_star2_flat do."""
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    m = _model(**({"spin1": 1.2, "spin2": 0.9, "beam_factor1": 1.3, "beam_factor2": 1.6}
                  if case == "async_beamed" else {}))
    t = 0.3 + (np.arange(600) % 2) if case == "one_phase_rank_one" else np.arange(1200) / 1200.0
    data = _data(m, t, 1.7, 0.6, seed=11)
    st, ref, rio = P.light_curve(m.pod, data, scale=True, rdata=True, diagnostics=True)
    assert st == 0
    rs = np.array(list(rio.sfac))
    for prec, tol_s, tol_c in ((L.LC_PREC_FP64, 1e-9, 1e-9), (L.LC_PREC_MIXED, 1e-5, 1e-5)):
        calc, io = engine.light_curve_comp(m, data, scale=True, rdata=True, diagnostics=True, precision=prec)
        s = np.array(list(io.sfac))
        print(f"iscale {case} precision {prec}: sfac {s} (oracle {rs}), curve rel err {rel_err(calc, ref):.2e}",
              flush=True)
        assert s[1] == 0 and s[2] == 0 and s[3] == 0
        assert abs(s[0] / rs[0] - 1) <= tol_s and abs(s[4] / rs[4] - 1) <= tol_s
        assert rel_err(calc, ref) <= tol_c
        assert abs(io.chisq / rio.chisq - 1) <= 1e-4 and abs(io.wnok / rio.wnok - 1) <= 1e-12  # summation order
        assert abs(io.wdwarf / rio.wdwarf - 1) <= 1e-5


def test_iscale_given_factors_and_batch_mix(engine):
    """Scale off: the caller's factors per star; a batch mixing iscale and
    ordinary instances gives each its own path."""
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    mi, mo = _model(), _model(iscale=0)
    data = L.make_data(np.arange(800) / 800.0, ferr=1e-3, flux=1.0)
    sf = [2.0, 7.0, 7.0, 7.0, 3.0]
    st, ref_i, _ = P.light_curve(mi.pod, data, scale=False, rdata=True, sfac=1.0)
    f1, f2 = P.star_curves(mi.pod, data)
    calc, io = engine.light_curve_comp(mi, data, scale=False, rdata=True, sfac=sf, diagnostics=False)
    assert rel_err(calc, 2.0 * f1 + 3.0 * f2) <= 1e-5
    calc_b, _, status = engine.light_curve_comp_batch([mi, mo, mi], data, scale=False, rdata=True)
    assert (status == 0).all()
    st, ref_o, _ = P.light_curve(mo.pod, data, scale=False, rdata=True)
    assert rel_err(calc_b[0], ref_i) <= 1e-5 and rel_err(calc_b[2], ref_i) <= 1e-5
    assert rel_err(calc_b[1], ref_o) <= 1e-5


def test_iscale_fast_path_vs_oracle(engine):
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    m = _model()
    data = _data(m, np.arange(3000) / 3000.0, 1.2, 0.9, seed=5)
    st, ref, rio = P.light_curve(m.pod, data, scale=True, rdata=True, max_model_points=400)
    assert st == 0
    calc, io = engine.light_curve_comp_fast(m, data, max_model_points=400, scale=True, rdata=True,
                                            diagnostics=False)
    assert rel_err(calc, ref) <= 1e-5
    assert abs(io.sfac[0] / rio.sfac[0] - 1) <= 1e-5 and abs(io.sfac[4] / rio.sfac[4] - 1) <= 1e-5
