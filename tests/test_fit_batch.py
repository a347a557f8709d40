"""Fit-loop batch (LM Jacobian + walkers, lcurve_re_amd/fit.py).

CPU tests run FitBatch over the FP64 oracle and pin it to facts that do
not come from fit.py itself:
* the step plan follows levmarq_solver.cpp:766-778, 800-850: central
  differences inside the limits, forward ones for parameters whose
  get_limit() interval excludes p - h (negative limb-darkening
  coefficients, limit (0, inf)), backward ones at the upper limit;
* the t0 column is the time derivative of the model: m(t; t0) = m(t - t0)
  so dr/dt0 = m'(t) / sigma without autoscale, checked against a central
  difference of the model in time (times shifted by +-1e-6, t0 fixed);
* walker chi**2 equals the oracle's own chi**2 (re_scale, re_scale.cpp:16-43).
GPU test: the batched engine against the oracle (Jacobian columns, chi**2).
"""
import json

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import fit as F
from lcurve_re_amd import workloads as W


def _problem(nphase=600, nlat=16, iangle=85.0):
    from oracle import pyoracle as P
    w = W.fit_batch(nphase=nphase)
    cfg = w["model"]
    for k in ("nlat1f", "nlat2f", "nlat1c", "nlat2c"):
        cfg["model_parameters"][k] = str(nlat)
    cfg["model_parameters"]["iangle"] = W._pp(iangle, 1)
    m = L.Model.from_json(json.dumps(cfg))
    base = L.make_data(w["time"], ferr=w["sigma"], weight=1.0)
    st, curve, _ = P.light_curve(m.pod, base, scale=False, rdata=False)
    assert st == 0
    flux = W.make_fit_flux(curve, w["sigma"], w["noise_seed"])
    data = L.make_data(w["time"], ferr=w["sigma"], weight=1.0, flux=flux)
    return m, data, w, curve


def test_step_plan_matches_reference_rules():
    m, data, _, _ = _problem(nphase=50)
    fb = F.FitBatch(m, data, evaluate=None)
    p = m.get_param()
    h = F.fd_steps(p, fb.limits)
    names = [m.get_name(i) for i in range(m.nvary())]
    kinds = dict(zip(names, (k for k, _ in fb.plan(p, h))))
    # ldc1_2 = -0.2 and ldc1_4 = -0.1 lie outside the limit (0, inf)
    assert kinds["ldc1_2"] == "f" and kinds["ldc1_4"] == "f"
    assert all(kinds[n] == "c" for n in names if n not in ("ldc1_2", "ldc1_4"))
    assert h[names.index("t0")] == F.FD_STEP_MIN          # p = 0
    assert abs(h[names.index("q")] / (0.5 * 1.220703125e-4) - 1.0) < 0.01
    # at the upper limit of iangle (90 deg) the step turns backward
    p2 = p.copy()
    p2[names.index("iangle")] = 90.0
    k2 = dict(zip(names, (k for k, _ in fb.plan(p2, F.fd_steps(p2, fb.limits)))))
    assert k2["iangle"] == "b"


def test_t0_column_is_time_derivative():
    from oracle import pyoracle as P
    m, _, w, curve = _problem(nphase=1000)
    # data in the model's own flux units, so that no autoscale is needed
    sig = w["sigma"] * float(np.mean(curve))
    t = w["time"]
    data = L.make_data(t, ferr=sig, weight=1.0, flux=curve)
    out = F.FitBatch(m, data, P.evaluator(), scale=False).jacobian()
    j = [m.get_name(i) for i in range(m.nvary())].index("t0")
    d = 1e-6
    _, cp, _ = P.light_curve(m.pod, L.make_data(t + d), scale=False, rdata=False)
    _, cm, _ = P.light_curve(m.pod, L.make_data(t - d), scale=False, rdata=False)
    expect = (cp - cm) / (2 * d) / sig
    got = out["J"][:, j]
    # face-by-face eclipse contacts make m piecewise smooth: most points agree
    # closely, a few near contacts see different face jumps at the two steps
    big = np.abs(expect) > 0.01 * np.percentile(np.abs(expect), 90)
    rel = np.abs(got[big] / expect[big] - 1.0)
    assert big.sum() > 0.5 * len(t)
    assert np.median(rel) < 1e-3 and np.percentile(rel, 90) < 1e-2, (np.median(rel), np.percentile(rel, 90))


def test_walker_chisq_matches_oracle():
    from oracle import pyoracle as P
    m, data, w, _ = _problem(nphase=300)
    walkers = F.walkers_around(m, 3, w["walker_seed"])
    fb = F.FitBatch(m, data, P.evaluator(), scale=True)
    out = fb.jacobian(walkers=walkers)
    assert out["instances_per_call"][0] == 1 + 2 * 10 + 2 + 3   # 10 central, 2 forward columns
    for k, v in enumerate(walkers):
        mm = m.copy()
        mm.set_param(v)
        st, _, io = P.light_curve(mm.pod, data, scale=True, rdata=True)
        assert st == 0
        assert abs(out["walker_chisq"][k] / io.chisq - 1.0) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [L.LC_PREC_FP64, L.LC_PREC_MIXED])
def test_jacobian_gpu_vs_oracle(precision, engine):
    from oracle import pyoracle as P
    m, data, w, _ = _problem(nphase=800, nlat=24)
    walkers = F.walkers_around(m, 8, w["walker_seed"])
    print(f"fit batch jacobian precision={precision}", flush=True)
    g = F.FitBatch(m, data, F.engine_evaluator(engine), scale=True, precision=precision).jacobian(walkers=walkers)
    o = F.FitBatch(m, data, P.evaluator(), scale=True).jacobian(walkers=walkers)
    # chi**2 ~ n: a relative model error e moves it by about 2 e / sigma
    # (sigma = 1.5e-3): 1e-7 for FP64 per face, 1e-5 at most in mixed mode
    ctol = 1e-6 if precision == L.LC_PREC_FP64 else 1e-4
    assert abs(g["chisq0"] / o["chisq0"] - 1.0) < ctol
    np.testing.assert_allclose(g["walker_chisq"], o["walker_chisq"], rtol=ctol)
    if precision != L.LC_PREC_FP64:
        # FP32 elements: curve errors of ~1e-7 relative, against parameter
        # steps of ~1e-4 |p|, leave finite-difference columns with errors of
        # tens of per cent (measured 0.4 for q), so the mixed mode serves
        # chi**2 (walkers) and the Jacobian is taken in FP64 (the default of
        # FitBatch)
        return
    names = [m.get_name(i) for i in range(m.nvary())]
    for j, n in enumerate(names):
        ref = o["J"][:, j]
        err = np.linalg.norm(g["J"][:, j] - ref) / max(np.linalg.norm(ref), 1e-300)
        assert err < 1e-3, (n, err)


def _fake_problem(fail):
    """A linear 'light curve' calc = A p over the fit model's parameters,
    with an evaluator that fails (status 2) or returns a non-finite curve
    wherever fail(j, delta) says so for the one perturbed parameter j."""
    cfg = W.fit_batch(nphase=50)["model"]
    m = L.Model.from_json(json.dumps(cfg))
    p0 = m.get_param()
    rng = np.random.default_rng(3)
    A = rng.standard_normal((50, len(p0)))
    data = L.make_data(np.arange(50) / 50.0, ferr=0.01, weight=1.0, flux=A @ p0 + 0.01)
    tried = []

    def evaluate(models, d, opts):
        calc = np.zeros((len(models), 50))
        st = np.zeros(len(models), np.int32)
        for k, mm in enumerate(models):
            p = mm.get_param()
            diff = np.nonzero(p != p0)[0]
            calc[k] = A @ p
            if len(diff) == 1:
                j = int(diff[0])
                tried.append((j, float(p[j] - p0[j])))
                f = fail(j, p[j] - p0[j])
                if f == "status":
                    st[k] = 2
                elif f == "nan":
                    calc[k, 7] = np.nan
        return calc, st
    return m, data, evaluate, tried, p0, A


@pytest.mark.parametrize("how", ["status", "nan"])
def test_jacobian_retries_like_the_reference(how):
    """levmarq_solver.cpp:800-866: a failing central column retries at
    h/2 .. h/16, then the one-sided step at h .. h/16, then backward at
    -h/16; a non-finite curve counts as a failed evaluation (:697)."""
    j_fail = 0   # q, central by the plan
    seen = []

    def fail(j, dlt):
        # every attempt on column j_fail fails but the 16th, the backward step
        if j != j_fail:
            return None
        seen.append(dlt)
        return None if len(seen) == 16 else how
    m, data, evaluate, tried, p0, A = _fake_problem(fail)
    fb = F.FitBatch(m, data, evaluate, scale=False)
    out = fb.jacobian()
    h = F.fd_steps(p0, m.get_limit())[j_fail]
    seq = [d for j, d in tried if j == j_fail]
    want = []
    for k in range(5):
        want += [h / 2 ** k, -h / 2 ** k]
    want += [h / 2 ** k for k in range(5)] + [-h / 16]
    assert len(seq) == len(want) and np.allclose(seq, want, rtol=1e-12, atol=0), (seq, want)
    # the last, backward step succeeded: for r = (f - A p) / sigma the
    # column is -A_j / sigma, up to the rounding of a step of h/16
    np.testing.assert_allclose(out["J"][:, j_fail], -A[:, j_fail] / 0.01, rtol=1e-3)
    assert out["kinds"][j_fail] == "b"


def test_jacobian_column_stays_zero_when_nothing_evaluates():
    m, data, evaluate, tried, p0, A = _fake_problem(lambda j, dlt: "status" if j == 1 else None)
    out = F.FitBatch(m, data, evaluate, scale=False).jacobian()
    assert np.all(out["J"][:, 1] == 0.0)
    assert np.any(out["J"][:, 0] != 0.0)
    # 5 central pairs, 5 forward steps, 1 backward step
    assert len([d for j, d in tried if j == 1]) == 10 + 5 + 1
