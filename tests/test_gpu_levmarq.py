"""The batched Levenberg-Marquardt driver on the GPU engine against the
same driver on the FP64 oracle (test_levmarq.py pins the driver itself):
an eclipsing binary fitted for inclination and both radii (flux scale
fitted in every evaluation) from a perturbed start, 10 iterations, with
noise of 1% of the maximum flux so that accept/reject decisions are not
decided by the face discretisation's own roughness of chi^2.  FP64 engine:
the same accept/reject sequence, parameters within 1e-6 relative.  Mixed
precision (FP32 per face): differences of step ~1e-4 |p| turn the FP32
element rounding (~1e-7 of the flux) into Jacobian errors ~1e-3, so the
path differs (FitBatch defaults to FP64 for this reason); it must still
reach a chi^2 within 1% of the oracle's and parameters within 1%."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem():
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    phys = {"q": 0.4, "iangle": 84.0, "r1": 0.12, "r2": 0.25, "t2": 4000.0}
    comp = {"nlat1f": 32, "nlat2f": 32, "nlat1c": 32, "nlat2c": 32}
    truth = L.Model.from_json(json.dumps(W.model_json(phys=phys, comp=comp)))
    t = np.arange(600) / 600.0
    st, clean, _ = P.light_curve(truth.pod, L.make_data(t, ferr=1.0, flux=1.0), scale=False, rdata=False)
    assert st == 0
    rng = np.random.default_rng(3)
    ferr = 1e-2 * np.abs(clean).max() * np.ones_like(t)
    data = L.make_data(t, ferr=ferr, flux=clean + ferr * rng.standard_normal(len(t)))
    start = dict(phys, iangle=83.0, r1=0.124, r2=0.245)
    vary = ("iangle", "r1", "r2")
    return lambda: L.Model.from_json(json.dumps(W.model_json(phys=start, comp=comp, vary=vary))), data


@pytest.mark.parametrize("prec", [1, 0])
def test_levmarq_engine_vs_oracle(engine, prec):
    from lcurve_re_amd import fit as F
    from oracle import pyoracle as P
    make, data = _problem()
    ref = F.levmarq(F.FitBatch(make(), data, P.evaluator(), scale=True), max_iter=10)
    got = F.levmarq(F.FitBatch(make(), data, F.engine_evaluator(engine), scale=True, precision=prec), max_iter=10)
    print(f"levmarq precision {prec}: engine {got['pars']} chisq {got['chisq']:.6f} it {got['iterations']} "
          f"calls {got['calls']} stop {got['stop']}; oracle {ref['pars']} chisq {ref['chisq']:.6f} "
          f"it {ref['iterations']} stop {ref['stop']}", flush=True)
    assert any(h[3] == "reject" for h in ref["history"])  # the batched candidates are exercised
    if prec == 1:
        assert [h[3] for h in got["history"]] == [h[3] for h in ref["history"]]
        np.testing.assert_allclose(got["pars"], ref["pars"], rtol=1e-6)
        assert abs(got["chisq"] / ref["chisq"] - 1) <= 1e-6
    else:
        np.testing.assert_allclose(got["pars"], ref["pars"], rtol=1e-2)
        assert abs(got["chisq"] / ref["chisq"] - 1) <= 1e-2
    assert ref["chisq"] < 1.1 * 600  # close to the noise level (600 points)


def test_levmarq_with_prior_engine_vs_oracle(engine):
    """The continuation over the prior scale (levmarq_solver.cpp:985-1530)
    with the physical prior's residuals (a K1 and an R1 constraint, R1 in
    tension with the data) on the engine and on the oracle: the same
    stages, balance factors and accept/reject sequence in FP64."""
    from lcurve_re_amd import fit as F
    from lcurve_re_amd import prior as PR
    from oracle import pyoracle as P
    make, data = _problem()
    m = make()
    pod = m.pod
    args = (84.0, pod.q.value, pod.velocity_scale.value, 0.12, pod.r2.value, pod.t1.value, pod.t2.value, 0.25)
    K1 = PR.lm_derived("K", *args)
    R1 = PR.lm_derived("R1", *args)
    cfg = {"use_priors": True, "true_period": 0.25, "prior_weight": 50.0,
           "priors": {"K1": f"{K1!r} {0.01 * K1!r} {0.01 * K1!r}", "R1": f"{1.02 * R1!r} {0.01 * R1!r} {0.01 * R1!r}"}}
    kw = dict(prior=PR.LMPrior(m, cfg), stages=2, max_iter=6)
    ref = F.levmarq(F.FitBatch(make(), data, P.evaluator(), scale=True), **kw)
    got = F.levmarq(F.FitBatch(make(), data, F.engine_evaluator(engine), scale=True), **kw)
    print(f"levmarq+prior: engine {got['pars']} sum_sq {got['chisq']:.6f} it {got['iterations']} calls "
          f"{got['calls']}; oracle {ref['pars']} sum_sq {ref['chisq']:.6f} it {ref['iterations']}; "
          f"eff {[s['eff'] for s in ref['stages']]}", flush=True)
    assert len(ref["stages"]) == 3 and ref["stages"][0]["eff"] == 0.0 and ref["stages"][-1]["eff"] > 0.0
    assert [h[3] for h in got["history"]] == [h[3] for h in ref["history"]]
    np.testing.assert_allclose([s["eff"] for s in got["stages"]], [s["eff"] for s in ref["stages"]], rtol=1e-6)
    np.testing.assert_allclose(got["pars"], ref["pars"], rtol=1e-6)
    assert abs(got["chisq"] / ref["chisq"] - 1) <= 1e-6


def test_laplace_modes_engine_vs_oracle(engine):
    """Mode grouping and Laplace weights (levmarq_solver.cpp:1640-1790) of
    the engine's multistart optima, computed with the engine and with the
    oracle: the same modes, lpost and log det Sigma (FP64)."""
    from lcurve_re_amd import fit as F
    from oracle import pyoracle as P
    make, data = _problem()
    starts = [[83.0, 0.124, 0.245], [85.0, 0.118, 0.252], [84.2, 0.121, 0.249]]
    fe = F.FitBatch(make(), data, F.engine_evaluator(engine), scale=True)
    runs, calls = F.levmarq_multistart(fe, starts, max_iter=8)
    got, gcalls = F.laplace_modes(fe, runs, dsteps=[0.05, 5e-4, 5e-4])
    ref, _ = F.laplace_modes(F.FitBatch(make(), data, P.evaluator(), scale=True), runs, dsteps=[0.05, 5e-4, 5e-4])
    print(f"modes: {[(m['pars'], m['n_runs'], m['weight']) for m in got]} calls {calls} + {gcalls}", flush=True)
    assert gcalls == 1 and sum(m["n_runs"] for m in got) == 3
    assert [m["run"] for m in got] == [m["run"] for m in ref]
    for a, b in zip(got, ref):
        assert abs(a["lpost"] - b["lpost"]) <= 1e-6 * abs(b["lpost"])
        assert abs(a["logdet"] - b["logdet"]) <= 1e-4 * max(1.0, abs(b["logdet"]))
        assert abs(a["weight"] - b["weight"]) <= 1e-4


def test_error_mcmc_engine_vs_oracle(engine):
    """The error sampling over the modes (levmarq_solver.cpp:2090-2500) on
    the engine and on the oracle from the same modes: the same chains in
    FP64 (each chain owns its generator, so only the evaluations differ)."""
    from lcurve_re_amd import fit as F
    from lcurve_re_amd import mcmc as M
    from oracle import pyoracle as P
    make, data = _problem()
    fe = F.FitBatch(make(), data, F.engine_evaluator(engine), scale=True)
    runs, _ = F.levmarq_multistart(fe, [[83.0, 0.124, 0.245], [84.2, 0.121, 0.249]], max_iter=8)
    modes, _ = F.laplace_modes(fe, runs, dsteps=[0.05, 5e-4, 5e-4])
    kw = dict(dsteps=[0.05, 5e-4, 5e-4], steps=150, min_steps=150, seed=5)
    got = M.error_mcmc(fe, modes, **kw)
    ref = M.error_mcmc(F.FitBatch(make(), data, P.evaluator(), scale=True), modes, **kw)
    print(f"error mcmc: per mode {got['per_mode']} acceptance {got['acceptance']:.3f} calls {got['calls']} "
          f"map {got['map_pars']}", flush=True)
    assert got["per_mode"] == ref["per_mode"] and got["acceptance"] > 0.01
    np.testing.assert_allclose(got["samples"], ref["samples"], rtol=1e-9)
    np.testing.assert_allclose(got["map_pars"], ref["map_pars"], rtol=1e-9)
