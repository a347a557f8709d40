"""The batched Levenberg-Marquardt driver (lcurve_re_amd/fit.py levmarq,
the inner loop of levmarq_solver.cpp:1225-1500).

Pinned by: on a toy nonlinear model (a exp(-b t) + c sin(2 pi t) + d) it
reaches the least-squares optimum that scipy's MINPACK lmder finds
(scipy.optimize.least_squares, method 'lm'); batched damping candidates
give exactly the iterates of the one-trial-at-a-time loop (candidates=1),
from a start that forces rejections; on a light curve evaluated by the
FP64 oracle it recovers the parameters the data were made with.  CPU only
(the GPU engine's version is in test_gpu_levmarq.py)."""
import json
import math

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import fit as F


class ToyModel:
    """Just the Model methods FitBatch uses."""

    def __init__(self, p, lims):
        self.p = np.array(p, float)
        self.lims = lims

    def copy(self):
        return ToyModel(self.p.copy(), self.lims)

    def get_param(self):
        return self.p.copy()

    def set_param(self, v):
        self.p = np.array(v, float)

    def get_limit(self):
        return self.lims

    def nvary(self):
        return len(self.p)


T = np.linspace(0.0, 3.0, 240)


def toy_curve(p):
    a, b, c, d = p
    return a * np.exp(-b * T) + c * np.sin(2 * np.pi * T) + d


def toy_eval(models, data, opts):
    return np.array([toy_curve(m.p) for m in models]), np.zeros(len(models), int)


def toy_problem():
    truth = np.array([2.0, 1.3, 0.4, 0.5])
    rng = np.random.default_rng(7)
    ferr = 0.01 + 0.01 * rng.random(len(T))
    flux = toy_curve(truth) + ferr * rng.standard_normal(len(T))
    weight = np.where(rng.random(len(T)) < 0.1, 0.0, 1.0 + rng.random(len(T)))
    data = L.make_data(T, ferr=ferr, flux=flux, weight=weight)
    lims = [(-10.0, 10.0), (0.01, 10.0), (-5.0, 5.0), (-5.0, 5.0)]
    return data, lims, ferr, flux, weight


def test_levmarq_reaches_the_minpack_optimum():
    from scipy.optimize import least_squares
    data, lims, ferr, flux, weight = toy_problem()
    start = np.array([1.0, 0.5, 0.0, 0.0])
    fb = F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False)
    res = F.levmarq(fb, max_iter=200)
    w = np.where(weight > 0, weight, 1.0)
    ref = least_squares(lambda p: np.sqrt(w) * (flux - toy_curve(p)) / ferr, start, method="lm",
                        xtol=1e-15, ftol=1e-15, gtol=1e-15)
    np.testing.assert_allclose(res["pars"], ref.x, rtol=1e-6, atol=1e-8)
    assert abs(res["chisq"] / float(np.sum(ref.fun ** 2)) - 1) <= 1e-10
    assert res["stop"] in ("ftol", "xtol", "noise floor"), res["stop"]


@pytest.mark.parametrize("k", [2, 4, 7])
def test_batched_candidates_give_the_sequential_iterates(k):
    data, lims, *_ = toy_problem()
    start = np.array([8.0, 5.0, -3.0, 3.0])  # far off: rejections happen
    seq = F.levmarq(F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False), tau0=1e-6, candidates=1)
    bat = F.levmarq(F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False), tau0=1e-6, candidates=k)
    assert any(h[3] == "reject" for h in seq["history"])
    assert seq["history"] == bat["history"]
    assert np.array_equal(seq["pars"], bat["pars"]) and seq["iterations"] == bat["iterations"]
    assert bat["calls"] < seq["calls"]


def test_max_fev_counts_and_stops_like_the_reference():
    """lm_max_fev (levmarq_solver.cpp:431, 649-693, 1050, 1494): every
    returned evaluation counts (the start point, the stage's starting point,
    2 per central column, each trial, the 8 column evaluations after an
    accepted trial); the test after an accepted or rejected iteration ends
    the fit.  Batched candidates count what the one-at-a-time loop counts,
    not what they evaluate."""
    data, lims, *_ = toy_problem()
    start = np.array([8.0, 5.0, -3.0, 3.0])
    fb = lambda: F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False)
    full = F.levmarq(fb(), tau0=1e-6, candidates=1, max_fev=10 ** 6)
    kinds = [h[3] for h in full["history"]]
    assert "reject" in kinds and kinds.count("accept") >= 3
    fev = 1 + 1 + 2 * 4
    after = []
    for kd in kinds:
        fev += 1 + (8 if kd in ("accept", "ftol") else 0)
        after.append(fev)
    assert full["fev"] == after[-1]
    assert fev < 200 * 5 and F.levmarq(fb(), tau0=1e-6)["stop"] == full["stop"]   # the default never binds here
    for k in (0, 3, 5):
        lim = after[k]
        for cand in (1, 4):
            r = F.levmarq(fb(), tau0=1e-6, candidates=cand, max_fev=lim)
            assert r["stop"] == "max function evaluations" and r["iterations"] == k + 1, (k, cand, r["stop"])
            assert r["history"] == full["history"][:k + 1] and r["fev"] == lim
        # one below: the same iteration already passes it only if the
        # previous count did not
        r = F.levmarq(fb(), tau0=1e-6, candidates=4, max_fev=lim - 1)
        assert r["iterations"] == (k + 1 if k == 0 or after[k - 1] < lim - 1 else k)


def test_levmarq_fits_an_oracle_light_curve():
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    phys = {"q": 0.4, "iangle": 84.0, "r1": 0.12, "r2": 0.25}
    comp = {"nlat1f": 16, "nlat2f": 16, "nlat1c": 16, "nlat2c": 16}
    truth = L.Model.from_json(json.dumps(W.model_json(phys=phys, comp=comp)))
    t = np.arange(300) / 300.0
    st, clean, _ = P.light_curve(truth.pod, L.make_data(t, ferr=1.0, flux=1.0), scale=False, rdata=False)
    assert st == 0
    ferr = 1e-3 * np.abs(clean).max() * np.ones_like(t)
    data = L.make_data(t, ferr=ferr, flux=clean)
    start = dict(phys, iangle=82.5, r1=0.125)
    m = L.Model.from_json(json.dumps(W.model_json(phys=start, comp=comp, vary=("iangle", "r1"))))
    fb = F.FitBatch(m, data, P.evaluator(), scale=False)
    res = F.levmarq(fb, max_iter=40)
    assert res["chisq"] < 1e-6 * float(np.sum(((clean - clean.mean()) / ferr) ** 2))
    np.testing.assert_allclose(res["pars"], [84.0, 0.12], rtol=2e-4)


def wave_curve(p):
    a, b = p
    return a * np.sin(b * T)


def wave_eval(models, data, opts):
    return np.array([wave_curve(m.p) for m in models]), np.zeros(len(models), int)


def test_multistart_runs_are_the_single_runs_and_find_the_global_minimum():
    """a sin(b t): chi^2 has many local minima in b; every start of the
    lockstep multistart ends where levmarq alone ends from it, in fewer
    engine calls than the runs one after another, and the best start finds
    the true b."""
    rng = np.random.default_rng(2)
    ferr = 0.05 * np.ones(len(T))
    data = L.make_data(T, ferr=ferr, flux=wave_curve([1.5, 7.0]) + ferr * rng.standard_normal(len(T)))
    lims = [(-10.0, 10.0), (0.1, 20.0)]
    starts = [np.array([1.0, b]) for b in (2.0, 4.5, 6.5, 9.0, 12.0)]
    fb = F.FitBatch(ToyModel(starts[0], lims), data, wave_eval, scale=False)
    res, calls = F.levmarq_multistart(fb, starts, max_iter=60)
    alone = [F.levmarq(fb, s, max_iter=60) for s in starts]
    for r in res:
        a = alone[r["start"]]
        assert np.array_equal(r["pars"], a["pars"]) and r["history"] == a["history"]
    assert calls < sum(a["calls"] for a in alone)
    assert len({round(r["pars"][1], 3) for r in res}) > 1  # distinct local minima
    assert abs(res[0]["pars"][1] - 7.0) < 0.01 and abs(res[0]["pars"][0] - 1.5) < 0.05


class ToyPrior:
    """One prior residual on b, the form LMPrior has: sqrt(weight) (b - mu) / sigma."""

    def __init__(self, mu, sigma, weight):
        self.mu, self.sigma, self.weight = mu, sigma, weight
        self.weight_sum = weight

    def __len__(self):
        return 1

    def __call__(self, p):
        return np.array([np.sqrt(self.weight) * (p[1] - self.mu) / self.sigma])


@pytest.mark.parametrize("weight", [2.0, 1e4])
def test_prior_continuation_reaches_the_augmented_optimum(weight):
    """levmarq_solver.cpp:985-1530: stages at prior scale 0, 1/6, .., 1, the
    balance factor min(chi2_LC / (n w), 1) at each stage's start; the last
    stage ends at the least-squares optimum of the data residuals plus the
    scaled prior residual (scipy MINPACK on the same augmented vector)."""
    from scipy.optimize import least_squares
    data, lims, ferr, flux, wt = toy_problem()
    w = np.where(wt > 0, wt, 1.0)
    pr = ToyPrior(1.5, 0.05, weight)
    start = np.array([1.0, 0.5, 0.0, 0.0])
    res = F.levmarq(F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False), prior=pr, max_iter=200)
    st = res["stages"]
    assert [s["scale"] for s in st] == [k / 6 for k in range(7)]
    assert st[0]["eff"] == 0.0 and np.array_equal(st[0]["start"], start)
    for s in st[1:]:
        chi_lc = float(np.sum(w * ((flux - toy_curve(s["start"])) / ferr) ** 2))
        assert s["eff"] ** 2 == pytest.approx(s["scale"] * min(chi_lc / weight, 1.0), rel=1e-12)
    eff = st[-1]["eff"]
    ref = least_squares(lambda p: np.concatenate([np.sqrt(w) * (flux - toy_curve(p)) / ferr, eff * pr(p)]),
                        st[-1]["start"], method="lm", xtol=1e-15, ftol=1e-15, gtol=1e-15)
    np.testing.assert_allclose(res["pars"], ref.x, rtol=1e-6, atol=1e-8)
    assert abs(res["chisq"] / float(np.sum(ref.fun ** 2)) - 1) <= 1e-10
    free = least_squares(lambda p: np.sqrt(w) * (flux - toy_curve(p)) / ferr, start, method="lm")
    assert abs(res["pars"][1] - 1.5) < abs(free.x[1] - 1.5)  # the prior pulls b toward 1.5
    assert res["iterations"] == sum(s["iterations"] for s in st)


def test_prior_candidates_give_the_sequential_iterates():
    data, lims, *_ = toy_problem()
    start = np.array([8.0, 5.0, -3.0, 3.0])
    pr = ToyPrior(1.5, 0.05, 10.0)
    seq = F.levmarq(F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False), prior=pr, tau0=1e-6,
                    candidates=1)
    bat = F.levmarq(F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False), prior=pr, tau0=1e-6,
                    candidates=4)
    assert seq["history"] == bat["history"] and np.array_equal(seq["pars"], bat["pars"])
    one = F.levmarq(F.FitBatch(ToyModel(start, lims), data, toy_eval, scale=False), prior=pr, tau0=1e-6,
                    continuation=False)
    assert [s["scale"] for s in one["stages"]] == [1.0]


# ------------------------------------------------------- modes of a multistart
XM = np.linspace(0.0, 1.0, 200)


def bimodal_curve(a):
    return a * a * (1.0 + 0.3 * a) * XM + 0.0075 * a * XM ** 2


def bimodal_eval(models, data, opts):
    return np.array([bimodal_curve(m.p[0]) for m in models]), np.zeros(len(models), int)


def test_laplace_modes_match_the_integrated_posterior_mass():
    """levmarq_solver.cpp:1640-1790 on y = a^2 (1 + 0.3 a) x + 0.0075 a x^2,
    whose chi^2 has three minima of different depth and width: the runs
    from nine starts fall into three modes, and each mode's weight is its
    share of the integral of exp(-0.5 chi^2 / s2) over its basin (scipy
    quadrature; the Laplace approximation is accurate for basins this
    narrow), s2 = chi^2_best / (N - 1)."""
    from scipy.integrate import quad
    rng = np.random.default_rng(11)
    sig = 0.02
    flux = bimodal_curve(0.9) + sig * rng.standard_normal(len(XM))
    data = L.make_data(XM, ferr=sig, flux=flux, weight=1.0)
    lims = [(-5.0, 5.0)]
    fb = F.FitBatch(ToyModel([0.5], lims), data, bimodal_eval, scale=False)
    starts = [[s] for s in (0.6, 0.8, 1.2, -0.8, -1.3, -1.0, -2.6, -3.3, -2.9)]
    runs, calls = F.levmarq_multistart(fb, starts, max_iter=200)
    modes, mcalls = F.laplace_modes(fb, runs, dsteps=[0.01])
    assert len(modes) == 3 and sorted(m["n_runs"] for m in modes) == [3, 3, 3] and mcalls == 1
    chi = lambda a: float(np.sum(((flux - bimodal_curve(a)) / sig) ** 2))
    s2 = min(chi(m["pars"][0]) for m in modes) / (len(XM) - 1)
    c0 = min(chi(m["pars"][0]) for m in modes)
    mass = []
    for m in modes:
        a0, w = m["pars"][0], math.sqrt(m["cov"][0, 0])
        mass.append(quad(lambda a: math.exp(-0.5 * (chi(a) - c0) / s2), a0 - 12 * w, a0 + 12 * w,
                         points=[a0], limit=200)[0])
    mass = np.array(mass) / sum(mass)
    got = np.array([m["weight"] for m in modes])
    assert got.max() / got.min() > 1.5  # the modes differ in mass
    np.testing.assert_allclose(got, mass, rtol=2e-3)
    assert abs(sum(got) - 1.0) < 1e-12 and list(got) == sorted(got, reverse=True)


# ------------------------------------------------------- error MCMC over modes
def test_error_mcmc_samples_the_gaussian_posterior_of_a_linear_fit():
    """levmarq_solver.cpp:2090-2500 on a linear model y = A p: the target
    exp(-0.5 chi^2 / s2) is N(p_hat, s2 (A^T W A)^-1), s2 = chi2_min /
    (N - k)."""
    from lcurve_re_amd import mcmc as M
    X = np.linspace(0.0, 1.0, 150)
    A = np.stack([np.ones_like(X), X, np.cos(5.0 * X)], axis=1)
    rng = np.random.default_rng(2)
    sig = 0.03 + 0.02 * rng.random(len(X))
    y = A @ np.array([0.5, 1.0, -0.2]) + sig * rng.standard_normal(len(X))
    data = L.make_data(X, ferr=sig, flux=y, weight=1.0)
    ev = lambda models, d, o: (np.array([A @ m.p for m in models]), np.zeros(len(models), int))
    fb = F.FitBatch(ToyModel([0.4, 0.9, 0.0], [(-10.0, 10.0)] * 3), data, ev, scale=False)
    runs, _ = F.levmarq_multistart(fb, [[0.4, 0.9, 0.0], [0.6, 1.1, -0.3]])
    modes, _ = F.laplace_modes(fb, runs, dsteps=[0.01] * 3)
    assert len(modes) == 1 and modes[0]["n_runs"] == 2
    r = M.error_mcmc(fb, modes, dsteps=[0.01] * 3, steps=12000, seed=4)
    W = 1.0 / sig ** 2
    F_ = A.T @ (A * W[:, None])
    phat = np.linalg.solve(F_, A.T @ (W * y))
    s2 = float(np.sum(W * (y - A @ phat) ** 2)) / (len(X) - 3)
    assert r["err_s2"] == pytest.approx(s2, rel=1e-8)
    C = s2 * np.linalg.inv(F_)
    sd = np.sqrt(np.diag(C))
    assert r["per_mode"][0][1] == 12000 and r["ok"]
    assert np.all(np.abs(r["samples"].mean(axis=0) - phat) < 0.15 * sd)
    np.testing.assert_allclose(r["samples"].std(axis=0), sd, rtol=0.15)
    assert 0.1 < r["acceptance"] < 0.5
    assert -0.5 * r["map_chisq"] >= -0.5 * float(np.sum(W * (y - A @ modes[0]["pars"]) ** 2)) - 1e-9


def test_error_mcmc_splits_steps_over_modes_by_weight():
    from lcurve_re_amd import mcmc as M
    rng = np.random.default_rng(11)
    sig = 0.02
    flux = bimodal_curve(0.9) + sig * rng.standard_normal(len(XM))
    data = L.make_data(XM, ferr=sig, flux=flux, weight=1.0)
    fb = F.FitBatch(ToyModel([0.5], [(-5.0, 5.0)]), data, bimodal_eval, scale=False)
    runs, _ = F.levmarq_multistart(fb, [[s] for s in (0.8, -1.3, -2.9)])
    modes, _ = F.laplace_modes(fb, runs, dsteps=[0.01])
    r = M.error_mcmc(fb, modes, dsteps=[0.01], steps=3000, seed=8)
    w = np.array([m["weight"] for m in modes])
    kept = [n for _, n, _ in r["per_mode"]]
    assert kept == [max(500, int(3000 * wk / w.sum())) for wk in w]
    o = 0
    for (k, n, _), m in zip(r["per_mode"], modes):
        s = r["samples"][o:o + n, 0]
        o += n
        assert np.all(np.abs(s - m["pars"][0]) < 10 * math.sqrt(m["cov"][0, 0]))  # each chain stays in its basin
    one = M.error_mcmc(fb, modes[:1], dsteps=[0.01], steps=kept[0], min_steps=kept[0], seed=8)
    assert np.array_equal(one["samples"], r["samples"][:kept[0]])  # a chain does not depend on the others
    bound = modes[0]["pars"][0] + 0.5 * math.sqrt(modes[0]["cov"][0, 0])
    capped = M.error_mcmc(fb, modes[:1], dsteps=[0.01], steps=1000, seed=3,
                          log_prior=lambda p: -1e30 if p[0] > bound else 0.0)
    assert capped["samples"][:, 0].max() <= bound


# ------------------------------------------------------- multistart starts
def test_cxx_uniform_draws_match_libstdcxx():
    """CxxMt19937Uniform against std::mt19937 + std::uniform_real_distribution
    compiled here with g++ (the generator the reference's multistart uses)."""
    import shutil
    import subprocess
    import tempfile
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = ("#include <cstdio>\n#include <random>\nint main(){for(int seed: {44, 7, 123456}){std::mt19937 g(seed);"
           "std::uniform_real_distribution<> u(0.0,1.0);for(int i=0;i<40;++i)std::printf(\"%.17g\\n\",u(g));}}\n")
    with tempfile.TemporaryDirectory() as d:
        open(f"{d}/u.cpp", "w").write(src)
        subprocess.run(["g++", "-O1", "-o", f"{d}/u", f"{d}/u.cpp"], check=True)
        want = [float(x) for x in subprocess.run([f"{d}/u"], check=True, capture_output=True,
                                                  text=True).stdout.split()]
    got = []
    for seed in (44, 7, 123456):
        u = F.CxxMt19937Uniform(seed)
        got += [u() for _ in range(40)]
    assert got == want


def test_multistart_starts_are_inside_the_limits_and_stratified():
    m = L.Model.from_json(json.dumps(__import__("lcurve_re_amd.workloads", fromlist=["x"]).model_json(
        vary=("q", "iangle", "r1", "r2", "t2"))))
    names = [m.get_name(i) for i in range(m.nvary())]
    p0 = np.array(m.get_param(), float)
    st = F.multistart_starts(m, n=8, span=0.5, seed=42)
    assert len(st) == 9 and np.array_equal(st[0], p0)
    lims, rng = m.get_limit(), np.asarray(m.get_range(), float)
    ii = names.index("iangle")
    ilo, ihi = max(lims[ii][0], 5.0), min(lims[ii][1], 90.0)
    for s, v in enumerate(st[1:]):
        for j in range(len(v)):
            assert lims[j][0] <= v[j] <= lims[j][1]
            if j != ii:
                assert abs(v[j] - p0[j]) <= 0.5 * rng[j] + 1e-15
        frac = (v[ii] - ilo) / (ihi - ilo) * 8 - s
        assert 0.2 <= frac <= 0.8
    again = F.multistart_starts(m, n=8, span=0.5, seed=42)
    assert all(np.array_equal(a, b) for a, b in zip(st, again))


def test_error_mcmc_reanchors_on_a_better_state():
    """levmarq_solver.cpp:2390-2470: a mode reported 5 sigma off the
    optimum; the first round's chains find states far better than it
    (gain > 0.5), so the sampling restarts from the MAP, and the second
    round, started there, gains little and ends it.  The kept samples are
    the second round's, of the right posterior."""
    from lcurve_re_amd import mcmc as M
    X = np.linspace(0.0, 1.0, 150)
    A = np.stack([np.ones_like(X), X], axis=1)
    rng = np.random.default_rng(5)
    sig = 0.05 * np.ones_like(X)
    y = A @ np.array([0.3, -0.7]) + sig * rng.standard_normal(len(X))
    data = L.make_data(X, ferr=sig, flux=y, weight=1.0)
    ev = lambda models, d, o: (np.array([A @ m.p for m in models]), np.zeros(len(models), int))
    fb = F.FitBatch(ToyModel([0.3, -0.7], [(-10.0, 10.0)] * 2), data, ev, scale=False)
    W = 1.0 / sig ** 2
    Fm = A.T @ (A * W[:, None])
    phat = np.linalg.solve(Fm, A.T @ (W * y))
    chi_min = float(np.sum(W * (y - A @ phat) ** 2))
    s2 = chi_min / (len(X) - 2)
    C = s2 * np.linalg.inv(Fm)
    off = phat + 5.0 * np.sqrt(np.diag(C))
    chi_off = float(np.sum(W * (y - A @ off) ** 2))
    mode = {"pars": off, "weight": 1.0, "cov": C, "chisq_lc": chi_min}
    r = M.error_mcmc(fb, [mode], dsteps=[0.01, 0.01], steps=6000, seed=2)
    assert r["reanchored"] and r["rounds"] == 2 and r["gains"][0] > 0.5 >= r["gains"][1]
    assert r["map_chisq"] < chi_off and r["map_chisq"] - chi_min < 5.0
    sd = np.sqrt(np.diag(C))
    assert np.all(np.abs(r["samples"].mean(axis=0) - phat) < 0.2 * sd)
    assert r["converged"]
    one = M.error_mcmc(fb, [mode], dsteps=[0.01, 0.01], steps=6000, seed=2, reanchor_rounds=0)
    assert one["rounds"] == 1 and not one["reanchored"]
    # the one round allowed still gained: the anchor is kept and the fit is
    # reported as not converged (levmarq_solver.cpp:2438-2458)
    assert not one["converged"] and one["gains"][0] > 0.5
    assert np.array_equal(one["map_pars"], off) and abs(one["map_chisq"] - chi_off) <= 1e-9 * chi_off
    assert one["best_chisq"] < one["map_chisq"]


def _linear_posterior(n=150, seed=5):
    from lcurve_re_amd import mcmc as M
    X = np.linspace(0.0, 1.0, n)
    A = np.stack([np.ones_like(X), X], axis=1)
    rng = np.random.default_rng(seed)
    sig = 0.05 * np.ones_like(X)
    y = A @ np.array([0.3, -0.7]) + sig * rng.standard_normal(len(X))
    data = L.make_data(X, ferr=sig, flux=y, weight=1.0)
    W = 1.0 / sig ** 2
    Fm = A.T @ (A * W[:, None])
    phat = np.linalg.solve(Fm, A.T @ (W * y))
    chi_min = float(np.sum(W * (y - A @ phat) ** 2))
    C = chi_min / (n - 2) * np.linalg.inv(Fm)
    return M, A, data, phat, chi_min, C


def test_error_mcmc_skips_a_mode_that_does_not_evaluate():
    """levmarq_solver.cpp:2412-2418: a mode whose start fails is skipped
    with no samples, the others are sampled."""
    M, A, data, phat, chi_min, C = _linear_posterior()
    bad = phat + np.array([3.0, 0.0])

    def ev(models, d, o):
        st = np.array([1 if m.p[0] > 2.0 else 0 for m in models])
        return np.array([A @ m.p for m in models]), st
    fb = F.FitBatch(ToyModel(list(phat), [(-10.0, 10.0)] * 2), data, ev, scale=False)
    modes = [{"pars": phat, "weight": 0.7, "cov": C, "chisq_lc": chi_min},
             {"pars": bad, "weight": 0.3, "cov": C, "chisq_lc": chi_min}]
    r = M.error_mcmc(fb, modes, dsteps=[0.01, 0.01], steps=2000, seed=3)
    kept = {k: n for k, n, _ in r["per_mode"]}
    assert kept[1] == 0 and kept[0] == 1400
    assert r["converged"] and r["ok"]
    assert np.array_equal(r["map_pars"], phat)


def test_error_mcmc_global_covariance_is_the_fallback_proposal():
    """A mode without its own covariance proposes from the LM covariance
    (pcov) rather than from the diagonal steps: with steps 1e-4 sigma the
    diagonal chain stays narrow even with its scale adapted to 1e3, the
    covariance chain samples the closed-form posterior."""
    M, A, data, phat, chi_min, C = _linear_posterior()
    ev = lambda models, d, o: (np.array([A @ m.p for m in models]), np.zeros(len(models), int))
    fb = F.FitBatch(ToyModel(list(phat), [(-10.0, 10.0)] * 2), data, ev, scale=False)
    sd = np.sqrt(np.diag(C))
    mode = {"pars": phat, "weight": 1.0, "cov": None, "chisq_lc": chi_min}
    tiny = M.error_mcmc(fb, [mode], dsteps=list(1e-4 * sd), steps=3000, seed=6)
    full = M.error_mcmc(fb, [mode], dsteps=list(1e-4 * sd), steps=3000, seed=6, cov=lambda p: C)
    assert np.all(tiny["samples"].std(axis=0) < 0.5 * sd)
    assert np.all(np.abs(full["samples"].std(axis=0) / sd - 1.0) < 0.2)
    assert 0.1 < full["acceptance"] < 0.6


def test_multistart_survives_starts_that_do_not_evaluate():
    """A start whose evaluation fails is a failed run (ok False) and the
    others go on; laplace_modes leaves it out."""
    data, lims, *_ = toy_problem()

    def ev(models, d, opts):
        calc, st = toy_eval(models, d, opts)
        return calc, np.array([1 if m.p[3] > 4.0 else 0 for m in models])

    fb = F.FitBatch(ToyModel([1.0, 0.5, 0.0, 0.0], lims), data, ev, scale=False)
    starts = [[1.0, 0.5, 0.0, 0.0], [1.0, 0.5, 0.0, 4.5], [2.2, 1.2, 0.3, 0.4]]
    runs, _ = F.levmarq_multistart(fb, starts, max_iter=50)
    bad = [r for r in runs if not r["ok"]]
    assert len(bad) == 1 and bad[0]["start"] == 1 and runs[-1] is bad[0]
    assert all(r["ok"] for r in runs[:2])
    modes, _ = F.laplace_modes(fb, runs, dsteps=[1e-3] * 4)
    assert sum(m["n_runs"] for m in modes) == 2
