"""The batched adaptive-Metropolis sampler (lcurve_re_amd/mcmc.py, the
reference's mcmc_solver.cpp:196-1175; the physical prior's own checks
are in test_prior.py).

Pinned by: on a linear model y = A p the chi^2 is exactly quadratic, so
after burn-in the chains must sample N(p_hat, s2 (A^T W A)^-1), with p_hat
the weighted least-squares solution and s2 the data-error rescale the
sampler fixes at the end of burn-in: their mean and covariance are checked
against that closed form; a chain of a multi-chain run is identical to the
same chain run alone (each chain owns its generator); the covariance
proposal switches on after the warm-up and the acceptance lands near the
Robbins-Monro target.  CPU only (engine vs oracle chains:
test_gpu_mcmc.py)."""
import numpy as np

import lcurve_re_amd as L
from lcurve_re_amd import mcmc as M

X = np.linspace(0.0, 1.0, 200)
A = np.stack([np.ones_like(X), X, np.sin(6.0 * X)], axis=1)


class LinModel:
    def __init__(self, p):
        self.p = np.array(p, float)

    def copy(self):
        return LinModel(self.p.copy())

    def get_param(self):
        return self.p.copy()

    def set_param(self, v):
        self.p = np.array(v, float)

    def get_limit(self):
        return [(-100.0, 100.0)] * 3

    def get_dstep(self):
        return np.array([0.01, 0.01, 0.01])

    def nvary(self):
        return 3


def lin_eval(models, data, opts):
    return np.array([A @ m.p for m in models]), np.zeros(len(models), int)


def lin_data():
    rng = np.random.default_rng(4)
    sig = 0.05 + 0.05 * rng.random(len(X))
    y = A @ np.array([1.0, -0.5, 0.3]) + sig * rng.standard_normal(len(X))
    return L.make_data(X, flux=y, ferr=sig, weight=1.0), y, sig


def test_chains_sample_the_gaussian_posterior():
    data, y, sig = lin_data()
    W = 1.0 / sig ** 2
    cov = np.linalg.inv(A.T @ (A * W[:, None]))
    phat = cov @ (A.T @ (W * y))
    r = M.adaptive_metropolis(LinModel([0.8, -0.3, 0.2]), data, lin_eval, nchains=8, nsteps=6000, seed=3)
    s2 = r["chi2_scale"]  # each chain's best chi^2 / dof: nearly the minimum, so nearly equal
    assert np.all((s2 > 0.7) & (s2 < 1.4)) and np.ptp(s2) < 0.01 * s2.mean()
    samples = r["chain"].reshape(-1, 3)
    sd = np.sqrt(np.diag(cov * s2.mean()))
    assert np.all(np.abs(samples.mean(axis=0) - phat) < 0.15 * sd), (samples.mean(axis=0), phat, sd)
    emp = np.cov(samples.T)
    np.testing.assert_allclose(np.sqrt(np.diag(emp)), sd, rtol=0.2)
    corr_e = emp / np.outer(np.sqrt(np.diag(emp)), np.sqrt(np.diag(emp)))
    corr_t = cov / np.outer(np.sqrt(np.diag(cov)), np.sqrt(np.diag(cov)))
    assert np.max(np.abs(corr_e - corr_t)) < 0.15
    assert r["cov_ready"].all()
    assert np.all(np.abs(r["acceptance"] - 0.234) < 0.12)


def test_a_chain_does_not_depend_on_the_others():
    data, *_ = lin_data()
    many = M.adaptive_metropolis(LinModel([0.8, -0.3, 0.2]), data, lin_eval, nchains=3, nsteps=1500, seed=10,
                                 cov_warmup=300)
    for k in range(3):
        one = M.adaptive_metropolis(LinModel([0.8, -0.3, 0.2]), data, lin_eval, nchains=1, nsteps=1500,
                                    seed=10 + k, cov_warmup=300)
        assert np.array_equal(many["chain"][k], one["chain"][0])
        assert np.array_equal(many["chain_chisq"][k], one["chain_chisq"][0])


def test_gaussian_prior_shifts_the_posterior_to_the_product():
    """A Gaussian prior on p0 (untempered, unscaled by s2) multiplies the
    tempered-free likelihood: the posterior is N(m, C) with C^-1 =
    A^T W A / s2 + e0 e0^T / s0^2 and m = C (A^T W y / s2 + e0 m0 / s0^2)."""
    data, y, sig = lin_data()
    W = 1.0 / sig ** 2
    F = A.T @ (A * W[:, None])
    m0, s0 = 1.1, 0.01
    prior = lambda p: -0.5 * ((p[0] - m0) / s0) ** 2
    r = M.adaptive_metropolis(LinModel([1.0, -0.5, 0.3]), data, lin_eval, nchains=8, nsteps=6000, seed=5,
                              prior=prior)
    s2 = r["chi2_scale"].mean()
    P = F / s2
    P[0, 0] += 1.0 / s0 ** 2
    C = np.linalg.inv(P)
    b = A.T @ (W * y) / s2
    b[0] += m0 / s0 ** 2
    m = C @ b
    samples = r["chain"].reshape(-1, 3)
    sd = np.sqrt(np.diag(C))
    assert np.all(np.abs(samples.mean(axis=0) - m) < 0.15 * sd), (samples.mean(axis=0), m, sd)
    np.testing.assert_allclose(samples.std(axis=0), sd, rtol=0.2)
    lp = r["chain_log_prior"].reshape(-1)
    np.testing.assert_allclose(lp, [prior(v) for v in samples], rtol=1e-12, atol=1e-300)


def test_proposals_the_prior_forbids_are_never_evaluated():
    data, *_ = lin_data()
    bound = 1.0
    seen = []

    def ev(models, d, opts):
        seen.extend(m.p[0] for m in models)
        return lin_eval(models, d, opts)

    prior = lambda p: -1e30 if p[0] > bound else 0.0
    r = M.adaptive_metropolis(LinModel([0.95, -0.5, 0.3]), data, ev, nchains=4, nsteps=800, seed=6, prior=prior)
    assert max(seen[4:]) <= bound  # every evaluation after the starting points
    assert r["chain"][..., 0].max() <= bound
    assert np.all(r["chain_log_prior"] == 0.0)
