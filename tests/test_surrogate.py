"""Learned surrogate (BF16 MLP on MFMA) and delayed-acceptance MCMC.

CPU tests pin the FP64 reference MLP of oracle/mlp_ref.py with central
finite differences, check bfloat16 rounding on known values, and check the
two-stage sampler on a Gaussian target whose moments are known exactly
(a biased, tempered surrogate must not move the posterior, only the number
of expensive evaluations).
GPU tests: the MFMA GEMM and the full forward/backward against the CPU
reference with the same bfloat16 roundings, training, and the surrogate
screening the engine's own posterior.
"""
import json

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import surrogate as S
from lcurve_re_amd import workloads as W


# --------------------------------------------------------------------- CPU
def test_reference_gradient_matches_finite_differences():
    from oracle import mlp_ref as R
    widths = [3, 7, 5, 4]
    rng = np.random.default_rng(3)
    p = rng.standard_normal(sum(widths[i] * widths[i + 1] + widths[i + 1] for i in range(3))) * 0.7
    x = rng.standard_normal((6, 3))
    t = rng.standard_normal((6, 4))
    loss, g = R.grad(p, widths, x, t)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd[i] = (R.grad(a, widths, x, t)[0] - R.grad(b, widths, x, t)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))
    assert abs(loss - np.mean((R.forward(p, widths, x)[0] - t) ** 2)) < 1e-15


def test_bf16_rounding():
    from oracle import mlp_ref as R
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38, 1.0 + 2.0 ** -7])
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    expect = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 2.9975e38, 1.0 + 2.0 ** -7])
    got = R.bf16(vals)
    assert np.all(np.abs(got - expect) <= np.abs(expect) * 2.0 ** -8)
    assert got[1] == 1.0 and got[2] == 1.0 + 2.0 ** -7


def test_encode_maps_prior_box_to_unit_cube():
    rng = np.random.Generator(np.random.PCG64(5))
    P = S.sample_prior(4000, rng)
    X = S.encode(P)
    assert X.shape == (4000, 12)
    j = [k for k in range(12) if S.NAMES[k] != "t0"]
    assert np.all(X[:, j] >= -1.0 - 1e-6) and np.all(X[:, j] <= 1.0 + 1e-6)
    assert np.all(np.abs(X[:, j].min(axis=0) + 1) < 0.05) and np.all(np.abs(X[:, j].max(axis=0) - 1) < 0.05)
    # r2 is drawn as a filling factor of the Roche lobe 1 - xl12(q)
    fill = P[:, 3] / np.array([1.0 - W.xl12(q) for q in P[:, 0]])
    assert np.all((fill >= 0.3 - 1e-9) & (fill <= 1.0 + 1e-9))


def test_model_array_sets_named_values():
    base = L.Model.from_json(json.dumps(W.model_json(vary=W.FIT_PARAMS)))
    P = np.array([[0.4, 84.0, 0.11, 0.21, 21000, 4100, 0.61, -0.21, 0.31, -0.11, 0.002, 0.09]])
    arr = S.model_array(base, P)
    m = base.copy()
    m.set_param(P[0])
    assert bytes(arr[0]) == bytes(m.pod)


def _gauss_problem():
    mean = np.array([1.0, -2.0])
    cov = np.array([[1.0, 0.8], [0.8, 2.0]])
    prec = np.linalg.inv(cov)

    def loglike(P):
        d = np.asarray(P) - mean
        return -0.5 * np.einsum("ij,jk,ik->i", d, prec, d)

    def legal(P):
        return np.ones(len(P), bool)

    return mean, cov, loglike, legal


@pytest.mark.parametrize("n_sub", [1, 5])
def test_delayed_acceptance_keeps_the_target(n_sub):
    mean, cov, loglike, legal = _gauss_problem()
    prec_s = np.linalg.inv(cov * 1.5)

    def sur(P):   # biased and too wide: a poor but informative screen
        d = np.asarray(P) - (mean + 0.3)
        return -0.5 * np.einsum("ij,jk,ik->i", d, prec_s, d)

    x0 = np.random.default_rng(1).standard_normal((32, 2)) + mean
    prop = 2.0 * cov
    plain, sp = S.run_chains(loglike, legal, x0, prop, 6000, seed=7)
    da, sd = S.run_chains(loglike, legal, x0, prop, 6000, seed=8, sur_loglike=sur, n_sub=n_sub)
    for ch in (plain, da):
        s = ch[500:].reshape(-1, 2)
        n_eff = S.ess(ch, 500).min()
        se = np.sqrt(np.diag(cov) / n_eff)
        assert np.all(np.abs(s.mean(axis=0) - mean) < 5 * se)
        assert np.allclose(np.cov(s.T), cov, rtol=0.12, atol=0.05)
    if n_sub == 1:
        # the screen rejects proposals without calling the expensive likelihood
        assert sd["evaluations"] < 0.7 * sp["evaluations"]
    else:
        # several surrogate moves per expensive evaluation: more accepted moves
        # per evaluation than the plain sampler
        assert sd["accepted"] / sd["evaluations"] > sp["accepted"] / sp["evaluations"]


@pytest.mark.parametrize("k_try", [1, 4])
def test_batched_first_stage_keeps_the_target(k_try):
    """run_chains_batched: the delayed-acceptance chain with each walker's
    first-stage trials screened k_try at a time; same target, and one
    expensive evaluation per round covers several chain steps."""
    mean, cov, loglike, legal = _gauss_problem()
    prec_s = np.linalg.inv(cov * 1.5)
    calls = []

    def counted(P):
        calls.append(len(P))
        return loglike(P)

    def sur(P):
        d = np.asarray(P) - (mean + 0.3)
        return -0.5 * np.einsum("ij,jk,ik->i", d, prec_s, d)

    x0 = np.random.default_rng(1).standard_normal((32, 2)) + mean
    ch, st = S.run_chains_batched(counted, legal, x0, 2.0 * cov, 6000, seed=9, sur_loglike=sur, k_try=k_try)
    assert ch.shape == (6000, 32, 2) and np.all(np.isfinite(ch))
    s = ch[500:].reshape(-1, 2)
    n_eff = S.ess(ch, 500).min()
    se = np.sqrt(np.diag(cov) / n_eff)
    assert np.all(np.abs(s.mean(axis=0) - mean) < 5 * se)
    assert np.allclose(np.cov(s.T), cov, rtol=0.12, atol=0.05)
    assert st["evaluations"] == sum(calls) and st["rounds"] == len(calls) - 1
    if k_try > 1:
        assert st["rounds"] < 0.6 * 6000


def test_batched_first_stage_is_the_stepwise_chain():
    """With an exact first stage (surrogate = target), the second stage
    accepts every first-stage acceptance; the chain's steps are then those
    of plain Metropolis driven by the first-stage draws: check the accepted
    fraction matches plain Metropolis with the same proposal."""
    mean, cov, loglike, legal = _gauss_problem()
    x0 = np.random.default_rng(2).standard_normal((64, 2)) + mean
    plain, sp = S.run_chains(loglike, legal, x0, 2.0 * cov, 3000, seed=3)
    ch, st = S.run_chains_batched(loglike, legal, x0, 2.0 * cov, 3000, seed=4, sur_loglike=loglike, k_try=6)
    assert st["accepted"] == st["surrogate_steps_accepted"] == st["evaluations"] - 64
    a0 = sp["accepted"] / sp["proposals"]
    a1 = st["accepted"] / st["proposals"]
    assert abs(a1 - a0) < 0.02, (a0, a1)


def test_autocorr_time_of_ar1():
    rng = np.random.default_rng(2)
    rho, n, W_ = 0.9, 20000, 8
    x = np.zeros((n, W_, 1))
    for i in range(1, n):
        x[i] = rho * x[i - 1] + rng.standard_normal((W_, 1))
    tau = S.autocorr_time(x)[0]
    assert abs(tau / ((1 + rho) / (1 - rho)) - 1) < 0.15


# --------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gemm_layer_matches_numpy():
    from oracle import mlp_ref as R
    widths = [40, 72]   # one linear layer; sizes off the 64 x 64 x 32 tiles
    rng = np.random.default_rng(11)
    p = (rng.standard_normal(40 * 72 + 72) * 0.3).astype(np.float32)
    x = rng.standard_normal((97, 40)).astype(np.float32)
    print("surrogate: single GEMM layer 97x40x72", flush=True)
    mlp = S.MLP(widths, max_batch=64)
    try:
        mlp.set_params(p)
        y = mlp.forward(x)
    finally:
        mlp.close()
    ref = R.forward(p.astype(np.float64), widths, x, bf=True)[0]
    assert np.max(np.abs(y - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [70, 256, 300])
def test_forward_backward_match_cpu_reference(batch):
    """Batches under 128 rows run on the 64 x 64 GEMM, larger ones on the
    128-row tiles with transposed operands (256: 16-byte loads where the
    widths allow; 300: rows that are not a multiple of 8, element loads and
    partial tiles)."""
    from oracle import mlp_ref as R
    widths = [12, 96, 80, 130]
    rng = np.random.default_rng(12)
    p = S.init_params(widths, 5)
    x = rng.uniform(-1, 1, (batch, 12)).astype(np.float32)
    t = rng.standard_normal((batch, 130)).astype(np.float32)
    print(f"surrogate: forward/backward 12-96-80-130, batch {batch}", flush=True)
    mlp = S.MLP(widths, max_batch=384)
    try:
        mlp.set_params(p)
        y = mlp.forward(x)
        loss, g = mlp.grad(x, t)
    finally:
        mlp.close()
    p64 = p.astype(np.float64)
    yr = R.forward(p64, widths, x, bf=True)[0]
    assert np.max(np.abs(y - yr)) <= 2e-3 * np.max(np.abs(yr))
    lr, gr = R.grad(p64, widths, x, t, bf=True)
    assert abs(loss / lr - 1) < 1e-4
    o = 0
    for l in range(1, len(widths)):
        k, n = widths[l - 1], widths[l]
        for part in (k * n, n):
            a, b = g[o:o + part], gr[o:o + part]
            assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), (l, part)
            o += part
    # and against exact FP64 arithmetic: within the bfloat16 rounding
    _, g64 = R.grad(p64, widths, x, t, bf=False)
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)


@pytest.mark.gpu
def test_training_reduces_loss():
    widths = [4, 64, 64, 16]
    rng = np.random.default_rng(13)
    x = rng.uniform(-1, 1, (4096, 4)).astype(np.float32)
    k = np.arange(16)[None, :]
    t = (np.sin(x[:, :1] * (1 + k / 4)) * x[:, 1:2] + 0.3 * x[:, 2:3] * np.cos(x[:, 3:4] * k / 3)).astype(np.float32)
    print("surrogate: toy training 4-64-64-16", flush=True)
    mlp = S.MLP(widths, max_batch=256)
    try:
        mlp.set_params(S.init_params(widths, 1))
        mlp.set_dataset(4096)
        mlp.load_dataset(0, x, t)
        before = mlp.eval_dataset(0, 4096)
        mlp.train(600, 256, 3e-3, 99)
        after = mlp.eval_dataset(0, 4096)
        y = mlp.forward(x[:512])
    finally:
        mlp.close()
    assert after < 0.05 * before, (before, after)
    assert abs(np.mean((y - t[:512]) ** 2) / after - 1) < 0.3


@pytest.mark.gpu
def test_chisq_kernel_matches_host():
    widths = [12, 64, 2048]
    rng = np.random.default_rng(21)
    mlp = S.MLP(widths, max_batch=64)
    try:
        mlp.set_params(S.init_params(widths, 2) * 0.1)
        mu = (0.05 * np.sin(np.arange(2048) * 2 * np.pi / 2048)).astype(np.float32)
        sur = S.Surrogate(mlp, mu, 0.02)
        P = S.sample_prior(40, np.random.Generator(np.random.PCG64(3)))
        P[:, 10] = rng.uniform(-0.01, 0.01, 40)
        t = np.sort(rng.uniform(0, 2, 900))
        f = 1.0 + 0.01 * rng.standard_normal(900)
        w = rng.uniform(0.5, 2, 900) / 1e-4
        w[::17] = 0.0
        print("surrogate: fused chi**2 kernel, 40 rows x 900 points", flush=True)
        sur.set_data((t, f, w))
        got = sur.chisq_gpu(P)
        ref = sur.chisq(P, (t, f, w))
    finally:
        mlp.close()
    np.testing.assert_allclose(got, ref, rtol=1e-9)


@pytest.mark.gpu
def test_surrogate_screens_engine_posterior(engine):
    """Small end-to-end run: a network trained on engine curves around the
    fit screens proposals of the detached-binary posterior; the two-stage
    chains keep the same posterior, and the first stage saves engine
    calls.  (Whether that buys effective samples depends on how well the
    network predicts chi**2 differences between neighbouring points; the
    bench measures it.)"""
    from lcurve_re_amd import fit as F
    cfg = W.fit_batch(nphase=1000)["model"]
    for key in ("nlat1f", "nlat2f", "nlat1c", "nlat2c"):
        cfg["model_parameters"][key] = "24"
    base = L.Model.from_json(json.dumps(cfg))
    t = np.arange(1000) / 1000.0
    curve = engine.light_curve_comp(base, L.make_data(t), scale=False, rdata=False, diagnostics=False)[0]
    data = L.make_data(t, ferr=1.5e-3, weight=1.0, flux=W.make_fit_flux(curve, 1.5e-3, 4))
    post = S.Posterior(engine, base, data)
    p0 = base.get_param()
    d = len(p0)
    J = F.FitBatch(base, data, F.engine_evaluator(engine), scale=True).jacobian()["J"]
    cov0 = np.linalg.inv(J.T @ J + np.diag(1.0 / (0.05 * np.abs(p0) + 1e-3) ** 2)) * (2.38 ** 2 / d)
    rng = np.random.default_rng(3)
    x0 = p0 + rng.standard_normal((32, d)) @ np.linalg.cholesky(cov0 / 100).T
    print("surrogate: tuning the plain proposal", flush=True)
    cov, x0, samples = S.tune_proposal(post.loglike, post.legal, x0, cov0 / 10, rounds=5, iters=40, seed=4)
    print("surrogate: local network on 8192 engine curves", flush=True)
    sur, info = S.train_local_surrogate(engine, base, samples, cov, 8192, 600, seed=5)
    print(f"surrogate: {info}", flush=True)
    assert info["local_valid_rms_relflux"] < 1e-4, info
    print("surrogate: plain, delayed-acceptance and surrogate-transition chains, 32 walkers x 150", flush=True)
    T, cal = S.calibrate_temperature(post, sur, x0, cov, 256, seed=6)
    print(f"surrogate: stage-1 temperature {T} {cal}", flush=True)
    sl = post.surrogate_loglike(sur, T)
    plain, sp = S.run_chains(post.loglike, post.legal, x0, cov, 150, seed=1)
    da, sd = S.run_chains(post.loglike, post.legal, x0, cov, 150, seed=2, sur_loglike=sl, n_sub=1)
    st, ss = S.run_chains(post.loglike, post.legal, x0, cov, 150, seed=3, sur_loglike=sl, n_sub=4)
    print(f"surrogate: plain {sp}\nsurrogate: DA {sd}\nsurrogate: ST {ss}", flush=True)
    # the first stage turns proposals away without an engine call
    assert sd["evaluations"] < sp["evaluations"]
    # and neither two-stage sampler moves the posterior
    a = plain[50:].reshape(-1, d)
    for ch in (da, st):
        diff = np.abs(ch[50:].reshape(-1, d).mean(axis=0) - a.mean(axis=0))
        assert np.all(diff < 1.0 * a.std(axis=0) + 1e-12), diff / a.std(axis=0)
