"""Learned surrogate of the light curve and delayed-acceptance MCMC.

The surrogate is a BF16 MLP (12 inputs -> 4 x 512 SiLU -> 2048 outputs)
trained on the matrix cores by the engine's own kernels (lc_mlp_* of the C
ABI, csrc/k_mlp.hip) on light curves that the engine itself computes.  It
maps the 12 fit parameters of the detached-binary problem (q, iangle, r1,
r2, t1, t2, ldc1_1..4, t0, gravity_dark2; the reference's vary order) to the
standardised log flux at 2048 phase bins k/2048.  t0 is applied as an exact
phase shift when the curve is mapped to the data times, so the network
sees t0 = 0.

The sampler is the two-stage Metropolis-Hastings of Christen & Fox (2005,
J. Comput. Graph. Stat. 14:795): a proposal y from x is first screened with
the surrogate posterior pi*, accepted with min(1, pi*(y)/pi*(x)); only then
is the engine run, and y is accepted with
min(1, pi(y) pi*(x) / (pi(x) pi*(y))).  For a symmetric proposal the chain
leaves the exact posterior pi invariant whatever the surrogate's errors;
the surrogate only decides how many engine evaluations are spent.  Both
stages run for all walkers at once (one MLP forward, one engine batch).
"""
import ctypes as C
import math
import time

import numpy as np

from . import LC_OK, LcurveError, _abi, lc_model, make_data

WIDTHS = [12, 512, 512, 512, 512, 2048]
NBIN = 2048

# Fit parameters (reference vary order, model.cpp:363-461) and the training
# prior: the batched-evaluation prior (q, i, r1, r2 as a Roche-lobe filling
# factor, T1, T2) extended with gravity_dark2, the star-1 limb darkening and
# t0, and with the filling factor reaching down to 0.3 so that the detached
# test binary (r2 = 0.2 at q = 0.5, f = 0.47) lies inside it.
NAMES = ["q", "iangle", "r1", "r2", "t1", "t2", "ldc1_1", "ldc1_2", "ldc1_3", "ldc1_4", "t0", "gravity_dark2"]
PRIOR = {"q": (0.1, 1.0), "iangle": (60.0, 90.0), "log_r1": (math.log(0.01), math.log(0.3)), "fill2": (0.3, 1.0),
         "t1": (8000.0, 40000.0), "t2": (3000.0, 8000.0), "ldc1_1": (0.5, 0.7), "ldc1_2": (-0.3, -0.1),
         "ldc1_3": (0.2, 0.4), "ldc1_4": (-0.2, 0.0), "t0": (-0.01, 0.01), "gravity_dark2": (0.05, 0.35)}
ENC = ["q", "iangle", "log_r1", "fill2", "t1", "t2", "ldc1_1", "ldc1_2", "ldc1_3", "ldc1_4", "t0", "gravity_dark2"]


def _lib():
    return _abi.load()


# ------------------------------------------------------------------ MLP
class MLP:
    """One lc_mlp: BF16 network with FP32 masters and Adam state on the GPU."""

    def __init__(self, widths=WIDTHS, max_batch=4096, device=0):
        self.widths = list(widths)
        w = (C.c_int32 * len(widths))(*widths)
        self._h = C.c_void_p()
        st = _lib().lc_mlp_create(int(device), w, len(widths), int(max_batch), C.byref(self._h))
        if st != LC_OK:
            raise LcurveError(st, "lc_mlp_create: " + _lib().lc_thread_error().decode())
        self.max_batch = max_batch

    def close(self):
        if self._h:
            _lib().lc_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != LC_OK:
            raise LcurveError(st, _lib().lc_mlp_last_error(self._h).decode())

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    @property
    def n_params(self):
        return int(_lib().lc_mlp_num_params(self._h))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(_lib().lc_mlp_set_params(self._h, self._fp(p), p.size))

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        self._chk(_lib().lc_mlp_get_params(self._h, self._fp(p), p.size))
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.widths[-1]), np.float32)
        self._chk(_lib().lc_mlp_forward(self._h, self._fp(x), x.shape[0], self._fp(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        self._chk(_lib().lc_mlp_grad(self._h, self._fp(x), self._fp(t), x.shape[0], self._fp(g), C.byref(loss)))
        return loss.value, g

    def set_dataset(self, nrows):
        self._chk(_lib().lc_mlp_set_dataset(self._h, int(nrows)))

    def load_dataset(self, row0, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        self._chk(_lib().lc_mlp_load_dataset(self._h, int(row0), self._fp(x), self._fp(t), x.shape[0]))

    def train(self, steps, batch, lr, seed):
        loss = C.c_double()
        self._chk(_lib().lc_mlp_train(self._h, int(steps), int(batch), float(lr), int(seed), C.byref(loss)))
        return loss.value

    def eval_dataset(self, row0, count):
        mse = C.c_double()
        self._chk(_lib().lc_mlp_eval_dataset(self._h, int(row0), int(count), C.byref(mse)))
        return mse.value

    def stats(self):
        f, ms = C.c_double(), C.c_double()
        self._chk(_lib().lc_mlp_stats(self._h, C.byref(f), C.byref(ms)))
        return {"gemm_flop": f.value, "train_ms": ms.value}

    def set_skip(self, S):
        """FP32 linear skip (inputs x outputs) added at inference; None removes it."""
        if S is None:
            self._chk(_lib().lc_mlp_set_skip(self._h, None, 0))
            return
        S = np.ascontiguousarray(S, np.float32)
        self._chk(_lib().lc_mlp_set_skip(self._h, self._fp(S), S.size))

    def set_likelihood(self, time_, flux, weight, mu, scale, period=1.0, log_output=True):
        t = np.ascontiguousarray(time_, np.float64)
        f = np.ascontiguousarray(flux, np.float64)
        w = np.ascontiguousarray(weight, np.float64)
        mu = np.ascontiguousarray(mu, np.float32)
        dp = C.POINTER(C.c_double)
        self._chk(_lib().lc_mlp_set_likelihood(self._h, t.ctypes.data_as(dp), f.ctypes.data_as(dp),
                                               w.ctypes.data_as(dp), t.size, self._fp(mu), float(scale),
                                               float(period), int(bool(log_output))))

    def chisq(self, x, t0):
        x = np.ascontiguousarray(x, np.float32)
        t0 = np.ascontiguousarray(t0, np.float64)
        out = np.zeros(x.shape[0], np.float64)
        dp = C.POINTER(C.c_double)
        self._chk(_lib().lc_mlp_chisq(self._h, self._fp(x), t0.ctypes.data_as(dp), x.shape[0],
                                      out.ctypes.data_as(dp)))
        return out


def init_params(widths, seed):
    """Normal weights with std sqrt(2 / fan_in), zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for l in range(1, len(widths)):
        k, n = widths[l - 1], widths[l]
        parts.append(rng.standard_normal(k * n) * math.sqrt(2.0 / k))
        parts.append(np.zeros(n))
    return np.concatenate(parts).astype(np.float32)


# ------------------------------------------------------------ parameters
def xl12_vec(q, spin=1.0):
    """L1 distance from star 1 for star 2 spinning at `spin` (roche.cpp:45-80), vectorised."""
    q = np.asarray(q, np.float64)
    ssq = spin * spin
    mu = q / (1.0 + q)
    a = [-1.0 + mu, 2.0 - 2.0 * mu, -ssq + mu, 3.0 * ssq + 2.0 * mu - 2.0, 1.0 - mu - 3.0 * ssq,
         np.full_like(q, ssq)]
    x = 1.0 / (1.0 + q)
    for _ in range(100):
        f = x * (x * (x * (x * (x * a[5] + a[4]) + a[3]) + a[2]) + a[1]) + a[0]
        df = x * (x * (x * (x * 5 * a[5] + 4 * a[4]) + 3 * a[3]) + 2 * a[2]) + a[1]
        x = x - f / df
    return x


def sample_prior(n, rng):
    """n x 12 parameter vectors (NAMES order) from PRIOR; t0 = 0."""
    u = {k: rng.uniform(lo, hi, n) for k, (lo, hi) in PRIOR.items()}
    P = np.zeros((n, len(NAMES)))
    P[:, 0] = u["q"]
    P[:, 1] = u["iangle"]
    P[:, 2] = np.exp(u["log_r1"])
    P[:, 3] = u["fill2"] * (1.0 - xl12_vec(u["q"]))
    P[:, 4] = u["t1"]
    P[:, 5] = u["t2"]
    for j, k in enumerate(["ldc1_1", "ldc1_2", "ldc1_3", "ldc1_4"]):
        P[:, 6 + j] = u[k]
    P[:, 10] = 0.0
    P[:, 11] = u["gravity_dark2"]
    return P


def encode(P):
    """Network inputs in [-1, 1] over the prior box (r1 as log r1, r2 as the
    filling factor of the Roche lobe)."""
    P = np.asarray(P, np.float64)
    raw = P.copy()
    raw[:, 2] = np.log(P[:, 2])
    raw[:, 3] = P[:, 3] / (1.0 - xl12_vec(P[:, 0]))
    X = np.empty_like(raw)
    for j, k in enumerate(ENC):
        lo, hi = PRIOR[k]
        X[:, j] = 2.0 * (raw[:, j] - lo) / (hi - lo) - 1.0
    return X.astype(np.float32)


def model_array(base, P, names=NAMES):
    """ctypes array of lc_model: `base` with the named values set from P."""
    P = np.asarray(P, np.float64)
    m = P.shape[0]
    arr = (lc_model * m)()
    sz = C.sizeof(lc_model)
    raw = np.frombuffer(arr, dtype=np.uint8).reshape(m, sz)
    raw[:] = np.frombuffer(C.string_at(C.addressof(base.pod), sz), dtype=np.uint8)
    for j, k in enumerate(names):
        off = getattr(lc_model, k).offset   # lc_pparam.value is its first field
        v = np.ndarray((m,), np.float64, buffer=arr, offset=off, strides=(sz,))
        v[:] = P[:, j]
    return arr


def evaluate(engine, base, P, data, scale=False, rdata=False, precision=0, max_model_points=0):
    """Engine light curves of the parameter rows P: (calc[m, n], chisq[m], status[m])."""
    pods = model_array(base, P)
    m, n = len(pods), len(data)
    calc = np.zeros((m, n), np.float64)
    io = (_abi.lc_scalars * m)()
    for i in range(m):
        for k in range(5):
            io[i].sfac[k] = 1.0
    status = np.zeros(m, np.int32)
    o = _abi.lc_eval_opts(int(scale), int(rdata), 0, 0, int(max_model_points), int(precision))
    st = _lib().lc_light_curve_comp_batch(engine._ctx, pods, m, data, n, C.byref(o),
                                          calc.ctypes.data_as(C.POINTER(C.c_double)), io,
                                          status.ctypes.data_as(C.POINTER(C.c_int32)))
    if st != LC_OK:
        engine._err(st)
    chisq = np.array([io[i].chisq for i in range(m)])
    return calc, chisq, status


# ------------------------------------------------------------ training data
def generate_curves(engine, base, n, seed, batch=4096, precision=0):
    """n (P, X, Y) rows: prior draws, network inputs and log(F / mean F) at
    the 2048 phase bins; draws the engine rejects are replaced."""
    rng = np.random.Generator(np.random.PCG64(seed))
    bins = make_data(np.arange(NBIN, dtype=np.float64) / NBIN)
    Ps, Ys, got, rejected = [], [], 0, 0
    while got < n:
        P = sample_prior(batch, rng)
        calc, _, status = evaluate(engine, base, P, bins, precision=precision)
        ok = (status == 0) & np.all(calc > 0, axis=1)
        rejected += int(np.count_nonzero(~ok))
        P, calc = P[ok], calc[ok]
        Y = np.log(calc / calc.mean(axis=1, keepdims=True))
        take = min(n - got, P.shape[0])
        Ps.append(P[:take])
        Ys.append(Y[:take].astype(np.float32))
        got += take
    P = np.concatenate(Ps)
    return P, encode(P), np.concatenate(Ys), rejected


class Surrogate:
    """Trained network plus its input encoding and output standardisation.

    Global networks see encode(P) (the prior box) and give the log flux at
    the 2048 phase bins, log(F / mean F) = y * scale + mu; local networks
    see (P - center) / spread and give the relative flux F / mean F =
    y * scale + mu itself (errors then weigh as they do in chi**2).  t0
    never enters the network: it shifts the phase of the data times
    exactly."""

    def __init__(self, mlp, mu, scale, center=None, spread=None, log_output=True):
        self.mlp = mlp
        self.mu = np.asarray(mu, np.float32)
        self.scale = float(scale)
        self.center = None if center is None else np.asarray(center, np.float64)
        self.spread = None if spread is None else np.asarray(spread, np.float64)
        self.log_output = log_output

    def inputs(self, P):
        P = np.array(P, np.float64)
        P[:, 10] = 0.0 if self.center is None else self.center[10]
        if self.center is None:
            return encode(P)
        return ((P - self.center) / self.spread).astype(np.float32)

    def output(self, P):
        """m x 2048 network curve at phases k / 2048 (t0 = 0): log(F / mean F)
        or F / mean F (log_output False)."""
        return self.mlp.forward(self.inputs(P)).astype(np.float64) * self.scale + self.mu

    def rel_flux(self, P):
        v = self.output(P)
        return np.exp(v) if self.log_output else v

    def flux_at(self, P, times, period=1.0):
        """m x n relative flux at the data times: the t0 of each row shifts
        the phase exactly; linear interpolation between the bins."""
        P = np.asarray(P, np.float64)
        yl = self.rel_flux(P)
        ph = ((np.asarray(times)[None, :] - P[:, 10:11]) / period) % 1.0
        x = ph * NBIN
        i0 = np.floor(x).astype(np.int64) % NBIN
        f = x - np.floor(x)
        i1 = (i0 + 1) % NBIN
        rows = np.arange(P.shape[0])[:, None]
        return yl[rows, i0] * (1.0 - f) + yl[rows, i1] * f

    def chisq(self, P, data_arrays):
        """chi**2 with the optimal scale factor (re_scale.cpp:16-43), on the
        host: the reference for chisq_gpu."""
        t, f, w = data_arrays
        m = self.flux_at(P, t)
        sw = np.sum(w * f * m, axis=1) / np.sum(w * m * m, axis=1)
        r = f[None, :] - sw[:, None] * m
        return np.sum(w * r * r, axis=1)

    def set_data(self, data_arrays, period=1.0):
        t, f, w = data_arrays
        self.mlp.set_likelihood(t, f, w, self.mu, self.scale, period, self.log_output)

    def chisq_gpu(self, P):
        """chi**2 of the rows of P against the data of set_data, on the GPU
        (network forward + one fused interpolation/scale/chi**2 kernel)."""
        P = np.asarray(P, np.float64)
        return self.mlp.chisq(self.inputs(P), P[:, 10])


def train_surrogate(engine, base, n_curves, seed=20261206, steps=3000, batch=4096, lr=1e-3, lr_final=2e-5,
                    n_valid=8192, widths=WIDTHS, log=None):
    """Generates n_curves engine curves from the prior, trains the network on
    them (Adam, cosine-decayed learning rate, minibatches drawn on the
    device) and measures the error on n_valid separately generated curves."""
    t0 = time.perf_counter()
    P, X, Y, rejected = generate_curves(engine, base, n_curves, seed)
    t_gen = time.perf_counter() - t0
    mu = Y.mean(axis=0)
    sc = float(np.std(Y - mu))
    T = (Y - mu) / sc
    mlp = MLP(widths, max_batch=batch)
    mlp.set_params(init_params(widths, seed))
    mlp.set_dataset(P.shape[0])
    for r0 in range(0, P.shape[0], 65536):
        mlp.load_dataset(r0, X[r0:r0 + 65536], T[r0:r0 + 65536])
    del X, T, Y
    chunk = 100
    losses = []
    t1 = time.perf_counter()
    for s0 in range(0, steps, chunk):
        frac = s0 / max(steps, 1)
        lr_s = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
        losses.append(mlp.train(min(chunk, steps - s0), batch, lr_s, seed + s0))
        if log:
            log(f"surrogate: step {min(s0 + chunk, steps)} loss {losses[-1]:.3e} lr {lr_s:.2e}")
    t_train = time.perf_counter() - t1
    st = mlp.stats()
    sur = Surrogate(mlp, mu, sc)
    Pv, _, Yv, _ = generate_curves(engine, base, n_valid, seed + 1)
    pred = sur.output(Pv)
    rms = float(np.sqrt(np.mean((pred - Yv) ** 2)))
    info = {"curves": int(n_curves), "rejected_draws": rejected, "gen_s": t_gen,
            "gen_curves_per_s": n_curves / t_gen if t_gen > 0 else None, "train_s": t_train,
            "steps": steps, "batch": batch, "epochs": steps * batch / n_curves,
            "final_loss": losses[-1] if losses else None, "valid_rms_logflux": rms,
            "gemm_tflops": st["gemm_flop"] / (st["train_ms"] * 1e-3) / 1e12 if st["train_ms"] > 0 else None}
    return sur, info


# ------------------------------------------------------------------ MCMC
class Posterior:
    """log posterior of the fit parameters: -chi**2 / 2 with the optimal
    scale factor, flat inside Model::is_not_legal, from the engine."""

    def __init__(self, engine, base, data, precision=0):
        from . import data_view
        self.engine, self.base, self.data, self.precision = engine, base, data, precision
        v = data_view(data)
        ok = v["weight"] > 0
        self.t = np.asarray(v["time"], np.float64)
        self.f = np.asarray(v["flux"], np.float64)
        self.w = np.where(ok, v["weight"] / np.where(ok, v["ferr"], 1.0) ** 2, 0.0)
        self.evaluations = 0

    def legal(self, P):
        P = np.asarray(P, np.float64)
        if P.shape[0] == 0:
            return np.zeros(0, bool)
        return ~self.base.is_not_legal_batch(P)

    def loglike(self, P):
        P = np.asarray(P, np.float64)
        if P.shape[0] == 0:
            return np.zeros(0)
        _, chisq, status = evaluate(self.engine, self.base, P, self.data, scale=True, rdata=True,
                                    precision=self.precision)
        self.evaluations += P.shape[0]
        return np.where(status == 0, -0.5 * chisq, -np.inf)

    def surrogate_loglike(self, sur, temperature=1.0):
        """-chi**2 / (2 T) of the surrogate curves, computed on the GPU."""
        sur.set_data((self.t, self.f, self.w))

        def f(P):
            if len(P) == 0:
                return np.zeros(0)
            return -0.5 * sur.chisq_gpu(P) / temperature
        return f


def run_chains(loglike, legal, x0, cov, n_iter, seed, sur_loglike=None, n_sub=1):
    """Random-walk Metropolis for W walkers at once (proposal N(0, cov)).

    Plain (sur_loglike None): every legal proposal is evaluated.  With a
    surrogate: each iteration first runs n_sub Metropolis steps on the
    surrogate posterior pi* from the current state, and the end point y is
    then judged by the engine with acceptance min(1, pi(y) pi*(x) /
    (pi(x) pi*(y))) (Christen & Fox 2005 for n_sub = 1; for n_sub > 1 the
    surrogate chain is a pi*-reversible proposal, Liu's surrogate
    transition, and the same ratio keeps pi invariant).  Walkers whose
    surrogate chain did not move skip the engine.  Returns the chain
    (n_iter, W, d) and counters."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.array(x0, np.float64)
    W, d = x.shape
    Lc = np.linalg.cholesky(cov)
    lx = loglike(x)
    ls = sur_loglike(x) if sur_loglike else None
    chain = np.empty((n_iter, W, d))
    n_eval, n_sub_acc, n_acc, t_sur = W, 0, 0, 0.0
    t0 = time.perf_counter()
    for it in range(n_iter):
        if sur_loglike is None:
            y = x + rng.standard_normal((W, d)) @ Lc.T
            cand = legal(y)
            lsy = None
        else:
            ts = time.perf_counter()
            y, lsy = x.copy(), ls.copy()
            for _ in range(n_sub):
                z = y + rng.standard_normal((W, d)) @ Lc.T
                ok = legal(z)
                lz = np.full(W, -np.inf)
                lz[ok] = sur_loglike(z[ok])
                with np.errstate(invalid="ignore"):
                    a = ok & (np.log(rng.random(W)) < lz - lsy)
                y[a], lsy[a] = z[a], lz[a]
                n_sub_acc += int(a.sum())
            cand = np.any(y != x, axis=1)
            t_sur += time.perf_counter() - ts
        u = rng.random(W)
        ly = np.full(W, -np.inf)
        if cand.any():
            ly[cand] = loglike(y[cand])
        n_eval += int(cand.sum())
        with np.errstate(invalid="ignore"):   # -inf - -inf: both invalid, rejected
            if sur_loglike is None:
                acc = cand & (np.log(u) < ly - lx)
            else:
                acc = cand & (np.log(u) < (ly - lx) - (lsy - ls))
                ls[acc] = lsy[acc]
        x[acc] = y[acc]
        lx[acc] = ly[acc]
        n_acc += int(acc.sum())
        chain[it] = x
    return chain, {"evaluations": n_eval, "surrogate_steps_accepted": n_sub_acc, "accepted": n_acc,
                   "proposals": n_iter * W, "n_sub": n_sub if sur_loglike else 0,
                   "surrogate_s": t_sur, "wall_s": time.perf_counter() - t0}


def run_chains_batched(loglike, legal, x0, cov, n_iter, seed, sur_loglike, k_try=4):
    """Delayed-acceptance Metropolis (Christen & Fox 2005, as run_chains with
    n_sub = 1) with the first-stage trials of a walker batched: until a
    walker's first stage accepts, every trial starts from the same state, so
    its next k_try trials y_1..y_k = x + N(0, cov) are drawn at once and
    screened in one surrogate call for all walkers; the first y_j that the
    first stage accepts goes to the engine (one call per round for all
    walkers), and y_1..y_{j-1} are the chain's first-stage rejections (the
    walker stays at x for those steps).  The chain is the one the step-by-
    step delayed-acceptance sampler draws from the same proposals; a round
    advances a walker by up to k_try steps for one engine evaluation.
    Returns the chain (n_iter, W, d) and counters (rounds = engine calls)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.array(x0, np.float64)
    W, d = x.shape
    Lc = np.linalg.cholesky(cov)
    lx = loglike(x)
    ls = sur_loglike(x)
    chain = np.empty((n_iter, W, d))
    step = np.zeros(W, int)
    n_eval, n_sur, n_sub_acc, n_acc, rounds, t_sur = W, W, 0, 0, 0, 0.0
    t0 = time.perf_counter()
    while True:
        act = np.nonzero(step < n_iter)[0]
        if len(act) == 0:
            break
        ts = time.perf_counter()
        z = x[act, None, :] + rng.standard_normal((len(act), k_try, d)) @ Lc.T
        u1 = np.log(rng.random((len(act), k_try)))
        u2 = np.log(rng.random(len(act)))
        flat = z.reshape(-1, d)
        ok = legal(flat)
        lz = np.full(len(flat), -np.inf)
        lz[ok] = sur_loglike(flat[ok])
        n_sur += int(ok.sum())
        lz = lz.reshape(len(act), k_try)
        t_sur += time.perf_counter() - ts
        room = np.minimum(k_try, n_iter - step[act])
        with np.errstate(invalid="ignore"):
            ok1 = (u1 < lz - ls[act, None]) & (np.arange(k_try)[None, :] < room[:, None])
        has = ok1.any(axis=1)
        j = np.where(has, np.argmax(ok1, axis=1), room)   # first-stage rejections before the first acceptance
        for t in range(k_try):                            # the walker stays at x for those steps
            rows = act[t < j]
            chain[step[rows] + t, rows] = x[rows]
        step[act] += j
        r = np.nonzero(has)[0]
        if len(r):
            pend = act[r]
            y = z[r, j[r]]
            lsy = lz[r, j[r]]
            n_sub_acc += len(r)
            ly = loglike(y)
            n_eval += len(r)
            rounds += 1
            with np.errstate(invalid="ignore"):
                acc = u2[r] < (ly - lx[pend]) - (lsy - ls[pend])
            a = pend[acc]
            x[a], lx[a], ls[a] = y[acc], ly[acc], lsy[acc]
            n_acc += int(acc.sum())
            chain[step[pend], pend] = x[pend]
            step[pend] += 1
    return chain, {"evaluations": n_eval, "surrogate_evaluations": n_sur, "surrogate_steps_accepted": n_sub_acc,
                   "accepted": n_acc, "proposals": n_iter * W, "n_sub": 1, "k_try": k_try, "rounds": rounds,
                   "surrogate_s": t_sur, "wall_s": time.perf_counter() - t0}


def calibrate_temperature(post, sur, x, cov, n, seed):
    """Stage-1 temperature from pilot proposals around the walkers x: the
    spread of (surrogate - engine) log-likelihood differences, at least 1.
    Fixed before sampling starts, so the chain stays exact."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Lc = np.linalg.cholesky(cov)
    i = rng.integers(0, x.shape[0], n)
    y = x[i] + rng.standard_normal((n, x.shape[1])) @ Lc.T
    ok = post.legal(y)
    y, xi = y[ok], x[i][ok]
    f = post.surrogate_loglike(sur, 1.0)
    d_sur = f(y) - f(xi)
    d_eng = post.loglike(y) - post.loglike(xi)
    good = np.isfinite(d_sur) & np.isfinite(d_eng)
    err = d_sur[good] - d_eng[good]
    robust = 1.4826 * float(np.median(np.abs(err - np.median(err))))   # MAD as a standard deviation
    return max(1.0, robust), {"pilot_pairs": int(good.sum()), "err_sd_robust": robust,
                              "corr": float(np.corrcoef(d_sur[good], d_eng[good])[0, 1])}


def tune_proposal(loglike, legal, x0, cov, rounds=4, iters=40, seed=0):
    """Adaptive pilot of plain Metropolis: after each round the proposal
    becomes 2.38^2 / d times the covariance of the walker positions of the
    round's second half.  Start from a small proposal and tight walkers:
    the rounds grow it to the posterior's scale.  Returns the tuned
    covariance, the walkers and the last round's second-half positions.
    (Pilot rounds are not part of any reported chain.)"""
    x = np.array(x0, np.float64)
    d = x.shape[1]
    samples = x
    for r in range(rounds):
        ch, _ = run_chains(loglike, legal, x, cov, iters, seed=seed + r)
        x = ch[-1]
        samples = ch[iters // 2:].reshape(-1, d)
        emp = np.cov(samples.T) + np.diag(1e-12 * np.abs(np.mean(samples, axis=0)) + 1e-30)
        cov = emp * (2.38 ** 2 / d)
    return cov, x, samples


def train_local_surrogate(engine, base, samples, cov, n, steps, seed, batch=2048, lr=1e-3, widths=WIDTHS,
                          log=None):
    """A network for the region a sampler explores: n engine curves at
    posterior samples (e.g. a pilot chain's positions) moved by one
    proposal step N(0, cov) -- where the sampler's first stage evaluates --
    legal draws only, t0 held at the samples' mean.  Inputs are
    (P - mean) / (3 std) of the samples; outputs the relative flux
    F / mean F at the 2048 bins.  Over so small a region the curve is
    nearly linear in the parameters: a least-squares linear map of the
    inputs becomes the network's FP32 skip, and the network is trained on
    what it leaves (standardised).  Returns the surrogate and the RMS error
    of its relative flux on 1024 fresh draws."""
    rng = np.random.Generator(np.random.PCG64(seed))
    samples = np.asarray(samples, np.float64)
    bins = make_data(np.arange(NBIN, dtype=np.float64) / NBIN)
    Lc = np.linalg.cholesky(np.asarray(cov))
    center = samples.mean(axis=0)
    spread = 3.0 * samples.std(axis=0)
    spread = np.where(spread > 0, spread, 1.0)

    def draw(cnt):
        Ps, Ys, got = [], [], 0
        while got < cnt:
            P = samples[rng.integers(0, samples.shape[0], 4096)] + rng.standard_normal((4096, samples.shape[1])) @ Lc.T
            P[:, 10] = center[10]
            P = P[~base.is_not_legal_batch(P)]
            if P.shape[0] == 0:
                continue
            Pz = P.copy()
            Pz[:, 10] = 0.0
            calc, _, status = evaluate(engine, base, Pz, bins)
            good = (status == 0) & np.all(calc > 0, axis=1)
            P, calc = P[good], calc[good]
            take = min(cnt - got, P.shape[0])
            Ps.append(P[:take])
            Ys.append((calc[:take] / calc[:take].mean(axis=1, keepdims=True)).astype(np.float32))
            got += take
        return np.concatenate(Ps), np.concatenate(Ys)

    t0 = time.perf_counter()
    P, Y = draw(n)
    t_gen = time.perf_counter() - t0
    mu = Y.mean(axis=0)
    mlp = MLP(widths, max_batch=max(batch, 4096))
    mlp.set_params(init_params(widths, seed))
    sur = Surrogate(mlp, mu, 1.0, center=center, spread=spread, log_output=False)
    X = sur.inputs(P)
    lin, *_ = np.linalg.lstsq(X.astype(np.float64), (Y - mu).astype(np.float64), rcond=None)
    R = (Y - mu) - (X.astype(np.float64) @ lin)
    sc = float(np.std(R))
    sur.scale = sc
    T = (R / sc).astype(np.float32)
    mlp.set_skip(lin / sc)
    lin_rms = sc
    mlp.set_dataset(P.shape[0])
    for r0 in range(0, P.shape[0], 65536):
        mlp.load_dataset(r0, X[r0:r0 + 65536], T[r0:r0 + 65536])
    t1 = time.perf_counter()
    for s0 in range(0, steps, 100):
        lr_s = 2e-5 + 0.5 * (lr - 2e-5) * (1.0 + math.cos(math.pi * s0 / max(steps, 1)))
        loss = mlp.train(min(100, steps - s0), min(batch, P.shape[0]), lr_s, seed + 7 + s0)
        if log:
            log(f"local surrogate: step {min(s0 + 100, steps)} loss {loss:.3e}")
    t_train = time.perf_counter() - t1
    Pv, Yv = draw(1024)
    rms = float(np.sqrt(np.mean((sur.output(Pv) - Yv) ** 2)))
    return sur, {"local_curves": int(n), "local_gen_s": t_gen, "local_train_s": t_train, "local_steps": steps,
                 "local_valid_rms_relflux": rms, "local_spread_relflux": float(np.std(Y - mu)),
                 "linear_part_rms_residual": lin_rms}


def autocorr_time(chain, c=5.0):
    """Integrated autocorrelation time per parameter of an ensemble chain
    (n, W, d): the walker-averaged normalised autocorrelation, summed up to
    Sokal's automatic window (the first M >= c tau(M))."""
    n, W, d = chain.shape
    taus = np.empty(d)
    nfft = 1 << int(math.ceil(math.log2(2 * n)))
    for j in range(d):
        x = chain[:, :, j] - chain[:, :, j].mean(axis=0)
        F = np.fft.rfft(x, n=nfft, axis=0)
        acf = np.fft.irfft(F * np.conj(F), axis=0)[:n].real
        acf = acf / np.where(acf[0] > 0, acf[0], 1.0)
        rho = acf.mean(axis=1)
        tau = 2.0 * np.cumsum(rho) - 1.0
        m = np.arange(n)
        win = np.argmax(m >= c * tau) if np.any(m >= c * tau) else n - 1
        taus[j] = max(tau[win], 1.0)
    return taus


def ess(chain, burn=0):
    """Effective sample size per parameter (walkers x steps / tau)."""
    ch = chain[burn:]
    return ch.shape[0] * ch.shape[1] / autocorr_time(ch)
