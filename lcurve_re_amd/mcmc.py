"""Adaptive-Metropolis MCMC over the batched engine: many chains advanced
in lockstep, the proposals of all chains evaluated in one batched call per
step.

Each chain follows the reference's sampler (mcmc_solver.cpp:196-1175):
a Gaussian proposal with per-parameter steps
dstep * exp(s), or, once the covariance is ready, (2.38^2/d) C + eps I
through its Cholesky factor times exp(s); proposals reflected off the
limits (at most 20 bounces, then clamped); an evaluation failure is a
rejection; log alpha = -0.5 (chi2' - chi2) / (T s2) + (lp' - lp), lp the
optional physical log prior (prior.py; never tempered, a proposal it puts
at <= -1e29 rejected before evaluation), with the annealing
temperature T cooling linearly from anneal_T0 to 1 over the first
anneal_steps and s2 = chi2_best / (N - k) fixed at the end of burn-in
(data-error rescale); Robbins-Monro scale adaptation every adapt_interval
steps (gamma = rate (1 + batch)^-decay, log scale clamped); Welford
covariance of the chain's states, reset once at cov_reset_step, its
Cholesky factor refreshed every adapt_interval steps once cov_n passes the
warm-up (the scale adaptation restarts when the covariance proposal is
switched on).  chi^2 is the reference's: sum over weight > 0 of
weight ((flux - fit) / ferr)^2 (light_curve_comp.cpp:574-584).

Each chain draws from its own numpy PCG64 generator (seed + chain index),
in the reference's order (the proposal's normals, then a uniform only when
log alpha < 0), so a chain does not depend on the evaluator or on the
other chains.
"""
import math

import numpy as np

from . import LC_PREC_FP64, data_view


def chisq(data, calc):
    v = data_view(data)
    w = v["weight"]
    ok = w > 0.0
    return float(np.sum(w[ok] * ((v["flux"][ok] - calc[ok]) / v["ferr"][ok]) ** 2))


class _Chain:
    def __init__(self, p, dsteps, lims, seed, cfg):
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self.p = np.array(p, float)
        self.npar = len(p)
        self.lims = lims
        self.d0 = np.array(dsteps, float)
        self.dsteps = self.d0.copy()
        self.cfg = cfg
        self.log_scale = 0.0
        self.batch = self.win_acc = self.win_n = 0
        self.cov_n = 0
        self.mean = np.zeros(self.npar)
        self.M2 = np.zeros((self.npar, self.npar))
        self.L = None
        self.cov_ready = self.cov_reset = False
        self.chi = self.best = math.inf
        self.best_p = self.p.copy()
        self.chi2_scale = 1.0
        self.accepted = 0
        self.lp = 0.0  # log prior of the current state

    def propose(self):
        sf = math.exp(self.log_scale)
        z = self.rng.standard_normal(self.npar)
        if self.cov_ready:
            prop = self.p + sf * (self.L @ z)
        else:
            prop = self.p + self.dsteps * z
        for i, (lo, hi) in enumerate(self.lims):
            b = 0
            while (prop[i] < lo or prop[i] > hi) and b < 20:
                prop[i] = 2 * lo - prop[i] if prop[i] < lo else 2 * hi - prop[i]
                b += 1
            prop[i] = min(max(prop[i], lo), hi)
        return prop

    def record(self, accepted):
        c = self.cfg
        if c["adapt"]:
            self.win_n += 1
            self.win_acc += 1 if accepted else 0
            if self.win_n >= c["adapt_interval"]:
                rate = self.win_acc / self.win_n
                gamma = c["adapt_rate"] * (1.0 + self.batch) ** (-c["adapt_decay"])
                self.log_scale = min(max(self.log_scale + gamma * (rate - c["target"]), c["log_min"]),
                                     c["log_max"])
                self.dsteps = self.d0 * math.exp(self.log_scale)
                self.win_acc = self.win_n = 0
                self.batch += 1
        if c["adapt_covariance"]:
            self.cov_n += 1
            delta = self.p - self.mean
            self.mean = self.mean + delta / self.cov_n
            self.M2 = self.M2 + np.outer(delta, self.p - self.mean)

    def update_cholesky(self, step):
        c = self.cfg
        if not c["adapt_covariance"]:
            return
        threshold = c["cov_min_after_reset"] if self.cov_reset else c["cov_warmup"]
        if self.cov_n >= threshold and (not self.cov_ready or step % c["adapt_interval"] == 0):
            L = None
            if self.cov_n >= self.npar + 1:
                C = c["cov_sd"] ** 2 * self.M2 / (self.cov_n - 1) + c["cov_epsilon"] * np.eye(self.npar)
                try:
                    L = np.linalg.cholesky(C)
                except np.linalg.LinAlgError:
                    L = None
            if L is not None:
                self.L = L
                if not self.cov_ready:
                    self.cov_ready = True
                    self.log_scale = 0.0
                    self.batch = self.win_acc = self.win_n = 0
                    self.dsteps = self.d0.copy()
            elif self.cov_ready:
                self.cov_ready = False


def adaptive_metropolis(model, data, evaluate, nchains=8, nsteps=2000, burn_in=None, seed=1, start=None,
                        scale=True, max_model_points=0, precision=LC_PREC_FP64, anneal=True, anneal_T0=10.0,
                        anneal_steps=None, adapt=True, target=None, adapt_interval=100, adapt_rate=1.0,
                        adapt_decay=0.6, adapt_min_stepscale=1e-4, adapt_max_stepscale=1e4, adapt_covariance=True,
                        cov_warmup=None, cov_epsilon=1e-6, cov_reset_step=None, scale_data_errors=True, prior=None,
                        log=None):
    """Runs nchains chains of nsteps (the reference's defaults where it
    has them: burn_in = nsteps / 4, anneal_steps = burn_in / 2, target 0.234
    (0.44 for one parameter), cov_warmup = max(50 d, 1000), covariance
    reset at 2/5 of the burn-in).  evaluate(models, data, opts) ->
    (calc[m, n], status[m]) as fit.engine_evaluator gives.  start: one
    vector or one per chain (default: the model's parameters).  prior:
    None or p -> log prior (prior.model_prior): it enters log alpha
    untempered, and a proposal with log prior <= -1e29 is rejected without
    an evaluation (mcmc_solver.cpp:1033-1056, 1097-1099).

    Returns dict: chain [nchains, nsteps - burn_in, d] and chain_chisq of
    the recorded states, acceptance per chain, best chi^2 and parameters,
    final scale factors, chi2_scale, and the number of batched calls."""
    npar = model.nvary()
    burn_in = nsteps // 4 if burn_in is None else burn_in
    anneal_steps = max(0, burn_in // 2 if anneal_steps is None else anneal_steps)
    cfg = dict(adapt=adapt, adapt_interval=adapt_interval, adapt_rate=adapt_rate, adapt_decay=adapt_decay,
               target=(0.44 if npar == 1 else 0.234) if target is None else target,
               log_min=math.log(adapt_min_stepscale), log_max=math.log(adapt_max_stepscale),
               adapt_covariance=adapt_covariance,
               cov_warmup=max(50 * npar, 1000) if cov_warmup is None else cov_warmup,
               cov_min_after_reset=max(10 * npar, 200), cov_epsilon=cov_epsilon,
               cov_sd=2.38 / math.sqrt(npar))
    cov_reset_step = burn_in * 2 // 5 if cov_reset_step is None else cov_reset_step
    lims = [tuple(x) for x in model.get_limit()]
    p0 = model.get_param() if start is None else start
    starts = [np.asarray(p0, float)] * nchains if np.ndim(p0) == 1 else [np.asarray(s, float) for s in p0]
    chains = [_Chain(starts[c], model.get_dstep(), lims, seed + c, cfg) for c in range(nchains)]
    opts = dict(scale=scale, rdata=True, max_model_points=max_model_points, precision=precision)
    ndata = len(data)

    def run(vectors):
        ms = []
        for v in vectors:
            m = model.copy()
            m.set_param(v)
            ms.append(m)
        calc, status = evaluate(ms, data, opts)
        return [chisq(data, calc[k]) if status[k] == 0 else None for k in range(len(vectors))]

    calls = 1
    for ch, x in zip(chains, run([c.p for c in chains])):
        if x is None:
            raise RuntimeError("a chain's starting point does not evaluate")
        ch.chi = ch.best = x
        ch.lp = prior(ch.p) if prior else 0.0
    keep = nsteps - burn_in
    rec = np.zeros((nchains, max(keep, 0), npar))
    rec_chi = np.zeros((nchains, max(keep, 0)))
    rec_lp = np.zeros((nchains, max(keep, 0)))
    for step in range(nsteps):
        for ch in chains:
            if adapt_covariance and cov_reset_step > 0 and step == cov_reset_step and not ch.cov_reset:
                ch.cov_n = 0
                ch.mean[:] = 0.0
                ch.M2[:] = 0.0
                ch.cov_ready = False
                ch.cov_reset = True
            if step == burn_in and scale_data_errors:
                dof = ndata - npar
                if dof > 0 and ch.best > 0.0:
                    ch.chi2_scale = ch.best / dof
        props = [ch.propose() for ch in chains]
        lps = [prior(v) if prior else 0.0 for v in props]
        live = [k for k in range(nchains) if lps[k] > -1e29]
        res = [None] * nchains
        if live:
            for k, x in zip(live, run([props[k] for k in live])):
                res[k] = x
            calls += 1
        T = 1.0
        if anneal and step < anneal_steps and anneal_steps > 0:
            frac = step / anneal_steps
            T = anneal_T0 * (1.0 - frac) + 1.0 * frac
        for k, ch in enumerate(chains):
            acc = False
            x = res[k]
            if x is not None:
                la = -0.5 * (x - ch.chi) / (T * ch.chi2_scale) + (lps[k] - ch.lp)
                if la >= 0.0 or math.log(ch.rng.random()) < la:
                    ch.p, ch.chi, ch.lp = props[k], x, lps[k]
                    ch.accepted += 1
                    acc = True
                    if x < ch.best:
                        ch.best, ch.best_p = x, props[k].copy()
            ch.record(acc)
            if x is not None:
                ch.update_cholesky(step)
            if step >= burn_in:
                rec[k, step - burn_in] = ch.p
                rec_chi[k, step - burn_in] = ch.chi
                rec_lp[k, step - burn_in] = ch.lp
        if log and (step % 100 == 0 or step == nsteps - 1):
            log(f"mcmc: step {step} accepted {np.mean([c.accepted for c in chains]) / (step + 1):.3f} "
                f"best {min(c.best for c in chains):.6g}")
    return {"chain": rec, "chain_chisq": rec_chi, "chain_log_prior": rec_lp, "acceptance": np.array([c.accepted / nsteps for c in chains]),
            "best_chisq": np.array([c.best for c in chains]), "best_pars": np.array([c.best_p for c in chains]),
            "scale_factor": np.array([math.exp(c.log_scale) for c in chains]),
            "cov_ready": np.array([c.cov_ready for c in chains]),
            "chi2_scale": np.array([c.chi2_scale for c in chains]), "calls": calls}


__all__ = ["adaptive_metropolis", "chisq"]


def _reflect(prop, lims):
    for i, (lo, hi) in enumerate(lims):
        b = 0
        while (prop[i] < lo or prop[i] > hi) and b < 20:
            prop[i] = 2 * lo - prop[i] if prop[i] < lo else 2 * hi - prop[i]
            b += 1
        prop[i] = min(max(prop[i], lo), hi)
    return prop


def error_mcmc(fb, modes, err_s2=None, log_prior=None, dsteps=None, steps=8000, min_steps=500,
               mode_min_weight=0.005, seed=43, reanchor_rounds=2, reanchor_tol=0.5, cov=None):
    """The LM program's error sampling over the posterior modes
    (levmarq_solver.cpp:2090-2500), on a fit.FitBatch and the modes of
    fit.laplace_modes.

    Modes of weight >= mode_min_weight are sampled (the heaviest if none
    is), each by one chain of max(min_steps, steps w / w_sampled) kept
    steps after max(300, that / 5) burn-in steps, from the mode's optimum:
    proposal p + exp(s) (2.38 / sqrt(d)) L z with L the Cholesky factor of
    the mode's covariance (else of `cov`, the LM covariance of the fit,
    else dsteps times z), reflected off
    the limits (at most 20 bounces, then clamped); a proposal with log
    prior <= -1e29 is rejected unevaluated, an evaluation failure is a
    rejection; log alpha = (-0.5 (chi2' - chi2) + (lp' - lp)) / err_s2
    with chi2 the LM's light-curve chi^2 (a zero weight counting as 1) and
    err_s2 by default chi2_LC / max(1, N - k) of the heaviest mode;
    Robbins-Monro during burn-in: every 50 steps s += (1 + batch)^-0.6
    (rate - 0.234), s in [log 1e-3, log 1e3].  A mode whose start does
    not evaluate is skipped for the round (0 samples).  The best state met
    (largest -0.5 chi2 + lp) is tracked from the heaviest mode's optimum,
    the first anchor.  Re-anchoring: when a round raises the best
    -chi2 + 2 lp by more than reanchor_tol, that state is adopted as the
    anchor (a callable `cov` is evaluated again there) and the sampling
    restarts (samples cleared, the chain that found it starting from it),
    at most reanchor_rounds times; when the last allowed round still gains
    more, the anchor is kept and the fit is reported as not converged
    (levmarq_solver.cpp:2435-2458).

    MI355X-side change: the chains of all modes advance in lockstep, one
    batched engine call per step for every chain still running, and each
    chain draws from its own PCG64 generator (seed + mode index) instead of
    one generator shared by chains run one after another.  Returns dict
    with samples (kept states of all chains, mode order), per_mode (mode
    index, kept, accepted; kept 0 for a skipped mode), acceptance,
    map_pars / map_chisq / map_lp (the adopted anchor, the reported
    solution), best_pars / best_chisq / best_lp (the best state met,
    adopted or not), converged (the last round gained at most
    reanchor_tol), ok (more than 100 samples and acceptance above 1%),
    calls, rounds, reanchored and the gain of each round.  cov: None, a
    matrix, or p -> matrix (or None).
    """
    npar = fb.npar
    lims = fb.limits
    data = fb.data
    ndata = len(data_view(data)["flux"])
    if err_s2 is None:
        err_s2 = max(modes[0]["chisq_lc"] / max(1, ndata - npar), 1e-30)
    if dsteps is None:
        dsteps = np.abs(np.asarray(fb.model.get_dstep(), float))
    diag = np.where(np.asarray(dsteps, float) > 0.0, dsteps, 1e-6)
    pick = [k for k, m in enumerate(modes) if m["weight"] >= mode_min_weight] or [0]
    wincl = sum(modes[k]["weight"] for k in pick)
    prop_sd = 2.38 / math.sqrt(npar)
    lp = (lambda v: log_prior(v)) if log_prior else (lambda v: 0.0)

    def chol(c):
        if c is None:
            return None
        try:
            return np.linalg.cholesky(np.asarray(c, float))
        except np.linalg.LinAlgError:
            return None

    def global_chol(p):
        return chol(cov(p) if callable(cov) else cov)

    chains = []
    for k in pick:
        m = modes[k]
        nst = max(min_steps, int(steps * m["weight"] / max(wincl, 1e-300)))
        chains.append({"mode": k, "p": np.array(m["pars"], float), "L": chol(m.get("cov")), "nst": nst,
                       "nbr": max(300, nst // 5), "step": 0, "rng": np.random.default_rng(seed + k),
                       "s": 0.0, "batch": 0, "bacc": 0, "bn": 0, "kept": [], "acc": 0})
    # the first anchor is the heaviest mode's optimum (the LM fit); a start
    # that does not evaluate leaves the best state unset
    anchor = np.array(modes[0]["pars"], float)
    calc, status = fb._run([anchor])
    calls = 1
    if status[0] == 0:
        anc_chisq = float(np.sum(residuals_lc(data, calc[0]) ** 2))
        anc_lp = -0.5 * anc_chisq + lp(anchor)
    else:
        anc_chisq, anc_lp = math.inf, -1e300
    Lg = global_chol(anchor) if cov is not None else None
    map_pars, map_chisq, map_lp, map_k = anchor.copy(), anc_chisq, anc_lp, -1
    gains, reanchored, converged, rnd = [], False, False, 0
    while True:
        round_start_lp = map_lp
        for k, c in enumerate(chains):
            c.update(p=(anchor if reanchored and k == map_k else modes[c["mode"]]["pars"]).copy(), step=0, s=0.0,
                     batch=0, bacc=0, bn=0, kept=[], acc=0)
        calc, status = fb._run([c["p"] for c in chains])
        calls += 1
        for k, c in enumerate(chains):
            # a mode whose start does not evaluate is skipped this round
            c["skip"] = status[k] != 0
            if c["skip"]:
                continue
            c["chi"] = float(np.sum(residuals_lc(data, calc[k]) ** 2))
            c["lp"] = lp(c["p"])
        while True:
            live = [c for c in chains if not c["skip"] and c["step"] < c["nst"] + c["nbr"]]
            if not live:
                break
            props, lps = [], []
            for c in live:
                z = c["rng"].standard_normal(npar)
                sf = math.exp(c["s"])
                Lc = c["L"] if c["L"] is not None else Lg
                prop = c["p"] + sf * prop_sd * (Lc @ z) if Lc is not None else c["p"] + sf * prop_sd * diag * z
                props.append(_reflect(prop, lims))
                lps.append(lp(props[-1]))
            ev = [i for i, v in enumerate(lps) if v > -1e29]
            res = [None] * len(live)
            if ev:
                calc, status = fb._run([props[i] for i in ev])
                calls += 1
                for j, i in enumerate(ev):
                    if status[j] == 0:
                        res[i] = float(np.sum(residuals_lc(data, calc[j]) ** 2))
            for i, c in enumerate(live):
                acc = False
                if res[i] is not None:
                    la = (-0.5 * (res[i] - c["chi"]) + (lps[i] - c["lp"])) / err_s2
                    if la >= 0.0 or math.log(c["rng"].random()) < la:
                        c["p"], c["chi"], c["lp"] = props[i], res[i], lps[i]
                        acc = True
                if acc and -0.5 * c["chi"] + c["lp"] > map_lp:
                    map_lp, map_pars, map_chisq = -0.5 * c["chi"] + c["lp"], c["p"].copy(), c["chi"]
                    map_k = chains.index(c)
                if c["step"] < c["nbr"]:
                    c["bn"] += 1
                    c["bacc"] += acc
                    if c["bn"] >= 50:
                        g = (1.0 + c["batch"]) ** -0.6
                        c["s"] = min(max(c["s"] + g * (c["bacc"] / c["bn"] - 0.234), math.log(1e-3)), math.log(1e3))
                        c["bacc"] = c["bn"] = 0
                        c["batch"] += 1
                else:
                    c["acc"] += acc
                    c["kept"].append(c["p"].copy())
                c["step"] += 1
        gains.append(2.0 * (map_lp - round_start_lp))
        if gains[-1] <= reanchor_tol:
            converged = True  # the round started at the optimum
            break
        if rnd >= reanchor_rounds:
            break  # still descending: the anchor is kept, not converged
        # adopt the best state (adopt_map): the anchor, its chain's start
        # and the global proposal covariance move there
        reanchored = True
        anchor, anc_chisq, anc_lp = map_pars.copy(), map_chisq, map_lp
        if callable(cov):
            Lg = global_chol(anchor)
        rnd += 1
    samples = np.array([v for c in chains for v in c["kept"]]).reshape(-1, npar)
    drawn = sum(len(c["kept"]) for c in chains)
    acc = sum(c["acc"] for c in chains) / drawn if drawn else 0.0
    return {"samples": samples, "per_mode": [(c["mode"], len(c["kept"]), c["acc"]) for c in chains],
            "acceptance": acc, "map_pars": anchor, "map_chisq": anc_chisq, "map_lp": anc_lp,
            "best_pars": map_pars, "best_chisq": map_chisq, "best_lp": map_lp, "converged": converged,
            "ok": len(samples) > 100 and acc > 0.01, "calls": calls, "err_s2": err_s2, "rounds": rnd + 1,
            "reanchored": reanchored, "gains": gains}


def residuals_lc(data, calc):
    """The LM's light-curve residuals (levmarq_solver.cpp:690-700)."""
    from .fit import residuals
    return residuals(data, calc)
