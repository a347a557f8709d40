"""Fit-loop batches: one Levenberg-Marquardt Jacobian plus MCMC walkers in
one batched evaluation.

The reference's LM solver differentiates the residuals
r_i = sqrt(w_i) (f_i - m_i) / sigma_i (levmarq_solver.cpp:670-700; a
non-positive sigma counts as 1e-10, a non-positive weight as 1) with
central differences of step h_j = fd_step_rel |p_j| + fd_step_min
(fd_step_rel = eps**(1/4), fd_step_min = 1e-10, at most half the span of
finite limits; levmarq_solver.cpp:427-441, 766-778).  A central difference
is used only when p_j +- h_j stay inside Model::get_limit(); otherwise a
one-sided difference (forward, or backward past the upper limit); a
failed evaluation halves the step, up to 4 times, and a central difference
that still fails falls back to the one-sided scheme
(levmarq_solver.cpp:800-880).  The reference
evaluates every column with its own light_curve_comp_fast call; here the
base point, all 2 x npar perturbations and the MCMC walkers go through one
light_curve_comp_batch call, and only columns whose evaluations failed are
retried (again batched) with halved steps.

Within a batch, perturbations that leave a star grid unchanged (t0, the
temperatures, limb darkening, gravity darkening) reuse the base point's
grids instead of rebuilding them (the engine's stage reuse).

`evaluate` is any callable (models, data, opts) -> (calc[m, n], status[m]);
engine_evaluator(engine) gives the GPU one.
"""
import math

import numpy as np

from . import LC_PREC_FP64, Model, data_view

FD_STEP_REL = math.sqrt(math.sqrt(np.finfo(np.float64).eps))
FD_STEP_MIN = 1e-10


def fd_steps(pars, limits, shrink=1.0):
    """h_j of levmarq_solver.cpp:766-778."""
    h = np.empty(len(pars))
    for j, p in enumerate(pars):
        hj = shrink * (FD_STEP_REL * abs(p) + FD_STEP_MIN)
        span = limits[j][1] - limits[j][0]
        if math.isfinite(span) and span > 0.0:
            hj = min(hj, 0.5 * span)
        h[j] = max(hj, FD_STEP_MIN)
    return h


def residuals(data, calc):
    """r_i = sqrt(w) (f - m) / sigma (levmarq_solver.cpp:690-700)."""
    v = data_view(data)
    sigma = np.where(v["ferr"] > 0.0, v["ferr"], 1e-10)
    w = np.where(v["weight"] > 0.0, v["weight"], 1.0)
    return np.sqrt(w) * (v["flux"] - calc) / sigma


class BaseEvalError(RuntimeError):
    """The base point of a Jacobian did not evaluate; status is its status
    (-1: the curve holds a non-finite value)."""

    def __init__(self, status):
        super().__init__(f"base evaluation failed with status {int(status)}")
        self.status = int(status)


def counted(status):
    """Evaluations that levmarq_solver.cpp:693 counts in fev_count: those
    that returned a curve, finite or not (status 0 or -1), not those that
    threw (an engine status > 0)."""
    return np.isin(np.asarray(status), (0, -1))


class FitBatch:
    """Jacobian + walker evaluation of one model against one data set.

    scale: fit the flux scale factor in every evaluation (re_scale), as the
    solvers do with autoscale on; max_model_points > 0 evaluates through
    light_curve_comp_fast.  precision defaults to FP64 per face: central
    differences of step ~1e-4 |p| (1e-10 for a parameter at 0, such as t0)
    resolve changes below the FP32 element resolution.
    """

    def __init__(self, model, data, evaluate, scale=True, max_model_points=0, precision=LC_PREC_FP64):
        self.model = model
        self.data = data
        self.evaluate = evaluate
        self.opts = dict(scale=scale, rdata=True, max_model_points=max_model_points, precision=precision)
        self.limits = [tuple(x) for x in model.get_limit()]
        self.npar = model.nvary()

    def _models(self, vectors):
        out = []
        for v in vectors:
            m = self.model.copy()
            m.set_param(v)
            out.append(m)
        return out

    def _run(self, vectors):
        """Evaluates the vectors; a curve with a non-finite value counts as a
        failed evaluation, as compute_residuals rejects it
        (levmarq_solver.cpp:697)."""
        calc, status = self.evaluate(self._models(vectors), self.data, self.opts)
        status = np.array(status, copy=True)
        if len(status):
            bad = (status == 0) & ~np.all(np.isfinite(calc), axis=1)
            status[bad] = -1
        return calc, status

    def plan(self, pars, h):
        """(kind, vectors) per column: kind 'c' central (+h, -h), 'f'/'b'
        one-sided forward/backward (one vector), None when no step fits."""
        cols = []
        for j in range(self.npar):
            lo, hi = self.limits[j]
            p = pars[j]
            if p + h[j] <= hi and p - h[j] >= lo:
                cols.append(("c", h[j]))
                continue
            hj, sign = h[j], 1.0
            if p + hj > hi:
                sign = -1.0
                if p - hj < lo:
                    hj, sign = 0.5 * (hi - lo), 1.0
                    if hj < 1e-15:
                        cols.append((None, 0.0))
                        continue
            cols.append(("f" if sign > 0 else "b", hj))
        return cols

    def _one_sided(self, pars, j, h):
        """The one-sided step of levmarq_solver.cpp:834-848 for column j:
        forward, backward past the upper limit, half the range forward when
        both sides leave the limits, (None, 0) when the range is empty."""
        lo, hi = self.limits[j]
        if pars[j] + h <= hi:
            return "f", h
        if pars[j] - h >= lo:
            return "b", h
        h = 0.5 * (hi - lo)
        return ("f", h) if h >= 1e-15 else (None, 0.0)

    def _vectors(self, pars, cols, which):
        vecs, index = [], []
        for j in which:
            kind, hj = cols[j]
            if kind == "c":
                for s in (1.0, -1.0):
                    v = np.array(pars, float)
                    v[j] = pars[j] + s * hj
                    vecs.append(v)
                index.append((j, len(vecs) - 2, len(vecs) - 1))
            elif kind in ("f", "b"):
                v = np.array(pars, float)
                v[j] = pars[j] + (hj if kind == "f" else -hj)
                vecs.append(v)
                index.append((j, len(vecs) - 1, None))
        return vecs, index

    def drive(self, gen):
        """Runs a request generator (yields lists of parameter vectors,
        receives (calc, status) for them) against this batch's evaluator;
        returns the generator's result."""
        try:
            req = next(gen)
            while True:
                req = gen.send(self._run(req))
        except StopIteration as e:
            return e.value

    def _resid(self, calc_row, vec, prior):
        r = residuals(self.data, calc_row)
        if prior is None:
            return r
        fn, eff = prior
        return np.concatenate([r, eff * fn(vec)])

    def jacobian(self, pars=None, walkers=(), shrink=1.0):
        """One batched call: base point, Jacobian columns, walkers.

        shrink scales the difference steps (the solver shrinks them when it
        stalls).  Returns dict with J (n x npar), resid0, chisq0, calc0,
        steps, walker_chisq (nan where a walker failed), walker_status, and
        the number of instances evaluated per call.
        """
        return self.drive(self.jacobian_requests(pars, walkers, shrink))

    def jacobian_requests(self, pars=None, walkers=(), shrink=1.0, prior=None):
        """jacobian() as a request generator (see drive).  prior: None or
        (fn, eff): the residual vector gains eff * fn(p) (prior residuals
        computed on the host, differenced with the same steps); eff may be
        a function of the base point's light-curve chi^2, evaluated once
        (its value is returned as prior_eff)."""
        pars = np.array(self.model.get_param() if pars is None else pars, float)
        h = fd_steps(pars, self.limits, shrink)
        cols = self.plan(pars, h)
        vecs, index = self._vectors(pars, cols, range(self.npar))
        walkers = [np.asarray(w, float) for w in walkers]
        batch = [pars] + vecs + walkers
        calc, status = yield batch
        status = np.asarray(status)
        n = calc.shape[1]
        if status[0] != 0:
            raise BaseEvalError(status[0])
        chisq_lc0 = float(np.sum(residuals(self.data, calc[0]) ** 2))
        if prior is not None and callable(prior[1]):
            prior = (prior[0], prior[1](chisq_lc0))
        r0 = self._resid(calc[0], pars, prior)
        J = np.zeros((len(r0), self.npar))
        calls = [len(batch)]
        # column evaluations that returned a curve (the reference's count)
        fev = int(np.sum(counted(status[1:1 + len(vecs)])))
        pending = []
        for (j, a, b) in index:
            kind, hj = cols[j]
            ok = status[1 + a] == 0 and (b is None or status[1 + b] == 0)
            if ok:
                J[:, j] = self._column(kind, hj, calc[1 + a], None if b is None else calc[1 + b], r0,
                                       vecs[a], None if b is None else vecs[b], prior)
            else:
                pending.append(j)
        # failed columns, retried in batches as levmarq_solver.cpp:800-866
        # does one by one: a central difference at h/2 .. h/16, then the
        # one-sided plan from the first step at h .. h/16, and a forward one
        # last tries backward at the smallest step; a column that never
        # evaluates stays zero
        first = {j: cols[j][1] for j in pending}
        tries = {j: 0 for j in pending}
        back = {j: False for j in pending}

        def advance(j):
            kind, hj = cols[j]
            if tries[j] < 4:
                tries[j] += 1
                cols[j] = (kind, 0.5 * hj)
                return True
            if kind == "c":
                okind, oh = self._one_sided(pars, j, first[j])
                if okind is None:
                    return False
                cols[j] = (okind, oh)
                tries[j] = 0
                return True
            if kind == "f" and not back[j]:
                back[j] = True
                cols[j] = ("b", hj)
                return True
            return False

        while pending:
            pending = [j for j in pending if advance(j)]
            if not pending:
                break
            vecs, index = self._vectors(pars, cols, pending)
            c2, s2 = yield vecs
            s2 = np.asarray(s2)
            calls.append(len(vecs))
            fev += int(np.sum(counted(s2)))
            nxt = []
            for (j, a, b) in index:
                kind, hj = cols[j]
                if s2[a] == 0 and (b is None or s2[b] == 0):
                    J[:, j] = self._column(kind, hj, c2[a], None if b is None else c2[b], r0,
                                           vecs[a], None if b is None else vecs[b], prior)
                else:
                    nxt.append(j)
            pending = nxt
        nw = len(walkers)
        wst = status[len(batch) - nw:] if nw else np.zeros(0, int)
        wchi = np.full(nw, np.nan)
        for k in range(nw):
            if wst[k] == 0:
                wchi[k] = float(np.sum(self._resid(calc[len(batch) - nw + k], walkers[k], prior) ** 2))
        steps = np.array([hj for _, hj in cols])
        return {"J": J, "resid0": r0, "chisq0": float(np.sum(r0 ** 2)), "chisq_lc0": chisq_lc0,
                "prior_eff": None if prior is None else prior[1], "calc0": calc[0], "steps": steps,
                "kinds": [k for k, _ in cols], "walker_chisq": wchi, "walker_status": wst,
                "instances_per_call": calls, "fev_columns": fev}

    def sum_squares(self, vectors, prior=None):
        """sum r_i^2 at each vector (nan where the evaluation failed)."""
        return self.drive(self.sum_squares_requests(vectors, prior))

    def sum_squares_requests(self, vectors, prior=None, with_status=False):
        if not len(vectors):
            return (np.zeros(0), np.zeros(0, int)) if with_status else np.zeros(0)
        vectors = [np.asarray(v, float) for v in vectors]
        calc, status = yield vectors
        ssq = np.array([float(np.sum(self._resid(calc[k], vectors[k], prior) ** 2)) if status[k] == 0 else np.nan
                        for k in range(len(vectors))])
        return (ssq, np.asarray(status)) if with_status else ssq

    def _column(self, kind, hj, ca, cb, r0, va, vb, prior):
        ra = self._resid(ca, va, prior)
        if kind == "c":
            d = (ra - self._resid(cb, vb, prior)) / (2.0 * hj)
        elif kind == "f":
            d = (ra - r0) / hj
        else:
            d = (r0 - ra) / hj
        return np.where(np.isfinite(d), d, 0.0)


def walkers_around(model, nwalk, seed, rel=1e-3):
    """nwalk legal parameter vectors p (1 + rel N(0,1)) (+ rel N(0,1) for
    parameters at 0), from a numpy PCG64 generator seeded with `seed`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = model.get_param()
    out = []
    while len(out) < nwalk:
        v = p * (1.0 + rel * rng.standard_normal(len(p))) + np.where(p == 0.0, rel * rng.standard_normal(len(p)), 0)
        if not model.is_not_legal(v):
            out.append(v)
    return out


SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def levmarq_requests(fb, pars=None, max_iter=200, ftol=SQRT_EPS, xtol=SQRT_EPS, prior=None, continuation=True,
                     stages=6, auto_balance=True, balance_target=1.0, max_fev=None, **kw):
    """Levenberg-Marquardt fit of fb.model's free parameters to fb.data
    (levmarq_solver.cpp:985-1530), as a request generator: it yields lists
    of parameter vectors to evaluate and receives (calc, status) for them
    (FitBatch.drive).  Without a prior it is one pass of the inner loop
    (lm_stage_requests).

    prior: None or prior.LMPrior, whose residuals extend the residual
    vector scaled by sqrt(s b).  With continuation the prior scale s steps
    through 0, 1/stages, ..., 1 (one inner-loop pass each, starting where
    the last ended, lambda, D and the difference steps fresh); a stage
    before the last runs at most max(20, max_iter // (stages + 1))
    iterations with ftol and xtol at least 1e-4.  b, the balance factor,
    is min(balance_target chi2_LC / (n_prior prior_weight), 1) at the
    stage's starting point with auto_balance, else 1 (0 while s is 0).
    MI355X-side change: chi2_LC comes from the base point of the stage's
    first Jacobian call instead of an evaluation of its own (the same
    engine result).  Iterations add up over the stages; calls counts
    batched engine calls.

    max_fev (default 200 (npar + 1), lm_max_fev of
    levmarq_solver.cpp:431): the fit ends with stop "max function
    evaluations" once the evaluations the one-at-a-time solver would have
    made since the start reach it, tested where it tests
    (levmarq_solver.cpp:1494: after an accepted, rejected or recovering
    iteration).  Counted as there (fev, returned): the start point, each
    stage's starting point and auto-balance point, each trial taken in
    order, each Jacobian column evaluation (retries included); evaluations
    that returned a curve, not those that failed.  The batched extras
    (trials past the accepted one, the Jacobian around a rejected first
    trial) are not counted.
    """
    p = np.array(fb.model.get_param() if pars is None else pars, float)
    fev = {"n": 1, "max": 200 * (fb.npar + 1) if max_fev is None else int(max_fev)}   # the start point
    if prior is None or len(prior) == 0:
        return (yield from lm_stage_requests(fb, p, max_iter=max_iter, ftol=ftol, xtol=xtol, fev=fev, **kw))
    sched = [s / stages for s in range(stages + 1)] if continuation else [1.0]
    total = {"iterations": 0, "calls": 0, "history": [], "stages": []}
    res = None
    for k, scale in enumerate(sched):
        final = k == len(sched) - 1

        def eff(chisq_lc, scale=scale):
            if auto_balance and scale > 0.0:
                raw = prior.weight_sum
                b = min(balance_target * chisq_lc / raw, 1.0) if raw > 0.0 and chisq_lc > 1e-30 else 1.0
            else:
                b = 1.0 if scale > 0.0 else 0.0
            return math.sqrt(scale * b)
        start = p
        if auto_balance and scale > 0.0:
            fev["n"] += 1   # the chi2_LC evaluation that sets the balance
        res = yield from lm_stage_requests(
            fb, p, prior=(prior, eff), max_iter=max_iter if final else max(20, max_iter // len(sched)),
            ftol=ftol if final else max(ftol, 1e-4), xtol=xtol if final else max(xtol, 1e-4), fev=fev, **kw)
        p = res["pars"]
        total["iterations"] += res["iterations"]
        total["calls"] += res["calls"]
        total["history"] += [(total["iterations"] - res["iterations"] + h[0],) + tuple(h[1:]) + (scale,)
                             for h in res["history"]]
        total["stages"].append({"scale": scale, "start": start, "eff": res["prior_eff"], "iterations": res["iterations"],
                                "stop": res["stop"], "sum_sq": res["chisq"]})
        if res["stop"] == "max function evaluations":
            break   # goto lm_done: no further stage
    return dict(res, iterations=total["iterations"], calls=total["calls"], history=total["history"],
                stages=total["stages"])


def lm_stage_requests(fb, p, prior=None, max_iter=200, ftol=SQRT_EPS, xtol=SQRT_EPS, gtol=0.0, tau0=1e-3,
                      lambda_conv=1.0, lambda_max=1e8, max_recoveries=3, fd_shrink_factor=0.1, candidates=4,
                      log=None, fev=None):
    """One pass of the LM inner loop (levmarq_solver.cpp:1225-1500) from p,
    with the prior residuals (fn, eff) when given (eff may be a function of
    the starting point's chi2_LC; see FitBatch.jacobian_requests).

    Damping as there: (J^T J + lambda D^2) delta = -J^T r with D_j the
    running maximum of the Jacobian column norms (update_scaling,
    levmarq_solver.cpp:935-950), the trial clamped to the parameter limits,
    gain ratio rho = (|r|^2 - |r_trial|^2) / (|J delta|^2 + 2 lambda
    |D delta|^2); rho > 0 accepts and scales lambda by
    max(1/3, 1 - (2 rho - 1)^3), otherwise lambda *= nu, nu *= 2; lambda is
    kept in [1e-20, lambda_max]; gtol, xtol (tested only while lambda <=
    lambda_conv) and ftol stop tests; a rejection at lambda_max shrinks the
    difference steps by fd_shrink_factor and rebuilds J (up to
    max_recoveries times), after which it ends the fit at the noise floor.

    Batching: while trials are rejected the damping sequence is fixed in
    advance, so one engine call evaluates `candidates` trial steps of that
    sequence together with the Jacobian around the first one; they are then
    taken in order exactly as the one-at-a-time loop would take them (the
    first accepted one ends the batch; if it is not the first, its Jacobian
    is one more call).  The iterates are those of the sequential loop
    (candidates=1).  Returns dict with pars, chisq (= |r|^2), iterations
    (trials, as the reference counts them), calls, stop reason, lambda,
    the iteration log and fev (the evaluation count; see
    levmarq_requests, whose dict {"n", "max"} fev is, shared over stages).
    """
    npar = fb.npar
    lims = fb.limits
    p = np.array(p, float)
    if fev is None:
        fev = {"n": 0, "max": 200 * (npar + 1)}
    shrink = 1.0
    jac = yield from fb.jacobian_requests(p, shrink=shrink, prior=prior)
    fev["n"] += 1 + jac["fev_columns"]   # the stage's starting point and its Jacobian
    calls = len(jac["instances_per_call"])
    if prior is not None:
        prior = (prior[0], jac["prior_eff"])
    D = None

    def scaling(J, D):
        cn = np.sqrt(np.sum(J * J, axis=0))
        if D is None:
            return np.where(cn > 0.0, cn, 1.0)
        return np.maximum(D, cn)

    J = jac["J"]
    D = scaling(J, D)
    r = jac["resid0"]
    sum_sq = float(np.sum(r * r))
    JtJ, Jtr = J.T @ J, J.T @ r
    lam, nu = tau0, 2.0
    it = 0
    fails = recoveries = 0
    stop = "max_iter"
    hist = []

    def solve(lam):
        A = JtJ + np.diag(lam * D * D)
        try:
            L = np.linalg.cholesky(A)
        except np.linalg.LinAlgError:
            return None
        return np.linalg.solve(L.T, np.linalg.solve(L, -Jtr))

    while it < max_iter:
        if gtol > 0.0:
            gmax = float(np.max(np.abs(Jtr) / np.maximum(1.0, D)))
            if gmax / max(1.0, math.sqrt(sum_sq)) < gtol:
                stop = "gtol"
                break
        # the trials of the damping sequence, in order, until `candidates`
        # of them are planned (Cholesky failures and the xtol test are steps
        # of the sequence too)
        plan, l, n_, k = [], lam, nu, it
        done = None
        while len(plan) < candidates and k < max_iter:
            d = solve(l)
            if d is None:
                plan.append(("chol", l, n_, None, None))
                l, n_ = l * n_, n_ * 2.0
                k += 1
                continue
            trial = np.array([min(max(p[j] + d[j], lims[j][0]), lims[j][1]) for j in range(npar)])
            d = trial - p
            step_norm = float(np.linalg.norm(d * D))
            if l <= lambda_conv:
                xnorm = float(np.linalg.norm(p * D))
                if xnorm > 0.0 and step_norm / xnorm < xtol:
                    plan.append(("xtol", l, n_, None, None))
                    break
            plan.append(("trial", l, n_, trial, d))
            k += 1
            if l * n_ >= lambda_max:  # a rejection here hits the ceiling: stall handling follows
                break
            l, n_ = l * n_, n_ * 2.0
        trials = [e[3] for e in plan if e[0] == "trial"]
        spec = None
        ssq = np.zeros(0)
        tst = np.zeros(0, int)
        if trials:
            try:
                spec = yield from fb.jacobian_requests(trials[0], walkers=trials[1:], shrink=shrink, prior=prior)
                calls += len(spec["instances_per_call"])
                ssq = np.concatenate([[spec["chisq0"]], spec["walker_chisq"]])
                tst = np.concatenate([[0], spec["walker_status"]])
            except BaseEvalError as e:  # the first trial failed to evaluate
                rest, rst = yield from fb.sum_squares_requests(trials[1:], prior, with_status=True)
                ssq = np.concatenate([[np.nan], rest])
                tst = np.concatenate([[e.status], rst])
                calls += 1 if len(trials) > 1 else 0
        ti = 0
        accepted = False
        for kind, l, n_, trial, d in plan:
            it += 1  # every pass of the loop counts, the one ending at xtol too
            if kind == "xtol":
                done = "xtol"
                break
            if kind == "chol":
                lam, nu = l * n_, n_ * 2.0
                fails += 1
                hist.append((it, sum_sq, lam, "cholesky"))
                if fails > 50:
                    done = "singular"
                    break
                continue
            s_t = ssq[ti]
            first = ti == 0
            fev["n"] += int(counted(tst[ti]))
            ti += 1
            if not np.isfinite(s_t):
                lam, nu = l * n_, n_ * 2.0
                fails += 1
                hist.append((it, sum_sq, lam, "eval_fail"))
                if fails > 50:
                    done = "max consecutive failures"
                    break
                continue  # as after a failed Cholesky: no clamp, no stall test
            Jd = J @ d
            pred = float(Jd @ Jd) + 2.0 * l * float(np.sum((D * d) ** 2))
            actual = sum_sq - s_t
            rho = actual / pred if pred > 0.0 else 0.0
            if rho > 0.0:
                p, sum_sq = trial, s_t
                lam = l * max(1.0 / 3.0, 1.0 - (2.0 * rho - 1.0) ** 3)
                nu = 2.0
                if first:
                    jac = spec
                else:
                    jac = yield from fb.jacobian_requests(p, shrink=shrink, prior=prior)
                    calls += len(jac["instances_per_call"])
                fev["n"] += jac["fev_columns"]
                J = jac["J"]
                r = jac["resid0"]
                D = scaling(J, D)
                JtJ, Jtr = J.T @ J, J.T @ r
                fails = 0
                status = "accept"
                if sum_sq > 0.0 and lam <= lambda_conv:
                    if abs(actual) / sum_sq < ftol and pred / sum_sq < ftol:
                        done = "ftol"
                        status = "ftol"
                lam = min(max(lam, 1e-20), lambda_max)
                hist.append((it, sum_sq, lam, status))
                accepted = True
                if not done and fev["n"] >= fev["max"]:
                    done = "max function evaluations"
                break
            lam, nu = l * n_, n_ * 2.0
            fails += 1
            lam = min(max(lam, 1e-20), lambda_max)
            hist.append((it, sum_sq, lam, "reject"))
            if lam >= lambda_max:
                if recoveries < max_recoveries:
                    recoveries += 1
                    shrink *= fd_shrink_factor
                    jac = yield from fb.jacobian_requests(p, shrink=shrink, prior=prior)
                    calls += len(jac["instances_per_call"])
                    fev["n"] += jac["fev_columns"]
                    J = jac["J"]
                    r = jac["resid0"]
                    D = scaling(J, None)
                    JtJ, Jtr = J.T @ J, J.T @ r
                    lam, nu = tau0, 2.0
                    fails = 0
                    hist.append((it, sum_sq, lam, "recover"))
                    if fev["n"] >= fev["max"]:
                        done = "max function evaluations"
                else:
                    done = "noise floor"
                break
            if fev["n"] >= fev["max"]:
                done = "max function evaluations"
                break
        if log:
            log(f"levmarq: iteration {it} sum_sq {sum_sq:.8g} lambda {lam:.3g} calls {calls}")
        if done:
            stop = done
            break
    return {"pars": p, "chisq": sum_sq, "chisq_lc": jac["chisq_lc0"],
            "iterations": it, "calls": calls, "stop": stop, "lambda": lam, "history": hist,
            "prior_eff": None if prior is None else prior[1], "fev": fev["n"]}


def levmarq(fb, pars=None, **kw):
    """Levenberg-Marquardt fit of fb.model's free parameters to fb.data; see
    levmarq_requests for the algorithm and the keyword arguments."""
    return fb.drive(levmarq_requests(fb, pars, **kw))


class CxxMt19937Uniform:
    """The draws of std::mt19937(seed) through libstdc++'s
    std::uniform_real_distribution<double>(0, 1): generate_canonical takes
    two 32-bit outputs w1, w2 and returns (w1 + w2 2^32) / 2^64 (the sum
    rounded to double; a result of 1 becomes the largest double below 1).
    numpy's legacy RandomState seeds MT19937 as std::mt19937 does and its
    full-range uint32 draws are the raw outputs."""

    def __init__(self, seed):
        self._rs = np.random.RandomState(int(seed) & 0xFFFFFFFF)

    def __call__(self):
        w1, w2 = (int(x) for x in self._rs.randint(0, 2 ** 32, size=2, dtype=np.uint64))
        r = (float(w1) + float(w2) * 4294967296.0) / 18446744073709551616.0
        return r if r < 1.0 else math.nextafter(1.0, 0.0)


def multistart_starts(model, pars=None, n=8, span=1.0, seed=42, constraints=None):
    """The multistart's starting vectors (levmarq_solver.cpp:1560-1606):
    the current parameters, then n draws; in each, every parameter but
    iangle uniform in [p - span range, p + span range] cut to its limits
    (kept when that is empty or the range is not positive), and iangle
    stratified over [max(lo, 5), min(hi, 90)]: start s at fraction
    (s + 0.2 + 0.6 u) / n.  Uniforms from std::mt19937(seed + 2) in the
    reference's order (CxxMt19937Uniform), so the starts are the
    reference's.  constraints (prior.Constraints) with K1, M1 and R1: q,
    velocity_scale and r1 of each drawn start come from
    prior.solve_consistent_params at its inclination, clamped to the
    limits, where it has a solution."""
    p0 = np.array(model.get_param() if pars is None else pars, float)
    npar = len(p0)
    lims = [tuple(x) for x in model.get_limit()]
    ranges = np.asarray(model.get_range(), float)
    names = [model.get_name(i) for i in range(npar)]
    ii = names.index("iangle") if "iangle" in names else -1
    u = CxxMt19937Uniform(seed + 2)
    starts = [p0.copy()]
    for s in range(n):
        st = p0.copy()
        for j in range(npar):
            if j == ii:
                continue
            rj = span * ranges[j]
            if not rj > 0.0:
                continue
            lo, hi = max(lims[j][0], p0[j] - rj), min(lims[j][1], p0[j] + rj)
            if hi > lo:
                st[j] = lo + (hi - lo) * u()
        if ii >= 0:
            ilo, ihi = max(lims[ii][0], 5.0), min(lims[ii][1], 90.0)
            st[ii] = ilo + (s + 0.2 + 0.6 * u()) / n * (ihi - ilo)
            if constraints is not None and constraints.K and constraints.M1 and constraints.R1:
                from .prior import solve_consistent_params
                sol = solve_consistent_params(st[ii], constraints.K[0], constraints.M1[0], constraints.R1[0],
                                              constraints.P_days)
                if sol is not None:
                    for name, val in zip(("q", "velocity_scale", "r1"), sol):
                        if name in names:
                            j = names.index(name)
                            st[j] = min(max(val, lims[j][0]), lims[j][1])
        starts.append(st)
    return starts


def run_lockstep(fb, gens):
    """Runs request generators (see FitBatch.drive) in lockstep: each
    round, the pending evaluations of every unfinished generator go to the
    evaluator in one batched call.  Returns their results in order and the
    number of batched calls."""
    out = [None] * len(gens)
    live, reqs = [], []
    for k, g in enumerate(gens):
        try:
            reqs.append(next(g))
            live.append((k, g))
        except StopIteration as e:
            out[k] = e.value
    calls = 0
    while live:
        batch = [v for r in reqs for v in r]
        calc, status = fb._run(batch)
        calls += 1
        status = np.asarray(status)
        nxt_g, nxt_r, o = [], [], 0
        for (k, g), r in zip(live, reqs):
            part = (calc[o:o + len(r)], status[o:o + len(r)])
            o += len(r)
            try:
                nxt_r.append(g.send(part))
                nxt_g.append((k, g))
            except StopIteration as e:
                out[k] = e.value
        live, reqs = nxt_g, nxt_r
    return out, calls


def levmarq_multistart(fb, starts, **kw):
    """Independent LM fits from several starting vectors (the reference's
    multistart, levmarq_solver.cpp:1550-1640; multistart_starts draws its
    starts),
    advanced in lockstep: each round, the pending evaluations of every
    unfinished fit go to the engine in one batched call.  Each fit is
    exactly the one levmarq() runs from its start.  Returns the results
    (levmarq's dicts, plus "start") sorted by chi^2, and the number of
    batched calls; a run whose evaluation failed has ok False (and chisq
    inf).  laplace_modes() groups them into modes."""
    def guarded(s0):
        # a start whose evaluation fails is a failed run, not the end of the
        # others (LMRun.ok, levmarq_solver.cpp:1052-1062)
        try:
            r = yield from levmarq_requests(fb, np.asarray(s0, float), **kw)
            return dict(r, ok=True)
        except RuntimeError as e:
            return {"ok": False, "error": str(e), "pars": np.asarray(s0, float), "chisq": math.inf}

    res, calls = run_lockstep(fb, [guarded(s0) for s0 in starts])
    out = [dict(r, start=k) for k, r in enumerate(res)]
    out.sort(key=lambda r: r["chisq"])
    return out, calls


def laplace_modes(fb, runs, dsteps, prior=None, log_prior=None, tol_dsteps=5.0, tol_frac=0.02):
    """Groups LM optima into modes and weights each by its Laplace
    posterior mass (levmarq_solver.cpp:1640-1790).

    runs: levmarq results (levmarq_multistart's; failed runs are left
    out).  Sorted by |r|^2, a run
    joins the first mode whose representative is within max(tol_dsteps
    |dstep_j|, tol_frac (|a_j| + |b_j|) / 2, 1e-12) in every parameter,
    else it opens a mode.  With s2 = max(chi2_LC(best) / max(1, N - k),
    1e-30), each mode gets lpost = (-0.5 chi2_LC + log_prior) / s2 and,
    from the central-difference Jacobian at its optimum with the data rows
    scaled by 1/sqrt(s2) (prior rows, when `prior` is given, at full scale
    times the best run's balance factor), Sigma = (J^T J)^-1 and log det
    Sigma; the weight is exp(lpost + 0.5 log det Sigma) normalised over
    the modes (the volume term only when every mode has its Sigma).
    log_prior: None or p -> log prior (the reference uses the physical
    prior with lm_error_mcmc_prior_weight, by default the balanced
    prior_weight).  MI355X-side change: the Jacobians of all modes are
    one batched engine call.  Returns the modes, heaviest first: dicts
    with pars, run (index into runs), n_runs, weight, lpost, logdet, cov
    (None when J^T J is not positive definite), chisq_lc; and calls.
    """
    order = sorted((k for k in range(len(runs)) if runs[k].get("ok", True)), key=lambda k: runs[k]["chisq"])
    if not order:
        raise ValueError("no LM run to group")
    dsteps = np.abs(np.asarray(dsteps, float))

    def same(a, b):
        thr = np.maximum(tol_dsteps * dsteps, tol_frac * 0.5 * (np.abs(a) + np.abs(b)))
        return bool(np.all(np.abs(a - b) <= np.maximum(thr, 1e-12)))

    modes = []
    for k in order:
        p = np.asarray(runs[k]["pars"], float)
        for m in modes:
            if same(p, m["pars"]):
                m["n_runs"] += 1
                break
        else:
            modes.append({"pars": p, "run": k, "n_runs": 1})
    best = runs[order[0]]
    ndata = len(data_view(fb.data)["flux"])
    s2 = max(best["chisq_lc"] / max(1, ndata - fb.npar), 1e-30)
    pr = None
    if prior is not None and len(prior):
        pr = (prior, best["stages"][-1]["eff"] if "stages" in best else 1.0)
    jacs, calls = run_lockstep(fb, [fb.jacobian_requests(m["pars"], prior=pr) for m in modes])
    dsc = 1.0 / math.sqrt(s2)
    for m, jac in zip(modes, jacs):
        m["chisq_lc"] = jac["chisq_lc0"]
        m["lpost"] = (-0.5 * jac["chisq_lc0"] + (log_prior(m["pars"]) if log_prior else 0.0)) / s2
        J = jac["J"].copy()
        J[:ndata] *= dsc
        A = J.T @ J
        m["cov"], m["logdet"] = None, 0.0
        try:
            L = np.linalg.cholesky(A)
            d = np.diag(L)
            if np.all(d * d > 1e-30):
                m["cov"] = np.linalg.inv(A)
                m["logdet"] = -2.0 * float(np.sum(np.log(d)))
        except np.linalg.LinAlgError:
            pass
    vol = all(m["cov"] is not None for m in modes)
    score = np.array([m["lpost"] + (0.5 * m["logdet"] if vol else 0.0) for m in modes])
    w = np.exp(np.maximum(score - score.max(), -700.0))
    w /= max(float(w.sum()), 1e-300)
    for m, wk in zip(modes, w):
        m["weight"] = float(wk)
    modes.sort(key=lambda m: -m["weight"])
    return modes, calls


def engine_evaluator(engine):
    """evaluate(models, data, opts) over Engine.light_curve_comp_batch."""
    def evaluate(models, data, opts):
        calc, _, status = engine.light_curve_comp_batch(models, data, scale=opts["scale"], rdata=opts["rdata"],
                                                        precision=opts["precision"],
                                                        max_model_points=opts["max_model_points"])
        return calc, status
    return evaluate


__all__ = ["FitBatch", "fd_steps", "residuals", "walkers_around", "engine_evaluator", "levmarq",
           "levmarq_multistart", "levmarq_requests", "Model"]
