"""The reference's physical prior for the samplers (physical_prior.h:15-242,
PhysicalPrior::compute): split-Gaussian constraints on quantities implied by
(inclination, q, velocity_scale, r1, r2, t1, t2) and the period -- K1, K2,
R1 and R2 in solar radii, masses from Kepler's third law, q, log g of both
stars, temperatures -- summed and scaled by prior_weight; -1e30 for
sin i < 1e-10, q <= 0, velocity_scale <= 0 or r1 outside (0, 1)."""
import math
from dataclasses import dataclass

import numpy as np

DEG2RAD = math.pi / 180.0
DAY2SEC = 86400.0
RSUN_KM = 695700.0
G_MSUN = 1.3271244e11  # km^3 s^-2 (IAU 2015 nominal)
LOGG_SUN = 4.4380

QUANTITIES = ("K", "K2", "M1", "M2", "M2min", "Mtotal", "q", "R1", "R2", "logg1", "logg2", "T1", "T2")


@dataclass
class Constraints:
    """ObservedConstraints: each quantity is None or (value, err_lo, err_hi)."""
    K: tuple = None
    K2: tuple = None
    M1: tuple = None
    M2: tuple = None
    M2min: tuple = None
    Mtotal: tuple = None
    q: tuple = None
    R1: tuple = None
    R2: tuple = None
    logg1: tuple = None
    logg2: tuple = None
    T1: tuple = None
    T2: tuple = None
    P_days: float = 1.0
    prior_weight: float = 1.0


def log_gauss(x, mu, sigma):
    if sigma <= 0.0:
        return 0.0
    z = (x - mu) / sigma
    return -0.5 * z * z - 0.5 * math.log(2.0 * math.pi) - math.log(sigma)


def log_split_gauss(x, mu, sig_lo, sig_hi):
    if sig_lo <= 0.0 and sig_hi <= 0.0:
        return 0.0
    sigma = sig_lo if x < mu else sig_hi
    if sigma <= 0.0:
        sigma = max(sig_lo, sig_hi)
    return log_gauss(x, mu, sigma)


def log_prior(i_deg, q, vs, r1, r2, t1, t2, c):
    """PhysicalPrior::compute (physical_prior.h:147-242)."""
    sin_i = math.sin(i_deg * DEG2RAD)
    if sin_i < 1e-10 or q <= 0.0 or vs <= 0.0 or r1 <= 0.0 or r1 >= 1.0:
        return -1e30
    P_s = c.P_days * DAY2SEC
    a_km = vs * P_s / (2.0 * math.pi)
    lp = 0.0

    def add(x, con):
        return log_split_gauss(x, con[0], con[1], con[2])

    if c.K:
        lp += add(vs * sin_i * q / (1.0 + q), c.K)
    if c.K2:
        lp += add(vs * sin_i / (1.0 + q), c.K2)
    R1 = r1 * a_km / RSUN_KM
    if c.R1:
        lp += add(R1, c.R1)
    R2 = r2 * a_km / RSUN_KM if 0.0 < r2 < 1.0 else 0.0
    if c.R2 and R2 > 0.0:
        lp += add(R2, c.R2)
    M_total = 4.0 * math.pi * math.pi * a_km * a_km * a_km / (G_MSUN * P_s * P_s)
    M1 = M_total / (1.0 + q)
    M2 = q * M1
    if c.M1:
        lp += add(M1, c.M1)
    if c.M2:
        lp += add(M2, c.M2)
    if c.M2min and M2 < c.M2min[0]:
        sig = c.M2min[1] if c.M2min[1] > 0.0 else c.M2min[2]
        if sig <= 0.0:
            sig = 0.01 * c.M2min[0]
        z = (M2 - c.M2min[0]) / sig
        lp += -0.5 * z * z
    if c.Mtotal:
        lp += add(M_total, c.Mtotal)
    if c.q:
        lp += add(q, c.q)
    if c.logg1 and M1 > 0.0 and R1 > 0.0:
        lp += add(LOGG_SUN + math.log10(M1) - 2.0 * math.log10(R1), c.logg1)
    if c.logg2 and M2 > 0.0 and R2 > 0.0:
        lp += add(LOGG_SUN + math.log10(M2) - 2.0 * math.log10(R2), c.logg2)
    if c.T1 and t1 > 0.0:
        lp += add(t1, c.T1)
    if c.T2 and t2 > 0.0:
        lp += add(t2, c.T2)
    return 0.0 + c.prior_weight * lp


# keys of the configuration's "priors" object (mcmc_solver.cpp:380-436)
PRIOR_KEYS = {"K1": "K", "vrad1_obs": "K", "K2": "K2", "M1": "M1", "m1": "M1", "M2": "M2", "M2_min": "M2min",
              "m2_min": "M2min", "M_total": "Mtotal", "q": "q", "R1": "R1", "r1": "R1", "R2": "R2",
              "logg1": "logg1", "logg2": "logg2", "T1": "T1", "T2": "T2"}


def parse_three(text):
    """Helpers::parseThreeDoubles (new_helpers.h:455-467): exactly three
    numbers separated by white space."""
    parts = str(text).split()
    if len(parts) != 3:
        raise ValueError("prior %r: expected three numbers (value, err_lo, err_hi)" % (text,))
    return tuple(float(x) for x in parts)


def constraints_from_config(config, sampler=True):
    """ObservedConstraints from a solver configuration, or None when
    use_priors is off (mcmc_solver.cpp:355-440): period from true_period,
    weight from prior_weight -- for the sampler mcmc_prior_weight wins when
    present -- and one constraint per entry of "priors", a missing error
    taking the other side's.  An unknown key raises ValueError (the
    reference stops with "Unknown prior")."""
    if not config.get("use_priors", False):
        return None
    c = Constraints(P_days=float(config.get("true_period", 1.0)),
                    prior_weight=float(config.get("prior_weight", 1.0)))
    if sampler:
        c.prior_weight = float(config.get("mcmc_prior_weight", c.prior_weight))
    for key, text in config.get("priors", {}).items():
        if key not in PRIOR_KEYS:
            raise ValueError("Unknown prior: %s" % key)
        val, lo, hi = parse_three(text)
        if hi <= 0.0:
            hi = lo
        if lo <= 0.0:
            lo = hi
        setattr(c, PRIOR_KEYS[key], (val, lo, hi))
    return c


def implied(i_deg, q, vs, r1, r2, P_days):
    """The quantities compute() constrains, by the same formulas: K1, K2
    (km/s), a (km), R1, R2 (solar radii; R2 0 unless 0 < r2 < 1), M1, M2,
    M_total (solar masses), logg1, logg2 (cgs)."""
    sin_i = math.sin(i_deg * DEG2RAD)
    P_s = P_days * DAY2SEC
    a_km = vs * P_s / (2.0 * math.pi)
    R1 = r1 * a_km / RSUN_KM
    R2 = r2 * a_km / RSUN_KM if 0.0 < r2 < 1.0 else 0.0
    M_total = 4.0 * math.pi * math.pi * a_km * a_km * a_km / (G_MSUN * P_s * P_s)
    M1 = M_total / (1.0 + q)
    M2 = q * M1
    lg = lambda M, R: LOGG_SUN + math.log10(M) - 2.0 * math.log10(R) if M > 0.0 and R > 0.0 else float("nan")
    return {"K1": vs * sin_i * q / (1.0 + q), "K2": vs * sin_i / (1.0 + q), "a_km": a_km, "R1": R1, "R2": R2,
            "M1": M1, "M2": M2, "M_total": M_total, "logg1": lg(M1, R1), "logg2": lg(M2, R2)}


def model_prior(model, constraints):
    """p -> log prior for the model's free-parameter vectors p: varied
    parameters from p, the others from the model (get_par,
    mcmc_solver.cpp:323-328 / levmarq_solver.cpp:323-328)."""
    get = _param_getter(model)

    def f(p):
        v = get(p)
        return log_prior(v["iangle"], v["q"], v["velocity_scale"], v["r1"], v["r2"], v["t1"], v["t2"], constraints)
    return f


__all__ = ["Constraints", "constraints_from_config", "implied", "log_prior", "model_prior", "log_gauss",
           "log_split_gauss", "parse_three", "LMPrior", "lm_derived", "lm_residual", "solve_consistent_params",
           "consistent_start"]


# ------------------------------------------------------------------
# The LM solver's form of the prior: one residual per constraint
# (levmarq_solver.cpp:195-320), with its own constants and definitions.
LM_G, LM_MSUN, LM_RSUN = 6.67430e-11, 1.98892e30, 6.9634e8


def lm_derived(key, i_deg, q, vs, r1, r2, t1, t2, P_days):
    """evaluate_derived (levmarq_solver.cpp:225-290): SI constants, M2_min
    as M2 sin^3 i, log g = log10(100 g) of g in m/s^2 (0 when g <= 0), R2
    from r2 whatever its value."""
    sin_i = math.sin(i_deg * (math.pi / 180.0))
    P_s = P_days * 86400.0
    a_m = vs * 1e3 * P_s / (2.0 * math.pi)
    a_rsun = a_m / LM_RSUN
    M_total = 4.0 * math.pi * math.pi * a_m * a_m * a_m / (LM_G * P_s * P_s * LM_MSUN)
    M1 = M_total / (1.0 + q)
    M2 = M_total * q / (1.0 + q)
    R1 = r1 * a_rsun
    R2 = r2 * a_rsun
    if key == "K":
        return vs * sin_i * q / (1.0 + q)
    if key == "K2":
        return vs * sin_i / (1.0 + q)
    if key == "M1":
        return M1
    if key == "M2":
        return M2
    if key == "M2min":
        return M2 * sin_i * sin_i * sin_i
    if key == "Mtotal":
        return M_total
    if key == "q":
        return q
    if key == "R1":
        return R1
    if key == "R2":
        return R2
    if key in ("logg1", "logg2"):
        M, R = (M1, R1) if key == "logg1" else (M2, R2)
        g = LM_G * M * LM_MSUN / ((R * LM_RSUN) * (R * LM_RSUN))
        return math.log10(g * 100.0) if g > 0 else 0.0
    if key == "T1":
        return t1
    if key == "T2":
        return t2
    raise KeyError(key)


def lm_residual(key, obs, weight, i_deg, q, vs, r1, r2, t1, t2, P_days):
    """prior_residual (levmarq_solver.cpp:292-318): sqrt(weight) (X - X_obs)
    / sigma, sigma err_hi when X >= X_obs else err_lo (the other when that
    one is 0; 1% of |X_obs| when both are)."""
    val, lo, hi = obs
    diff = lm_derived(key, i_deg, q, vs, r1, r2, t1, t2, P_days) - val
    sigma = hi if diff >= 0.0 else lo
    if sigma <= 0.0:
        sigma = max(lo, hi)
    if sigma <= 0.0:
        sigma = 0.01 * abs(val) + 1e-30
    return math.sqrt(weight) * diff / sigma


class LMPrior:
    """The LM's prior residuals for a model's free-parameter vectors, in
    the order of the configuration's "priors" object; weight is the plain
    prior_weight (levmarq_solver.cpp:457-550).  Calling it with p gives
    the unscaled residual vector; weight_sum is what the balance factor
    divides by (levmarq_solver.cpp:1100-1104)."""

    def __init__(self, model, config):
        c = constraints_from_config(config, sampler=False)
        self.specs = []
        if c is not None:
            for key in config.get("priors", {}):
                name = PRIOR_KEYS[key]
                self.specs.append((name, getattr(c, name)))
        self.P_days = c.P_days if c is not None else 1.0
        self.weight = c.prior_weight if c is not None else 1.0
        self.weight_sum = self.weight * len(self.specs)
        self._get = _param_getter(model)

    def __len__(self):
        return len(self.specs)

    def __call__(self, p):
        v = self._get(p)
        return np.array([lm_residual(name, obs, self.weight, v["iangle"], v["q"], v["velocity_scale"], v["r1"],
                                     v["r2"], v["t1"], v["t2"], self.P_days) for name, obs in self.specs])


def _param_getter(model):
    names = [model.get_name(i) for i in range(model.nvary())]
    fixed = {k: getattr(model.pod, k).value for k in ("iangle", "q", "velocity_scale", "r1", "r2", "t1", "t2")}
    idx = {n: j for j, n in enumerate(names) if n in fixed}

    def get(p):
        v = dict(fixed)
        for n, j in idx.items():
            v[n] = float(p[j])
        return v
    return get


def solve_consistent_params(i_deg, K, M1, R1, P_days):
    """PhysicalPrior::solve_consistent_params (physical_prior.h:283-318):
    the q, velocity_scale and r1 for which an orbit at inclination i_deg
    has K1 = K, M1 and R1 (solar units) -- q the root of (1 + q)^2 / q^3 =
    2 pi G M1 sin^3 i / (K^3 P) by geometric bisection on [1e-4, 1e4] (200
    steps or a ratio below 1 + 1e-12), vs = K (1 + q) / (q sin i), r1 = R1
    / a.  None when sin i < 0.01 or r1 falls outside (0, 1)."""
    sin_i = math.sin(i_deg * DEG2RAD)
    if sin_i < 0.01:
        return None
    P_s = P_days * DAY2SEC
    rhs = 2.0 * math.pi * G_MSUN * M1 * sin_i ** 3 / (K * K * K * P_s)
    lo, hi = 1e-4, 1e4
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if (1.0 + mid) * (1.0 + mid) / (mid * mid * mid) > rhs:
            lo = mid
        else:
            hi = mid
        if hi / lo < 1.0 + 1e-12:
            break
    q = math.sqrt(lo * hi)
    vs = K * (1.0 + q) / (q * sin_i)
    r1 = R1 * RSUN_KM / (vs * P_s / (2.0 * math.pi))
    if r1 <= 0.0 or r1 >= 1.0:
        return None
    return q, vs, r1


def consistent_start(model, pars, c):
    """The LM's auto_consistent_init (levmarq_solver.cpp:580-640): with K1,
    M1 and R1 constraints, when the starting K1 is more than 3 sigma from
    the observed one, q, velocity_scale and r1 are replaced by
    solve_consistent_params at the starting inclination (else at the
    first of 85, 80, .., 30 degrees that has a solution, which then also
    becomes the inclination), clamped to the limits.  Returns the new
    vector (the old one when nothing applies)."""
    p = np.array(pars, float)
    if c is None or not (c.K and c.M1 and c.R1):
        return p
    names = [model.get_name(i) for i in range(model.nvary())]
    lims = model.get_limit()
    v = _param_getter(model)(p)
    K_impl = v["velocity_scale"] * math.sin(v["iangle"] * DEG2RAD) * v["q"] / (1.0 + v["q"])
    sig = c.K[1] if K_impl < c.K[0] else c.K[2]
    if sig <= 0.0:
        sig = max(c.K[1], c.K[2])
    if sig <= 0.0:
        sig = 0.1 * c.K[0]
    if abs((K_impl - c.K[0]) / sig) <= 3.0:
        return p
    i_deg = v["iangle"]
    sol = solve_consistent_params(i_deg, c.K[0], c.M1[0], c.R1[0], c.P_days)
    if sol is None:
        for i_try in np.arange(85.0, 29.0, -5.0):
            sol = solve_consistent_params(float(i_try), c.K[0], c.M1[0], c.R1[0], c.P_days)
            if sol is not None:
                i_deg = float(i_try)
                break
    if sol is None:
        return p
    for name, val in zip(("iangle", "q", "velocity_scale", "r1"), (i_deg,) + sol):
        if name in names:
            p[names.index(name)] = val
    return np.array([min(max(x, lims[j][0]), lims[j][1]) for j, x in enumerate(p)])
