"""The physical prior (lcurve_re_amd/prior.py, a restatement of the
reference's physical_prior.h:147-242 and of the configuration reading in
mcmc_solver.cpp:355-440).

Pinned by physics and closed forms, not by retyping the formulas: an orbit
of one solar mass at 1 au has Kepler's period (365.2569 d for the IAU 2015
GM_sun) and orbital speed; K1 + K2 = vs sin i and K1 / K2 = q; M2 / M1 = q;
a star of one solar mass and radius has log g = 4.438; a Gaussian
constraint equals scipy's normal log density, the split one takes the
matching side's sigma; constraints add, the weight scales them, and the
one-sided M2_min penalty is zero above the bound.  CPU only."""
import math

import numpy as np
import pytest
from scipy import stats

from lcurve_re_amd import prior as PR

AU_KM = 1.495978707e8
GM_SUN = 1.3271244e11


def _kepler_orbit():
    """vs (relative orbital speed) and P of a circular orbit of M_total = 1 at 1 au."""
    P_s = 2.0 * math.pi * math.sqrt(AU_KM ** 3 / GM_SUN)
    return 2.0 * math.pi * AU_KM / P_s, P_s / 86400.0


def test_kepler_orbit_of_the_sun():
    vs, P = _kepler_orbit()
    assert abs(P - 365.2569) < 2e-3 and abs(vs - 29.7847) < 1e-3
    d = PR.implied(90.0, 1e-6, vs, 0.1, 0.01, P)
    assert d["M_total"] == pytest.approx(1.0, rel=1e-12)
    assert d["a_km"] == pytest.approx(AU_KM, rel=1e-12)
    assert d["R1"] == pytest.approx(0.1 * AU_KM / 695700.0, rel=1e-12)


@pytest.mark.parametrize("i,q,vs", [(90.0, 0.3, 400.0), (63.0, 0.8, 150.0), (20.0, 2.0, 55.0)])
def test_velocities_and_masses_split_by_q(i, q, vs):
    d = PR.implied(i, q, vs, 0.2, 0.3, 0.7)
    assert d["K1"] + d["K2"] == pytest.approx(vs * math.sin(math.radians(i)), rel=1e-14)
    assert d["K1"] / d["K2"] == pytest.approx(q, rel=1e-14)
    assert d["M1"] + d["M2"] == pytest.approx(d["M_total"], rel=1e-14)
    assert d["M2"] / d["M1"] == pytest.approx(q, rel=1e-14)
    assert d["R2"] / d["R1"] == pytest.approx(1.5, rel=1e-14)


def test_solar_log_g():
    # choose vs and P so that M1 = 1 and R1 = 1 solar unit: a = 1 / r1 R_sun
    r1, q = 0.25, 0.5
    a = 695700.0 / r1
    P_s = 2.0 * math.pi * math.sqrt(a ** 3 / (GM_SUN * (1.0 + q)))
    vs = 2.0 * math.pi * a / P_s
    d = PR.implied(90.0, q, vs, r1, 0.2, P_s / 86400.0)
    assert d["M1"] == pytest.approx(1.0, rel=1e-12) and d["R1"] == pytest.approx(1.0, rel=1e-12)
    assert d["logg1"] == pytest.approx(4.438, abs=1e-12)
    # 10^(4.438) cgs = G M_sun / R_sun^2 with G = 6.67430e-8, M_sun = 1.98892e33 g, R_sun = 6.957e10 cm
    assert abs(10 ** 4.438 / (6.67430e-8 * 1.98892e33 / 6.957e10 ** 2) - 1) < 1e-3


def test_gaussian_terms_are_normal_log_densities():
    for x in (-1.0, 0.3, 2.5):
        assert PR.log_gauss(x, 0.5, 0.7) == pytest.approx(stats.norm.logpdf(x, 0.5, 0.7), rel=1e-13)
        lo = PR.log_split_gauss(x, 0.5, 0.2, 0.9)
        assert lo == pytest.approx(stats.norm.logpdf(x, 0.5, 0.2 if x < 0.5 else 0.9), rel=1e-13)
    assert PR.log_gauss(1.0, 0.0, 0.0) == 0.0
    assert PR.log_split_gauss(1.0, 0.0, 0.0, 0.0) == 0.0
    assert PR.log_split_gauss(-1.0, 0.0, 0.0, 0.4) == pytest.approx(stats.norm.logpdf(-1.0, 0.0, 0.4), rel=1e-13)


def _pars():
    return dict(i_deg=81.0, q=0.45, vs=380.0, r1=0.18, r2=0.27, t1=15000.0, t2=3500.0)


def test_constraints_add_and_the_weight_scales_them():
    p = _pars()
    d = PR.implied(p["i_deg"], p["q"], p["vs"], p["r1"], p["r2"], 0.3)
    K = (d["K1"] * 1.01, d["K1"] * 0.02, d["K1"] * 0.04)
    T2 = (3400.0, 100.0, 50.0)
    lg = (d["logg1"] + 0.05, 0.1, 0.1)
    c = lambda **kw: PR.Constraints(P_days=0.3, **kw)
    args = [p[k] for k in ("i_deg", "q", "vs", "r1", "r2", "t1", "t2")]
    lK = PR.log_prior(*args, c(K=K))
    lT = PR.log_prior(*args, c(T2=T2))
    lg1 = PR.log_prior(*args, c(logg1=lg))
    assert lK == pytest.approx(stats.norm.logpdf(d["K1"], K[0], K[1]), rel=1e-12)  # below the value: err_lo
    assert lT == pytest.approx(stats.norm.logpdf(3500.0, 3400.0, 50.0), rel=1e-12)  # above: err_hi
    assert lg1 == pytest.approx(stats.norm.logpdf(d["logg1"], lg[0], 0.1), rel=1e-12)
    both = PR.log_prior(*args, c(K=K, T2=T2, logg1=lg))
    assert both == pytest.approx(lK + lT + lg1, rel=1e-13)
    w = PR.log_prior(*args, PR.Constraints(P_days=0.3, K=K, T2=T2, logg1=lg, prior_weight=7.5))
    assert w == pytest.approx(7.5 * both, rel=1e-13)
    assert PR.log_prior(*args, c()) == 0.0


def test_m2_min_is_one_sided():
    p = _pars()
    args = [p[k] for k in ("i_deg", "q", "vs", "r1", "r2", "t1", "t2")]
    M2 = PR.implied(p["i_deg"], p["q"], p["vs"], p["r1"], p["r2"], 0.3)["M2"]
    assert PR.log_prior(*args, PR.Constraints(P_days=0.3, M2min=(0.9 * M2, 0.05, 0.05))) == 0.0
    got = PR.log_prior(*args, PR.Constraints(P_days=0.3, M2min=(1.1 * M2, 0.05, 0.0)))
    assert got == pytest.approx(-0.5 * (0.1 * M2 / 0.05) ** 2, rel=1e-12)
    got = PR.log_prior(*args, PR.Constraints(P_days=0.3, M2min=(1.1 * M2, 0.0, 0.0)))
    assert got == pytest.approx(-0.5 * (0.1 * M2 / (0.011 * M2)) ** 2, rel=1e-12)


@pytest.mark.parametrize("bad", [dict(i_deg=0.0), dict(q=0.0), dict(vs=-1.0), dict(r1=0.0), dict(r1=1.0)])
def test_unphysical_parameters_are_rejected(bad):
    p = _pars()
    p.update(bad)
    args = [p[k] for k in ("i_deg", "q", "vs", "r1", "r2", "t1", "t2")]
    assert PR.log_prior(*args, PR.Constraints()) == -1e30


def test_r2_outside_the_unit_interval_drops_its_constraints():
    p = _pars()
    p["r2"] = -1.0  # the reference's convention for a point-like star 2
    args = [p[k] for k in ("i_deg", "q", "vs", "r1", "r2", "t1", "t2")]
    assert PR.log_prior(*args, PR.Constraints(R2=(0.1, 0.01, 0.01), logg2=(5.0, 0.1, 0.1))) == 0.0


def test_constraints_from_config():
    cfg = {"use_priors": True, "true_period": 0.25, "prior_weight": 3.0, "mcmc_prior_weight": 1.0,
           "priors": {"K1": "120 5 0", "m2_min": "0.1 0 0.02", "logg1": "7.8 0.1 0.2"}}
    c = PR.constraints_from_config(cfg)
    assert c.P_days == 0.25 and c.prior_weight == 1.0
    assert c.K == (120.0, 5.0, 5.0) and c.M2min == (0.1, 0.02, 0.02) and c.logg1 == (7.8, 0.1, 0.2)
    assert PR.constraints_from_config(cfg, sampler=False).prior_weight == 3.0
    assert PR.constraints_from_config({"priors": {"K1": "1 1 1"}}) is None
    with pytest.raises(ValueError):
        PR.constraints_from_config({"use_priors": True, "priors": {"vsini": "1 1 1"}})
    with pytest.raises(ValueError):
        PR.constraints_from_config({"use_priors": True, "priors": {"K1": "1 1"}})


def test_model_prior_takes_varied_parameters_from_the_vector():
    import json

    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    m = L.Model.from_json(json.dumps(W.model_json(vary=("q", "iangle", "r1"))))
    names = [m.get_name(i) for i in range(m.nvary())]
    assert "q" in names and "iangle" in names
    c = PR.Constraints(P_days=0.2, K=(200.0, 10.0, 10.0), T1=(15000.0, 500.0, 500.0))
    f = PR.model_prior(m, c)
    v = np.array(m.get_param(), float)
    base = {k: getattr(m.pod, k).value for k in ("iangle", "q", "velocity_scale", "r1", "r2", "t1", "t2")}
    assert f(v) == PR.log_prior(*base.values(), c)
    v2 = v.copy()
    v2[names.index("q")] *= 1.2
    base["q"] *= 1.2
    assert f(v2) == pytest.approx(PR.log_prior(*base.values(), c), rel=1e-15)


# ---------------------------------------------------------------- LM form
def test_lm_kepler_and_derived_quantities():
    """The LM's constants are SI (G = 6.67430e-11, M_sun = 1.98892e30 kg,
    R_sun = 6.9634e8 m): a circular orbit of one such solar mass at 1 au
    gives M_total 1; K1 + K2 = vs sin i; M2_min = M2 sin^3 i; log g of one
    solar mass and radius is the solar 4.438 to the precision of these
    constants."""
    GM = PR.LM_G * PR.LM_MSUN / 1e9  # km^3 s^-2
    P_s = 2.0 * math.pi * math.sqrt(AU_KM ** 3 / GM)
    vs, P = 2.0 * math.pi * AU_KM / P_s, P_s / 86400.0
    args = (60.0, 0.4, vs, 0.1, 0.2, 5000.0, 3000.0, P)
    assert PR.lm_derived("Mtotal", *args) == pytest.approx(1.0, rel=1e-12)
    s = math.sin(math.radians(60.0))
    assert PR.lm_derived("K", *args) + PR.lm_derived("K2", *args) == pytest.approx(vs * s, rel=1e-14)
    assert PR.lm_derived("M2min", *args) == pytest.approx(PR.lm_derived("M2", *args) * s ** 3, rel=1e-14)
    assert PR.lm_derived("R1", *args) == pytest.approx(0.1 * AU_KM * 1e3 / PR.LM_RSUN, rel=1e-12)
    # M1 = 1, R1 = 1 solar unit
    q, r1 = 0.5, 0.25
    a = PR.LM_RSUN / 1e3 / r1
    P_s = 2.0 * math.pi * math.sqrt(a ** 3 / (GM * (1.0 + q)))
    args = (90.0, q, 2.0 * math.pi * a / P_s, r1, 0.2, 0.0, 0.0, P_s / 86400.0)
    assert PR.lm_derived("M1", *args) == pytest.approx(1.0, rel=1e-12)
    assert abs(PR.lm_derived("logg1", *args) - 4.438) < 2e-3


def test_lm_residual_sides_and_fallbacks():
    args = (80.0, 0.3, 300.0, 0.2, 0.3, 12000.0, 4000.0, 0.5)
    assert PR.lm_residual("T1", (11000.0, 100.0, 400.0), 4.0, *args) == pytest.approx(2.0 * 1000.0 / 400.0)
    assert PR.lm_residual("T1", (13000.0, 100.0, 400.0), 4.0, *args) == pytest.approx(2.0 * -1000.0 / 100.0)
    assert PR.lm_residual("T2", (3000.0, 0.0, 0.0), 1.0, *args) == pytest.approx(1000.0 / 30.0)


def test_lm_prior_from_a_configuration():
    import json

    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    m = L.Model.from_json(json.dumps(W.model_json(vary=("q", "iangle", "r1"))))
    cfg = {"use_priors": True, "true_period": 0.3, "prior_weight": 9.0, "mcmc_prior_weight": 1.0,
           "priors": {"T1": "15000 300 300", "K1": "150 10 20"}}
    lp = PR.LMPrior(m, cfg)
    assert len(lp) == 2 and lp.weight == 9.0 and lp.weight_sum == 18.0
    names = [m.get_name(i) for i in range(m.nvary())]
    v = np.array(m.get_param(), float)
    v[names.index("q")] = 0.31
    r = lp(v)
    pod = m.pod
    base = (pod.iangle.value, 0.31, pod.velocity_scale.value, pod.r1.value, pod.r2.value, pod.t1.value,
            pod.t2.value, 0.3)
    assert r[0] == PR.lm_residual("T1", (15000.0, 300.0, 300.0), 9.0, *base)
    assert r[1] == PR.lm_residual("K", (150.0, 10.0, 20.0), 9.0, *base)
    assert len(PR.LMPrior(m, {"use_priors": False})) == 0


# ------------------------------------------------- consistent starting point
@pytest.mark.parametrize("i_deg,K,M1,R1,P", [(85.0, 120.0, 0.5, 0.2, 0.25), (60.0, 45.0, 1.2, 1.1, 2.0),
                                             (30.0, 200.0, 0.6, 0.015, 0.1)])
def test_consistent_params_reproduce_the_targets(i_deg, K, M1, R1, P):
    """physical_prior.h:283-318: the forward quantities of the solved q,
    velocity_scale and r1 (implied(), Kepler's law) are the targets."""
    sol = PR.solve_consistent_params(i_deg, K, M1, R1, P)
    assert sol is not None
    q, vs, r1 = sol
    d = PR.implied(i_deg, q, vs, r1, 0.1, P)
    assert d["K1"] == pytest.approx(K, rel=1e-12)
    assert d["M1"] == pytest.approx(M1, rel=1e-10)  # q bisected to a ratio of 1e-12
    assert d["R1"] == pytest.approx(R1, rel=1e-12)


def test_consistent_params_fail_where_the_reference_does():
    assert PR.solve_consistent_params(0.5, 100.0, 1.0, 1.0, 1.0) is None  # sin i < 0.01
    assert PR.solve_consistent_params(85.0, 100.0, 1.0, 1e4, 1.0) is None  # r1 >= 1


def test_consistent_start_and_multistart():
    import json

    import lcurve_re_amd as L
    from lcurve_re_amd import fit as F
    from lcurve_re_amd import workloads as W
    m = L.Model.from_json(json.dumps(W.model_json(vary=("q", "iangle", "r1", "velocity_scale"))))
    names = [m.get_name(i) for i in range(m.nvary())]
    c = PR.Constraints(P_days=0.25, K=(90.0, 2.0, 2.0), M1=(0.47, 0.02, 0.02), R1=(0.2, 0.01, 0.01))
    p = PR.consistent_start(m, m.get_param(), c)
    v = dict(zip(names, p))
    d = PR.implied(v["iangle"], v["q"], v["velocity_scale"], v["r1"], 0.1, 0.25)
    assert d["K1"] == pytest.approx(90.0, rel=1e-12) and d["M1"] == pytest.approx(0.47, rel=1e-10)
    assert np.array_equal(PR.consistent_start(m, p, c), p)  # within 3 sigma: unchanged
    for st in F.multistart_starts(m, pars=p, n=6, span=0.1, seed=1, constraints=c)[1:]:
        v = dict(zip(names, st))
        d = PR.implied(v["iangle"], v["q"], v["velocity_scale"], v["r1"], 0.1, 0.25)
        assert d["K1"] == pytest.approx(90.0, rel=1e-12) and d["R1"] == pytest.approx(0.2, rel=1e-12)
