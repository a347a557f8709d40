"""Synthetic binaries and data sets of the benchmark workloads.

Every choice the workload descriptions leave open is fixed here and
recorded in the returned dictionaries (limb-darkening law of star 2,
velocity scale, seeds, infill latitudes, ...).

* ``detached_small``   -- q=0.5, i=85, r1=0.10, r2=0.20, T=20000/4000 K,
  beta=0.25/0.08, Claret (0.6,-0.2,0.3,-0.1) for star 1, nlat 40 both
  stars, 1000 phases k/1000.
* ``semidetached_large`` -- q=0.3, i=88, r1=0.015, r2 at 98% of its Roche
  lobe, T=30000/3500 K, irradiation on, beaming 1.2/1.5, coarse nlat 200,
  fine nlat 400 with an infill strip, 20000 phases, 5 sub-exposures each.
* ``batch_prior``      -- 512 draws of (q, i, r1, r2 filling factor, T1, T2,
  nlat) from a seeded generator, 5000 phases each.
* ``fit_batch``        -- the detached curve at 8000 phases plus noise, and
  the 89 parameter vectors of one LM iteration with 64 walkers.
"""
import json
import math

import numpy as np

# Pparam strings are "value range dstep vary [defined]" (lcurve.h:169-176)


def _pp(v, vary=0, rng=0.1, dstep=1e-4):
    return f"{v!r} {rng!r} {dstep!r} {int(vary)} 1"


BASE_PHYS = {
    "q": 0.5, "iangle": 85.0, "r1": 0.10, "r2": 0.20, "cphi3": 0.01, "cphi4": 0.055,
    "spin1": 1.0, "spin2": 1.0, "t1": 20000.0, "t2": 4000.0,
    "ldc1_1": 0.6, "ldc1_2": -0.2, "ldc1_3": 0.3, "ldc1_4": -0.1,
    "ldc2_1": 0.5315, "ldc2_2": 0.0, "ldc2_3": 0.0, "ldc2_4": 0.0,
    "velocity_scale": 300.0, "beam_factor1": 0.0, "beam_factor2": 0.0,
    "t0": 0.0, "period": 1.0, "pdot": 0.0, "deltat": 0.0,
    "gravity_dark1": 0.25, "gravity_dark2": 0.08, "absorb": 0.0,
    "slope": 0.0, "quad": 0.0, "cube": 0.0, "third": 0.0,
    "rdisc1": 0.0, "rdisc2": 0.0, "height_disc": 0.0, "beta_disc": 0.0, "temp_disc": 0.0,
    "texp_disc": 0.0, "lin_limb_disc": 0.0, "quad_limb_disc": 0.0, "temp_edge": 0.0,
    "absorb_edge": 0.0, "radius_spot": 0.0, "length_spot": 0.0, "height_spot": 0.0,
    "expon_spot": 0.0, "epow_spot": 0.0, "angle_spot": 0.0, "yaw_spot": 0.0, "temp_spot": 0.0,
    "tilt_spot": 0.0, "cfrac_spot": 0.0,
}

BASE_COMP = {
    "delta_phase": "1e-07", "nlat1f": "40", "nlat2f": "40", "nlat1c": "40", "nlat2c": "40",
    "npole": "0", "nlatfill": "0", "nlngfill": "0", "lfudge": "0", "llo": "90", "lhi": "-90",
    "phase1": "0.1", "phase2": "0.4", "wavelength": "547.7", "roche1": "1", "roche2": "1",
    "eclipse1": "1", "eclipse2": "1", "glens1": "0", "use_radii": "1", "tperiod": "1",
    "gdark_bolom1": "1", "gdark_bolom2": "1", "mucrit1": "0", "mucrit2": "0",
    "limb1": "Claret", "limb2": "Poly", "mirror": "0", "add_disc": "0", "nrad": "40",
    "opaque": "0", "add_spot": "0", "nspot": "0", "iscale": "0",
}


def model_json(phys=None, comp=None, vary=()):
    """Reference-format configuration ({"model_parameters": {...}})."""
    p = dict(BASE_PHYS)
    p.update(phys or {})
    c = dict(BASE_COMP)
    c.update({k: str(v) for k, v in (comp or {}).items()})
    mp = {k: _pp(v, k in vary) for k, v in p.items()}
    mp.update(c)
    return {"model_parameters": mp}


def ring_layout_exceeds_numface(nlat):
    """True when the rings of a grid without infill hold more faces than
    numface() allots: add_faces counts max(16, int(2 pi sin(theta) / dtheta))
    per ring (set_star_grid.cpp:261-273), numface max(16, int(2 pi / dphi))
    with dphi = dtheta / sin(theta) (numface.cpp:41-50), and the two roundings
    differ for some nlat (153, 297, 315, ...).  The reference writes the extra
    faces past the end of its vector."""
    pi = 3.14159265358979323846
    dtheta = pi / nlat
    n_num = n_ring = 0
    for i in range(int(math.ceil(pi / dtheta))):
        theta = pi * (i + 0.5) / nlat
        n_num += max(16, int(2 * pi / (dtheta / math.sin(theta))))
    for nt in range(nlat):
        sint = math.sin(0.0 + (pi - 0.0) * (nt + 0.5) / nlat)
        n_ring += max(16, int(2 * pi * sint / dtheta))
    return n_ring > n_num


def xl12(q, spin=1.0):
    """L1 distance from star 1 for an asynchronous star 2 (roche.cpp:45-80)."""
    ssq = spin * spin
    mu = q / (1.0 + q)
    a = [-1.0 + mu, 2.0 - 2.0 * mu, -ssq + mu, 3.0 * ssq + 2.0 * mu - 2.0, 1.0 - mu - 3.0 * ssq, ssq]
    x, xold, n = 1.0 / (1.0 + q), 0.0, 0
    while n < 1000 and abs(x - xold) > 1e-12 * abs(x):
        xold = x
        f = x * (x * (x * (x * (x * a[5] + a[4]) + a[3]) + a[2]) + a[1]) + a[0]
        df = x * (x * (x * (x * 5 * a[5] + 4 * a[4]) + 3 * a[3]) + 2 * a[2]) + a[1]
        x -= f / df
        n += 1
    return x


def detached_small(nlat=40, nphase=1000):
    cfg = model_json(comp={"nlat1f": nlat, "nlat2f": nlat, "nlat1c": nlat, "nlat2c": nlat})
    t = np.arange(nphase, dtype=np.float64) / nphase
    return {"name": "detached_small", "model": cfg, "time": t, "expose": 0.0, "ndiv": 1,
            "ferr": 1.0, "weight": 1.0, "flux": 0.0}


def semidetached_large(nlat_coarse=200, nlat_fine=400, nphase=20000, ndiv=5):
    r2 = 0.98 * (1.0 - xl12(0.3, 1.0))
    cfg = model_json(
        phys={"q": 0.3, "iangle": 88.0, "r1": 0.015, "r2": r2, "t1": 30000.0, "t2": 3500.0,
              "absorb": 1.0, "beam_factor1": 1.2, "beam_factor2": 1.5, "velocity_scale": 300.0},
        comp={"nlat1c": nlat_coarse, "nlat2c": nlat_coarse, "nlat1f": nlat_fine, "nlat2f": nlat_fine,
              "npole": 1, "nlatfill": 1, "nlngfill": 1, "phase1": 0.1, "phase2": 0.4})
    t = np.arange(nphase, dtype=np.float64) / nphase
    return {"name": "semidetached_large", "model": cfg, "time": t, "expose": 1.0 / nphase, "ndiv": ndiv,
            "ferr": 1.0, "weight": 1.0, "flux": 0.0}


def batch_prior(m=512, nphase=5000, seed=20261203, nlat_lo=40, nlat_hi=200):
    """m parameter vectors from the batched-evaluation prior (numpy PCG64
    generator seeded with `seed`); draws with r1 >= RL1 are redrawn."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cfgs, params = [], []
    while len(cfgs) < m:
        q = rng.uniform(0.1, 1.0)
        inc = rng.uniform(60.0, 90.0)
        r1 = math.exp(rng.uniform(math.log(0.01), math.log(0.3)))
        f = rng.uniform(0.5, 1.0)
        t1 = rng.uniform(8000.0, 40000.0)
        t2 = rng.uniform(3000.0, 8000.0)
        nlat = int(rng.integers(nlat_lo, nlat_hi + 1))
        rl1 = _xl11(q)
        if r1 >= rl1:
            continue
        r2 = f * (1.0 - xl12(q, 1.0))
        cfgs.append(model_json(phys={"q": q, "iangle": inc, "r1": r1, "r2": r2, "t1": t1, "t2": t2},
                               comp={"nlat1f": nlat, "nlat2f": nlat, "nlat1c": nlat, "nlat2c": nlat}))
        params.append((q, inc, r1, r2, t1, t2, nlat))
    t = np.arange(nphase, dtype=np.float64) / nphase
    return {"name": "batch_prior", "models": cfgs, "params": params, "time": t, "expose": 0.0, "ndiv": 1,
            "ferr": 1.0, "weight": 1.0, "flux": 0.0, "seed": seed, "generator": "numpy PCG64"}


def _xl11(q, spin=1.0):
    ssq = spin * spin
    mu = q / (1.0 + q)
    a = [-1.0 + mu, 2.0 - 2.0 * mu, -1.0 + mu, ssq + 2.0 * mu, -2.0 * ssq - mu, ssq]
    x, xold, n = 1.0 / (1.0 + q), 0.0, 0
    while n < 1000 and abs(x - xold) > 1e-12 * abs(x):
        xold = x
        f = x * (x * (x * (x * (x * a[5] + a[4]) + a[3]) + a[2]) + a[1]) + a[0]
        df = x * (x * (x * (x * 5 * a[5] + 4 * a[4]) + 3 * a[3]) + 2 * a[2]) + a[1]
        x -= f / df
        n += 1
    return x


# Free parameters of the LM fit, in the reference's vary order
# (model.cpp:363-461)
FIT_PARAMS = ["q", "iangle", "r1", "r2", "t1", "t2", "ldc1_1", "ldc1_2", "ldc1_3", "ldc1_4", "t0",
              "gravity_dark2"]


def fit_batch(nphase=8000, nwalk=64, sigma=1.5e-3, noise_seed=20261204, walker_seed=20261205):
    """The detached binary with the 12 LM parameters free, evaluated at
    nphase phases k/nphase; the data are its own light curve normalised to
    mean 1 plus N(0, sigma^2) noise (make_fit_flux), fitted with autoscale.
    One batch = 1 base point + 2 x 12 central differences + nwalk walkers."""
    cfg = model_json(vary=FIT_PARAMS)
    t = np.arange(nphase, dtype=np.float64) / nphase
    return {"name": "fit_batch", "model": cfg, "time": t, "expose": 0.0, "ndiv": 1, "ferr": sigma,
            "weight": 1.0, "sigma": sigma, "noise_seed": noise_seed, "walker_seed": walker_seed, "nwalk": nwalk,
            "generator": "numpy PCG64"}


def make_fit_flux(curve, sigma=1.5e-3, seed=20261204):
    """curve / mean(curve) + N(0, sigma^2) from a PCG64 generator."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = np.asarray(curve, np.float64)
    return c / np.mean(c) + sigma * rng.standard_normal(c.size)


def to_json_text(cfg):
    return json.dumps(cfg)
