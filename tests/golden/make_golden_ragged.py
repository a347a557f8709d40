"""Writes tests/golden/flat_ragged.npz: flat_sum_star{1,2}_multi of the
reference (oracle/_ref/libref.so, compiled from its own sources) on FlatGrids
whose faces carry several eclipse ranges, as star faces do when an opaque
disc's ranges follow the star's own (light_curve_comp.cpp:286-326;
FlatGrid::build, comp_light_flat.cpp:26-65, puts such faces last).

The star grids are the reference's own add_faces output for a detached
binary at nlat 45 (about 2600 faces per star, so the face classes straddle
the 1024-face blocks of the sums); face fluxes come from the oracle's
continuum.  The extra ranges are drawn from a fixed seed: a quarter of the
faces get one more range, a tenth two more, widths 0.02-0.4 in phase with
egress past 1 allowed, as the reference's ranges are stored.

Run:  python tests/golden/make_golden_ragged.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from lcurve_re_amd import _abi, workloads as W  # noqa: E402
from oracle import pyoracle as P  # noqa: E402
import make_golden as MG  # noqa: E402

NLAT = 45


def extra_ranges(n, rng):
    """face index -> list of extra (ingress, egress) ranges."""
    extra = {}
    u = rng.random(n)
    for i in range(n):
        k = 2 if u[i] < 0.10 else (1 if u[i] < 0.35 else 0)
        if k:
            lo = rng.random(k)
            w = rng.uniform(0.02, 0.4, k)
            extra[i] = list(zip(lo, lo + w))
    return extra


def ref_sum(R, star, flat, ph, ldc, beam, spin, VFAC, XCOFM, glens1, rlens1):
    out = np.zeros(ph.n)
    R.ref_flat_sum_multi_ragged(star, flat.n, flat.n0, flat.n1, P.dptr(flat.dx), P.dptr(flat.dy),
                                P.dptr(flat.dz), P.dptr(flat.px), P.dptr(flat.py), P.dptr(flat.pz),
                                P.dptr(flat.flux), P.dptr(flat.in1), P.dptr(flat.out1), P.iptr(flat.moff),
                                P.dptr(flat.min_), P.dptr(flat.mout_), ph.n, P.dptr(ph.ex), P.dptr(ph.ey),
                                P.dptr(ph.ez), P.dptr(ph.phin), ldc.c1, ldc.c2, ldc.c3, ldc.c4, ldc.mucrit,
                                ldc.type, beam, spin, VFAC, XCOFM, int(glens1), rlens1, P.dptr(out))
    return out


# (star, ltype, beam, glens, mucrit)
VARIANTS = [(0, 0, 0.0, 0, 0.0), (0, 1, 0.0, 0, 0.0), (0, 1, 1.3, 0, 0.0), (1, 0, 1.7, 0, 0.0),
            (1, 1, 0.0, 1, 0.0), (1, 0, 0.0, 0, 0.15)]


def main():
    import lcurve_re_amd as L
    R = P.ref()
    P.oracle()
    cfg = W.model_json({"iangle": 84.0}, {"nlat1f": NLAT, "nlat2f": NLAT, "nlat1c": NLAT, "nlat2c": NLAT})
    m = L.Model.from_json(cfg).pod
    f1, _ = MG.ref_grid(R, m, 0, 1)
    f2, _ = MG.ref_grid(R, m, 1, 1)
    P.star_continuum(m, f1, f2)
    out = {}
    rng = np.random.Generator(np.random.PCG64(20261204))
    extras = []
    for key, f in (("s0", f1), ("s1", f2)):
        for k in ("px", "py", "pz", "dx", "dy", "dz", "area", "gravity", "necl", "ingress", "egress", "flux"):
            out[f"{key}_{k}"] = getattr(f, k)
        ex = extra_ranges(f.n, rng)
        idx = np.array([i for i in sorted(ex) for _ in ex[i]], np.int32)
        rin = np.array([a for i in sorted(ex) for a, _ in ex[i]])
        rout = np.array([b for i in sorted(ex) for _, b in ex[i]])
        out[f"{key}_extra_face"], out[f"{key}_extra_in"], out[f"{key}_extra_out"] = idx, rin, rout
        extras.append(ex)
    phs = np.concatenate([np.arange(613) / 613.0, rng.uniform(-0.08, 0.08, 48), 0.5 + rng.uniform(-0.08, 0.08, 48)])
    ph = P.phases_for(m.iangle.value, phs)
    for k in ("ex", "ey", "ez", "phin"):
        out[f"ph_{k}"] = getattr(ph, k)
    VFAC = np.float32(m.velocity_scale.value) / (2.99792458e8 / 1.0e3)
    XCOFM = m.q.value / (1.0 + m.q.value)
    out["VFAC_XCOFM"] = np.array([VFAC, XCOFM])
    for i, (star, lt, beam, glens, muc) in enumerate(VARIANTS):
        fl = P.Flat(f1 if star == 0 else f2, extras[star])
        c = (0.6, -0.2, 0.3, -0.1) if lt == 1 else (0.45, 0.12, -0.05, 0.02)
        ldc = _abi.lc_ldc(*c, muc, lt, 0)
        spin = m.spin1.value if star == 0 else m.spin2.value
        rl = 3.0e-4 if glens else 0.0
        out[f"sum{i}"] = ref_sum(R, star, fl, ph, ldc, beam, spin, VFAC, XCOFM, glens, rl)
        out[f"sum{i}_spec"] = np.array([star, lt, beam, glens, rl, spin, *c, muc])
        print(f"sum{i}: n {fl.n} n0 {fl.n0} n1 {fl.n1} ranges {int(fl.moff[-1])}")
    path = os.path.join(HERE, "flat_ragged.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
