"""Writes the golden vectors of tests/golden/*.npz from the reference
compiled from its own sources (oracle/_ref/libref.so, built by
`make -C oracle ref` where the reference tree is present).

The reference's Lcurve::Model cannot be constructed in this build (its
model.cpp needs a newer nlohmann/json than the image holds), so the script
drives the functions that take explicit arguments: it derives the
set_star_grid preamble (set_star_grid.cpp:52-207) from reference calls
(xl11, xl12, ref_sphere, envelope, numface, face) and then calls the
reference add_faces, flat_sum_star{1,2}_multi, comp_star1/2, re_scale,
comp_radius1/2, planck and dlpdlt.

Run:  python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lcurve_re_amd import _abi, workloads as W  # noqa: E402
from oracle import pyoracle as P  # noqa: E402

PI = 3.14159265358979323846264


def d(x):
    return C.c_double(x)


def preamble(R, m, star, fine):
    """set_star_grid.cpp:52-207 through reference calls."""
    q, spin1, spin2 = m.q.value, m.spin1.value, m.spin2.value
    r1, r2 = m.r1.value, m.r2.value
    nlat = [[m.nlat1c, m.nlat1f], [m.nlat2c, m.nlat2f]][star][int(fine)]
    nlatfill = 0 if star == 0 else (m.nlatfill if fine else 0)
    nlngfill = 0 if star == 0 else (m.nlngfill if fine else 0)
    latlo = 0 if star == 0 else m.llo
    lathi = 0 if star == 0 else m.lhi
    rl1 = R.ref_xl11(q, spin1)
    if r1 < 0:
        r1 = rl1
    rl2 = 1.0 - R.ref_xl12(q, spin2)
    if r2 < 0:
        r2 = rl2
    assert r1 <= rl1 and r2 <= rl2
    ffac1, ffac2 = r1 / rl1, r2 / rl2
    rref1, pref1, rref2, pref2 = (C.c_double() for _ in range(4))
    R.ref_ref_sphere(q, 0, spin1, ffac1, C.byref(rref1), C.byref(pref1))
    R.ref_ref_sphere(q, 1, spin2, ffac2, C.byref(rref2), C.byref(pref2))
    infill = bool(m.npole) and star == 1 and (nlatfill > 0 or nlngfill > 0) and r2 > r1
    thelo = thehi = 0.0
    if infill:
        rangle = np.pi * m.iangle.value / 180.0
        cosi = np.cos(rangle)
        ratio = (cosi + r1) / r2
        if ratio >= 1.0:
            thehi = PI / 2.0 + rangle
        else:
            llo, lhi = 0.0, PI / 2.0
            x, y = C.c_double(), C.c_double()
            while lhi > llo + 1.0e-7:
                lmid = (llo + lhi) / 2.0
                R.ref_envelope(rangle, lmid, r1, C.byref(x), C.byref(y))
                if x.value * x.value + y.value * y.value < r2 * r2:
                    llo = lmid
                else:
                    lhi = lmid
            sini = np.sin(rangle)
            thehi = max(PI / 2.0 - np.pi * latlo / 180.0,
                        min(PI / 2.0 + rangle, np.arccos(y.value * sini / r2) + np.pi * m.lfudge / 180.0))
        ratio = (cosi - r1) / r2
        if ratio >= 1.0:
            infill = False
        elif ratio <= -1.0:
            thelo = 0.0
        else:
            thelo = min(PI / 2.0 - np.pi * lathi / 180.0,
                        max(0.0, PI / 2.0 - np.arccos(ratio) + rangle - np.pi * m.lfudge / 180.0))
    nface = R.ref_numface(nlat, int(infill), thelo, thehi, nlatfill, nlngfill)
    acc = m.delta_phase / 10.0
    gref = 1.0
    roche_self = m.roche1 if star == 0 else m.roche2
    if roche_self:
        dirn = np.array([-1.0, 0.0, 0.0] if star == 0 else [1.0, 0.0, 0.0])
        pv, dv = np.zeros(3), np.zeros(3)
        r, g = C.c_double(), C.c_double()
        st = R.ref_face(q, star, spin1 if star == 0 else spin2, P.dptr(dirn),
                        rref1.value if star == 0 else rref2.value, pref1.value if star == 0 else pref2.value,
                        acc, P.dptr(pv), P.dptr(dv), C.byref(r), C.byref(g))
        assert st == 0
        gref = g.value
    return dict(nlat=nlat, nlatfill=nlatfill, nlngfill=nlngfill, infill=infill, thelo=thelo, thehi=thehi,
                nface=nface, r1=r1, r2=r2, rref1=rref1.value, pref1=pref1.value, rref2=rref2.value,
                pref2=pref2.value, ffac1=ffac1, ffac2=ffac2, gref=gref)


def ref_grid(R, m, star, fine):
    p = preamble(R, m, star, fine)
    f = P.Faces(p["nface"])
    nf = C.c_int(0)
    dtheta = PI / p["nlat"]
    if p["infill"]:
        segs = [(0.0, p["thelo"], 0, 0), (p["thelo"], p["thehi"], p["nlatfill"], p["nlngfill"]),
                (p["thehi"], PI, 0, 0)]
    else:
        segs = [(0.0, PI, 0, 0)]
    for tlo, thi, nlf, nlg in segs:
        st = R.ref_add_faces(p["nface"], C.byref(nf), tlo, thi, dtheta, nlf, nlg, int(m.npole), star,
                             m.q.value, m.iangle.value, p["r1"], p["r2"], p["rref1"], p["rref2"], int(m.roche1),
                             int(m.roche2), m.spin1.value, m.spin2.value, int(m.eclipse1 if star == 0 else m.eclipse2),
                             p["gref"], p["pref1"], p["pref2"], p["ffac1"], p["ffac2"], m.delta_phase,
                             P.dptr(f.px), P.dptr(f.py), P.dptr(f.pz), P.dptr(f.dx), P.dptr(f.dy), P.dptr(f.dz),
                             P.fptr(f.area), P.fptr(f.gravity), P.iptr(f.necl), P.dptr(f.ingress),
                             P.dptr(f.egress))
        assert st == 0, R.ref_last_error()
    f.gravity[nf.value:] = 1.0  # Point() default for any slot add_faces leaves
    return f, p


def ref_flat_sum(R, star, flat, ph, ldc, beam, spin, VFAC, XCOFM, glens1, rlens1):
    out = np.zeros(ph.n)
    R.ref_flat_sum_multi(star, flat.n, flat.n0, flat.n1, P.dptr(flat.dx), P.dptr(flat.dy), P.dptr(flat.dz),
                         P.dptr(flat.px), P.dptr(flat.py), P.dptr(flat.pz), P.dptr(flat.flux), P.dptr(flat.in1),
                         P.dptr(flat.out1), ph.n, P.dptr(ph.ex), P.dptr(ph.ey), P.dptr(ph.ez), P.dptr(ph.phin),
                         ldc.c1, ldc.c2, ldc.c3, ldc.c4, ldc.mucrit, ldc.type, beam, spin, VFAC, XCOFM,
                         int(glens1), rlens1, P.dptr(out))
    return out


CASES = {
    # detached, both Roche, total eclipse of star 1 (the small benchmark at nlat 12)
    "detached": dict(comp={"nlat1f": 12, "nlat2f": 12, "nlat1c": 12, "nlat2c": 12}),
    # semi-detached, tiny star 1, infill strip on star 2
    "semidetached": dict(phys={"q": 0.3, "iangle": 88.0, "r1": 0.015, "r2": 0.98 * (1 - W.xl12(0.3)),
                               "t1": 30000.0, "t2": 3500.0, "absorb": 1.0, "beam_factor1": 1.2,
                               "beam_factor2": 1.5},
                         comp={"nlat1f": 14, "nlat2f": 14, "nlat1c": 10, "nlat2c": 10, "npole": 1,
                               "nlatfill": 1, "nlngfill": 1}),
    # asynchronous rotation, partial eclipses
    "async": dict(phys={"q": 0.8, "iangle": 75.0, "spin1": 1.2, "spin2": 0.9, "r1": 0.25, "r2": 0.3},
                  comp={"nlat1f": 10, "nlat2f": 10, "nlat1c": 10, "nlat2c": 10}),
    # both stars spherical: closed-form sphere eclipses
    "spheres": dict(phys={"iangle": 89.0}, comp={"roche1": 0, "roche2": 0, "nlat1f": 12, "nlat2f": 12,
                                                 "nlat1c": 12, "nlat2c": 12}),
    # low mass ratio, small primary, grazing
    "lowq": dict(phys={"q": 0.15, "iangle": 65.0, "r1": 0.02, "r2": 0.7 * (1 - W.xl12(0.15))},
                 comp={"nlat1f": 12, "nlat2f": 12, "nlat1c": 12, "nlat2c": 12}),
}


def flux_variants(case):
    """(star, ltype, beam, glens) combinations summed for a case."""
    v = [(0, 0, 0.0, 0), (0, 1, 0.0, 0), (1, 0, 0.0, 0), (1, 1, 0.0, 0), (0, 1, 1.3, 0), (1, 0, 1.7, 0)]
    if case == "spheres":
        v += [(1, 1, 0.0, 1), (1, 0, 1.5, 1)]
    return v


def main():
    R = P.ref()
    P.oracle()
    for name, spec in CASES.items():
        cfg = W.model_json(spec.get("phys"), spec.get("comp"))
        import lcurve_re_amd as L
        m = L.Model.from_json(cfg).pod
        out = {"model_json": np.frombuffer(json.dumps(cfg).encode(), np.uint8)}
        grids = {}
        for star in (0, 1):
            for fine in (1, 0):
                f, p = ref_grid(R, m, star, fine)
                key = f"s{star}{'f' if fine else 'c'}"
                grids[key] = f
                for k in ("px", "py", "pz", "dx", "dy", "dz", "area", "gravity", "necl", "ingress", "egress"):
                    out[f"{key}_{k}"] = getattr(f, k)
                out[f"{key}_gref"] = np.array([p["gref"]])
        # fluxes for the sums: the oracle continuum on the reference grids
        f1, f2 = grids["s0f"].copy(), grids["s1f"].copy()
        P.star_continuum(m, f1, f2)
        out["s0f_flux"], out["s1f_flux"] = f1.flux, f2.flux
        # sub-phases: uniform grid plus points clustered at contact phases
        rng = np.random.Generator(np.random.PCG64(7))
        phs = np.concatenate([np.arange(257) / 257.0, rng.uniform(-0.08, 0.08, 64), 0.5 + rng.uniform(-0.08, 0.08, 64),
                              [1e-9, -1e-9, 0.25, 0.75, 1.0 - 1e-12]])
        ph = P.phases_for(m.iangle.value, phs)
        out["phase"] = phs
        for k in ("ex", "ey", "ez", "phin"):
            out[f"ph_{k}"] = getattr(ph, k)
        VFAC = np.float32(m.velocity_scale.value) / (2.99792458e8 / 1.0e3)
        XCOFM = m.q.value / (1.0 + m.q.value)
        out["VFAC_XCOFM"] = np.array([VFAC, XCOFM])
        for i, (star, lt, beam, glens) in enumerate(flux_variants(name)):
            fl = P.Flat(f1 if star == 0 else f2)
            c = (0.6, -0.2, 0.3, -0.1) if lt == 1 else (0.45, 0.12, -0.05, 0.02)
            ldc = _abi.lc_ldc(*c, 0.0, lt, 0)
            spin = m.spin1.value if star == 0 else m.spin2.value
            rl = 3.0e-4 if glens else 0.0
            out[f"sum{i}"] = ref_flat_sum(R, star, fl, ph, ldc, beam, spin, VFAC, XCOFM, glens, rl)
            out[f"sum{i}_spec"] = np.array([star, lt, beam, glens, rl, spin, *c])
        # Point-path sums (comp_star1/2, ndiv=1 and ndiv=5) at a few phases
        pts = []
        for star, f in ((0, f1), (1, f2)):
            for phase, expose, ndiv in ((0.02, 0.0, 1), (0.3, 0.0, 1), (0.47, 0.01, 5), (0.9, 0.002, 3)):
                ldc = (0.6, -0.2, 0.3, -0.1, 0.0, 1) if star == 0 else (0.5315, 0.0, 0.0, 0.0, 0.0, 0)
                beam = 1.2 if star == 0 else 0.0
                val = R.ref_comp_star(star, f.n, P.dptr(f.px), P.dptr(f.py), P.dptr(f.pz), P.dptr(f.dx),
                                      P.dptr(f.dy), P.dptr(f.dz), P.fptr(f.flux), P.iptr(f.necl), P.dptr(f.ingress),
                                      P.dptr(f.egress), *ldc, m.iangle.value, phase, expose, ndiv, m.q.value, beam,
                                      np.float32(m.velocity_scale.value), 0, 0.0, 0.0, 0.5)
                pts.append([star, phase, expose, ndiv, beam, val])
        out["point_sums"] = np.array(pts)
        # comp_radius on the fine grids
        out["radius"] = np.array([R.ref_comp_radius(s, f.n, P.dptr(f.px), P.dptr(f.py), P.dptr(f.pz),
                                                    P.dptr(f.dx), P.dptr(f.dy), P.dptr(f.dz), P.fptr(f.area))
                                  for s, f in ((0, f1), (1, f2))])
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path), "bytes")

    # scalar functions
    qs = np.array([0.05, 0.1, 0.3, 0.5, 0.8, 1.0, 1.7, 3.0])
    spins = np.array([0.5, 0.9, 1.0, 1.3])
    xl = np.array([[q, s, R.ref_xl1(q), R.ref_xl11(q, s), R.ref_xl12(q, s)] for q in qs for s in spins])
    temps = np.array([3000.0, 4000.0, 5777.0, 12000.0, 30000.0, 100.0])
    waves = np.array([300.0, 547.7, 786.5, 2200.0])
    pl = np.array([[w, t, R.ref_planck(w, t), R.ref_dlpdlt(w, t)] for w in waves for t in temps])
    nf = np.array([[n, R.ref_numface(n, 0, 0.0, 0.0, 0, 0)] for n in range(1, 401)])
    # re_scale
    rng = np.random.Generator(np.random.PCG64(11))
    n = 300
    time = np.linspace(0, 1, n)
    flux = 1.0 + 0.1 * rng.standard_normal(n)
    ferr = 0.05 + 0.01 * rng.random(n)
    weight = np.where(rng.random(n) < 0.1, 0.0, 1.0 + rng.random(n))
    fit = 0.5 + 0.05 * rng.standard_normal(n)
    fit2 = fit.copy()
    chisq, wnok = C.c_double(), C.c_double()
    s = R.ref_re_scale(n, P.dptr(time), P.dptr(flux), P.dptr(ferr), P.dptr(weight), P.dptr(fit2), C.byref(chisq),
                       C.byref(wnok))
    np.savez_compressed(os.path.join(HERE, "scalars.npz"), xl=xl, planck=pl, numface=nf,
                        rs_time=time, rs_flux=flux, rs_ferr=ferr, rs_weight=weight, rs_fit=fit, rs_fit_out=fit2,
                        rs_scale=np.array([s, chisq.value, wnok.value]))
    print("wrote scalars.npz")


if __name__ == "__main__":
    main()
