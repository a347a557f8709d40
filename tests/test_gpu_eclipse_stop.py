"""Eclipse searches that end once no tested phase can tell (GPU).

The batched engine stops an ingress/egress bisection as soon as its bracket
lies within one integer and either holds none of the phases at which that
instance's eclipse ranges are ever compared (its sub-phases, the Ginterp
phases, phase 0.5), or spans only phases at which the face is not seen
(mu <= mucrit, where the reference's element is zero whatever the range
says).  Every comparison (ecl_range, comp_light_flat.cpp:171-173) then has
the same answer as with the reference's full bisection to delta_phase, or
multiplies a zero, so in FP64 mode light curves and scalars must be
bit-identical to an engine that always bisects to the end
(LCURVE_AMD_FULL_ECLIPSE).  In mixed mode a face whose stopped range puts a
tested phase on the other side joins another of the flux kernel's compacted
lists, which changes only the grouping of the FP32 partial sums.
"""
import json

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture()
def engines(monkeypatch):
    """(full searches in one launch, early stop in one launch, early stop
    in rounds).  By default only calls of 2^17 faces and more stop early
    (LCURVE_AMD_ECL_STOP_MIN_FACES) and only those of 2^20 faces and more
    run in rounds (LCURVE_AMD_ECL_ROUNDS_MIN, in faces x 2); both are
    lowered here so that these small batches take every path."""
    monkeypatch.setenv("LCURVE_AMD_FULL_ECLIPSE", "1")
    monkeypatch.setenv("LCURVE_AMD_ECL_ROUNDS_MIN", str(1 << 40))
    monkeypatch.setenv("LCURVE_AMD_ECL_STOP_MIN_FACES", "0")  # small batches stop early too
    full = L.Engine(0)
    monkeypatch.delenv("LCURVE_AMD_FULL_ECLIPSE")
    fast = L.Engine(0)
    monkeypatch.setenv("LCURVE_AMD_ECL_ROUNDS_MIN", "0")
    rounds = L.Engine(0)
    es = (full, fast, rounds)
    for e in es:
        e.set_stage_cache(0)
    yield es
    for e in es:
        e.close()


def _models():
    w = W.batch_prior(m=5, nphase=10, nlat_lo=24, nlat_hi=40)
    ms = [L.Model.from_json(c) for c in w["models"]]
    # fine/coarse grids with Ginterp switching, beaming, an exposure split
    ms.append(L.Model.from_json(json.dumps(W.model_json(
        phys={"q": 0.3, "iangle": 86.0, "r1": 0.05, "r2": 0.9 * (1.0 - W.xl12(0.3, 1.0)), "t1": 25000.0,
              "t2": 4000.0, "absorb": 1.0, "beam_factor1": 1.2, "t0": 0.013},
        comp={"nlat1c": 24, "nlat2c": 24, "nlat1f": 48, "nlat2f": 48, "phase1": 0.1, "phase2": 0.4}))))
    # a grazing eclipse and an asynchronous companion
    ms.append(L.Model.from_json(json.dumps(W.model_json(
        phys={"iangle": 74.0, "spin2": 0.7, "t0": -0.002},
        comp={"nlat1f": 36, "nlat2f": 36, "nlat1c": 36, "nlat2c": 36}))))
    # the same geometry at another t0: one grid built for two instances
    ms.append(ms[0].copy().set("t0", 0.0031))
    return ms


@pytest.mark.parametrize("prec", [L.LC_PREC_FP64, L.LC_PREC_MIXED])
@pytest.mark.parametrize("nphase,expose,ndiv", [(2000, 0.0, 1), (500, 0.004, 5), (3, 0.0, 1)])
def test_early_stop_matches_full_searches(engines, prec, nphase, expose, ndiv):
    full = engines[0]
    models = _models()
    t = np.sort(np.random.default_rng(nphase).uniform(-0.3, 1.7, nphase))
    data = L.make_data(t, expose=expose, ndiv=ndiv, flux=1.0, ferr=1e-3)
    print(f"eclipse stop: {len(models)} models, {nphase} points, ndiv {ndiv}, precision {prec}", flush=True)
    kw = dict(scale=True, rdata=True, diagnostics=True, precision=prec)
    ca, ia, sa = full.light_curve_comp_batch(models, data, **kw)
    assert not sa.any(), sa
    for e in engines[1:]:
        cb, ib, sb = e.light_curve_comp_batch(models, data, **kw)
        assert list(sa) == list(sb), (sa, sb)
        if prec == L.LC_PREC_FP64:
            assert np.array_equal(ca, cb)
            for a, b in zip(ia, ib):
                assert (a.chisq, a.wnok, a.wdwarf, a.logg1, a.logg2, a.rv1, a.rv2, a.sfac[0]) == \
                       (b.chisq, b.wnok, b.wdwarf, b.logg1, b.logg2, b.rv1, b.rv2, b.sfac[0])
        else:
            rel = float(np.max(np.abs(ca - cb) / np.max(np.abs(ca), axis=1, keepdims=True)))
            print(f"   mixed: largest difference {rel:.3e} of each curve's maximum", flush=True)
            assert rel <= 1e-7
            for a, b in zip(ia, ib):
                # logg and rv do not depend on the ranges; wdwarf carries the fitted scale
                assert (a.wnok, a.logg1, a.logg2, a.rv1, a.rv2) == (b.wnok, b.logg1, b.logg2, b.rv1, b.rv2)
                for x, y in ((a.chisq, b.chisq), (a.sfac[0], b.sfac[0]), (a.wdwarf, b.wdwarf)):
                    assert abs(x - y) <= 1e-6 * abs(x), (x, y)


def test_rounds_give_the_same_grid(engines):
    """set_star_grid bisects to delta_phase; in rounds or in one launch the
    eclipse ranges are the same bits."""
    from oracle import pyoracle as P  # face buffers only
    full, _, rounds = engines
    m = _models()[-2]
    for star in (0, 1):
        n = m.grid_size(star, True)
        a, b = P.Faces(n), P.Faces(n)
        full.set_star_grid(m, star, True, a.view())
        rounds.set_star_grid(m, star, True, b.view())
        assert np.count_nonzero(a.necl) > 0
        for k in ("necl", "ingress", "egress", "px", "dx", "area"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), (star, k)
