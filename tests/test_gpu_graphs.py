"""Batched calls replayed as captured kernel graphs (GPU).

The second call of a launch shape captures its kernels as one hipGraph and
later calls of that shape replay it.  Every replay must give the eager
result bit for bit while the data values change between calls, with other
shapes and plain eager work in between.
"""
import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture()
def pair():
    g, e = L.Engine(0), L.Engine(0)
    e.set_graphs(0)
    for x in (g, e):
        x.set_stage_cache(0)
    yield g, e
    g.close()
    e.close()


def _same(a, b):
    ca, ia, sa = a
    cb, ib, sb = b
    assert list(sa) == list(sb)
    assert np.array_equal(ca, cb)
    assert [(x.chisq, x.wnok, x.sfac[0]) for x in ia] == [(x.chisq, x.wnok, x.sfac[0]) for x in ib]


@pytest.mark.parametrize("prec", [L.LC_PREC_MIXED, L.LC_PREC_FP64])
def test_replays_match_eager(pair, prec):
    g, e = pair
    w = W.batch_prior(m=4, nphase=700, nlat_lo=20, nlat_hi=36)
    models = [L.Model.from_json(c) for c in w["models"]]
    rng = np.random.default_rng(5)
    print(f"graphs: 4 models x 700 phases, precision {prec}", flush=True)
    for it in range(6):
        flux = 1.0 + 0.01 * rng.standard_normal(len(w["time"]))
        data = L.make_data(w["time"], flux=flux, ferr=1e-3)
        kw = dict(scale=True, precision=prec)
        _same(g.light_curve_comp_batch(models, data, **kw), e.light_curve_comp_batch(models, data, **kw))
        if it == 3:  # another shape and a single call in between
            _same(g.light_curve_comp_batch(models[:2], data, **kw), e.light_curve_comp_batch(models[:2], data, **kw))
            c1, _ = g.light_curve_comp(models[1], data, scale=False, precision=prec)
            c2, _ = e.light_curve_comp(models[1], data, scale=False, precision=prec)
            assert np.array_equal(c1, c2)
    p = g.profile()
    assert p["graph_captures"] >= 1 and p["graph_launches"] >= 4, p
    assert e.profile()["graph_launches"] == 0


def test_new_times_recapture(pair):
    """Data times enter the kernels by value (the time range): another set
    of times with the same count is another shape."""
    g, e = pair
    m = [L.Model.from_json(W.detached_small(nlat=24)["model"])]
    for k in range(4):
        t = np.linspace(0.0, 1.0 + 0.1 * (k // 2), 300, endpoint=False)
        data = L.make_data(t, flux=1.0, ferr=1e-3)
        _same(g.light_curve_comp_batch(m, data, scale=True), e.light_curve_comp_batch(m, data, scale=True))
    assert g.profile()["graph_launches"] >= 1
