"""Stage reuse of star grids (the reference's exact-match grid cache,
light_curve_comp.cpp:34-117, 156-207), on the GPU.

A reused grid is a bit-for-bit copy of the one the face and eclipse kernels
would build, so every result must be identical to an evaluation with reuse
switched off; the profile counters show that the reuse happened.
"""
import json

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import workloads as W

pytestmark = pytest.mark.gpu


def _model(**phys):
    return L.Model.from_json(json.dumps(W.model_json(phys=phys, comp={"nlat1f": 32, "nlat2f": 32,
                                                                      "nlat1c": 32, "nlat2c": 32})))


def _data(n=500):
    return L.make_data(np.arange(n) / n, ferr=1e-3, weight=1.0, flux=1.0)


@pytest.fixture()
def cold():
    e = L.Engine(0)
    e.set_stage_cache(0)
    yield e
    e.close()


def test_batch_dedupe_is_bit_identical(cold):
    base = dict(q=0.45, iangle=84.0, r1=0.12, r2=0.22)
    models = [_model(**base), _model(**base, t0=0.01), _model(**base, t1=22000.0), _model(q=0.6, iangle=80.0),
              _model(**base, gravity_dark2=0.1)]
    data = _data()
    print("stage reuse: batch of 5, three sharing the first geometry", flush=True)
    cold.reset_profile()
    calc, io, st = cold.light_curve_comp_batch(models, data, scale=True)
    assert not st.any()
    prof = cold.profile()
    # each instance has two grids (fine = coarse): 4 geometries are copies
    assert prof["grids_reused"] == 6 and prof["grids_computed"] == 4, prof
    for i, m in enumerate(models):
        c1, _, s1 = cold.light_curve_comp_batch([m], data, scale=True)
        assert s1[0] == 0
        assert np.array_equal(c1[0], calc[i]), i


def test_cross_call_reuse(cold):
    m = _model(q=0.4, iangle=86.0, r1=0.1, r2=0.25)
    data = _data()
    ref, _, _ = cold.light_curve_comp_batch([m], data)
    e = L.Engine(0)
    try:
        e.set_stage_cache(1 << 20)
        print("stage reuse: same geometry in consecutive calls", flush=True)
        a, _, sa = e.light_curve_comp_batch([m], data)
        p1 = e.profile()
        m2 = m.copy().set("t1", 25000.0).set("t0", 0.003)
        b, _, sb = e.light_curve_comp_batch([m2], data)
        p2 = e.profile()
        assert p1["grids_computed"] == 2 and p2["grids_computed"] == 2 and p2["grids_reused"] == 2, (p1, p2)
        ref2, _, _ = cold.light_curve_comp_batch([m2], data)
        assert np.array_equal(a, ref) and np.array_equal(b, ref2)
        m3 = m.copy().set("q", 0.41)   # new geometry: rebuilt
        e.light_curve_comp_batch([m3], data)
        assert e.profile()["grids_computed"] == 4
    finally:
        e.close()


def test_cache_smaller_than_a_grid(cold):
    m = _model()
    data = _data(200)
    e = L.Engine(0)
    try:
        e.set_stage_cache(100)   # fewer faces than one grid: nothing kept
        print("stage reuse: capacity below one grid", flush=True)
        a, _, _ = e.light_curve_comp_batch([m], data)
        b, _, _ = e.light_curve_comp_batch([m], data)
        assert e.profile()["grids_reused"] == 0
        assert np.array_equal(a, b)
    finally:
        e.close()


def test_autotune_picks_a_candidate_and_keeps_results(cold):
    models = [_model(q=0.3 + 0.1 * k, iangle=80.0 + 2 * k) for k in range(4)]
    data = _data(3000)
    ref, _, _ = cold.light_curve_comp_batch(models, data)
    print("autotune: 4 models x 3000 phases", flush=True)
    R, chunk = cold.autotune(models, data)
    assert R in (1, 2, 4) and chunk in (256, 512, 1024, 2048)
    got, _, st = cold.light_curve_comp_batch(models, data)
    assert not st.any()
    # another tile shape changes the FP32 summation grouping only
    assert np.max(np.abs(got - ref)) <= 1e-6 * np.max(np.abs(ref))
