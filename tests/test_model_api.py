"""Lcurve::Model parameter API (lc_model_* of the C ABI, host only).

Expected values are restated here from the reference source, independently
of lcurve_re_amd/csrc/model.cpp:
* vary order: model.cpp:363-461 (nvary) -- the order get_param/set_param/
  get_name/get_range/get_dstep/get_limit/is_not_legal all follow;
* get_limit pairs: model.cpp:1286-1410;
* is_not_legal bounds: model.cpp:788-1158 (r1/r2 against xl11 and
  1 - xl12 of the model's current q and spins);
* parse rules and messages: model.cpp:10-360, lcurve.h:169-176.
"""
import json
import math

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import workloads as W

INF = math.inf

# model.cpp:363-461, with use_radii true and no disc/spot/star-spot groups
VARY_ORDER = ["q", "iangle", "r1", "r2", "spin1", "spin2", "t1", "t2",
              "ldc1_1", "ldc1_2", "ldc1_3", "ldc1_4", "ldc2_1", "ldc2_2", "ldc2_3", "ldc2_4",
              "velocity_scale", "beam_factor1", "beam_factor2", "t0", "period", "pdot", "deltat",
              "gravity_dark1", "gravity_dark2", "absorb", "slope", "quad", "cube", "third"]

# model.cpp:1286-1325
LIMITS = {"q": (0.01, 100), "iangle": (0, 90), "t0": (-INF, INF), "pdot": (-INF, INF), "absorb": (0, 1)}

# model.cpp:788-935: (lowest legal, highest legal, open-below)
LEGAL = {"q": (0.0, 100.0, True), "iangle": (0.0, 90.0, False), "spin1": (0.0, 1000.0, True),
         "spin2": (0.0, 1000.0, True), "t1": (500.0, 1e6, True), "t2": (500.0, 1e6, True),
         "ldc1_1": (-1.0, 1.0, False), "ldc2_4": (-1.0, 1.0, False), "period": (1e-3, 100.0, True),
         "deltat": (-1.0, 1.0, True), "gravity_dark1": (-1.0, 1.0, False), "gravity_dark2": (-1.0, 1.0, False),
         "absorb": (0.0, 10.0, False), "slope": (-2.0, 2.0, False), "quad": (-2.0, 2.0, False),
         "cube": (-2.0, 2.0, False)}


def _model(vary=(), phys=None, comp=None, edit=None):
    cfg = W.model_json(phys=phys, comp=comp, vary=vary)
    if edit:
        edit(cfg["model_parameters"])
    return L.Model.from_json(json.dumps(cfg)), cfg


def _pparam_fields(s):
    parts = s.split()
    return float(parts[0]), float(parts[1]), float(parts[2])


def test_parse_values_and_flags():
    m, cfg = _model()
    mp = cfg["model_parameters"]
    for name, val in W.BASE_PHYS.items():
        p = getattr(m.pod, name)
        v, r, d = _pparam_fields(mp[name])
        assert (p.value, p.range, p.dstep) == (v, r, d), name
        assert p.vary == 0 and p.defined == 1, name
    assert m.pod.delta_phase == 1e-7 and m.pod.wavelength == 547.7
    assert (m.pod.nlat1f, m.pod.nlat2c, m.pod.npole) == (40, 40, 0)
    assert (m.pod.limb1, m.pod.limb2) == (L.LC_LDC_CLARET, L.LC_LDC_POLY)
    assert m.nvary() == 0


def test_pparam_defined_defaults_true_and_explicit_false():
    def edit(mp):
        mp["q"] = "0.4 0.05 0.001 1"          # no 'defined' field -> true
        mp["t2"] = "3000 10 1 0 0"            # explicit defined = 0
    m, _ = _model(edit=edit)
    assert m.pod.q.vary == 1 and m.pod.q.defined == 1 and m.pod.q.value == 0.4
    assert m.pod.t2.defined == 0


def test_optional_parameters_default():
    def edit(mp):
        for k in ("pdot", "third", "temp_edge", "absorb_edge"):
            del mp[k]
    m, _ = _model(edit=edit)
    for k in ("pdot", "third", "temp_edge", "absorb_edge"):   # "0 1.e-10 1.e-10 0 0"
        p = getattr(m.pod, k)
        assert (p.value, p.range, p.dstep, p.vary, p.defined) == (0.0, 1e-10, 1e-10, 0, 0), k


@pytest.mark.parametrize("edit, needle", [
    (lambda cfg: cfg.pop("model_parameters"), "model_parameters"),
    (lambda cfg: cfg["model_parameters"].update(q=0.5), "Malformed model parameter value for parameter 'q'"),
    (lambda cfg: cfg["model_parameters"].pop("iangle"), "not initialised"),
    (lambda cfg: cfg["model_parameters"].update(limb1="Linear"), "limb1"),
    (lambda cfg: cfg["model_parameters"].update(glens1="1"), "glens1"),
    (lambda cfg: cfg["model_parameters"].update(q="abc"), "Pparam"),
])
def test_parse_errors(edit, needle):
    cfg = W.model_json()
    edit(cfg)
    with pytest.raises(L.LcurveError) as e:
        L.Model.from_json(json.dumps(cfg))
    assert needle in str(e.value)


def test_vary_order_matches_reference():
    m, cfg = _model(vary=VARY_ORDER)
    assert m.nvary() == len(VARY_ORDER)
    assert [m.get_name(i) for i in range(m.nvary())] == VARY_ORDER
    mp = cfg["model_parameters"]
    np.testing.assert_array_equal(m.get_param(), [_pparam_fields(mp[k])[0] for k in VARY_ORDER])
    np.testing.assert_array_equal(m.get_range(), [_pparam_fields(mp[k])[1] for k in VARY_ORDER])
    np.testing.assert_array_equal(m.get_dstep(), [_pparam_fields(mp[k])[2] for k in VARY_ORDER])


def test_fit_parameter_subset_order():
    m, _ = _model(vary=W.FIT_PARAMS)
    assert [m.get_name(i) for i in range(m.nvary())] == W.FIT_PARAMS


def test_cphi_parameterisation_replaces_radii():
    m, _ = _model(vary=["q", "r1", "r2", "cphi3", "cphi4"], comp={"use_radii": "0"})
    assert [m.get_name(i) for i in range(m.nvary())] == ["q", "cphi3", "cphi4"]


def test_set_param_round_trip():
    m, _ = _model(vary=W.FIT_PARAMS)
    v = m.get_param() * (1.0 + 1e-3) + 1e-4
    m.set_param(v)
    np.testing.assert_array_equal(m.get_param(), v)
    for k, x in zip(W.FIT_PARAMS, v):
        assert getattr(m.pod, k).value == x


def test_get_limit_matches_reference():
    m, _ = _model(vary=VARY_ORDER)
    lim = dict(zip(VARY_ORDER, m.get_limit()))
    for k in VARY_ORDER:
        assert tuple(lim[k]) == LIMITS.get(k, (0, INF)), k


@pytest.mark.parametrize("name", sorted(LEGAL))
def test_is_not_legal_bounds(name):
    m, _ = _model(vary=[name])
    lo, hi, open_below = LEGAL[name]
    assert m.is_not_legal([lo]) == open_below, (name, lo)
    assert not m.is_not_legal([hi]), (name, hi)
    assert m.is_not_legal([hi + abs(hi) * 1e-9 + 1e-12]), name
    assert m.is_not_legal([lo - abs(lo) * 1e-9 - 1e-12]), name


def test_is_not_legal_radii_use_current_roche_lobes():
    m, _ = _model(vary=["r1", "r2"])
    xl11 = W._xl11(0.5, 1.0)
    rl2 = 1.0 - W.xl12(0.5, 1.0)
    assert not m.is_not_legal([0.5 * xl11, rl2])
    assert m.is_not_legal([xl11 * (1 + 1e-9), 0.1])       # r1 >= xl11
    assert m.is_not_legal([0.1, rl2 * (1 + 1e-9)])        # r2 > 1 - xl12
    assert m.is_not_legal([0.0, 0.1])


def test_is_not_legal_batch_is_the_single_test_per_row():
    """lc_model_is_not_legal_batch: the per-vector test on every row, at the
    edges of each limit and at random points inside and outside."""
    m, _ = _model(vary=VARY_ORDER)
    rng = np.random.default_rng(4)
    p = m.get_param()
    rows = [p]
    for j, (lo, hi) in enumerate(m.get_limit()):
        for v in (lo, hi, lo - 1e-9 * max(1.0, abs(lo)), hi + 1e-9 * max(1.0, abs(hi))):
            if np.isfinite(v):
                q = p.copy()
                q[j] = v
                rows.append(q)
    rows += list(p * (1 + 0.3 * rng.standard_normal((400, len(p)))))
    P = np.array(rows)
    want = np.array([m.is_not_legal(r) for r in P])
    got = m.is_not_legal_batch(P)
    assert want.any() and not want.all()
    np.testing.assert_array_equal(got, want)
    assert m.is_not_legal_batch(P[:, :-1]).all()   # too few values
    assert m.is_not_legal_batch(np.zeros((0, len(p)))).shape == (0,)
