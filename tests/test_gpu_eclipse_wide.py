"""Bisections that share several lanes per search (k_eclipse_bisect_wide:
each step evaluates fblink at every point the next 2-4 bisection steps can
visit, then replays the one-lane path) against the one-lane loop of the
same engine (LCURVE_AMD_ECL_WIDE=0): eclipse flags, ingress and egress of
the golden star grids, and the light curves of the small detached
workload and of the disc / bright-spot binaries (whose components are
eclipsed through the same search), all bit for bit.  The one-lane loop is
pinned to the reference by tests/test_gpu_parity.py."""
import json
import os

import numpy as np
import pytest

from conftest import CASES, golden_model, load_golden

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "curves.npz")
WIDE = (4, 8, 16)


def _engine(wide):
    import lcurve_re_amd as L
    keys = ("LCURVE_AMD_ECL_WIDE",)
    old = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["LCURVE_AMD_ECL_WIDE"] = str(wide)  # fixed for calls of every size
        return L.Engine(0)  # the context reads the setting when it is created
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def engines():
    es = {w: _engine(w) for w in (0,) + WIDE}
    yield es
    for e in es.values():
        e.close()


@pytest.mark.parametrize("case", CASES)
def test_star_grid_eclipses_bit_identical(case, engines):
    from oracle import pyoracle as P
    g = load_golden(case)
    m = golden_model(g)
    for star in (0, 1):
        got = {}
        for w, eng in engines.items():
            f = P.Faces(m.grid_size(star, True))
            eng.set_star_grid(m, star, True, f.view())
            got[w] = f
        ref = got[0]
        print(f"case={case} star={star} faces={ref.n} eclipsed={int(np.count_nonzero(ref.necl))}", flush=True)
        for w in WIDE:
            assert np.array_equal(got[w].necl, ref.necl), (case, star, w)
            assert np.array_equal(got[w].ingress, ref.ingress), (case, star, w)
            assert np.array_equal(got[w].egress, ref.egress), (case, star, w)


def test_small_detached_curves_bit_identical(engines):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    w = W.detached_small()
    m = L.Model.from_json(json.dumps(w["model"]))
    data = L.make_data(w["time"], ferr=1.0, flux=0.0)
    for prec in (L.LC_PREC_FP64, L.LC_PREC_MIXED):
        out = {k: e.light_curve_comp(m, data, scale=False, rdata=True, precision=prec)[0]
               for k, e in engines.items()}
        print(f"detached_small precision {prec}: curve range {out[0].min():.6g} .. {out[0].max():.6g}", flush=True)
        for k in WIDE:
            assert np.array_equal(out[k], out[0]), (prec, k)


def test_component_curves_bit_identical(engines):
    import lcurve_re_amd as L
    gold = dict(np.load(GOLD, allow_pickle=False))
    for case in ("cv_disc_spot", "cv_disc_spot_opaque"):
        m = L.Model.from_json(json.loads(bytes(gold[f"{case}_model_json"]).decode()))
        data = L.make_data(gold["pts_time"], expose=gold["pts_expose"], ndiv=gold["pts_ndiv"],
                           ferr=gold["pts_ferr"], weight=gold["pts_weight"], flux=gold[f"{case}_dflux"])
        out = {k: e.light_curve_comp(m, data, scale=False, rdata=True, precision=L.LC_PREC_FP64)[0]
               for k, e in engines.items()}
        print(f"components {case}: {len(out[0])} points", flush=True)
        for k in WIDE:
            assert np.array_equal(out[k], out[0]), (case, k)
