"""The engine against the FP64 oracle (and, for the detached workload, the
reference's own golden curve) at the shapes of the five benchmark
workloads.  Each curve's error is reported two ways: max |gpu - ref| /
max |ref| and the pointwise max |gpu - ref| / |ref|; both must stay within
1e-5 (the oracle is pinned bit for bit to the reference compiled from its
sources: test_oracle_golden*.py, test_oracle_vs_reference_live.py).

* detached workload: q 0.5, i 85, nlat 40, 1000 phases, FP64 and FP32 per
  face, diagnostics included;
* semi-detached workload at full shape: coarse nlat 200, fine nlat 400 with
  the infill strip, 20000 points x 5 sub-phases, irradiation, beaming;
* batched workload: 48 draws of the 512 (the 16 with the most faces and
  the first 32) at 5000 phases in one batch;
* fit workload: the full LM Jacobian (1 + 22 central + 2 one-sided
  instances) and 64 walkers at
  8000 phases: Jacobian columns, chi**2 and the walkers' chi**2;
* surrogate network at its shape 12 -> 4 x 512 -> 2048: forward and the
  parameter gradient against the FP64 reference MLP at batch 32.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "curves.npz")


def errors(calc, ref):
    calc, ref = np.asarray(calc, np.float64), np.asarray(ref, np.float64)
    d = np.abs(calc - ref)
    scale = np.max(np.abs(ref))
    nz = np.abs(ref) > 1e-12 * scale
    return float(d.max() / scale), float(np.max(d[nz] / np.abs(ref[nz])))


def check(tag, calc, ref, tol):
    e_max, e_pt = errors(calc, ref)
    print(f"{tag}: max-normalised {e_max:.3e}, pointwise {e_pt:.3e}", flush=True)
    assert e_max <= tol and e_pt <= tol, (tag, e_max, e_pt)


def test_detached_workload_vs_reference_golden(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    g = np.load(GOLD, allow_pickle=False)
    ref, rio = g["detached_small_calc"], g["detached_small_io"]
    w = W.detached_small()
    m = L.Model.from_json(w["model"])
    data = L.make_data(w["time"], ferr=1.0, weight=1.0)
    for prec, tol in ((L.LC_PREC_FP64, 1e-9), (L.LC_PREC_MIXED, 1e-5)):
        print(f"detached workload nlat 40 x 1000 phases, precision {prec}", flush=True)
        calc, io = engine.light_curve_comp(m, data, scale=False, rdata=True, diagnostics=True, precision=prec)
        check(f"detached precision {prec}", calc, ref, tol)
        got = [io.wdwarf, io.logg1, io.logg2, io.rv1, io.rv2]
        want = [rio[5], rio[8], rio[9], rio[10], rio[11]]
        for k, (a, b) in enumerate(zip(got, want)):
            assert abs(a - b) <= 1e-5 * abs(b), (k, a, b)


def test_semidetached_workload_full_shape(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    w = W.semidetached_large()
    m = L.Model.from_json(w["model"])
    data = L.make_data(w["time"], expose=w["expose"], ndiv=w["ndiv"], ferr=1.0, weight=1.0)
    print("semi-detached workload nlat 200/400 + infill, 20000 x 5: oracle", flush=True)
    st, ref, rio = P.light_curve(m.pod, data, scale=False, rdata=True, diagnostics=True)
    assert st == 0
    for prec in (L.LC_PREC_MIXED, L.LC_PREC_FP64):
        print(f"semi-detached workload on the GPU, precision {prec}", flush=True)
        calc, io = engine.light_curve_comp(m, data, scale=False, rdata=True, diagnostics=True, precision=prec)
        check(f"semidetached precision {prec}", calc, ref, 1e-5)
        for k in ("logg1", "logg2", "rv1", "rv2", "wdwarf"):
            assert abs(getattr(io, k) / getattr(rio, k) - 1) <= 1e-5, k


def test_batched_workload_draws(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    w = W.batch_prior(m=512)
    nlat = np.array([p[-1] for p in w["params"]])
    top = [int(i) for i in np.argsort(-nlat, kind="stable")[:16]]
    pick = top + [i for i in range(len(nlat)) if i not in top][:32]
    models = [L.Model.from_json(w["models"][i]) for i in pick]
    data = L.make_data(w["time"], ferr=1.0, weight=1.0)
    print(f"batched workload: 48 draws x 5000 phases in one batch, nlat {sorted(nlat[pick])[-3:]} at the top",
          flush=True)
    calc, _, status = engine.light_curve_comp_batch(models, data, scale=False, rdata=True,
                                                    precision=L.LC_PREC_MIXED)
    worst = (0.0, 0.0)
    for k, (i, m) in enumerate(zip(pick, models)):
        st, ref, _ = P.light_curve(m.pod, data, scale=False, rdata=True)
        assert status[k] == st, (i, status[k], st)
        if st:
            continue
        e = errors(calc[k], ref)
        worst = (max(worst[0], e[0]), max(worst[1], e[1]))
        assert e[0] <= 1e-5 and e[1] <= 1e-5, (i, int(nlat[i]), e)
    print(f"batched workload: worst max-normalised {worst[0]:.3e}, pointwise {worst[1]:.3e}", flush=True)


def test_fit_workload_full_shape(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import fit as F
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    w = W.fit_batch()
    m = L.Model.from_json(w["model"])
    plain = L.make_data(w["time"], ferr=w["sigma"], weight=1.0)
    st, curve, _ = P.light_curve(m.pod, plain, scale=False, rdata=False)
    assert st == 0
    data = L.make_data(w["time"], ferr=w["sigma"], weight=1.0, flux=W.make_fit_flux(curve, w["sigma"],
                                                                                     w["noise_seed"]))
    walkers = F.walkers_around(m, w["nwalk"], w["walker_seed"])
    print("fit workload: LM Jacobian (12 columns) + 64 walkers x 8000 phases (FP64 per face)", flush=True)
    g = F.FitBatch(m, data, F.engine_evaluator(engine), scale=True).jacobian(walkers=walkers)
    o = F.FitBatch(m, data, P.evaluator(), scale=True).jacobian(walkers=walkers)
    # 1 base point, 2 per central column, 1 per one-sided column (limb
    # darkening below its lower limit 0), 64 walkers
    assert sum(g["instances_per_call"]) == 1 + sum(2 if k == "c" else 1 for k in g["kinds"]) + 64
    assert abs(g["chisq0"] / o["chisq0"] - 1.0) < 1e-6
    np.testing.assert_allclose(g["walker_chisq"], o["walker_chisq"], rtol=1e-6)
    worst = 0.0
    for j in range(m.nvary()):
        ref = o["J"][:, j]
        err = float(np.linalg.norm(g["J"][:, j] - ref) / max(np.linalg.norm(ref), 1e-300))
        worst = max(worst, err)
        assert err < 1e-3, (m.get_name(j), err)
    check("fit workload base curve", g["calc0"], o["calc0"], 1e-5)
    print(f"fit workload: worst Jacobian column {worst:.3e}", flush=True)


@pytest.mark.parametrize("batch", [32, 320])
def test_surrogate_network_full_shape(batch):
    from lcurve_re_amd import surrogate as S
    from oracle import mlp_ref as R
    widths = S.WIDTHS
    assert widths == [12, 512, 512, 512, 512, 2048]
    rng = np.random.default_rng(21)
    p = S.init_params(widths, 9)
    x = rng.uniform(-1, 1, (batch, 12)).astype(np.float32)
    t = rng.standard_normal((batch, 2048)).astype(np.float32)
    print(f"surrogate network 12-4x512-2048, batch {batch}: forward and gradient", flush=True)
    mlp = S.MLP(widths, max_batch=max(64, batch))
    try:
        mlp.set_params(p)
        y = mlp.forward(x)
        loss, gr = mlp.grad(x, t)
    finally:
        mlp.close()
    p64 = p.astype(np.float64)
    yr = R.forward(p64, widths, x, bf=True)[0]
    e = float(np.max(np.abs(y - yr)) / np.max(np.abs(yr)))
    lr, gref = R.grad(p64, widths, x, t, bf=True)
    print(f"surrogate network: forward max-normalised {e:.3e}, loss {loss:.6g} vs {lr:.6g}", flush=True)
    assert e <= 2e-3                     # bf16 inputs of each layer, fp32 accumulation
    assert abs(loss / lr - 1) < 1e-4
    o = 0
    for l in range(1, len(widths)):
        k, n = widths[l - 1], widths[l]
        for part in (k * n, n):
            a, b = gr[o:o + part], gref[o:o + part]
            rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            assert rel <= 1e-2, (l, part, rel)
            o += part
    assert o == len(gr)
