"""Benchmark: light curves per second on one MI355X.

Workload (default ``batch_prior``): 512 synthetic binaries drawn from the
batched-evaluation prior (q~U[0.1,1], i~U[60,90] deg, r1~logU[0.01,0.3],
r2 = f*RL2 with f~U[0.5,1], T1~U[8000,40000] K, T2~U[3000,8000] K,
nlat~U{40..200} per instance, Claret star 1 / Poly star 2), each evaluated
at 5000 phases.  One step is one batched light_curve_comp of all 512
instances through the C ABI: face grids, FP64 eclipse searches, continuum,
flux sums, scatter and chi**2 -- everything recomputed (no stage reuse).

Prints one JSON line.  Extra keys: ``roofline`` for the flux-sum kernel
(FP32 VALU bound; algorithmic FLOP per face x sub-phase from the reference
element arithmetic, timed with HIP events on the engine stream),
``cpu_baseline`` (the FP64 CPU oracle on a bounded sample of the same
workload) and ``max_rel_err`` of the GPU curves against that oracle sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (packed FP32 instructions)
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector peak
PEAK_HBM_GBS = 8000.0
# VALU issue ceiling without packed FP32 (this build emits none): one wave64
# VALU instruction per SIMD every 4 cycles, 256 CUs x 4 SIMDs x 16 lanes x
# 2 FLOP (an FMA) x 2.4 GHz
PEAK_VALU_ISSUE_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    if name == "batch_prior":
        w = W.batch_prior(m=512 + 64)
        models = [L.Model.from_json(c) for c in w["models"]]
        cfg = {"workload": "batch_prior", "instances": 512, "phases": len(w["time"]),
               "nlat": "U{40..200}", "seed": w["seed"], "generator": w["generator"]}
    elif name == "semidetached_large":
        w = W.semidetached_large()
        models = [L.Model.from_json(w["model"])]
        cfg = {"workload": "semidetached_large", "instances": 1, "phases": len(w["time"]), "ndiv": w["ndiv"],
               "nlat_fine": 400, "nlat_coarse": 200}
    else:
        w = W.detached_small()
        models = [L.Model.from_json(w["model"])]
        cfg = {"workload": "detached_small", "instances": 1, "phases": len(w["time"]), "nlat": 40}
    data = L.make_data(w["time"], expose=w["expose"], ndiv=w["ndiv"], ferr=w["ferr"], weight=w["weight"],
                       flux=w["flux"])
    return models, data, cfg


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def curve_errors(calc, ref):
    """(max |d| / max |ref|, max |d| / |ref| pointwise over |ref| > 1e-12 max)."""
    calc, ref = np.asarray(calc, np.float64), np.asarray(ref, np.float64)
    d = np.abs(calc - ref)
    scale = np.max(np.abs(ref))
    nz = np.abs(ref) > 1e-12 * scale
    return float(d.max() / scale), float(np.max(d[nz] / np.abs(ref[nz])))


def cpu_baseline(models, data, gpu_calc, budget_s=15.0):
    """The CPU path timed on this host's cores over the first instances of
    the workload until `budget_s` elapses: the reference's own
    light_curve_comp compiled from its sources (oracle/_ref/libref.so, with
    its grid cache off, LCURVE_NO_GRID_CACHE, as in an MCMC where every
    proposal moves the geometry), then the FP64 oracle over the same
    instances as a cross-check of the port's speed.  Returns the baseline,
    the GPU curves' errors against the reference and the oracle timing."""
    from oracle import pyoracle as P
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    use_ref = P.have_ref()
    if use_ref:
        os.environ["LCURVE_NO_GRID_CACHE"] = "1"
        P.ref()
    P.oracle().orc_set_threads(threads)
    from lcurve_re_amd.workloads import ring_layout_exceeds_numface
    run = P.ref_light_curve if use_ref else P.light_curve
    t0 = time.perf_counter()
    done, e_max, e_pt, skipped = 0, 0.0, 0.0, 0
    sample = []
    for i, m in enumerate(models):
        pod = m.pod
        if use_ref and any(ring_layout_exceeds_numface(n) for n in (pod.nlat1f, pod.nlat1c, pod.nlat2f, pod.nlat2c)):
            # the reference writes such grids past the end of their vectors
            skipped += 1
            continue
        sample.append(m)
        st, ref, _ = run(pod, data, scale=False, rdata=True, diagnostics=False)
        if st == 0 and gpu_calc is not None:
            a, b = curve_errors(gpu_calc[i], ref)
            e_max, e_pt = max(e_max, a), max(e_pt, b)
        done += 1
        if time.perf_counter() - t0 > budget_s and done >= 2:
            break
    # a workload of few instances is evaluated again until the budget is used
    passes = 1
    while sample and done == len(sample) * passes and time.perf_counter() - t0 < budget_s:
        for m in sample:
            run(m.pod, data, scale=False, rdata=True, diagnostics=False)
            done += 1
        passes += 1
    el = time.perf_counter() - t0
    base = {"value": done / el, "unit": "light curves/s", "cores": threads,
            "kind": "reference" if use_ref else "port", "cpu_model": cpu_model(), "host_cores": os.cpu_count(),
            "sample": f"first {len(sample)} of {len(models)} instances of the same workload"
                      + (f" ({skipped} with nlat 153 passed over)" if skipped else "")
                      + (f", {passes} passes ({done} curves)" if passes > 1 else "") + ", "
                      + ("the reference's light_curve_comp compiled from its sources (oracle/_ref, -O3, OpenMP, "
                         "grid cache off)" if use_ref else "FP64 oracle (OpenMP, -O3 -ffp-contract=off)")
                      + f", {el:.1f} s"}
    port = None
    if use_ref:
        t1 = time.perf_counter()
        for _ in range(passes):
            for m in sample:
                P.light_curve(m.pod, data, scale=False, rdata=True, diagnostics=False)
        elp = time.perf_counter() - t1
        port = {"value": done / elp, "unit": "light curves/s", "cores": threads,
                "sample": f"the same {done} curves through the FP64 oracle, {elp:.1f} s"}
    return base, {"max_normalised": e_max, "pointwise": e_pt,
                  "against": "reference" if use_ref else "oracle", "instances": done}, port


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, scaled by a
    plain copy's readings), or None when there is none for this kernel."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    want = kernel.replace("True", "true").replace("False", "false")
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch "):
            name, val = line[len("traffic_bytes_per_launch "):].rsplit("=", 1)
            if name.strip() == want:
                return float(val)
    return None


def valu_counters(workload, kernel):
    """VALU instructions per launch of `kernel` (all, and transcendental
    FP32) and its lane utilisation, from the committed rocprofv3 counter
    summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"valu_counters_{workload}.md")
    if not os.path.exists(path):
        return None
    want = kernel.replace("True", "true").replace("False", "false")
    out = {}
    for line in open(path):
        for key in ("valu_insts_per_launch", "valu_trans_per_launch", "valu_lane_utilisation",
                    "fp32_ops_per_launch", "fp64_ops_per_launch"):
            if line.startswith(key + " "):
                name, val = line[len(key) + 1:].rsplit("=", 1)
                if name.strip() == want:
                    out[key] = float(val)
    return out if "valu_insts_per_launch" in out else None


def valu_calibration(blocks=8192, iters=4096, reps=5):
    """Issue rate of kernels with a known VALU instruction count
    (lcurve_re_amd/csrc/k_calib.hip): FP32 FMAs, FMAs with transcendental
    reciprocals, FP64 FMAs.  Instructions per second as FMA-equivalent
    TFLOP/s (wave64 instructions x 64 lanes x 2)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "lcurve_re_amd", "_lib", "liblcurve_calib.so"))
    lib.lcc_valu_issue.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    lib.lcc_valu_issue.restype = C.c_int
    out = {}
    for kind, name in ((0, "fma_f32"), (1, "fma_f32_with_rcp_f32_8to1"), (3, "fma_f32_with_rcp_f32_8to2"),
                       (2, "fma_f64")):
        ms, per_lane = C.c_double(), C.c_double()
        st = lib.lcc_valu_issue(kind, blocks, iters, reps, C.byref(ms), C.byref(per_lane))
        if st:
            raise RuntimeError(f"lcc_valu_issue({kind}): HIP error {st}")
        waves = blocks * 256 // 64
        inst = waves * per_lane.value
        fma = waves * iters * 8
        out[name] = {"ms": ms.value, "valu_wave_insts": inst, "fma_wave_insts": fma,
                     "valu_issue_tflops": inst * 128 / (ms.value * 1e-3) / 1e12,
                     "fma_tflops": fma * 128 / (ms.value * 1e-3) / 1e12}
    return out


def copy_calibration(mib=256, reps=4):
    """Plain device-to-device copies of `mib` MiB: their FETCH_SIZE/WRITE_SIZE
    counter readings against the known byte counts give the factor that turns
    the counters of the flux kernel into HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n * 4, "bytes_written": n * 4, "reps": reps,
                                           "checksum": float(y[::4096].sum().item())}}), flush=True)


def mlp_train_timing(steps=300, batch=4096, rows=65536):
    """Adam steps of the surrogate network (12 -> 4 x 512 -> 2048, BF16 on the
    matrix cores) on a synthetic data set resident on the device: time per
    step and the GEMM work per second of training wall time."""
    from lcurve_re_amd import surrogate as S
    widths = S.WIDTHS
    mlp = S.MLP(widths, max_batch=batch)
    try:
        mlp.set_params(S.init_params(widths, 3))
        rng = np.random.default_rng(1)
        mlp.set_dataset(rows)
        mlp.load_dataset(0, rng.uniform(-1, 1, (rows, widths[0])).astype(np.float32),
                         0.1 * rng.standard_normal((rows, widths[-1])).astype(np.float32))
        mlp.train(5, batch, 1e-3, 7)
        st0 = mlp.stats()
        t0 = time.perf_counter()
        loss = mlp.train(steps, batch, 1e-3, 8)
        el = time.perf_counter() - t0
        st = mlp.stats()
    finally:
        mlp.close()
    flop = st["gemm_flop"] - st0["gemm_flop"]
    return {"network": widths, "batch": batch, "steps": steps, "ms_per_step": el / steps * 1e3,
            "gemm_flop_per_step": flop / steps, "gemm_tflops_over_wall": flop / el / 1e12, "final_loss": loss}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batch_prior",
                    choices=["batch_prior", "semidetached_large", "detached_small", "fit_batch", "surrogate_mcmc"])
    ap.add_argument("--curves", type=int, default=1 << 20, help="surrogate_mcmc: training light curves")
    ap.add_argument("--train-steps", type=int, default=3000, help="surrogate_mcmc: Adam steps")
    ap.add_argument("--mcmc-steps", type=int, default=200000, help="surrogate_mcmc: walker steps in total")
    ap.add_argument("--local-curves", type=int, default=65536, help="surrogate_mcmc: curves for the local network")
    ap.add_argument("--refine-steps", type=int, default=2000, help="surrogate_mcmc: Adam steps of the local network")
    ap.add_argument("--n-sub", type=int, default=5, help="surrogate_mcmc: surrogate steps per engine evaluation")
    ap.add_argument("--k-try", type=int, default=6,
                    help="surrogate_mcmc: first-stage trials screened per walker and round (batched sampler)")
    ap.add_argument("--precision", type=int, default=None,
                    help="0: FP32 per face (default), 1: FP64 per face (default for fit_batch)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default flux tile shape")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (HBM counter calibration)")
    ap.add_argument("--mlp-train", action="store_true",
                    help="run only Adam steps of the surrogate network on synthetic data (GEMM timing)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="run only the VALU issue-rate kernels of known instruction count")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is designed, not built", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    if args.mlp_train:
        print(json.dumps({"mlp_train": mlp_train_timing()}), flush=True)
        return
    if args.valu_calibration:
        print(json.dumps({"valu_calibration": valu_calibration()}), flush=True)
        return

    import lcurve_re_amd as L
    eng = L.Engine(0)
    eng.set_stage_cache(0)   # cold: no grid kept from an earlier step
    if args.workload == "fit_batch":
        out = run_fit_batch(L, eng, args)
    elif args.workload == "surrogate_mcmc":
        out = run_surrogate_mcmc(L, eng, args)
    else:
        out = run_batch_workload(L, eng, args)
    eng.close()
    print(json.dumps(out), flush=True)


def roofline_from_profile(prof, steps, precision, workload, R=None):
    """Roofline of the dominant flux kernel: the variant (limb-darkening law
    x beaming x lensing x mucrit) with the most launch time.

    achieved = the FP32 + FP64 arithmetic the kernel issued per launch, from
    rocprofv3 counters (profiles/valu_counters_<workload>.md: (2 FMA + ADD +
    MUL + TRANS + INT32) x 64 for FP32, (2 FMA + ADD + MUL) x 64 for FP64,
    AMD's TOTAL_32_OPS / TOTAL_64_OPS), divided by the launch time measured
    here with HIP events; peak = 78.6 TFLOP/s, the VALU ceiling without
    packed FP32 (which this build does not emit; FP64 has the same peak).
    valu_issue_* = non-transcendental VALU instructions (x 64 lanes x 2,
    FMA-equivalent) against one wave64 instruction per SIMD per 4 cycles at
    2.4 GHz.  alg_* = the algorithmic count of the reference's element
    arithmetic over every face x sub-phase element, including the elements
    the kernel proves zero and skips (culled_fraction)."""
    var_ms = prof["flux_var_ms"]
    v = int(np.argmax(var_ms))
    launches = max(int(prof["flux_var_launches"][v]), 1)
    avg_ms = var_ms[v] / launches
    flop_per_launch = prof["flux_var_flops"][v] / launches
    alg = flop_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    if R is None:
        R = int(prof["last_rphase"])  # the shape the engine used
    if not precision and (v & 6) == 0 and os.environ.get("LCURVE_AMD_FLUX_WAVE", "1") != "0":
        # no beaming, no lensing: the per-wave tile kernel
        kname = f"k_flux_wave<{v & 1}, {bool(v & 8)}, {R}>"
    else:
        kname = (f"k_flux_{'fp64' if precision else 'mixed'}<{v & 1}, {bool(v & 2)}, {bool(v & 4)}"
                 + ("" if precision else f", {bool(v & 8)}") + f", {R}>")
    c_elem = {0: 24, 1: 26, 2: 39, 3: 41}[v & 3]
    evald = prof["flux_var_evaluated"][v] / launches
    traffic = measured_traffic(workload, kname)
    vc = valu_counters(workload, kname)
    t = avg_ms * 1e-3
    achieved = issue = None
    if vc and t > 0:
        achieved = (vc["fp32_ops_per_launch"] + vc["fp64_ops_per_launch"]) / t / 1e12
        issue = (vc["valu_insts_per_launch"] - vc["valu_trans_per_launch"]) * 128 / t / 1e12
    alg_elems = flop_per_launch / c_elem
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_ISSUE_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_VALU_ISSUE_TFLOPS if achieved is not None else None, "traffic": traffic,
            "achieved_source": f"profiles/valu_counters_{workload}.md" if vc else None,
            "traffic_source": f"profiles/hbm_traffic_{workload}.md" if traffic else None,
            "hbm_frac": traffic / t / (PEAK_HBM_GBS * 1e9) if traffic and t > 0 else None,
            "valu_issue_tflops_fma_equivalent": issue,
            "valu_issue_frac": issue / PEAK_VALU_ISSUE_TFLOPS if issue is not None else None,
            "valu_trans_insts_per_launch": vc["valu_trans_per_launch"] if vc else None,
            "valu_lane_utilisation": vc["valu_lane_utilisation"] if vc else None,
            "packed_fp32_peak": PEAK_FP32_TFLOPS,
            "kernel": kname, "launches_per_step": launches / steps, "avg_launch_ms": avg_ms,
            "alg_tflops": alg, "alg_frac_of_packed_fp32_peak": alg / PEAK_FP32_TFLOPS if not precision else None,
            "alg_flop_per_launch": flop_per_launch, "flop_per_element": c_elem,
            "alg_elements_per_launch": alg_elems,
            # elements left after tile culling (provably zero ones skipped)
            "evaluated_elements_per_launch": evald if not precision else None,
            "culled_fraction": 1.0 - evald / alg_elems if not precision and alg_elems > 0 else None,
            "flux_stage_alg_tflops": prof["flux_flops"] / (prof["flux_ms_total"] * 1e-3) / 1e12
            if prof["flux_ms_total"] > 0 else 0.0,
            "stage_ms_per_step": {k: prof[k] / steps for k in
                                  ("grid_ms_total", "eclipse_ms_total", "cont_ms_total", "flux_ms_total",
                                   "other_ms_total", "host_plan_ms_total", "call_ms_total")},
            "eclipse_candidates_per_step": prof["eclipse_candidates"] / steps,
            "eclipse_hits_per_step": prof["eclipse_hits"] / steps,
            "grids_computed_per_step": prof["grids_computed"] / steps,
            "grids_reused_per_step": prof["grids_reused"] / steps,
            "continua_reused_per_step": prof["continua_reused"] / steps,
            "eclipse_searches_bisected_per_step": prof["eclipse_bisected"] / steps}


def plain_gpu(L, eng, models, data, precision, batched, batched_rate, passes=2):
    """The same instances evaluated one engine call each (no batching across
    instances, the way a caller of light_curve_comp evaluates them), timed
    over `passes` passes after one untimed pass; light curves/s, the
    batched rate over it, and the largest difference from the batched
    curves relative to each curve's largest value."""
    packs = [L.ModelBatch([m]) for m in models]
    out = np.zeros((1, len(data)), np.float64)
    worst = 0.0
    for k, p in enumerate(packs):
        eng.light_curve_comp_batch(p, data, precision=precision, out=out)
        worst = max(worst, float(np.max(np.abs(out[0] - batched[k])) / max(np.max(np.abs(batched[k])), 1e-300)))
    t0 = time.perf_counter()
    for _ in range(passes):
        for p in packs:
            eng.light_curve_comp_batch(p, data, precision=precision, out=out)
    el = time.perf_counter() - t0
    rate = len(models) * passes / el
    log(f"bench: plain GPU, one call per instance: {rate:.1f} light curves/s")
    return {"value": rate, "unit": "light curves/s", "calls_per_pass": len(models), "passes": passes,
            "batched_over_plain": batched_rate / rate, "batched_output": "host memory",
            "max_rel_diff_vs_batched": worst}


def eager_profile(eng, step, nsteps):
    """Profile of nsteps calls with graph replay off: a replayed graph is
    timed as a whole, so per-kernel and per-stage times come from eager
    calls of the same step (same kernels, same shapes)."""
    eng.set_graphs(0)
    eng.reset_profile()
    for _ in range(nsteps):
        step()
    prof = eng.profile()
    eng.set_graphs(1)
    return prof


def run_batch_workload(L, eng, args):
    if args.precision is None:
        args.precision = L.LC_PREC_MIXED
    models, data, cfg = build_workload(args.workload)
    M = cfg["instances"]
    if len(models) > M:
        # draws the engine rejects (a geometry error, as the reference would
        # throw) are replaced by the next draws of the same generator
        pool, models = models[M:], models[:M]
        redrawn = 0
        while True:
            _, _, status = eng.light_curve_comp_batch(models, data, precision=args.precision)
            bad = [i for i in range(M) if status[i] != 0]
            if not bad:
                break
            for i in bad:
                models[i] = pool.pop(0)
                redrawn += 1
        cfg["redrawn"] = redrawn
    R = None
    if not args.precision and not args.no_autotune:
        # flux tile shape for this workload, timed before the measured steps
        R, chunk = eng.autotune(models, data, precision=args.precision)
        cfg["flux_tile"] = {"subphases_per_lane": R, "face_chunk": chunk, "autotuned": True}
    log(f"bench: workload {cfg}")
    # the models packed once into the C layout and the output in GPU memory
    # (lc_light_curve_comp_batch_dev), as a caller evaluating the same
    # instances repeatedly on the device would keep them; the same steps with
    # the curves copied back to host memory are timed after them
    packed = L.ModelBatch(models)
    calc = np.zeros((len(models), len(data)), np.float64)
    dcalc = L.DeviceBuffer((len(models), len(data)))

    def step_dev():
        return eng.light_curve_comp_batch_dev(packed, data, dcalc.ptr, precision=args.precision)

    def step_host():
        return eng.light_curve_comp_batch(packed, data, precision=args.precision, out=calc)[1:]

    # one untimed call of the workload's launch shape before the warm-up, so
    # that the engine's graph capture (on a shape's second call) is not in
    # the timed steps whatever --warmup is
    step_dev()
    cfg["graph_prime_calls"] = 1
    for i in range(args.warmup):
        io, status = step_dev()
        log(f"bench: warmup {i} done, failed instances: {int(np.count_nonzero(status))}")
    eng.reset_profile()
    t0 = time.perf_counter()
    for i in range(args.steps):
        io, status = step_dev()
    el = time.perf_counter() - t0
    timed = eng.profile()
    nok = int(np.count_nonzero(status == 0))
    log(f"bench: {args.steps} steps in {el:.3f} s (curves left in GPU memory); profile {timed}")
    cfg["graph_replays_per_step"] = timed["graph_launches"] / args.steps
    cfg["output"] = "device memory (lc_light_curve_comp_batch_dev)"
    step_host()   # the host-output launch shape's own capture
    for i in range(max(1, args.warmup)):
        step_host()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step_host()
    el_host = time.perf_counter() - t0
    host_out = {"value": M * args.steps / el_host, "unit": "light curves/s", "ms_per_step": el_host / args.steps * 1e3,
                "what": "the same steps with the curves copied to host memory (lc_light_curve_comp_batch; "
                        "the PCIe download is inside the step)"}
    log(f"bench: host output {host_out}")
    assert np.array_equal(dcalc.to_host()[status == 0], calc[status == 0]), "device and host curves differ"
    nprof = max(1, min(args.steps, 3))
    prof = eager_profile(eng, step_dev, nprof)
    roofline = roofline_from_profile(prof, nprof, args.precision, cfg["workload"])
    # one call per instance returns host curves, so it is set against the
    # batched steps with host output
    plain = plain_gpu(L, eng, models, data, args.precision, calc, host_out["value"])
    cpu, err, port = (None, None, None)
    if not args.no_cpu:
        cpu, err, port = cpu_baseline(models, data, calc, args.cpu_budget)
    return {"metric": "light curves/sec on 1 MI355X (batched engine, <=1e-5 rel err vs FP64 oracle)",
            "value": M * args.steps / el, "unit": "light curves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "fp64 per face and accumulation" if args.precision else
            "fp32 per face, fp64 geometry/accumulation",
            "data": "synthetic", "config": cfg, "instances_ok": nok, "roofline": roofline,
            "cpu_baseline": cpu, "cpu_oracle_port": port, "plain_gpu": plain, "host_output": host_out,
            "max_rel_err": err["max_normalised"] if err else None, "rel_err": err}


PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def run_surrogate_mcmc(L, eng, args):
    """Surrogate pipeline on the fit problem: a network trained on engine
    curves from the prior (MFMA training), a network trained on engine
    curves around the fit, then plain, delayed-acceptance and
    surrogate-transition MCMC (64 walkers, --mcmc-steps walker steps) on
    the 8000-point data."""
    from scipy import stats as SS

    from lcurve_re_amd import fit as F
    from lcurve_re_amd import surrogate as S
    from lcurve_re_amd import workloads as W
    w = W.fit_batch()
    base = L.Model.from_json(w["model"])
    eng.set_stage_cache(1 << 22)   # MCMC: walkers revisit geometries (t0/T/LD moves)
    t_all = time.perf_counter()
    sur_g, info = S.train_surrogate(eng, base, args.curves, steps=args.train_steps, log=log)
    log(f"bench: prior-trained surrogate {info}")
    plain_data = L.make_data(w["time"], ferr=w["sigma"], weight=1.0)
    curve, _ = eng.light_curve_comp(base, plain_data, scale=False, rdata=False, diagnostics=False,
                                    precision=L.LC_PREC_FP64)
    data = L.make_data(w["time"], ferr=w["sigma"], weight=1.0,
                       flux=W.make_fit_flux(curve, w["sigma"], w["noise_seed"]))
    post = S.Posterior(eng, base, data)
    p0 = base.get_param()
    d = len(p0)
    J = F.FitBatch(base, data, F.engine_evaluator(eng), scale=True).jacobian()["J"]
    # Fisher covariance, regularised (the four limb-darkening coefficients
    # are nearly degenerate), as the starting point of the adaptive pilot
    cov = np.linalg.inv(J.T @ J + np.diag(1.0 / (0.05 * np.abs(p0) + 1e-3) ** 2)) * (2.38 ** 2 / d)
    nw = w["nwalk"]
    rng = np.random.Generator(np.random.PCG64(w["walker_seed"]))
    x0 = p0 + rng.standard_normal((nw, d)) @ np.linalg.cholesky(cov / 100).T
    x0 = np.array([x if not base.is_not_legal(x) else p0 for x in x0])
    # adaptive pilot of the plain sampler: the proposal every sampler uses
    t_pilot = time.perf_counter()
    cov, x0, samples = S.tune_proposal(post.loglike, post.legal, x0, cov / 10, rounds=6, iters=60, seed=21)
    info["pilot_s"] = time.perf_counter() - t_pilot
    # the prior-trained network cannot resolve this posterior (its log-flux
    # error is ~1e-2 against a 1.5e-3 noise); the screen is a network trained
    # on engine curves where the sampler's first stage looks: pilot
    # positions moved by one proposal step
    sur, linfo = S.train_local_surrogate(eng, base, samples, cov, args.local_curves, args.refine_steps,
                                         seed=31, log=log)
    info.update(linfo)
    T_cal, cal = S.calibrate_temperature(post, sur, x0, cov, 512, seed=5)
    # stage-1 temperature: short pilot runs (not reported) at a few values,
    # keeping the one with the most accepted moves per second
    pilot = {}
    for Tc in sorted({1.0, 2.0, 4.0, round(T_cal, 3)}):
        _, stp = S.run_chains(post.loglike, post.legal, x0, cov, 60, seed=41,
                              sur_loglike=post.surrogate_loglike(sur, Tc), n_sub=1)
        pilot[Tc] = stp["accepted"] / stp["wall_s"]
    T = max(pilot, key=pilot.get)
    cal["pilot_accepted_per_s"] = {str(k): v for k, v in pilot.items()}
    # the same for the sampler with batched first-stage trials
    pilot_b = {}
    for Tc in sorted({1.0, 2.0, 4.0, round(T_cal, 3)}):
        _, stp = S.run_chains_batched(post.loglike, post.legal, x0, cov, 60, seed=43,
                                      sur_loglike=post.surrogate_loglike(sur, Tc), k_try=args.k_try)
        pilot_b[Tc] = stp["accepted"] / stp["wall_s"]
    T_b = max(pilot_b, key=pilot_b.get)
    cal["pilot_accepted_per_s_batched"] = {str(k): v for k, v in pilot_b.items()}
    log(f"bench: local surrogate {linfo}; stage-1 temperature {T:.3g} {cal}")
    n_iter = max(args.mcmc_steps // nw, 20)
    burn = n_iter // 10
    res = {}
    sl = post.surrogate_loglike(sur, T)
    for name, f, nsub in (("plain", None, 0), ("delayed_acceptance", sl, 1),
                          ("surrogate_transition", sl, args.n_sub),
                          ("delayed_acceptance_batched", post.surrogate_loglike(sur, T_b), -1)):
        eng.reset_profile()
        if nsub < 0:
            chain, st = S.run_chains_batched(post.loglike, post.legal, x0, cov, n_iter, seed=11, sur_loglike=f,
                                             k_try=args.k_try)
        else:
            chain, st = S.run_chains(post.loglike, post.legal, x0, cov, n_iter, seed=11, sur_loglike=f,
                                     n_sub=max(nsub, 1))
        e = S.ess(chain, burn)
        res[name] = {"chain": chain, "stats": st, "ess_min": float(e.min()), "ess": e,
                     "ess_per_s": float(e.min()) / st["wall_s"], "accept": st["accepted"] / st["proposals"],
                     "engine_evals_per_step": st["evaluations"] / st["proposals"],
                     "engine_call_ms": eng.profile()["call_ms_total"]}
        log(f"bench: {name}: {st} ess_min {e.min():.1f}")
    a = res["plain"]["chain"][burn:].reshape(-1, d)
    ks, shift = {}, {}
    sd = a.std(axis=0)
    others = ("delayed_acceptance", "surrogate_transition", "delayed_acceptance_batched")
    for other in others:
        b = res[other]["chain"][burn:].reshape(-1, d)
        kk = {}
        for j, name in enumerate(S.NAMES):
            # thin each walker's chain by twice its integrated
            # autocorrelation time (at one time samples still correlate
            # by ~1/e, which a two-sample KS test reads as a difference)
            ta = max(1, int(2 * (n_iter - burn) * nw / res["plain"]["ess"][j]))
            tb = max(1, int(2 * (n_iter - burn) * nw / res[other]["ess"][j]))
            ca = res["plain"]["chain"][burn::ta, :, j].ravel()
            cb = res[other]["chain"][burn::tb, :, j].ravel()
            kk[name] = float(SS.ks_2samp(ca, cb).pvalue)
        ks[other] = kk
        shift[other] = (np.abs(a.mean(axis=0) - b.mean(axis=0)) / np.where(sd > 0, sd, 1)).tolist()
    best = max(others, key=lambda k: res[k]["ess_per_s"])
    gain = res[best]["ess_per_s"] / res["plain"]["ess_per_s"]
    out_stats = {k: {kk: vv for kk, vv in v.items() if kk not in ("chain", "ess")} for k, v in res.items()}
    return {"metric": "effective samples/sec of surrogate-screened MCMC vs plain GPU MCMC",
            "value": res[best]["ess_per_s"], "unit": "effective samples/s", "n_gpus": 1,
            "steps": n_iter, "warmup": burn, "ms_per_step": res[best]["stats"]["wall_s"] / n_iter * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 MLP (fp32 accumulate); engine fp32 per face, fp64 accumulation",
            "data": "synthetic", "config": {"workload": "surrogate_mcmc", "walkers": nw, "walker_steps": n_iter * nw,
                                            "phases": len(w["time"]), "network": S.WIDTHS,
                                            "training_curves": args.curves, "prior": S.PRIOR,
                                            "n_sub": args.n_sub, "k_try": args.k_try},
            "surrogate": info, "stage1_temperature": T, "stage1_temperature_batched": T_b, "calibration": cal,
            "chains": out_stats,
            "best": best, "gain_ess_per_s": gain, "posterior_ks_pvalue": ks, "posterior_mean_shift_in_sd": shift,
            "roofline": {"bound": "mfma", "kernel": "k_gemm (training steps: all kernels of a step)",
                         "achieved": info["gemm_tflops"], "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                         "frac": (info["gemm_tflops"] or 0.0) / PEAK_BF16_TFLOPS, "traffic": None},
            "total_s": time.perf_counter() - t_all}


def run_fit_batch(L, eng, args):
    """One LM Jacobian (base + 2 x 12 central differences) and 64 MCMC
    walkers per step, all in one batched evaluation of 8000 phases."""
    from lcurve_re_amd import fit as F
    from lcurve_re_amd import workloads as W
    w = W.fit_batch()
    model = L.Model.from_json(w["model"])
    prec = L.LC_PREC_FP64 if args.precision is None else args.precision
    plain = L.make_data(w["time"], ferr=w["sigma"], weight=1.0)
    curve, _ = eng.light_curve_comp(model, plain, scale=False, rdata=False, diagnostics=False,
                                    precision=L.LC_PREC_FP64)
    data = L.make_data(w["time"], ferr=w["sigma"], weight=1.0,
                       flux=W.make_fit_flux(curve, w["sigma"], w["noise_seed"]))
    walkers = F.walkers_around(model, w["nwalk"], w["walker_seed"])
    fb = F.FitBatch(model, data, F.engine_evaluator(eng), scale=True, precision=prec)
    cfg = {"workload": "fit_batch", "phases": len(w["time"]),
           "free_parameters": [model.get_name(i) for i in range(model.nvary())], "walkers": len(walkers),
           "nlat": 40, "sigma": w["sigma"], "noise_seed": w["noise_seed"], "walker_seed": w["walker_seed"],
           "autoscale": True, "precision": "fp64" if prec == L.LC_PREC_FP64 else "mixed"}
    log(f"bench: workload {cfg}")
    out = fb.jacobian(walkers=walkers)
    # 1 base point + 2 per central column + 1 per one-sided column + walkers
    cfg["instances"] = out["instances_per_call"][0]
    cfg["difference_kinds"] = dict(zip(cfg["free_parameters"], out["kinds"]))
    for i in range(args.warmup):
        out = fb.jacobian(walkers=walkers)
    eng.reset_profile()
    t0 = time.perf_counter()
    ninst = 0
    for i in range(args.steps):
        out = fb.jacobian(walkers=walkers)
        ninst += sum(out["instances_per_call"])
    el = time.perf_counter() - t0
    timed = eng.profile()
    log(f"bench: {args.steps} steps in {el:.3f} s; profile {timed}")
    cfg["graph_replays_per_step"] = timed["graph_launches"] / args.steps
    nprof = max(1, min(args.steps, 3))
    prof = eager_profile(eng, lambda: fb.jacobian(walkers=walkers), nprof)
    roofline = roofline_from_profile(prof, nprof, prec, "fit_batch")
    # the same step with grid reuse inside the batch switched off
    os.environ["LCURVE_AMD_DEDUPE"] = "0"
    e2 = L.Engine(0)
    e2.set_stage_cache(0)
    fb2 = F.FitBatch(model, data, F.engine_evaluator(e2), scale=True, precision=prec)
    fb2.jacobian(walkers=walkers)
    t1 = time.perf_counter()
    for i in range(args.steps):
        fb2.jacobian(walkers=walkers)
    el2 = time.perf_counter() - t1
    e2.close()
    del os.environ["LCURVE_AMD_DEDUPE"]
    # a Levenberg-Marquardt solve of the same fit (up to 100 iterations from
    # a start 0.1% off in every parameter): batched damping candidates and
    # one trial per engine call take the same steps; calls and time differ
    start = F.walkers_around(model, 1, w["walker_seed"] + 1, rel=1e-3)[0]
    lm = {"start_chisq": float(fb.sum_squares([start])[0])}
    lm_pars = None
    for k in (4, 1):
        tl = time.perf_counter()
        r = F.levmarq(F.FitBatch(model, data, F.engine_evaluator(eng), scale=True, precision=prec), pars=start,
                      max_iter=100, candidates=k)
        lm_pars = r["pars"] if lm_pars is None else lm_pars
        lm[f"candidates_{k}"] = {"s": time.perf_counter() - tl, "engine_calls": r["calls"],
                                 "iterations": r["iterations"], "chisq": r["chisq"], "stop": r["stop"],
                                 "rejections": sum(1 for h in r["history"] if h[3] == "reject")}
    # multistart: 8 starts 0.1% off the LM solution (the workload's own
    # start has two limb-darkening coefficients below their lower limit 0,
    # where every clamped trial is rejected), up to 30 iterations each,
    # lockstep (one engine call per round for all starts) and one after
    # another
    m_lm = model.copy()
    m_lm.set_param(lm_pars)
    starts = F.walkers_around(m_lm, 8, w["walker_seed"] + 2, rel=1e-3)
    fbm = F.FitBatch(model, data, F.engine_evaluator(eng), scale=True, precision=prec)
    tl = time.perf_counter()
    ms_res, ms_calls = F.levmarq_multistart(fbm, starts, max_iter=30)
    t_lock = time.perf_counter() - tl
    tl = time.perf_counter()
    seq_calls = sum(F.levmarq(fbm, s0, max_iter=30)["calls"] for s0 in starts)
    t_seq = time.perf_counter() - tl
    lm["multistart_8"] = {"lockstep_s": t_lock, "lockstep_engine_calls": ms_calls, "sequential_s": t_seq,
                          "sequential_engine_calls": seq_calls, "best_chisq": float(ms_res[0]["chisq"])}
    # the multistart's optima grouped into modes (Laplace masses; every
    # mode's Jacobian in one call), then the error sampling over them (one
    # chain per mode, 400 kept steps shared by mass, at least 100 each,
    # burn-in 300), all chains' proposals of a step in one call
    from lcurve_re_amd import mcmc as MC
    tl = time.perf_counter()
    modes, mode_calls = F.laplace_modes(fbm, ms_res, dsteps=model.get_dstep())
    t_modes = time.perf_counter() - tl
    tl = time.perf_counter()
    emc = MC.error_mcmc(fbm, modes, steps=400, min_steps=100, seed=w["walker_seed"] + 3)
    t_emc = time.perf_counter() - tl
    lm["modes"] = {"s": t_modes, "engine_calls": mode_calls, "modes": len(modes),
                   "weights": [round(m["weight"], 6) for m in modes]}
    lm["error_mcmc"] = {"s": t_emc, "engine_calls": emc["calls"], "kept": [n for _, n, _ in emc["per_mode"]],
                        "acceptance": emc["acceptance"], "ok": emc["ok"], "converged": emc["converged"],
                        "chain_steps_per_s": sum(
                            n + max(300, n // 5) for _, n, _ in emc["per_mode"]) / t_emc}
    # the LM with physical priors (K1 and R1 at 1% around the solution's
    # implied values, R1 moved by 2%): continuation over 7 prior scales
    from lcurve_re_amd import prior as PRI
    mp = model.copy()
    mp.set_param(lm_pars)
    pod = mp.pod
    pargs = (pod.iangle.value, pod.q.value, pod.velocity_scale.value, pod.r1.value, pod.r2.value, pod.t1.value,
             pod.t2.value, 1.0)
    K1, R1 = PRI.lm_derived("K", *pargs), PRI.lm_derived("R1", *pargs)
    pcfg = {"use_priors": True, "true_period": 1.0, "prior_weight": 10.0,
            "priors": {"K1": f"{K1!r} {0.01 * K1!r} {0.01 * K1!r}", "R1": f"{1.02 * R1!r} {0.01 * R1!r} {0.01 * R1!r}"}}
    tl = time.perf_counter()
    rp = F.levmarq(F.FitBatch(model, data, F.engine_evaluator(eng), scale=True, precision=prec), pars=lm_pars,
                   max_iter=60, prior=PRI.LMPrior(model, pcfg))
    lm["with_prior"] = {"s": time.perf_counter() - tl, "engine_calls": rp["calls"], "iterations": rp["iterations"],
                        "stages": len(rp["stages"]), "sum_sq": rp["chisq"], "chisq_lc": rp["chisq_lc"],
                        "final_eff": rp["stages"][-1]["eff"]}
    log(f"bench: levmarq {lm}")
    # the adaptive-Metropolis sampler: 64 chains, 60 steps, all proposals of
    # a step in one batched call, started at the LM solution (inside the
    # limits: the workload's limb-darkening start is not) with initial steps
    # 2.38 / sqrt(d) times each parameter's conditional Gauss-Newton
    # deviation there, 1 / sqrt((J^T J)_ii) (the model file's 1e-4 are far
    # wider than this posterior, and the limb-darkening coefficients are too
    # degenerate for marginal deviations)
    mmc = model.copy()
    mmc.set_param(lm_pars)
    Jl = fb.jacobian(lm_pars)["J"]
    sd = 1.0 / np.sqrt(np.maximum(np.diag(Jl.T @ Jl), 1e-300))
    for i, name in enumerate(cfg["free_parameters"]):
        getattr(mmc.pod, name).dstep = 2.38 / math.sqrt(model.nvary()) * float(sd[i])
    tm = time.perf_counter()
    mc = MC.adaptive_metropolis(mmc, data, F.engine_evaluator(eng), nchains=64, nsteps=60, burn_in=20,
                                seed=w["walker_seed"], precision=prec, cov_warmup=30, adapt_interval=10)
    elm = time.perf_counter() - tm
    mcmc = {"chains": 64, "steps": 60, "s": elm, "chain_steps_per_s": 64 * 60 / elm, "engine_calls": mc["calls"],
            "acceptance_mean": float(mc["acceptance"].mean()), "start": "LM solution",
            "best_chisq_min": float(mc["best_chisq"].min())}
    log(f"bench: mcmc {mcmc}")
    cpu, jerr = None, None
    if not args.no_cpu:
        from oracle import pyoracle as P
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        P.oracle().orc_set_threads(threads)
        tc = time.perf_counter()
        ref = F.FitBatch(model, data, P.evaluator(), scale=True).jacobian(walkers=walkers)
        elc = time.perf_counter() - tc
        cpu = {"value": sum(ref["instances_per_call"]) / elc, "unit": "light curves/s", "cores": threads,
               "kind": "port", "sample": f"one full step ({sum(ref['instances_per_call'])} instances x "
                                         f"{len(w['time'])} phases), FP64 oracle, {elc:.1f} s"}
        col = [float(np.linalg.norm(out["J"][:, j] - ref["J"][:, j]) / max(np.linalg.norm(ref["J"][:, j]), 1e-300))
               for j in range(model.nvary())]
        jerr = {"jacobian_col_rel_err_max": max(col), "jacobian_col_rel_err": dict(zip(cfg["free_parameters"], col)),
                "chisq0_rel_err": abs(out["chisq0"] / ref["chisq0"] - 1.0),
                "walker_chisq_rel_err_max": float(np.nanmax(np.abs(out["walker_chisq"] / ref["walker_chisq"] - 1.0))),
                "chisq0": out["chisq0"], "chisq0_oracle": ref["chisq0"]}
    return {"metric": "light curves/sec on 1 MI355X (LM Jacobian + 64 walkers per batch, vs FP64 oracle)",
            "value": ninst / el, "unit": "light curves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "fp64 per face and accumulation" if prec == L.LC_PREC_FP64 else
            "fp32 per face, fp64 geometry/accumulation",
            "data": "synthetic", "config": cfg, "roofline": roofline,
            "no_grid_reuse": {"value": ninst / el2 if el2 > 0 else None, "ms_per_step": el2 / args.steps * 1e3},
            "levmarq": lm, "mcmc": mcmc, "cpu_baseline": cpu, "parity": jerr}


if __name__ == "__main__":
    main()
