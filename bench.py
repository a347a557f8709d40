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
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0


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


def cpu_baseline(models, data, gpu_calc, budget_s=15.0):
    """FP64 CPU oracle over the first instances until `budget_s` elapses."""
    from oracle import pyoracle as P
    lib = P.oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    lib.orc_set_threads(threads)
    t0 = time.perf_counter()
    done, errs = 0, []
    for i, m in enumerate(models):
        st, ref, _ = P.light_curve(m.pod, data, scale=False, rdata=True, diagnostics=False)
        if st == 0 and gpu_calc is not None:
            errs.append(float(np.max(np.abs(gpu_calc[i] - ref)) / np.max(np.abs(ref))))
        done += 1
        if time.perf_counter() - t0 > budget_s and done >= 2:
            break
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "light curves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(models)} instances of the same workload, FP64 oracle "
                      f"(OpenMP, -O3 -ffp-contract=off), {el:.1f} s"}, (max(errs) if errs else None)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batch_prior",
                    choices=["batch_prior", "semidetached_large", "detached_small"])
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is designed, not built", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return

    import lcurve_re_amd as L
    models, data, cfg = build_workload(args.workload)
    eng = L.Engine(0)
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
    log(f"bench: workload {cfg}")
    calc = None
    for i in range(args.warmup):
        calc, io, status = eng.light_curve_comp_batch(models, data, precision=args.precision)
        log(f"bench: warmup {i} done, failed instances: {int(np.count_nonzero(status))}")
    eng.reset_profile()
    t0 = time.perf_counter()
    for i in range(args.steps):
        calc, io, status = eng.light_curve_comp_batch(models, data, precision=args.precision)
    el = time.perf_counter() - t0
    prof = eng.profile()
    nok = int(np.count_nonzero(status == 0))
    log(f"bench: {args.steps} steps in {el:.3f} s; profile {prof}")
    # dominant flux kernel: the variant (limb-darkening law x beaming x
    # lensing x mucrit) with the most launch time in the timed region
    var_ms = prof["flux_var_ms"]
    v = int(np.argmax(var_ms))
    launches = max(int(prof["flux_var_launches"][v]), 1)
    avg_ms = var_ms[v] / launches
    flop_per_launch = prof["flux_var_flops"][v] / launches
    achieved = flop_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    R = int(os.environ.get("LCURVE_AMD_FLUX_R", "2"))
    kname = (f"k_flux_{'fp64' if args.precision else 'mixed'}<{v & 1}, {bool(v & 2)}, {bool(v & 4)}"
             + ("" if args.precision else f", {bool(v & 8)}") + f", {R}>")
    c_elem = {0: 24, 1: 26, 2: 39, 3: 41}[v & 3]
    evald = prof["flux_var_evaluated"][v] / launches
    traffic = measured_traffic(cfg["workload"], kname)
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "traffic_source": f"profiles/hbm_traffic_{cfg['workload']}.md" if traffic else None,
                "hbm_frac": traffic / (avg_ms * 1e-3) / (PEAK_HBM_GBS * 1e9) if traffic and avg_ms > 0 else None,
                "kernel": kname, "launches_per_step": launches / args.steps, "avg_launch_ms": avg_ms,
                "alg_flop_per_launch": flop_per_launch, "flop_per_element": c_elem,
                "alg_elements_per_launch": flop_per_launch / c_elem,
                # elements left after tile culling (provably zero ones skipped):
                # the rate the VALUs actually ran at
                "evaluated_elements_per_launch": evald,
                "evaluated_tflops": evald * c_elem / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0,
                "flux_stage_tflops": prof["flux_flops"] / (prof["flux_ms_total"] * 1e-3) / 1e12
                if prof["flux_ms_total"] > 0 else 0.0,
                "stage_ms_per_step": {k: prof[k] / args.steps for k in
                                      ("grid_ms_total", "eclipse_ms_total", "cont_ms_total", "flux_ms_total",
                                       "other_ms_total", "host_plan_ms_total", "call_ms_total")},
                "eclipse_candidates_per_step": prof["eclipse_candidates"] / args.steps,
                "eclipse_hits_per_step": prof["eclipse_hits"] / args.steps}
    cpu, err = (None, None)
    if not args.no_cpu:
        cpu, err = cpu_baseline(models, data, calc, args.cpu_budget)
    out = {"metric": "light curves/sec on 1 MI355X (batched engine, <=1e-5 rel err vs FP64 oracle)",
           "value": M * args.steps / el, "unit": "light curves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32 per face, fp64 geometry/accumulation",
           "data": "synthetic", "config": cfg, "instances_ok": nok, "roofline": roofline,
           "cpu_baseline": cpu, "max_rel_err": err}
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
