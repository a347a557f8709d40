// k_calib.hip -- measurement kernels for the roofline of the flux kernels.
//
// The flux kernels are bound by VALU instruction issue.  These kernels issue
// a known number of VALU instructions per lane with no memory traffic and no
// dependency stalls (eight independent chains per lane, eight waves per
// SIMD), so instructions / time is the issue ceiling the hardware reaches,
// measured on the same box and clock as the flux kernels:
//   kind 0: v_fma_f32 only                       (8 per iteration)
//   kind 1: v_fma_f32 + v_rcp_f32, 8 : 1          (transcendental co-issue)
//   kind 2: v_fma_f64 only                       (8 per iteration)
//   kind 3: v_fma_f32 + v_rcp_f32, 8 : 2
// Built into its own library (lcurve_re_amd/_lib/liblcurve_calib.so), used
// by bench.py --valu-calibration; the engine never loads it.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;

template <int KIND>
__global__ __launch_bounds__(kThreads) void k_valu_issue(int iters, float a, float b, double ad, double bd,
                                                         float *out, double *outd) {
    const int tid = blockIdx.x * kThreads + threadIdx.x;
    if constexpr (KIND == 2) {
        double x[8];
        for (int k = 0; k < 8; k++) x[k] = ad + 1e-3 * (tid + k);
        for (int i = 0; i < iters; i++) {
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = __builtin_fma(x[k], ad, bd);
        }
        double s = 0.;
        for (int k = 0; k < 8; k++) s += x[k];
        outd[tid] = s;
    } else {
        float x[8], y[2];
        for (int k = 0; k < 8; k++) x[k] = a + 1e-3f * (float)(tid + k);
        y[0] = 1.5f + 1e-3f * (float)tid;
        y[1] = 0.75f + 1e-3f * (float)tid;
        for (int i = 0; i < iters; i++) {
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = __builtin_fmaf(x[k], a, b);
            if constexpr (KIND == 1 || KIND == 3) y[0] = __builtin_amdgcn_rcpf(y[0]);
            if constexpr (KIND == 3) y[1] = __builtin_amdgcn_rcpf(y[1]);
        }
        float s = y[0] + y[1];
        for (int k = 0; k < 8; k++) s += x[k];
        out[tid] = s;
    }
}

}  // namespace

extern "C" {

// Runs kind `kind` (0, 1, 2, 3) on `blocks` workgroups of 256 lanes for `iters`
// iterations, `reps` times; returns the shortest launch in *ms and the VALU
// instructions per lane of the loop in *valu_per_lane (the loop's FMA/RCP
// count; set-up and the final sum add ~20).  0 on success, else the HIP
// error code.
int lcc_valu_issue(int kind, int blocks, int iters, int reps, double *ms, double *valu_per_lane) {
    if (kind < 0 || kind > 3 || blocks <= 0 || iters <= 0 || reps <= 0 || !ms || !valu_per_lane) return -1;
    const size_t n = (size_t)blocks * kThreads;
    float *out = nullptr;
    double *outd = nullptr;
    hipError_t e = hipMalloc((void **)&out, n * sizeof(float));
    if (e != hipSuccess) return (int)e;
    e = hipMalloc((void **)&outd, n * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(out);
        return (int)e;
    }
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    double best = 1e300;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        (void)hipEventRecord(t0, 0);
        if (kind == 0)
            k_valu_issue<0><<<blocks, kThreads>>>(iters, 0.999f, 1e-4f, 0.999, 1e-4, out, outd);
        else if (kind == 1)
            k_valu_issue<1><<<blocks, kThreads>>>(iters, 0.999f, 1e-4f, 0.999, 1e-4, out, outd);
        else if (kind == 3)
            k_valu_issue<3><<<blocks, kThreads>>>(iters, 0.999f, 1e-4f, 0.999, 1e-4, out, outd);
        else
            k_valu_issue<2><<<blocks, kThreads>>>(iters, 0.999f, 1e-4f, 0.999, 1e-4, out, outd);
        (void)hipEventRecord(t1, 0);
        e = hipEventSynchronize(t1);
        if (e == hipSuccess) e = hipGetLastError();
        float t = 0.f;
        (void)hipEventElapsedTime(&t, t0, t1);
        if (t < best) best = t;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(out);
    (void)hipFree(outd);
    *ms = best;
    *valu_per_lane = (double)iters * (kind == 1 ? 9.0 : kind == 3 ? 10.0 : 8.0);
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
