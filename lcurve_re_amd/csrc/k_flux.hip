// k_flux.hip -- per-phase limb-darkened, beamed flux sums
// (flat_sum_star{1,2}_multi, comp_light_flat.cpp:241-465).
//
// Work decomposition: a task is (batch, phase tile, face chunk).  Each lane
// owns R sub-phases of the tile and keeps their sums in registers while the
// workgroup streams the chunk's faces; face records are read with uniform
// (scalar) loads, so every element costs VALU work only.  Partial sums per
// (chunk, sub-phase) are merged in chunk order by k_flux_reduce, which
// keeps results independent of scheduling.
//
// Two element precisions:
//   mixed: FP32 per-face arithmetic, FP32 partial sums over 32 faces, FP64
//          accumulation of those partials;
//   fp64 : the reference element arithmetic in FP64 with 1024-face chunks,
//          faces ascending, chunks merged in order -- the reduction order of
//          comp_light_flat.cpp:254-307, so sums are bit-identical to it.
// Both skip whole faces that are provably invisible for every sub-phase of
// the tile (mu <= mucrit) and eclipse tests that cannot trigger inside the
// tile's phase range; skipped elements are exact zeros in the reference.
#include "lc_kernels.h"

namespace lck {
namespace {

constexpr int NT = 256;  // threads per workgroup (4 waves)

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

struct TileInfo {
    float ecx, ecy, ecz, rad;  // centre earth vector and max |e - e_c| over the tile
    double pmin, pmax;         // phin range of the tile
};

// Tile statistics for the visibility and eclipse shortcuts.
template <int R>
__device__ __forceinline__ TileInfo tile_info(const double *ex, const double *ey, const double *ez, const double *ph,
                                              const bool *ok, double cx, double cy, double cz) {
    __shared__ float s_rad[NT / 64];
    __shared__ double s_min[NT / 64], s_max[NT / 64];
    float rad = 0.f;
    double pmin = 2.0, pmax = -1.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (ok[r]) {
            double a = ex[r] - cx, b = ey[r] - cy, c = ez[r] - cz;
            rad = fmaxf(rad, (float)sqrt(a * a + b * b + c * c));
            pmin = fmin(pmin, ph[r]);
            pmax = fmax(pmax, ph[r]);
        }
    }
    rad = wave_max(rad);
    pmin = wave_min_d(pmin);
    pmax = wave_max_d(pmax);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_rad[w] = rad;
        s_min[w] = pmin;
        s_max[w] = pmax;
    }
    __syncthreads();
    TileInfo t;
    t.rad = s_rad[0];
    t.pmin = s_min[0];
    t.pmax = s_max[0];
#pragma unroll
    for (int k = 1; k < NT / 64; k++) {
        t.rad = fmaxf(t.rad, s_rad[k]);
        t.pmin = fmin(t.pmin, s_min[k]);
        t.pmax = fmax(t.pmax, s_max[k]);
    }
    t.ecx = (float)cx;
    t.ecy = (float)cy;
    t.ecz = (float)cz;
    return t;
}

// Loads the R sub-phases of this lane and the tile centre.
template <int R>
__device__ __forceinline__ void load_phases(const BatchDev &B, int S, int count, int tile0, const int *lists,
                                            const PhaseArrays &pa, double *ex, double *ey, double *ez, double *ph,
                                            bool *ok, double &cx, double &cy, double &cz) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        int e = tile0 + r * NT + (int)threadIdx.x;
        ok[r] = e < count;
        int j = ok[r] ? (B.list_off < 0 ? e : lists[B.list_off + e]) : 0;
        size_t idx = (size_t)B.inst * S + j;
        ex[r] = ok[r] ? pa.ex[idx] : 0.;
        ey[r] = ok[r] ? pa.ey[idx] : 0.;
        ez[r] = ok[r] ? pa.ez[idx] : 0.;
        ph[r] = ok[r] ? pa.phin[idx] : 0.5;
    }
    int live = min(count - tile0, R * NT);
    int em = tile0 + (live - 1) / 2;
    int jm = B.list_off < 0 ? em : lists[B.list_off + em];
    size_t im = (size_t)B.inst * S + jm;
    cx = pa.ex[im];
    cy = pa.ey[im];
    cz = pa.ez[im];
}

// eclipse test of one range (comp_light_flat.cpp:171-173)
__device__ __forceinline__ bool ecl_range(double phin, double in, double out) {
    return (phin >= in && phin <= out) || phin <= out - 1.0;
}

// 0: no sub-phase of the tile can be eclipsed, 1: some can, 2: all are.
__device__ __forceinline__ int tile_eclipse_class(const TileInfo &t, double in, double out) {
    bool any = (t.pmax >= in && t.pmin <= out) || (t.pmin <= out - 1.0);
    if (!any) return 0;
    bool all = (t.pmin >= in && t.pmax <= out) || (t.pmax <= out - 1.0);
    return all ? 2 : 1;
}

template <int LT>
__device__ __forceinline__ float imu_f(float m, float c1, float c2, float c3, float c4, float base) {
    // m already clipped to <= 1; caller masks m <= 0
    if (LT == 0) {
        float om = 1.f - m;
        return 1.f - om * (c1 + om * (c2 + om * (c3 + om * c4)));
    }
    float s = sqrtf(fmaxf(m, 0.f));
    return base + s * (c1 + s * (c2 + s * (c3 + s * c4)));
}

template <int LT, bool BEAM, bool GLENS, int R>
__global__ __launch_bounds__(NT) void k_flux_mixed(const TaskDev *__restrict__ tasks,
                                                   const BatchDev *__restrict__ batches,
                                                   const GridDev *__restrict__ grids, FlatArrays fl, PhaseArrays pa,
                                                   int S, const int *__restrict__ lists,
                                                   const int *__restrict__ counts, double *__restrict__ partial) {
    const TaskDev T = tasks[blockIdx.x];
    const BatchDev B = batches[T.batch];
    const GridDev &G = grids[B.grid];
    const int count = B.count_idx < 0 ? S : counts[B.count_idx];
    const int TILE = R * NT;
    const int tile0 = T.tile * TILE;
    if (tile0 >= count) return;

    double dex[R], dey[R], dez[R], ph[R];
    bool ok[R];
    double cx, cy, cz;
    load_phases<R>(B, S, count, tile0, lists, pa, dex, dey, dez, ph, ok, cx, cy, cz);
    const TileInfo ti = tile_info<R>(dex, dey, dez, ph, ok, cx, cy, cz);
    float ex[R], ey[R], ez[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        ex[r] = (float)dex[r];
        ey[r] = (float)dey[r];
        ez[r] = (float)dez[r];
    }

    const float c1 = (float)G.ldc[0], c2 = (float)G.ldc[1], c3 = (float)G.ldc[2], c4 = (float)G.ldc[3];
    const float base = (float)(1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]));
    const float mucrit = (float)G.mucrit;
    const float beam = (float)G.beam;
    const float rlens = (float)G.rlens1;
    const float cull = (float)G.mucrit - 1.e-5f - ti.rad;

    const int n0 = fl.n0[B.grid];
    const int f_lo = T.chunk * B.chunk;
    const int f_hi = min(G.nface, f_lo + B.chunk);
    const float4 *__restrict__ rec0 = fl.rec0 + G.face_off;
    const float4 *__restrict__ rec1 = fl.rec1 + G.face_off;
    const float4 *__restrict__ rec2 = fl.rec2 + G.face_off;
    const double2 *__restrict__ eclr = fl.ecl + G.face_off;

    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.;

    for (int fb = f_lo; fb < f_hi; fb += 32) {
        float a32[R];
#pragma unroll
        for (int r = 0; r < R; r++) a32[r] = 0.f;
        const int fe = min(fb + 32, f_hi);
        for (int f = fb; f < fe; f++) {
            const float4 a = rec0[f];
            // every sub-phase of the tile sees mu <= mucrit: exact zero
            if (fmaf(a.z, ti.ecz, fmaf(a.y, ti.ecy, a.x * ti.ecx)) <= cull) continue;
            int ecls = 0;
            double ein = 0., eout = 0.;
            if (f >= n0) {
                const double2 io = eclr[f];
                ein = io.x;
                eout = io.y;
                ecls = tile_eclipse_class(ti, ein, eout);
                if (ecls == 2) continue;
            }
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f), p = make_float4(0.f, 0.f, 0.f, 0.f);
            if (BEAM) v = rec1[f];
            if (GLENS) p = rec2[f];
#pragma unroll
            for (int r = 0; r < R; r++) {
                float mu = fmaf(ez[r], a.z, fmaf(ey[r], a.y, ex[r] * a.x));
                float magn = 1.f;
                if (GLENS) {
                    float d = -(p.x * ex[r] + p.y * ey[r] + p.z * ez[r]);
                    float qx = p.x + d * ex[r], qy = p.y + d * ey[r], qz = p.z + d * ez[r];
                    float pp = sqrtf(qx * qx + qy * qy + qz * qz);
                    float hp = 0.5f * pp;
                    float rd = rlens * d;
                    float pd = (hp * hp > 25.f * rd) ? pp + rd / pp : hp + sqrtf(hp * hp + rd);
                    float mg = pd * pd / (pd - hp) / hp / 4.f;
                    magn = d > 0.f ? mg : 1.f;
                }
                float val;
                if (!BEAM) {
                    float m = fminf(mu, 1.f);
                    float im = imu_f<LT>(m, c1, c2, c3, c4, base);
                    val = (mu > 0.f) ? mu * magn * a.w * im : 0.f;
                } else {
                    float vr = -(ex[r] * v.x + ey[r] * v.y);
                    float mud = mu - mu * vr - v.z;
                    float m = fminf(mud, 1.f);
                    float im = imu_f<LT>(m, c1, c2, c3, c4, base);
                    val = (mud > 0.f) ? mu * magn * a.w * (1.f - beam * vr) * im : 0.f;
                }
                bool vis = mu > mucrit;
                if (ecls == 1) vis = vis && !ecl_range(ph[r], ein, eout);
                a32[r] += vis ? val : 0.f;
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] += (double)a32[r];
    }

    double *out = partial + B.part_off + (size_t)T.chunk * (size_t)(B.ntile * TILE);
#pragma unroll
    for (int r = 0; r < R; r++) {
        int e = tile0 + r * NT + (int)threadIdx.x;
        if (ok[r]) out[e] = acc[r];
    }
}

// FP64 element in the reference's operand order (comp_light_flat.cpp:81-168)
template <int LT, bool BEAM, bool GLENS>
__device__ __forceinline__ double elem64(double ex, double ey, double ez, double dx, double dy, double dz, double px,
                                         double py, double pz, double flux, const GridDev &G, double base) {
    double mu = ex * dx + ey * dy + ez * dz;
    if (!(mu > G.mucrit)) return 0.0;
    double magn = 1.0;
    if (GLENS) {
        double d = -(px * ex + py * ey + pz * ez);
        double qx = px + d * ex, qy = py + d * ey, qz = pz + d * ez;
        double p = sqrt(qx * qx + qy * qy + qz * qz);
        double ph = 0.5 * p;
        double rd = G.rlens1 * d;
        double pd = (ph * ph > 25.0 * rd) ? p + rd / p : ph + sqrt(ph * ph + rd);
        double mg = pd * pd / (pd - ph) / ph / 4.0;
        magn = d > 0.0 ? mg : 1.0;
    }
    const double c1 = G.ldc[0], c2 = G.ldc[1], c3 = G.ldc[2], c4 = G.ldc[3];
    double arg, lead;
    if (!BEAM) {
        arg = mu;
        lead = GLENS ? mu * magn * flux : mu * flux;
    } else {
        double vx = -G.VFAC * G.spin_beam * py;
        double vy = G.star == 0 ? G.VFAC * (G.spin_beam * px - G.XCOFM)
                                : G.VFAC * (G.spin_beam * (px - 1.0) + 1.0 - G.XCOFM);
        double vr = -(ex * vx + ey * vy);
        double vn = dx * vx + dy * vy;
        arg = mu - mu * vr - vn;
        lead = GLENS ? mu * magn * flux * (1.0 - G.beam * vr) : mu * flux * (1.0 - G.beam * vr);
    }
    if (!(arg > 0.0)) return 0.0;
    double m = arg < 1.0 ? arg : 1.0;
    double im;
    if (LT == 0) {
        double om = 1.0 - m;
        im = 1.0 - om * (c1 + om * (c2 + om * (c3 + om * c4)));
    } else {
        double msq = sqrt(m > 0.0 ? m : 0.0);
        im = base + msq * (c1 + msq * (c2 + msq * (c3 + msq * c4)));
    }
    return lead * im;
}

template <int LT, bool BEAM, bool GLENS, int R>
__global__ __launch_bounds__(NT) void k_flux_fp64(const TaskDev *__restrict__ tasks,
                                                  const BatchDev *__restrict__ batches,
                                                  const GridDev *__restrict__ grids, FlatArrays fl, PhaseArrays pa,
                                                  int S, const int *__restrict__ lists,
                                                  const int *__restrict__ counts, double *__restrict__ partial) {
    const TaskDev T = tasks[blockIdx.x];
    const BatchDev B = batches[T.batch];
    const GridDev &G = grids[B.grid];
    const int count = B.count_idx < 0 ? S : counts[B.count_idx];
    const int TILE = R * NT;
    const int tile0 = T.tile * TILE;
    if (tile0 >= count) return;

    double ex[R], ey[R], ez[R], ph[R];
    bool ok[R];
    double cx, cy, cz;
    load_phases<R>(B, S, count, tile0, lists, pa, ex, ey, ez, ph, ok, cx, cy, cz);
    const TileInfo ti = tile_info<R>(ex, ey, ez, ph, ok, cx, cy, cz);
    const double cull = G.mucrit - 1.e-5 - (double)ti.rad;
    const double base = 1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]);

    const int n0 = fl.n0[B.grid];
    const int f_lo = T.chunk * B.chunk;
    const int f_hi = min(G.nface, f_lo + B.chunk);
    const size_t o = (size_t)G.face_off;

    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0;
    for (int f = f_lo; f < f_hi; f++) {
        const double dx = fl.fdx[o + f], dy = fl.fdy[o + f], dz = fl.fdz[o + f];
        if (dx * cx + dy * cy + dz * cz <= cull) continue;
        int ecls = 0;
        double ein = 0., eout = 0.;
        if (f >= n0) {
            const double2 io = fl.ecl[o + f];
            ein = io.x;
            eout = io.y;
            ecls = tile_eclipse_class(ti, ein, eout);
            if (ecls == 2) continue;
        }
        const double px = fl.fpx[o + f], py = fl.fpy[o + f], pz = fl.fpz[o + f], flux = fl.fflux[o + f];
#pragma unroll
        for (int r = 0; r < R; r++) {
            double v = elem64<LT, BEAM, GLENS>(ex[r], ey[r], ez[r], dx, dy, dz, px, py, pz, flux, G, base);
            if (ecls == 1 && ecl_range(ph[r], ein, eout)) v = 0.0;
            if (v != 0.0) acc[r] += v;
        }
    }
    double *out = partial + B.part_off + (size_t)T.chunk * (size_t)(B.ntile * TILE);
#pragma unroll
    for (int r = 0; r < R; r++) {
        int e = tile0 + r * NT + (int)threadIdx.x;
        if (ok[r]) out[e] = acc[r];
    }
}

// Merge chunk partials in chunk order; out[(inst*2 + star)*S + j].
__global__ void k_flux_reduce(const BatchDev *__restrict__ batches, int S, int tile_len, const int *__restrict__ lists,
                              const int *__restrict__ counts, const double *__restrict__ partial,
                              double *__restrict__ out) {
    const BatchDev B = batches[blockIdx.y];
    const int count = B.count_idx < 0 ? S : counts[B.count_idx];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const size_t stride = (size_t)B.ntile * tile_len;
    const double *p = partial + B.part_off + e;
    double s = 0.0;
    for (int c = 0; c < B.nchunk; c++) s += p[(size_t)c * stride];
    const int j = B.list_off < 0 ? e : lists[B.list_off + e];
    out[((size_t)B.inst * 2 + B.star) * S + j] = s;
}

template <int R>
void launch_variant(hipStream_t s, int variant, int fp64, const TaskDev *tasks, int ntask, const BatchDev *batches,
                    const GridDev *grids, FlatArrays fl, PhaseArrays pa, int S, const int *lists, const int *counts,
                    double *partial) {
#define LC_FLUX_CASE(V, LT, BM, GL)                                                                                 \
    case V:                                                                                                         \
        if (fp64)                                                                                                   \
            hipLaunchKernelGGL((k_flux_fp64<LT, BM, GL, R>), dim3(ntask), dim3(NT), 0, s, tasks, batches, grids,   \
                               fl, pa, S, lists, counts, partial);                                                 \
        else                                                                                                        \
            hipLaunchKernelGGL((k_flux_mixed<LT, BM, GL, R>), dim3(ntask), dim3(NT), 0, s, tasks, batches, grids,  \
                               fl, pa, S, lists, counts, partial);                                                 \
        break;
    switch (variant) {
        LC_FLUX_CASE(0, 0, false, false)
        LC_FLUX_CASE(1, 1, false, false)
        LC_FLUX_CASE(2, 0, true, false)
        LC_FLUX_CASE(3, 1, true, false)
        LC_FLUX_CASE(4, 0, false, true)
        LC_FLUX_CASE(5, 1, false, true)
        LC_FLUX_CASE(6, 0, true, true)
        LC_FLUX_CASE(7, 1, true, true)
    default: break;
    }
#undef LC_FLUX_CASE
}

}  // namespace

// variant = ltype | beam << 1 | glens << 2; rphase in {1, 2, 4}
void launch_flux(hipStream_t s, int variant, int fp64, int rphase, const TaskDev *tasks, int ntask,
                 const BatchDev *batches, const GridDev *grids, FlatArrays fl, PhaseArrays pa, int S,
                 const int *lists, const int *counts, double *partial) {
    if (ntask <= 0) return;
    if (rphase == 1)
        launch_variant<1>(s, variant, fp64, tasks, ntask, batches, grids, fl, pa, S, lists, counts, partial);
    else if (rphase == 4)
        launch_variant<4>(s, variant, fp64, tasks, ntask, batches, grids, fl, pa, S, lists, counts, partial);
    else
        launch_variant<2>(s, variant, fp64, tasks, ntask, batches, grids, fl, pa, S, lists, counts, partial);
}

void launch_flux_reduce(hipStream_t s, const BatchDev *batches, int nbatch, int S, int tile_len, const int *lists,
                        const int *counts, const double *partial, double *out) {
    if (nbatch <= 0 || S <= 0) return;
    hipLaunchKernelGGL(k_flux_reduce, dim3((S + 255) / 256, nbatch), dim3(256), 0, s, batches, S, tile_len, lists,
                       counts, partial, out);
}

}  // namespace lck
