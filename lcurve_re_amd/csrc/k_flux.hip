// k_flux.hip -- per-phase limb-darkened, beamed flux sums
// (flat_sum_star{1,2}_multi, comp_light_flat.cpp:241-465).
//
// Work decomposition: a task is (batch, phase tile, face chunk).  Each lane
// owns R sub-phases of the tile and keeps their sums in registers.  The
// workgroup walks its face chunk in stages of 256 faces: every thread loads
// one face record (coalesced, prefetched one stage ahead), decides once for
// the whole tile whether the face can contribute, and the contributing
// faces are compacted -- in ascending face order -- into LDS.  All waves
// then sweep the compacted records with broadcast LDS reads, so the inner
// loop is pure VALU work on registers.  Partial sums per (chunk, sub-phase)
// are merged in chunk order by k_flux_reduce, independent of scheduling.
//
// A face is skipped for a tile when it is provably invisible for every
// sub-phase (mu <= mucrit: with e_c the tile's centre earth vector and R
// the largest |e - e_c|, mu(e) <= d.e_c + R) or eclipsed at every sub-phase
// of the tile; skipped elements are exact zeros in the reference sums.
//
// Two element precisions:
//   mixed: FP32 per-face arithmetic, FP32 partial sums over at most 32
//          faces, FP64 accumulation of those partials;
//   fp64 : the reference element arithmetic in FP64 with 1024-face chunks,
//          faces ascending, chunks merged in order -- the reduction order of
//          comp_light_flat.cpp:254-307, so the sums are bit-identical.
#include "lc_kernels.h"

namespace lck {
namespace {

constexpr int NT = 256;  // threads per workgroup (4 waves of 64)
constexpr int FB = 256;  // faces per stage

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
    double cx, cy, cz;  // centre earth vector
    float rad;          // max |e - e_c| over the tile
    double pmin, pmax;  // phin range of the tile
    bool zflat;         // every sub-phase has e_z == cz exactly (one inclination)
};

// Tile statistics for the visibility and eclipse shortcuts.
template <int R>
__device__ __forceinline__ TileInfo tile_info(const double *ex, const double *ey, const double *ez, const double *ph,
                                              const bool *ok, double cx, double cy, double cz) {
    __shared__ float s_rad[NT / 64];
    __shared__ double s_min[NT / 64], s_max[NT / 64];
    __shared__ int s_zd[NT / 64];
    float rad = 0.f;
    double pmin = 2.0, pmax = -1.0;
    bool zd = false;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (ok[r]) {
            double a = ex[r] - cx, b = ey[r] - cy, c = ez[r] - cz;
            rad = fmaxf(rad, (float)sqrt(a * a + b * b + c * c));
            pmin = fmin(pmin, ph[r]);
            pmax = fmax(pmax, ph[r]);
            zd = zd || ez[r] != cz;
        }
    }
    rad = wave_max(rad);
    pmin = wave_min_d(pmin);
    pmax = wave_max_d(pmax);
    const bool wzd = __ballot(zd) != 0ull;
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_rad[w] = rad;
        s_min[w] = pmin;
        s_max[w] = pmax;
        s_zd[w] = wzd ? 1 : 0;
    }
    __syncthreads();
    TileInfo t;
    t.rad = s_rad[0];
    t.pmin = s_min[0];
    t.pmax = s_max[0];
    int zdiff = s_zd[0];
#pragma unroll
    for (int k = 1; k < NT / 64; k++) {
        t.rad = fmaxf(t.rad, s_rad[k]);
        t.pmin = fmin(t.pmin, s_min[k]);
        t.pmax = fmax(t.pmax, s_max[k]);
        zdiff |= s_zd[k];
    }
    t.zflat = zdiff == 0;
    t.cx = cx;
    t.cy = cy;
    t.cz = cz;
    __syncthreads();
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

// The per-tile decision for one face record (both kernels of the Claret
// law make it with this function, so they agree face by face): act = it
// can contribute to some sub-phase (d.e_c above cull), cls = its eclipse
// class for the tile, full = seen at every sub-phase and not eclipsed in
// the tile (d.e_c above full_cut), to the moment sums.
struct FaceDecision {
    bool act, full;
    int cls;
    float dc;
};
__device__ __forceinline__ FaceDecision decide_face(const float4 &a0, const double2 &io, bool in, bool ecl,
                                                    float ecx, float ecy, float ecz, float cull, float full_cut,
                                                    const TileInfo &ti, bool mom) {
    FaceDecision r;
    r.dc = fmaf(a0.z, ecz, fmaf(a0.y, ecy, a0.x * ecx));
    r.act = in && r.dc > cull;
    r.cls = 0;
    if (r.act && ecl) {
        r.cls = tile_eclipse_class(ti, io.x, io.y);
        r.act = r.cls != 2;
    }
    r.full = mom && r.act && r.cls == 0 && r.dc > full_cut;
    return r;
}

// Full faces of a Claret-law tile that go to the Taylor moments: the tile has
// one inclination (e - e_c has no z part) and |d.(e - e_c)| <= rad |d_xy|
// stays below TAY_RHO mu_c, where the Taylor series of mu^(3/2) and
// mu^(5/2) truncated after degree 9 is within 1.3e-8 of the terms.
constexpr float TAY_RHO = 0.3f;
__device__ __forceinline__ bool taylor_face(const float4 &a0, float dc, const TileInfo &ti) {
    const float dxy2 = fmaf(a0.y, a0.y, a0.x * a0.x);
    return ti.zflat && ti.rad * ti.rad * dxy2 <= (TAY_RHO * TAY_RHO) * (dc * dc);
}

// Order-preserving compaction of the active faces of a stage: returns the
// slot of this thread's face (if active) and the stage's active count.
__device__ __forceinline__ int compact_slot(bool act, int &total, int *s_cnt) {
    const unsigned long long m = __ballot(act);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_cnt[w] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; k++) {
        base += k < w ? s_cnt[k] : 0;
        tot += s_cnt[k];
    }
    total = tot;
    return base + pre;
}

struct Coef32 {
    float c1, c2, c3, c4, base, mucrit, beam, rlens;
    // POLY: mu * I(mu) = mu (q1 + mu (q2 + mu (q3 + mu (q4 + mu q5))))
    float q1, q2, q3, q4, q5;
};

// FP32 element without the face flux: returns the factor that multiplies
// the face's brightness*area (comp_light_flat.cpp:110-168).  mu is clipped
// to [0, 1]; for mucrit <= 0 the clip alone reproduces the reference's
// masks (mu <= 0 gives zero), otherwise (MUC) mu > mucrit is tested too.
// Square roots and reciprocals use the hardware instructions (v_sqrt_f32,
// v_rcp_f32: about 1 ulp) instead of the correctly rounded expansions,
// which cost ~15 extra VALU operations each; the element error stays far
// inside the 1e-5 budget of this mode.
template <int LT, bool BEAM, bool GLENS, bool MUC>
__device__ __forceinline__ float elem32(float ex, float ey, float ez, const float4 &a, const float4 &v,
                                        const float4 &p, const Coef32 &K) {
#pragma clang fp contract(fast)
    const float mu = fmaf(ez, a.z, fmaf(ey, a.y, ex * a.x));
    float val;
    if (!BEAM) {
        const float m = __builtin_amdgcn_fmed3f(mu, 0.f, 1.f);
        float im;
        if (LT == 0) {
            // one Horner chain for mu I(mu): the polynomial vanishes at 0, so
            // the clip to [0, 1] also zeroes faces turned away
            val = m * fmaf(m, fmaf(m, fmaf(m, fmaf(m, K.q5, K.q4), K.q3), K.q2), K.q1);
        } else {
            const float s = __builtin_amdgcn_sqrtf(m);
            im = fmaf(s, fmaf(s, fmaf(s, fmaf(s, K.c4, K.c3), K.c2), K.c1), K.base);
            val = m * im;
        }
        if (MUC) val = mu > K.mucrit ? val : 0.f;
    } else {
        // v = (vx, vy, vn): vr_ref = -(ex vx + ey vy)
        const float vr = fmaf(ey, v.y, ex * v.x);
        const float mud = fmaf(mu, vr, mu) - v.z;
        const float md = __builtin_amdgcn_fmed3f(mud, 0.f, 1.f);
        float im;
        if (LT == 0) {
            const float om = 1.f - md;
            im = fmaf(-om, fmaf(om, fmaf(om, fmaf(om, K.c4, K.c3), K.c2), K.c1), 1.f);
        } else {
            const float s = __builtin_amdgcn_sqrtf(md);
            im = fmaf(s, fmaf(s, fmaf(s, fmaf(s, K.c4, K.c3), K.c2), K.c1), K.base);
        }
        const float g = fmaf(K.beam, vr, 1.f);
        val = (mud > 0.f && mu > K.mucrit) ? mu * g * im : 0.f;
    }
    if (GLENS) {
        const float d = -(p.x * ex + p.y * ey + p.z * ez);
        const float qx = fmaf(d, ex, p.x), qy = fmaf(d, ey, p.y), qz = fmaf(d, ez, p.z);
        const float pp = __builtin_amdgcn_sqrtf(qx * qx + qy * qy + qz * qz);
        const float hp = 0.5f * pp;
        const float rd = K.rlens * d;
        const float pd = (hp * hp > 25.f * rd) ? pp + rd * __builtin_amdgcn_rcpf(pp)
                                               : hp + __builtin_amdgcn_sqrtf(hp * hp + rd);
        const float mg = 0.25f * pd * pd * __builtin_amdgcn_rcpf((pd - hp) * hp);
        val *= d > 0.f ? mg : 1.f;
    }
    return val;
}

// One compacted face record against this lane's R sub-phases.
template <int LT, bool BEAM, bool GLENS, bool MUC, bool ECL, int R>
__device__ __forceinline__ void face_elems(const float4 *s_a0, const float4 *s_a1, const float4 *s_a2,
                                           const double2 *s_io, int k, const float *ex, const float *ey,
                                           const float *ez, const double *ph, const Coef32 &K, float *a32) {
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 a = s_a0[k];
    const float4 v = BEAM ? s_a1[k] : z4;
    const float4 p = GLENS ? s_a2[k] : z4;
    if (ECL) {
        const double2 e = s_io[k];
#pragma unroll
        for (int r = 0; r < R; r++) {
            float val = elem32<LT, BEAM, GLENS, MUC>(ex[r], ey[r], ez[r], a, v, p, K);
            val = ecl_range(ph[r], e.x, e.y) ? 0.f : val;
            a32[r] = fmaf(a.w, val, a32[r]);
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++)
            a32[r] = fmaf(a.w, elem32<LT, BEAM, GLENS, MUC>(ex[r], ey[r], ez[r], a, v, p, K), a32[r]);
    }
}

// Sweeps n compacted records; FP32 partials over 32 faces are added into
// the FP64 accumulators.  n is uniform across the workgroup (it comes from
// the LDS counters), so the loop bounds live in scalar registers and the
// face loop is unrolled by 4 to batch the LDS reads.
template <int LT, bool BEAM, bool GLENS, bool MUC, bool ECL, int R, int UNR = 4>
__device__ __forceinline__ void sweep(const float4 *s_a0, const float4 *s_a1, const float4 *s_a2, const double2 *s_io,
                                      int n_, const float *ex, const float *ey, const float *ez, const double *ph,
                                      const Coef32 &K, double *acc) {
    const int n = __builtin_amdgcn_readfirstlane(n_);
    for (int k0 = 0; k0 < n; k0 += 32) {
        float a32[R];
#pragma unroll
        for (int r = 0; r < R; r++) a32[r] = 0.f;
        const int k1 = min(k0 + 32, n);
        int k = k0;
        for (; k + UNR <= k1; k += UNR) {
#pragma unroll
            for (int u = 0; u < UNR; u++)
                face_elems<LT, BEAM, GLENS, MUC, ECL, R>(s_a0, s_a1, s_a2, s_io, k + u, ex, ey, ez, ph, K, a32);
        }
        for (; k < k1; k++)
            face_elems<LT, BEAM, GLENS, MUC, ECL, R>(s_a0, s_a1, s_a2, s_io, k, ex, ey, ez, ph, K, a32);
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] += (double)a32[r];
    }
}

// Polynomial limb darkening without beaming or lensing: for faces seen (mu
// above 0 and mucrit) and not eclipsed at every sub-phase of a tile,
// mu I(mu) = Q(d.e) is a polynomial of degree 5 in e (Q has no constant
// term), so their sum is
//   sum_i f_i Q(d_i.e) = sum_alpha q_|alpha| (|alpha|! / alpha!) M_alpha e^alpha,
//   M_alpha = sum_i f_i d_i^alpha,
// over the 55 monomials of degree 1..5.  The moments M are FP64 sums over
// the face chunk (each wave owns the monomials k = wave + 4t); the tile's
// sub-phases then evaluate the polynomial in FP64.  Terms reach ~2e4 times
// the result, so FP64 keeps the evaluation error near 1e-11 relative.  The
// Claret law, with half-integer powers of mu, goes through the Taylor
// moments below instead.
constexpr int NMOM = 55, MOM_PER_WAVE = 14;
constexpr unsigned char MOM_A[NMOM] = {1,0,0,2,1,1,0,0,0,3,2,2,1,1,1,0,0,0,0,4,3,3,2,2,2,1,1,1,1,0,0,0,0,0,
                                        5,4,4,3,3,3,2,2,2,2,1,1,1,1,1,0,0,0,0,0,0};
constexpr unsigned char MOM_B[NMOM] = {0,1,0,0,1,0,2,1,0,0,1,0,2,1,0,3,2,1,0,0,1,0,2,1,0,3,2,1,0,4,3,2,1,0,
                                        0,1,0,2,1,0,3,2,1,0,4,3,2,1,0,5,4,3,2,1,0};
constexpr unsigned char MOM_C[NMOM] = {0,0,1,0,0,1,0,1,2,0,0,1,0,1,2,0,1,2,3,0,0,1,0,1,2,0,1,2,3,0,1,2,3,4,
                                        0,0,1,0,1,2,0,1,2,3,0,1,2,3,4,0,1,2,3,4,5};
__constant__ unsigned char c_mom_n[NMOM] = {1,1,1,2,2,2,2,2,2,3,3,3,3,3,3,3,3,3,3,4,4,4,4,4,4,4,4,4,4,4,4,4,4,4,
                                            5,5,5,5,5,5,5,5,5,5,5,5,5,5,5,5,5,5,5,5,5};
__constant__ double c_mom_mult[NMOM] = {1.,1.,1.,1.,2.,2.,1.,2.,1.,1.,3.,3.,3.,6.,3.,1.,3.,3.,1.,1.,4.,4.,6.,
                                        12.,6.,4.,12.,12.,4.,1.,4.,6.,4.,1.,1.,5.,5.,10.,20.,10.,10.,30.,30.,
                                        10.,5.,20.,30.,20.,5.,1.,5.,10.,10.,5.,1.};

// m[t] += f x^a y^b z^c for monomial k = W + 4t of record (x, y, z, f)
template <int W, int NM>
__device__ __forceinline__ void mom_add(const float4 &rec, double (&m)[MOM_PER_WAVE]) {
    const double x = rec.x, y = rec.y, z = rec.z;
    double xp[6], yp[6], zp[6];
    xp[0] = rec.w;
    yp[0] = 1.0;
    zp[0] = 1.0;
#pragma unroll
    for (int i = 1; i < 6; i++) {
        xp[i] = xp[i - 1] * x;
        yp[i] = yp[i - 1] * y;
        zp[i] = zp[i - 1] * z;
    }
#pragma unroll
    for (int t = 0; t < MOM_PER_WAVE; t++) {
        if (W + 4 * t < NM)
            m[t] = fma(xp[MOM_A[W + 4 * t]], yp[MOM_B[W + 4 * t]] * zp[MOM_C[W + 4 * t]], m[t]);
    }
}
template <int W, int NM>
__device__ __forceinline__ void mom_sweep(const float4 *s_m, int n, double (&m)[MOM_PER_WAVE]) {
    for (int j = (int)(threadIdx.x & 63); j < n; j += 64) mom_add<W, NM>(s_m[j], m);
}


// Claret law without beaming or lensing, the full faces that taylor_face
// selects.  With delta = d.(e - e_c) = d_x dX + d_y dY (the tile's
// sub-phases share e_z), mu I(mu) = H(mu) = base mu + c1 mu^(3/2) + c2 mu^2 +
// c3 mu^(5/2) + c4 mu^3 is expanded about mu_c = d.e_c:
//   H(mu_c + delta) = sum_k T_k(mu_c) delta^k,  T_k = H^(k)(mu_c) / k!,
// exact in its integer powers and truncated after k = 9 in the half-integer
// ones (|delta| <= 0.3 mu_c: relative error below 1.3e-8 of those terms).
// Summed over the faces,
//   sum_i f_i H(mu_i) = sum_{k,j} C(k,j) M_kj dX^(k-j) dY^j,
//   M_kj = sum_i f_i T_k(mu_ci) d_xi^(k-j) d_yi^j,
// 55 moments in FP64 (wave w owns monomials w + 4t) over the chunk's
// selected faces, reduced over the lanes and evaluated in FP64 at every
// sub-phase of the tile.  The other full faces of the tile (|delta| above
// 0.3 mu_c, or several inclinations in the tile) are summed face by face.
// Tried: a separate kernel for these moments (its per-task set-up and
// reductions were latency bound: 6.8 against 6.7 ms per batched step for
// the two kernels together).
constexpr int TAY_K = 9, NTAY = 55, TAY_PER_WAVE = 14;
constexpr unsigned char TAY_KK[NTAY] = {0,1,1,2,2,2,3,3,3,3,4,4,4,4,4,5,5,5,5,5,5,6,6,6,6,6,6,6,7,7,7,7,7,7,7,7,
                                        8,8,8,8,8,8,8,8,8,9,9,9,9,9,9,9,9,9,9};
constexpr unsigned char TAY_JJ[NTAY] = {0,0,1,0,1,2,0,1,2,3,0,1,2,3,4,0,1,2,3,4,5,0,1,2,3,4,5,6,0,1,2,3,4,5,6,7,
                                        0,1,2,3,4,5,6,7,8,0,1,2,3,4,5,6,7,8,9};
__constant__ double c_tay_binom[NTAY] = {1.,1.,1.,1.,2.,1.,1.,3.,3.,1.,1.,4.,6.,4.,1.,1.,5.,10.,10.,5.,1.,1.,6.,
                                         15.,20.,15.,6.,1.,1.,7.,21.,35.,35.,21.,7.,1.,1.,8.,28.,56.,70.,56.,28.,
                                         8.,1.,1.,9.,36.,84.,126.,126.,84.,36.,9.,1.};

// m[t] += f T_k(mu_c) d_x^(k-j) d_y^j for this wave's monomials (k, j) =
// (TAY_KK, TAY_JJ)[W + 4t], one face, k ascending: the mu^(3/2) and
// mu^(5/2) terms by the binomial recurrence
//   C(p, k+1) mu^(p-k-1) = C(p, k) mu^(p-k) (p - k) / ((k + 1) mu),
// the integer powers' coefficients I_k added for k <= 3.  1/sqrt(mu_c)
// from the hardware estimate and one Newton step (relative error ~1e-14).
template <int W>
__device__ __forceinline__ void tay_add(const float4 &rec, double cx, double cy, double cz, const double *cl,
                                        double (&m)[TAY_PER_WAVE]) {
    const double x = rec.x, y = rec.y, z = rec.z, f = rec.w;
    const double mu = fma(z, cz, fma(y, cy, x * cx));
    double q = __builtin_amdgcn_rsq(mu);
    q = q * fma(-0.5 * mu, q * q, 1.5);
    const double r = q * q;
    const double base = cl[4], c1 = cl[0], c2 = cl[1], c3 = cl[2], c4 = cl[3];
    const double fm15 = f * mu * (mu * q);  // f mu^(3/2)
    double p15 = c1 * fm15, p25 = c3 * (fm15 * mu);
    double xp[TAY_K + 1], yp[TAY_K + 1];
    xp[0] = yp[0] = 1.0;
#pragma unroll
    for (int i = 1; i <= TAY_K; i++) {
        xp[i] = xp[i - 1] * x;
        yp[i] = yp[i - 1] * y;
    }
#pragma unroll
    for (int k = 0; k <= TAY_K; k++) {
        double T = p15 + p25;
        if (k == 0) T = fma(f * mu, fma(mu, fma(mu, c4, c2), base), T);
        if (k == 1) T = fma(f, fma(mu, fma(3.0 * c4, mu, 2.0 * c2), base), T);
        if (k == 2) T = fma(f, fma(3.0 * c4, mu, c2), T);
        if (k == 3) T = fma(f, c4, T);
#pragma unroll
        for (int t = 0; t < TAY_PER_WAVE; t++) {
            if (W + 4 * t < NTAY && TAY_KK[W + 4 * t] == k) {
                const int j = TAY_JJ[W + 4 * t];
                m[t] = fma(T, xp[k - j] * yp[j], m[t]);
            }
        }
        p15 *= r * ((1.5 - k) / (k + 1));
        p25 *= r * ((2.5 - k) / (k + 1));
    }
}

template <int W>
__device__ __forceinline__ void tay_sweep(const float4 *s_t, int n, double cx, double cy, double cz, const double *cl,
                                          double (&m)[TAY_PER_WAVE]) {
    for (int j = (int)(threadIdx.x & 63); j < n; j += 64) tay_add<W>(s_t[j], cx, cy, cz, cl, m);
}

template <int LT, bool BEAM, bool GLENS, bool MUC, int R>
__global__ __launch_bounds__(NT, 4) void k_flux_mixed(const TaskDev *__restrict__ tasks,
                                                   const BatchDev *__restrict__ batches,
                                                   const GridDev *__restrict__ grids, FlatArrays fl, PhaseArrays pa,
                                                   int S, const int *__restrict__ lists,
                                                   const int *__restrict__ counts, double *__restrict__ partial) {
    // compacted records: [0, FB) faces seen at every sub-phase of the tile,
    // [FB, 2FB) faces whose eclipse starts or ends inside the tile
    __shared__ float4 s_a0[2 * FB];
    __shared__ float4 s_a1[BEAM ? 2 * FB : 1];
    __shared__ float4 s_a2[GLENS ? 2 * FB : 1];
    __shared__ double2 s_io[FB];
    __shared__ int s_cnt[3][NT / 64];
    constexpr bool MOM = !BEAM && !GLENS;  // moment sums (above)
    // Claret law: the full faces taylor_face selects go through the Taylor
    // moments (the others of the full class through the direct sweep)
    constexpr bool TAYF = LT == 1 && MOM;
    constexpr int NM = LT == 0 ? NMOM : 0;  // 3-variable monomials: degree 5 (polynomial law)
    __shared__ float4 s_m[MOM ? FB : 1];
    __shared__ double s_coef[MOM ? NMOM : 1];

    const TaskDev T = tasks[blockIdx.x];
    const BatchDev B = batches[T.batch];
    const GridDev &G = grids[B.grid];
    const int count = B.count_idx < 0 ? S : counts[B.count_idx];
    const int TILE = R * NT;
    const int tile0 = T.tile * TILE;
    if (tile0 >= count) return;
    const int live = min(count - tile0, TILE);

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
    Coef32 K;
    K.c1 = (float)G.ldc[0];
    K.c2 = (float)G.ldc[1];
    K.c3 = (float)G.ldc[2];
    K.c4 = (float)G.ldc[3];
    K.base = (float)(1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]));
    {
        // mu I(mu) = mu - sum_j c_j mu (1 - mu)^j expanded in powers of mu
        const double c1 = G.ldc[0], c2 = G.ldc[1], c3 = G.ldc[2], c4 = G.ldc[3];
        K.q1 = (float)(1.0 - (c1 + c2 + c3 + c4));
        K.q2 = (float)(c1 + 2.0 * c2 + 3.0 * c3 + 4.0 * c4);
        K.q3 = (float)(-(c2 + 3.0 * c3 + 6.0 * c4));
        K.q4 = (float)(c3 + 4.0 * c4);
        K.q5 = (float)(-c4);
    }
    K.mucrit = (float)G.mucrit;
    K.beam = (float)G.beam;
    K.rlens = (float)G.rlens1;
    const float ecx = (float)ti.cx, ecy = (float)ti.cy, ecz = (float)ti.cz;
    const float cull = (float)G.mucrit - 1.e-5f - ti.rad;

    const int n0 = fl.n0[B.grid];
    const int f_lo = T.chunk * B.chunk;
    const int f_hi = min(G.nface, f_lo + B.chunk);
    const float4 *__restrict__ rec0 = fl.rec0 + G.face_off;
    const float4 *__restrict__ rec1 = fl.rec1 + G.face_off;
    const float4 *__restrict__ rec2 = fl.rec2 + G.face_off;
    const double2 *__restrict__ eclr = fl.ecl + G.face_off;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);

    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.;
    unsigned long long evc = 0;  // elements evaluated by this workgroup
    double mom[MOM_PER_WAVE];
#pragma unroll
    for (int t = 0; t < MOM_PER_WAVE; t++) mom[t] = 0.;
    // d.e_c above this: mu > max(0, mucrit) at every sub-phase, no clip or mask
    const float full_cut = fmaxf(0.f, (float)G.mucrit) + 1.e-5f + ti.rad;
    // Claret coefficients for the Taylor moments: c1..c4, base
    const double tcl[5] = {G.ldc[0], G.ldc[1], G.ldc[2], G.ldc[3],
                           1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3])};

    // stage prefetch registers (clamped index keeps the loads unconditional)
    int f = f_lo + (int)threadIdx.x;
    const int flast = f_hi - 1;
    float4 pa0 = rec0[min(f, flast)];
    float4 pa1 = BEAM ? rec1[min(f, flast)] : z4;
    float4 pa2 = GLENS ? rec2[min(f, flast)] : z4;
    double2 pio = eclr[min(f, flast)];

    for (int fb = f_lo; fb < f_hi; fb += FB) {
        const float4 a0 = pa0, a1 = pa1, a2 = pa2;
        const double2 io = pio;
        const bool in = f < f_hi;
        const bool ecl = f >= n0;
        const int fn = f + FB;
        if (fb + FB < f_hi) {  // prefetch the next stage
            const int fc = min(fn, flast);
            pa0 = rec0[fc];
            if (BEAM) pa1 = rec1[fc];
            if (GLENS) pa2 = rec2[fc];
            pio = eclr[fc];
        }
        const FaceDecision fd = decide_face(a0, io, in, ecl, ecx, ecy, ecz, cull, full_cut, ti, MOM);
        const bool act = fd.act, full = fd.full;  // full: to the moment sums
        const int cls = fd.cls;
        const bool tay = TAYF && full && taylor_face(a0, fd.dc, ti);
        const bool tomom = TAYF ? tay : full;  // to the moment list s_m
        int tot0, tot1, totm = 0;
        const int slot0 = compact_slot(act && cls == 0 && !tomom, tot0, s_cnt[0]);
        const int slot1 = compact_slot(act && cls == 1, tot1, s_cnt[1]);
        const int slotm = MOM ? compact_slot(tomom, totm, s_cnt[2]) : 0;
        if (tomom) {
            s_m[slotm] = a0;
        } else if (act) {
            const int slot = cls == 0 ? slot0 : FB + slot1;
            s_a0[slot] = a0;
            if (BEAM) s_a1[slot] = a1;
            if (GLENS) s_a2[slot] = a2;
            if (cls) s_io[slot1] = io;
        }
        __syncthreads();
        evc += (unsigned long long)(tot0 + tot1 + (TAYF ? 0 : totm)) * (unsigned long long)live;
        if (TAYF && totm) {
            const int nm = __builtin_amdgcn_readfirstlane(totm);
            switch (threadIdx.x >> 6) {
            case 0: tay_sweep<0>(s_m, nm, ti.cx, ti.cy, ti.cz, tcl, mom); break;
            case 1: tay_sweep<1>(s_m, nm, ti.cx, ti.cy, ti.cz, tcl, mom); break;
            case 2: tay_sweep<2>(s_m, nm, ti.cx, ti.cy, ti.cz, tcl, mom); break;
            default: tay_sweep<3>(s_m, nm, ti.cx, ti.cy, ti.cz, tcl, mom); break;
            }
        } else if (MOM && totm) {
            const int nm = __builtin_amdgcn_readfirstlane(totm);
            switch (threadIdx.x >> 6) {
            case 0: mom_sweep<0, NM>(s_m, nm, mom); break;
            case 1: mom_sweep<1, NM>(s_m, nm, mom); break;
            case 2: mom_sweep<2, NM>(s_m, nm, mom); break;
            default: mom_sweep<3, NM>(s_m, nm, mom); break;
            }
        }
        sweep<LT, BEAM, GLENS, MUC, false, R>(s_a0, s_a1, s_a2, s_io, tot0, ex, ey, ez, ph, K, acc);
        if (tot1)
            sweep<LT, BEAM, GLENS, MUC, true, R>(s_a0 + FB, BEAM ? s_a1 + FB : s_a1, GLENS ? s_a2 + FB : s_a2, s_io,
                                                 tot1, ex, ey, ez, ph, K, acc);
        __syncthreads();
        f = fn;
    }

    if (TAYF) {
        // the Taylor moments summed over each wave's lanes, times C(k, j);
        // then every sub-phase evaluates the polynomial in dX, dY
        const int w = (int)(threadIdx.x >> 6);
#pragma unroll
        for (int t = 0; t < TAY_PER_WAVE; t++) {
            double v = mom[t];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            const int k = w + 4 * t;
            if ((threadIdx.x & 63) == 0 && k < NTAY) s_coef[k] = v * c_tay_binom[k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
            // the FP64 direction is loaded again (registers, as below)
            const int el = tile0 + r * NT + (int)threadIdx.x;
            const int j = ok[r] ? (B.list_off < 0 ? el : lists[B.list_off + el]) : 0;
            const size_t idx = (size_t)B.inst * S + j;
            const double X = ok[r] ? pa.ex[idx] - ti.cx : 0., Y = ok[r] ? pa.ey[idx] - ti.cy : 0.;
            double xp[TAY_K + 1], yp[TAY_K + 1];
            xp[0] = yp[0] = 1.0;
#pragma unroll
            for (int i = 1; i <= TAY_K; i++) {
                xp[i] = xp[i - 1] * X;
                yp[i] = yp[i - 1] * Y;
            }
            double e = 0.;
#pragma unroll
            for (int k = 0; k < NTAY; k++) e = fma(s_coef[k], xp[TAY_KK[k] - TAY_JJ[k]] * yp[TAY_JJ[k]], e);
            acc[r] += e;
        }
    } else if (MOM) {
        // each wave's moments summed over its lanes, scaled into the
        // polynomial's coefficients; then every sub-phase evaluates it
        const double c1 = G.ldc[0], c2 = G.ldc[1], c3 = G.ldc[2], c4 = G.ldc[3];
        const double base = 1.0 - (c1 + c2 + c3 + c4);
        // coefficient of mu^n of the polynomial part
        const double q[6] = {0., base, LT == 0 ? c1 + 2.0 * c2 + 3.0 * c3 + 4.0 * c4 : c2,
                             LT == 0 ? -(c2 + 3.0 * c3 + 6.0 * c4) : c4, LT == 0 ? c3 + 4.0 * c4 : 0.,
                             LT == 0 ? -c4 : 0.};
        const int w = (int)(threadIdx.x >> 6);
#pragma unroll
        for (int t = 0; t < MOM_PER_WAVE; t++) {
            double v = mom[t];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            const int k = w + 4 * t;
            if ((threadIdx.x & 63) == 0 && k < NM) s_coef[k] = v * q[c_mom_n[k]] * c_mom_mult[k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
            // Claret: the FP64 direction is loaded again, since keeping it
            // live through the face loop costs registers its sweep needs
            double x = dex[r], y = dey[r], z = dez[r];
            if (LT == 1) {
                const int el = tile0 + r * NT + (int)threadIdx.x;
                const int j = ok[r] ? (B.list_off < 0 ? el : lists[B.list_off + el]) : 0;
                const size_t idx = (size_t)B.inst * S + j;
                x = ok[r] ? pa.ex[idx] : 0.;
                y = ok[r] ? pa.ey[idx] : 0.;
                z = ok[r] ? pa.ez[idx] : 0.;
            }
            double xp[6], yp[6], zp[6];
            xp[0] = yp[0] = zp[0] = 1.0;
#pragma unroll
            for (int i = 1; i < 6; i++) {
                xp[i] = xp[i - 1] * x;
                yp[i] = yp[i - 1] * y;
                zp[i] = zp[i - 1] * z;
            }
            double e = 0.;
#pragma unroll
            for (int k = 0; k < NM; k++) e = fma(s_coef[k], xp[MOM_A[k]] * yp[MOM_B[k]] * zp[MOM_C[k]], e);
            acc[r] += e;
        }
    }
    double *out = partial + B.part_off + (size_t)T.chunk * (size_t)(B.ntile * TILE);
#pragma unroll
    for (int r = 0; r < R; r++) {
        int e = tile0 + r * NT + (int)threadIdx.x;
        if (ok[r]) out[e] = acc[r];
    }
    // one counter update per workgroup, spread over 64 slots per variant
    if (threadIdx.x == 0 && evc)
        atomicAdd(fl.evald + (LT | (BEAM ? 2 : 0) | (GLENS ? 4 : 0) | (MUC ? 8 : 0)) * 64 + (blockIdx.x & 63), evc);
}

// ----------------------------------------------------------- per-wave tiles
// k_flux_wave: the flux sums of grids without beaming or lensing (the
// moment variants of k_flux_mixed), with the task's phase tile split
// between its waves.  Wave w owns the R*64 consecutive sub-phases
// [tile0 + w R 64, tile0 + (w + 1) R 64) and makes every per-tile decision
// (cull, eclipse class, moments or direct sum) against its own centre and
// radius, a quarter of the workgroup tile's.  A face crossing the limb then
// costs direct elements only in the wave whose sub-phases it crosses in,
// and the moment sums accept faces four times closer to the limb, so the
// direct elements of a batched launch drop by about half (1/4 per tile
// width, twice as many moment faces).  There is no workgroup barrier: each
// wave streams the chunk's records itself (64 faces per step, one per
// lane; the four waves of a workgroup read the same lines, which the CU's
// cache serves, and the XCD-aware block order below keeps a chunk's phase
// tiles on one XCD's L2), compacts the faces it sweeps into its own LDS lists and
// queues its moment faces in an LDS ring, from which it sums them a full
// wave at a time (Claret: 32 faces, two lanes each; polynomial law: 64).
//
// Moments of the Claret law: Taylor moments as in k_flux_mixed, in delta =
// d_x dX + d_y dY, here of degree 8 with |delta| <= 0.27 mu_c (the same
// accuracy; a ninth degree would not fit the registers of four waves per
// SIMD).  The lower half of the wave accumulates the 25 monomials (k, j)
// with j <= k - j, the upper half runs the same code with d_x and d_y
// exchanged, which gives the mirrored monomials (k, k - j); the
// self-mirrored ones (2j = k) are taken from the lower half only.  Polynomial law: mu I(mu) = Q(mu) of degree 5, so its
// Taylor series about mu_c = d.e_c ends at degree 5 and is exact:
//   sum_i f_i Q(d_i.e) = sum_{k<=5, j<=k} C(k, j) M_kj dX^(k-j) dY^j,
//   M_kj = sum_i f_i Q^(k)(mu_ci)/k! d_xi^(k-j) d_yi^j,
// 21 FP64 moments per lane, one lane per face.  Both need the wave's
// sub-phases to share e_z (one inclination); otherwise those faces are
// summed directly.
constexpr int WQ = 128;                     // moment queue entries per wave
// Claret Taylor moments of the per-wave tiles: degree 8 with |delta| <=
// 0.27 mu_c (remainder of the mu^(3/2) series below 2.3e-8 of the term,
// mu^(5/2) below 9e-9), 25 monomials per lane
constexpr int WK = 8;
constexpr float TAY_RHO_W = 0.27f;
constexpr int NTA = 25;
constexpr unsigned char TA_K[NTA] = {0,1,2,2,3,3,4,4,4,5,5,5,6,6,6,6,7,7,7,7,8,8,8,8,8};
constexpr unsigned char TA_J[NTA] = {0,0,0,1,0,1,0,1,2,0,1,2,0,1,2,3,0,1,2,3,0,1,2,3,4};
constexpr int NPM = 21;                     // polynomial-law moments, k <= 5
// binomial C(n, k) and Taylor coefficients of mu^(3/2), mu^(5/2):
// C(3/2, k), C(5/2, k)
constexpr double BIN15[TAY_K + 1] = {1., 1.5, 0.375, -0.0625, 0.0234375, -0.01171875, 0.0068359375,
                                     -0.00439453125, 0.0030212402343750, -0.0021820068359375};
constexpr double BIN25[TAY_K + 1] = {1., 2.5, 1.875, 0.3125, -0.0390625, 0.01171875, -0.00488281250,
                                     0.00244140625, -0.0013732910156250, 0.0008392333984375};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// a value equal in every active lane, moved to scalar registers
__device__ __forceinline__ double uni(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// set bits of m below this lane
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
__device__ __forceinline__ float unif(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// m[t] += f T_k(mu_c) u^(k-j) v^j for the 25 lower-half monomials
// (k, j) = (TA_K, TA_J)[t], T_k = C(3/2, k) c1 mu^(3/2-k) + C(5/2, k) c3
// mu^(5/2-k) + the integer powers' coefficients (k <= 3).  (u, v) = (d_x,
// d_y), or (d_y, d_x) in the upper half of the wave.  The binomials are
// dyadic fractions with short mantissas (FP64 literal operands).
__device__ __forceinline__ void tay_add_half(const float4 &rec, bool swap, double cx, double cy, double cz,
                                             double c1, double c2, double c3, double c4, double base,
                                             double (&m)[NTA]) {
    const double x = rec.x, y = rec.y, z = rec.z, f = rec.w;
    const double mu = fma(z, cz, fma(y, cy, x * cx));
    double q = __builtin_amdgcn_rsq(mu);
    q = q * fma(-0.5 * mu, q * q, 1.5);
    const double r = q * q;
    const double u = swap ? y : x, v = swap ? x : y;
    // u^(k-j) v^j = u^k (v/u)^j: one ratio instead of two power tables
    // (registers); with u = 0 every monomial of the half but k = 0 is zero
    double ru = __builtin_amdgcn_rcp(u);
    ru = ru * fma(-u, ru, 2.0);
    const double wr = u == 0.0 ? 0.0 : v * ru;
    const double fm15 = f * mu * (mu * q);  // f mu^(3/2)
    double s1 = fm15 * c1, s3 = fm15 * mu * c3;  // f c1 mu^(3/2 - k), f c3 mu^(5/2 - k)
    double uk = 1.0;  // u^k
#pragma unroll
    for (int k = 0; k <= WK; k++) {
        double T = fma(BIN15[k], s1, BIN25[k] * s3);
        if (k == 0) T = fma(f * mu, fma(mu, fma(mu, c4, c2), base), T);
        if (k == 1) T = fma(f, fma(mu, fma(3.0 * c4, mu, 2.0 * c2), base), T);
        if (k == 2) T = fma(f, fma(3.0 * c4, mu, c2), T);
        if (k == 3) T = fma(f, c4, T);
        double t = k == 0 ? T : T * uk;
#pragma unroll
        for (int tt = 0; tt < NTA; tt++) {
            if (TA_K[tt] == k) {
                if (TA_J[tt] > 0) t *= wr;
                m[tt] += t;
            }
        }
        uk *= u;
        s1 *= r;
        s3 *= r;
    }
}

// m[k(k+1)/2 + j] += f Q^(k)(mu_c)/k! d_x^(k-j) d_y^j, Q(mu) = sum_n q_n mu^n
// (n = 1..5, qc[n-1]); the Taylor shift by synthetic division.
__device__ __forceinline__ void poly_add(const float4 &rec, double cx, double cy, double cz, const double *qc,
                                         double (&m)[NPM]) {
    const double x = rec.x, y = rec.y, z = rec.z, f = rec.w;
    const double mu = fma(z, cz, fma(y, cy, x * cx));
    double b[6] = {0., qc[0], qc[1], qc[2], qc[3], qc[4]};
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int k = 4; k >= i; k--) b[k] = fma(mu, b[k + 1], b[k]);
    double up[6], vp[6];
    up[0] = f;
    vp[0] = 1.0;
    vp[1] = y;
#pragma unroll
    for (int i = 1; i < 6; i++) up[i] = up[i - 1] * x;
#pragma unroll
    for (int i = 2; i < 6; i++) vp[i] = vp[i - 1] * y;
#pragma unroll
    for (int k = 0; k <= 5; k++)
#pragma unroll
        for (int j = 0; j <= k; j++) {
            const int t = k * (k + 1) / 2 + j;
            m[t] = j == 0 ? fma(b[k], up[k], m[t]) : fma(b[k] * vp[j], up[k - j], m[t]);
        }
}

template <int LT, bool MUC, int R>
__global__ __launch_bounds__(NT, 4) void k_flux_wave(const TaskDev *__restrict__ tasks, int ntask,
                                                  const BatchDev *__restrict__ batches,
                                                  const GridDev *__restrict__ grids, FlatArrays fl, PhaseArrays pa,
                                                  int S, const int *__restrict__ lists,
                                                  const int *__restrict__ counts, double *__restrict__ partial) {
    constexpr int NW = NT / 64;
    constexpr int NMW = LT == 1 ? NTA : NPM;    // moments per lane
    constexpr int QB = LT == 1 ? 32 : 64;       // moment faces per queue pass
    // faces per unrolled sweep step: the Claret moments hold 50 registers
    // through the face loop, so its sweeps batch fewer LDS reads
    constexpr int SWU = LT == 1 && R > 1 ? 2 : 4;
    __shared__ float4 s_d0[NW][64];             // direct faces, no eclipse edge in the wave's tile
    __shared__ float4 s_d1[NW][64];             // faces whose eclipse starts or ends inside it
    __shared__ double2 s_io[NW][64];            // their ranges
    __shared__ float4 s_q[NW][WQ];              // moment queue (ring)
    __shared__ double s_cf[NW][NTAY];           // reduced moments x binomials

    // wave-uniform values are kept in scalar registers (readfirstlane)
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    // XCD-aware order: workgroups are dispatched to the 8 XCDs in turn
    // (block b to XCD b % 8), so XCD x takes the x-th eighth of the task
    // table in order; tasks are chunk-major (the tiles of one face chunk are
    // consecutive), so one XCD runs a chunk's tiles back to back and its L2
    // serves the chunk's records to all of them
    const int per_xcd = (ntask + 7) >> 3;
    const int ti_ = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (ti_ >= ntask) return;
    const TaskDev T = tasks[ti_];
    const BatchDev B = batches[T.batch];
    const GridDev &G = grids[B.grid];
    const int count = B.count_idx < 0 ? S : counts[B.count_idx];
    const int TILE = R * NT;
    const int sub0 = T.tile * TILE + w * (R * 64);  // this wave's first element
    if (sub0 >= count) return;                      // no workgroup barrier below
    const int live = min(count - sub0, R * 64);

    // this wave's sub-phases: element sub0 + r*64 + lane
    float ex[R], ey[R], ez[R];
    double ph[R];
    bool ok[R];
    TileInfo ti;
    {
        const int em = sub0 + (live - 1) / 2;
        const int jm = B.list_off < 0 ? em : lists[B.list_off + em];
        const size_t im = (size_t)B.inst * S + jm;
        ti.cx = uni(pa.ex[im]);
        ti.cy = uni(pa.ey[im]);
        ti.cz = uni(pa.ez[im]);
        float rad = 0.f;
        double pmin = 2.0, pmax = -1.0;
        bool zd = false;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int el = sub0 + r * 64 + lane;
            ok[r] = el < count;
            const int j = ok[r] ? (B.list_off < 0 ? el : lists[B.list_off + el]) : 0;
            const size_t idx = (size_t)B.inst * S + j;
            const double dx = ok[r] ? pa.ex[idx] : 0., dy = ok[r] ? pa.ey[idx] : 0., dz = ok[r] ? pa.ez[idx] : 0.;
            ph[r] = ok[r] ? pa.phin[idx] : 0.5;
            ex[r] = (float)dx;
            ey[r] = (float)dy;
            ez[r] = (float)dz;
            if (ok[r]) {
                const double a = dx - ti.cx, b = dy - ti.cy, c = dz - ti.cz;
                rad = fmaxf(rad, (float)sqrt(a * a + b * b + c * c));
                pmin = fmin(pmin, ph[r]);
                pmax = fmax(pmax, ph[r]);
                zd = zd || dz != ti.cz;
            }
        }
        ti.rad = unif(wave_max(rad));
        ti.pmin = uni(wave_min_d(pmin));
        ti.pmax = uni(wave_max_d(pmax));
        ti.zflat = __ballot(zd) == 0ull;
    }
    Coef32 K;
    K.c1 = (float)G.ldc[0];
    K.c2 = (float)G.ldc[1];
    K.c3 = (float)G.ldc[2];
    K.c4 = (float)G.ldc[3];
    K.base = (float)(1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]));
    const double c1 = G.ldc[0], c2 = G.ldc[1], c3 = G.ldc[2], c4 = G.ldc[3];
    const double base = uni(1.0 - (c1 + c2 + c3 + c4));
    // mu I(mu) = mu - sum_j c_j mu (1 - mu)^j expanded in powers of mu
    const double qc[5] = {base, uni(c1 + 2.0 * c2 + 3.0 * c3 + 4.0 * c4), uni(-(c2 + 3.0 * c3 + 6.0 * c4)),
                          uni(c3 + 4.0 * c4), -c4};
    K.q1 = (float)qc[0];
    K.q2 = (float)qc[1];
    K.q3 = (float)qc[2];
    K.q4 = (float)qc[3];
    K.q5 = (float)qc[4];
    K.mucrit = (float)G.mucrit;
    K.beam = 0.f;
    K.rlens = 0.f;
    const float ecx = (float)ti.cx, ecy = (float)ti.cy, ecz = (float)ti.cz;
    const float cull = unif((float)G.mucrit - 1.e-5f - ti.rad);
    const float full_cut = unif(fmaxf(0.f, (float)G.mucrit) + 1.e-5f + ti.rad);
    // the polynomial law's moments need one inclination in the tile
    const bool pmom = LT == 0 && ti.zflat;

    const int n0 = fl.n0[B.grid];
    const int f_lo = T.chunk * B.chunk;
    const int f_hi = min(G.nface, f_lo + B.chunk);
    const float4 *__restrict__ rec0 = fl.rec0 + G.face_off;
    const double2 *__restrict__ eclr = fl.ecl + G.face_off;

    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.;
    double mom[NMW];
#pragma unroll
    for (int t = 0; t < NMW; t++) mom[t] = 0.;
    unsigned long long evc = 0;
    int qhead = 0, qn = 0;

    auto moment_pass = [&](int nq) {  // nq <= QB queued faces from qhead
        if constexpr (LT == 1) {
            const int fi = lane & 31;
            if (fi < nq)
                tay_add_half(s_q[w][(qhead + fi) & (WQ - 1)], lane >= 32, ti.cx, ti.cy, ti.cz, c1, c2, c3, c4, base,
                             mom);
        } else {
            if (lane < nq) poly_add(s_q[w][(qhead + lane) & (WQ - 1)], ti.cx, ti.cy, ti.cz, qc, mom);
        }
    };

    int f = f_lo + lane;
    const int flast = f_hi - 1;
    float4 pa0 = rec0[min(f, flast)];
    double2 pio = f >= n0 ? eclr[min(f, flast)] : make_double2(0., 0.);
    for (int fb = f_lo; fb < f_hi; fb += 64) {
        const float4 a0 = pa0;
        const double2 io = pio;
        const bool in = f < f_hi;
        const bool ecl = f >= n0;
        const int fn = f + 64;
        if (fb + 64 < f_hi) {  // prefetch the next step
            pa0 = rec0[min(fn, flast)];
            pio = fn >= n0 ? eclr[min(fn, flast)] : make_double2(0., 0.);
        }
        const FaceDecision fd = decide_face(a0, io, in, ecl, ecx, ecy, ecz, cull, full_cut, ti, true);
        const bool tomom =
            fd.full && (LT == 1 ? ti.zflat && ti.rad * ti.rad * fmaf(a0.y, a0.y, a0.x * a0.x) <=
                                                  (TAY_RHO_W * TAY_RHO_W) * (fd.dc * fd.dc)
                                : pmom);
        const bool d0 = fd.act && fd.cls == 0 && !tomom, d1 = fd.act && fd.cls == 1;
        const unsigned long long b0 = __ballot(d0), b1 = __ballot(d1), bm = __ballot(tomom);
        wave_sync();  // the previous step's list reads are done
        if (d0) s_d0[w][lanes_below(b0)] = a0;
        if (d1) {
            const int k = lanes_below(b1);
            s_d1[w][k] = a0;
            s_io[w][k] = io;
        }
        if (tomom) s_q[w][(qhead + qn + lanes_below(bm)) & (WQ - 1)] = a0;
        wave_sync();
        const int n_d0 = __popcll(b0), n_d1 = __popcll(b1);
        qn += __popcll(bm);
        evc += (unsigned long long)(n_d0 + n_d1) * (unsigned long long)live;
        sweep<LT, false, false, MUC, false, R, SWU>(s_d0[w], s_d0[w], s_d0[w], s_io[w], n_d0, ex, ey, ez, ph, K, acc);
        if (n_d1)
            sweep<LT, false, false, MUC, true, R, SWU>(s_d1[w], s_d1[w], s_d1[w], s_io[w], n_d1, ex, ey, ez, ph, K, acc);
        while (qn >= QB) {
            moment_pass(QB);
            qhead = (qhead + QB) & (WQ - 1);
            qn -= QB;
        }
        f = fn;
    }
    if (qn > 0) moment_pass(qn);

    // moments summed over the lanes (Claret: each half of the wave), times
    // the binomials, into this wave's coefficient table; then every
    // sub-phase evaluates the polynomial in (dX, dY)
    if constexpr (LT == 1) {
#pragma unroll
        for (int t = 0; t < NTA; t++) {
            double v = mom[t];
            for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            const int k = TA_K[t], j = TA_J[t];
            const double c = v * c_tay_binom[k * (k + 1) / 2 + j];
            if (lane == 0) s_cf[w][k * (k + 1) / 2 + j] = c;
            if (lane == 32 && 2 * j != k) s_cf[w][k * (k + 1) / 2 + (k - j)] = c;
        }
    } else {
        if (pmom) {
#pragma unroll
            for (int t = 0; t < NPM; t++) {
                double v = mom[t];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) s_cf[w][t] = v * c_tay_binom[t];
            }
        }
    }
    wave_sync();
    if (LT == 1 || pmom) {
        constexpr int KMAX = LT == 1 ? WK : 5;
#pragma unroll
        for (int r = 0; r < R; r++) {
            // the FP64 direction is loaded again (registers)
            const int el = sub0 + r * 64 + lane;
            const int j = ok[r] ? (B.list_off < 0 ? el : lists[B.list_off + el]) : 0;
            const size_t idx = (size_t)B.inst * S + j;
            const double X = ok[r] ? pa.ex[idx] - ti.cx : 0., Y = ok[r] ? pa.ey[idx] - ti.cy : 0.;
            // sum_k sum_j cf[k, j] X^(k-j) Y^j = sum_a X^a Q_a(Y),
            // Q_a(Y) = sum_b cf[a + b, b] Y^b: Horner in Y inside Horner in X
            double e = 0.;
#pragma unroll
            for (int a = KMAX; a >= 0; a--) {
                double qa = 0.;
#pragma unroll
                for (int b = KMAX - a; b >= 0; b--) qa = fma(qa, Y, s_cf[w][(a + b) * (a + b + 1) / 2 + b]);
                e = fma(e, X, qa);
            }
            acc[r] += e;
        }
    }
    double *out = partial + B.part_off + (size_t)T.chunk * (size_t)(B.ntile * TILE);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int el = sub0 + r * 64 + lane;
        if (ok[r]) out[el] = acc[r];
    }
    if (lane == 0 && evc) atomicAdd(fl.evald + (LT | (MUC ? 8 : 0)) * 64 + (blockIdx.x & 63), evc);
}

// FP64 element in the reference's operand order (comp_light_flat.cpp:81-168);
// zero when the reference's masks make it zero.
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
    __shared__ double s_d[7][FB];  // dx dy dz px py pz flux
    __shared__ double2 s_io[FB];
    __shared__ int s_cls[FB];
    __shared__ int s_cnt[NT / 64];

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
    for (int fb = f_lo; fb < f_hi; fb += FB) {
        const int f = fb + (int)threadIdx.x;
        bool act = false;
        int cls = 0;
        double d[7] = {0., 0., 0., 0., 0., 0., 0.};
        double2 io = make_double2(0., 0.);
        if (f < f_hi) {
            d[0] = fl.fdx[o + f];
            d[1] = fl.fdy[o + f];
            d[2] = fl.fdz[o + f];
            act = d[0] * cx + d[1] * cy + d[2] * cz > cull;
            if (act && f >= n0) {
                io = fl.ecl[o + f];
                cls = tile_eclipse_class(ti, io.x, io.y);
                act = cls != 2;
            }
            if (act) {
                d[3] = fl.fpx[o + f];
                d[4] = fl.fpy[o + f];
                d[5] = fl.fpz[o + f];
                d[6] = fl.fflux[o + f];
            }
        }
        int tot;
        const int slot = compact_slot(act, tot, s_cnt);
        if (act) {
#pragma unroll
            for (int k = 0; k < 7; k++) s_d[k][slot] = d[k];
            s_io[slot] = io;
            s_cls[slot] = cls;
        }
        __syncthreads();
        for (int k = 0; k < tot; k++) {
            const double dx = s_d[0][k], dy = s_d[1][k], dz = s_d[2][k];
            const double px = s_d[3][k], py = s_d[4][k], pz = s_d[5][k], flux = s_d[6][k];
            const int c = s_cls[k];
            const double2 e = s_io[k];
#pragma unroll
            for (int r = 0; r < R; r++) {
                double v = elem64<LT, BEAM, GLENS>(ex[r], ey[r], ez[r], dx, dy, dz, px, py, pz, flux, G, base);
                if (c == 1 && ecl_range(ph[r], e.x, e.y)) v = 0.0;
                if (v != 0.0) acc[r] += v;
            }
        }
        __syncthreads();
    }
    double *out = partial + B.part_off + (size_t)T.chunk * (size_t)(B.ntile * TILE);
#pragma unroll
    for (int r = 0; r < R; r++) {
        int e = tile0 + r * NT + (int)threadIdx.x;
        if (ok[r]) out[e] = acc[r];
    }
}

// Flux sums of a grid with an opaque disc's ranges: one thread per routed
// sub-phase, the grid's faces in the reference's FlatGrid order (faces with
// no eclipse range, with one, with more; each class in index order,
// comp_light_flat.cpp:26-65), FP64 elements summed in 1024-face blocks that
// are merged in order (comp_light_flat.cpp:242-307): the reference's sums
// bit for bit.  A face is hidden when any of its ranges (the star's, then
// the disc's) holds the phase, as its multi-range test does.  Every thread
// of a workgroup reads the same face at the same time.  One partial row.
template <int LT, bool BEAM, bool GLENS>
__global__ __launch_bounds__(256) void k_flux_ragged(const int *__restrict__ order, const BatchDev *__restrict__ batches,
                                                    const GridDev *__restrict__ grids, FaceArrays fa, PhaseArrays pa,
                                                    int S, const int *__restrict__ lists,
                                                    const int *__restrict__ counts, double *__restrict__ partial) {
    const BatchDev B = batches[order[blockIdx.y]];
    const GridDev &G = grids[B.grid];
    const int count = B.count_idx < 0 ? S : counts[B.count_idx];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const int j = B.list_off < 0 ? e : lists[B.list_off + e];
    const size_t idx = (size_t)B.inst * S + j;
    const double ex = pa.ex[idx], ey = pa.ey[idx], ez = pa.ez[idx], ph = pa.phin[idx];
    const double base = 1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]);
    double total = 0., blk = 0.;
    int pos = 0;
    for (int cls = 0; cls < 3; cls++) {
        for (int k = 0; k < G.nface; k++) {
            const int f = G.face_off + k;
            const int se = fa.ecl[f] ? 1 : 0, nd = fa.dn[f];
            if (min(se + nd, 2) != cls) continue;
            bool hid = se && ecl_range(ph, fa.ein[f], fa.eout[f]);
            for (int r = 0; r < nd && !hid; r++) {
                const double2 d = fa.dr[2 * f + r];
                hid = ecl_range(ph, d.x, d.y);
            }
            if (!hid)
                blk += elem64<LT, BEAM, GLENS>(ex, ey, ez, fa.dx[f], fa.dy[f], fa.dz[f], fa.px[f], fa.py[f], fa.pz[f],
                                               (double)fa.flux[f], G, base);
            if (++pos == 1024) {
                total += blk;
                blk = 0.;
                pos = 0;
            }
        }
    }
    partial[B.part_off + e] = total + blk;
}

// Faces with several eclipse ranges at the end of a flat grid (the flat-sum
// entry points): one thread per sub-phase and chunk, continuing the chunk's
// partial in face order as the reference's block loop does.
template <int LT, bool BEAM, bool GLENS>
__global__ __launch_bounds__(256) void k_flux_ragged_tail(const GridDev *__restrict__ grid,
                                                         const BatchDev *__restrict__ batch, FlatArrays fl,
                                                         const int *__restrict__ moff,
                                                         const double2 *__restrict__ ranges, PhaseArrays pa, int S,
                                                         int n1, int nface, int chunk, int c_first, int nchunk_main,
                                                         size_t stride, double *__restrict__ partial) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S) return;
    const GridDev &G = *grid;
    const int c = c_first + (int)blockIdx.y;
    const double ex = pa.ex[e], ey = pa.ey[e], ez = pa.ez[e], ph = pa.phin[e];
    const double base = 1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]);
    double *out = partial + batch->part_off + (size_t)c * stride + e;
    double acc = c < nchunk_main ? *out : 0.0;
    const int lo = max(c * chunk, n1), hi = min((c + 1) * chunk, nface);
    for (int i = lo; i < hi; i++) {
        bool hid = false;
        for (int j = moff[i - n1]; j < moff[i - n1 + 1] && !hid; j++) hid = ecl_range(ph, ranges[j].x, ranges[j].y);
        if (!hid)
            acc += elem64<LT, BEAM, GLENS>(ex, ey, ez, fl.fdx[i], fl.fdy[i], fl.fdz[i], fl.fpx[i], fl.fpy[i],
                                           fl.fpz[i], fl.fflux[i], G, base);
    }
    *out = acc;
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
void launch_variant(hipStream_t s, int variant, int fp64, int wave, const TaskDev *tasks, int ntask,
                    const BatchDev *batches, const GridDev *grids, FlatArrays fl, PhaseArrays pa, int S,
                    const int *lists, const int *counts, double *partial) {
    if (!fp64 && wave && (variant & 6) == 0) {  // no beaming, no lensing: per-wave tiles
        const bool mc = (variant & 8) != 0;
        const int grid8 = ((ntask + 7) >> 3) << 3;  // a whole number of blocks per XCD
        if (variant & 1) {
            if (mc)
                hipLaunchKernelGGL((k_flux_wave<1, true, R>), dim3(grid8), dim3(NT), 0, s, tasks, ntask, batches, grids,
                                   fl, pa, S, lists, counts, partial);
            else
                hipLaunchKernelGGL((k_flux_wave<1, false, R>), dim3(grid8), dim3(NT), 0, s, tasks, ntask, batches, grids,
                                   fl, pa, S, lists, counts, partial);
        } else {
            if (mc)
                hipLaunchKernelGGL((k_flux_wave<0, true, R>), dim3(grid8), dim3(NT), 0, s, tasks, ntask, batches, grids,
                                   fl, pa, S, lists, counts, partial);
            else
                hipLaunchKernelGGL((k_flux_wave<0, false, R>), dim3(grid8), dim3(NT), 0, s, tasks, ntask, batches, grids,
                                   fl, pa, S, lists, counts, partial);
        }
        return;
    }
#define LC_FLUX_CASE(V, LT, BM, GL)                                                                                 \
    case V:                                                                                                         \
        if (fp64)                                                                                                   \
            hipLaunchKernelGGL((k_flux_fp64<LT, BM, GL, R>), dim3(ntask), dim3(NT), 0, s, tasks, batches, grids,   \
                               fl, pa, S, lists, counts, partial);                                                 \
        else if (muc)                                                                                               \
            hipLaunchKernelGGL((k_flux_mixed<LT, BM, GL, true, R>), dim3(ntask), dim3(NT), 0, s, tasks, batches,   \
                               grids, fl, pa, S, lists, counts, partial);                                          \
        else                                                                                                        \
            hipLaunchKernelGGL((k_flux_mixed<LT, BM, GL, false, R>), dim3(ntask), dim3(NT), 0, s, tasks, batches,  \
                               grids, fl, pa, S, lists, counts, partial);                                          \
        break;
    const bool muc = (variant & 8) != 0;
    switch (variant & 7) {
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

// variant = ltype | beam << 1 | glens << 2 | (mucrit > 0) << 3; rphase in {1, 2, 4}
void launch_flux(hipStream_t s, int variant, int fp64, int rphase, int wave, const TaskDev *tasks, int ntask,
                 const BatchDev *batches, const GridDev *grids, FlatArrays fl, PhaseArrays pa, int S,
                 const int *lists, const int *counts, double *partial) {
    if (ntask <= 0) return;
    if (rphase == 1)
        launch_variant<1>(s, variant, fp64, wave, tasks, ntask, batches, grids, fl, pa, S, lists, counts, partial);
    else if (rphase == 4)
        launch_variant<4>(s, variant, fp64, wave, tasks, ntask, batches, grids, fl, pa, S, lists, counts, partial);
    else
        launch_variant<2>(s, variant, fp64, wave, tasks, ntask, batches, grids, fl, pa, S, lists, counts, partial);
}

__global__ __launch_bounds__(256) void k_make_tasks(const int *__restrict__ first, const int *__restrict__ order,
                                                   int nb, const BatchDev *__restrict__ batches, int ntask,
                                                   TaskDev *__restrict__ tasks) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntask) return;
    int lo = 0, hi = nb - 1;  // last k with first[k] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int b = order[lo], local = t - first[lo], ntile = batches[b].ntile;
    TaskDev T;
    T.batch = b;
    T.chunk = local / ntile;  // chunk-major: a chunk's tiles are consecutive
    T.tile = local - T.chunk * ntile;
    T.pad = 0;
    tasks[t] = T;
}

void launch_make_tasks(hipStream_t s, const int *first, const int *order, int nb, const BatchDev *batches,
                       int ntask, TaskDev *tasks) {
    if (ntask <= 0 || nb <= 0) return;
    hipLaunchKernelGGL(k_make_tasks, dim3((ntask + 255) / 256), dim3(256), 0, s, first, order, nb, batches, ntask,
                       tasks);
}

void launch_flux_ragged_tail(hipStream_t s, int variant, const GridDev *grid, const BatchDev *batch, FlatArrays fl,
                             const int *moff, const double2 *ranges, PhaseArrays pa, int S, int n1, int nface,
                             int chunk, int c_first, int nchunk, int nchunk_main, size_t stride, double *partial) {
    if (S <= 0 || nchunk <= c_first) return;
    const dim3 g((S + 255) / 256, nchunk - c_first);
#define LC_TAIL_CASE(V, LT, BM, GL)                                                                               \
    case V:                                                                                                     \
        hipLaunchKernelGGL((k_flux_ragged_tail<LT, BM, GL>), g, dim3(256), 0, s, grid, batch, fl, moff, ranges, pa, \
                           S, n1, nface, chunk, c_first, nchunk_main, stride, partial);                           \
        break;
    switch (variant & 7) {
        LC_TAIL_CASE(0, 0, false, false)
        LC_TAIL_CASE(1, 1, false, false)
        LC_TAIL_CASE(2, 0, true, false)
        LC_TAIL_CASE(3, 1, true, false)
        LC_TAIL_CASE(4, 0, false, true)
        LC_TAIL_CASE(5, 1, false, true)
        LC_TAIL_CASE(6, 0, true, true)
        LC_TAIL_CASE(7, 1, true, true)
    default: break;
    }
#undef LC_TAIL_CASE
}

void launch_flux_ragged(hipStream_t s, int variant, const int *order, int nb, int max_count, const BatchDev *batches,
                        const GridDev *grids, FaceArrays fa, PhaseArrays pa, int S, const int *lists,
                        const int *counts, double *partial) {
    if (nb <= 0 || max_count <= 0) return;
    const dim3 grid((max_count + 255) / 256, nb);
#define LC_RAGGED_CASE(V, LT, BM, GL)                                                                            \
    case V:                                                                                                    \
        hipLaunchKernelGGL((k_flux_ragged<LT, BM, GL>), grid, dim3(256), 0, s, order, batches, grids, fa, pa, S, \
                           lists, counts, partial);                                                            \
        break;
    switch (variant & 7) {
        LC_RAGGED_CASE(0, 0, false, false)
        LC_RAGGED_CASE(1, 1, false, false)
        LC_RAGGED_CASE(2, 0, true, false)
        LC_RAGGED_CASE(3, 1, true, false)
        LC_RAGGED_CASE(4, 0, false, true)
        LC_RAGGED_CASE(5, 1, false, true)
        LC_RAGGED_CASE(6, 0, true, true)
        LC_RAGGED_CASE(7, 1, true, true)
    }
#undef LC_RAGGED_CASE
}

void launch_flux_reduce(hipStream_t s, const BatchDev *batches, int nbatch, int S, int tile_len, const int *lists,
                        const int *counts, const double *partial, double *out) {
    if (nbatch <= 0 || S <= 0) return;
    hipLaunchKernelGGL(k_flux_reduce, dim3((S + 255) / 256, nbatch), dim3(256), 0, s, batches, S, tile_len, lists,
                       counts, partial, out);
}

}  // namespace lck
