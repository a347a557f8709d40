// k_phase.hip -- orbital phases and sub-exposures, fine/coarse routing,
// scatter of weighted sums back to data points, scale factor and chi**2
// (light_curve_comp.cpp:350-586, re_scale.cpp:16-43).
#include <algorithm>

#include "lc_kernels.h"
#include "lc_math.h"

namespace lck {
using namespace lcx;

namespace {

__device__ __forceinline__ int gint_type(double phase, double phase1, double phase2) {
    double pnorm = phase - floor(phase);
    if (pnorm <= phase1 || pnorm >= 1. - phase1) return 1;
    if ((pnorm > phase1 && pnorm < phase2) || (pnorm > 1. - phase2 && pnorm < 1. - phase1)) return 2;
    return 3;
}
__device__ __forceinline__ double gscale1(const InstDev &I, double phase) {
    double pnorm = phase - floor(phase);
    if (pnorm <= I.phase1 || pnorm >= 1. - I.phase1) return 1.;
    return (I.scale11 * (1. - I.phase1 - pnorm) + I.scale12 * (pnorm - I.phase1)) / (1 - 2. * I.phase1);
}
__device__ __forceinline__ double gscale2(const InstDev &I, double phase) {
    double pnorm = phase - floor(phase);
    if (pnorm >= I.phase2 && pnorm <= 1. - I.phase2) return 1.;
    if (pnorm < 0.5) return (I.scale22 * (I.phase2 - pnorm) + I.scale21 * (pnorm + I.phase2)) / (2. * I.phase2);
    return (I.scale21 * (1. + I.phase2 - pnorm) + I.scale22 * (pnorm - 1. + I.phase2)) / (2. * I.phase2);
}

// One work item per (instance, data point): ephemeris Newton steps, the
// deltat term, the slope/quad/cube factor, then every sub-exposure's earth
// vector, normalised phase, trapezoid weights and grid type
// (light_curve_comp.cpp:358-376, 425-455).
// Orbital phase of a data point: four Newton steps on
// t0 + phase (P + Pdot phase) = time, then the deltat term
// (light_curve_comp.cpp:358-376).
__device__ __forceinline__ double point_phase(const InstDev &I, double time) {
    double phase = (time - I.t0) / I.period;
    for (int it = 0; it < 4; it++)
        phase -= (I.t0 + phase * (I.period + I.pdot * phase) - time) / (I.period + 2. * I.pdot * phase);
    phase += I.deltat / I.period / 2. * (cos(2. * PI * phase) - 1.);
    return phase;
}
// Phase of sub-exposure nd of ndiv (light_curve_comp.cpp:425-455);
// expose in units of the period.
__device__ __forceinline__ double sub_phase(double phase, double expose, int ndiv, int nd) {
    if (ndiv == 1) return phase;
    return phase + expose * (nd - (ndiv - 1) / 2.0) / (ndiv - 1);
}

__global__ __launch_bounds__(256) void k_phases(const InstDev *__restrict__ insts, DataDev d, PhaseArrays pa) {
    const int np = blockIdx.x * blockDim.x + threadIdx.x;
    const int inst = blockIdx.y;
    if (np >= d.n) return;
    const InstDev &I = insts[inst];
    if (!I.active) return;
    const double time = d.time[np];
    const double phase = point_phase(I, time);
    const double frac = (time - d.middle) / d.range;
    const double slfac = 1. + frac * (I.slope + frac * (I.quad + frac * I.cube));
    const double expose = d.expose[np] / I.period;
    const int ndiv = d.ndiv[np];
    const double norm = (double)max(1, ndiv - 1);
    const size_t j0 = (size_t)inst * d.S + d.sub_off[np];
    for (int nd = 0; nd < ndiv; ++nd) {
        const double phi = sub_phase(phase, expose, ndiv, nd);
        const double wgt = (ndiv == 1) ? 1.0 : ((nd == 0 || nd == ndiv - 1) ? 0.5 : 1.0);
        V3 e = set_earth(I.cosi, I.sini, phi);
        const size_t j = j0 + nd;
        pa.ex[j] = e.x;
        pa.ey[j] = e.y;
        pa.ez[j] = e.z;
        pa.phin[j] = phi - floor(phi);
        pa.w1[j] = slfac * wgt * gscale1(I, phi) / norm;
        pa.w2[j] = slfac * wgt * gscale2(I, phi) / norm;
        pa.type[j] = (signed char)gint_type(phi, I.phase1, I.phase2);
    }
}

// Bins of [0, 1) holding a phase at which the instance's eclipse ranges
// are tested, with the arithmetic of k_phases (phin), of the Point-path
// sums (the Ginterp phases) and of the diagnostics (0.5).
__global__ __launch_bounds__(256) void k_phase_bits(const InstDev *__restrict__ insts, DataDev d, unsigned *bits,
                                                   int lb) {
    const int np = blockIdx.x * blockDim.x + threadIdx.x;
    const int inst = blockIdx.y;
    const InstDev &I = insts[inst];
    if (!I.active) return;
    unsigned *b = bits + ((size_t)inst << (lb - 5));
    const double B = (double)(1u << lb);
    auto mark = [&](double phi) {
        const double ph = phi - floor(phi);
        const int bin = min((int)(ph * B), (int)(1u << lb) - 1);
        atomicOr(&b[bin >> 5], 1u << (bin & 31));
    };
    if (np == 0) {
        for (int k = 0; k < 8; k++) mark(gint_phase(k, I));
        mark(0.5);
    }
    if (np >= d.n) return;
    const double phase = point_phase(I, d.time[np]);
    const double expose = d.expose[np] / I.period;
    const int ndiv = d.ndiv[np];
    for (int nd = 0; nd < ndiv; ++nd) mark(sub_phase(phase, expose, ndiv, nd));
}

// Fine/coarse sub-phase lists of one (instance, star) in ascending order
// (the PhaseBatch routing of light_curve_comp.cpp:448-453).  Only runs for
// stars whose coarse grid is not an alias of the fine one.
__global__ __launch_bounds__(1024) void k_route(const InstDev *__restrict__ insts, int S, PhaseArrays pa, int *lists,
                                               int *counts) {
    __shared__ int lds_w[16];
    const int star = blockIdx.x, inst = blockIdx.y;
    const InstDev &I = insts[inst];
    if (!I.active || (star == 0 ? I.copy1 : I.copy2)) return;
    int *fine = lists + ((size_t)inst * 2 + star) * 2 * S;
    int *coarse = fine + S;
    int nf = 0, nc = 0;
    for (int c = 0; c < S; c += 1024) {
        const int j = c + (int)threadIdx.x;
        bool isf = false;
        if (j < S) {
            int t = pa.type[(size_t)inst * S + j];
            isf = star == 0 ? (t == 1) : (t == 3);
        }
        int tot;
        int ex = block_count_excl<1024>(isf, lds_w, tot);
        if (j < S) {
            if (isf)
                fine[nf + ex] = j;
            else
                coarse[nc + ((int)threadIdx.x - ex)] = j;
        }
        nf += tot;
        nc += min(1024, S - c) - tot;
    }
    if (threadIdx.x == 0) {
        counts[(inst * 2 + star) * 2 + 0] = nf;
        counts[(inst * 2 + star) * 2 + 1] = nc;
    }
}

template <int NTH>
__device__ __forceinline__ double block_sum(double v, double *lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int s = NTH / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    double r = lds[0];
    __syncthreads();
    return r;
}

// Weighted sums back to data points in the reference's order (star 1 fine
// then coarse, star 2 fine then coarse, each ascending; then `third`),
// followed by re_scale or the fixed scale factor and chi**2.
// scal[inst*8]: 0 sfac, 1 chisq, 2 wnok, 3 chisq/wnok written.
// Scatter of the sub-phase sums onto the data points, third light,
// optimal scale and chi**2 (light_curve_comp.cpp:476-586, re_scale.cpp:16-43),
// in three kernels over blocks of SCAT_PTS or 4 SCAT_PTS points so that one large
// instance fills the GPU:
//   k_scatter_sum    calc before scaling, and per block the scale sums;
//   k_scatter_scale  the scale (every block sums the instance's partials in
//                    block order, so all agree), scaled calc, chi**2 sums;
//   k_scatter_final  the instance's chi**2 in block order and its scalars.
// The point's value: the weighted star sums (Ginterp choosing the fine or
// coarse grid per sub-phase) in the order star 1 fine, star 1 coarse,
// star 2 fine, star 2 coarse, plus third light (light_curve_comp.cpp:
// 476-498); with SPLIT the two stars apart and no third light, the iscale
// components (light_curve_comp.cpp:378-416).
// The disc, edge and bright spot at one data point: each one's
// sub-exposure sums with the trapezoid weights over max(1, ndiv - 1), their
// total times the slope factor (comp_light.cpp:322-407,
// light_curve_comp.cpp:483-497).
__device__ __forceinline__ double comp_point(const InstDev &I, const DataDev &d, const double *comp, int np) {
    const double frac = (d.time[np] - d.middle) / d.range;
    const double slfac = 1. + frac * (I.slope + frac * (I.quad + frac * I.cube));
    const int j0 = d.sub_off[np], ndiv = d.ndiv[np];
    double c[3];
    for (int kind = 0; kind < 3; kind++) {
        const double *sk = comp + I.comp_off + (size_t)kind * d.S + j0;
        double sum = 0.0;
        for (int nd = 0; nd < ndiv; nd++) {
            const double wgt = (ndiv == 1) ? 1.0 : ((nd == 0 || nd == ndiv - 1) ? 0.5 : 1.0);
            sum += wgt * sk[nd];
        }
        c[kind] = sum / max(1, ndiv - 1);
    }
    return slfac * (c[0] + c[1] + c[2]);
}

// The three component columns of the iscale fit at one data point: each
// one's sub-exposure sum as in comp_point, times the slope factor
// (light_curve_comp.cpp:390-403)
__device__ __forceinline__ void comp_cols(const InstDev &I, const DataDev &d, const double *comp, int np, double *c) {
    const double frac = (d.time[np] - d.middle) / d.range;
    const double slfac = 1. + frac * (I.slope + frac * (I.quad + frac * I.cube));
    const int j0 = d.sub_off[np], ndiv = d.ndiv[np];
    for (int kind = 0; kind < 3; kind++) {
        const double *sk = comp + I.comp_off + (size_t)kind * d.S + j0;
        double sum = 0.0;
        for (int nd = 0; nd < ndiv; nd++) {
            const double wgt = (ndiv == 1) ? 1.0 : ((nd == 0 || nd == ndiv - 1) ? 0.5 : 1.0);
            sum += wgt * sk[nd];
        }
        c[kind] = slfac * (sum / max(1, ndiv - 1));
    }
}

template <bool SPLIT>
__device__ __forceinline__ void point_sums(const InstDev &I, const DataDev &d, const PhaseArrays &pa, size_t base,
                                           const double *o1, const double *o2, const double *comp, int np,
                                           double &v1, double &v2) {
    const int j0 = d.sub_off[np], nd = d.ndiv[np];
    double v = 0.;
    for (int k = 0; k < nd; k++) {
        size_t j = j0 + k;
        if (I.copy1 || pa.type[base + j] == 1) v += pa.w1[base + j] * o1[j];
    }
    if (!I.copy1)
        for (int k = 0; k < nd; k++) {
            size_t j = j0 + k;
            if (pa.type[base + j] != 1) v += pa.w1[base + j] * o1[j];
        }
    if (SPLIT) {
        v1 = v;
        v = 0.;
    }
    for (int k = 0; k < nd; k++) {
        size_t j = j0 + k;
        if (I.copy2 || pa.type[base + j] == 3) v += pa.w2[base + j] * o2[j];
    }
    if (!I.copy2)
        for (int k = 0; k < nd; k++) {
            size_t j = j0 + k;
            if (pa.type[base + j] != 3) v += pa.w2[base + j] * o2[j];
        }
    if (SPLIT) {
        v2 = I.star2 ? v : 0.;  // star 2 is a component only when t2 > 0
    } else {
        if (I.has_comp) v += comp_point(I, d, comp, np);
        v1 = v + I.third;
    }
}

__global__ __launch_bounds__(256) void k_scatter_sum(const InstDev *__restrict__ insts, DataDev d, PhaseArrays pa,
                                                    const double *__restrict__ out, const double *__restrict__ comp,
                                                    double *calc, double *part, int pts) {
    __shared__ double lds[256];
    const int inst = blockIdx.y, nblk = gridDim.x;
    const InstDev &I = insts[inst];
    if (!I.active) return;
    const size_t base = (size_t)inst * d.S;
    const double *o1 = out + ((size_t)inst * 2 + 0) * d.S;
    const double *o2 = out + ((size_t)inst * 2 + 1) * d.S;
    double *c = calc + (size_t)inst * d.n;
    double sdy = 0., syy = 0., wn = 0., s12 = 0., s22 = 0., b2 = 0.;
    const bool five = I.iscale && I.has_comp;  // star 1, disc, edge, spot, star 2
    double g5[20];
#pragma unroll
    for (int q = 0; q < 20; q++) g5[q] = 0.;
    const int p1 = min(d.n, (int)(blockIdx.x + 1) * pts);
    for (int np = blockIdx.x * pts + threadIdx.x; np < p1; np += 256) {
        if (I.iscale) {
            // rows of the weighted design matrix: weight > 0 and
            // sigma = ferr / sqrt(weight) > 0 (svdfit.cpp:34-53)
            double v1, v2;
            point_sums<true>(I, d, pa, base, o1, o2, comp, np, v1, v2);
            c[np] = v1;
            if (d.weight[np] > 0.) {
                wn += d.weight[np];
                const double z = d.ferr[np] / sqrt(d.weight[np]);
                if (z > 0.) {
                    const double is = 1. / z, u1 = is * v1, u2 = is * v2, yb = is * d.flux[np];
                    sdy += u1 * yb;
                    syy += u1 * u1;
                    s12 += u1 * u2;
                    s22 += u2 * u2;
                    b2 += u2 * yb;
                    if (five) {
                        double cc[3];
                        comp_cols(I, d, comp, np, cc);
                        const double u[5] = {u1, is * cc[0], is * cc[1], is * cc[2], u2};
                        int q = 0;
#pragma unroll
                        for (int a = 0; a < 5; a++)
#pragma unroll
                            for (int b = a; b < 5; b++) g5[q++] += u[a] * u[b];
#pragma unroll
                        for (int a = 0; a < 5; a++) g5[15 + a] += u[a] * yb;
                    }
                }
            }
            continue;
        }
        double v, unused;
        point_sums<false>(I, d, pa, base, o1, o2, comp, np, v, unused);
        c[np] = v;
        if (d.weight[np] > 0.) {
            double wgt = d.weight[np] / sqr(d.ferr[np]);
            sdy += wgt * d.flux[np] * v;
            syy += wgt * v * v;
            wn += d.weight[np];
        }
    }
    if (!I.scale) return;
    sdy = block_sum<256>(sdy, lds);
    syy = block_sum<256>(syy, lds);
    wn = block_sum<256>(wn, lds);
    double *pp = part + ((size_t)inst * nblk + blockIdx.x) * SCAT_PARTS;
    if (I.iscale) {
        s12 = block_sum<256>(s12, lds);
        s22 = block_sum<256>(s22, lds);
        b2 = block_sum<256>(b2, lds);
        if (threadIdx.x == 0) {
            pp[5] = s12;
            pp[6] = s22;
            pp[7] = b2;
        }
        if (five)
            for (int q = 0; q < 20; q++) {
                const double t = block_sum<256>(g5[q], lds);
                if (threadIdx.x == 0) pp[8 + q] = t;
            }
    }
    if (threadIdx.x == 0) {
        pp[0] = sdy;
        pp[1] = syy;
        pp[2] = wn;
    }
}

// Least squares for the two star factors through the 2x2 Gram matrix of
// the weighted design columns: its eigenvalues are the squared singular
// values, so an eigenvalue below 1e-10 of the largest is a singular value
// below 1e-5 of the largest, which Subs::svdfit sets to zero
// (svdfit.cpp:58-64); the factors are then the sum over the kept
// eigenvectors v of v (v . g) / lambda, g = (b1, b2).  The disc, edge and
// spot columns are zero: their singular values are zero and their factors 0.
__device__ __forceinline__ void iscale_factors(double s11, double s12, double s22, double b1, double b2, double &a0,
                                               double &a4) {
    a0 = a4 = 0.;
    const double h = 0.5 * (s11 + s22), dd = 0.5 * (s11 - s22);
    const double l1 = h + sqrt(dd * dd + s12 * s12);
    if (!(l1 > 0.)) return;  // no rows, or both columns zero
    const double det = s11 * s22 - s12 * s12;
    const double l2 = fmax(0., det / l1);
    if (l2 > 0. && l2 >= 1e-10 * l1) {
        a0 = (s22 * b1 - s12 * b2) / det;
        a4 = (s11 * b2 - s12 * b1) / det;
        return;
    }
    // one singular value kept: the eigenvector of l1
    double v0 = s11 >= s22 ? l1 - s22 : s12, v1 = s11 >= s22 ? s12 : l1 - s11;
    const double nv = sqrt(v0 * v0 + v1 * v1);
    v0 /= nv;
    v1 /= nv;
    const double f = (v0 * b1 + v1 * b2) / l1;
    a0 = v0 * f;
    a4 = v1 * f;
}

// Least squares for the five iscale columns (star 1, disc, edge, spot,
// star 2) through their Gram matrix G (packed upper triangle) and g: cyclic
// Jacobi for its eigenvalues (the squared singular values of the weighted
// design matrix) and eigenvectors; directions with an eigenvalue below
// 1e-10 of the largest (singular value below 1e-5, svdfit.cpp:58-64) are
// dropped, the factors are the sum of v (v . g) / lambda over the rest.
__device__ void iscale_factors5(const double *G, const double *g, double *a) {
    double A[5][5], V[5][5];
    int q = 0;
    for (int i = 0; i < 5; i++)
        for (int j = i; j < 5; j++) A[i][j] = A[j][i] = G[q++];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) V[i][j] = i == j ? 1. : 0.;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0., dia = 0.;
        for (int i = 0; i < 5; i++) {
            dia += A[i][i] * A[i][i];
            for (int j = i + 1; j < 5; j++) off += A[i][j] * A[i][j];
        }
        if (!(off > 1e-32 * dia)) break;
        for (int p = 0; p < 4; p++)
            for (int r = p + 1; r < 5; r++) {
                if (A[p][r] == 0.) continue;
                const double th = (A[r][r] - A[p][p]) / (2. * A[p][r]);
                const double t = (th >= 0. ? 1. : -1.) / (fabs(th) + sqrt(th * th + 1.));
                const double c = 1. / sqrt(t * t + 1.), s = t * c;
                for (int k = 0; k < 5; k++) {  // A <- J^T A J
                    const double akp = A[k][p], akr = A[k][r];
                    A[k][p] = c * akp - s * akr;
                    A[k][r] = s * akp + c * akr;
                }
                for (int k = 0; k < 5; k++) {
                    const double apk = A[p][k], ark = A[r][k];
                    A[p][k] = c * apk - s * ark;
                    A[r][k] = s * apk + c * ark;
                }
                for (int k = 0; k < 5; k++) {
                    const double vkp = V[k][p], vkr = V[k][r];
                    V[k][p] = c * vkp - s * vkr;
                    V[k][r] = s * vkp + c * vkr;
                }
            }
    }
    double lmax = 0.;
    for (int i = 0; i < 5; i++) lmax = fmax(lmax, A[i][i]);
    for (int k = 0; k < 5; k++) a[k] = 0.;
    if (!(lmax > 0.)) return;
    for (int i = 0; i < 5; i++) {
        const double l = A[i][i];
        if (!(l > 0.) || l < 1e-10 * lmax) continue;
        double vg = 0.;
        for (int k = 0; k < 5; k++) vg += V[k][i] * g[k];
        const double f = vg / l;
        for (int k = 0; k < 5; k++) a[k] += V[k][i] * f;
    }
}

// The instance's factors (star 1 / all, star 2; with iscale and
// components also the disc, edge and spot factors a13) from the block
// partials, summed in block order.
__device__ __forceinline__ void inst_scale(const InstDev &I, const double *pi, int nblk, double &a0, double &a4,
                                           double *a13, double &wn, bool &want_chi) {
    a0 = I.sfac0;
    a4 = I.sfac4;
    a13[0] = I.sfac1;
    a13[1] = I.sfac2;
    a13[2] = I.sfac3;
    want_chi = I.rdata != 0;
    wn = 0.;
    if (!I.scale) return;
    double sdy = 0., syy = 0., s12 = 0., s22 = 0., b2 = 0.;
    for (int b = 0; b < nblk; b++) {
        const double *q = pi + b * SCAT_PARTS;
        sdy += q[0];
        syy += q[1];
        wn += q[2];
        if (I.iscale) {
            s12 += q[5];
            s22 += q[6];
            b2 += q[7];
        }
    }
    if (I.iscale && I.has_comp) {
        double G[20], a[5];
        for (int q = 0; q < 20; q++) G[q] = 0.;
        for (int b = 0; b < nblk; b++)
            for (int q = 0; q < 20; q++) G[q] += pi[b * SCAT_PARTS + 8 + q];
        iscale_factors5(G, G + 15, a);
        a0 = a[0];
        a13[0] = a[1];
        a13[1] = a[2];
        a13[2] = a[3];
        a4 = a[4];
        want_chi = true;
        return;
    }
    if (I.iscale) {
        iscale_factors(syy, s12, s22, sdy, b2, a0, a4);
        a13[0] = a13[1] = a13[2] = 0.;
        want_chi = true;  // light_curve_comp.cpp:539-547
        return;
    }
    a0 = a4 = wn > 0.0 ? sdy / syy : 1.;
    want_chi = wn > 0.0;
}

__global__ __launch_bounds__(256) void k_scatter_scale(const InstDev *__restrict__ insts, DataDev d, PhaseArrays pa,
                                                      const double *__restrict__ out, const double *__restrict__ comp,
                                                      double *calc, double *part, int pts) {
    __shared__ double lds[256];
    const int inst = blockIdx.y, nblk = gridDim.x;
    const InstDev &I = insts[inst];
    if (!I.active) return;
    double *pi = part + (size_t)inst * nblk * SCAT_PARTS;
    double a0, a4, wn, a13[3];
    bool want_chi;
    inst_scale(I, pi, nblk, a0, a4, a13, wn, want_chi);
    const size_t base = (size_t)inst * d.S;
    const double *o1 = out + ((size_t)inst * 2 + 0) * d.S;
    const double *o2 = out + ((size_t)inst * 2 + 1) * d.S;
    // iscale: sfac[1..3] multiply the zero disc, edge and spot columns
    // (0 after a fit; the caller's values otherwise, as written there)
    const double zc = I.scale ? 0. : I.zcols;
    double *c = calc + (size_t)inst * d.n;
    double chi = 0., wn2 = 0.;
    const int p1 = min(d.n, (int)(blockIdx.x + 1) * pts);
    for (int np = blockIdx.x * pts + threadIdx.x; np < p1; np += 256) {
        double v;
        if (I.iscale && I.has_comp) {
            // sfac[0] f0 + sfac[1] f1 + sfac[2] f2 + sfac[3] f3 (+ sfac[4] f4),
            // light_curve_comp.cpp:555-558
            double v1, v2, cc[3];
            point_sums<true>(I, d, pa, base, o1, o2, comp, np, v1, v2);
            comp_cols(I, d, comp, np, cc);
            v = a0 * v1 + a13[0] * cc[0] + a13[1] * cc[1] + a13[2] * cc[2];
            if (I.star2) v += a4 * v2;
        } else if (I.iscale) {
            double v1, v2;
            point_sums<true>(I, d, pa, base, o1, o2, comp, np, v1, v2);
            v = a0 * v1 + zc;
            if (I.star2) v += a4 * v2;
        } else {
            v = c[np];
            if (!I.scale || wn > 0.0) v *= a0;
        }
        c[np] = v;
        if (want_chi && d.weight[np] > 0.) {
            chi += d.weight[np] * sqr((d.flux[np] - v) / d.ferr[np]);
            wn2 += d.weight[np];
        }
    }
    chi = block_sum<256>(chi, lds);
    wn2 = block_sum<256>(wn2, lds);
    if (threadIdx.x == 0) {
        pi[blockIdx.x * SCAT_PARTS + 3] = chi;
        pi[blockIdx.x * SCAT_PARTS + 4] = wn2;
    }
}

__global__ void k_scatter_final(const InstDev *__restrict__ insts, int nblk, const double *part, double *scal) {
    const int inst = blockIdx.x;
    const InstDev &I = insts[inst];
    if (!I.active || threadIdx.x != 0) return;
    const double *pi = part + (size_t)inst * nblk * SCAT_PARTS;
    double a0, a4, wn, a13[3];
    bool want_chi;
    inst_scale(I, pi, nblk, a0, a4, a13, wn, want_chi);
    double chi = 0., wn2 = 0.;
    for (int b = 0; b < nblk; b++) {
        chi += pi[b * SCAT_PARTS + 3];
        wn2 += pi[b * SCAT_PARTS + 4];
    }
    double *s = scal + (size_t)inst * 8;
    s[0] = a0;
    s[4] = a4;
    s[5] = a13[0];
    s[6] = a13[1];
    s[7] = a13[2];
    if (I.scale && !I.iscale) {
        s[1] = chi;
        s[2] = wn;
        s[3] = 1.;
    } else {
        s[1] = chi;
        s[2] = wn2;
        s[3] = want_chi ? 1. : 0.;
    }
}

__global__ __launch_bounds__(256) void k_regions(ByteRegions R) {
    const int r = blockIdx.y;
    const size_t n4 = R.n[r] >> 2;
    unsigned *d = static_cast<unsigned *>(R.dst[r]);
    const unsigned *src = static_cast<const unsigned *>(R.src[r]);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        d[i] = src ? src[i] : 0u;
}

}  // namespace

void launch_regions(hipStream_t s, const ByteRegions &r, int nreg) {
    if (nreg <= 0) return;
    size_t mx = 0;
    for (int k = 0; k < nreg; k++) mx = std::max(mx, r.n[k]);
    const int bx = (int)std::max<size_t>(1, std::min<size_t>(4096, (mx / 4 + 255) / 256));
    hipLaunchKernelGGL(k_regions, dim3(bx, nreg), dim3(256), 0, s, r);
}

void launch_phases(hipStream_t s, const InstDev *insts, int ninst, DataDev d, PhaseArrays pa) {
    if (ninst <= 0 || d.n <= 0) return;
    hipLaunchKernelGGL(k_phases, dim3((d.n + 255) / 256, ninst), dim3(256), 0, s, insts, d, pa);
}

void launch_phase_bits(hipStream_t s, const InstDev *insts, int ninst, DataDev d, unsigned *bits, int lb) {
    if (ninst <= 0) return;
    hipLaunchKernelGGL(k_phase_bits, dim3((std::max(d.n, 1) + 255) / 256, ninst), dim3(256), 0, s, insts, d, bits, lb);
}

void launch_route(hipStream_t s, const InstDev *insts, int ninst, int S, PhaseArrays pa, int *lists, int *counts) {
    if (ninst <= 0) return;
    hipLaunchKernelGGL(k_route, dim3(2, ninst), dim3(1024), 0, s, insts, S, pa, lists, counts);
}

void launch_scatter(hipStream_t s, const InstDev *insts, int ninst, DataDev d, PhaseArrays pa, const double *out,
                    const double *comp, double *calc, double *scal, double *part) {
    if (ninst <= 0) return;
    // four points per thread, or one where that leaves the device short of
    // workgroups (small calls: the blocks' sequential loops set the time)
    const int pts = (long long)((d.n + 4 * SCAT_PTS - 1) / (4 * SCAT_PTS)) * ninst >= 1024 ? 4 * SCAT_PTS : SCAT_PTS;
    const int nblk = std::max(1, (d.n + pts - 1) / pts);
    hipLaunchKernelGGL(k_scatter_sum, dim3(nblk, ninst), dim3(256), 0, s, insts, d, pa, out, comp, calc, part, pts);
    hipLaunchKernelGGL(k_scatter_scale, dim3(nblk, ninst), dim3(256), 0, s, insts, d, pa, out, comp, calc, part, pts);
    hipLaunchKernelGGL(k_scatter_final, dim3(ninst), dim3(64), 0, s, insts, nblk, part, scal);
}

}  // namespace lck
