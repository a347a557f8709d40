// k_phase.hip -- orbital phases and sub-exposures, fine/coarse routing,
// scatter of weighted sums back to data points, scale factor and chi**2
// (light_curve_comp.cpp:350-586, re_scale.cpp:16-43).
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
// in three kernels over blocks of SCAT_PTS points so that one large
// instance fills the GPU:
//   k_scatter_sum    calc before scaling, and per block the scale sums;
//   k_scatter_scale  the scale (every block sums the instance's partials in
//                    block order, so all agree), scaled calc, chi**2 sums;
//   k_scatter_final  the instance's chi**2 in block order and its scalars.
__device__ __forceinline__ double point_value(const InstDev &I, const DataDev &d, const PhaseArrays &pa, size_t base,
                                              const double *o1, const double *o2, int np) {
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
    for (int k = 0; k < nd; k++) {
        size_t j = j0 + k;
        if (I.copy2 || pa.type[base + j] == 3) v += pa.w2[base + j] * o2[j];
    }
    if (!I.copy2)
        for (int k = 0; k < nd; k++) {
            size_t j = j0 + k;
            if (pa.type[base + j] != 3) v += pa.w2[base + j] * o2[j];
        }
    return v + I.third;
}

__global__ __launch_bounds__(256) void k_scatter_sum(const InstDev *__restrict__ insts, DataDev d, PhaseArrays pa,
                                                    const double *__restrict__ out, double *calc, double *part) {
    __shared__ double lds[256];
    const int inst = blockIdx.y, nblk = gridDim.x;
    const InstDev &I = insts[inst];
    if (!I.active) return;
    const size_t base = (size_t)inst * d.S;
    const double *o1 = out + ((size_t)inst * 2 + 0) * d.S;
    const double *o2 = out + ((size_t)inst * 2 + 1) * d.S;
    double *c = calc + (size_t)inst * d.n;
    double sdy = 0., syy = 0., wn = 0.;
    const int p1 = min(d.n, (int)(blockIdx.x + 1) * SCAT_PTS);
    for (int np = blockIdx.x * SCAT_PTS + threadIdx.x; np < p1; np += 256) {
        const double v = point_value(I, d, pa, base, o1, o2, np);
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
    if (threadIdx.x == 0) {
        double *pp = part + ((size_t)inst * nblk + blockIdx.x) * 5;
        pp[0] = sdy;
        pp[1] = syy;
        pp[2] = wn;
    }
}

// the instance's scale from the block partials, summed in block order
__device__ __forceinline__ void inst_scale(const InstDev &I, const double *pi, int nblk, double &scale, double &wn,
                                           bool &want_chi) {
    scale = I.sfac0;
    want_chi = I.rdata != 0;
    wn = 0.;
    if (I.scale) {
        double sdy = 0., syy = 0.;
        for (int b = 0; b < nblk; b++) {
            sdy += pi[b * 5 + 0];
            syy += pi[b * 5 + 1];
            wn += pi[b * 5 + 2];
        }
        scale = wn > 0.0 ? sdy / syy : 1.;
        want_chi = wn > 0.0;
    }
}

__global__ __launch_bounds__(256) void k_scatter_scale(const InstDev *__restrict__ insts, DataDev d, double *calc,
                                                      double *part) {
    __shared__ double lds[256];
    const int inst = blockIdx.y, nblk = gridDim.x;
    const InstDev &I = insts[inst];
    if (!I.active) return;
    double *pi = part + (size_t)inst * nblk * 5;
    double scale, wn;
    bool want_chi;
    inst_scale(I, pi, nblk, scale, wn, want_chi);
    double *c = calc + (size_t)inst * d.n;
    double chi = 0., wn2 = 0.;
    const int p1 = min(d.n, (int)(blockIdx.x + 1) * SCAT_PTS);
    for (int np = blockIdx.x * SCAT_PTS + threadIdx.x; np < p1; np += 256) {
        double v = c[np];
        if (!I.scale || wn > 0.0) v *= scale;
        c[np] = v;
        if (want_chi && d.weight[np] > 0.) {
            chi += d.weight[np] * sqr((d.flux[np] - v) / d.ferr[np]);
            wn2 += d.weight[np];
        }
    }
    chi = block_sum<256>(chi, lds);
    wn2 = block_sum<256>(wn2, lds);
    if (threadIdx.x == 0) {
        pi[blockIdx.x * 5 + 3] = chi;
        pi[blockIdx.x * 5 + 4] = wn2;
    }
}

__global__ void k_scatter_final(const InstDev *__restrict__ insts, int nblk, const double *part, double *scal) {
    const int inst = blockIdx.x;
    const InstDev &I = insts[inst];
    if (!I.active || threadIdx.x != 0) return;
    const double *pi = part + (size_t)inst * nblk * 5;
    double scale, wn;
    bool want_chi;
    inst_scale(I, pi, nblk, scale, wn, want_chi);
    double chi = 0., wn2 = 0.;
    for (int b = 0; b < nblk; b++) {
        chi += pi[b * 5 + 3];
        wn2 += pi[b * 5 + 4];
    }
    double *s = scal + (size_t)inst * 8;
    s[0] = scale;
    if (I.scale) {
        s[1] = chi;
        s[2] = wn;
        s[3] = 1.;
    } else {
        s[1] = chi;
        s[2] = wn2;
        s[3] = want_chi ? 1. : 0.;
    }
}

}  // namespace

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
                    double *calc, double *scal, double *part) {
    if (ninst <= 0) return;
    const int nblk = std::max(1, (d.n + SCAT_PTS - 1) / SCAT_PTS);
    hipLaunchKernelGGL(k_scatter_sum, dim3(nblk, ninst), dim3(256), 0, s, insts, d, pa, out, calc, part);
    hipLaunchKernelGGL(k_scatter_scale, dim3(nblk, ninst), dim3(256), 0, s, insts, d, calc, part);
    hipLaunchKernelGGL(k_scatter_final, dim3(ninst), dim3(64), 0, s, insts, nblk, part, scal);
}

}  // namespace lck
