// k_geom.hip -- face grid, eclipse search, continuum, FlatGrid build,
// Ginterp and diagnostic reductions (FP64).
//
// Face grid: one work item per half-ring face solves the radius (the ring
// solve of face.cpp:71-161, falling back to the scalar Roche::face when it
// fails, set_star_grid.cpp:455-471) and writes the face and its mirror
// partner (set_star_grid.cpp:493-510).  Faces whose line of sight can cross
// the companion's reference sphere are appended to a candidate list; the
// eclipse kernel then runs the FP64 ingress/egress search
// (ingress_egress.cpp:35-97) over that compacted list only.
#include <algorithm>

#include "lc_comp.h"
#include "lc_kernels.h"
#include "lc_math.h"

namespace lck {
using namespace lcx;

namespace {

__device__ __forceinline__ void set_status(int *inst_status, int inst, int st) {
    if (st) atomicCAS(&inst_status[inst], 0, st);
}

// Last index i < n with key(i) <= t, for keys ascending in i and t
// ascending across the active lanes of a wave: one binary search on the
// wave's first t (wave-uniform, so scalar loads) and a short forward scan
// per lane -- a wave's 64 consecutive items span only a few rings or grids.
template <class Key>
__device__ __forceinline__ int wave_lookup(int n, int t, Key key) {
    const int t0 = __builtin_amdgcn_readfirstlane(t);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(mid) <= t0)
            lo = mid;
        else
            hi = mid - 1;
    }
    int r = lo;
    while (r + 1 < n && key(r + 1) <= t) r++;
    return r;
}

// Eclipse of one face by the companion (set_star_grid.cpp:372-386).  For a
// Roche companion only the reference-sphere test runs here: true makes the
// face a candidate for the full search.
__device__ __forceinline__ bool eclipse_pretest(const GridDev &G, const V3 &pos, int face, FaceArrays fa) {
    fa.ecl[face] = 0;
    fa.ein[face] = 0.;
    fa.eout[face] = 0.;
    if (!G.eclipse) return false;
    const V3 cofm_other = G.star == PRIMARY ? v3(1., 0., 0.) : v3(0., 0., 0.);
    double phi1, phi2, lam1, lam2;
    if (G.roche_other) return sphere_eclipse_test(G.cosi, G.sini, pos, cofm_other, G.rref_other);
    if (sphere_eclipse(G.cosi, G.sini, pos, cofm_other, G.r_other, phi1, phi2, lam1, lam2)) {
        fa.ecl[face] = 1;
        fa.ein[face] = phi1;
        fa.eout[face] = phi2;
    }
    return false;
}

#ifndef LC_FACES_NT
#define LC_FACES_NT 256  // 1024 (one list atomic per 2048 faces) halves the occupancy: 0.94 vs 0.66 ms
#endif
__global__ __launch_bounds__(LC_FACES_NT) void k_faces(const GridDev *__restrict__ grids, const RingDev *__restrict__ rings,
                                              int nring, int nhalf_total, FaceArrays fa, int2 *cand, int *cand_count,
                                              int *inst_status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // eclipse candidates of this item (the face and its mirror partner),
    // appended below with one atomic per workgroup: a per-face atomic on the
    // one counter serialises the whole launch
    bool c1 = false, c2 = false;
    int f1 = 0, f2 = 0, gidx = 0;
    if (t < nhalf_total) {
        const int lo = wave_lookup(nring, t, [&](int k) { return rings[k].half_off; });
        const RingDev R = rings[lo];
        const GridDev &G = grids[R.grid];
        const int np = t - R.half_off;
        const int nphi = R.nphi;
        double phi = R.phi1 + R.dphi_span * (np + 0.5) / nphi;
        double sinp, cosp;
        sincos(phi, &sinp, &cosp);
        V3 dirn = G.npole ? v3(R.sint * cosp, R.sint * sinp, R.cost) : v3(R.cost, R.sint * cosp, R.sint * sinp);
        const V3 cofm = G.star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
        V3 posn, dvec;
        double rad, gravity;
        if (G.roche_self) {
            RayPot rp = make_raypot(G.q, G.star, G.spin_self, dirn, false);
            int fail = solve_radius(rp, G.rref_self, G.pref_self, G.acc, rad);
            if (fail) {
                int st = face_scalar(G.q, G.star, G.spin_self, dirn, G.rref_self, G.pref_self, G.acc, posn, dvec, rad,
                                     gravity);
                if (st) {
                    set_status(inst_status, G.inst, ST_FACE);
                    posn = add(cofm, scl(G.rref_self, dirn));
                    dvec = dirn;
                    rad = G.rref_self;
                    gravity = 1.;
                }
            } else {
                posn = add(cofm, scl(rad, dirn));
                dvec = drpot_star(G.star, G.q, G.spin_self, posn);
                gravity = len(dvec);
                dvec = v3(dvec.x / gravity, dvec.y / gravity, dvec.z / gravity);
            }
        } else {
            rad = G.r_self;
            posn = add(cofm, scl(rad, dirn));
            dvec = dirn;
            gravity = 1.;
        }
        double area = ((R.dphi_span / nphi * rad * R.sint) * (R.tspan_nl * rad)) / dot(dirn, dvec);
        const float farea = (float)area, fgrav = (float)(gravity / G.gref);

        // faces past numface() are not part of the star (see plan_grid)
        const int fend = G.face_off + G.nface;
        int f = R.face_off + np;
        if (f < fend) {
            fa.px[f] = posn.x; fa.py[f] = posn.y; fa.pz[f] = posn.z;
            fa.dx[f] = dvec.x; fa.dy[f] = dvec.y; fa.dz[f] = dvec.z;
            fa.area[f] = farea;
            fa.grav[f] = fgrav;
            fa.flux[f] = 0.f;
            c1 = eclipse_pretest(G, posn, f, fa);
            f1 = f;
        }
        gidx = R.grid;

        int np2 = nphi - 1 - np;
        if (np2 != np && R.face_off + np2 < fend) {
            V3 posn2 = posn, dvec2 = dvec;
            if (G.npole) {
                posn2.y = -posn.y;
                dvec2.y = -dvec.y;
            } else {
                posn2.z = -posn.z;
                dvec2.z = -dvec.z;
            }
            f2 = R.face_off + np2;
            fa.px[f2] = posn2.x; fa.py[f2] = posn2.y; fa.pz[f2] = posn2.z;
            fa.dx[f2] = dvec2.x; fa.dy[f2] = dvec2.y; fa.dz[f2] = dvec2.z;
            fa.area[f2] = farea;
            fa.grav[f2] = fgrav;
            fa.flux[f2] = 0.f;
            c2 = eclipse_pretest(G, posn2, f2, fa);
        }
    }
    // candidates in lane order within the workgroup; workgroups take their
    // ranges in arrival order (the search's results do not depend on it)
    constexpr int NW = LC_FACES_NT / 64;
    __shared__ int s_w[NW], s_base;
    const unsigned long long b1 = __ballot(c1), b2 = __ballot(c2);
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (lane == 0) s_w[w] = __popcll(b1) + __popcll(b2);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int v = 0; v < NW; v++) tot += s_w[v];
        s_base = tot ? atomicAdd(cand_count, tot) : 0;
    }
    __syncthreads();
    int k = s_base + __popcll(b1 & below) + __popcll(b2 & below);
    for (int v = 0; v < w; v++) k += s_w[v];
    if (c1) cand[k++] = make_int2(f1, gidx);  // the search needs the face's grid
    if (c2) cand[k] = make_int2(f2, gidx);
}

// Eclipse pre-test of component elements (disc, edge, bright spot): every
// element sits in two slots, one per eclipsing star, whose pseudo-grids
// name that star as the companion; candidates go to the search as star
// faces do (star_eclipse.cpp:15-25 for each star in turn).
__global__ __launch_bounds__(256) void k_comp_pretest(const GridDev *__restrict__ grids, int ngrid, int nface,
                                                      FaceArrays fa, int2 *cand, int *cand_count) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    bool c = false;
    int g = 0;
    if (f < nface) {
        g = wave_lookup(ngrid, f, [&](int k) { return grids[k].face_off; });
        c = eclipse_pretest(grids[g], v3(fa.px[f], fa.py[f], fa.pz[f]), f, fa);
    }
    __shared__ int s_w[4], s_base;
    const unsigned long long b = __ballot(c);
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    if (lane == 0) s_w[w] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        s_base = tot ? atomicAdd(cand_count, tot) : 0;
    }
    __syncthreads();
    int k = s_base + __popcll(b & ((1ull << lane) - 1ull));
    for (int v = 0; v < w; v++) k += s_w[v];
    if (c) cand[k] = make_int2(f, g);
}

// FP64 ingress/egress search (ingress_egress.cpp:35-97) over the compacted
// candidates, in three kernels so that each keeps little live state:
//   k_eclipse_potmin  reference-sphere window and pot_min's first point for
//                     every candidate, then the full pot_min for the ones
//                     that need it (compacted); hits go to a second list;
//   k_eclipse_bisect  one work item per (hit, side): the ingress or egress
//                     bisection with fblink at every step;
//   k_eclipse_final   normalises the phases and marks the face eclipsed.
// Counts live on the device, so the chain needs no host round trip; every
// grid-stride loop ends once its list is consumed.
// Pass FIRST (every candidate): the sphere window and pot_min's first point
// (pot_min_start); a potential there at or below the surface value is a hit
// with pot_min's own result, the other candidates go to a list.  Second
// pass (that list, compacted, so the long searches share waves): the full
// pot_min.  Hits and list entries are appended with one atomic per
// workgroup and pass.
template <bool FIRST, int NT>
__global__ __launch_bounds__(NT) void k_eclipse_potmin(const GridDev *__restrict__ grids, FaceArrays fa,
                                                       const int2 *__restrict__ cand,
                                                       const int *__restrict__ cand_count, EclipseWork ew,
                                                       int *inst_status) {
    const int n = FIRST ? *cand_count : ew.acount[ECL_ROUNDS];
    int *more_list = ew.alist[1];
    constexpr int NW = NT / 64;
    __shared__ int s_w[2][NW], s_base[2];
    // the loop's trip count is the same for the whole workgroup (its
    // barriers), items past n are inactive
    for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const int t = i0 + (int)threadIdx.x;
        bool hit = false, more = false;
        int ci = 0, f = 0, gidx = 0;
        double phi1, phi2, lam2, lam1, phi, lam, ACC;
        if (t < n) {
            ci = FIRST ? t : more_list[t];
            const int2 c = cand[ci];
            f = c.x;
            gidx = c.y;
            const GridDev &G = grids[gidx];
            const int other = G.star == PRIMARY ? SECONDARY : PRIMARY;
            const V3 cofm = other == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
            const V3 r = v3(fa.px[f], fa.py[f], fa.pz[f]);
            if (FIRST) {
                fa.ecl[f] = 0;
                fa.ein[f] = 0.;
                fa.eout[f] = 0.;
            }
            if (sphere_eclipse(G.cosi, G.sini, r, cofm, G.rref_other, phi1, phi2, lam1, lam2)) {
                ACC = 2. * sqrt(2. * TWOPI * (lam2 - lam1) * G.delta);
                if (FIRST) {
                    V3 earth, rr;
                    const double pot = pot_min_start(G.q, G.cosi, G.sini, other, G.spin_other, r, phi1, phi2, lam1,
                                                     lam2, phi, lam, earth, rr);
                    hit = pot <= G.pref_other;
                    more = !hit;
                } else {
                    const int st = pot_min(G.q, G.cosi, G.sini, other, G.spin_other, r, phi1, phi2, lam1, lam2,
                                           G.rref_other, G.pref_other, ACC, phi, lam, hit);
                    if (st) {
                        set_status(inst_status, G.inst, st);
                        hit = false;
                    }
                }
            }
        }
        const unsigned long long bh = __ballot(hit), bm = FIRST ? __ballot(more) : 0ull;
        const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (lane == 0) {
            s_w[0][w] = __popcll(bh);
            s_w[1][w] = __popcll(bm);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int th = 0, tm = 0;
            for (int v = 0; v < NW; v++) {
                th += s_w[0][v];
                tm += s_w[1][v];
            }
            s_base[0] = th ? atomicAdd(ew.hit_count, th) : 0;
            s_base[1] = tm ? atomicAdd(&ew.acount[ECL_ROUNDS], tm) : 0;
        }
        __syncthreads();
        int bh0 = s_base[0], bm0 = s_base[1];
        for (int v = 0; v < w; v++) {
            bh0 += s_w[0][v];
            bm0 += s_w[1][v];
        }
        if (hit) {
            const int k = bh0 + __popcll(bh & below);
            ew.face[k] = f;
            ew.grid[k] = gidx;
            ew.phi[k] = phi;
            ew.lim[2 * k] = phi1;
            ew.lim[2 * k + 1] = phi2;
            ew.acc[k] = ACC;
        }
        if (FIRST && more) more_list[bm0 + __popcll(bm & below)] = ci;
        __syncthreads();  // s_w and s_base are rewritten by the next pass
    }
}

// True when no phase at which the instance's eclipse ranges are tested can
// lie in the bisection bracket [lo, hi] (raw phases).  Every test is a
// comparison of such a phase with the normalised ingress or egress
// (ecl_range, comp_light_flat.cpp:171-173; Point::visible), and the
// normalisation x - floor(x) and the egress + 1 shift are monotone, so any
// two values inside a clear bracket give the same answer to every test: the
// search may end there instead of halving on to delta_phase.  Conservative:
// the bracket may not contain an integer, margins of 1e-12 cover the
// roundings of the normalisation and of the shift, and brackets wider than
// 64 bins are not looked at.
__device__ __forceinline__ bool bracket_clear(const unsigned *bits, int lb, double lo, double hi) {
    const double F = floor(lo);
    if (floor(hi) != F) return false;
    const double a = (lo - F) - 1e-12, b = (hi - F) + 1e-12;
    if (!(a >= 0.0) || !(b < 1.0)) return false;
    const double B = (double)(1u << lb);
    if ((b - a) * B > 64.0) return false;
    const int b0 = (int)(a * B), b1 = (int)(b * B);
    for (int w = b0 >> 5; w <= (b1 >> 5); w++) {
        unsigned m = bits[w];
        if (w == (b0 >> 5)) m &= ~0u << (b0 & 31);
        if (w == (b1 >> 5)) m &= ~0u >> (31 - (b1 & 31));
        if (m) return false;
    }
    return true;
}

// The visibility test of an early-stopping search (see k_eclipse_bisect):
// FP32 direction, the phase of its largest mu and the threshold.
struct SeenTest {
    float dx, dy, dz, ph_top, thr;
    bool never;  // never seen at any phase
    __device__ void init(const FaceArrays &fa, int f, const GridDev &G) {
        const double ddx = fa.dx[f], ddy = fa.dy[f], ddz = fa.dz[f];
        thr = (float)G.mucrit - 1e-5f;
        never = G.sini * sqrt(ddx * ddx + ddy * ddy) + ddz * G.cosi <= (double)thr;
        ph_top = (float)(atan2(-ddy, ddx) / TWOPI);
        dx = (float)ddx;
        dy = (float)ddy;
        dz = (float)ddz;
    }
    __device__ float mu(const V3 &e) const { return fmaf(dz, (float)e.z, fmaf(dy, (float)e.y, dx * (float)e.x)); }
    // no phase of [min(a, b), max(a, b)] shows the face (mu_a, mu_b: mu at a, b)
    __device__ bool unseen(double a, double b, float mu_a, float mu_b) const {
        if (never) return true;
        const double lo = fmin(a, b), hi = fmax(a, b);
        if (fmaxf(mu_a, mu_b) > thr) return false;
        return (double)ph_top + ceil(lo - (double)ph_top) > hi;  // the maximum is not inside
    }
};

// Ahead of round 0 of calls with early stops: every search (items
// 2*hit + side) whose first bracket [phi, phi1 or phi2] never shows its face
// ends here with the bracket's midpoint; the others go to the list round 0
// runs over (alist[1], count acount[ECL_ROUNDS + 1]).  A workgroup takes
// OPEN_NT * OPEN_PER consecutive searches per pass and appends its open
// ones with one atomic (appends of one atomic per wave on the one counter
// serialised this kernel: 230k atomics, 1.6 ms on the batched workload).
constexpr int OPEN_NT = 1024, OPEN_PER = 4;
__global__ __launch_bounds__(OPEN_NT) void k_eclipse_open(const GridDev *__restrict__ grids, FaceArrays fa,
                                                          EclipseWork ew) {
    constexpr int NW = OPEN_NT / 64;
    __shared__ int s_cnt[OPEN_PER * NW];
    __shared__ int s_base;
    const int n = *ew.hit_count * 2;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    for (int i0 = blockIdx.x * OPEN_NT * OPEN_PER; i0 < n; i0 += gridDim.x * OPEN_NT * OPEN_PER) {
        bool open[OPEN_PER];
        unsigned long long m[OPEN_PER];
#pragma unroll
        for (int p = 0; p < OPEN_PER; p++) {
            const int i = i0 + p * OPEN_NT + (int)threadIdx.x;
            open[p] = false;
            if (i < n) {
                const int k = i >> 1, f = ew.face[k];
                const GridDev &G = grids[ew.grid[k]];
                open[p] = true;
                if (G.ecl_stop) {
                    SeenTest v;
                    v.init(fa, f, G);
                    const double a = ew.phi[k], b = ew.lim[i];
                    if (v.unseen(a, b, v.mu(set_earth(G.cosi, G.sini, a)), v.mu(set_earth(G.cosi, G.sini, b)))) {
                        const double x = (a + b) / 2.;
                        ew.res[i] = x - floor(x);
                        open[p] = false;
                    }
                }
            }
            m[p] = __ballot(open[p]);
            if (lane == 0) s_cnt[p * NW + w] = __popcll(m[p]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // exclusive prefix in (pass, wave) order, one atomic
            int t = 0;
            for (int k = 0; k < OPEN_PER * NW; k++) {
                const int c = s_cnt[k];
                s_cnt[k] = t;
                t += c;
            }
            s_base = t ? atomicAdd(&ew.acount[ECL_ROUNDS + 1], t) : 0;
        }
        __syncthreads();
        const int base = s_base;
#pragma unroll
        for (int p = 0; p < OPEN_PER; p++) {
            const int i = i0 + p * OPEN_NT + (int)threadIdx.x;
            if (open[p]) {
                const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m[p] >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)m[p], 0u));
                ew.alist[1][base + s_cnt[p * NW + w] + below] = i;
            }
        }
        __syncthreads();  // s_cnt and s_base are rewritten by the next pass
    }
}

#ifndef LC_FIRST_NT
#define LC_FIRST_NT 512  // workgroup size of the first pot_min pass (4x fewer list atomics than 128)
#endif
#ifndef LC_ECL_WAVES
#define LC_ECL_WAVES 3  // up to 168 VGPRs: three waves per SIMD, no spills
#endif
// One round of the ingress/egress bisections: every open search (the
// items 2*hit + side; all of them in round 0) takes up to nstep more steps;
// a finished search stores its result, an open one its bracket and its
// index in the next round's list.  Splitting the searches into rounds keeps
// lanes whose search has ended from idling until the longest search of
// their wave ends.
__global__ __launch_bounds__(128, LC_ECL_WAVES) void k_eclipse_bisect(const GridDev *__restrict__ grids, int ngrid,
                                                                     FaceArrays fa, EclipseWork ew,
                                                                     int *inst_status, int round, int nstep,
                                                                     int open_list) {
    // round 0 runs over every search, or over the list k_eclipse_open left
    const int *in = round == 0 ? (open_list ? ew.alist[1] : nullptr) : ew.alist[(round - 1) & 1];
    const int n = round == 0 ? (open_list ? ew.acount[ECL_ROUNDS + 1] : *ew.hit_count * 2) : ew.acount[round - 1];
    int *out = ew.alist[round & 1];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int i = in ? in[t] : t;
        const int k = i >> 1;
        const int f = ew.face[k];
        const GridDev &G = grids[ew.grid[k]];
        // the grid's constants in registers for the whole search
        const int other = G.star == PRIMARY ? SECONDARY : PRIMARY;
        const double q = G.q, spin = G.spin_other, cosi = G.cosi, sini = G.sini, delta = G.delta;
        const double rref = G.rref_other, pref = G.pref_other;
        const V3 r = v3(fa.px[f], fa.py[f], fa.pz[f]);
        const double ACC = ew.acc[k];
        double pin = round > 0 ? ew.b_in[i] : ew.phi[k], pout = round > 0 ? ew.b_out[i] : ew.lim[i];
        const unsigned *bits = G.ecl_stop ? ew.phase_bits + ((size_t)G.inst << (ew.bits_lb - 5)) : nullptr;
        // The face's visibility mu(phase) = d.e(phase) (the same early-stop
        // scope): where mu <= mucrit the reference's elements are zero
        // whatever the ranges say (comp_light_flat.cpp:110-168, Point path
        // likewise), so a search over whose bracket the face is never seen
        // can stop with any value inside it: the eclipsed phases (the
        // projection of [ingress, egress] onto [0, 1), which the
        // normalisation and the egress + 1 shift represent) then differ from
        // the reference's only inside the bracket.  mu = sini |d_xy|
        // cos(2 pi (phase - ph_top)) + d_z cosi has one maximum per orbit,
        // at ph_top; over a bracket without it the largest mu is at an end.
        // Kept in FP32: the 1e-5 margin covers that rounding, the FP32 mu of
        // the mixed flux kernels, and a maximum misplaced by the FP32 ph_top
        // (mu is flat there: ~1e-13).
        SeenTest vis;
        float mu_in = 0.f, mu_out = 0.f;
        if (bits) {
            vis.init(fa, f, G);
            mu_in = vis.mu(set_earth(cosi, sini, pin));
            mu_out = vis.mu(set_earth(cosi, sini, pout));
        }
        int st = ST_OK;
        bool done = false;
        for (int s = 0; s < nstep; s++) {
            if (!(fabs(pin - pout) > delta)) break;
            const double pmid = (pin + pout) / 2.;
            const V3 em = set_earth(cosi, sini, pmid);
            bool h;
            st = fblink(q, other, spin, ACC, em, r, rref, pref, h);
            if (st) break;
            const float mu_mid = bits ? vis.mu(em) : 0.f;
            if (h) {
                pin = pmid;
                mu_in = mu_mid;
            } else {
                pout = pmid;
                mu_out = mu_mid;
            }
            if (bits && (bracket_clear(bits, ew.bits_lb, fmin(pin, pout), fmax(pin, pout)) ||
                         vis.unseen(pin, pout, mu_in, mu_out))) {
                done = true;
                break;
            }
        }
        if (st) {
            set_status(inst_status, G.inst, st);
            ew.res[i] = 0.;
        } else if (done || !(fabs(pin - pout) > delta)) {
            const double x = (pin + pout) / 2.;
            ew.res[i] = x - floor(x);
        } else {
            ew.b_in[i] = pin;
            ew.b_out[i] = pout;
            out[atomicAdd(&ew.acount[round], 1)] = i;
        }
    }
}

// The bisections of calls too small to fill the device (no early stops, one
// round): there the longest chain of fblink calls sets the time.  GL lanes
// share a search.  Each step evaluates fblink at the GL - 1 points the next
// LV = log2(GL) bisection steps can visit, in heap order (node 1: the
// midpoint; node j's children: 2j where the point is not hidden, so pout
// moves, and 2j + 1 where it is, so pin moves), one point per lane; then
// every lane of the group replays the LV steps from the ballots.  A node's
// point is formed by the same FP64 midpoints as in the one-lane loop, and
// only the nodes on the replayed path count (their status too), so the
// result equals that loop's bit for bit; the others are discarded work.
// Every wave runs whole groups and leaves its loops together (the shuffles
// and ballots see all 64 lanes).
template <int GL>
__global__ __launch_bounds__(128, LC_ECL_WAVES) void k_eclipse_bisect_wide(const GridDev *__restrict__ grids,
                                                                          FaceArrays fa, EclipseWork ew,
                                                                          int *inst_status) {
    static_assert(GL == 4 || GL == 8 || GL == 16, "lanes per search");
    constexpr int LV = GL == 4 ? 2 : (GL == 8 ? 3 : 4);
    constexpr int PER_WAVE = 64 / GL;
    const int n = *ew.hit_count * 2;
    const int lane = (int)(threadIdx.x & 63);
    const int node_j = lane & (GL - 1);  // this lane's node; 0 evaluates nothing
    const int gbase = lane & ~(GL - 1);
    const int depth = node_j ? 31 - __clz(node_j) : 0;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
    for (int i0 = wave * PER_WAVE; i0 < n; i0 += nwave * PER_WAVE) {
        const int i = i0 + lane / GL;
        const bool valid = i < n;
        int other = PRIMARY, inst = 0;
        double q = 0., spin = 0., cosi = 0., sini = 0., delta = 0., rref = 0., pref = 0., ACC = 0.;
        double pin = 0., pout = 0.;
        V3 r = v3(0., 0., 0.);
        if (valid) {
            const int k = i >> 1;
            const int f = ew.face[k];
            const GridDev &G = grids[ew.grid[k]];
            other = G.star == PRIMARY ? SECONDARY : PRIMARY;
            inst = G.inst;
            q = G.q;
            spin = G.spin_other;
            cosi = G.cosi;
            sini = G.sini;
            delta = G.delta;
            rref = G.rref_other;
            pref = G.pref_other;
            r = v3(fa.px[f], fa.py[f], fa.pz[f]);
            ACC = ew.acc[k];
            pin = ew.phi[k];
            pout = ew.lim[i];
        }
        bool live = valid && fabs(pin - pout) > delta;
        int st_fail = ST_OK;
        while (__ballot(live)) {
            bool h = false;
            int st = ST_OK;
            if (live && node_j) {
                double a = pin, b = pout;  // walk from the root to this node
                for (int d = depth - 1; d >= 0; d--) {
                    const double m = (a + b) / 2.;
                    if ((node_j >> d) & 1)
                        a = m;
                    else
                        b = m;
                }
                st = fblink(q, other, spin, ACC, set_earth(cosi, sini, (a + b) / 2.), r, rref, pref, h);
            }
            const unsigned long long hm = __ballot(h), em = __ballot(st != ST_OK);
            int node = 1;
#pragma unroll
            for (int d = 0; d < LV; d++) {
                const int st_node = __shfl(st, gbase + node, 64);
                if (live) {
                    if (!(fabs(pin - pout) > delta)) {
                        live = false;
                    } else if ((em >> (gbase + node)) & 1ull) {
                        st_fail = st_node;
                        live = false;
                    } else {
                        const double pmid = (pin + pout) / 2.;
                        const bool hh = (hm >> (gbase + node)) & 1ull;
                        if (hh)
                            pin = pmid;
                        else
                            pout = pmid;
                        node = 2 * node + (hh ? 1 : 0);
                    }
                }
            }
            live = live && fabs(pin - pout) > delta;
        }
        if (valid && node_j == 0) {
            if (st_fail) {
                set_status(inst_status, inst, st_fail);
                ew.res[i] = 0.;
            } else {
                const double x = (pin + pout) / 2.;
                ew.res[i] = x - floor(x);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_eclipse_final(FaceArrays fa, EclipseWork ew) {
    const int n = *ew.hit_count;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int f = ew.face[k];
        const double ingress = ew.res[2 * k];
        double egress = ew.res[2 * k + 1];
        if (egress < ingress) egress++;
        fa.ecl[f] = 1;
        fa.ein[f] = ingress;
        fa.eout[f] = egress;
    }
}

// Faces past the ring layout of a grid, up to numface(): no ring writes
// them, so they keep the default face (zeros, gravity 1).  One workgroup
// row per gap (blockIdx.y).
__global__ __launch_bounds__(256) void k_fill_gaps(const int2 *__restrict__ gaps, FaceArrays fa) {
    const int2 g = gaps[blockIdx.y];
    for (int i = (int)threadIdx.x; i < g.y; i += (int)blockDim.x) {
        const int f = g.x + i;
        fa.px[f] = 0.; fa.py[f] = 0.; fa.pz[f] = 0.;
        fa.dx[f] = 0.; fa.dy[f] = 0.; fa.dz[f] = 0.;
        fa.area[f] = 0.f;
        fa.grav[f] = 1.f;
        fa.ecl[f] = 0;
        fa.ein[f] = 0.;
        fa.eout[f] = 0.;
    }
}

// Stage reuse: a grid whose geometry inputs equal those of another grid
// of the batch, or of a grid kept from an earlier call, takes a copy of its
// faces instead of recomputing them (the reference's exact-match cache,
// light_curve_comp.cpp:34-117, 156-207).  blockIdx.y selects the copy.
__global__ __launch_bounds__(256) void k_copy_faces(const FaceCopy *__restrict__ items, FaceArrays src,
                                                    FaceArrays dst) {
    const FaceCopy c = items[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
        const int a = c.src + i, b = c.dst + i;
        dst.px[b] = src.px[a];
        dst.py[b] = src.py[a];
        dst.pz[b] = src.pz[a];
        dst.dx[b] = src.dx[a];
        dst.dy[b] = src.dy[a];
        dst.dz[b] = src.dz[a];
        dst.area[b] = src.area[a];
        dst.grav[b] = src.grav[a];
        dst.ecl[b] = src.ecl[a];
        dst.ein[b] = src.ein[a];
        dst.eout[b] = src.eout[a];
    }
}

__global__ __launch_bounds__(256) void k_copy_flux(const FaceCopy *__restrict__ items, const float *__restrict__ src,
                                                   float *__restrict__ dst) {
    const FaceCopy c = items[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x)
        dst[c.dst + i] = src[c.src + i];
}

// Continuum brightness * area of one face (set_star_continuum.cpp:103-296)
__global__ __launch_bounds__(256) void k_continuum(const GridDev *__restrict__ grids, int ngrid, int nface_total,
                                                  FaceArrays fa) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface_total) return;
    const GridDev &G = grids[wave_lookup(ngrid, f, [&](int g) { return grids[g].face_off; })];
    if (G.cont_copy) return;  // fluxes copied from an identical grid afterwards
    const V3 posn = v3(fa.px[f], fa.py[f], fa.pz[f]);
    const V3 dirn = v3(fa.dx[f], fa.dy[f], fa.dz[f]);
    const V3 cofm_other = G.star == PRIMARY ? v3(1., 0., 0.) : v3(0., 0., 0.);
    V3 vec = sub(cofm_other, posn);
    double r = len(vec);
    double mu = dot(dirn, vec) / r;
    double t = G.t_base;
    for (int k = 0; k < G.nspot; k++) {
        V3 sd = v3(G.spot_dir[k][0], G.spot_dir[k][1], G.spot_dir[k][2]);
        double dist;
        if (G.star == PRIMARY) {
            dist = rad2deg(acos(dot(posn, sd) / len(posn)));
        } else {
            V3 off = sub(posn, v3(1., 0., 0.));
            dist = rad2deg(acos(dot(off, sd) / len(off)));
        }
        const double EFAC = 2.3548200450309493;  // sqrt(8 ln 2), constants.h:96
        double tref = G.spot_running ? t : G.t_base;
        t += (G.spot_tcen[k] - tref) * exp(-sqr(dist / (G.spot_fwhm[k] / EFAC)) / 2.);
    }
    if (G.uespot) {
        V3 evec = v3(posn.x, posn.y, 0);
        double evecl = len(evec);
        if (evecl > 0.) {
            double rd = len(posn);
            double clat = dot(posn, evec) / evecl / rd;
            clat = dmin(1.0, dmax(-1., clat));
            double dlat = rad2deg(acos(clat));
            V3 ue = v3(G.ue_dir[0], G.ue_dir[1], G.ue_dir[2]);
            double clong = dot(ue, evec) / evecl;
            clong = dmin(1.0, dmax(-1., clong));
            double dlong = rad2deg(acos(clong));
            if (dlat <= G.ue_lathw && dlong <= G.ue_longhw)
                t = G.ue_temp;
            else if (dlat <= G.ue_lathw)
                t += (G.ue_temp - G.t_base) * exp(-(dlong - G.ue_longhw) / G.ue_taper);
            else if (dlong <= G.ue_longhw)
                t += (G.ue_temp - G.t_base) * exp(-(dlat - G.ue_lathw) / G.ue_taper);
            else
                t += (G.ue_temp - G.t_base) * exp(-(dlat - G.ue_lathw) / G.ue_taper) *
                     exp(-(dlong - G.ue_longhw) / G.ue_taper);
        }
    }
    const double ro = G.r_cont_other;
    double temp, geom;
    const double g = (double)fa.grav[f];
    // x^4 as two squarings and x^(1/4) as two square roots instead of pow
    // (set_star_continuum.cpp:172-192, 271-290 call pow): each differs from pow by
    // an ulp or so, and the two pow calls were half the kernel's time
    const double tg = t * pow(g, G.gdcbol);
    const double tg2 = tg * tg, ti2 = G.t_irr * G.t_irr;
    if (mu >= ro) {
        geom = sqr(ro / r) * mu;
        temp = sqrt(sqrt(tg2 * tg2 + G.absorb * (ti2 * ti2) * geom));
    } else if (mu > -ro) {
        double x0 = -mu / ro;
        geom = sqr(ro / r) * ro * (sqrt(1. - x0 * x0) * (2 + x0 * x0) / 3 - x0 * (PI / 2 - asin(x0))) / PI;
        temp = sqrt(sqrt(tg2 * tg2 + G.absorb * (ti2 * ti2) * geom));
    } else {
        geom = 0.;
        temp = tg;
    }
    const double area = (double)fa.area[f];
    float flux = (float)(area * planck(G.wave, temp));
    if (G.mirror) flux = (float)((double)flux + area * geom * planck(G.wave, G.t_mirror));
    fa.flux[f] = flux;
}

// Deterministic block sums (fixed tree order).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    double r = lds[0];
    __syncthreads();
    return r;
}

// Stable partition of every grid into never-eclipsed / eclipsed faces
// (FlatGrid::build, comp_light_flat.cpp:26-65) and the packed float4
// records of the mixed-precision flux kernel, in two passes over segments
// of FLAT_SEG faces: k_flat_count counts each segment's never-eclipsed
// faces, k_flat_build places a segment's faces after those of the grid's
// earlier segments.  The result is the single-pass partition's.
__global__ __launch_bounds__(256) void k_flat_count(const GridDev *__restrict__ grids,
                                                    const FlatSeg *__restrict__ segs, FaceArrays fa, int *seg_cnt) {
    __shared__ int lds_w[4];
    const FlatSeg S = segs[blockIdx.x];
    const int base = grids[S.grid].face_off + S.start;
    int cnt = 0;
    for (int i = threadIdx.x; i < S.n; i += 256) cnt += fa.ecl[base + i] ? 0 : 1;
    // per-wave sums, then the workgroup's
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0) lds_w[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) seg_cnt[blockIdx.x] = lds_w[0] + lds_w[1] + lds_w[2] + lds_w[3];
}

__global__ __launch_bounds__(1024) void k_flat_build(const GridDev *__restrict__ grids,
                                                     const FlatSeg *__restrict__ segs,
                                                     const int *__restrict__ seg_cnt, FaceArrays fa, FlatArrays fl,
                                                     int fp64) {
    __shared__ int lds_w[16];
    const FlatSeg S = segs[blockIdx.x];
    const GridDev &G = grids[S.grid];
    // never-eclipsed faces of the whole grid and of its earlier segments
    int n0 = 0, before0 = 0;
    for (int k = 0; k < S.nseg; k++) {
        const int c = seg_cnt[S.seg0 + k];
        n0 += c;
        before0 += S.seg0 + k < (int)blockIdx.x ? c : 0;
    }
    if ((int)blockIdx.x == S.seg0 && threadIdx.x == 0) fl.n0[S.grid] = n0;
    const int before1 = S.start - before0;  // eclipsed faces of the earlier segments
    const int base = G.face_off;
    int done0 = 0, done1 = 0;
    for (int c = 0; c < S.n; c += 1024) {
        const bool in = c + (int)threadIdx.x < S.n;
        const int i = S.start + c + (int)threadIdx.x;
        const bool plain = in && !fa.ecl[base + i];
        int tot;
        const int ex = block_count_excl<1024>(plain, lds_w, tot);
        if (in) {
            const int pos = plain ? before0 + done0 + ex : n0 + before1 + done1 + ((int)threadIdx.x - ex);
            const int f = base + i, p = base + pos;
            double dx = fa.dx[f], dy = fa.dy[f], dz = fa.dz[f];
            float flux = fa.flux[f];
            fl.rec0[p] = make_float4((float)dx, (float)dy, (float)dz, flux);
            // positions only where a record needs them (a third of the reads)
            const bool need_pos = G.beam != 0.0 || G.glens || fp64;
            double px = 0., py = 0., pz = 0.;
            if (need_pos) {
                px = fa.px[f];
                py = fa.py[f];
                pz = fa.pz[f];
            }
            if (G.beam != 0.0) {  // read only by the beaming variants of the flux kernel
                double vx = -G.VFAC * G.spin_beam * py;
                double vy = G.star == PRIMARY ? G.VFAC * (G.spin_beam * px - G.XCOFM)
                                              : G.VFAC * (G.spin_beam * (px - 1.0) + 1.0 - G.XCOFM);
                double vn = dx * vx + dy * vy;
                fl.rec1[p] = make_float4((float)vx, (float)vy, (float)vn, 0.f);
            }
            if (G.glens) fl.rec2[p] = make_float4((float)px, (float)py, (float)pz, 0.f);  // lensing variants
            // the flux kernels read a face's range only past n0 (the
            // faces that can be eclipsed), so plain faces skip it
            if (!plain) fl.ecl[p] = make_double2(fa.ein[f], fa.eout[f]);
            if (fp64) {
                fl.fdx[p] = dx; fl.fdy[p] = dy; fl.fdz[p] = dz;
                fl.fpx[p] = px; fl.fpy[p] = py; fl.fpz[p] = pz;
                fl.fflux[p] = (double)flux;
            }
        }
        done0 += tot;
        done1 += min(1024, S.n - c) - tot;
    }
}

// LDC::imu (lcurve.h:216-231)
__device__ __forceinline__ double ldc_imu(double mu, const GridDev &G) {
    if (mu <= 0) return 0.;
    mu = dmin(mu, 1.);
    double ommu = 1. - mu, im = 1.;
    if (G.ltype == 0) {
        im -= ommu * (G.ldc[0] + ommu * (G.ldc[1] + ommu * (G.ldc[2] + ommu * G.ldc[3])));
    } else {
        im -= G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3];
        double msq = sqrt(mu);
        im += msq * (G.ldc[0] + msq * (G.ldc[1] + msq * (G.ldc[2] + msq * G.ldc[3])));
    }
    return im;
}

// hidden at the normalised phase ph by one of the opaque disc's ranges of
// the face (Point::visible over the ranges appended after the star's own)
__device__ __forceinline__ bool disc_hidden(const GridDev &G, const FaceArrays &fa, int f, double ph) {
    if (G.disc_comp < 0) return false;
    for (int r = 0; r < fa.dn[f]; r++) {
        const double2 d = fa.dr[2 * f + r];
        if ((ph >= d.x && ph <= d.y) || ph <= d.y - 1.0) return true;
    }
    return false;
}

// Point-path element at one phase with spin = 1 for beaming
// (comp_light.cpp:21-92 as called by comp_star1/comp_star2)
__device__ double point_elem(const GridDev &G, FaceArrays fa, int f, const V3 &earth, double phi) {
    double ph = phi - floor(phi);
    // the star's own range only: the reference sets the grid-switching
    // scales before it appends an opaque disc's ranges
    // (light_curve_comp.cpp:205-250 before 286-326)
    if (fa.ecl[f] && ((ph >= fa.ein[f] && ph <= fa.eout[f]) || ph <= fa.eout[f] - 1.0)) return 0.0;
    V3 dirn = v3(fa.dx[f], fa.dy[f], fa.dz[f]);
    double mu = dot(earth, dirn);
    if (!(mu > G.mucrit)) return 0.0;
    double flux = (double)fa.flux[f];
    double magn = 1.0;
    if (G.star == SECONDARY && G.glens) {
        V3 s = v3(fa.px[f], fa.py[f], fa.pz[f]);
        double d = -dot(s, earth);
        if (d > 0.0) {
            double p = len(add(s, scl(d, earth)));
            double hp = 0.5 * p;
            double rd = G.rlens1 * d;
            double pd = (hp * hp > 25.0 * rd) ? p + rd / p : hp + sqrt(hp * hp + rd);
            magn = pd * pd / (pd - hp) / hp / 4.0;
        }
    }
    if (G.beam == 0.0) {
        if (G.star == PRIMARY) return mu * flux * ldc_imu(mu, G);
        return mu * magn * flux * ldc_imu(mu, G);
    }
    const double spin = 1.0;
    double vx = -G.VFAC * spin * fa.py[f];
    double vy = G.star == PRIMARY ? G.VFAC * (spin * fa.px[f] - G.XCOFM)
                                  : G.VFAC * (spin * (fa.px[f] - 1.0) + 1.0 - G.XCOFM);
    double vr = -(earth.x * vx + earth.y * vy);
    double vn = dirn.x * vx + dirn.y * vy;
    double mud = mu - mu * vr - vn;
    if (G.star == PRIMARY) return mu * flux * (1.0 - G.beam * vr) * ldc_imu(mud, G);
    return mu * magn * flux * (1.0 - G.beam * vr) * ldc_imu(mud, G);
}

__device__ __forceinline__ int gint_type(double phase, double phase1, double phase2) {
    double pnorm = phase - floor(phase);
    if (pnorm <= phase1 || pnorm >= 1. - phase1) return 1;
    if ((pnorm > phase1 && pnorm < phase2) || (pnorm > 1. - phase2 && pnorm < 1. - phase1)) return 2;
    return 3;
}


// Raw Point-path sums for the Ginterp scale factors: block (c, k, inst)
// sums chunk c of the grid that type(phi_k) selects into
// part[(inst * 8 + k) * nchunk + c] (zero past the grid's end).
__global__ __launch_bounds__(256) void k_point_sums(const GridDev *__restrict__ grids,
                                                   const InstDev *__restrict__ insts, FaceArrays fa, int nchunk,
                                                   double *part) {
    __shared__ double lds[256];
    const int c = blockIdx.x, k = blockIdx.y, inst = blockIdx.z;
    const InstDev &I = insts[inst];
    const bool star1 = k < 4;
    if (!I.active || (star1 && I.copy1) || (!star1 && I.copy2)) return;
    double phi = gint_phase(k, I);
    int ty = gint_type(phi, I.phase1, I.phase2);
    int gi = star1 ? (ty == 1 ? I.grid1f : I.grid1c) : (ty == 3 ? I.grid2f : I.grid2c);
    const GridDev &G = grids[gi];
    V3 earth = set_earth(I.cosi, I.sini, phi);
    const int i1 = min(G.nface, (c + 1) * PSUM_FACES);
    double s = 0.0;
    for (int i = c * PSUM_FACES + (int)threadIdx.x; i < i1; i += 256) s += point_elem(G, fa, G.face_off + i, earth, phi);
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) part[((size_t)inst * 8 + k) * nchunk + c] = s;
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

// Ginterp scale factors in the reference's update order
// (light_curve_comp.cpp:214-254): each comp_star value is
// (0 + 1 * scale(phi) * s) / 1 with the scales updated so far.
__global__ void k_gint_scales(InstDev *insts, int ninst, int nchunk, const double *part) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    InstDev I = insts[i];
    if (!I.active) return;
    I.scale11 = I.scale12 = I.scale21 = I.scale22 = 1.;
    double s[8];
    for (int k = 0; k < 8; k++) {  // the chunk sums in chunk order
        const bool used = k < 4 ? !I.copy1 : !I.copy2;
        s[k] = 0.;
        if (used)
            for (int c = 0; c < nchunk; c++) s[k] += part[((size_t)i * 8 + k) * nchunk + c];
    }
    if (!I.copy1) {
        double ff = (0.0 + 1.0 * gscale1(I, gint_phase(0, I)) * s[0]) / 1;
        double fc = (0.0 + 1.0 * gscale1(I, gint_phase(1, I)) * s[1]) / 1;
        I.scale11 = ff / fc;
        ff = (0.0 + 1.0 * gscale1(I, gint_phase(2, I)) * s[2]) / 1;
        fc = (0.0 + 1.0 * gscale1(I, gint_phase(3, I)) * s[3]) / 1;
        I.scale12 = ff / fc;
    }
    if (!I.copy2) {
        double ff = (0.0 + 1.0 * gscale2(I, gint_phase(4, I)) * s[4]) / 1;
        double fc = (0.0 + 1.0 * gscale2(I, gint_phase(5, I)) * s[5]) / 1;
        I.scale21 = ff / fc;
        ff = (0.0 + 1.0 * gscale2(I, gint_phase(6, I)) * s[6]) / 1;
        fc = (0.0 + 1.0 * gscale2(I, gint_phase(7, I)) * s[7]) / 1;
        I.scale22 = ff / fc;
    }
    insts[i] = I;
}

// Diagnostic sums of one grid: [0] sum flux*gravity (float product,
// comp_gravity.cpp), [1] sum flux, [2] sum area*rcosa/r^3 and [3] sum
// area*rcosa (comp_radius.cpp), [4] the phase-0.5 flat sum with spin = 1
// for the white-dwarf contribution (comp_light_flat.cpp:543-571).
__global__ __launch_bounds__(1024) void k_diag(const GridDev *__restrict__ grids, const int *__restrict__ gidx,
                                              FaceArrays fa, double *sums) {
    __shared__ double lds[1024];
    const GridDev &G = grids[gidx[blockIdx.x]];
    const V3 cofm = G.star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    const V3 earth = set_earth(G.cosi, G.sini, 0.5);
    double s[5] = {0., 0., 0., 0., 0.};
    for (int i = threadIdx.x; i < G.nface; i += 1024) {
        const int f = G.face_off + i;
        float fl = fa.flux[f];
        s[0] += (double)(fl * fa.grav[f]);
        s[1] += (double)fl;
        V3 vec = sub(v3(fa.px[f], fa.py[f], fa.pz[f]), cofm);
        double r = len(vec);
        double rcosa = dot(v3(fa.dx[f], fa.dy[f], fa.dz[f]), vec);
        s[2] += fa.area[f] * rcosa / pow(r, 3.0);
        s[3] += fa.area[f] * rcosa;
        if (G.star == PRIMARY) {
            // flat element (comp_light_flat.cpp:110-127), eclipse by range
            double mu = earth.x * fa.dx[f] + earth.y * fa.dy[f] + earth.z * fa.dz[f];
            bool hidden = (fa.ecl[f] && ((0.5 >= fa.ein[f] && 0.5 <= fa.eout[f]) || 0.5 <= fa.eout[f] - 1.0)) ||
                          disc_hidden(G, fa, f, 0.5);
            if (mu > G.mucrit && !hidden) {
                double m = dmin(mu, 1.0), im;
                if (G.ltype == 0) {
                    double om = 1.0 - m;
                    im = 1.0 - om * (G.ldc[0] + om * (G.ldc[1] + om * (G.ldc[2] + om * G.ldc[3])));
                } else {
                    double base = 1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]);
                    double msq = sqrt(dmax(m, 0.0));
                    im = base + msq * (G.ldc[0] + msq * (G.ldc[1] + msq * (G.ldc[2] + msq * G.ldc[3])));
                }
                double val;
                if (G.beam == 0.0) {
                    val = mu * (double)fl * (mu > 0.0 ? im : 0.0);
                } else {
                    double vx = -G.VFAC * 1.0 * fa.py[f];
                    double vy = G.VFAC * (1.0 * fa.px[f] - G.XCOFM);
                    double vr = -(earth.x * vx + earth.y * vy);
                    double vn = fa.dx[f] * vx + fa.dy[f] * vy;
                    double mud = mu - mu * vr - vn;
                    double md = dmin(mud, 1.0), imd;
                    if (G.ltype == 0) {
                        double om = 1.0 - md;
                        imd = 1.0 - om * (G.ldc[0] + om * (G.ldc[1] + om * (G.ldc[2] + om * G.ldc[3])));
                    } else {
                        double base = 1.0 - (G.ldc[0] + G.ldc[1] + G.ldc[2] + G.ldc[3]);
                        double msq = sqrt(dmax(md, 0.0));
                        imd = base + msq * (G.ldc[0] + msq * (G.ldc[1] + msq * (G.ldc[2] + msq * G.ldc[3])));
                    }
                    val = mu * (double)fl * (1.0 - G.beam * vr) * (mud > 0.0 ? imd : 0.0);
                }
                s[4] += val;
            }
        }
    }
    for (int k = 0; k < 5; k++) {
        double t = block_sum<1024>(s[k], lds);
        if (threadIdx.x == 0) sums[blockIdx.x * 8 + k] = t;
    }
}

}  // namespace

void launch_faces(hipStream_t s, const GridDev *grids, const RingDev *rings, int nring, int nhalf_total,
                  FaceArrays fa, int2 *cand, int *cand_count, int *inst_status) {
    if (nhalf_total <= 0) return;
    hipLaunchKernelGGL(k_faces, dim3((nhalf_total + LC_FACES_NT - 1) / LC_FACES_NT), dim3(LC_FACES_NT), 0, s, grids,
                       rings, nring, nhalf_total,
                       fa, cand, cand_count, inst_status);
}

void launch_eclipse(hipStream_t s, const GridDev *grids, int ngrid, FaceArrays fa, const int2 *cand,
                    const int *cand_count, int max_cand, EclipseWork ew, int *inst_status, bool rounds,
                    bool stop, int wide) {
    if (max_cand <= 0 || ngrid <= 0) return;
    const int blocks = std::min((max_cand + 127) / 128, 16384);
    hipLaunchKernelGGL((k_eclipse_potmin<true, LC_FIRST_NT>), dim3(std::min((max_cand + LC_FIRST_NT - 1) / LC_FIRST_NT, 16384)),
                       dim3(LC_FIRST_NT), 0, s, grids, fa, cand, cand_count, ew, inst_status);
    hipLaunchKernelGGL((k_eclipse_potmin<false, 128>), dim3(blocks), dim3(128), 0, s, grids, fa, cand, cand_count, ew,
                       inst_status);
    const int blocks2 = std::min((2 * max_cand + 127) / 128, 16384);
    if (stop)
        hipLaunchKernelGGL(k_eclipse_open, dim3(std::min((2 * max_cand + OPEN_NT * OPEN_PER - 1) / (OPEN_NT * OPEN_PER), 4096)),
                           dim3(OPEN_NT), 0, s, grids, fa, ew);
    if (!rounds && !stop && (wide == 4 || wide == 8 || wide == 16)) {
        const int bw = (int)std::min<long long>((2LL * max_cand * wide + 127) / 128, 16384);
        if (wide == 4)
            hipLaunchKernelGGL(k_eclipse_bisect_wide<4>, dim3(bw), dim3(128), 0, s, grids, fa, ew, inst_status);
        else if (wide == 8)
            hipLaunchKernelGGL(k_eclipse_bisect_wide<8>, dim3(bw), dim3(128), 0, s, grids, fa, ew, inst_status);
        else
            hipLaunchKernelGGL(k_eclipse_bisect_wide<16>, dim3(bw), dim3(128), 0, s, grids, fa, ew, inst_status);
        hipLaunchKernelGGL(k_eclipse_final, dim3(std::min((max_cand + 255) / 256, 4096)), dim3(256), 0, s, fa, ew);
        return;
    }
    const int nround = rounds ? ECL_ROUNDS : 1;
    for (int round = 0; round < nround; round++) {
        const int nstep = round == nround - 1 ? (1 << 30) : (round == 0 ? ECL_FIRST_STEPS : ECL_STEPS);
        hipLaunchKernelGGL(k_eclipse_bisect, dim3(blocks2), dim3(128), 0, s, grids, ngrid, fa, ew, inst_status, round,
                           nstep, stop ? 1 : 0);
    }
    hipLaunchKernelGGL(k_eclipse_final, dim3(std::min((max_cand + 255) / 256, 4096)), dim3(256), 0, s, fa, ew);
}

void launch_copy_faces(hipStream_t s, const FaceCopy *items, int nitem, FaceArrays src, FaceArrays dst) {
    for (int k = 0; k < nitem; k += 65535)
        hipLaunchKernelGGL(k_copy_faces, dim3(32, std::min(65535, nitem - k)), dim3(256), 0, s, items + k, src, dst);
}

void launch_copy_flux(hipStream_t s, const FaceCopy *items, int nitem, const float *src, float *dst) {
    for (int k = 0; k < nitem; k += 65535)
        hipLaunchKernelGGL(k_copy_flux, dim3(16, std::min(65535, nitem - k)), dim3(256), 0, s, items + k, src, dst);
}

void launch_comp_pretest(hipStream_t s, const GridDev *grids, int ngrid, int nface, FaceArrays fa, int2 *cand,
                         int *cand_count) {
    if (nface <= 0 || ngrid <= 0) return;
    hipLaunchKernelGGL(k_comp_pretest, dim3((nface + 255) / 256), dim3(256), 0, s, grids, ngrid, nface, fa, cand,
                       cand_count);
}

void launch_fill_gaps(hipStream_t s, const int2 *gaps, int ngap, FaceArrays fa) {
    for (int k = 0; k < ngap; k += 65535)
        hipLaunchKernelGGL(k_fill_gaps, dim3(1, std::min(65535, ngap - k)), dim3(256), 0, s, gaps + k, fa);
}

void launch_continuum(hipStream_t s, const GridDev *grids, int ngrid, int nface_total, FaceArrays fa) {
    if (nface_total <= 0) return;
    hipLaunchKernelGGL(k_continuum, dim3((nface_total + 255) / 256), dim3(256), 0, s, grids, ngrid, nface_total, fa);
}

void launch_flat_build(hipStream_t s, const GridDev *grids, const FlatSeg *segs, int nseg, int *seg_cnt,
                       FaceArrays fa, FlatArrays fl, int fp64) {
    if (nseg <= 0) return;
    hipLaunchKernelGGL(k_flat_count, dim3(nseg), dim3(256), 0, s, grids, segs, fa, seg_cnt);
    hipLaunchKernelGGL(k_flat_build, dim3(nseg), dim3(1024), 0, s, grids, segs, seg_cnt, fa, fl, fp64);
}

void launch_point_sums(hipStream_t s, const GridDev *grids, const InstDev *insts, int ninst, FaceArrays fa,
                       int nchunk, double *part) {
    if (ninst <= 0) return;
    hipLaunchKernelGGL(k_point_sums, dim3(nchunk, 8, ninst), dim3(256), 0, s, grids, insts, fa, nchunk, part);
}

void launch_gint_scales(hipStream_t s, InstDev *insts, int ninst, int nchunk, const double *part) {
    if (ninst <= 0) return;
    hipLaunchKernelGGL(k_gint_scales, dim3((ninst + 63) / 64), dim3(64), 0, s, insts, ninst, nchunk, part);
}

void launch_diag(hipStream_t s, const GridDev *grids, const int *gidx, int ng, FaceArrays fa, double *sums) {
    if (ng <= 0) return;
    hipLaunchKernelGGL(k_diag, dim3(ng), dim3(1024), 0, s, grids, gidx, fa, sums);
}

}  // namespace lck
