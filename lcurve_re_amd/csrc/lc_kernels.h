// lc_kernels.h -- device-side records of the engine and the host launch
// wrappers of its kernels.  All records are plain structs shared by host
// planning and the kernels (HBM layouts are structure-of-arrays except the
// packed float4 flux records).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lck {

// One star grid of one model instance (fine or coarse).
struct GridDev {
    int inst, star, fine, nface;
    int face_off;            // first face in the global face arrays
    int ring_off, nring;     // rings of this grid in the ring table
    int half_off, nhalf;     // first / number of half-ring work items
    int roche_self, roche_other, eclipse, npole;
    // geometry (set_star_grid.cpp:52-207 derived values)
    double q, spin_self, spin_other;
    double rref_self, pref_self, rref_other, pref_other;
    double r_self, r_other, acc, delta, cosi, sini, gref;
    // continuum (set_star_continuum.cpp)
    double t_base;          // T1 for star 1, |T2| for star 2
    double t_irr;           // temperature of the irradiating star in pow(T, 4)
    double t_mirror;        // temperature of the mirror term
    double gdcbol, absorb, wave, r_cont_other;
    int mirror, nspot, uespot, spot_running;  // spot_running: star 2 offsets use the running T
    double spot_dir[3][3], spot_fwhm[3], spot_tcen[3];
    double ue_dir[3], ue_longhw, ue_lathw, ue_taper, ue_temp;
    // flux sums (comp_light_flat.cpp:110-168)
    double beam, spin_beam, VFAC, XCOFM, rlens1;
    double ldc[4], mucrit;
    int ltype, glens;
    // 1: the eclipse search may end once its bracket holds none of the
    // instance's phases (a grid used by one instance and kept by no cache)
    int ecl_stop;
    // 1: the fluxes equal those of another grid of the call or of the cache
    // (same faces, same continuum inputs) and are copied after k_continuum,
    // which skips this grid
    int cont_copy;
    // the instance's opaque disc hides parts of this star (index of its
    // CompDev, -1 without): the disc's ranges are in FaceArrays dn/dr and
    // the grid's flux sums run in k_flux_ragged
    int disc_comp;
};

// One latitude ring (set_star_grid.cpp:262-320).
struct RingDev {
    double sint, cost, phi1, dphi_span, tspan_nl;
    int nphi, face_off, half_off, grid;
};

// Per-instance parameters for phases, scatter and scaling.
struct InstDev {
    double t0, period, pdot, deltat, slope, quad, cube, third, cosi, sini;
    double phase1, phase2, scale11, scale12, scale21, scale22;
    double sfac0, sfac4;  // the caller's factors (scale off): star 1 / all, star 2 (iscale)
    double sfac1, sfac2, sfac3;  // iscale, scale off: the caller's disc, edge and spot factors
    double zcols;         // iscale, scale off: sfac[1..3] times the disc, edge and spot columns (0)
    int copy1, copy2, grid1f, grid2f, grid1c, grid2c;
    int scale, rdata, active, iscale;
    int star2;           // iscale: star 2 is a component (t2 > 0)
    int has_comp;        // disc, edge or bright spot present
    long long comp_off;  // first of its disc, edge and spot sums (3 x S)
};

// Phases at which light_curve_comp.cpp:220-254 samples the Point-path sums.
__device__ __forceinline__ double gint_phase(int k, const InstDev &I) {
    switch (k) {
    case 0: return 0.9999999999 * I.phase1;
    case 1: return 1.0000000001 * I.phase1;
    case 2: return 1. - 0.9999999999 * I.phase1;
    case 3: return 1. - 1.0000000001 * I.phase1;
    case 4: return 1 - 1.0000000001 * I.phase2;
    case 5: return 1 - 0.9999999999 * I.phase2;
    case 6: return 1.0000000001 * I.phase2;
    default: return 0.9999999999 * I.phase2;
    }
}

struct FaceArrays {
    double *px, *py, *pz, *dx, *dy, *dz;
    float *area, *grav, *flux;
    int *ecl;
    double *ein, *eout;
    // an opaque disc's eclipse ranges of the face (disc_eclipse: 0 to 2,
    // appended after the star's own range, light_curve_comp.cpp:286-326);
    // null when no grid of the call has one
    int *dn;
    double2 *dr;  // [2 * face]
};

// Flux-stage copy of every grid, in the reference's FlatGrid order
// (never-eclipsed faces first, comp_light_flat.cpp:26-65).
struct FlatArrays {
    float4 *rec0;   // dx, dy, dz, flux
    float4 *rec1;   // vx, vy, vn (beaming), 0
    float4 *rec2;   // px, py, pz (lensing), 0
    double2 *ecl;   // ingress, egress
    double *fdx, *fdy, *fdz, *fpx, *fpy, *fpz, *fflux;  // FP64 copies
    int *n0;        // never-eclipsed count per grid
    unsigned long long *evald;  // [16 x 64] face x sub-phase elements evaluated, per variant
};

// Work lists of the three-kernel eclipse search (k_geom.hip).
struct EclipseWork {
    int *hit_count;  // faces pot_min found eclipsed
    int *face;       // [hit] face index
    int *grid;       // [hit] grid of that face
    double *phi;     // [hit] phase inside the eclipse (pot_min)
    double *lim;     // [2*hit] sphere window phi1, phi2
    double *acc;     // [hit] position accuracy of fblink
    double *res;     // [2*hit + side] ingress (side 0) / egress (side 1) in [0, 1)
    // bisection in rounds: brackets of unfinished items (indexed like res),
    // the item lists of consecutive rounds (alternating) and their counts
    double *b_in, *b_out;
    int *alist[2];   // alist[1] first holds the candidates pot_min has to search (k_eclipse_first)
    int *acount;     // [ECL_COUNTS]: open searches after each round, the pot_min list count, the
                     // searches left open by the visibility test ahead of round 0
    // per instance, a bitmap of the 2^bits_lb bins of [0, 1) that hold a
    // phase at which eclipse ranges are tested (bin = floor(phase * 2^lb))
    const unsigned *phase_bits;
    int bits_lb;
};

// Bisection rounds for searches that may stop early: the first runs
// ECL_FIRST_STEPS steps of every search, the second ECL_STEPS more on the
// searches still open (compacted), the last one runs them to the end.
constexpr int ECL_ROUNDS = 3, ECL_FIRST_STEPS = 10, ECL_STEPS = 3;
constexpr int ECL_COUNTS = ECL_ROUNDS + 2;

struct DataDev {
    const double *time, *expose, *flux, *ferr, *weight;
    const int *ndiv, *sub_off;
    int n, S;
    double middle, range;
};

struct PhaseArrays {
    double *ex, *ey, *ez, *phin, *w1, *w2;  // [inst * S + j]
    signed char *type;                      // Ginterp::type(phi)
};

struct BatchDev {
    int inst, star, grid;
    int list_off;    // offset of the sub-phase index list, -1: identity
    int count_idx;   // index into the device counts, -1: all S sub-phases
    int nchunk, chunk, ntile;
    long long part_off;
};

struct TaskDev {
    int batch, tile, chunk, pad;
};

// One face-range copy of the geometry arrays (stage reuse of a grid).
struct FaceCopy {
    int src, dst, n, pad;
};

// ------------------------------------------------------------ launches
void launch_faces(hipStream_t s, const GridDev *grids, const RingDev *rings, int nring, int nhalf_total,
                  FaceArrays fa, int2 *cand, int *cand_count, int *inst_status);
// rounds: run the bisections in ECL_ROUNDS launches (else one)
// stop: some grids may end their searches early (GridDev::ecl_stop); their
// searches over a face that is never seen in the first bracket end before
// round 0, which then runs over the list of the others.
void launch_eclipse(hipStream_t s, const GridDev *grids, int ngrid, FaceArrays fa, const int2 *cand,
                    const int *cand_count, int max_cand, EclipseWork ew, int *inst_status, bool rounds,
                    bool stop, int wide);
// Copies positions, normals, areas, gravities and eclipse ranges of whole
// grids (not the fluxes, which the continuum kernel recomputes).
void launch_copy_faces(hipStream_t s, const FaceCopy *items, int nitem, FaceArrays src, FaceArrays dst);
void launch_continuum(hipStream_t s, const GridDev *grids, int ngrid, int nface_total, FaceArrays fa);
// Copies the fluxes of whole grids (FaceCopy ranges).
void launch_copy_flux(hipStream_t s, const FaceCopy *items, int nitem, const float *src, float *dst);
// Exclusive prefix count of a 0/1 flag over the workgroup in thread order
// (NT a multiple of 64), from wave ballots; the total through `total`.
// lds_w holds NT / 64 ints.
template <int NT>
__device__ __forceinline__ int block_count_excl(bool flag, int *lds_w, int &total) {
    const unsigned long long b = __ballot(flag);
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int in_wave = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) lds_w[w] = __popcll(b);
    __syncthreads();
    int before = 0, tot = 0;
    for (int k = 0; k < NT / 64; k++) {
        const int c = lds_w[k];
        before += k < w ? c : 0;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return before + in_wave;
}

// One segment of a grid for the FlatGrid build: faces [start, start + n)
// of grid `grid` (start counted from the grid's first face); the grid's
// segments are seg0 .. seg0 + nseg - 1.  A grid without faces has one
// empty segment.
constexpr int FLAT_SEG = 8192;
struct FlatSeg {
    int grid, start, n, seg0, nseg, pad[3];
};
// Faces between a grid's ring layout and numface() (first face, count):
// the default face (zeros, gravity 1).
void launch_fill_gaps(hipStream_t s, const int2 *gaps, int ngap, FaceArrays fa);
void launch_flat_build(hipStream_t s, const GridDev *grids, const FlatSeg *segs, int nseg, int *seg_cnt,
                       FaceArrays fa, FlatArrays fl, int fp64);
// Ginterp single-phase sums in chunks of PSUM_FACES faces: part holds
// ninst * 8 * nchunk partial sums, reduced in chunk order by
// launch_gint_scales.
constexpr int PSUM_FACES = 16384;
void launch_point_sums(hipStream_t s, const GridDev *grids, const InstDev *insts, int ninst, FaceArrays fa,
                       int nchunk, double *part);
void launch_gint_scales(hipStream_t s, InstDev *insts, int ninst, int nchunk, const double *part);
void launch_phases(hipStream_t s, const InstDev *insts, int ninst, DataDev d, PhaseArrays pa);
// Byte regions (sizes multiples of 4) copied from src, or zeroed where src
// is null, in one launch: the clears ahead of a call's kernels and the
// gather of its results into one block for a single download.
constexpr int MAX_REGIONS = 12;
struct ByteRegions {
    void *dst[MAX_REGIONS];
    const void *src[MAX_REGIONS];
    size_t n[MAX_REGIONS];
};
void launch_regions(hipStream_t s, const ByteRegions &r, int nreg);
// Marks in bits (ninst x 2^lb bits, cleared by the caller) the bin of every
// phase at which an instance's eclipse ranges are tested: its sub-phases,
// the eight Ginterp phases and phase 0.5 (diagnostics).
void launch_phase_bits(hipStream_t s, const InstDev *insts, int ninst, DataDev d, unsigned *bits, int lb);
void launch_route(hipStream_t s, const InstDev *insts, int ninst, int S, PhaseArrays pa, int *lists, int *counts);
// variant = ltype | beaming << 1 | lensing << 2 | (mucrit > 0) << 3;
// rphase = sub-phases per lane
// wave: mixed-precision variants without beaming or lensing run k_flux_wave
// (per-wave phase tiles) instead of k_flux_mixed.
void launch_flux(hipStream_t s, int variant, int fp64, int rphase, int wave, const TaskDev *tasks, int ntask,
                 const BatchDev *batches, const GridDev *grids, FlatArrays fl, PhaseArrays pa, int S,
                 const int *lists, const int *counts, double *partial);
// Expands the task table on the device: the tasks of batch order[k] are
// first[k] .. first[k+1]-1 (first[nb] = ntask), numbered chunk-major
// (chunk * ntile + tile).
void launch_make_tasks(hipStream_t s, const int *first, const int *order, int nb, const BatchDev *batches,
                       int ntask, TaskDev *tasks);
void launch_flux_reduce(hipStream_t s, const BatchDev *batches, int nbatch, int S, int threads_per_tile,
                        const int *lists, const int *counts, const double *partial, double *out);
// Flux sums of batches whose grid has an opaque disc's ranges (several
// eclipse ranges per face): batches order[0 .. nb), FP64 in the reference's
// order (faces with 0, 1, more ranges; 1024-face blocks), one partial row.
void launch_flux_ragged(hipStream_t s, int variant, const int *order, int nb, int max_count, const BatchDev *batches,
                        const GridDev *grids, FaceArrays fa, PhaseArrays pa, int S, const int *lists,
                        const int *counts, double *partial);
// Continues the chunk partials of one flat grid (batch 0, no routing) over
// its faces [n1, nface) that have several eclipse ranges: chunks
// c_first .. nchunk-1 of `chunk` faces; a chunk below nchunk_main starts
// from its partial, a later one from zero; faces in index order, FP64
// elements, hidden when any of ranges[moff[i-n1] .. moff[i-n1+1]) holds
// the phase (comp_light_flat.cpp:294-307).  stride = doubles per chunk row.
void launch_flux_ragged_tail(hipStream_t s, int variant, const GridDev *grid, const BatchDev *batch, FlatArrays fl,
                             const int *moff, const double2 *ranges, PhaseArrays pa, int S, int n1, int nface,
                             int chunk, int c_first, int nchunk, int nchunk_main, size_t stride, double *partial);
// part: workspace of ninst * ceil(n / SCAT_PTS) * SCAT_PARTS doubles (blocks
// of SCAT_PTS points, one per thread, or 4 SCAT_PTS in calls that fill the device)
// slots 0-7: scale sums, chi**2, the two-star iscale sums; 8-27: the
// five-column iscale Gram matrix (15) and right-hand side (5)
constexpr int SCAT_PTS = 256, SCAT_PARTS = 28;
// comp: the disc, edge and spot sums of instances with has_comp (else unused)
void launch_scatter(hipStream_t s, const InstDev *insts, int ninst, DataDev d, PhaseArrays pa, const double *out,
                    const double *comp, double *calc, double *scal, double *part);
void launch_diag(hipStream_t s, const GridDev *grids, const int *gidx, int ng, FaceArrays fa, double *sums);

}  // namespace lck
