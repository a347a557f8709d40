// lc_comp.h -- the disc, the disc's outer edge and the bright spot
// (light_curve_comp.cpp:272-333, 483-497): host planning of their elements
// and the kernels that find their eclipses and sum their light.
#pragma once
#include <string>
#include <vector>

#include "../../include/lcurve_amd.h"
#include "lc_kernels.h"

namespace lck {

// One instance's components: faces [off, off + n_disc + n_edge + n_spot) of
// the component arrays, disc faces first, then the edge, then the spot.
struct CompDev {
    int inst, off, n_disc, n_edge, n_spot;
    int opaque;              // the disc eclipses what lies behind it
    long long sum_off;       // first of the instance's 3 x S sums
    double lin, quad;        // disc and edge limb darkening (comp_light.cpp:95-107)
    double rdisc1, rdisc2, beta, height, iangle;  // disc_eclipse inputs (disc_eclipse.cpp:55)
};

// Per component face: the normal and flux the sums read, and the eclipse
// ranges (up to 4: the opaque disc gives at most 2, each star 1).
constexpr int COMP_MAX_RANGES = 4;
struct CompArrays {
    double *dx, *dy, *dz;
    float *flux;
    int *nrange;
    double2 *range;          // [face * COMP_MAX_RANGES + k]
    double *sums;            // [sum_off + kind * S + j]: disc, edge, spot at sub-phase j
};

// Elements of one instance's components, built on the host.
struct CompHost {
    std::vector<double> px, py, pz, dx, dy, dz;
    std::vector<float> flux;
    int n_disc = 0, n_edge = 0, n_spot = 0;
    // the stars as eclipsers (star_eclipse.cpp:15-25)
    double q = 0., spin1 = 1., spin2 = 1., r1 = 0., r2 = 0., rref1 = 0., pref1 = 0., rref2 = 0., pref2 = 0.;
    double cosi = 1., sini = 0., delta = 0.;
    int roche1 = 1, roche2 = 1, eclipse1 = 1, eclipse2 = 1;
    double lin = 0., quad = 0., rdisc1 = 0., rdisc2 = 0., beta = 0., height = 0., iangle = 0.;
    int opaque = 0;
    int size() const { return n_disc + n_edge + n_spot; }
};

// set_disc_grid + set_disc_edge(outer, not visual) + set_disc_continuum +
// set_edge_continuum, and set_bright_spot_grid (set_disc_grid.cpp,
// set_disc_continuum.cpp, set_bright_spot_grid.cpp): 0 or an lc_status with
// the reference's message in err.
int plan_components(const lc_model &m, CompHost &out, std::string &err);

// Candidates of the star eclipse searches over the component faces' two
// slots (one per eclipsing star), appended to cand.
void launch_comp_pretest(hipStream_t s, const GridDev *grids, int ngrid, int nface, FaceArrays fa, int2 *cand,
                         int *cand_count);
// Eclipse ranges of every component face: the opaque disc's (disc_eclipse),
// then the searches' results in slot (by star 1) and slot + nc (by star 2).
void launch_comp_build(hipStream_t s, const CompDev *comps, int ncomp, int nc, FaceArrays slots, CompArrays ca,
                       int *inst_status);
// Sums of the disc, edge and spot elements at every sub-phase of the
// instances that have components (comp_light.cpp:322-407, FP64).
void launch_comp_sums(hipStream_t s, const CompDev *comps, int ncomp, int S, PhaseArrays pa, CompArrays ca);
// An opaque disc's eclipse ranges of the star faces (light_curve_comp.cpp:
// 286-326): for each grid gidx[k] (with disc_comp >= 0), dn/dr of its faces.
void launch_disc_star_ranges(hipStream_t s, const GridDev *grids, const int *gidx, int ng, int max_face,
                             const CompDev *comps, FaceArrays fa, int *inst_status);

}  // namespace lck
