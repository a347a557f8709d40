// comp.hip -- disc, disc edge and bright spot (light_curve_comp.cpp:272-333,
// 483-497).
//
// Host: the elements of each component (positions, normals, areas, fluxes)
// follow set_disc_grid / set_disc_edge / set_bright_spot_grid and the two
// continuum routines operation for operation, so they carry the same float
// areas and fluxes.  The bright spot sits where the gas stream from L1
// crosses radius_spot; the stream is integrated here with classical RK4 at
// a fixed step of 1e-4 and the crossing bisected in time to 1e-13 in radius
// (the reference integrates with Bulirsch-Stoer to 1e-8 relative accuracy
// and bisects to 1e-10, so the two spot positions agree to the reference's
// own integration accuracy, not to the bit).
//
// Device: the star eclipses of component elements run through the engine's
// FP64 eclipse search (two slots per element, one per eclipsing star, fed
// by k_comp_pretest in k_geom.hip); k_comp_build gathers each element's
// ranges, adding the opaque disc's own (disc_eclipse.cpp, restated below);
// k_comp_sums adds the elements at every sub-phase in the reference's
// order, in FP64 (comp_light.cpp:95-119, 322-407).
#include <algorithm>
#include <cmath>
#include <string>

#include "lc_comp.h"
#include "lc_math.h"

namespace lck {
using namespace lcx;

namespace {

// lcurve.h:399-410
void r1r2_of(const lc_model &m, double &rr1, double &rr2) {
    if (m.use_radii) {
        rr1 = m.r1.value;
        rr2 = m.r2.value;
    } else {
        double sini = std::sin(deg2rad(m.iangle.value));
        double r2pr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * M_PI_ * m.cphi4.value), 2));
        double r2mr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * M_PI_ * m.cphi3.value), 2));
        rr1 = (r2pr1 - r2mr1) / 2.0;
        rr2 = (r2pr1 + r2mr1) / 2.0;
    }
}

// Roche::rocacc (rocacc.cpp:14-34): acceleration in the rotating frame
V3 rocacc(double q, const V3 &r, const V3 &v) {
    const double f1 = 1.0 / (1.0 + q), f2 = f1 * q;
    const double yzsq = sqr(r.y) + sqr(r.z);
    const double r1sq = sqr(r.x) + yzsq, r2sq = sqr(r.x - 1.0) + yzsq;
    const double fm1 = f1 / (r1sq * std::sqrt(r1sq)), fm2 = f2 / (r2sq * std::sqrt(r2sq)), fm3 = fm1 + fm2;
    return v3(-fm3 * r.x + fm2 + 2.0 * v.y + r.x - f2, -fm3 * r.y - 2.0 * v.x + r.y, -fm3 * r.z);
}

void rk4_step(double q, V3 &r, V3 &v, double h) {
    const V3 a1 = rocacc(q, r, v);
    const V3 r2 = add(r, scl(0.5 * h, v)), v2 = add(v, scl(0.5 * h, a1));
    const V3 a2 = rocacc(q, r2, v2);
    const V3 r3 = add(r, scl(0.5 * h, v2)), v3_ = add(v, scl(0.5 * h, a2));
    const V3 a3 = rocacc(q, r3, v3_);
    const V3 r4 = add(r, scl(h, v3_)), v4 = add(v, scl(h, a3));
    const V3 a4 = rocacc(q, r4, v4);
    r = add(r, scl(h / 6.0, add(add(v, scl(2.0, v2)), add(scl(2.0, v3_), v4))));
    v = add(v, scl(h / 6.0, add(add(a1, scl(2.0, a2)), add(scl(2.0, a3), a4))));
}

// Roche::strinit (strining.cpp:14-32) then Roche::stradv (stradv.cpp:27-80):
// the stream from just inside L1 to where it first crosses radius rad.
int stream_point(double q, double rad, V3 &pos, std::string &err) {
    const double SMALL = 1.e-5;
    int st = 0;
    const double rl1 = lagrange_l1(0, q, 1.0, &st);
    if (st) {
        err = "Roche::xl1: q <= 0 or no convergence";
        return LC_E_PARAM;
    }
    const double mu = q / (1.0 + q);
    const double a = (1.0 - mu) / std::pow(rl1, 3) + mu / std::pow((1. - rl1), 3);
    const double lambda1 = std::sqrt(((a - 2.0) + std::sqrt(a * (9.0 * a - 8.0))) / 2.0);
    const double m1 = (lambda1 * lambda1 - 2.0 * a - 1.0) / 2.0 / lambda1;
    V3 r = v3(rl1 - SMALL, -m1 * SMALL, 0.0), v = v3(-lambda1 * SMALL, -lambda1 * m1 * SMALL, 0.0);
    const double H = 1e-4, TMAX = 10.;
    const double rinit = len(r);
    double rnow = rinit, time = 0.;
    V3 ro = r, vo = v;
    while ((rinit > rad && rnow > rad) || (rinit < rad && rnow < rad)) {
        ro = r;
        vo = v;
        rk4_step(q, r, v, H);
        time += H;
        rnow = len(r);
        if (time > TMAX) {
            err = "Roche::stradv: taken too long without crossing radius = " + std::to_string(rad) +
                  " in Roche::stradv.";
            return LC_E_PARAM;
        }
    }
    // bisect the last step in time
    double lo = 0., hi = H;
    const bool above = len(ro) > rad;
    for (int it = 0; it < 80 && hi - lo > 0.; it++) {
        const double mid = 0.5 * (lo + hi);
        V3 rr = ro, vv = vo;
        rk4_step(q, rr, vv, mid);
        const double rm = len(rr);
        if ((rm > rad) == above)
            lo = mid;
        else
            hi = mid;
        r = rr;
        if (std::fabs(rm - rad) < 1e-13) break;
    }
    pos = r;
    return 0;
}

// set_edge_continuum (set_disc_continuum.cpp:49-88): one element's
// temperature with the donor's irradiation, times its area.
double edge_flux(double tedge, double r2, double t2, double absorb, double wave, const V3 &posn, const V3 &dirn,
                 float area) {
    const V3 vec = sub(v3(1., 0., 0.), posn);
    const double r = len(vec);
    const double mu = dot(dirn, vec) / r;
    double temp, geom;
    if (mu >= r2) {
        geom = sqr(r2 / r) * mu;
        temp = std::pow(std::pow(tedge, 4) + absorb * std::pow(t2, 4) * geom, 0.25);
    } else if (mu > -r2) {
        const double x0 = -mu / r2;
        geom = sqr(r2 / r) * r2 * (std::sqrt(1. - x0 * x0) * (2 + x0 * x0) / 3 - x0 * (PI / 2 - std::asin(x0))) / PI;
        temp = std::pow(std::pow(tedge, 4) + absorb * std::pow(t2, 4) * geom, 0.25);
    } else {
        temp = tedge;
    }
    return (double)area * planck(wave, temp);
}

}  // namespace

int plan_components(const lc_model &m, CompHost &c, std::string &err) {
    double r1, r2;
    r1r2_of(m, r1, r2);
    int st = 0;
    const double rl1 = lagrange_l1(1, m.q.value, m.spin1.value, &st);
    const double rl2 = 1. - lagrange_l1(2, m.q.value, m.spin2.value, &st);
    if (st) {
        err = "Roche::xl11/xl12: q <= 0 or no convergence";
        return LC_E_PARAM;
    }
    const char *who = m.add_disc ? "set_disc_grid" : "set_bright_spot_grid";
    if (r1 < 0)
        r1 = rl1;
    else if (r1 > rl1) {
        err = std::string(who) + ": the primary star is larger than its Roche lobe!";
        return LC_E_ROCHE_OVERFLOW;
    }
    if (r2 < 0)
        r2 = rl2;
    else if (r2 > rl2) {
        err = std::string(who) + ": the secondary star is larger than its Roche lobe!";
        return LC_E_ROCHE_OVERFLOW;
    }
    const double ffac1 = r1 / rl1, ffac2 = r2 / rl2;
    c.q = m.q.value;
    c.spin1 = m.spin1.value;
    c.spin2 = m.spin2.value;
    c.r1 = r1;
    c.r2 = r2;
    ref_sphere(c.q, 0, c.spin1, ffac1, c.rref1, c.pref1);
    ref_sphere(c.q, 1, c.spin2, ffac2, c.rref2, c.pref2);
    const double ri = deg2rad(m.iangle.value);
    c.cosi = std::cos(ri);
    c.sini = std::sin(ri);
    c.delta = m.delta_phase;
    c.roche1 = m.roche1;
    c.roche2 = m.roche2;
    c.eclipse1 = m.eclipse1;
    c.eclipse2 = m.eclipse2;
    c.iangle = m.iangle.value;
    c.opaque = m.add_disc && m.opaque;
    c.lin = m.lin_limb_disc.value;
    c.quad = m.quad_limb_disc.value;
    const double rdisc1 = m.rdisc1.value > 0. ? m.rdisc1.value : r1;
    const double rdisc2 = m.rdisc2.value > 0. ? m.rdisc2.value : m.radius_spot.value;
    c.rdisc1 = rdisc1;
    c.rdisc2 = rdisc2;
    c.beta = m.beta_disc.value;
    c.height = m.height_disc.value;
    auto push = [&](const V3 &p, const V3 &d, double flux) {
        c.px.push_back(p.x); c.py.push_back(p.y); c.pz.push_back(p.z);
        c.dx.push_back(d.x); c.dy.push_back(d.y); c.dz.push_back(d.z);
        c.flux.push_back((float)flux);
    };
    const double EFAC = 1.0000001;
    const double wave = m.wavelength;
    if (m.add_disc) {
        if (m.nrad <= 0) {
            err = "set_disc_grid: nrad must be positive";
            return LC_E_PARAM;
        }
        // set_disc_grid.cpp:66-130 and set_disc_continuum (set_disc_continuum.cpp:23-32)
        const double drad = (rdisc2 - rdisc1) / m.nrad, drrad = rdisc2 / m.nrad;
        const double bright = planck(wave, m.temp_disc.value);
        for (int i = 0; i < m.nrad; i++) {
            const double rad = rdisc1 + (rdisc2 - rdisc1) * (i + 0.5) / m.nrad;
            int ntheta = int(std::ceil(TWOPI * rad / drrad));
            ntheta = ntheta > 8 ? ntheta : 8;
            const double h = EFAC * m.height_disc.value * std::pow(rad, m.beta_disc.value);
            const double tanp = m.beta_disc.value * h / rad;
            const double cosp = 1. / std::sqrt(1 + sqr(tanp));
            const double sinp = tanp * cosp;
            const float area = (float)(TWOPI / ntheta * rad * drad);
            for (int j = 0; j < ntheta; j++) {
                const double theta = TWOPI * j / ntheta;
                const double sint = std::sin(theta), cost = std::cos(theta);
                const V3 posn = v3(rad * cost, rad * sint, h);
                const V3 dirn = v3(-cost * sinp, -sint * sinp, cosp);
                push(posn, dirn, bright * std::pow(len(posn) / rdisc2, m.texp_disc.value) * (double)area);
            }
        }
        c.n_disc = (int)c.px.size();
        // set_disc_edge(mdl, true, edge, false) (set_disc_grid.cpp:160-264) and
        // set_edge_continuum with the donor's r2, |t2| and absorb_edge
        const double size = rdisc2 / m.nrad;
        const double RAD = rdisc2;
        const double H = m.height_disc.value * std::pow(RAD, m.beta_disc.value);
        const int NOUT = int(2 * H / size) + 2;
        const int nt = int(std::ceil(TWOPI * RAD / size));
        const int NTHETA = nt > 8 ? nt : 8;
        const double area = TWOPI / NTHETA * RAD * 2 * H / (NOUT - 1);
        const double t2 = std::fabs(m.t2.value);
        for (int i = 0; i < NTHETA; i++) {
            const double theta = TWOPI * i / NTHETA;
            const double sint = std::sin(theta), cost = std::cos(theta);
            const V3 dirn = v3(cost, sint, 0);
            const V3 top = v3(EFAC * RAD * cost, EFAC * RAD * sint, EFAC * H);
            const float a_top = (float)(area / 2.);
            push(top, dirn, edge_flux(m.temp_edge.value, r2, t2, m.absorb_edge.value, wave, top, dirn, a_top));
            for (int j = 0; j < NOUT - 1; j++) {
                const V3 posn = v3(EFAC * RAD * cost, EFAC * RAD * sint, -H + 2 * H * j / (NOUT - 1));
                push(posn, dirn,
                     edge_flux(m.temp_edge.value, r2, t2, m.absorb_edge.value, wave, posn, dirn, (float)area));
            }
        }
        c.n_edge = (int)c.px.size() - c.n_disc;
    }
    if (m.add_spot) {
        // set_bright_spot_grid.cpp:49-100
        if (m.nspot < 2) {
            err = "set_bright_spot_grid: nspot must be at least 2";
            return LC_E_PARAM;
        }
        V3 bspot;
        const int sst = stream_point(m.q.value, m.radius_spot.value, bspot, err);
        if (sst) return sst;
        const double theta = std::atan2(bspot.y, bspot.x) + TWOPI / 4. + deg2rad(m.angle_spot.value);
        const double alpha = deg2rad(m.yaw_spot.value);
        const double tilt = deg2rad(m.tilt_spot.value);
        const V3 bvec = v3(std::cos(theta), std::sin(theta), 0);
        const V3 pvec = v3(0, 0, 1);
        const V3 tvec = v3(std::sin(tilt) * std::sin(theta + alpha), -std::sin(tilt) * std::cos(theta + alpha),
                           std::cos(tilt));
        const double expon = m.expon_spot.value, epow = m.epow_spot.value;
        const double BMAX = std::pow(expon / epow, 1. / epow);
        const double SFAC = 20. + BMAX;
        const float AREA = (float)(SFAC * m.length_spot.value * m.height_spot.value / (m.nspot - 1));
        const double BRIGHT = planck(wave, m.temp_spot.value);
        std::vector<V3> pos(m.nspot);
        std::vector<double> br(m.nspot);
        for (int i = 0; i < m.nspot; i++) {
            const double dist = SFAC * i / (m.nspot - 1);
            pos[i] = add(bspot, scl(m.length_spot.value * (dist - BMAX), bvec));
            br[i] = BRIGHT * std::pow(dist / BMAX, expon) * std::exp(expon / epow - std::pow(dist, epow));
        }
        for (int i = 0; i < m.nspot; i++)  // the tilted strip
            push(pos[i], tvec, br[i] * (1 - m.cfrac_spot.value) * (double)AREA);
        for (int i = 0; i < m.nspot; i++)  // the parallel strip
            push(pos[i], pvec, br[i] * m.cfrac_spot.value * (double)AREA);
        c.n_spot = 2 * m.nspot;
    }
    return 0;
}

namespace {

// disc_eclipse.cpp:305-366
enum Circle { ABOVE, INSIDE, OUTSIDE, SEPARATE, CROSSING };

__device__ double cut_phase(double rxy, double rcone, double radius, int &bad) {
    if (rxy + rcone <= radius || rxy >= radius + rcone || rcone >= radius + rxy) bad = 1;
    return acos((sqr(rxy) + sqr(rcone) - sqr(radius)) / (2. * rcone * rxy)) / TWOPI;
}

__device__ Circle circle_eclipse(double rxy, double z, double zcirc, double radius, double tani, double &phase,
                                 int &bad) {
    if (z >= zcirc) return ABOVE;
    const double RCONE = tani * (zcirc - z);
    if (RCONE >= rxy + radius) return OUTSIDE;
    if (rxy >= RCONE + radius) return SEPARATE;
    if (rxy + RCONE <= radius) return INSIDE;
    phase = cut_phase(rxy, RCONE, radius, bad);
    return CROSSING;
}

// Roche::disc_eclipse (disc_eclipse.cpp:55-303): up to two (ingress,
// egress) pairs during which the flared disc hides the point.  bad = 1 where
// the reference throws (beta < 1, inconsistent circle cuts).
__device__ int disc_eclipse(double sini, double cosi, double rdisc1, double rdisc2, double beta, double height,
                            const V3 &r, double2 *out, int &bad) {
    if (beta < 1) {
        bad = 1;
        return 0;
    }
    const double HOUT = height * pow(rdisc2, beta);
    if (r.z >= HOUT) return 0;
    if (cosi == 0.) {
        if (fabs(r.z) < HOUT) {
            const double rxy = sqrt(sqr(r.x) + sqr(r.y));
            if (rxy <= rdisc2) {
                out[0] = make_double2(0., 1.);
            } else {
                const double subtend = asin(rdisc2 / rxy) / TWOPI;
                const double pcen = atan2(r.y, -r.x) / TWOPI;
                double ingress = pcen - subtend;
                ingress = ingress - floor(ingress);
                out[0] = make_double2(ingress, ingress + 2. * subtend);
            }
            return 1;
        }
        return 0;
    }
    const double RXY = sqrt(sqr(r.x) + sqr(r.y));
    if (rdisc1 < RXY && RXY < rdisc2 && fabs(r.z) < height * pow(RXY, beta)) {
        out[0] = make_double2(0., 1.1);
        return 1;
    }
    const double TANI = sini / cosi;
    Circle result;
    double phase = 0., ingress, egress;
    if (RXY < rdisc2 && r.z >= height * pow(dmax(rdisc1, RXY), beta)) {
        result = circle_eclipse(RXY, r.z, HOUT, rdisc2, TANI, phase, bad);
        if (result == OUTSIDE) {
            out[0] = make_double2(0., 1.1);
            return 1;
        } else if (result == CROSSING) {
            const double phi0 = atan2(r.y, -r.x) / TWOPI;
            ingress = phi0 + phase;
            ingress -= floor(ingress);
            egress = ingress + 1 - 2. * phase;
            out[0] = make_double2(ingress, egress);
            return 1;
        }
        return 0;
    }
    const double RCONE_LO = dmax(0., TANI * (-HOUT - r.z));
    if (RCONE_LO >= RXY + rdisc2) return 0;
    const double RCONE_HI = TANI * (HOUT - r.z);
    if (RXY >= RCONE_HI + rdisc2) return 0;
    double eclipse_phase = 0;
    if (RXY + RCONE_LO <= rdisc2) {
        eclipse_phase = 0.5;
    } else if (RXY <= rdisc2) {
        eclipse_phase = cut_phase(RXY, RCONE_LO, rdisc2, bad);
    } else {
        if (sqr(RCONE_HI) + sqr(rdisc2) >= sqr(RXY) && sqr(RCONE_LO) + sqr(rdisc2) <= sqr(RXY)) {
            eclipse_phase = asin(rdisc2 / RXY) / TWOPI;
        } else if (sqr(RCONE_HI) + sqr(rdisc2) < sqr(RXY)) {
            eclipse_phase = cut_phase(RXY, RCONE_HI, rdisc2, bad);
        } else {
            eclipse_phase = cut_phase(RXY, RCONE_LO, rdisc2, bad);
        }
    }
    const double HIN = height * pow(rdisc1, beta);
    double appear_phase = -1;
    // the rims the line of sight has to thread, innermost last
    auto thread = [&](double zc, double rc, bool first) {
        if (first) {
            result = circle_eclipse(RXY, r.z, zc, rc, TANI, phase, bad);
            if (result == INSIDE)
                appear_phase = 0.5;
            else if (result == CROSSING)
                appear_phase = phase;
        } else if (appear_phase > 0) {
            result = circle_eclipse(RXY, r.z, zc, rc, TANI, phase, bad);
            if (result == CROSSING)
                appear_phase = dmin(appear_phase, phase);
            else if (result != INSIDE)
                appear_phase = -1.;
        }
    };
    if (r.z < -HOUT) {
        thread(-HOUT, rdisc2, true);
        thread(-HIN, rdisc1, false);
        thread(HIN, rdisc1, false);
        thread(HOUT, rdisc2, false);
    } else if (RXY < rdisc1) {
        if (r.z < -HIN) {
            thread(-HIN, rdisc1, true);
            thread(HIN, rdisc1, false);
            thread(HOUT, rdisc2, false);
        } else if (r.z < HIN) {
            thread(HIN, rdisc1, true);
            thread(HOUT, rdisc2, false);
        }
    }
    const double phi0 = atan2(r.y, -r.x) / TWOPI;
    if (appear_phase <= 0) {
        ingress = phi0 - eclipse_phase;
        ingress -= floor(ingress);
        out[0] = make_double2(ingress, ingress + 2. * eclipse_phase);
        return 1;
    } else if (appear_phase < eclipse_phase) {
        ingress = phi0 - eclipse_phase;
        ingress -= floor(ingress);
        out[0] = make_double2(ingress, ingress + (eclipse_phase - appear_phase));
        ingress = phi0 + appear_phase;
        ingress -= floor(ingress);
        out[1] = make_double2(ingress, ingress + (eclipse_phase - appear_phase));
        return 2;
    }
    return 0;
}

// One work item per component face: normal, flux and eclipse ranges.
__global__ __launch_bounds__(256) void k_comp_build(const CompDev *__restrict__ comps, int ncomp, int nc,
                                                   FaceArrays slots, CompArrays ca, int *inst_status) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nc) return;
    int lo = 0, hi = ncomp - 1;  // the component set holding face f
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (comps[mid].off <= f)
            lo = mid;
        else
            hi = mid - 1;
    }
    const CompDev &C = comps[lo];
    const int k = f - C.off;
    ca.dx[f] = slots.dx[f];
    ca.dy[f] = slots.dy[f];
    ca.dz[f] = slots.dz[f];
    ca.flux[f] = slots.flux[f];
    double2 *rg = ca.range + (size_t)f * COMP_MAX_RANGES;
    int nr = 0, bad = 0;
    if (C.opaque && k < C.n_disc + C.n_edge) {
        // the reference's disc_eclipse works from sin and cos of
        // Constants::PI / 180 * iangle (disc_eclipse.cpp:62-66)
        const double ang = PI / 180. * C.iangle;
        nr = disc_eclipse(sin(ang), cos(ang), C.rdisc1, C.rdisc2, C.beta, C.height,
                          v3(slots.px[f], slots.py[f], slots.pz[f]), rg, bad);
    }
    if (slots.ecl[f]) rg[nr++] = make_double2(slots.ein[f], slots.eout[f]);
    if (slots.ecl[nc + f]) rg[nr++] = make_double2(slots.ein[nc + f], slots.eout[nc + f]);
    ca.nrange[f] = nr;
    if (bad) atomicCAS(&inst_status[C.inst], 0, LC_E_ECLIPSE_SEARCH);
}

// The opaque disc's ranges of each star face of the grids gidx[]
// (blockIdx.y): disc_eclipse at the face position, as light_curve_comp.cpp:
// 286-326 appends them to the star's own.
__global__ __launch_bounds__(256) void k_disc_star_ranges(const GridDev *__restrict__ grids, const int *__restrict__ gidx,
                                                         const CompDev *__restrict__ comps, FaceArrays fa,
                                                         int *inst_status) {
    const GridDev &G = grids[gidx[blockIdx.y]];
    const CompDev &C = comps[G.disc_comp];
    const double ang = PI / 180. * C.iangle;  // disc_eclipse.cpp:62-66
    const double si = sin(ang), co = cos(ang);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < G.nface; k += gridDim.x * blockDim.x) {
        const int f = G.face_off + k;
        int bad = 0;
        const int nr =
            disc_eclipse(si, co, C.rdisc1, C.rdisc2, C.beta, C.height, v3(fa.px[f], fa.py[f], fa.pz[f]), fa.dr + 2 * f,
                         bad);
        fa.dn[f] = nr;
        if (bad) atomicCAS(&inst_status[G.inst], 0, LC_E_ECLIPSE_SEARCH);
    }
}

constexpr int CT = 256;  // threads (sub-phases) per workgroup and faces per LDS stage

// Disc, edge and spot sums at the sub-phases of one instance
// (comp_light.cpp:95-119 elements, 322-407 faces in order).
__global__ __launch_bounds__(CT) void k_comp_sums(const CompDev *__restrict__ comps, int S, PhaseArrays pa,
                                                 CompArrays ca) {
    __shared__ double s_d[3][CT];
    __shared__ float s_f[CT];
    __shared__ int s_n[CT];
    __shared__ double2 s_r[COMP_MAX_RANGES][CT];
    const CompDev C = comps[blockIdx.y];
    const int j = blockIdx.x * CT + threadIdx.x;
    const bool act = j < S;
    const size_t jj = (size_t)C.inst * S + (act ? j : 0);
    const double ex = pa.ex[jj], ey = pa.ey[jj], ez = pa.ez[jj], phi = pa.phin[jj];
    const int ends[3] = {C.n_disc, C.n_disc + C.n_edge, C.n_disc + C.n_edge + C.n_spot};
    double sum[3] = {0., 0., 0.};
    const int ntot = ends[2];
    for (int c0 = 0; c0 < ntot; c0 += CT) {
        const int cn = min(CT, ntot - c0);
        __syncthreads();
        if ((int)threadIdx.x < cn) {
            const int f = C.off + c0 + threadIdx.x;
            s_d[0][threadIdx.x] = ca.dx[f];
            s_d[1][threadIdx.x] = ca.dy[f];
            s_d[2][threadIdx.x] = ca.dz[f];
            s_f[threadIdx.x] = ca.flux[f];
            const int nr = ca.nrange[f];
            s_n[threadIdx.x] = nr;
            for (int r = 0; r < nr; r++) s_r[r][threadIdx.x] = ca.range[(size_t)f * COMP_MAX_RANGES + r];
        }
        __syncthreads();
        for (int i = 0; i < cn; i++) {
            const int k = c0 + i;
            const int kind = k < ends[0] ? 0 : (k < ends[1] ? 1 : 2);
            // Point::visible (lcurve.h:81-89)
            bool vis = true;
            for (int r = 0; r < s_n[i]; r++) {
                const double2 p = s_r[r][i];
                if ((phi >= p.x && phi <= p.y) || phi <= p.y - 1.0) vis = false;
            }
            double v = 0.;
            if (vis) {
                const double mu = ex * s_d[0][i] + ey * s_d[1][i] + ez * s_d[2][i];
                if (kind < 2) {
                    if (mu > 0.0) {
                        const double om = 1.0 - mu;
                        v = mu * (double)s_f[i] * (1.0 - om * (C.lin + C.quad * om));
                    }
                } else {
                    v = (mu > 0.0) ? mu * (double)s_f[i] : 0.0;
                }
            }
            sum[kind] += v;
        }
    }
    if (act)
        for (int kind = 0; kind < 3; kind++) ca.sums[(size_t)C.sum_off + (size_t)kind * S + j] = sum[kind];
}

}  // namespace

void launch_comp_build(hipStream_t s, const CompDev *comps, int ncomp, int nc, FaceArrays slots, CompArrays ca,
                       int *inst_status) {
    if (ncomp <= 0 || nc <= 0) return;
    hipLaunchKernelGGL(k_comp_build, dim3((nc + 255) / 256), dim3(256), 0, s, comps, ncomp, nc, slots, ca,
                       inst_status);
}

void launch_disc_star_ranges(hipStream_t s, const GridDev *grids, const int *gidx, int ng, int max_face,
                             const CompDev *comps, FaceArrays fa, int *inst_status) {
    if (ng <= 0 || max_face <= 0) return;
    for (int k = 0; k < ng; k += 65535)
        hipLaunchKernelGGL(k_disc_star_ranges, dim3(std::min((max_face + 255) / 256, 64), std::min(65535, ng - k)),
                           dim3(256), 0, s, grids, gidx + k, comps, fa, inst_status);
}

void launch_comp_sums(hipStream_t s, const CompDev *comps, int ncomp, int S, PhaseArrays pa, CompArrays ca) {
    if (ncomp <= 0 || S <= 0) return;
    hipLaunchKernelGGL(k_comp_sums, dim3((S + CT - 1) / CT, ncomp), dim3(CT), 0, s, comps, S, pa, ca);
}

}  // namespace lck
