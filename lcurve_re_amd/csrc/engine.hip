// engine.hip -- host side of the MI355X light-curve engine and its C ABI.
//
// One evaluation (or a batch of model instances sharing one data set) runs
// as one launch sequence on the context's stream:
//
//   faces -> eclipse search -> continuum -> [Ginterp sums] -> FlatGrid build
//   -> phases -> [fine/coarse routing] -> flux sums -> chunk merge
//   -> scatter + scale + chi**2 -> [diagnostics]
//
// The host only plans: per grid it derives the scalars of set_star_grid
// (set_star_grid.cpp:52-207) and the latitude-ring table of add_faces
// (set_star_grid.cpp:256-320), which fixes every face's slot in HBM; all
// per-face and per-phase work runs on the GPU.  Tables are uploaded once per
// call; results come back in one copy at the end.
#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <exception>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include <unistd.h>

#include "../../include/lcurve_amd.h"
#include "lc_comp.h"
#include "lc_host.h"
#include "lc_kernels.h"
#include "lc_math.h"
#include "lc_pool.h"

using namespace lck;
using lcpool::parallel_for;
using lcx::V3;

double lcx_host_l1(int kind, double q, double spin, int *st) { return lcx::lagrange_l1(kind, q, spin, st); }

namespace {

struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw EngineError(LC_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

const double PI = lcx::PI;
const double TWOPI = lcx::TWOPI;

// ------------------------------------------------------------- arenas
// Device memory grows geometrically and is reused across calls, so the
// steady state of repeated evaluations allocates nothing.
struct DevArena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void reset() { used = 0; }
    void reserve(size_t bytes) {
        if (bytes <= cap) return;
        if (base) HIPCHK(hipFree(base));
        base = nullptr;
        size_t nc = std::max(bytes, cap + cap / 2);
        HIPCHK(hipMalloc(&base, nc));
        cap = nc;
    }
    template <class T> T *take(size_t n) {
        size_t a = (used + 255) & ~(size_t)255;
        T *p = reinterpret_cast<T *>(base + a);
        used = a + n * sizeof(T);
        if (used > cap) throw EngineError(LC_E_DEVICE, "device arena overflow");
        return p;
    }
    ~DevArena() {
        if (base) (void)hipFree(base);
    }
};

struct HostPinned {
    char *base = nullptr;
    size_t cap = 0;
    void reserve(size_t bytes) {
        if (bytes <= cap) return;
        if (base) HIPCHK(hipHostFree(base));
        base = nullptr;
        size_t nc = std::max(bytes, cap + cap / 2);
        HIPCHK(hipHostMalloc((void **)&base, nc, hipHostMallocDefault));
        cap = nc;
    }
    ~HostPinned() {
        if (base) (void)hipHostFree(base);
    }
};

// Byte-size accounting helper: sizes computed in one pass, then taken.
struct Sizer {
    size_t total = 0;
    template <class T> void add(size_t n) { total = ((total + 255) & ~(size_t)255) + n * sizeof(T); }
};

// ---------------------------------------------------------- planning
struct ModelScalars {
    double r1, r2;          // get_r1r2 (lcurve.h:399-410)
    double cosi, sini;      // of deg2rad(iangle)
};

void get_r1r2(const lc_model &m, double &rr1, double &rr2) {
    if (m.use_radii) {
        rr1 = m.r1.value;
        rr2 = m.r2.value;
    } else {
        double sini = std::sin(lcx::deg2rad(m.iangle.value));
        double r2pr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * M_PI * m.cphi4.value), 2.0));
        double r2mr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * M_PI * m.cphi3.value), 2.0));
        rr1 = (r2pr1 - r2mr1) / 2.0;
        rr2 = (r2pr1 + r2mr1) / 2.0;
    }
}

double l1(int kind, double q, double spin) {
    int st = 0;
    double x = lcx::lagrange_l1(kind, q, spin, &st);
    if (st) throw EngineError(LC_E_PARAM, kind == 0 ? "Roche::xl1: q <= 0 or no convergence"
                                                    : (kind == 1 ? "Roche::xl11(double, double): q <= 0 or no convergence"
                                                                 : "Roche::xl12(double,double): q <= 0 or no convergence"));
    return x;
}

// numface (numface.cpp:14-53)
int numface(int nlat, bool infill, double thelo, double thehi, int nlatfill, int nlngfill) {
    int nface = 0;
    double theta, dphi, dtheta = PI / nlat;
    if (infill) {
        int nl1 = int(std::ceil(thelo / dtheta));
        for (int i = 0; i < nl1; i++) {
            theta = thelo * (i + 0.5) / nl1;
            dphi = dtheta / std::sin(theta);
            nface += std::max(16, int(2 * PI / dphi));
        }
        int nl2 = int(std::ceil((1 + nlatfill) * (thehi - thelo) / dtheta));
        for (int i = 0; i < nl2; i++) {
            theta = thelo + (thehi - thelo) * (i + 0.5) / nl2;
            dphi = dtheta / std::sin(theta) / (1 + nlngfill);
            nface += std::max(8, int(PI / dphi));
        }
        int nl3 = int(std::ceil((thehi - thelo) / dtheta));
        for (int i = 0; i < nl3; i++) {
            theta = thelo + (thehi - thelo) * (i + 0.5) / nl3;
            dphi = dtheta / std::sin(theta);
            nface += std::max(8, int(PI / dphi));
        }
        int nl4 = int(std::ceil((PI - thehi) / dtheta));
        for (int i = 0; i < nl4; i++) {
            theta = thehi + (PI - thehi) * (i + 0.5) / nl4;
            dphi = dtheta / std::sin(theta);
            nface += std::max(16, int(2 * PI / dphi));
        }
    } else {
        nlat = int(std::ceil(PI / dtheta));
        for (int i = 0; i < nlat; i++) {
            theta = PI * (i + 0.5) / nlat;
            dphi = dtheta / std::sin(theta);
            nface += std::max(16, int(2 * PI / dphi));
        }
    }
    return nface;
}

struct RingHost {
    double sint, cost, phi1, dphi_span, tspan_nl;
    int nphi;
};

struct GridPlan {
    GridDev dev;
    std::vector<RingHost> rings;
    int ring_faces;  // faces laid out by the rings (<= nface)
    int ring_halves;  // half-ring positions of the rings (face kernel work items)
};

// Exact-match key of a grid's geometry: every GridDev field the face and
// eclipse kernels read, plus the ring table.  Equal keys give bit-identical
// faces and eclipse ranges, so a grid with a known key is copied instead of
// recomputed.  (The reference keys its cache on the Model fields that feed
// set_star_grid, light_curve_comp.cpp:34-41; keying on the derived values
// also catches parameter changes that leave a grid unchanged.)
// The key is split in two: a short signature of the GridDev fields and the
// ring count (hashed and compared first) and the ring table itself, which
// is compared only when two signatures are equal.
struct GeoSig {
    double dv[14];
    int iv[9];
    int pad;  // zero: no implicit padding, so memcmp and the hash see only values
    bool operator==(const GeoSig &o) const { return std::memcmp(this, &o, sizeof(GeoSig)) == 0; }
};
struct GeoSigHash {
    size_t operator()(const GeoSig &k) const {
        uint64_t w[sizeof(GeoSig) / 8], h = 0x9e3779b97f4a7c15ull;
        std::memcpy(w, &k, sizeof(w));
        for (uint64_t x : w) {
            h = (h ^ x) * 0xff51afd7ed558ccdull;
            h ^= h >> 32;
        }
        return (size_t)h;
    }
};
static_assert(sizeof(GeoSig) == 14 * 8 + 10 * 4, "GeoSig has no implicit padding");
// Exact-match key of a grid's continuum: every GridDev field k_continuum
// reads besides the faces themselves.  Equal faces (equal GeoSig and ring
// table) with equal keys give bit-identical fluxes, so those are copied
// instead of recomputed (the reference's ContKey, light_curve_comp.cpp:43-52,
// 205-268, keyed here on the derived temperatures, exponents and spot
// directions).  Compared byte for byte, so -0.0 and 0.0 count as different
// inputs and the grid is simply recomputed.
struct ContSig {
    double v[36];
    bool operator==(const ContSig &o) const { return std::memcmp(this, &o, sizeof(ContSig)) == 0; }
};
ContSig cont_sig(const GridDev &G) {
    ContSig c;
    std::memset(&c, 0, sizeof(c));
    int k = 0;
    auto put = [&](double x) { c.v[k++] = x; };
    put(G.star); put(G.t_base); put(G.t_irr); put(G.t_mirror); put(G.gdcbol); put(G.absorb); put(G.wave);
    put(G.r_cont_other); put(G.mirror); put(G.nspot); put(G.uespot); put(G.spot_running);
    for (int j = 0; j < G.nspot && j < 3; j++) {
        put(G.spot_dir[j][0]); put(G.spot_dir[j][1]); put(G.spot_dir[j][2]);
        put(G.spot_fwhm[j]); put(G.spot_tcen[j]);
    }
    if (G.uespot) {
        put(G.ue_dir[0]); put(G.ue_dir[1]); put(G.ue_dir[2]);
        put(G.ue_longhw); put(G.ue_lathw); put(G.ue_taper); put(G.ue_temp);
    }
    return c;
}

GeoSig geom_sig(const GridPlan &gp) {
    const GridDev &G = gp.dev;
    GeoSig k;
    std::memset(&k, 0, sizeof(k));
    const int iv[9] = {G.star, G.fine,    G.nface,        G.roche_self, G.roche_other,
                       G.eclipse, G.npole, gp.ring_faces, (int)gp.rings.size()};
    const double dv[14] = {G.q,          G.spin_self, G.spin_other, G.rref_self, G.pref_self,
                           G.rref_other, G.pref_other, G.r_self,    G.r_other,   G.acc,
                           G.delta,      G.cosi,      G.sini,       G.gref};
    std::memcpy(k.iv, iv, sizeof(iv));
    std::memcpy(k.dv, dv, sizeof(dv));
    return k;
}

bool same_rings(const std::vector<RingHost> &a, const std::vector<RingHost> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (a[i].sint != b[i].sint || a[i].cost != b[i].cost || a[i].phi1 != b[i].phi1 ||
            a[i].dphi_span != b[i].dphi_span || a[i].tspan_nl != b[i].tspan_nl || a[i].nphi != b[i].nphi)
            return false;
    return true;
}

// Rings of one add_faces segment (set_star_grid.cpp:256-320), appended to
// `rings`.
void add_segment(std::vector<RingHost> &rings, int nlat, int star, double tlo, double thi, int nlf, int nlg) {
    const double dtheta = PI / nlat;
    const bool inf = (nlf > 0) || (nlg > 0);
    int nl, nl1 = 0, nl2 = 0;
    if (inf) {
        nl1 = int(std::ceil((1 + nlf) * (thi - tlo) / dtheta));
        nl2 = int(std::ceil((thi - tlo) / dtheta));
        nl = nl1 + nl2;
    } else {
        nl = int(std::ceil((thi - tlo) / dtheta));
    }
    rings.reserve(rings.size() + std::max(nl, 0));
    for (int nt = 0; nt < nl; nt++) {
        RingHost R;
        double theta, phi1, phi2;
        int nls;
        if (inf) {
            if (nt < nl1) {
                theta = tlo + (thi - tlo) * (nt + 0.5) / nl1;
                R.sint = std::sin(theta);
                R.cost = std::cos(theta);
                R.nphi = std::max(8, int(PI * R.sint * (1 + nlg) / dtheta));
                nls = nl1;
                phi1 = star == 0 ? -PI / 2. : +PI / 2.;
                phi2 = star == 0 ? +PI / 2. : +3. * PI / 2.;
            } else {
                theta = tlo + (thi - tlo) * (nt - nl1 + 0.5) / nl2;
                R.sint = std::sin(theta);
                R.cost = std::cos(theta);
                R.nphi = std::max(8, int(PI * R.sint / dtheta));
                nls = nl2;
                phi1 = star == 0 ? +PI / 2. : -PI / 2.;
                phi2 = star == 0 ? +3. * PI / 2. : +PI / 2.;
            }
        } else {
            theta = tlo + (thi - tlo) * (nt + 0.5) / nl;
            R.sint = std::sin(theta);
            R.cost = std::cos(theta);
            R.nphi = std::max(16, int(2 * PI * R.sint / dtheta));
            nls = nl;
            phi1 = 0.;
            phi2 = TWOPI;
        }
        R.phi1 = phi1;
        R.dphi_span = phi2 - phi1;
        R.tspan_nl = (thi - tlo) / nls;
        rings.push_back(R);
    }
}

// A grid without infill has the ring table and face count of its nlat
// alone (the one segment 0..pi, the same for both stars): those are
// computed once per nlat in the process and copied.
struct PlainGrid {
    std::vector<RingHost> rings;
    int nface, faces, halves;  // numface(); sums of nphi and (nphi + 1) / 2 over the rings
};
const PlainGrid &plain_grid(int nlat) {
    static std::shared_mutex mu;
    static std::unordered_map<int, std::unique_ptr<PlainGrid>> memo;
    {
        std::shared_lock<std::shared_mutex> lk(mu);
        auto it = memo.find(nlat);
        if (it != memo.end()) return *it->second;
    }
    auto pg = std::make_unique<PlainGrid>();
    pg->nface = numface(nlat, false, 0., 0., 0, 0);
    add_segment(pg->rings, nlat, 0, 0., PI, 0, 0);
    pg->faces = pg->halves = 0;
    for (const RingHost &R : pg->rings) {
        pg->faces += R.nphi;
        pg->halves += (R.nphi + 1) / 2;
    }
    std::unique_lock<std::shared_mutex> lk(mu);
    auto &slot = memo[nlat];
    if (!slot) slot = std::move(pg);
    return *slot;
}

// Everything set_star_grid derives before add_faces, plus the ring table
// of every add_faces segment (set_star_grid.cpp:49-320).
GridPlan plan_grid(const lc_model &m, int star, bool fine) {
    GridPlan gp;
    GridDev &G = gp.dev;
    std::memset(&G, 0, sizeof(G));
    G.disc_comp = -1;
    const double q = m.q.value;
    double r1, r2;
    get_r1r2(m, r1, r2);
    G.star = star;
    G.fine = fine ? 1 : 0;
    G.eclipse = star == 0 ? m.eclipse1 : m.eclipse2;
    const int nlat = star == 0 ? (fine ? m.nlat1f : m.nlat1c) : (fine ? m.nlat2f : m.nlat2c);
    const int nlatfill = star == 0 ? 0 : (fine ? m.nlatfill : 0);
    const int nlngfill = star == 0 ? 0 : (fine ? m.nlngfill : 0);
    const double latlo = star == 0 ? 0 : m.llo;
    const double lathi = star == 0 ? 0 : m.lhi;
    if (nlat < 0) throw EngineError(LC_E_PARAM, "set_star_grid: number of latitude strips must not be negative");

    const double rl1 = l1(1, q, m.spin1.value);
    if (r1 < 0)
        r1 = rl1;
    else if (r1 > rl1)
        throw EngineError(LC_E_ROCHE_OVERFLOW, "set_star_grid: the primary star is larger than its Roche lobe!");
    const double rl2 = 1. - l1(2, q, m.spin2.value);
    if (r2 < 0)
        r2 = rl2;
    else if (r2 > rl2)
        throw EngineError(LC_E_ROCHE_OVERFLOW, "set_star_grid: the secondary star is larger than its Roche lobe!");
    if (m.glens1 && star == 1 && m.roche1 && m.eclipse2)
        throw EngineError(LC_E_LENSING, "set_star_grid: cannot have gravitational lensing, eclipse and Roche lobe "
                                        "geometry at the same time");
    if (m.glens1 && star == 1) {
        double gm = std::pow(1000. * m.velocity_scale.value, 3.0) * m.tperiod * lcx::DAY / TWOPI;
        double a = std::pow(gm / lcx::sqr(TWOPI / lcx::DAY / m.tperiod), 1. / 3.);
        double rlens1 = 4. * gm / (1. + q) / a / lcx::sqr(lcx::C_SI);
        r1 -= rlens1 / (r1 / (1. - r2 / 2.));
        if (r1 < 0)
            throw EngineError(LC_E_LENSING, "set_star_grid: gravitational lensing correction more than the current "
                                            "program can cope with.");
    }
    const double ffac1 = r1 / rl1, ffac2 = r2 / rl2;
    double rref1, pref1, rref2, pref2;
    lcx::ref_sphere(q, 0, m.spin1.value, ffac1, rref1, pref1);
    lcx::ref_sphere(q, 1, m.spin2.value, ffac2, rref2, pref2);

    bool infill = m.npole && star == 1 && (nlatfill > 0 || nlngfill > 0) && r2 > r1;
    double thelo = 0, thehi = 0;
    if (infill) {
        double rangle = lcx::deg2rad(m.iangle.value);
        double cosi = std::cos(rangle);
        double ratio = (cosi + r1) / r2;
        if (ratio >= 1.) {
            thehi = PI / 2. + rangle;
        } else {
            // bisect where the lower envelope swept by star 1 crosses the
            // limb of star 2 (set_star_grid.cpp:125-144, envelope 528-538)
            double llo = 0., lhi = PI / 2., ex = 0., ey = 0.;
            while (lhi > llo + 1.e-7) {
                double lmid = (llo + lhi) / 2.;
                double si = std::sin(rangle), co = std::cos(rangle);
                double sl = std::sin(lmid), cl = std::cos(lmid);
                double norm = std::sqrt(co * co + si * si * cl * cl);
                ex = sl + r1 * co * sl / norm;
                ey = -co * cl - r1 * cl / norm;
                if (ex * ex + ey * ey < r2 * r2)
                    llo = lmid;
                else
                    lhi = lmid;
            }
            double sini = std::sin(rangle);
            thehi = std::max(PI / 2. - lcx::deg2rad(latlo),
                             std::min(PI / 2. + rangle, std::acos(ey * sini / r2) + lcx::deg2rad(m.lfudge)));
        }
        ratio = (cosi - r1) / r2;
        if (ratio >= 1.)
            infill = false;
        else if (ratio <= -1.)
            thelo = 0.;
        else
            thelo = std::min(PI / 2. - lcx::deg2rad(lathi),
                             std::max(0., PI / 2. - std::acos(ratio) + rangle - lcx::deg2rad(m.lfudge)));
    }
    const PlainGrid *plain = !infill && nlat > 0 ? &plain_grid(nlat) : nullptr;
    G.nface = plain ? plain->nface : numface(nlat, infill, thelo, thehi, nlatfill, nlngfill);

    const double delta = m.delta_phase;
    const double ACC = delta / 10.;
    G.gref = 1.;
    if ((star == 0 && m.roche1) || (star == 1 && m.roche2)) {
        V3 pv, dv;
        double rad, g;
        int st = star == 0 ? lcx::face_scalar(q, 0, m.spin1.value, lcx::v3(-1., 0., 0.), rref1, pref1, ACC, pv, dv,
                                              rad, g)
                           : lcx::face_scalar(q, 1, m.spin2.value, lcx::v3(1., 0., 0.), rref2, pref2, ACC, pv, dv,
                                              rad, g);
        if (st) throw EngineError(LC_E_FACE, "Roche::face error: could not solve for the reference gravity");
        G.gref = g;
    }

    // ring tables of the add_faces segments (set_star_grid.cpp:256-320)
    if (plain) {
        gp.rings = plain->rings;
    } else if (nlat > 0) {
        add_segment(gp.rings, nlat, star, 0., thelo, 0, 0);
        add_segment(gp.rings, nlat, star, thelo, thehi, nlatfill, nlngfill);
        add_segment(gp.rings, nlat, star, thehi, PI, 0, 0);
    }
    // The rings may hold one or two faces more than numface() allots (nlat =
    // 153, 297, 315, 333, 339, ...): the reference writes those past the end
    // of its vector, so they are not part of the star; the face kernel
    // drops every face at or beyond nface.
    if (plain) {
        gp.ring_faces = plain->faces;
        gp.ring_halves = plain->halves;
    } else {
        gp.ring_faces = gp.ring_halves = 0;
        for (auto &R : gp.rings) {
            gp.ring_faces += R.nphi;
            gp.ring_halves += (R.nphi + 1) / 2;
        }
    }
    gp.ring_faces = std::min(gp.ring_faces, G.nface);

    // geometry scalars of the face and eclipse kernels
    const double ri = lcx::deg2rad(m.iangle.value);
    G.cosi = std::cos(ri);
    G.sini = std::sin(ri);
    G.q = q;
    G.npole = m.npole ? 1 : 0;
    G.acc = ACC;
    G.delta = delta;
    G.roche_self = star == 0 ? m.roche1 : m.roche2;
    G.roche_other = star == 0 ? m.roche2 : m.roche1;
    G.spin_self = star == 0 ? m.spin1.value : m.spin2.value;
    G.spin_other = star == 0 ? m.spin2.value : m.spin1.value;
    G.rref_self = star == 0 ? rref1 : rref2;
    G.pref_self = star == 0 ? pref1 : pref2;
    G.rref_other = star == 0 ? rref2 : rref1;
    G.pref_other = star == 0 ? pref2 : pref1;
    G.r_self = star == 0 ? r1 : r2;
    G.r_other = star == 0 ? r2 : r1;
    return gp;
}

// Continuum and flux-sum constants of a grid (set_star_continuum.cpp:30-101,
// 205-237; light_curve_comp.cpp:141-154, 419-423).
void fill_physics(const lc_model &m, GridDev &G) {
    const int star = G.star;
    const double q = m.q.value;
    double r1, r2;
    get_r1r2(m, r1, r2);
    double rl1 = l1(0, q, 0.);
    if (r1 < 0) r1 = rl1;
    double rl2 = 1. - rl1;
    if (r2 < 0) r2 = rl2;
    const double T1 = m.t1.value, T2 = m.t2.value;
    G.wave = m.wavelength;
    G.absorb = m.absorb.value;
    G.mirror = m.mirror ? 1 : 0;
    if (star == 0) {
        G.r_cont_other = r2;
        G.t_base = T1;
        G.t_irr = T2;
        G.t_mirror = std::fabs(T2);
        G.gdcbol = m.gdark_bolom1 ? m.gravity_dark1.value : m.gravity_dark1.value / lcx::dlpdlt(m.wavelength, T1);
        G.spot_running = 0;
    } else {
        G.r_cont_other = r1;
        G.t_base = std::fabs(T2);
        G.t_irr = T1;
        G.t_mirror = T1;
        G.gdcbol = m.gdark_bolom2 ? m.gravity_dark2.value
                                  : m.gravity_dark2.value / lcx::dlpdlt(m.wavelength, std::fabs(T2));
        G.spot_running = 1;
    }
    const lc_pparam *spots[5][4] = {{&m.stsp11_long, &m.stsp11_lat, &m.stsp11_fwhm, &m.stsp11_tcen},
                                    {&m.stsp12_long, &m.stsp12_lat, &m.stsp12_fwhm, &m.stsp12_tcen},
                                    {&m.stsp13_long, &m.stsp13_lat, &m.stsp13_fwhm, &m.stsp13_tcen},
                                    {&m.stsp21_long, &m.stsp21_lat, &m.stsp21_fwhm, &m.stsp21_tcen},
                                    {&m.stsp22_long, &m.stsp22_lat, &m.stsp22_fwhm, &m.stsp22_tcen}};
    G.nspot = 0;
    for (int g = star == 0 ? 0 : 3; g < (star == 0 ? 3 : 5); g++) {
        bool on = spots[g][0]->defined && spots[g][1]->defined && spots[g][2]->defined && spots[g][3]->defined;
        if (!on) continue;
        double clong = std::cos(lcx::deg2rad(spots[g][0]->value));
        double slong = std::sin(lcx::deg2rad(spots[g][0]->value));
        double clat = std::cos(lcx::deg2rad(spots[g][1]->value));
        double slat = std::sin(lcx::deg2rad(spots[g][1]->value));
        int k = G.nspot++;
        G.spot_dir[k][0] = clat * clong;
        G.spot_dir[k][1] = clat * slong;
        G.spot_dir[k][2] = slat;
        G.spot_fwhm[k] = spots[g][2]->value;
        G.spot_tcen[k] = spots[g][3]->value;
    }
    G.uespot = 0;
    if (star == 0 && m.uesp_long1.defined && m.uesp_long2.defined && m.uesp_lathw.defined && m.uesp_taper.defined &&
        m.uesp_temp.defined) {
        G.uespot = 1;
        G.ue_longhw = (m.uesp_long2.value - m.uesp_long1.value) / 2.;
        double lcen = lcx::deg2rad((m.uesp_long1.value + m.uesp_long2.value) / 2.);
        G.ue_dir[0] = std::cos(lcen);
        G.ue_dir[1] = std::sin(lcen);
        G.ue_dir[2] = 0.;
        G.ue_lathw = m.uesp_lathw.value;
        G.ue_taper = m.uesp_taper.value;
        G.ue_temp = m.uesp_temp.value;
    }
    // flux sums
    const float vs = (float)m.velocity_scale.value;
    G.VFAC = vs / (lcx::C_SI / 1.e3);
    G.XCOFM = q / (1.0 + q);
    G.beam = star == 0 ? m.beam_factor1.value : m.beam_factor2.value;
    // With iscale each star's curve is a separate component summed by
    // comp_star1_flat / comp_star2_flat, which use spin = 1 in the beaming
    // term (comp_light_flat.cpp:565-568, 596-599, light_curve_comp.cpp:389-413)
    G.spin_beam = m.iscale ? 1.0 : (star == 0 ? m.spin1.value : m.spin2.value);
    const lc_pparam *l = star == 0 ? &m.ldc1_1 : &m.ldc2_1;
    const lc_pparam *ls[4] = {star == 0 ? &m.ldc1_1 : &m.ldc2_1, star == 0 ? &m.ldc1_2 : &m.ldc2_2,
                              star == 0 ? &m.ldc1_3 : &m.ldc2_3, star == 0 ? &m.ldc1_4 : &m.ldc2_4};
    (void)l;
    for (int k = 0; k < 4; k++) G.ldc[k] = ls[k]->value;
    G.mucrit = star == 0 ? m.mucrit1 : m.mucrit2;
    G.ltype = star == 0 ? m.limb1 : m.limb2;
    G.glens = (star == 1 && m.glens1) ? 1 : 0;
    G.rlens1 = 0.;
    if (m.glens1) {
        double gm = std::pow(1000. * m.velocity_scale.value, 3.0) * m.tperiod * lcx::DAY / TWOPI;
        double a = std::pow(gm / (TWOPI / lcx::DAY / m.tperiod * TWOPI / lcx::DAY / m.tperiod), 1. / 3.);
        G.rlens1 = 4. * gm / (1. + q) / a / (lcx::C_SI * lcx::C_SI);
    }
}

int flux_variant(const GridDev &G) {
    return (G.ltype ? 1 : 0) | (G.beam != 0.0 ? 2 : 0) | (G.glens ? 4 : 0) | (G.mucrit > 0.0 ? 8 : 0);
}

struct InstPlan {
    int status = LC_OK;
    std::string msg;
    int g[4] = {-1, -1, -1, -1};  // 1f, 2f, 1c, 2c
    bool copy1 = true, copy2 = true;
    double gref_diag[2] = {0., 0.};
    double gscale = 0.;
};

}  // namespace

// Device copies of grids kept across calls, keyed by every input of the
// face and eclipse kernels.  A bump allocator over one allocation: when a
// new grid does not fit, the whole cache is dropped and refilled.
constexpr size_t DEFAULT_GEO_CACHE_FACES = size_t(1) << 22;  // 4M faces, ~300 MB

struct GeoCache {
    struct Entry {
        int off, n;
    };
    struct Rec {
        std::vector<RingHost> rings;
        Entry e;
        ContSig cont;  // continuum of the fluxes kept with the faces
    };
    size_t cap = 0;  // faces (0: no reuse across calls)
    size_t used = 0;
    void *mem = nullptr;
    FaceArrays pool{};
    std::unordered_map<GeoSig, std::vector<Rec>, GeoSigHash> map;  // by geom_sig
    void clear() {
        map.clear();
        used = 0;
    }
    void release() {
        clear();
        if (mem) (void)hipFree(mem);
        mem = nullptr;
        std::memset(&pool, 0, sizeof(pool));
    }
    // bytes per cached face: positions and normals (6 x 8), area, gravity,
    // eclipse flag, flux (4 x 4), ingress and egress (2 x 8)
    static constexpr size_t BYTES_PER_FACE = 6 * 8 + 4 * 4 + 2 * 8;
    // false when the pool cannot be allocated (reuse across calls is then off)
    bool ensure_pool() {
        if (mem) return true;
        if (cap == 0) return false;
        if (hipMalloc(&mem, cap * BYTES_PER_FACE + 16 * 256) != hipSuccess) {
            (void)hipGetLastError();
            mem = nullptr;
            cap = 0;
            clear();
            return false;
        }
        char *p = static_cast<char *>(mem);
        auto take = [&](size_t bytes) {
            char *r = p;
            p += (bytes + 255) & ~(size_t)255;
            return r;
        };
        pool.px = reinterpret_cast<double *>(take(cap * 8));
        pool.py = reinterpret_cast<double *>(take(cap * 8));
        pool.pz = reinterpret_cast<double *>(take(cap * 8));
        pool.dx = reinterpret_cast<double *>(take(cap * 8));
        pool.dy = reinterpret_cast<double *>(take(cap * 8));
        pool.dz = reinterpret_cast<double *>(take(cap * 8));
        pool.area = reinterpret_cast<float *>(take(cap * 4));
        pool.grav = reinterpret_cast<float *>(take(cap * 4));
        pool.ecl = reinterpret_cast<int *>(take(cap * 4));
        pool.ein = reinterpret_cast<double *>(take(cap * 8));
        pool.eout = reinterpret_cast<double *>(take(cap * 8));
        pool.flux = reinterpret_cast<float *>(take(cap * 4));
        return true;
    }
};

// ================================================================== ctx
struct lc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevArena dev;
    HostPinned pin;
    lc_profile prof;
    hipEvent_t ev[16];
    hipEvent_t ev_var[17];  // boundaries between flux-variant launches
    // flux tile shape; 0: chosen per call (R = 2 with beaming, else 1; the
    // chunk from the number of flux tasks); last_* hold the last call's
    int flux_rphase = 0;
    int flux_chunk = 0;
    int last_rphase = 1, last_chunk = 1024;
    long long last_part = 0;  // the last call's chunk partial sums (doubles)
    // mixed-precision sums without beaming or lensing: per-wave phase tiles
    // (k_flux_wave); 0: the workgroup-tile kernel k_flux_mixed
    int flux_wave = 1;
    bool dedupe = true;  // build grids equal within one batch once
    bool ecl_stop = true;  // eclipse searches end once no data phase can tell
    // below this many faces per call the searches are latency bound: the
    // longest one sets the time and stopping others early saves nothing
    long long ecl_stop_min_faces = 1LL << 17;
    long long ecl_round_min = 1LL << 21;  // faces x 2 from which early-stopping searches run in rounds
    // lanes per search of the bisections without early stops
    // (k_eclipse_bisect_wide; 0 or 1: one lane per search); -1: 16, 8 or 4,
    // the most that keep faces x lanes within ecl_wide_budget, else one
    // (measured on the small detached model at nlat 40 ... 120: the wide
    // search pays off while it does not oversubscribe the device)
    int ecl_wide = -1;
    long long ecl_wide_budget = 1LL << 18;
    int wide_lanes(long long nface) const {
        if (ecl_wide >= 0) return ecl_wide;
        for (int g : {16, 8, 4})
            if (nface * g <= ecl_wide_budget) return g;
        return 0;
    }
    GeoCache geo;
    // kernel sequence of the last repeated launch shape, as a graph
    bool graphs = true;
    std::string g_key, g_seen;
    hipGraph_t g_graph = nullptr;
    hipGraphExec_t g_exec = nullptr;
    void drop_graph() {
        if (g_exec) (void)hipGraphExecDestroy(g_exec);
        if (g_graph) (void)hipGraphDestroy(g_graph);
        g_exec = nullptr;
        g_graph = nullptr;
        g_key.clear();
    }
};

namespace {

const char *status_msg(int st) {
    switch (st) {
    case LC_E_FACE: return "Roche::face error: could not solve for a face radius";
    case LC_E_ECLIPSE_SEARCH: return "Roche::ingress_egress: eclipse search failed (pot_min/linmin/dbrent)";
    default: return "device evaluation failed";
    }
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.;
    return ms;
}

// Algorithmic FLOP per face x sub-phase element (POLY/CLARET x beaming):
// the counts the roofline uses.
double elem_flops(const GridDev &G) {
    static const double c[4] = {24., 26., 39., 41.};
    return c[(G.ltype ? 1 : 0) + (G.beam != 0.0 ? 2 : 0)];
}

struct EvalRequest {
    const lc_model *models;
    size_t m;
    const lc_datum *data;
    size_t n;
    lc_eval_opts opts;
    double *calc;       // m x n (host)
    lc_scalars *io;     // m
    int32_t *status;    // m
    double *dev_calc = nullptr;  // m x n on the device instead of calc
};

// Runs the whole pipeline for a batch of instances sharing the data set.
void run_batch(lc_ctx *ctx, const EvalRequest &rq) {
    const auto t_call0 = std::chrono::steady_clock::now();
    const size_t M = rq.m, n = rq.n;
    const lc_eval_opts &o = rq.opts;
    const bool fp64 = o.precision == LC_PREC_FP64;

    // ---------------- data layout (shared by all instances)
    std::vector<int> sub_off(n + 1);
    sub_off[0] = 0;
    for (size_t i = 0; i < n; i++) sub_off[i + 1] = sub_off[i] + std::max(0, rq.data[i].ndiv);
    const int S = sub_off[n];
    double xmin = rq.data[0].time, xmax = rq.data[0].time;
    for (size_t i = 1; i < n; i++) {
        xmin = rq.data[i].time > xmin ? xmin : rq.data[i].time;
        xmax = rq.data[i].time < xmax ? xmax : rq.data[i].time;
    }
    const double middle = (xmin + xmax) / 2., range = (xmax - xmin) / 2.;

    // ---------------- per instance planning (host threads), then the grids
    // are numbered in instance order
    std::vector<InstPlan> ip(M);
    std::vector<InstDev> insts(M);
    std::vector<std::vector<GridPlan>> local(M);
    std::vector<CompHost> chost(M);  // disc, edge and bright spot elements
    parallel_for(M, [&](size_t i) {
        const lc_model &m = rq.models[i];
        InstPlan &P = ip[i];
        InstDev &I = insts[i];
        std::memset(&I, 0, sizeof(I));
        try {
            double r1, r2;
            get_r1r2(m, r1, r2);
            double rl2 = 1. - l1(2, m.q.value, m.spin2.value);
            if (r2 < 0)
                r2 = rl2;
            else if (r2 > rl2)
                throw EngineError(LC_E_ROCHE_OVERFLOW,
                                  "light_curve_comp: the secondary star is larger than its Roche lobe!");
            std::vector<GridPlan> &L = local[i];
            L.push_back(plan_grid(m, 0, true));
            L.push_back(plan_grid(m, 1, true));
            P.copy1 = (m.nlat1f == m.nlat1c);
            P.copy2 = (m.nlat2f == m.nlat2c) && (!m.npole || r1 >= r2 || (m.nlatfill == 0 && m.nlngfill == 0));
            if (!P.copy1) L.push_back(plan_grid(m, 0, false));
            if (!P.copy2) L.push_back(plan_grid(m, 1, false));
            for (auto &gp : L) {
                gp.dev.inst = (int)i;
                fill_physics(m, gp.dev);
            }
            // instance parameters
            const double ri = lcx::deg2rad(m.iangle.value);
            I.t0 = m.t0.value;
            I.period = m.period.value;
            I.pdot = m.pdot.value;
            I.deltat = m.deltat.value;
            I.slope = m.slope.value;
            I.quad = m.quad.value;
            I.cube = m.cube.value;
            I.third = m.third.value;
            I.cosi = std::cos(ri);
            I.sini = std::sin(ri);
            I.phase1 = m.phase1;
            I.phase2 = m.phase2;
            I.scale11 = I.scale12 = I.scale21 = I.scale22 = 1.;
            I.sfac0 = rq.io[i].sfac[0];
            // iscale: each star scaled on its own (light_curve_comp.cpp:378-416, 511-573)
            I.iscale = m.iscale ? 1 : 0;
            I.star2 = m.t2.value > 0. ? 1 : 0;
            I.sfac4 = rq.io[i].sfac[4];
            I.zcols = rq.io[i].sfac[1] * 0. + rq.io[i].sfac[2] * 0. + rq.io[i].sfac[3] * 0.;
            I.sfac1 = rq.io[i].sfac[1];
            I.sfac2 = rq.io[i].sfac[2];
            I.sfac3 = rq.io[i].sfac[3];
            I.copy1 = P.copy1;
            I.copy2 = P.copy2;
            I.scale = o.scale;
            I.rdata = o.rdata;
            if (m.add_disc || m.add_spot) {
                std::string cerr;
                const int cst = plan_components(m, chost[i], cerr);
                if (cst) throw EngineError(cst, cerr);
                I.has_comp = chost[i].size() > 0 ? 1 : 0;
            }
            I.active = 1;
        } catch (const EngineError &e) {
            P.status = e.code;
            P.msg = e.what();
            I.active = 0;
            local[i].clear();
        }
    });
    std::vector<GridPlan> grids;
    for (size_t i = 0; i < M; i++) {
        if (ip[i].status != LC_OK) continue;
        InstPlan &P = ip[i];
        InstDev &I = insts[i];
        for (auto &gp : local[i]) {
            P.g[gp.dev.star + (gp.dev.fine ? 0 : 2)] = (int)grids.size();
            grids.push_back(std::move(gp));
        }
        if (P.copy1) P.g[2] = P.g[0];
        if (P.copy2) P.g[3] = P.g[1];
        I.grid1f = P.g[0];
        I.grid2f = P.g[1];
        I.grid1c = P.g[2];
        I.grid2c = P.g[3];
        if (o.info) {
            std::fprintf(stderr, "Number of points for star 1 (fine) = %d\n", grids[P.g[0]].dev.nface);
            std::fprintf(stderr, "Number of points for star 2 (fine) = %d\n", grids[P.g[1]].dev.nface);
            std::fprintf(stderr, "Number of points for star 1 (coarse) = %d\n", grids[P.g[2]].dev.nface);
            std::fprintf(stderr, "Number of points for star 2 (coarse) = %d\n", grids[P.g[3]].dev.nface);
        }
    }

    // ---------------- stage reuse: grids equal to an earlier grid of the
    // batch (dup_of) or to a grid kept from an earlier call (hit_off) are
    // copied; the rest are computed (and offered to the cache).
    const int ngrid = (int)grids.size();
    std::vector<int> dup_of(ngrid, -1), hit_off(ngrid, -1);
    std::vector<char> hit_flux(ngrid, 0);  // the cache record also holds this grid's fluxes
    std::vector<GeoSig> sigs(ngrid);
    std::vector<ContSig> csig(ngrid);
    parallel_for((size_t)ngrid, [&](size_t g) {
        sigs[g] = geom_sig(grids[g]);
        csig[g] = cont_sig(grids[g].dev);
    });
    {
        std::unordered_map<GeoSig, std::vector<int>, GeoSigHash> first;  // computed grids by signature
        first.reserve((size_t)ngrid);
        for (int g = 0; g < ngrid; g++) {
            if (ctx->dedupe) {
                auto it = first.find(sigs[g]);
                if (it != first.end())
                    for (int h : it->second)
                        if (same_rings(grids[h].rings, grids[g].rings)) {
                            dup_of[g] = h;
                            break;
                        }
                if (dup_of[g] >= 0) continue;
            }
            first[sigs[g]].push_back(g);
            auto c = ctx->geo.map.find(sigs[g]);
            if (c != ctx->geo.map.end())
                for (const auto &rec : c->second)
                    if (rec.e.n == grids[g].dev.nface && same_rings(rec.rings, grids[g].rings)) {
                        hit_off[g] = rec.e.off;
                        hit_flux[g] = rec.cont == csig[g];
                        break;
                    }
        }
    }
    auto computed = [&](int g) { return dup_of[g] < 0 && hit_off[g] < 0; };

    // ---------------- global tables
    std::vector<GridDev> gdev(ngrid);
    int face_cur = 0, half_cur = 0, ring_cur = 0;
    for (int g = 0; g < ngrid; g++) {
        GridPlan &gp = grids[g];
        GridDev &G = gp.dev;
        G.face_off = face_cur;
        G.ring_off = ring_cur;
        G.nring = computed(g) ? (int)gp.rings.size() : 0;
        G.half_off = half_cur;
        const int nh = G.nring ? gp.ring_halves : 0;
        G.nhalf = nh;
        half_cur += nh;
        ring_cur += G.nring;
        face_cur += G.nface;
        gdev[g] = G;
    }
    // ring table of grid g, written by the planning threads straight into
    // the pinned staging block (below)
    auto fill_rings = [&](size_t g, RingDev *rdev) {
        const GridDev &G = gdev[g];
        int fo = G.face_off, ho = G.half_off;
        for (int k = 0; k < G.nring; k++) {
            const RingHost &R = grids[g].rings[k];
            RingDev &d = rdev[G.ring_off + k];
            d.sint = R.sint;
            d.cost = R.cost;
            d.phi1 = R.phi1;
            d.dphi_span = R.dphi_span;
            d.tspan_nl = R.tspan_nl;
            d.nphi = R.nphi;
            d.face_off = fo;
            d.half_off = ho;
            d.grid = (int)g;
            fo += R.nphi;
            ho += (R.nphi + 1) / 2;
        }
    };
    const int nface_total = face_cur, nhalf_total = half_cur, nring = ring_cur;
    if ((long long)face_cur > (1LL << 30)) throw EngineError(LC_E_ARG, "batch too large: more than 2^30 faces");

    // copies of reused grids, and computed grids offered to the cache
    std::vector<FaceCopy> copy_hit, copy_dup, copy_out;
    std::vector<std::pair<int, GeoCache::Entry>> inserts;  // grid, pool slot
    for (int g = 0; g < ngrid; g++) {
        if (hit_off[g] >= 0) copy_hit.push_back(FaceCopy{hit_off[g], gdev[g].face_off, gdev[g].nface, 0});
        if (dup_of[g] >= 0) copy_dup.push_back(FaceCopy{gdev[dup_of[g]].face_off, gdev[g].face_off, gdev[g].nface, 0});
    }
    if (ctx->geo.cap > 0) {
        size_t need = 0;
        for (int g = 0; g < ngrid; g++)
            if (computed(g) && (size_t)gdev[g].nface <= ctx->geo.cap) need += gdev[g].nface;
        if (need && ctx->geo.used + need > ctx->geo.cap) ctx->geo.clear();
        for (int g = 0; g < ngrid; g++) {
            if (!computed(g) || gdev[g].nface == 0) continue;
            if (ctx->geo.used + (size_t)gdev[g].nface > ctx->geo.cap) continue;
            GeoCache::Entry e{(int)ctx->geo.used, gdev[g].nface};
            ctx->geo.used += gdev[g].nface;
            inserts.push_back({g, e});
            copy_out.push_back(FaceCopy{gdev[g].face_off, e.off, e.n, 0});
        }
        if (!copy_out.empty() && !ctx->geo.ensure_pool()) {
            copy_out.clear();
            inserts.clear();
        }
    }

    // continuum reuse: a grid whose faces are copied and whose continuum
    // inputs equal its source's takes the source's fluxes too -- from the
    // cache, or from a grid of this call after k_continuum computed it; new
    // grids offered to the cache take their fluxes along
    std::vector<FaceCopy> flux_hit, flux_dup, flux_out;
    for (int g = 0; g < ngrid; g++)
        if (hit_off[g] >= 0 && hit_flux[g]) {
            flux_hit.push_back(FaceCopy{hit_off[g], gdev[g].face_off, gdev[g].nface, 0});
            gdev[g].cont_copy = 1;
        }
    for (int g = 0; g < ngrid; g++) {
        const int h = dup_of[g];
        if (h < 0 || !(csig[g] == csig[h])) continue;
        if (gdev[h].cont_copy)  // the source takes cached fluxes: so does this copy
            flux_hit.push_back(FaceCopy{hit_off[h], gdev[g].face_off, gdev[g].nface, 0});
        else
            flux_dup.push_back(FaceCopy{gdev[h].face_off, gdev[g].face_off, gdev[g].nface, 0});
        gdev[g].cont_copy = 1;
    }
    for (const auto &ins : inserts)
        flux_out.push_back(FaceCopy{gdev[ins.first].face_off, ins.second.off, ins.second.n, 0});

    // disc, edge and bright spot: their elements in one table (instance
    // order), each in two slots of the component face arrays -- [off, off+n)
    // searched for eclipses by star 1, [NC + off, ...) by star 2 -- through
    // pseudo-grids that name that star as the companion
    std::vector<CompDev> comps;
    int NC = 0;
    for (size_t i = 0; i < M; i++) {
        if (ip[i].status != LC_OK || !insts[i].has_comp) continue;
        const CompHost &c = chost[i];
        CompDev C;
        std::memset(&C, 0, sizeof(C));
        C.inst = (int)i;
        C.off = NC;
        C.n_disc = c.n_disc;
        C.n_edge = c.n_edge;
        C.n_spot = c.n_spot;
        C.opaque = c.opaque;
        C.lin = c.lin;
        C.quad = c.quad;
        C.rdisc1 = c.rdisc1;
        C.rdisc2 = c.rdisc2;
        C.beta = c.beta;
        C.height = c.height;
        C.iangle = c.iangle;
        C.sum_off = (long long)comps.size() * 3 * S;
        insts[i].comp_off = C.sum_off;
        comps.push_back(C);
        NC += c.size();
    }
    const int ncomp = (int)comps.size();
    // an opaque disc hides parts of both stars: their grids get the disc's
    // ranges (k_disc_star_ranges) and ragged flux sums (k_flux_ragged)
    std::vector<int> dgrids;
    int dmax_face = 0;
    for (int k = 0; k < ncomp; k++) {
        if (!comps[k].opaque) continue;
        const InstPlan &P = ip[comps[k].inst];
        for (int q = 0; q < 4; q++) {
            const int g = P.g[q];
            if (g < 0 || gdev[g].disc_comp >= 0) continue;
            gdev[g].disc_comp = k;
            if (gdev[g].nface > 0) {
                dgrids.push_back(g);
                dmax_face = std::max(dmax_face, gdev[g].nface);
            }
        }
    }
    const bool any_disc = !dgrids.empty();
    std::vector<GridDev> cgrids;
    std::vector<double> cpx, cpy, cpz, cdx, cdy, cdz;
    std::vector<float> cflux;
    if (ncomp) {
        cgrids.resize(2 * ncomp);
        for (int by = 0; by < 2; by++)
            for (int k = 0; k < ncomp; k++) {
                const CompHost &c = chost[comps[k].inst];
                GridDev &G = cgrids[by * ncomp + k];
                std::memset(&G, 0, sizeof(G));
                G.disc_comp = -1;
                G.inst = comps[k].inst;
                G.nface = c.size();
                G.face_off = by * NC + comps[k].off;
                // by star 1: the element plays a star-2 face (companion star
                // 1), by star 2 a star-1 face (star_eclipse.cpp:15-25)
                G.star = by == 0 ? 1 : 0;
                G.eclipse = by == 0 ? c.eclipse1 : c.eclipse2;
                G.roche_other = by == 0 ? c.roche1 : c.roche2;
                G.q = c.q;
                G.spin_other = by == 0 ? c.spin1 : c.spin2;
                G.rref_other = by == 0 ? c.rref1 : c.rref2;
                G.pref_other = by == 0 ? c.pref1 : c.pref2;
                G.r_other = by == 0 ? c.r1 : c.r2;
                G.cosi = c.cosi;
                G.sini = c.sini;
                G.delta = c.delta;
                G.acc = c.delta / 10.;
                G.cont_copy = 1;
            }
        cpx.resize(2 * (size_t)NC); cpy.resize(2 * (size_t)NC); cpz.resize(2 * (size_t)NC);
        cdx.resize(NC); cdy.resize(NC); cdz.resize(NC); cflux.resize(NC);
        for (const CompDev &C : comps) {
            const CompHost &c = chost[C.inst];
            for (int k = 0; k < c.size(); k++) {
                const size_t f = (size_t)C.off + k;
                cpx[f] = cpx[NC + f] = c.px[k];
                cpy[f] = cpy[NC + f] = c.py[k];
                cpz[f] = cpz[NC + f] = c.pz[k];
                cdx[f] = c.dx[k];
                cdy[f] = c.dy[k];
                cdz[f] = c.dz[k];
                cflux[f] = c.flux[k];
            }
        }
    }
    const size_t CN = (size_t)std::max(NC, 1), CN2 = 2 * CN;

    // Eclipse searches that may stop early: grids used by one instance only
    // (no in-batch copies) and offered to no cache, whose ranges are then
    // tested only at that instance's phases.
    bool any_ecl_stop = false;
    {
        std::vector<char> shared(ngrid, 0);
        for (int g = 0; g < ngrid; g++)
            if (dup_of[g] >= 0) shared[dup_of[g]] = 1;
        for (const auto &ins : inserts) shared[ins.first] = 1;
        const bool stop = ctx->ecl_stop && nface_total >= ctx->ecl_stop_min_faces;
        for (int g = 0; g < ngrid; g++) {
            gdev[g].ecl_stop = (stop && computed(g) && !shared[g]) ? 1 : 0;
            any_ecl_stop = any_ecl_stop || gdev[g].ecl_stop;
        }
    }
    // phase bins: about 64 bins per phase, 2^10 .. 2^22 per instance
    int bits_lb = 10;
    while (bits_lb < 22 && (double)(1u << bits_lb) < 64.0 * (S + 9)) bits_lb++;
    const size_t bits_words = any_ecl_stop ? (std::max<size_t>(M, 1) << (bits_lb - 5)) : 1;

    // FlatGrid build segments
    std::vector<FlatSeg> fsegs;
    for (int g = 0; g < ngrid; g++) {
        const int nf = gdev[g].nface, ns = std::max(1, (nf + FLAT_SEG - 1) / FLAT_SEG), s0 = (int)fsegs.size();
        for (int k = 0; k < ns; k++)
            fsegs.push_back(FlatSeg{g, k * FLAT_SEG, std::min(FLAT_SEG, nf - k * FLAT_SEG), s0, ns, {0, 0, 0}});
    }

    // gaps between the ring layout and numface() hold default faces
    std::vector<int2> gaps;  // (first face, count)
    for (int g = 0; g < ngrid; g++)
        if (computed(g) && grids[g].ring_faces < gdev[g].nface)
            gaps.push_back(make_int2(gdev[g].face_off + grids[g].ring_faces, gdev[g].nface - grids[g].ring_faces));

    // ---------------- flux batches and tasks
    int R = ctx->flux_rphase;
    if (R <= 0) {  // beaming elements cost twice as much: two sub-phases per lane pay off
        R = 1;
        for (int g = 0; g < ngrid; g++)
            if (gdev[g].beam != 0.0) R = 2;
    }
    const int TILE = R * 256;
    const int ntile = S > 0 ? (S + TILE - 1) / TILE : 0;
    auto count_tasks = [&](int ch) {
        long long t = 0;
        for (size_t i = 0; i < M; i++) {
            if (ip[i].status != LC_OK) continue;
            for (int star = 0; star < 2; star++) {
                const bool copy = star == 0 ? ip[i].copy1 : ip[i].copy2;
                for (int part = 0; part < (copy ? 1 : 2); part++)
                    t += (long long)ntile * ((gdev[ip[i].g[star + (part ? 2 : 0)]].nface + ch - 1) / ch);
            }
        }
        return t;
    };
    int chunk = fp64 ? 1024 : ctx->flux_chunk;
    if (chunk <= 0) {  // the largest face chunk that still gives every CU several tasks
        chunk = 2048;
        while (chunk > 64 && count_tasks(chunk) < 2048) chunk /= 2;
    }
    ctx->last_rphase = R;
    ctx->last_chunk = chunk;
    ctx->prof.last_rphase = R;
    ctx->prof.last_chunk = chunk;
    const bool any_route = std::any_of(ip.begin(), ip.end(), [](const InstPlan &p) {
        return p.status == LC_OK && (!p.copy1 || !p.copy2);
    });
    std::vector<BatchDev> batches;
    std::vector<std::vector<int>> batches_by_var(16), ragged_by_var(8);
    long long part_total = 0;
    double alg_flops = 0., alg_elems = 0., alg_bytes = 0.;
    double var_flops[16] = {0.};
    for (size_t i = 0; i < M; i++) {
        if (ip[i].status != LC_OK || S == 0) continue;
        for (int star = 0; star < 2; star++) {
            bool copy = star == 0 ? ip[i].copy1 : ip[i].copy2;
            for (int part = 0; part < (copy ? 1 : 2); part++) {
                int gi = ip[i].g[star + (part ? 2 : 0)];
                const GridDev &G = gdev[gi];
                if (G.nface == 0) continue;
                BatchDev B;
                B.inst = (int)i;
                B.star = star;
                B.grid = gi;
                B.list_off = copy ? -1 : (int)(((size_t)i * 2 + star) * 2 * S + (part ? S : 0));
                B.count_idx = copy ? -1 : (int)((i * 2 + star) * 2 + part);
                B.chunk = chunk;
                B.nchunk = (G.nface + chunk - 1) / chunk;
                B.ntile = ntile;
                B.part_off = part_total;
                part_total += (long long)B.nchunk * ntile * TILE;
                int var = flux_variant(G);
                if (G.disc_comp >= 0) {  // one partial row, k_flux_ragged
                    part_total -= (long long)B.nchunk * ntile * TILE;
                    B.nchunk = 1;
                    part_total += (long long)ntile * TILE;
                    ragged_by_var[var & 7].push_back((int)batches.size());
                    batches.push_back(B);
                    continue;
                }
                int bi = (int)batches.size();
                batches.push_back(B);
                batches_by_var[var].push_back(bi);
                if (copy) {  // algorithmic work: every face x every sub-phase of the grid
                    alg_elems += (double)G.nface * S;
                    alg_flops += (double)G.nface * S * elem_flops(G);
                    alg_bytes += (double)G.nface * 40. + (double)S * 32.;
                    var_flops[var] += (double)G.nface * S * elem_flops(G);
                }
            }
        }
    }
    // the (batch, tile, chunk) task table is expanded on the device from
    // each batch's first task; a variant's tasks are contiguous
    std::vector<int> task_first, task_order;
    std::vector<std::pair<int, int>> var_range(16, {0, 0});
    int ntask = 0;
    for (int v = 0; v < 16; v++) {
        const int v0 = ntask;
        for (int b : batches_by_var[v]) {
            task_order.push_back(b);
            task_first.push_back(ntask);
            ntask += batches[b].ntile * batches[b].nchunk;
        }
        var_range[v] = {v0, ntask - v0};
    }
    task_first.push_back(ntask);
    std::vector<int> ragged_order;
    std::vector<std::pair<int, int>> ragged_range(8, {0, 0});
    for (int v = 0; v < 8; v++) {
        ragged_range[v] = {(int)ragged_order.size(), (int)ragged_by_var[v].size()};
        ragged_order.insert(ragged_order.end(), ragged_by_var[v].begin(), ragged_by_var[v].end());
    }

    // diagnostics grids: star 1 fine, star 2 fine, and the wdwarf grid
    std::vector<int> diag_idx;
    std::vector<std::array<int, 3>> diag_of(M, std::array<int, 3>{-1, -1, -1});
    if (o.diagnostics)
        for (size_t i = 0; i < M; i++) {
            if (ip[i].status != LC_OK) continue;
            const lc_model &m = rq.models[i];
            double pn = 0.5;
            bool t1 = (pn <= m.phase1 || pn >= 1. - m.phase1);
            int wgrid = t1 ? ip[i].g[0] : ip[i].g[2];
            diag_of[i][0] = (int)diag_idx.size();
            diag_idx.push_back(ip[i].g[0]);
            diag_of[i][1] = (int)diag_idx.size();
            diag_idx.push_back(ip[i].g[1]);
            if (wgrid == ip[i].g[0]) {
                diag_of[i][2] = diag_of[i][0];
            } else {
                diag_of[i][2] = (int)diag_idx.size();
                diag_idx.push_back(wgrid);
            }
        }

    // ---------------- device memory
    const size_t FN = (size_t)std::max(nface_total, 1);
    const size_t SM = (size_t)std::max(S, 1) * std::max<size_t>(M, 1);
    Sizer sz;
    sz.add<GridDev>(std::max(ngrid, 1));
    sz.add<RingDev>(std::max(nring, 1));
    sz.add<InstDev>(std::max<size_t>(M, 1));
    sz.add<int>(M);  // status
    for (int k = 0; k < 6; k++) sz.add<double>(FN);  // px..dz
    for (int k = 0; k < 3; k++) sz.add<float>(FN);   // area grav flux
    sz.add<int>(FN);
    sz.add<double>(FN);
    sz.add<double>(FN);
    sz.add<int2>(FN);  // cand
    // components: uploaded tables, slots, their eclipse search, records, sums
    sz.add<CompDev>(std::max(ncomp, 1));
    sz.add<GridDev>(std::max(2 * ncomp, 1));
    for (int k = 0; k < 3; k++) sz.add<double>(CN2);
    for (int k = 0; k < 3; k++) sz.add<double>(CN);
    sz.add<float>(CN);
    sz.add<int>(CN2);
    sz.add<double>(CN2);
    sz.add<double>(CN2);
    sz.add<int2>(CN2);
    sz.add<int>(1);                                     // candidate count
    sz.add<int>(1);                                     // hit count
    for (int k = 0; k < 2; k++) sz.add<int>(CN2);       // face, grid
    for (int k = 0; k < 2; k++) sz.add<int>(2 * CN2);   // round lists
    for (int k = 0; k < 6; k++) sz.add<double>(2 * CN2);  // phi, lim, acc, res, b_in, b_out
    sz.add<int>(ECL_COUNTS);
    for (int k = 0; k < 3; k++) sz.add<double>(CN);
    sz.add<float>(CN);
    sz.add<int>(CN);
    sz.add<double2>(CN * COMP_MAX_RANGES);
    sz.add<double>(std::max<size_t>((size_t)ncomp * 3 * S, 1));
    sz.add<int>(1);
    sz.add<int>(1);      // eclipse work lists
    sz.add<int>(FN);
    sz.add<int>(FN);
    for (int k = 0; k < 6; k++) sz.add<double>(FN);
    for (int k = 0; k < 4; k++) sz.add<double>(FN);  // bisection brackets
    for (int k = 0; k < 4; k++) sz.add<int>(FN);     // bisection round lists
    sz.add<int>(ECL_COUNTS);
    sz.add<float4>(FN);
    sz.add<float4>(FN);
    sz.add<float4>(FN);
    sz.add<double2>(FN);
    if (fp64)
        for (int k = 0; k < 7; k++) sz.add<double>(FN);
    sz.add<int>(std::max(ngrid, 1));
    sz.add<FlatSeg>(fsegs.size());
    sz.add<int>(fsegs.size());
    sz.add<unsigned>(bits_words);
    sz.add<unsigned long long>(16 * 64);
    sz.add<FaceCopy>(std::max<size_t>(copy_hit.size(), 1));
    sz.add<FaceCopy>(std::max<size_t>(copy_dup.size(), 1));
    sz.add<FaceCopy>(std::max<size_t>(copy_out.size(), 1));
    sz.add<FaceCopy>(std::max<size_t>(flux_hit.size(), 1));
    sz.add<FaceCopy>(std::max<size_t>(flux_dup.size(), 1));
    sz.add<FaceCopy>(std::max<size_t>(flux_out.size(), 1));
    sz.add<int2>(std::max<size_t>(gaps.size(), 1));
    for (int k = 0; k < 5; k++) sz.add<double>(std::max<size_t>(n, 1));
    sz.add<int>(std::max<size_t>(n, 1));
    sz.add<int>(n + 1);
    for (int k = 0; k < 6; k++) sz.add<double>(SM);
    sz.add<signed char>(SM);
    if (any_route) {
        sz.add<int>(SM * 4);
        sz.add<int>(std::max<size_t>(M, 1) * 4);
    }
    sz.add<double>(SM * 2);  // out
    sz.add<double>((size_t)std::max<long long>(part_total, 1));
    ctx->last_part = part_total;
    sz.add<BatchDev>(std::max<size_t>(batches.size(), 1));
    sz.add<TaskDev>(std::max(ntask, 1));
    sz.add<int>(task_first.size());
    sz.add<int>(std::max<size_t>(task_order.size(), 1));
    sz.add<double>(std::max<size_t>(M, 1) * n);  // calc
    sz.add<double>(std::max<size_t>(M, 1) * 8);  // scalars
    int max_nface = 1;
    for (int g = 0; g < ngrid; g++) max_nface = std::max(max_nface, gdev[g].nface);
    const int psum_chunks = any_route ? (max_nface + PSUM_FACES - 1) / PSUM_FACES : 1;
    sz.add<double>(std::max<size_t>(M, 1) * 8 * psum_chunks);  // Ginterp partial sums
    const size_t scat_part = std::max<size_t>(M, 1) * std::max<size_t>(1, (n + SCAT_PTS - 1) / SCAT_PTS) * SCAT_PARTS;
    sz.add<double>(scat_part);
    sz.add<int>(std::max<size_t>(diag_idx.size(), 1));
    sz.add<double>(std::max<size_t>(diag_idx.size(), 1) * 8);
    sz.add<int>(std::max<size_t>(ragged_order.size(), 1));
    sz.add<int>(std::max<size_t>(dgrids.size(), 1));
    if (any_disc) {
        sz.add<int>(FN);
        sz.add<double2>(2 * FN);
    }
    // results gathered into one block for a single download: scalars,
    // diagnostics sums, instance records, statuses, the eclipse counts, the
    // evaluated-element counters, the routed sub-phase counts and, last (left
    // out when the curves stay on the device), calc
    auto up8 = [](size_t b) { return (b + 7) & ~(size_t)7; };
    const bool dev_out = rq.dev_calc != nullptr;
    const size_t o_scal = 0, o_diag = o_scal + M * 8 * 8,
                 o_insts = up8(o_diag + diag_idx.size() * 8 * 8), o_status = up8(o_insts + M * sizeof(InstDev)),
                 o_counts = up8(o_status + M * sizeof(int)), o_evald = o_counts + 16,
                 o_route = o_evald + 16 * 64 * sizeof(unsigned long long),
                 o_calc = up8(o_route + (any_route ? M * 4 * sizeof(int) : 0)),
                 down_total = dev_out ? o_calc : o_calc + M * n * 8;
    sz.add<char>(down_total);
    ctx->dev.reserve(sz.total + 4096);
    ctx->dev.reset();
    DevArena &A = ctx->dev;
    // everything uploaded from the host first: the staging block has the
    // same layout and goes up in one copy
    GridDev *d_grids = A.take<GridDev>(std::max(ngrid, 1));
    RingDev *d_rings = A.take<RingDev>(std::max(nring, 1));
    InstDev *d_insts = A.take<InstDev>(std::max<size_t>(M, 1));
    int *d_status = A.take<int>(M);
    FaceCopy *d_copy_hit = A.take<FaceCopy>(std::max<size_t>(copy_hit.size(), 1));
    FaceCopy *d_copy_dup = A.take<FaceCopy>(std::max<size_t>(copy_dup.size(), 1));
    FaceCopy *d_copy_out = A.take<FaceCopy>(std::max<size_t>(copy_out.size(), 1));
    FaceCopy *d_flux_hit = A.take<FaceCopy>(std::max<size_t>(flux_hit.size(), 1));
    FaceCopy *d_flux_dup = A.take<FaceCopy>(std::max<size_t>(flux_dup.size(), 1));
    FaceCopy *d_flux_out = A.take<FaceCopy>(std::max<size_t>(flux_out.size(), 1));
    int2 *d_gaps = A.take<int2>(std::max<size_t>(gaps.size(), 1));
    double *d_time = A.take<double>(std::max<size_t>(n, 1));
    double *d_expose = A.take<double>(std::max<size_t>(n, 1));
    double *d_flux = A.take<double>(std::max<size_t>(n, 1));
    double *d_ferr = A.take<double>(std::max<size_t>(n, 1));
    double *d_weight = A.take<double>(std::max<size_t>(n, 1));
    int *d_ndiv = A.take<int>(std::max<size_t>(n, 1));
    int *d_sub_off = A.take<int>(n + 1);
    BatchDev *d_batches = A.take<BatchDev>(std::max<size_t>(batches.size(), 1));
    int *d_task_first = A.take<int>(task_first.size());
    int *d_task_order = A.take<int>(std::max<size_t>(task_order.size(), 1));
    int *d_diag_idx = A.take<int>(std::max<size_t>(diag_idx.size(), 1));
    int *d_ragged = A.take<int>(std::max<size_t>(ragged_order.size(), 1));
    int *d_dgrids = A.take<int>(std::max<size_t>(dgrids.size(), 1));
    FlatSeg *d_fsegs = A.take<FlatSeg>(fsegs.size());
    int *d_seg_cnt = A.take<int>(fsegs.size());
    CompDev *d_comps = A.take<CompDev>(std::max(ncomp, 1));
    GridDev *d_cgrids = A.take<GridDev>(std::max(2 * ncomp, 1));
    FaceArrays cs;  // component slots: positions (both halves), normals and fluxes (first half)
    std::memset(&cs, 0, sizeof(cs));
    cs.px = A.take<double>(CN2); cs.py = A.take<double>(CN2); cs.pz = A.take<double>(CN2);
    cs.dx = A.take<double>(CN); cs.dy = A.take<double>(CN); cs.dz = A.take<double>(CN);
    cs.flux = A.take<float>(CN);
    FaceArrays fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.px = A.take<double>(FN); fa.py = A.take<double>(FN); fa.pz = A.take<double>(FN);
    fa.dx = A.take<double>(FN); fa.dy = A.take<double>(FN); fa.dz = A.take<double>(FN);
    fa.area = A.take<float>(FN); fa.grav = A.take<float>(FN); fa.flux = A.take<float>(FN);
    fa.ecl = A.take<int>(FN);
    fa.ein = A.take<double>(FN);
    fa.eout = A.take<double>(FN);
    if (any_disc) {
        fa.dn = A.take<int>(FN);
        fa.dr = A.take<double2>(2 * FN);
    }
    int2 *d_cand = A.take<int2>(FN);
    int *d_cand_count = A.take<int>(1);
    EclipseWork ew;
    std::memset(&ew, 0, sizeof(ew));  // padding too: the struct is part of the launch key
    ew.hit_count = A.take<int>(1);
    ew.face = A.take<int>(FN);
    ew.grid = A.take<int>(FN);
    ew.phi = A.take<double>(FN);
    ew.lim = A.take<double>(2 * FN);
    ew.acc = A.take<double>(FN);
    ew.res = A.take<double>(2 * FN);
    ew.b_in = A.take<double>(2 * FN);
    ew.b_out = A.take<double>(2 * FN);
    ew.alist[0] = A.take<int>(2 * FN);
    ew.alist[1] = A.take<int>(2 * FN);
    ew.acount = A.take<int>(ECL_COUNTS);
    unsigned *d_bits = A.take<unsigned>(bits_words);
    ew.phase_bits = d_bits;
    ew.bits_lb = bits_lb;
    FlatArrays fl;
    std::memset(&fl, 0, sizeof(fl));
    fl.rec0 = A.take<float4>(FN);
    fl.rec1 = A.take<float4>(FN);
    fl.rec2 = A.take<float4>(FN);
    fl.ecl = A.take<double2>(FN);
    if (fp64) {
        fl.fdx = A.take<double>(FN); fl.fdy = A.take<double>(FN); fl.fdz = A.take<double>(FN);
        fl.fpx = A.take<double>(FN); fl.fpy = A.take<double>(FN); fl.fpz = A.take<double>(FN);
        fl.fflux = A.take<double>(FN);
    }
    fl.n0 = A.take<int>(std::max(ngrid, 1));
    fl.evald = A.take<unsigned long long>(16 * 64);
    PhaseArrays pa;
    pa.ex = A.take<double>(SM); pa.ey = A.take<double>(SM); pa.ez = A.take<double>(SM);
    pa.phin = A.take<double>(SM); pa.w1 = A.take<double>(SM); pa.w2 = A.take<double>(SM);
    pa.type = A.take<signed char>(SM);
    int *d_lists = nullptr, *d_counts = nullptr;
    if (any_route) {
        d_lists = A.take<int>(SM * 4);
        d_counts = A.take<int>(std::max<size_t>(M, 1) * 4);
    }
    double *d_out = A.take<double>(SM * 2);
    double *d_partial = A.take<double>((size_t)std::max<long long>(part_total, 1));
    TaskDev *d_tasks = A.take<TaskDev>(std::max(ntask, 1));
    double *d_calc = dev_out ? rq.dev_calc : A.take<double>(std::max<size_t>(M, 1) * n);
    double *d_scal = A.take<double>(std::max<size_t>(M, 1) * 8);
    double *d_gsums = A.take<double>(std::max<size_t>(M, 1) * 8 * psum_chunks);
    double *d_scat_part = A.take<double>(scat_part);
    double *d_diag = A.take<double>(std::max<size_t>(diag_idx.size(), 1) * 8);
    cs.ecl = A.take<int>(CN2);
    cs.ein = A.take<double>(CN2);
    cs.eout = A.take<double>(CN2);
    int2 *d_ccand = A.take<int2>(CN2);
    int *d_ccount = A.take<int>(1);
    EclipseWork cew;
    std::memset(&cew, 0, sizeof(cew));
    cew.hit_count = A.take<int>(1);
    cew.face = A.take<int>(CN2);
    cew.grid = A.take<int>(CN2);
    cew.alist[0] = A.take<int>(2 * CN2);
    cew.alist[1] = A.take<int>(2 * CN2);
    cew.phi = A.take<double>(CN2 * 2);
    cew.lim = A.take<double>(2 * CN2);
    cew.acc = A.take<double>(2 * CN2);
    cew.res = A.take<double>(2 * CN2);
    cew.b_in = A.take<double>(2 * CN2);
    cew.b_out = A.take<double>(2 * CN2);
    cew.acount = A.take<int>(ECL_COUNTS);
    CompArrays ca;
    ca.dx = A.take<double>(CN); ca.dy = A.take<double>(CN); ca.dz = A.take<double>(CN);
    ca.flux = A.take<float>(CN);
    ca.nrange = A.take<int>(CN);
    ca.range = A.take<double2>(CN * COMP_MAX_RANGES);
    ca.sums = A.take<double>(std::max<size_t>((size_t)ncomp * 3 * S, 1));
    char *d_down = A.take<char>(down_total);

    // ---------------- uploads through one pinned staging block
    std::vector<int> ndiv(n);
    std::vector<double> col[5];
    for (int k = 0; k < 5; k++) col[k].resize(n);
    for (size_t i = 0; i < n; i++) {
        col[0][i] = rq.data[i].time;
        col[1][i] = rq.data[i].expose;
        col[2][i] = rq.data[i].flux;
        col[3][i] = rq.data[i].ferr;
        col[4][i] = rq.data[i].weight;
        ndiv[i] = std::max(0, rq.data[i].ndiv);
    }
    std::vector<int> st0(M);
    for (size_t i = 0; i < M; i++) st0[i] = ip[i].status;
    struct Up {
        void *dst;
        const void *src;
        size_t bytes;
    };
    std::vector<Up> ups = {
        {d_grids, gdev.data(), gdev.size() * sizeof(GridDev)},
        {d_rings, nullptr, (size_t)nring * sizeof(RingDev)},  // filled in place (fill_rings)
        {d_insts, insts.data(), insts.size() * sizeof(InstDev)},
        {d_status, st0.data(), st0.size() * sizeof(int)},
        {d_time, col[0].data(), n * 8}, {d_expose, col[1].data(), n * 8}, {d_flux, col[2].data(), n * 8},
        {d_ferr, col[3].data(), n * 8}, {d_weight, col[4].data(), n * 8},
        {d_ndiv, ndiv.data(), n * sizeof(int)},
        {d_sub_off, sub_off.data(), (n + 1) * sizeof(int)},
        {d_batches, batches.data(), batches.size() * sizeof(BatchDev)},
        {d_task_first, task_first.data(), task_first.size() * sizeof(int)},
        {d_task_order, task_order.data(), task_order.size() * sizeof(int)},
        {d_diag_idx, diag_idx.data(), diag_idx.size() * sizeof(int)},
        {d_ragged, ragged_order.data(), ragged_order.size() * sizeof(int)},
        {d_dgrids, dgrids.data(), dgrids.size() * sizeof(int)},
        {d_fsegs, fsegs.data(), fsegs.size() * sizeof(FlatSeg)},
        {d_copy_hit, copy_hit.data(), copy_hit.size() * sizeof(FaceCopy)},
        {d_copy_dup, copy_dup.data(), copy_dup.size() * sizeof(FaceCopy)},
        {d_copy_out, copy_out.data(), copy_out.size() * sizeof(FaceCopy)},
        {d_flux_hit, flux_hit.data(), flux_hit.size() * sizeof(FaceCopy)},
        {d_flux_dup, flux_dup.data(), flux_dup.size() * sizeof(FaceCopy)},
        {d_flux_out, flux_out.data(), flux_out.size() * sizeof(FaceCopy)},
        {d_gaps, gaps.data(), gaps.size() * sizeof(int2)},
        {d_comps, comps.data(), comps.size() * sizeof(CompDev)},
        {d_cgrids, cgrids.data(), cgrids.size() * sizeof(GridDev)},
        {cs.px, cpx.data(), cpx.size() * 8}, {cs.py, cpy.data(), cpy.size() * 8}, {cs.pz, cpz.data(), cpz.size() * 8},
        {cs.dx, cdx.data(), cdx.size() * 8}, {cs.dy, cdy.data(), cdy.size() * 8}, {cs.dz, cdz.data(), cdz.size() * 8},
        {cs.flux, cflux.data(), cflux.size() * sizeof(float)},
    };
    size_t up_end = 0;  // the uploaded arrays lead the arena
    for (auto &u : ups)
        if (u.bytes) up_end = std::max(up_end, (size_t)(static_cast<char *>(u.dst) - A.base) + u.bytes);
    ctx->pin.reserve(std::max(up_end, down_total) + 4096);
    hipStream_t s = ctx->stream;
    ctx->prof.host_plan_ms_total +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call0).count();
    if (up_end) {
        // the staging block mirrors the arena's first up_end bytes; it is
        // filled in pieces of up to 1 MiB by the planning threads
        struct Piece {
            char *dst;
            const char *src;
            size_t bytes;
        };
        const size_t PIECE = (size_t)1 << 20;
        std::vector<Piece> pieces;
        for (auto &u : ups)
            if (u.src)
                for (size_t o = 0; o < u.bytes; o += PIECE)
                    pieces.push_back(Piece{ctx->pin.base + (static_cast<char *>(u.dst) - A.base) + o,
                                           static_cast<const char *>(u.src) + o, std::min(PIECE, u.bytes - o)});
        RingDev *pin_rings = reinterpret_cast<RingDev *>(ctx->pin.base + (reinterpret_cast<char *>(d_rings) - A.base));
        const size_t np = pieces.size();
        parallel_for(
            np + (size_t)ngrid,
            [&](size_t k) {
                if (k < np)
                    std::memcpy(pieces[k].dst, pieces[k].src, pieces[k].bytes);
                else
                    fill_rings(k - np, pin_rings);
            },
            4);
        HIPCHK(hipMemcpyAsync(A.base, ctx->pin.base, up_end, hipMemcpyHostToDevice, s));
    }
    launch_fill_gaps(s, d_gaps, (int)gaps.size(), fa);  // one launch for all gaps

    DataDev dd;
    dd.time = d_time; dd.expose = d_expose; dd.flux = d_flux; dd.ferr = d_ferr; dd.weight = d_weight;
    dd.ndiv = d_ndiv; dd.sub_off = d_sub_off; dd.n = (int)n; dd.S = S; dd.middle = middle; dd.range = range;

    // ---------------- launches
    hipEvent_t *ev = ctx->ev;
    // the clears of the kernel sequence (its first kernel, so a captured
    // graph holds them too) and the gather of its results (its last)
    ByteRegions zero, gather;
    std::memset(&zero, 0, sizeof(zero));
    std::memset(&gather, 0, sizeof(gather));
    int nzero = 0, ngather = 0;
    auto add_zero = [&](void *p, size_t b) {
        if (nzero == MAX_REGIONS) throw EngineError(LC_E_DEVICE, "internal: too many clear regions");
        zero.dst[nzero] = p;
        zero.n[nzero++] = b;
    };
    auto add_gather = [&](size_t off, const void *src, size_t b) {
        if (!b) return;
        if (ngather == MAX_REGIONS) throw EngineError(LC_E_DEVICE, "internal: too many gather regions");
        gather.dst[ngather] = d_down + off;
        gather.src[ngather] = src;
        gather.n[ngather++] = b;
    };
    add_zero(d_cand_count, sizeof(int));
    add_zero(ew.hit_count, sizeof(int));
    add_zero(ew.acount, ECL_COUNTS * sizeof(int));
    if (ncomp) {
        add_zero(d_ccount, sizeof(int));
        add_zero(cew.hit_count, sizeof(int));
        add_zero(cew.acount, ECL_COUNTS * sizeof(int));
    }
    add_zero(d_out, SM * 2 * sizeof(double));
    add_zero(fl.evald, 16 * 64 * sizeof(unsigned long long));
    if (any_ecl_stop) add_zero(d_bits, bits_words * sizeof(unsigned));
    if (!dev_out) add_gather(o_calc, d_calc, M * n * 8);
    add_gather(o_scal, d_scal, M * 8 * 8);
    add_gather(o_diag, d_diag, diag_idx.size() * 8 * 8);
    add_gather(o_insts, d_insts, M * sizeof(InstDev));
    add_gather(o_status, d_status, M * sizeof(int));
    add_gather(o_counts, d_cand_count, sizeof(int));
    add_gather(o_counts + 4, ew.hit_count, sizeof(int));
    if (any_ecl_stop) add_gather(o_counts + 8, ew.acount + ECL_ROUNDS + 1, sizeof(int));
    add_gather(o_evald, fl.evald, 16 * 64 * sizeof(unsigned long long));
    if (any_route) add_gather(o_route, d_counts, M * 4 * sizeof(int));
    const bool ecl_rounds = any_ecl_stop && 2LL * nface_total >= ctx->ecl_round_min;

    // ---------------- launches: kernels only, so the sequence can be a graph
    std::vector<int> var_used;
    for (int v = 0; v < 16; v++)
        if (var_range[v].second) var_used.push_back(v);
    auto issue = [&](bool timed) {
        auto mark = [&](hipEvent_t e) {
            if (timed) HIPCHK(hipEventRecord(e, s));
        };
        launch_regions(s, zero, nzero);
        if (any_ecl_stop) launch_phase_bits(s, d_insts, (int)M, dd, d_bits, bits_lb);
        mark(ev[0]);
        launch_faces(s, d_grids, d_rings, nring, nhalf_total, fa, d_cand, d_cand_count, d_status);
        mark(ev[1]);
        launch_eclipse(s, d_grids, ngrid, fa, d_cand, d_cand_count, nface_total, ew, d_status, ecl_rounds,
                       any_ecl_stop, ctx->wide_lanes(nface_total));
        if (ncomp) {  // the components' eclipses by the stars, then their ranges
            launch_comp_pretest(s, d_cgrids, 2 * ncomp, 2 * NC, cs, d_ccand, d_ccount);
            launch_eclipse(s, d_cgrids, 2 * ncomp, cs, d_ccand, d_ccount, 2 * NC, cew, d_status, false, false,
                           ctx->wide_lanes(2LL * NC));
            launch_comp_build(s, d_comps, ncomp, NC, cs, ca, d_status);
        }
        // reused grids: cache hits first (a duplicate may copy from one), then
        // duplicates within the batch, then the new grids into the cache
        if (!copy_hit.empty()) launch_copy_faces(s, d_copy_hit, (int)copy_hit.size(), ctx->geo.pool, fa);
        if (!copy_dup.empty()) launch_copy_faces(s, d_copy_dup, (int)copy_dup.size(), fa, fa);
        if (!copy_out.empty()) launch_copy_faces(s, d_copy_out, (int)copy_out.size(), fa, ctx->geo.pool);
        if (any_disc)  // the opaque disc's ranges of the star faces, copied grids included
            launch_disc_star_ranges(s, d_grids, d_dgrids, (int)dgrids.size(), dmax_face, d_comps, fa, d_status);
        mark(ev[2]);
        launch_continuum(s, d_grids, ngrid, nface_total, fa);
        // fluxes of grids with a known continuum: cached ones first (the
        // pool slots new grids take below may be ones read here), then
        // copies within the call, then the new grids' fluxes into the cache
        if (!flux_hit.empty()) launch_copy_flux(s, d_flux_hit, (int)flux_hit.size(), ctx->geo.pool.flux, fa.flux);
        if (!flux_dup.empty()) launch_copy_flux(s, d_flux_dup, (int)flux_dup.size(), fa.flux, fa.flux);
        if (!flux_out.empty()) launch_copy_flux(s, d_flux_out, (int)flux_out.size(), fa.flux, ctx->geo.pool.flux);
        mark(ev[3]);
        if (any_route) {
            launch_point_sums(s, d_grids, d_insts, (int)M, fa, psum_chunks, d_gsums);
            launch_gint_scales(s, d_insts, (int)M, psum_chunks, d_gsums);
        }
        launch_flat_build(s, d_grids, d_fsegs, (int)fsegs.size(), d_seg_cnt, fa, fl, fp64 ? 1 : 0);
        launch_phases(s, d_insts, (int)M, dd, pa);
        if (ncomp) launch_comp_sums(s, d_comps, ncomp, S, pa, ca);
        if (any_route) launch_route(s, d_insts, (int)M, S, pa, d_lists, d_counts);
        launch_make_tasks(s, d_task_first, d_task_order, (int)task_order.size(), d_batches, ntask, d_tasks);
        mark(ev[4]);
        for (size_t k = 0; k < var_used.size(); k++) {
            const int v = var_used[k];
            if (k == 0) mark(ctx->ev_var[0]);
            launch_flux(s, v, fp64 ? 1 : 0, R, ctx->flux_wave, d_tasks + var_range[v].first, var_range[v].second, d_batches, d_grids,
                        fl, pa, S, d_lists, d_counts, d_partial);
            mark(ctx->ev_var[k + 1]);
        }
        for (int v = 0; v < 8; v++)
            launch_flux_ragged(s, v, d_ragged + ragged_range[v].first, ragged_range[v].second, S, d_batches, d_grids,
                               fa, pa, S, d_lists, d_counts, d_partial);
        mark(ev[5]);
        launch_flux_reduce(s, d_batches, (int)batches.size(), S, TILE, d_lists, d_counts, d_partial, d_out);
        launch_scatter(s, d_insts, (int)M, dd, pa, d_out, ca.sums, d_calc, d_scal, d_scat_part);
        if (!diag_idx.empty()) launch_diag(s, d_grids, d_diag_idx, (int)diag_idx.size(), fa, d_diag);
        launch_regions(s, gather, ngather);
    };
    // A graph replays the exact launches (grids, blocks, arguments) of its
    // capture: the key holds every size, flag and device pointer they use.
    bool graphed = false;
    if (ctx->graphs) {
        std::string key;
        auto put = [&key](const void *p, size_t nb) { key.append(static_cast<const char *>(p), nb); };
        const long long sizes[] = {(long long)M, (long long)n, S, ngrid, nring, nhalf_total, nface_total,
                                   (long long)copy_hit.size(), (long long)copy_dup.size(), (long long)copy_out.size(),
                                   (long long)flux_hit.size(), (long long)flux_dup.size(), (long long)flux_out.size(),
                                   ncomp, NC,
                                   any_route, any_ecl_stop, ecl_rounds, bits_lb, psum_chunks,
                                   (long long)fsegs.size(), (long long)task_order.size(), ntask, R, fp64,
                                   (long long)batches.size(), TILE, (long long)diag_idx.size(),
                                   (long long)ragged_order.size(), (long long)dgrids.size(), dmax_face};
        put(sizes, sizeof(sizes));
        for (int v = 0; v < 16; v++) put(&var_range[v], sizeof(var_range[v]));
        for (int v = 0; v < 8; v++) put(&ragged_range[v], sizeof(ragged_range[v]));
        const void *ptrs[] = {d_grids, d_rings, d_insts, d_status, d_cand, d_cand_count, d_copy_hit, d_copy_dup,
                              d_copy_out, d_flux_hit, d_flux_dup, d_flux_out, d_comps, d_cgrids, d_ccand, d_ccount, d_gsums, d_fsegs, d_seg_cnt, d_lists, d_counts, d_task_first, d_task_order,
                              d_batches, d_tasks, d_partial, d_out, d_calc, d_scal, d_scat_part, d_diag_idx, d_diag,
                              d_bits, d_ragged, d_dgrids, d_down};
        put(ptrs, sizeof(ptrs));
        put(&ctx->geo.pool, sizeof(ctx->geo.pool));
        put(&dd, sizeof(dd));
        put(&fa, sizeof(fa));
        put(&fl, sizeof(fl));
        put(&pa, sizeof(pa));
        put(&ew, sizeof(ew));
        put(&cs, sizeof(cs));
        put(&cew, sizeof(cew));
        put(&ca, sizeof(ca));
        if (ctx->g_exec && key == ctx->g_key) {
            graphed = true;
        } else if (key == ctx->g_seen) {  // second call of this shape: capture it
            ctx->drop_graph();
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                issue(false);
            } catch (...) {  // leave the stream out of capture mode
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(s, &ctx->g_graph));
            HIPCHK(hipGraphInstantiate(&ctx->g_exec, ctx->g_graph, nullptr, nullptr, 0));
            ctx->g_key = key;
            ctx->prof.graph_captures++;
            graphed = true;
        } else {
            ctx->g_seen = key;
        }
    }
    if (graphed) {
        HIPCHK(hipEventRecord(ev[0], s));
        HIPCHK(hipGraphLaunch(ctx->g_exec, s));
        HIPCHK(hipEventRecord(ev[5], s));
    } else {
        issue(true);
    }
    HIPCHK(hipGetLastError());

    // ---------------- the download: one copy of the gathered block
    char *pb = ctx->pin.base;
    HIPCHK(hipMemcpyAsync(pb, d_down, down_total, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ev[6], s));
    HIPCHK(hipStreamSynchronize(s));
    const double *h_calc = reinterpret_cast<const double *>(pb + o_calc);
    const double *h_scal = reinterpret_cast<const double *>(pb + o_scal);
    const double *h_diag = reinterpret_cast<const double *>(pb + o_diag);
    const InstDev *h_insts = reinterpret_cast<const InstDev *>(pb + o_insts);
    int *h_status = reinterpret_cast<int *>(pb + o_status);
    const int *h_counts = reinterpret_cast<const int *>(pb + o_counts);
    const unsigned long long *h_evald = reinterpret_cast<const unsigned long long *>(pb + o_evald);
    const int *cnt = reinterpret_cast<const int *>(pb + o_route);  // routed sub-phase counts
    ctx->prof.eclipse_candidates += h_counts[0];
    ctx->prof.eclipse_hits += h_counts[1];
    ctx->prof.eclipse_bisected += any_ecl_stop ? h_counts[2] : 2.0 * h_counts[1];

    if (any_route) {  // algorithmic work of the fine/coarse batches
        for (const BatchDev &B : batches) {
            if (B.count_idx < 0) continue;
            const GridDev &G = gdev[B.grid];
            const double c = cnt[B.count_idx];
            alg_elems += (double)G.nface * c;
            alg_flops += (double)G.nface * c * elem_flops(G);
            alg_bytes += (double)G.nface * 40. + c * 32.;
            var_flops[flux_variant(G)] += (double)G.nface * c * elem_flops(G);
        }
    }
    if (graphed) {
        ctx->prof.graph_launches++;
        ctx->prof.graph_ms_total += elapsed(ev[0], ev[5]);
    } else {
        ctx->prof.grid_ms_total += elapsed(ev[0], ev[1]);
        ctx->prof.eclipse_ms_total += elapsed(ev[1], ev[2]);
        ctx->prof.cont_ms_total += elapsed(ev[2], ev[3]);
        ctx->prof.other_ms_total += elapsed(ev[3], ev[4]) + elapsed(ev[5], ev[6]);
        ctx->prof.flux_ms_total += elapsed(ev[4], ev[5]);
        ctx->prof.flux_launches += (int64_t)var_used.size();
        for (size_t k = 0; k < var_used.size(); k++) {
            const int v = var_used[k];
            ctx->prof.flux_var_ms[v] += elapsed(ctx->ev_var[k], ctx->ev_var[k + 1]);
            ctx->prof.flux_var_flops[v] += var_flops[v];
            ctx->prof.flux_var_launches[v] += 1;
        }
        for (int k = 0; k < 16 * 64; k++) ctx->prof.flux_var_evaluated[k / 64] += (double)h_evald[k];
        ctx->prof.flux_elements += alg_elems;
        ctx->prof.flux_flops += alg_flops;
        ctx->prof.flux_bytes += alg_bytes;
    }

    // ---------------- stage reuse bookkeeping
    for (int g = 0; g < ngrid; g++) {  // a copy inherits its source's failure
        if (dup_of[g] < 0) continue;
        const int si = gdev[dup_of[g]].inst, di = gdev[g].inst;
        if (h_status[si] && !h_status[di]) h_status[di] = h_status[si];
    }
    for (const auto &ins : inserts) {
        const int inst = gdev[ins.first].inst;
        if (ip[inst].status == LC_OK && !h_status[inst])
            ctx->geo.map[sigs[ins.first]].push_back(GeoCache::Rec{grids[ins.first].rings, ins.second,
                                                                  csig[ins.first]});
    }
    for (int g = 0; g < ngrid; g++) {
        if (computed(g))
            ctx->prof.grids_computed += 1;
        else
            ctx->prof.grids_reused += 1;
        if (gdev[g].cont_copy) ctx->prof.continua_reused += 1;
    }

    // ---------------- finalise per instance (on the planning threads)
    parallel_for(M, [&](size_t i) {
        int st = ip[i].status;
        if (st == LC_OK && h_status[i]) st = h_status[i];
        rq.status[i] = st;
        if (st != LC_OK) {
            if (ip[i].msg.empty()) ip[i].msg = status_msg(st);
            return;
        }
        if (!dev_out) std::memcpy(rq.calc + i * n, h_calc + i * n, n * 8);
        lc_scalars &io = rq.io[i];
        const double *sc = h_scal + i * 8;
        if (o.scale) {
            if (rq.models[i].iscale) {  // star 1, disc, edge, spot (zero columns without them), star 2
                io.sfac[0] = sc[0];
                io.sfac[1] = sc[5];
                io.sfac[2] = sc[6];
                io.sfac[3] = sc[7];
                io.sfac[4] = sc[4];
            } else {
                for (int k = 0; k < 5; k++) io.sfac[k] = sc[0];
            }
            io.chisq = sc[1];
            io.wnok = sc[2];
        } else if (o.rdata) {
            io.chisq = sc[1];
            io.wnok = sc[2];
        }
        const lc_model &m = rq.models[i];
        if (o.diagnostics) {
            const InstDev &I = h_insts[i];
            double pn = 0.5;
            double s1 = 1.;
            if (!(pn <= I.phase1 || pn >= 1. - I.phase1))
                s1 = (I.scale11 * (1. - I.phase1 - pn) + I.scale12 * (pn - I.phase1)) / (1 - 2. * I.phase1);
            const double wd = s1 * h_diag[diag_of[i][2] * 8 + 4];
            io.wdwarf = wd * (o.scale ? sc[0] : io.sfac[0]);
            // comp_gravity1/2 (comp_gravity.cpp) and comp_radius1/2
            double gm = std::pow(1000. * m.velocity_scale.value, 3.0) * m.tperiod * lcx::DAY / TWOPI;
            double a = std::pow(gm / lcx::sqr(TWOPI / lcx::DAY / m.tperiod), 1. / 3.);
            double gscale = 100 * gm / lcx::sqr(a);
            double r1, r2;
            get_r1r2(m, r1, r2);
            double rl2 = 1. - l1(2, m.q.value, m.spin2.value);
            if (m.r2.value < 0) r2 = rl2;
            for (int star = 0; star < 2; star++) {
                double gref;
                const double ACC = m.delta_phase / 10.;
                if (star == 0 ? m.roche1 : m.roche2) {
                    double ffac, rref, pref;
                    if (star == 0) {
                        double rl1 = l1(2, m.q.value, m.spin1.value);  // the reference uses xl12 here
                        ffac = r1 / rl1;
                    } else {
                        ffac = r2 / rl2;
                    }
                    lcx::ref_sphere(m.q.value, star, star == 0 ? m.spin1.value : m.spin2.value, ffac, rref, pref);
                    V3 pv, dv;
                    double rad;
                    int fst = lcx::face_scalar(m.q.value, star, star == 0 ? m.spin1.value : m.spin2.value,
                                               lcx::v3(star == 0 ? -1. : 1., 0., 0.), rref, pref, ACC, pv, dv, rad,
                                               gref);
                    if (fst) gref = 0.;
                    gref *= gscale;
                } else {
                    gref = star == 0 ? gscale / (1 + m.q.value) / lcx::sqr(r1)
                                     : gscale * m.q.value / (1 + m.q.value) / lcx::sqr(r2);
                }
                const double *dg = h_diag + diag_of[i][star] * 8;
                double lg = ((gref > 0) & (dg[0] > 0) & (dg[1] > 0)) ? std::log10(gref * dg[0] / dg[1]) : 0.;
                double rv = std::pow(dg[3] / dg[2], 1. / 3.);
                if (star == 0) {
                    io.logg1 = lg;
                    io.rv1 = rv;
                } else {
                    io.logg2 = lg;
                    io.rv2 = rv;
                }
            }
        } else {
            io.wdwarf = 0.;
            io.logg1 = io.logg2 = io.rv1 = io.rv2 = 0.;
        }
    });
    for (size_t i = 0; i < M; i++) {
        if (rq.status[i] == LC_OK)
            ctx->prof.evaluations++;
        else if (M == 1)
            ctx->err = ip[i].msg;
    }
    ctx->prof.call_ms_total +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call0).count();
}

// light_curve_comp_fast: model at max_model_points evenly indexed points,
// linear interpolation in time, chi**2 over all points
// (light_curve_comp_fast.cpp:6-84).
void run_eval(lc_ctx *ctx, const EvalRequest &rq) {
    const int Mp = rq.opts.max_model_points;
    if (Mp <= 0 || !(rq.n > (size_t)Mp)) {
        run_batch(ctx, rq);
        return;
    }
    if (Mp < 2) throw EngineError(LC_E_ARG, "light_curve_comp_fast: max_model_points must be at least 2");
    const size_t n = rq.n;
    std::vector<lc_datum> sample;
    std::vector<size_t> sidx;
    for (int i = 0; i < Mp; ++i) {
        size_t idx = (size_t)i * (n - 1) / (size_t)(Mp - 1);
        sample.push_back(rq.data[idx]);
        sidx.push_back(idx);
    }
    std::vector<double> mcalc(rq.m * sample.size());
    EvalRequest sub = rq;
    sub.data = sample.data();
    sub.n = sample.size();
    sub.calc = mcalc.data();
    run_batch(ctx, sub);
    for (size_t k = 0; k < rq.m; k++) {
        if (rq.status[k] != LC_OK) continue;
        const double *mc = mcalc.data() + k * sample.size();
        double *c = rq.calc + k * n;
        for (size_t i = 0; i < n; ++i) {
            auto it = std::lower_bound(sidx.begin(), sidx.end(), i);
            if (it == sidx.begin()) {
                c[i] = mc[0];
            } else if (it == sidx.end()) {
                c[i] = mc[sample.size() - 1];
            } else {
                size_t i2x = (size_t)(it - sidx.begin()), i1x = i2x - 1;
                size_t i1 = sidx[i1x], i2 = sidx[i2x];
                double t = (rq.data[i].time - rq.data[i1].time) / (rq.data[i2].time - rq.data[i1].time);
                c[i] = mc[i1x] * (1.0 - t) + mc[i2x] * t;
            }
        }
        if (rq.opts.rdata) {
            double wnok = 0., chisq = 0.;
            for (size_t i = 0; i < n; ++i)
                if (rq.data[i].weight > 0.) {
                    wnok += rq.data[i].weight;
                    chisq += rq.data[i].weight * lcx::sqr((rq.data[i].flux - c[i]) / rq.data[i].ferr);
                }
            rq.io[k].wnok = wnok;
            rq.io[k].chisq = chisq;
        }
    }
}

// Per-device address of the no-allocation entry point's scratch array
// (defined with that entry point below).
void noalloc_scratch_init(int device);

template <class F> lc_status guarded(lc_ctx *ctx, F f) {
    try {
        return (lc_status)f();
    } catch (const EngineError &e) {
        if (ctx) ctx->err = e.what();
        lch::set_thread_error(e.what());
        return (lc_status)e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        lch::set_thread_error(e.what());
        return LC_E_DEVICE;
    }
}

}  // namespace

// =================================================================== ABI
extern "C" {

const char *lc_version(void) { return "lcurve_re_amd 0.1 (gfx950)"; }

lc_status lc_ctx_create(int device, lc_ctx **out) {
    if (!out) return LC_E_ARG;
    *out = nullptr;
    return guarded(nullptr, [&]() -> int {
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (ndev <= 0) throw EngineError(LC_E_DEVICE, "no HIP device available");
        if (device < 0 || device >= ndev) device = 0;
        HIPCHK(hipSetDevice(device));
        noalloc_scratch_init(device);
        std::unique_ptr<lc_ctx> c(new lc_ctx());
        c->device = device;
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
        for (auto &e : c->ev_var) HIPCHK(hipEventCreate(&e));
        std::memset(&c->prof, 0, sizeof(c->prof));
        if (const char *r = std::getenv("LCURVE_AMD_FLUX_R")) c->flux_rphase = std::atoi(r);
        if (c->flux_rphase != 1 && c->flux_rphase != 2 && c->flux_rphase != 4) c->flux_rphase = 0;
        if (const char *ch = std::getenv("LCURVE_AMD_FLUX_CHUNK")) c->flux_chunk = std::max(32, std::atoi(ch));
        if (const char *fw = std::getenv("LCURVE_AMD_FLUX_WAVE")) c->flux_wave = std::atoi(fw) != 0;
        // grids kept across calls, as the reference's per-thread cache
        // (disabled by the same environment variable)
        c->geo.cap = std::getenv("LCURVE_NO_GRID_CACHE") ? 0 : DEFAULT_GEO_CACHE_FACES;
        if (const char *d = std::getenv("LCURVE_AMD_DEDUPE")) c->dedupe = std::atoi(d) != 0;
        if (std::getenv("LCURVE_AMD_FULL_ECLIPSE")) c->ecl_stop = false;
        if (const char *g = std::getenv("LCURVE_AMD_GRAPHS")) c->graphs = std::atoi(g) != 0;
        if (const char *r = std::getenv("LCURVE_AMD_ECL_ROUNDS_MIN")) c->ecl_round_min = std::atoll(r);
        if (const char *w = std::getenv("LCURVE_AMD_ECL_WIDE")) c->ecl_wide = std::atoi(w);
        if (const char *r = std::getenv("LCURVE_AMD_ECL_STOP_MIN_FACES")) c->ecl_stop_min_faces = std::atoll(r);
        *out = c.release();
        return LC_OK;
    });
}

void lc_ctx_destroy(lc_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->ev_var)
        if (e) (void)hipEventDestroy(e);
    ctx->geo.release();
    ctx->drop_graph();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *lc_last_error(const lc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

lc_status lc_ctx_get_profile(const lc_ctx *ctx, lc_profile *out) {
    if (!ctx || !out) return LC_E_ARG;
    *out = ctx->prof;
    return LC_OK;
}

lc_status lc_ctx_reset_profile(lc_ctx *ctx) {
    if (!ctx) return LC_E_ARG;
    std::memset(&ctx->prof, 0, sizeof(ctx->prof));
    return LC_OK;
}

lc_status lc_ctx_set_stage_cache(lc_ctx *ctx, size_t max_faces) {
    if (!ctx) return LC_E_ARG;
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->geo.release();
        ctx->geo.cap = max_faces;
        return LC_OK;
    });
}

lc_status lc_ctx_set_graphs(lc_ctx *ctx, int on) {
    if (!ctx) return LC_E_ARG;
    ctx->graphs = on != 0;
    ctx->g_seen.clear();
    if (!ctx->graphs) {
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        ctx->drop_graph();
    }
    return LC_OK;
}

lc_status lc_ctx_autotune(lc_ctx *ctx, const lc_model *models, size_t m, const lc_datum *data, size_t n,
                          const lc_eval_opts *opts, int32_t *best_rphase, int32_t *best_chunk) {
    if (!ctx) return LC_E_ARG;
    if (!models || !data || !opts || m == 0 || n == 0) {
        ctx->err = "lc_ctx_autotune: empty or null workload";
        return LC_E_ARG;
    }
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        std::vector<double> calc(m * n);
        std::vector<lc_scalars> io(m);
        std::vector<int32_t> status(m);
        for (auto &s : io)
            for (double &f : s.sfac) f = 1.0;
        EvalRequest rq{models, m, data, n, *opts, calc.data(), io.data(), status.data()};
        const lc_profile saved = ctx->prof;
        const int r0 = ctx->flux_rphase, c0 = ctx->flux_chunk;
        // candidates are timed from the stage events of eager calls (a
        // replayed graph is timed only as a whole): the flux sums and the
        // stage after them, whose merge of the chunk partials grows as the
        // chunks shrink
        const bool graphs = ctx->graphs;
        ctx->graphs = false;
        struct Restore {  // the context's own settings if a timing run throws
            lc_ctx *c;
            bool graphs;
            int r, ch;
            lc_profile prof;
            bool armed = true;
            ~Restore() {
                if (!armed) return;
                c->graphs = graphs;
                c->flux_rphase = r;
                c->flux_chunk = ch;
                c->prof = prof;
            }
        } restore{ctx, graphs, r0, c0, saved};
        run_eval(ctx, rq);  // warm the arenas once
        auto time_of = [&](int r, int c) {
            ctx->flux_rphase = r;
            ctx->flux_chunk = c;
            double ms = 0.;
            for (int rep = 0; rep < 2; rep++) {
                const double f0 = ctx->prof.flux_ms_total + ctx->prof.other_ms_total;
                run_eval(ctx, rq);
                ms += ctx->prof.flux_ms_total + ctx->prof.other_ms_total - f0;
            }
            return ms;
        };
        // a candidate replaces the current shape only when it is more than
        // 2% faster, so timing noise does not flip between near-equal shapes;
        // the current shape is reported as what it resolved to here
        double best = time_of(r0, c0) / 1.02;
        int br = ctx->last_rphase, bc = ctx->last_chunk;
        // the partial sums grow as 1/chunk: chunks below 256 faces pay off
        // only for small calls, so a candidate whose partial buffer would
        // pass 4 GiB is not tried (the arena keeps its largest size)
        const double part_bytes_x_chunk = 8.0 * (double)ctx->last_part * (double)bc;
        for (int r : {1, 2, 4})
            for (int c : {64, 128, 256, 512, 1024, 2048, 4096, 8192}) {
                if (r == br && c == bc) continue;
                if (part_bytes_x_chunk / c > 4.0 * 1024 * 1024 * 1024) continue;
                const double ms = time_of(r, c);
                if (ms < best) {
                    best = ms;
                    br = r;
                    bc = c;
                }
            }
        ctx->flux_rphase = br;
        ctx->flux_chunk = bc;
        ctx->prof = saved;
        ctx->graphs = graphs;
        restore.armed = false;
        if (best_rphase) *best_rphase = br;
        if (best_chunk) *best_chunk = bc;
        return LC_OK;
    });
}

lc_status lc_star_grid_size(const lc_model *m, int star, int fine, size_t *nface) {
    if (!m || !nface || (star != 0 && star != 1)) return LC_E_ARG;
    return guarded(nullptr, [&]() -> int {
        GridPlan gp = plan_grid(*m, star, fine != 0);
        *nface = (size_t)gp.dev.nface;
        return LC_OK;
    });
}

lc_status lc_light_curve_comp_batch(lc_ctx *ctx, const lc_model *models, size_t m, const lc_datum *data, size_t n,
                                    const lc_eval_opts *opts, double *calc, lc_scalars *io, int32_t *status) {
    if (!ctx) return LC_E_ARG;
    if (!models || !data || !opts || !calc || !io || !status) {
        ctx->err = "light_curve_comp: null argument (an empty data set has no data pointer)";
        return LC_E_ARG;
    }
    if (m == 0) return LC_OK;
    return guarded(ctx, [&]() -> int {
        if (n == 0) throw EngineError(LC_E_ARG, "light_curve_comp: empty data");
        HIPCHK(hipSetDevice(ctx->device));
        EvalRequest rq{models, m, data, n, *opts, calc, io, status};
        run_eval(ctx, rq);
        return LC_OK;
    });
}

lc_status lc_light_curve_comp_batch_dev(lc_ctx *ctx, const lc_model *models, size_t m, const lc_datum *data,
                                        size_t n, const lc_eval_opts *opts, double *dev_calc, lc_scalars *io,
                                        int32_t *status) {
    if (!ctx) return LC_E_ARG;
    if (!models || !data || !opts || !dev_calc || !io || !status) {
        ctx->err = "light_curve_comp_batch_dev: null argument";
        return LC_E_ARG;
    }
    if (m == 0) return LC_OK;
    return guarded(ctx, [&]() -> int {
        if (n == 0) throw EngineError(LC_E_ARG, "light_curve_comp: empty data");
        if (opts->max_model_points > 0 && n > (size_t)opts->max_model_points)
            throw EngineError(LC_E_ARG, "light_curve_comp_batch_dev: no light_curve_comp_fast sampling "
                                        "(its interpolation runs on the host)");
        HIPCHK(hipSetDevice(ctx->device));
        hipPointerAttribute_t attr;
        const hipError_t qa = hipPointerGetAttributes(&attr, dev_calc);
        if (qa != hipSuccess) (void)hipGetLastError();  // a host pointer: clear the query's error
        if (qa != hipSuccess || attr.type != hipMemoryTypeDevice || attr.device != ctx->device)
            throw EngineError(LC_E_ARG, "light_curve_comp_batch_dev: dev_calc is not device memory of this "
                                        "context's GPU");
        EvalRequest rq{models, m, data, n, *opts, nullptr, io, status};
        rq.dev_calc = dev_calc;
        run_batch(ctx, rq);
        return LC_OK;
    });
}

lc_status lc_light_curve_comp(lc_ctx *ctx, const lc_model *model, const lc_datum *data, size_t n,
                              const lc_eval_opts *opts, double *calc, lc_scalars *io) {
    if (!ctx) return LC_E_ARG;
    if (!model || !data || !opts || !calc || !io) {
        ctx->err = "light_curve_comp: null argument (an empty data set has no data pointer)";
        return LC_E_ARG;
    }
    int32_t st = LC_OK;
    lc_status r = lc_light_curve_comp_batch(ctx, model, 1, data, n, opts, calc, io, &st);
    if (r != LC_OK) return r;
    return (lc_status)st;
}

}  // extern "C"

// ------------------------------------------------------- stage entry points
namespace {

template <class T> T *dev_upload(DevArena &A, const T *src, size_t n, hipStream_t s) {
    T *d = A.take<T>(std::max<size_t>(n, 1));
    if (n) HIPCHK(hipMemcpyAsync(d, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

// flat_sum_star{1,2}_multi on caller-supplied FlatGrid / PhaseBatch arrays.
int flat_sum_stage(lc_ctx *ctx, int star, const lc_flat_grid *g, const lc_phase_batch *pb, const lc_ldc *ldc,
                   double beam, double spin, double VFAC, double XCOFM, int glens1, double rlens1, int precision,
                   double *out) {
    if (!g || !pb || !ldc || !out) throw EngineError(LC_E_ARG, "flat_sum: null argument");
    if (g->n0 > g->n1 || g->n1 > g->n) throw EngineError(LC_E_ARG, "flat_sum: need n0 <= n1 <= n");
    const size_t nf = g->n, S = pb->n, n1 = g->n1, nr = nf - n1;  // faces [n1, n) have several ranges
    size_t nrange = 0;
    if (nr) {
        if (!g->moff || g->moff[0] != 0) throw EngineError(LC_E_ARG, "flat_sum: moff must start at 0");
        for (size_t r = 0; r < nr; r++)
            if (g->moff[r + 1] < g->moff[r]) throw EngineError(LC_E_ARG, "flat_sum: moff must not decrease");
        nrange = (size_t)g->moff[nr];
        if (nrange && (!g->min_ || !g->mout_)) throw EngineError(LC_E_ARG, "flat_sum: null min_/mout_");
    }
    if (S == 0) return LC_OK;
    if (nf == 0) {
        std::fill(out, out + S, 0.0);
        return LC_OK;
    }
    if (nf > (size_t)(1 << 30) || S > (size_t)(1 << 28)) throw EngineError(LC_E_ARG, "flat_sum: arrays too large");
    const bool fp64 = precision == LC_PREC_FP64;
    GridDev G;
    std::memset(&G, 0, sizeof(G));
    G.disc_comp = -1;
    G.star = star;
    G.nface = (int)n1;  // the chunked kernels sum faces [0, n1); the rest after them
    G.beam = beam;
    G.spin_beam = spin;
    G.VFAC = VFAC;
    G.XCOFM = XCOFM;
    G.rlens1 = rlens1;
    G.glens = (star == 1 && glens1) ? 1 : 0;
    G.ldc[0] = ldc->c1; G.ldc[1] = ldc->c2; G.ldc[2] = ldc->c3; G.ldc[3] = ldc->c4;
    G.mucrit = ldc->mucrit;
    G.ltype = ldc->type;
    std::vector<float4> r0(nf), r1(nf), r2(nf);
    std::vector<double2> ec(nf);
    for (size_t i = 0; i < nf; i++) {
        r0[i] = make_float4((float)g->dx[i], (float)g->dy[i], (float)g->dz[i], (float)g->flux[i]);
        double vx = -VFAC * spin * g->py[i];
        double vy = star == 0 ? VFAC * (spin * g->px[i] - XCOFM) : VFAC * (spin * (g->px[i] - 1.0) + 1.0 - XCOFM);
        double vn = g->dx[i] * vx + g->dy[i] * vy;
        r1[i] = make_float4((float)vx, (float)vy, (float)vn, 0.f);
        r2[i] = make_float4((float)g->px[i], (float)g->py[i], (float)g->pz[i], 0.f);
        ec[i] = i >= g->n0 && i < n1 ? make_double2(g->in1[i - g->n0], g->out1[i - g->n0]) : make_double2(0., 0.);
    }
    const int R = ctx->flux_rphase > 0 ? ctx->flux_rphase : (beam != 0.0 ? 2 : 1), TILE = 256 * R;
    const int chunk = fp64 ? 1024 : (ctx->flux_chunk > 0 ? ctx->flux_chunk : 1024);
    BatchDev B;
    B.inst = 0; B.star = star; B.grid = 0; B.list_off = -1; B.count_idx = -1;
    B.chunk = chunk;
    B.nchunk = (int)((nf + chunk - 1) / chunk);
    B.ntile = (int)((S + TILE - 1) / TILE);
    B.part_off = 0;
    const int nchunk_main = (int)((n1 + chunk - 1) / chunk);
    std::vector<TaskDev> tasks;
    for (int t = 0; t < B.ntile; t++)
        for (int c = 0; c < nchunk_main; c++) tasks.push_back(TaskDev{0, t, c, 0});
    const size_t part = (size_t)B.nchunk * B.ntile * TILE;
    std::vector<double2> rng(nrange);
    for (size_t j = 0; j < nrange; j++) rng[j] = make_double2(g->min_[j], g->mout_[j]);
    ctx->dev.reserve(nf * (3 * 16 + 16 + 7 * 8 + 8) + S * 8 * 10 + part * 8 + tasks.size() * 16 + nr * 4 +
                     nrange * 16 + 65536);
    ctx->dev.reset();
    DevArena &A = ctx->dev;
    hipStream_t s = ctx->stream;
    FlatArrays fl;
    std::memset(&fl, 0, sizeof(fl));
    fl.rec0 = dev_upload(A, r0.data(), nf, s);
    fl.rec1 = dev_upload(A, r1.data(), nf, s);
    fl.rec2 = dev_upload(A, r2.data(), nf, s);
    fl.ecl = dev_upload(A, ec.data(), nf, s);
    if (fp64 || nr) {
        fl.fdx = dev_upload(A, g->dx, nf, s); fl.fdy = dev_upload(A, g->dy, nf, s); fl.fdz = dev_upload(A, g->dz, nf, s);
        fl.fpx = dev_upload(A, g->px, nf, s); fl.fpy = dev_upload(A, g->py, nf, s); fl.fpz = dev_upload(A, g->pz, nf, s);
        fl.fflux = dev_upload(A, g->flux, nf, s);
    }
    int n0 = (int)g->n0;
    fl.n0 = dev_upload(A, &n0, 1, s);
    fl.evald = A.take<unsigned long long>(16 * 64);
    HIPCHK(hipMemsetAsync(fl.evald, 0, 16 * 64 * sizeof(unsigned long long), s));
    PhaseArrays pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.ex = dev_upload(A, pb->ex, S, s);
    pa.ey = dev_upload(A, pb->ey, S, s);
    pa.ez = dev_upload(A, pb->ez, S, s);
    pa.phin = dev_upload(A, pb->phin, S, s);
    GridDev *d_grid = dev_upload(A, &G, 1, s);
    BatchDev *d_batch = dev_upload(A, &B, 1, s);
    TaskDev *d_tasks = dev_upload(A, tasks.data(), tasks.size(), s);
    double *d_partial = A.take<double>(part);
    double *d_out = A.take<double>(2 * S);
    const int *d_moff = nr ? dev_upload(A, g->moff, nr + 1, s) : nullptr;
    const double2 *d_rng = nrange ? dev_upload(A, rng.data(), nrange, s) : nullptr;
    HIPCHK(hipMemsetAsync(d_out, 0, 2 * S * sizeof(double), s));
    HIPCHK(hipEventRecord(ctx->ev[4], s));
    launch_flux(s, flux_variant(G), fp64 ? 1 : 0, R, ctx->flux_wave, d_tasks, (int)tasks.size(), d_batch, d_grid, fl, pa, (int)S,
                nullptr, nullptr, d_partial);
    if (nr)
        launch_flux_ragged_tail(s, flux_variant(G), d_grid, d_batch, fl, d_moff, d_rng, pa, (int)S, (int)n1, (int)nf,
                                chunk, (int)(n1 / chunk), B.nchunk, nchunk_main, (size_t)B.ntile * TILE, d_partial);
    HIPCHK(hipEventRecord(ctx->ev[5], s));
    launch_flux_reduce(s, d_batch, 1, (int)S, TILE, nullptr, nullptr, d_partial, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_out + (size_t)star * S, S * sizeof(double), hipMemcpyDeviceToHost, s));
    unsigned long long h_evald[16 * 64];
    HIPCHK(hipMemcpyAsync(h_evald, fl.evald, sizeof(h_evald), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int k = 0; k < 16 * 64; k++) ctx->prof.flux_var_evaluated[k / 64] += (double)h_evald[k];
    const double ms = elapsed(ctx->ev[4], ctx->ev[5]);
    const int v = flux_variant(G);
    ctx->prof.flux_ms_total += ms;
    ctx->prof.flux_launches += 1;
    ctx->prof.flux_var_ms[v] += ms;
    ctx->prof.flux_var_flops[v] += (double)nf * S * elem_flops(G);
    ctx->prof.flux_var_launches[v] += 1;
    ctx->prof.flux_elements += (double)nf * S;
    ctx->prof.flux_flops += (double)nf * S * elem_flops(G);
    ctx->prof.flux_bytes += (double)nf * 40. + (double)S * 32.;
    return LC_OK;
}

// ------------------------------------------------ flux sums without allocation
// Work memory of lc_flat_sum_multi_noalloc: one static device array, which
// every GPU the library runs on holds until the process ends.  Host code
// reaches it only through the per-device address that noalloc_scratch_init
// reads with hipGetSymbolAddress (the array's name is a host-side stand-in,
// not a device address).  Calls on one GPU share the array: calls that
// overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::string g_scratch_err[NOALLOC_MAX_DEVICES];

// Called once per device by lc_ctx_create, with that device current.  A
// failure leaves the address unset: the other entry points work, and
// lc_flat_sum_multi_noalloc reports the error.
void noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_scratch_ptr[device] = p;
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        throw EngineError(LC_E_DEVICE, "flat_sum_noalloc: device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device])
        throw EngineError(LC_E_DEVICE, "flat_sum_noalloc: no scratch address for this device (" +
                                           g_scratch_err[device] + ")");
    return g_scratch_ptr[device];
}

// Flux records of the caller's device-resident FlatGrid, the same values
// flat_sum_stage builds on the host (float casts of the FP64 arrays; the
// beaming velocities in FP64, comp_light_flat.cpp:110-168), the task table
// (tile-major), the evaluated-element counters cleared, and the descriptors.
__global__ void k_pack_flat_dev(lc_flat_grid g, int star, double spin, double VFAC, double XCOFM, int ntask,
                                int nchunk, GridDev G, BatchDev B, float4 *rec0, float4 *rec1, float4 *rec2,
                                double2 *ecl, TaskDev *tasks, unsigned long long *evald, GridDev *d_grid,
                                BatchDev *d_batch, int *d_n0) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.n) {
        const double dx = g.dx[i], dy = g.dy[i], dz = g.dz[i], px = g.px[i], py = g.py[i], pz = g.pz[i];
        rec0[i] = make_float4((float)dx, (float)dy, (float)dz, (float)g.flux[i]);
        const double vx = -VFAC * spin * py;
        const double vy = star == 0 ? VFAC * (spin * px - XCOFM) : VFAC * (spin * (px - 1.0) + 1.0 - XCOFM);
        const double vn = dx * vx + dy * vy;
        rec1[i] = make_float4((float)vx, (float)vy, (float)vn, 0.f);
        rec2[i] = make_float4((float)px, (float)py, (float)pz, 0.f);
        ecl[i] = i >= g.n0 ? make_double2(g.in1[i - g.n0], g.out1[i - g.n0]) : make_double2(0., 0.);
    }
    if (i < (size_t)ntask) tasks[i] = TaskDev{0, (int)(i / nchunk), (int)(i % nchunk), 0};
    if (i < 16 * 64) evald[i] = 0ull;
    if (i == 0) {
        *d_grid = G;
        *d_batch = B;
        *d_n0 = (int)g.n0;
    }
}

__global__ void k_zero_doubles(double *out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.0;
}

// flat_sum_stage on device-resident arrays with the work memory carved from
// the device's scratch array: the same records, tile shape and kernels, so
// the same sums bit for bit.
int flat_sum_noalloc_stage(lc_ctx *ctx, int star, const lc_flat_grid *g, const lc_phase_batch *pb,
                           const lc_ldc *ldc, double beam, double spin, double VFAC, double XCOFM, int glens1,
                           double rlens1, int precision, double *out) {
    if (!g || !pb || !ldc || !out) throw EngineError(LC_E_ARG, "flat_sum_noalloc: null argument");
    if (g->n1 != g->n) throw EngineError(LC_E_UNSUPPORTED, "flat_sum_noalloc: faces with several eclipse ranges");
    if (g->n0 > g->n1) throw EngineError(LC_E_ARG, "flat_sum_noalloc: n0 > n1");
    const size_t nf = g->n, S = pb->n;
    if (S == 0) return LC_OK;
    if (!pb->ex || !pb->ey || !pb->ez || !pb->phin)
        throw EngineError(LC_E_ARG, "flat_sum_noalloc: null sub-phase array");
    if (nf > 0 && (!g->dx || !g->dy || !g->dz || !g->px || !g->py || !g->pz || !g->flux ||
                   (g->n1 > g->n0 && (!g->in1 || !g->out1))))
        throw EngineError(LC_E_ARG, "flat_sum_noalloc: null face array");
    const bool fp64 = precision == LC_PREC_FP64;
    const int R = ctx->flux_rphase > 0 ? ctx->flux_rphase : (beam != 0.0 ? 2 : 1), TILE = 256 * R;
    const int chunk = fp64 ? 1024 : (ctx->flux_chunk > 0 ? ctx->flux_chunk : 1024);
    const size_t nchunk = nf == 0 ? 1 : (nf + chunk - 1) / chunk;
    const size_t ntile = (S + TILE - 1) / TILE;
    const size_t ntask = nchunk * ntile;
    const size_t part = nchunk * ntile * (size_t)TILE;
    // layout of the scratch array, each piece 256-byte aligned
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_grid = 0, o_batch = o_grid + up(sizeof(GridDev)), o_n0 = o_batch + up(sizeof(BatchDev));
    const size_t o_evald = o_n0 + 256, o_tasks = o_evald + up(16 * 64 * sizeof(unsigned long long));
    const size_t o_rec0 = o_tasks + up(ntask * sizeof(TaskDev)), o_rec1 = o_rec0 + up(nf * sizeof(float4));
    const size_t o_rec2 = o_rec1 + up(nf * sizeof(float4)), o_ecl = o_rec2 + up(nf * sizeof(float4));
    const size_t o_part = o_ecl + up(nf * sizeof(double2)), need = o_part + part * sizeof(double);
    if (nf > (size_t)(1 << 30) || S > (size_t)(1 << 28) || need > NOALLOC_SCRATCH_BYTES)
        throw EngineError(LC_E_ARG, "flat_sum_noalloc: " + std::to_string(nf) + " faces x " + std::to_string(S) +
                                        " sub-phases need " + std::to_string(need) +
                                        " bytes of work memory; the library's scratch holds " +
                                        std::to_string(NOALLOC_SCRATCH_BYTES) +
                                        " (at most about 64 MiB = 64 x faces + 8 x faces / 1024 x sub-phases)");
    unsigned char *base = static_cast<unsigned char *>(noalloc_scratch(ctx->device));
    hipStream_t s = ctx->stream;
    if (nf == 0) {
        // no faces: every sum is zero (a fill kernel, no memset node)
        hipLaunchKernelGGL(k_zero_doubles, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, out, S);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        return LC_OK;
    }
    GridDev G;
    std::memset(&G, 0, sizeof(G));
    G.disc_comp = -1;
    G.star = star;
    G.nface = (int)nf;
    G.beam = beam;
    G.spin_beam = spin;
    G.VFAC = VFAC;
    G.XCOFM = XCOFM;
    G.rlens1 = rlens1;
    G.glens = (star == 1 && glens1) ? 1 : 0;
    G.ldc[0] = ldc->c1; G.ldc[1] = ldc->c2; G.ldc[2] = ldc->c3; G.ldc[3] = ldc->c4;
    G.mucrit = ldc->mucrit;
    G.ltype = ldc->type;
    BatchDev B;
    std::memset(&B, 0, sizeof(B));
    B.inst = 0; B.star = 0; B.grid = 0; B.list_off = -1; B.count_idx = -1;   // star 0: out[j] directly
    B.chunk = chunk;
    B.nchunk = (int)nchunk;
    B.ntile = (int)ntile;
    B.part_off = 0;
    GridDev *d_grid = reinterpret_cast<GridDev *>(base + o_grid);
    BatchDev *d_batch = reinterpret_cast<BatchDev *>(base + o_batch);
    int *d_n0 = reinterpret_cast<int *>(base + o_n0);
    FlatArrays fl;
    std::memset(&fl, 0, sizeof(fl));
    fl.rec0 = reinterpret_cast<float4 *>(base + o_rec0);
    fl.rec1 = reinterpret_cast<float4 *>(base + o_rec1);
    fl.rec2 = reinterpret_cast<float4 *>(base + o_rec2);
    fl.ecl = reinterpret_cast<double2 *>(base + o_ecl);
    fl.n0 = d_n0;
    fl.evald = reinterpret_cast<unsigned long long *>(base + o_evald);
    if (fp64) {
        fl.fdx = const_cast<double *>(g->dx); fl.fdy = const_cast<double *>(g->dy);
        fl.fdz = const_cast<double *>(g->dz); fl.fpx = const_cast<double *>(g->px);
        fl.fpy = const_cast<double *>(g->py); fl.fpz = const_cast<double *>(g->pz);
        fl.fflux = const_cast<double *>(g->flux);
    }
    TaskDev *d_tasks = reinterpret_cast<TaskDev *>(base + o_tasks);
    double *d_partial = reinterpret_cast<double *>(base + o_part);
    PhaseArrays pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.ex = const_cast<double *>(pb->ex);
    pa.ey = const_cast<double *>(pb->ey);
    pa.ez = const_cast<double *>(pb->ez);
    pa.phin = const_cast<double *>(pb->phin);
    const size_t nthreads = std::max<size_t>({nf, ntask, (size_t)(16 * 64)});
    hipLaunchKernelGGL(k_pack_flat_dev, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, s, *g, star, spin,
                       VFAC, XCOFM, (int)ntask, (int)nchunk, G, B, fl.rec0, fl.rec1, fl.rec2, fl.ecl, d_tasks,
                       fl.evald, d_grid, d_batch, d_n0);
    launch_flux(s, flux_variant(G), fp64 ? 1 : 0, R, ctx->flux_wave, d_tasks, (int)ntask, d_batch, d_grid, fl, pa, (int)S, nullptr,
                nullptr, d_partial);
    launch_flux_reduce(s, d_batch, 1, (int)S, TILE, nullptr, nullptr, d_partial, out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return LC_OK;
}

// Uploads a set of grids' face tables and runs faces + eclipse search.
int star_grid_stage(lc_ctx *ctx, const lc_model *m, int star, int fine, lc_star_faces *out) {
    if (!m || !out || (star != 0 && star != 1)) throw EngineError(LC_E_ARG, "set_star_grid: bad argument");
    GridPlan gp = plan_grid(*m, star, fine != 0);
    fill_physics(*m, gp.dev);
    GridDev G = gp.dev;
    if ((size_t)G.nface != out->n) throw EngineError(LC_E_ARG, "set_star_grid: output size differs from the grid size");
    const int nf = G.nface;
    std::vector<RingDev> rings;
    int fo = 0, ho = 0;
    for (auto &R : gp.rings) {
        RingDev d;
        d.sint = R.sint; d.cost = R.cost; d.phi1 = R.phi1; d.dphi_span = R.dphi_span; d.tspan_nl = R.tspan_nl;
        d.nphi = R.nphi; d.face_off = fo; d.half_off = ho; d.grid = 0;
        rings.push_back(d);
        fo += R.nphi;
        ho += (R.nphi + 1) / 2;
    }
    G.face_off = 0; G.ring_off = 0; G.nring = (int)rings.size(); G.half_off = 0; G.nhalf = ho; G.inst = 0;
    if (nf == 0) return LC_OK;
    const size_t FN = (size_t)nf;
    ctx->dev.reserve(FN * (6 * 8 + 3 * 4 + 4 + 16 + 8 + 4 + 4 + 6 * 8 + 4 * 8 + 4 * 4) +
                     rings.size() * sizeof(RingDev) + 65536);
    ctx->dev.reset();
    DevArena &A = ctx->dev;
    hipStream_t s = ctx->stream;
    FaceArrays fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.px = A.take<double>(FN); fa.py = A.take<double>(FN); fa.pz = A.take<double>(FN);
    fa.dx = A.take<double>(FN); fa.dy = A.take<double>(FN); fa.dz = A.take<double>(FN);
    fa.area = A.take<float>(FN); fa.grav = A.take<float>(FN); fa.flux = A.take<float>(FN);
    fa.ecl = A.take<int>(FN); fa.ein = A.take<double>(FN); fa.eout = A.take<double>(FN);
    int2 *d_cand = A.take<int2>(FN);
    int *d_count = A.take<int>(1);
    int *d_status = A.take<int>(1);
    EclipseWork ew;
    ew.hit_count = A.take<int>(1);
    ew.face = A.take<int>(FN);
    ew.grid = A.take<int>(FN);
    ew.phi = A.take<double>(FN);
    ew.lim = A.take<double>(2 * FN);
    ew.acc = A.take<double>(FN);
    ew.res = A.take<double>(2 * FN);
    ew.b_in = A.take<double>(2 * FN);
    ew.b_out = A.take<double>(2 * FN);
    ew.alist[0] = A.take<int>(2 * FN);
    ew.alist[1] = A.take<int>(2 * FN);
    ew.acount = A.take<int>(ECL_COUNTS);
    ew.phase_bits = nullptr;  // no data: full searches (G.ecl_stop = 0)
    ew.bits_lb = 10;
    GridDev *d_grid = dev_upload(A, &G, 1, s);
    RingDev *d_rings = dev_upload(A, rings.data(), rings.size(), s);
    HIPCHK(hipMemsetAsync(d_count, 0, sizeof(int), s));
    HIPCHK(hipMemsetAsync(d_status, 0, sizeof(int), s));
    HIPCHK(hipMemsetAsync(ew.hit_count, 0, sizeof(int), s));
    HIPCHK(hipMemsetAsync(ew.acount, 0, ECL_COUNTS * sizeof(int), s));
    if (gp.ring_faces < nf) {
        int gap = nf - gp.ring_faces, g0 = gp.ring_faces;
        for (double *p : {fa.px, fa.py, fa.pz, fa.dx, fa.dy, fa.dz, fa.ein, fa.eout})
            HIPCHK(hipMemsetAsync(p + g0, 0, gap * 8, s));
        HIPCHK(hipMemsetAsync(fa.area + g0, 0, gap * 4, s));
        HIPCHK(hipMemsetAsync(fa.ecl + g0, 0, gap * 4, s));
        HIPCHK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(fa.grav + g0), 0x3f800000u, gap, s));  // 1.0f
    }
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    launch_faces(s, d_grid, d_rings, (int)rings.size(), ho, fa, d_cand, d_count, d_status);
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    // full searches gain nothing from rounds; LCURVE_AMD_ECL_ROUNDS_MIN=0 forces them (tests)
    launch_eclipse(s, d_grid, 1, fa, d_cand, d_count, nf, ew, d_status, ctx->ecl_round_min <= 0, false,
                   ctx->wide_lanes(nf));
    HIPCHK(hipEventRecord(ctx->ev[2], s));
    HIPCHK(hipGetLastError());
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, d_status, sizeof(int), hipMemcpyDeviceToHost, s));
    struct Dn { void *dst; const void *src; size_t b; };
    Dn dns[] = {{out->px, fa.px, FN * 8}, {out->py, fa.py, FN * 8}, {out->pz, fa.pz, FN * 8},
                {out->dx, fa.dx, FN * 8}, {out->dy, fa.dy, FN * 8}, {out->dz, fa.dz, FN * 8},
                {out->area, fa.area, FN * 4}, {out->gravity, fa.grav, FN * 4}, {out->necl, fa.ecl, FN * 4},
                {out->ingress, fa.ein, FN * 8}, {out->egress, fa.eout, FN * 8}};
    for (auto &d : dns)
        if (d.dst) HIPCHK(hipMemcpyAsync(d.dst, d.src, d.b, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (out->flux) std::fill(out->flux, out->flux + FN, 0.f);
    ctx->prof.grid_ms_total += elapsed(ctx->ev[0], ctx->ev[1]);
    ctx->prof.eclipse_ms_total += elapsed(ctx->ev[1], ctx->ev[2]);
    if (st) throw EngineError(st, status_msg(st));
    return LC_OK;
}

int continuum_stage(lc_ctx *ctx, const lc_model *m, lc_star_faces *s1, lc_star_faces *s2) {
    if (!m || !s1 || !s2) throw EngineError(LC_E_ARG, "set_star_continuum: null argument");
    GridDev G[2];
    for (int k = 0; k < 2; k++) {
        std::memset(&G[k], 0, sizeof(GridDev));
        G[k].disc_comp = -1;
        G[k].star = k;
        fill_physics(*m, G[k]);
    }
    G[0].nface = (int)s1->n;
    G[0].face_off = 0;
    G[1].nface = (int)s2->n;
    G[1].face_off = (int)s1->n;
    const size_t FN = s1->n + s2->n;
    if (FN == 0) return LC_OK;
    ctx->dev.reserve(FN * (6 * 8 + 3 * 4) + 65536);
    ctx->dev.reset();
    DevArena &A = ctx->dev;
    hipStream_t s = ctx->stream;
    FaceArrays fa;
    std::memset(&fa, 0, sizeof(fa));
    double **dp[6] = {&fa.px, &fa.py, &fa.pz, &fa.dx, &fa.dy, &fa.dz};
    double *src1[6] = {s1->px, s1->py, s1->pz, s1->dx, s1->dy, s1->dz};
    double *src2[6] = {s2->px, s2->py, s2->pz, s2->dx, s2->dy, s2->dz};
    for (int k = 0; k < 6; k++) {
        *dp[k] = A.take<double>(FN);
        if (s1->n) HIPCHK(hipMemcpyAsync(*dp[k], src1[k], s1->n * 8, hipMemcpyHostToDevice, s));
        if (s2->n) HIPCHK(hipMemcpyAsync(*dp[k] + s1->n, src2[k], s2->n * 8, hipMemcpyHostToDevice, s));
    }
    fa.area = A.take<float>(FN);
    fa.grav = A.take<float>(FN);
    fa.flux = A.take<float>(FN);
    if (s1->n) {
        HIPCHK(hipMemcpyAsync(fa.area, s1->area, s1->n * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(fa.grav, s1->gravity, s1->n * 4, hipMemcpyHostToDevice, s));
    }
    if (s2->n) {
        HIPCHK(hipMemcpyAsync(fa.area + s1->n, s2->area, s2->n * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(fa.grav + s1->n, s2->gravity, s2->n * 4, hipMemcpyHostToDevice, s));
    }
    GridDev *d_grids = dev_upload(A, G, 2, s);
    HIPCHK(hipEventRecord(ctx->ev[2], s));
    // grids with no faces still take part in the binary search by offset
    launch_continuum(s, d_grids, s2->n ? 2 : 1, (int)FN, fa);
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    HIPCHK(hipGetLastError());
    if (s1->n) HIPCHK(hipMemcpyAsync(s1->flux, fa.flux, s1->n * 4, hipMemcpyDeviceToHost, s));
    if (s2->n) HIPCHK(hipMemcpyAsync(s2->flux, fa.flux + s1->n, s2->n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    ctx->prof.cont_ms_total += elapsed(ctx->ev[2], ctx->ev[3]);
    return LC_OK;
}

}  // namespace

extern "C" {

lc_status lc_flat_sum_star1_multi(lc_ctx *ctx, const lc_flat_grid *g, const lc_phase_batch *pb, const lc_ldc *ldc,
                                  double beam, double spin, double VFAC, double XCOFM, int precision, double *out) {
    if (!ctx) return LC_E_ARG;
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        return flat_sum_stage(ctx, 0, g, pb, ldc, beam, spin, VFAC, XCOFM, 0, 0., precision, out);
    });
}

lc_status lc_flat_sum_star2_multi(lc_ctx *ctx, const lc_flat_grid *g, const lc_phase_batch *pb, const lc_ldc *ldc,
                                  double beam, double spin, double VFAC, double XCOFM, int glens1, double rlens1,
                                  int precision, double *out) {
    if (!ctx) return LC_E_ARG;
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        return flat_sum_stage(ctx, 1, g, pb, ldc, beam, spin, VFAC, XCOFM, glens1, rlens1, precision, out);
    });
}

lc_status lc_flat_sum_multi_noalloc(lc_ctx *ctx, int star, const lc_flat_grid *g, const lc_phase_batch *pb,
                                     const lc_ldc *ldc, double beam, double spin, double VFAC, double XCOFM,
                                     int glens1, double rlens1, int precision, double *out) {
    if (!ctx || (star != 0 && star != 1)) return LC_E_ARG;
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        return flat_sum_noalloc_stage(ctx, star, g, pb, ldc, beam, spin, VFAC, XCOFM, glens1, rlens1, precision,
                                      out);
    });
}

lc_status lc_set_star_grid(lc_ctx *ctx, const lc_model *m, int star, int fine, lc_star_faces *out) {
    if (!ctx) return LC_E_ARG;
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        return star_grid_stage(ctx, m, star, fine, out);
    });
}

lc_status lc_set_star_continuum(lc_ctx *ctx, const lc_model *m, lc_star_faces *star1, lc_star_faces *star2) {
    if (!ctx) return LC_E_ARG;
    return guarded(ctx, [&]() -> int {
        HIPCHK(hipSetDevice(ctx->device));
        return continuum_stage(ctx, m, star1, star2);
    });
}

}  // extern "C"
