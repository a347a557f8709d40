/*
 * lc_oracle.cpp -- FP64 CPU/OpenMP restatement of the lcurve_re light-curve
 * path (TEST INFRASTRUCTURE ONLY; see lc_oracle.h).
 *
 * Reference paths are relative to the lcurve_re repository root.  Every
 * expression keeps the reference's operand order; build with
 * -ffp-contract=off (oracle/Makefile) so the rounding matches the reference
 * compiled the same way.  Reduction orders follow the reference:
 *   - flux sums: 1024-face blocks, faces ascending inside a block, blocks
 *     merged in ascending order (comp_light_flat.cpp:241-308);
 *   - Point-path sums: faces ascending (comp_light.cpp:221-315);
 *   - diagnostics: 8192-face blocks (comp_gravity.cpp, comp_radius.cpp).
 * Where the reference is schedule dependent (the first error message of a
 * parallel grid build, set_star_grid.cpp:512-522) the oracle reports the
 * failure of the lowest face index instead.
 */
#include "lc_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------ constants
// src/lcurve_base/constants.h:15-96 (note the non-CODATA H and K)
const double C_SI = 2.99792458e8;
const double H_SI = 6.6262e-34;
const double K_SI = 1.3806e-23;
const double PI = 3.14159265358979323846264;
const double TWOPI = 2. * PI;
const float DAY = 86400.;
const double EFAC = std::sqrt(8. * std::log(2.));

// src/new_subs.h:199-206
inline double deg2rad(double deg) { return M_PI * deg / 180.; }
inline double rad2deg(double rad) { return 180. * rad / M_PI; }
template <class X> inline X sqr(const X &a) { return a * a; }
// src/new_subs.h:96-99
inline double sign_of(double a, double b) { return b >= 0. ? std::fabs(a) : -std::fabs(a); }

enum { PRIMARY = 0, SECONDARY = 1 };

struct V3 {
    double x, y, z;
};
inline V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
// Vec3 arithmetic, src/new_subs.h:261-356
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(const V3 &a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline double sqrlen(const V3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 add(const V3 &a, const V3 &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(const V3 &a, const V3 &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 scl(double c, const V3 &a) { return v3(c * a.x, c * a.y, c * a.z); }

// Errors carry an lc_status code.
struct OracleError : std::runtime_error {
    int code;
    OracleError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

thread_local std::string g_last_error;

// ------------------------------------------------------------ Roche L1
// src/lroche_base/roche.cpp:8-42
double xl11(double q, double spin) {
    const int NMAX = 1000;
    const double EPS = 1.e-12;
    if (q <= 0.) throw OracleError(LC_E_PARAM, "Roche::xl11(double, double): q <= 0.");
    double ssq = spin * spin;
    double mu = q / (1. + q);
    double a1 = -1. + mu;
    double a2 = 2. - 2. * mu, d1 = 1. * a2;
    double a3 = -1. + mu, d2 = 2. * a3;
    double a4 = ssq + 2. * mu, d3 = 3. * a4;
    double a5 = -2. * ssq - mu, d4 = 4. * a5;
    double a6 = ssq, d5 = 5. * a6;
    int n = 0;
    double xold = 0., x = 1. / (1. + q);
    while (n < NMAX && std::fabs(x - xold) > EPS * std::fabs(x)) {
        xold = x;
        double f = x * (x * (x * (x * (x * a6 + a5) + a4) + a3) + a2) + a1;
        double df = x * (x * (x * (x * d5 + d4) + d3) + d2) + d1;
        x -= f / df;
        n++;
    }
    if (n == NMAX) throw OracleError(LC_E_PARAM, "Roche::xl11(double, double): exceeded maximum iterations");
    return x;
}

// src/lroche_base/roche.cpp:45-80
double xl12(double q, double spin) {
    const int NMAX = 1000;
    const double EPS = 1.e-12;
    if (q <= 0.) throw OracleError(LC_E_PARAM, "Roche::xl12(double, double): q <= 0.");
    double ssq = spin * spin;
    double mu = q / (1. + q);
    double a1 = -1. + mu;
    double a2 = 2. - 2. * mu, d1 = 1. * a2;
    double a3 = -ssq + mu, d2 = 2. * a3;
    double a4 = 3. * ssq + 2. * mu - 2., d3 = 3. * a4;
    double a5 = 1. - mu - 3. * ssq, d4 = 4. * a5;
    double a6 = ssq, d5 = 5. * a6;
    int n = 0;
    double xold = 0., x = 1. / (1. + q);
    while (n < NMAX && std::fabs(x - xold) > EPS * std::fabs(x)) {
        xold = x;
        double f = x * (x * (x * (x * (x * a6 + a5) + a4) + a3) + a2) + a1;
        double df = x * (x * (x * (x * d5 + d4) + d3) + d2) + d1;
        x -= f / df;
        n++;
    }
    if (n == NMAX) throw OracleError(LC_E_PARAM, "Roche::xl12(double,double): exceeded maximum iterations");
    return x;
}

// src/lroche_base/xl1.cpp:42-76
double xl1(double q) {
    const int NMAX = 1000;
    const double EPS = 1.e-12;
    if (q <= 0.) throw OracleError(LC_E_PARAM, "Roche::xl1(double): q <= 0.");
    double mu = q / (1. + q);
    double a1 = -1. + mu;
    double a2 = 2. - 2. * mu, d1 = 1. * a2;
    double a3 = -1. + mu, d2 = 2. * a3;
    double a4 = 1. + 2. * mu, d3 = 3. * a4;
    double a5 = -2. - mu, d4 = 4. * a5;
    double a6 = 1., d5 = 5. * a6;
    int n = 0;
    double xold = 0., x = 1. / (1. + q);
    while (n < NMAX && std::fabs(x - xold) > EPS * std::fabs(x)) {
        xold = x;
        double f = x * (x * (x * (x * (x * a6 + a5) + a4) + a3) + a2) + a1;
        double df = x * (x * (x * (x * d5 + d4) + d3) + d2) + d1;
        x -= f / df;
        n++;
    }
    if (n == NMAX) throw OracleError(LC_E_PARAM, "Roche::xl1(double): exceeded maximum iterations");
    return x;
}

// ------------------------------------------------------- Roche potential
// src/lroche_base/rpot.cpp:37-48
double rpot1(double q, double spin, const V3 &p) {
    double mu = q / (1 + q);
    double comp = 1. - mu;
    double x2y2 = sqr(p.x) + sqr(p.y);
    double z2 = sqr(p.z);
    double r1sq = x2y2 + z2;
    double r1 = std::sqrt(r1sq);
    double r2 = std::sqrt(r1sq + 1. - 2. * p.x);
    return (-comp / r1 - mu / r2 - sqr(spin) * x2y2 / 2. + mu * p.x);
}

// src/lroche_base/rpot.cpp:58-86
double rpot2(double q, double spin, const V3 &p) {
    const double mu = q / (1 + q);
    const double comp = 1. - mu;
    const double x2y2 = p.x * p.x + p.y * p.y;
    const double r1sq = x2y2 + p.z * p.z;
    const double r1 = std::sqrt(r1sq);
    const double r2 = std::sqrt(r1sq + 1.0 - 2.0 * p.x);
    const double spin_sq = spin * spin;
    return -comp / r1 - mu / r2 - spin_sq * (0.5 + 0.5 * x2y2 - p.x) - comp * p.x;
}

inline double rpot_star(int star, double q, double spin, const V3 &p) {
    return star == PRIMARY ? rpot1(q, spin, p) : rpot2(q, spin, p);
}

// src/lroche_base/drpot.cpp:40-57
V3 drpot1(double q, double spin, const V3 &p) {
    double r1sq = sqrlen(p);
    double r1 = std::sqrt(r1sq);
    double r2sq = r1sq + 1. - 2. * p.x;
    double r2 = std::sqrt(r2sq);
    double mu = q / (1 + q);
    double mu1 = mu / r2 / r2sq;
    double comp = (1. - mu) / r1 / r1sq;
    double ssq = sqr(spin);
    return v3(comp * p.x + mu1 * (p.x - 1.) - ssq * p.x + mu,
              comp * p.y + mu1 * p.y - ssq * p.y,
              comp * p.z + mu1 * p.z);
}

// src/lroche_base/drpot.cpp:68-85
V3 drpot2(double q, double spin, const V3 &p) {
    double r1sq = sqrlen(p);
    double r1 = std::sqrt(r1sq);
    double r2sq = r1sq + 1. - 2. * p.x;
    double r2 = std::sqrt(r2sq);
    double mu = q / (1 + q);
    double mu1 = mu / r2 / r2sq;
    double comp = (1. - mu) / r1 / r1sq;
    double ssq = sqr(spin);
    return v3(comp * p.x + mu1 * (p.x - 1.) - ssq * (p.x - 1.) + mu - 1.,
              comp * p.y + mu1 * p.y - ssq * p.y,
              comp * p.z + mu1 * p.z);
}

inline V3 drpot_star(int star, double q, double spin, const V3 &p) {
    return star == PRIMARY ? drpot1(q, spin, p) : drpot2(q, spin, p);
}

// src/lroche_base/set_earth.cpp:39-46
inline V3 set_earth(double cosi, double sini, double phase) {
    double rp = TWOPI * phase;
    double cosp = std::cos(rp), sinp = std::sin(rp);
    return v3(sini * cosp, -sini * sinp, cosi);
}

// src/lroche_base/rpot_val.cpp:21-30
inline double rpot_val(double q, int star, double spin, const V3 &earth, const V3 &p, double lam) {
    V3 r = add(p, scl(lam, earth));
    return rpot_star(star, q, spin, r);
}

// src/lroche_base/rpot_grad.cpp:25-40
inline void rpot_grad(double q, int star, double spin, const V3 &earth, const V3 &p, double lam,
                      double &dphi, double &dlam) {
    V3 r = add(p, scl(lam, earth));
    V3 d = drpot_star(star, q, spin, r);
    V3 ed = v3(earth.y, -earth.x, 0.);
    dphi = TWOPI * lam * dot(d, ed);
    dlam = dot(d, earth);
}

// src/lroche_base/rpot_val_grad.cpp:23-44
inline void rpot_val_grad(double q, int star, double spin, const V3 &earth, const V3 &p, double lam,
                          double &pot, double &dphi, double &dlam) {
    V3 r = add(p, scl(lam, earth));
    V3 d = drpot_star(star, q, spin, r);
    pot = rpot_star(star, q, spin, r);
    V3 ed = v3(earth.y, -earth.x, 0.);
    dphi = TWOPI * lam * dot(d, ed);
    dlam = dot(d, earth);
}

// src/lroche_base/ref_sphere.cpp:30-41
void ref_sphere(double q, int star, double spin, double ffac, double &rref, double &pref) {
    if (star == PRIMARY) {
        double tref = xl11(q, spin);
        rref = std::min(1.001 * ffac, 1.0) * tref;
        pref = rpot1(q, spin, v3(ffac * tref, 0., 0.));
    } else {
        double tref = 1. - xl12(q, spin);
        rref = std::min(1.001 * ffac, 1.0) * tref;
        pref = rpot2(q, spin, v3(1. - ffac * tref, 0., 0.));
    }
}

// ----------------------------------------------------------- face solve
// Potential along a ray, the scalar path of face.cpp:34-62 (RayPot): the
// distance to the star's own centre is a square root.
struct RayPotScalar {
    double mu, comp, spin_sq, cx, dx, dy, dz;
    bool primary;
    RayPotScalar(double q, int star, double spin, const V3 &d)
        : mu(q / (1 + q)), comp(1. - mu), spin_sq(spin * spin), cx(star == PRIMARY ? 0. : 1.),
          dx(d.x), dy(d.y), dz(d.z), primary(star == PRIMARY) {}
    double operator()(double rad) const {
        double px = cx + rad * dx;
        double py = rad * dy;
        double pz = rad * dz;
        double x2y2 = px * px + py * py;
        double r1sq = x2y2 + pz * pz;
        double r1 = std::sqrt(r1sq);
        double r2 = std::sqrt(r1sq + 1. - 2. * px);
        if (primary) return -comp / r1 - mu / r2 - spin_sq * x2y2 / 2. + mu * px;
        return -comp / r1 - mu / r2 - spin_sq * (0.5 + 0.5 * x2y2 - px) - comp * px;
    }
};

// src/lroche_base/face.cpp:163-220
void face(double q, int star, double spin, const V3 &dirn, double rref, double pref, double acc,
          V3 &pvec, V3 &dvec, double &r, double &g) {
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    const RayPotScalar rp(q, star, spin, dirn);
    double tref = rp(rref);
    if (tref < pref)
        throw OracleError(LC_E_FACE, "Roche::face error: point at reference radius appears to have a lower "
                                     "potential than the reference");
    double r1 = rref / 2, r2 = rref;
    tref = pref + 1;
    const int MAXSEARCH = 30;
    for (int i = 0; i < MAXSEARCH && tref > pref; i++) {
        r1 = r2 / 2;
        tref = rp(r1);
        if (tref > pref) r2 = r1;
    }
    if (tref > pref)
        throw OracleError(LC_E_FACE, "Roche::face error: could not find a radius with a potential below the "
                                     "reference potential; probably bad inputs.");
    const int MAXCHOP = 100;
    int nchop = 0;
    while (r2 - r1 > acc && nchop < MAXCHOP) {
        r = (r1 + r2) / 2.;
        if (rp(r) < pref)
            r1 = r;
        else
            r2 = r;
        nchop++;
    }
    if (nchop == MAXCHOP)
        throw OracleError(LC_E_FACE, "Roche::face error: reached maximum number of binary chops = 100");
    r = (r1 + r2) / 2.;
    pvec = add(cofm, scl(r, dirn));
    dvec = drpot_star(star, q, spin, pvec);
    g = len(dvec);
    dvec = v3(dvec.x / g, dvec.y / g, dvec.z / g);
}

// src/lroche_base/face.cpp:71-161: the ring solve used by set_star_grid.
// Along a ray of unit direction the distance to the star's own centre is
// taken as r itself (no square root); only the companion distance needs one.
void face_chop_batch(double q, int star, double spin, const double *dx, const double *dy, int n,
                     double rref, double pref, double acc, double *rad, unsigned char *fail) {
    const double mu = q / (1 + q);
    const double comp = 1. - mu;
    const double spin_sq = spin * spin;
    const double cx = star == PRIMARY ? 0. : 1.;
    const bool primary = star == PRIMARY;
    auto pot = [&](int i, double r) -> double {
        double rdx = r * dx[i];
        double px = cx + rdx;
        double py = r * dy[i];
        double x2y2 = px * px + py * py;
        double rr = r * r;
        double down = r;
        double dcomp = std::sqrt(rr + 1. + (primary ? -2. : 2.) * rdx);
        if (primary) return -comp / down - mu / dcomp - spin_sq * x2y2 / 2. + mu * px;
        return -comp / dcomp - mu / down - spin_sq * (0.5 + 0.5 * x2y2 - px) - comp * px;
    };
    // Each lane runs independently here: lock-step masking in the reference
    // only groups lanes, it does not change any lane's sequence.
    for (int i = 0; i < n; i++) {
        fail[i] = pot(i, rref) < pref ? 1 : 0;
        double r1 = rref / 2, r2 = rref, tref = pref + 1.;
        const int MAXSEARCH = 30;
        for (int it = 0; it < MAXSEARCH; it++) {
            if (!fail[i] && tref > pref) {
                r1 = r2 / 2;
                double t = pot(i, r1);
                tref = t;
                if (t > pref) r2 = r1;
            }
        }
        if (!fail[i] && tref > pref) fail[i] = 2;
        const int MAXCHOP = 100;
        for (int nchop = 0; nchop < MAXCHOP; nchop++) {
            if (!fail[i] && r2 - r1 > acc) {
                double r = (r1 + r2) / 2.;
                if (pot(i, r) < pref)
                    r1 = r;
                else
                    r2 = r;
            }
        }
        if (!fail[i] && r2 - r1 > acc) fail[i] = 3;
        rad[i] = (r1 + r2) / 2.;
    }
}

// ------------------------------------------------------- sphere eclipse
// src/lroche_base/sphere_eclipse.cpp:33-75
bool sphere_eclipse(double cosi, double sini, const V3 &r, const V3 &c, double rsphere, double &phi1,
                    double &phi2, double &lam1, double &lam2) {
    const V3 d = sub(r, c);
    const double PDIST = std::sqrt(sqr(d.x) + sqr(d.y));
    const double BQUAD = d.z * cosi - PDIST * sini;
    if (BQUAD >= 0.) return false;
    const double CQUAD = sqrlen(d) - sqr(rsphere);
    double fac = sqr(BQUAD) - CQUAD;
    if (fac <= 0.) return false;
    fac = std::sqrt(fac);
    lam2 = -BQUAD + fac;
    lam1 = std::max(0., CQUAD / lam2);
    if (CQUAD < 0.) {
        phi1 = 0.;
        phi2 = 1.;
    } else {
        double delta = std::acos((cosi * d.z + std::sqrt(CQUAD)) / (sini * PDIST));
        double phi = std::atan2(d.y, -d.x);
        phi1 = (phi - delta) / TWOPI;
        phi1 -= std::floor(phi1);
        phi2 = phi1 + 2. * delta / TWOPI;
    }
    return true;
}

// src/lroche_base/sphere_eclipse.cpp:94-116
bool sphere_eclipse_earth(const V3 &earth, const V3 &r, const V3 &c, double rsphere, double &lam1,
                          double &lam2) {
    const V3 d = sub(r, c);
    const double BQUAD = dot(earth, d);
    if (BQUAD >= 0.) return false;
    const double CQUAD = sqrlen(d) - sqr(rsphere);
    double fac = sqr(BQUAD) - CQUAD;
    if (fac <= 0.) return false;
    fac = std::sqrt(fac);
    lam2 = -BQUAD + fac;
    lam1 = std::max(0., CQUAD / lam2);
    return true;
}

// ----------------------------------------------------------------- dbrent
// src/dbrent.cpp:19-126 (NR dbrent with the early exit on fref)
template <class F, class DF>
double dbrent(double ax, double bx, double cx, F &func, DF &dfunc, double acc, bool stopfast, double fref,
              double &xmin) {
    double a = (ax < cx ? ax : cx);
    double b = (ax > cx ? ax : cx);
    double x, w, v;
    x = w = v = bx;
    double fw, fv, fx;
    fw = fv = fx = func(x);
    if (stopfast && fx < fref) {
        xmin = x;
        return fx;
    }
    double dw, dv, dx;
    dw = dv = dx = dfunc(x);
    const int ITMAX = 100;
    double xm, tol1, tol2, e = 0., d1, d2, u1, u2, olde, d = 0, u, fu, du;
    bool ok1, ok2;
    for (int iter = 0; iter < ITMAX; iter++) {
        xm = 0.5 * (a + b);
        tol1 = acc;
        tol2 = 2. * tol1;
        if (std::fabs(x - xm) <= (tol2 - 0.5 * (b - a))) {
            xmin = x;
            return fx;
        }
        if (std::fabs(e) > tol1) {
            d2 = d1 = 2. * (b - a);
            if (dw != dx) d1 = (w - x) * dx / (dx - dw);
            if (dv != dx) d2 = (v - x) * dx / (dx - dv);
            u1 = x + d1;
            u2 = x + d2;
            ok1 = (a - u1) * (u1 - b) > 0. && dx * d1 <= 0.;
            ok2 = (a - u2) * (u2 - b) > 0. && dx * d2 <= 0.;
            olde = e;
            e = d;
            if (ok1 || ok2) {
                if (ok1 && ok2)
                    d = (std::fabs(d1) < std::fabs(d2) ? d1 : d2);
                else if (ok1)
                    d = d1;
                else
                    d = d2;
                if (std::fabs(d) <= std::fabs(0.5 * olde)) {
                    u = x + d;
                    if (u - a < tol2 || b - u < tol2) d = sign_of(tol1, xm - x);
                } else {
                    d = 0.5 * (e = (dx >= 0. ? a - x : b - x));
                }
            } else {
                d = 0.5 * (e = (dx >= 0. ? a - x : b - x));
            }
        } else {
            d = 0.5 * (e = (dx >= 0. ? a - x : b - x));
        }
        if (std::fabs(d) >= tol1) {
            u = x + d;
            fu = func(u);
            if (stopfast && fu < fref) {
                xmin = u;
                return fu;
            }
        } else {
            u = x + sign_of(tol1, d);
            fu = func(u);
            if (stopfast && fu < fref) {
                xmin = u;
                return fu;
            }
            if (fu > fx) {
                xmin = x;
                return fx;
            }
        }
        du = dfunc(u);
        if (fu <= fx) {
            if (u >= x)
                a = x;
            else
                b = x;
            v = w; fv = fw; dv = dw;
            w = x; fw = fx; dw = dx;
            x = u; fx = fu; dx = du;
        } else {
            if (u < x)
                a = u;
            else
                b = u;
            if (fu <= fw || w == x) {
                v = w; fv = fw; dv = dw;
                w = u; fw = fu; dw = du;
            } else if (fu < fv || v == x || v == w) {
                v = u; fv = fu; dv = du;
            }
        }
    }
    throw OracleError(LC_E_ECLIPSE_SEARCH, "Subs::dbrent: too many iterations");
}

// --------------------------------------------------------- line functors
// roche.h:258-324 (Rpot, Drpot): potential / directional derivative along a
// line in (phase, lambda) space.
struct PhiLamLine {
    double q;
    int star;
    double spin, cosi, sini;
    V3 p;
    double phi, dphi, lam, dlam;
    double operator()(double x) const {
        return rpot_val(q, star, spin, set_earth(cosi, sini, phi + dphi * x), p, lam + dlam * x);
    }
    double deriv(double x) const {
        double dp, dl;
        rpot_grad(q, star, spin, set_earth(cosi, sini, phi + dphi * x), p, lam + dlam * x, dp, dl);
        return (dp * dphi + dl * dlam);
    }
};
struct PhiLamLineD {
    const PhiLamLine *f;
    double operator()(double x) const { return f->deriv(x); }
};

// roche.h:328-365 (Rlpot, Dlrpot): potential along the line of sight.
struct LamLine {
    double q;
    int star;
    double spin;
    V3 earth, p;
    double operator()(double x) const { return rpot_val(q, star, spin, earth, p, x); }
};
struct LamLineD {
    double q;
    int star;
    double spin;
    V3 earth, p;
    double operator()(double x) const {
        double dp, dl;
        rpot_grad(q, star, spin, earth, p, x, dp, dl);
        return dl;
    }
};

// ----------------------------------------------------------------- linmin
// src/lroche_base/pot_min.cpp:127-420
void linmin(double q, int star, double spin, double cosi, double sini, const V3 &p, double &phi, double &lam,
            double dphi, double dlam, double phi1, double phi2, double lam1, double lam2, double pref, double acc,
            double &pmin, bool &jammed) {
    PhiLamLine func = {q, star, spin, cosi, sini, p, phi, dphi, lam, dlam};
    PhiLamLineD dfunc = {&func};

    double xmax = 1.e30;
    double xend;
    int nbound = 0;
    if (dphi != 0.) {
        xend = (phi1 - phi) / dphi;
        if (xend > 0. && xend < xmax) { nbound = 1; xmax = xend; }
        xend = (phi2 - phi) / dphi;
        if (xend > 0. && xend < xmax) { nbound = 2; xmax = xend; }
    }
    if (dlam != 0.) {
        xend = (lam1 - lam) / dlam;
        if (xend > 0. && xend < xmax) { nbound = 3; xmax = xend; }
        xend = (lam2 - lam) / dlam;
        if (xend > 0. && xend < xmax) { nbound = 4; xmax = xend; }
    }

    double xa = 0., fa = func(xa);
    double xb = 1.e-8 * xmax, fb = func(xb);
    int nten = 0;
    const int NTEN = 7;
    while ((fb >= fa || xa == xb) && nten < NTEN) {
        nten++;
        xb *= 10.;
        fb = func(xb);
    }
    if (fb <= pref) {
        phi += dphi * xb;
        lam += dlam * xb;
        pmin = fb;
        return;
    }
    if (fb >= fa) {
        pmin = fa;
        return;
    }

    bool bracketted = false;
    double xc = 0., fc = 0., dc, xbold = xb, fbold = fb;
    const int NTRY = 5;
    xmax -= xb;
    for (int n = 1; n <= NTRY; n++) {
        xc = xbold + xmax * n / NTRY;
        fc = func(xc);
        if (fc <= pref) {
            phi += dphi * xc;
            lam += dlam * xc;
            pmin = fc;
            return;
        }
        if (fc < fb) {
            xb = xc;
            fb = fc;
        } else {
            bracketted = true;
            break;
        }
    }

    jammed = false;
    if (!bracketted) {
        dc = dfunc(xc);
        if (dc > 0.) {
            xb = xbold;
            double db = dfunc(xb);
            if (db < 0.) {
                double xm = (xb + xc) / 2.;
                while ((xc - xb) > 1.e-8 * xc) {
                    xm = (xb + xc) / 2.;
                    if (dfunc(xm) > 0.)
                        xc = xm;
                    else
                        xb = xm;
                }
                double fm = func(xm);
                if (fm <= pref) {
                    phi += dphi * xm;
                    lam += dlam * xm;
                    pmin = fm;
                    return;
                }
                if (fm < fc && fm < fbold) {
                    xb = xm;
                    bracketted = true;
                } else {
                    throw OracleError(LC_E_ECLIPSE_SEARCH, "Roche::linmin: failed to bracket minimum, error 3");
                }
            } else {
                throw OracleError(LC_E_ECLIPSE_SEARCH, "Roche::linmin: failed to bracket minimum, error 1");
            }
        } else {
            // trapped on a boundary: continue along it
            jammed = true;
            phi += dphi * xmax;
            lam += dlam * xmax;
            xmax = 1.;
            rpot_grad(q, star, spin, set_earth(cosi, sini, phi), p, lam, dphi, dlam);
            switch (nbound) {
            case 1:
                phi = phi1;
                dphi = 0.;
                dlam = dlam > 0. ? lam1 - lam : lam2 - lam;
                break;
            case 2:
                phi = phi2;
                dphi = 0.;
                dlam = dlam > 0. ? lam1 - lam : lam2 - lam;
                break;
            case 3:
                lam = lam1;
                dlam = 0.;
                dphi = dphi > 0. ? phi1 - phi : phi2 - phi;
                break;
            case 4:
                lam = lam2;
                dlam = 0.;
                dphi = dphi > 0. ? phi1 - phi : phi2 - phi;
                break;
            default:
                break;
            }
            func.phi = phi;
            func.dphi = dphi;
            func.lam = lam;
            func.dlam = dlam;

            nten = 0;
            xb = 1.e-6;
            fb = func(xb);
            while (fb >= fa && nten < NTEN) {
                xb *= 10.;
                nten++;
                fb = func(xb);
            }
            if (fb <= pref) {
                phi += dphi * xb;
                lam += dlam * xb;
                pmin = fb;
                return;
            }
            if (fb >= fa) {
                pmin = fa;
                return;
            }
            bracketted = false;
            for (int n = 1; n <= NTRY; n++) {
                xc = xmax * n / NTRY;
                fc = func(xc);
                if (fc <= pref) {
                    phi += dphi * xc;
                    lam += dlam * xc;
                    pmin = fc;
                    return;
                }
                if (fc < fb) {
                    xb = xc;
                    fb = fc;
                } else {
                    bracketted = true;
                    break;
                }
            }
            if (!bracketted) {
                if (dfunc(xc) > 0.) {
                    throw OracleError(LC_E_ECLIPSE_SEARCH, "Roche::linmin failed to bracket minimum, error 2");
                } else {
                    phi += dphi * xmax;
                    lam += dlam * xmax;
                    pmin = fc;
                    return;
                }
            }
        }
    }

    double xmin;
    double xacc = acc / std::sqrt(sqr(TWOPI * dphi) + sqr(dlam));
    pmin = dbrent(xa, xb, xc, func, dfunc, xacc, true, pref, xmin);
    phi += dphi * xmin;
    lam += dlam * xmin;
}

// ---------------------------------------------------------------- pot_min
// src/lroche_base/pot_min.cpp:38-100 (conjugate gradients in phase, lambda)
bool pot_min(double q, double cosi, double sini, int star, double spin, const V3 &p, double phi1, double phi2,
             double lam1, double lam2, double rref, double pref, double acc, double &phi, double &lam) {
    double dphi, dlam, gdphi, gdlam, hdphi, hdlam;
    phi = (phi1 + phi2) / 2.;
    lam = (lam1 + lam2) / 2.;
    double rp = TWOPI * phi;
    double cosp = std::cos(rp), sinp = std::sin(rp);
    V3 earth = v3(sini * cosp, -sini * sinp, cosi);
    double pot;
    rpot_val_grad(q, star, spin, earth, p, lam, pot, dphi, dlam);
    if (pot <= pref) return true;
    gdphi = -dphi;
    gdlam = -dlam;
    dphi = hdphi = gdphi;
    dlam = hdlam = gdlam;
    const int ITMAX = 200;
    const double DELPHI = q / (1. + q) * sqr(acc / rref) / 2.;
    double pmin, gam, dgg, gg;
    // linmin leaves `jammed` untouched on its early exits; those exits return
    // pmin equal to the current potential, so pot_min stops either way.
    bool jammed = false;
    for (int its = 0; its < ITMAX; its++) {
        linmin(q, star, spin, cosi, sini, p, phi, lam, dphi, dlam, phi1, phi2, lam1, lam2, pref, acc, pmin, jammed);
        if (pmin <= pref) return true;
        if (jammed || std::fabs(pmin - pot) < DELPHI) return false;
        pot = pmin;
        rp = TWOPI * phi;
        cosp = std::cos(rp);
        sinp = std::sin(rp);
        earth = v3(sini * cosp, -sini * sinp, cosi);
        rpot_grad(q, star, spin, earth, p, lam, dphi, dlam);
        gg = gdphi * gdphi + gdlam * gdlam;
        if (gg == 0.) return false;
        dgg = (dphi + gdphi) * dphi + (dlam + gdlam) * dlam;
        gam = dgg / gg;
        gdphi = -dphi;
        gdlam = -dlam;
        dphi = hdphi = gdphi + gam * hdphi;
        dlam = hdlam = gdlam + gam * hdlam;
    }
    throw OracleError(LC_E_ECLIPSE_SEARCH, "Roche::pot_min: too many iterations.");
}

// ----------------------------------------------------------------- fblink
// src/lroche_base/fblink.cpp:34-104
bool fblink(double q, int star, double spin, double acc, const V3 &earth, const V3 &p, double rref, double pref) {
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    double lam1, lam2;
    if (!sphere_eclipse_earth(earth, p, cofm, rref, lam1, lam2)) return false;
    if (lam1 == 0.) return true;
    LamLine func = {q, star, spin, earth, p};
    int nstep = 1;
    double step = (lam2 - lam1);
    double f1 = 0., f2 = 0., flam = 1., lam = lam1;
    while (step > acc) {
        lam = lam1 + step / 2.;
        for (int n = 0; n < nstep; n++) {
            flam = func(lam);
            if (flam <= pref) return true;
            if (nstep == 1) {
                f1 = func(lam1);
                f2 = func(lam2);
            }
            if (flam < f1 && flam < f2) break;
            lam += step;
        }
        if (flam < f1 && flam < f2) break;
        step /= 2.;
        nstep *= 2;
    }
    if (flam < f1 && flam < f2) {
        LamLineD dfunc = {q, star, spin, earth, p};
        double xmin;
        flam = dbrent(lam1, lam, lam2, func, dfunc, acc, true, pref, xmin);
        return (flam < pref);
    }
    return false;
}

// --------------------------------------------------------- ingress_egress
// src/lroche_base/ingress_egress.cpp:35-97
bool ingress_egress(double q, int star, double spin, double iangle, double delta, const V3 &r, double &ingress,
                    double &egress, double rref, double pref) {
    double ri = deg2rad(iangle);
    double cosi = std::cos(ri), sini = std::sin(ri);
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    double phi1, phi2, lam1, lam2, phi, lam;
    if (!sphere_eclipse(cosi, sini, r, cofm, rref, phi1, phi2, lam1, lam2)) return false;
    const double ACC = 2. * std::sqrt(2. * TWOPI * (lam2 - lam1) * delta);
    if (!pot_min(q, cosi, sini, star, spin, r, phi1, phi2, lam1, lam2, rref, pref, ACC, phi, lam)) return false;
    double pin = phi, pout = phi1, pmid;
    while (std::fabs(pin - pout) > delta) {
        pmid = (pin + pout) / 2.;
        if (fblink(q, star, spin, ACC, set_earth(cosi, sini, pmid), r, rref, pref))
            pin = pmid;
        else
            pout = pmid;
    }
    ingress = (pin + pout) / 2.;
    ingress = ingress - std::floor(ingress);
    pin = phi;
    pout = phi2;
    while (std::fabs(pin - pout) > delta) {
        pmid = (pin + pout) / 2.;
        if (fblink(q, star, spin, ACC, set_earth(cosi, sini, pmid), r, rref, pref))
            pin = pmid;
        else
            pout = pmid;
    }
    egress = (pin + pout) / 2.;
    egress = egress - std::floor(egress);
    if (egress < ingress) egress++;
    return true;
}

// ---------------------------------------------------------------- numface
// src/lcurve_base/numface.cpp:14-53
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

// ------------------------------------------------------------- planck etc
// src/new_subs.cpp:125-136
double planck(double wave, double temp) {
    const double FAC1 = 2.e27 * H_SI * C_SI;
    const double FAC2 = 1.e9 * H_SI * C_SI / K_SI;
    double efac = FAC2 / (wave * temp);
    if (efac > 40.) return FAC1 * std::exp(-efac) / (wave * sqr(wave));
    return FAC1 / (std::exp(efac) - 1.) / (wave * sqr(wave));
}
// src/new_subs.cpp:158-164
double dlpdlt(double wave, double temp) {
    const double FAC2 = 1.e9 * H_SI * C_SI / K_SI;
    double efac = FAC2 / (wave * temp);
    return efac / (1. - std::exp(-efac));
}

// ------------------------------------------------------------------ model
// lcurve.h:399-410
void get_r1r2(const lc_model &m, double &rr1, double &rr2) {
    if (m.use_radii) {
        rr1 = m.r1.value;
        rr2 = m.r2.value;
    } else {
        double sini = std::sin(deg2rad(m.iangle.value));
        double r2pr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * M_PI * m.cphi4.value), 2.0));
        double r2mr1 = std::sqrt(1.0 - std::pow(sini * std::cos(2.0 * M_PI * m.cphi3.value), 2.0));
        rr1 = (r2pr1 - r2mr1) / 2.0;
        rr2 = (r2pr1 + r2mr1) / 2.0;
    }
}

// One face of a star grid (Lcurve::Point, lcurve.h:49-90) with at most one
// eclipse range (stars are only eclipsed by their companion).
struct Face {
    V3 posn, dirn;
    float area, gravity, flux;
    int necl;
    double ein, eout;
};

// Ring decoding of add_faces (set_star_grid.cpp:262-320)
struct Ring {
    double theta, sint, cost, phi1, phi2;
    int nphi, nl, off;
};

// envelope (set_star_grid.cpp:528-538)
void envelope(double rangle, double lambda, double r1, double &ex, double &ey) {
    double sini = std::sin(rangle);
    double cosi = std::cos(rangle);
    double sinl = std::sin(lambda);
    double cosl = std::cos(lambda);
    double norm = std::sqrt(cosi * cosi + sini * sini * cosl * cosl);
    ex = sinl + r1 * cosi * sinl / norm;
    ey = -cosi * cosl - r1 * cosl / norm;
}

// Everything set_star_grid derives from the model before add_faces
// (set_star_grid.cpp:52-207).
struct GridSetup {
    int star;
    bool eclipse, infill;
    int nlat, nlatfill, nlngfill, nface;
    double r1, r2, rref1, pref1, rref2, pref2, ffac1, ffac2, thelo, thehi, gref;
};

GridSetup grid_setup(const lc_model &m, int star, bool fine, bool want_gref) {
    GridSetup s;
    s.star = star;
    double r1, r2;
    get_r1r2(m, r1, r2);
    s.eclipse = star == PRIMARY ? m.eclipse1 : m.eclipse2;
    s.nlat = star == PRIMARY ? (fine ? m.nlat1f : m.nlat1c) : (fine ? m.nlat2f : m.nlat2c);
    s.nlatfill = star == PRIMARY ? 0 : (fine ? m.nlatfill : 0);
    s.nlngfill = star == PRIMARY ? 0 : (fine ? m.nlngfill : 0);
    double latlo = star == PRIMARY ? 0 : m.llo;
    double lathi = star == PRIMARY ? 0 : m.lhi;
    if (s.nlat < 0) throw OracleError(LC_E_PARAM, "set_star_grid: nlat must not be negative");

    double rl1 = xl11(m.q.value, m.spin1.value);
    if (r1 < 0)
        r1 = rl1;
    else if (r1 > rl1)
        throw OracleError(LC_E_ROCHE_OVERFLOW, "set_star_grid: the primary star is larger than its Roche lobe!");
    double rl2 = 1. - xl12(m.q.value, m.spin2.value);
    if (r2 < 0)
        r2 = rl2;
    else if (r2 > rl2)
        throw OracleError(LC_E_ROCHE_OVERFLOW, "set_star_grid: the secondary star is larger than its Roche lobe!");
    if (m.glens1 && star == SECONDARY && m.roche1 && m.eclipse2)
        throw OracleError(LC_E_LENSING, "set_star_grid: cannot have gravitational lensing, eclipse and Roche "
                                        "lobe geometry at the same time");
    if (m.glens1 && star == SECONDARY) {
        double gm = std::pow(1000. * m.velocity_scale.value, 3.0) * m.tperiod * DAY / TWOPI;
        double a = std::pow(gm / sqr(TWOPI / DAY / m.tperiod), 1. / 3.);
        double rlens1 = 4. * gm / (1. + m.q.value) / a / sqr(C_SI);
        r1 -= rlens1 / (r1 / (1. - r2 / 2.));
        if (r1 < 0)
            throw OracleError(LC_E_LENSING, "set_star_grid: gravitational lensing correction more than the current "
                                            "program can cope with.");
    }
    s.ffac1 = r1 / rl1;
    ref_sphere(m.q.value, PRIMARY, m.spin1.value, s.ffac1, s.rref1, s.pref1);
    s.ffac2 = r2 / rl2;
    ref_sphere(m.q.value, SECONDARY, m.spin2.value, s.ffac2, s.rref2, s.pref2);

    s.infill = m.npole && (star == SECONDARY) && (s.nlatfill > 0 || s.nlngfill > 0) && r2 > r1;
    s.thelo = 0;
    s.thehi = 0;
    if (s.infill) {
        double rangle = deg2rad(m.iangle.value);
        double cosi = std::cos(rangle);
        double ratio = (cosi + r1) / r2;
        if (ratio >= 1.) {
            s.thehi = PI / 2. + rangle;
        } else {
            double llo = 0., lhi = PI / 2.;
            double ex = 0., ey = 0.;
            while (lhi > llo + 1.e-7) {
                double lmid = (llo + lhi) / 2.;
                envelope(rangle, lmid, r1, ex, ey);
                if (ex * ex + ey * ey < r2 * r2)
                    llo = lmid;
                else
                    lhi = lmid;
            }
            double sini = std::sin(rangle);
            s.thehi = std::max(PI / 2. - deg2rad(latlo),
                               std::min(PI / 2. + rangle, std::acos(ey * sini / r2) + deg2rad(m.lfudge)));
        }
        ratio = (cosi - r1) / r2;
        if (ratio >= 1.) {
            s.infill = false;
        } else if (ratio <= -1.) {
            s.thelo = 0.;
        } else {
            s.thelo = std::min(PI / 2. - deg2rad(lathi),
                               std::max(0., PI / 2. - std::acos(ratio) + rangle - deg2rad(m.lfudge)));
        }
    }
    s.nface = numface(s.nlat, s.infill, s.thelo, s.thehi, s.nlatfill, s.nlngfill);
    s.r1 = r1;
    s.r2 = r2;
    s.gref = 1.;
    if (want_gref) {
        const double ACC = m.delta_phase / 10.;
        V3 posn, dvec;
        double rad, gref;
        if (star == PRIMARY && m.roche1) {
            face(m.q.value, PRIMARY, m.spin1.value, v3(-1., 0., 0.), s.rref1, s.pref1, ACC, posn, dvec, rad, gref);
            s.gref = gref;
        } else if (star == SECONDARY && m.roche2) {
            face(m.q.value, SECONDARY, m.spin2.value, v3(1., 0., 0.), s.rref2, s.pref2, ACC, posn, dvec, rad, gref);
            s.gref = gref;
        }
    }
    return s;
}

// Lcurve::add_faces (set_star_grid.cpp:226-523) restated for one
// theta segment [tlo, thi).  Faces land at their ring offsets in `star`.
void add_faces(const lc_model &m, const GridSetup &s, std::vector<Face> &star, int &nface, double tlo, double thi,
               double dtheta, int nlatfill, int nlngfill, int &err_face, int &err_code, std::string &err_msg) {
    const int which = s.star;
    const double q = m.q.value, iangle = m.iangle.value;
    const double spin1 = m.spin1.value, spin2 = m.spin2.value;
    const bool npole = m.npole != 0, roche1 = m.roche1 != 0, roche2 = m.roche2 != 0;
    const double delta = m.delta_phase;
    const double r1 = s.r1, r2 = s.r2;
    double ri = deg2rad(iangle);
    double cosi = std::cos(ri), sini = std::sin(ri);
    const V3 cofm1 = v3(0., 0., 0.), cofm2 = v3(1., 0., 0.);
    const double ACC = delta / 10.;

    bool infill = (nlatfill > 0) || (nlngfill > 0);
    int nlat, nlat1 = 0, nlat2 = 0;
    if (infill) {
        nlat1 = int(std::ceil((1 + nlatfill) * (thi - tlo) / dtheta));
        nlat2 = int(std::ceil((thi - tlo) / dtheta));
        nlat = nlat1 + nlat2;
    } else {
        nlat = int(std::ceil((thi - tlo) / dtheta));
    }

    std::vector<Ring> rings(nlat > 0 ? nlat : 0);
    for (int nt = 0; nt < nlat; nt++) {
        Ring &R = rings[nt];
        R.off = nface;
        if (infill) {
            if (nt < nlat1) {
                R.theta = tlo + (thi - tlo) * (nt + 0.5) / nlat1;
                R.sint = std::sin(R.theta);
                R.cost = std::cos(R.theta);
                R.nphi = std::max(8, int(PI * R.sint * (1 + nlngfill) / dtheta));
                R.nl = nlat1;
                if (which == PRIMARY) {
                    R.phi1 = -PI / 2.;
                    R.phi2 = +PI / 2.;
                } else {
                    R.phi1 = +PI / 2.;
                    R.phi2 = +3. * PI / 2.;
                }
            } else {
                R.theta = tlo + (thi - tlo) * (nt - nlat1 + 0.5) / nlat2;
                R.sint = std::sin(R.theta);
                R.cost = std::cos(R.theta);
                R.nphi = std::max(8, int(PI * R.sint / dtheta));
                R.nl = nlat2;
                if (which == PRIMARY) {
                    R.phi1 = +PI / 2.;
                    R.phi2 = +3. * PI / 2.;
                } else {
                    R.phi1 = -PI / 2.;
                    R.phi2 = +PI / 2.;
                }
            }
        } else {
            R.theta = tlo + (thi - tlo) * (nt + 0.5) / nlat;
            R.sint = std::sin(R.theta);
            R.cost = std::cos(R.theta);
            R.nphi = std::max(16, int(2 * PI * R.sint / dtheta));
            R.nl = nlat;
            R.phi1 = 0.;
            R.phi2 = TWOPI;
        }
        nface += R.nphi;
    }
    // For a few nlat (153, 297, 315, 333, 339, ...) the rings hold one or two
    // faces more than numface() allots; the reference writes those past the
    // end of its vector (set_star_grid.cpp:490, 508), so they are never part
    // of the star.  Faces at or beyond star.size() are dropped here.
    const int nstar = (int)star.size();

    const bool roche_solve = (which == PRIMARY && roche1) || (which == SECONDARY && roche2);
    const double rref_self = which == PRIMARY ? s.rref1 : s.rref2;
    const double pref_self = which == PRIMARY ? s.pref1 : s.pref2;
    const double spin_self = which == PRIMARY ? spin1 : spin2;
    const V3 cofm_self = which == PRIMARY ? cofm1 : cofm2;

    // eclipse of a face by the companion (set_star_grid.cpp:372-386)
    auto compute_eclipse = [&](const V3 &pos, Face &f) {
        double ingress = 0., egress = 0., lam1, lam2;
        bool ecl = false;
        if (s.eclipse) {
            if (which == PRIMARY) {
                if (roche2)
                    ecl = ingress_egress(q, SECONDARY, spin2, iangle, delta, pos, ingress, egress, s.rref2, s.pref2);
                else
                    ecl = sphere_eclipse(cosi, sini, pos, cofm2, r2, ingress, egress, lam1, lam2);
            } else {
                if (roche1)
                    ecl = ingress_egress(q, PRIMARY, spin1, iangle, delta, pos, ingress, egress, s.rref1, s.pref1);
                else
                    ecl = sphere_eclipse(cosi, sini, pos, cofm1, r1, ingress, egress, lam1, lam2);
            }
        }
        f.necl = ecl ? 1 : 0;
        f.ein = ecl ? ingress : 0.;
        f.eout = ecl ? egress : 0.;
    };

#pragma omp parallel for schedule(dynamic)
    for (int nt = 0; nt < nlat; nt++) {
        const Ring &R = rings[nt];
        const int nphi = R.nphi;
        const int nhalf = (nphi + 1) / 2;
        std::vector<double> bdx(nhalf), bdy(nhalf), brad(nhalf);
        std::vector<unsigned char> bfail(nhalf);
        for (int np = 0; np < nhalf; np++) {
            double phi = R.phi1 + (R.phi2 - R.phi1) * (np + 0.5) / nphi;
            double sinp = std::sin(phi), cosp = std::cos(phi);
            if (npole) {
                bdx[np] = R.sint * cosp;
                bdy[np] = R.sint * sinp;
            } else {
                bdx[np] = R.cost;
                bdy[np] = R.sint * cosp;
            }
        }
        if (roche_solve)
            face_chop_batch(q, which, spin_self, bdx.data(), bdy.data(), nhalf, rref_self, pref_self, ACC,
                            brad.data(), bfail.data());

        for (int np = 0; np < nhalf; np++) {
            const int idx = R.off + np;
            try {
                double phi = R.phi1 + (R.phi2 - R.phi1) * (np + 0.5) / nphi;
                double sinp = std::sin(phi), cosp = std::cos(phi);
                V3 dirn = npole ? v3(R.sint * cosp, R.sint * sinp, R.cost) : v3(R.cost, R.sint * cosp, R.sint * sinp);
                V3 posn, dvec;
                double rad, gravity;
                if (roche_solve) {
                    if (bfail[np]) {
                        face(q, which, spin_self, dirn, rref_self, pref_self, ACC, posn, dvec, rad, gravity);
                    } else {
                        rad = brad[np];
                        posn = add(cofm_self, scl(rad, dirn));
                        dvec = drpot_star(which, q, spin_self, posn);
                        gravity = len(dvec);
                        dvec = v3(dvec.x / gravity, dvec.y / gravity, dvec.z / gravity);
                    }
                } else {
                    rad = which == PRIMARY ? r1 : r2;
                    posn = add(cofm_self, scl(rad, dirn));
                    dvec = dirn;
                    gravity = 1.;
                }
                double area = ((R.phi2 - R.phi1) / nphi * rad * R.sint) * ((thi - tlo) / R.nl * rad) / dot(dirn, dvec);
                if (idx < nstar) {
                    Face &f = star[idx];
                    f.posn = posn;
                    f.dirn = dvec;
                    f.area = (float)area;
                    f.gravity = (float)(gravity / s.gref);
                    f.flux = 0.f;
                    compute_eclipse(posn, f);
                }
                int np2 = nphi - 1 - np;
                if (np2 != np && R.off + np2 < nstar) {
                    Face &g = star[R.off + np2];
                    V3 posn2 = posn, dvec2 = dvec;
                    if (npole) {
                        posn2.y = -posn.y;
                        dvec2.y = -dvec.y;
                    } else {
                        posn2.z = -posn.z;
                        dvec2.z = -dvec.z;
                    }
                    g.posn = posn2;
                    g.dirn = dvec2;
                    g.area = (float)area;
                    g.gravity = (float)(gravity / s.gref);
                    g.flux = 0.f;
                    compute_eclipse(posn2, g);
                }
            } catch (const OracleError &e) {
#pragma omp critical(orc_grid_err)
                {
                    if (err_code == 0 || idx < err_face) {
                        err_face = idx;
                        err_code = e.code;
                        err_msg = e.what();
                    }
                }
            }
        }
    }
}

// Lcurve::set_star_grid (set_star_grid.cpp:49-224)
void set_star_grid(const lc_model &m, int star, bool fine, std::vector<Face> &out) {
    GridSetup s = grid_setup(m, star, fine, true);
    Face zero;
    std::memset(&zero, 0, sizeof(zero));
    zero.gravity = 1.f; // Point() default (lcurve.h:53)
    out.assign(s.nface, zero);
    int nface = 0;
    double dtheta = PI / s.nlat;
    int err_face = 1 << 30, err_code = 0;
    std::string err_msg;
    if (s.infill) {
        add_faces(m, s, out, nface, 0., s.thelo, dtheta, 0, 0, err_face, err_code, err_msg);
        add_faces(m, s, out, nface, s.thelo, s.thehi, dtheta, s.nlatfill, s.nlngfill, err_face, err_code, err_msg);
        add_faces(m, s, out, nface, s.thehi, PI, dtheta, 0, 0, err_face, err_code, err_msg);
    } else {
        add_faces(m, s, out, nface, 0., PI, dtheta, 0, 0, err_face, err_code, err_msg);
    }
    if (err_code) throw OracleError(err_code, err_msg);
}

// --------------------------------------------------------------- continuum
struct SpotDir {
    bool on;
    V3 dir;
};

SpotDir spot_dir(const lc_pparam &lg, const lc_pparam &lt, const lc_pparam &fw, const lc_pparam &tc) {
    SpotDir sd;
    sd.on = lg.defined && lt.defined && fw.defined && tc.defined;
    sd.dir = v3(0., 0., 0.);
    if (sd.on) {
        double clong = std::cos(deg2rad(lg.value));
        double slong = std::sin(deg2rad(lg.value));
        double clat = std::cos(deg2rad(lt.value));
        double slat = std::sin(deg2rad(lt.value));
        sd.dir = v3(clat * clong, clat * slong, slat);
    }
    return sd;
}

// Lcurve::set_star_continuum (set_star_continuum.cpp:26-297)
void set_star_continuum(const lc_model &m, std::vector<Face> &star1, std::vector<Face> &star2) {
    double r1, r2;
    get_r1r2(m, r1, r2);
    double rl1 = xl1(m.q.value);
    if (r1 < 0) r1 = rl1;
    double rl2 = 1. - rl1;
    if (r2 < 0) r2 = rl2;
    const V3 cofm2 = v3(1., 0., 0.);
    const double GDCBOL1 =
        m.gdark_bolom1 ? m.gravity_dark1.value : m.gravity_dark1.value / dlpdlt(m.wavelength, m.t1.value);
    const SpotDir s11 = spot_dir(m.stsp11_long, m.stsp11_lat, m.stsp11_fwhm, m.stsp11_tcen);
    const SpotDir s12 = spot_dir(m.stsp12_long, m.stsp12_lat, m.stsp12_fwhm, m.stsp12_tcen);
    const SpotDir s13 = spot_dir(m.stsp13_long, m.stsp13_lat, m.stsp13_fwhm, m.stsp13_tcen);
    const bool is_uespot = m.uesp_long1.defined && m.uesp_long2.defined && m.uesp_lathw.defined &&
                           m.uesp_taper.defined && m.uesp_temp.defined;
    double longhw = 0.;
    V3 uespot = v3(0., 0., 0.);
    if (is_uespot) {
        longhw = (m.uesp_long2.value - m.uesp_long1.value) / 2.;
        double lcen = deg2rad((m.uesp_long1.value + m.uesp_long2.value) / 2.);
        uespot = v3(std::cos(lcen), std::sin(lcen), 0.);
    }
    const double T1 = m.t1.value, T2 = m.t2.value;
    const double absorb = m.absorb.value, wave = m.wavelength;

    const int nelem1 = (int)star1.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nelem1; i++) {
        Face &f = star1[i];
        double mu, r, temp, geom;
        V3 vec = sub(cofm2, f.posn);
        r = len(vec);
        mu = dot(f.dirn, vec) / r;
        double t1 = T1;
        const SpotDir *sp[3] = {&s11, &s12, &s13};
        const lc_pparam *fw[3] = {&m.stsp11_fwhm, &m.stsp12_fwhm, &m.stsp13_fwhm};
        const lc_pparam *tc[3] = {&m.stsp11_tcen, &m.stsp12_tcen, &m.stsp13_tcen};
        for (int k = 0; k < 3; k++) {
            if (sp[k]->on) {
                double dist = rad2deg(std::acos(dot(f.posn, sp[k]->dir) / len(f.posn)));
                t1 += (tc[k]->value - T1) * std::exp(-sqr(dist / (fw[k]->value / EFAC)) / 2.);
            }
        }
        if (is_uespot) {
            V3 evec = v3(f.posn.x, f.posn.y, 0);
            double evecl = len(evec);
            if (evecl > 0.) {
                double rd = len(f.posn);
                double clat = dot(f.posn, evec) / evecl / rd;
                clat = std::min(1.0, std::max(-1., clat));
                double dlat = rad2deg(std::acos(clat));
                double clong = dot(uespot, evec) / evecl;
                clong = std::min(1.0, std::max(-1., clong));
                double dlong = rad2deg(std::acos(clong));
                const double lathw = m.uesp_lathw.value, taper = m.uesp_taper.value, ut = m.uesp_temp.value;
                if (dlat <= lathw && dlong <= longhw) {
                    t1 = ut;
                } else if (dlat <= lathw) {
                    t1 += (ut - T1) * std::exp(-(dlong - longhw) / taper);
                } else if (dlong <= longhw) {
                    t1 += (ut - T1) * std::exp(-(dlat - lathw) / taper);
                } else {
                    t1 += (ut - T1) * std::exp(-(dlat - lathw) / taper) * std::exp(-(dlong - longhw) / taper);
                }
            }
        }
        if (mu >= r2) {
            geom = sqr(r2 / r) * mu;
            temp = std::pow(std::pow(t1 * std::pow(double(f.gravity), GDCBOL1), 4.0) + absorb * std::pow(T2, 4.0) * geom,
                            0.25);
        } else if (mu > -r2) {
            double x0 = -mu / r2;
            geom = sqr(r2 / r) * r2 * (std::sqrt(1. - x0 * x0) * (2 + x0 * x0) / 3 - x0 * (PI / 2 - std::asin(x0))) / PI;
            temp = std::pow(std::pow(t1 * std::pow(double(f.gravity), GDCBOL1), 4.0) + absorb * std::pow(T2, 4.0) * geom,
                            0.25);
        } else {
            geom = 0.;
            temp = t1 * std::pow(double(f.gravity), GDCBOL1);
        }
        f.flux = (float)(f.area * planck(wave, temp));
        if (m.mirror) f.flux = (float)(f.flux + f.area * geom * planck(wave, std::fabs(T2)));
    }

    const V3 cofm1 = v3(0., 0., 0.);
    const double GDCBOL2 =
        m.gdark_bolom2 ? m.gravity_dark2.value : m.gravity_dark2.value / dlpdlt(m.wavelength, std::fabs(T2));
    const SpotDir s21 = spot_dir(m.stsp21_long, m.stsp21_lat, m.stsp21_fwhm, m.stsp21_tcen);
    const SpotDir s22 = spot_dir(m.stsp22_long, m.stsp22_lat, m.stsp22_fwhm, m.stsp22_tcen);
    const int nelem2 = (int)star2.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nelem2; i++) {
        Face &f = star2[i];
        double mu, r, temp, geom;
        V3 vec = sub(cofm1, f.posn);
        r = len(vec);
        mu = dot(f.dirn, vec) / r;
        double t2 = std::fabs(T2);
        const SpotDir *sp[2] = {&s21, &s22};
        const lc_pparam *fw[2] = {&m.stsp21_fwhm, &m.stsp22_fwhm};
        const lc_pparam *tc[2] = {&m.stsp21_tcen, &m.stsp22_tcen};
        for (int k = 0; k < 2; k++) {
            if (sp[k]->on) {
                V3 off = sub(f.posn, cofm2);
                double dist = rad2deg(std::acos(dot(off, sp[k]->dir) / len(off)));
                t2 += (tc[k]->value - t2) * std::exp(-sqr(dist / (fw[k]->value / EFAC)) / 2.);
            }
        }
        if (mu >= r1) {
            geom = sqr(r1 / r) * mu;
            temp = std::pow(std::pow(t2 * std::pow(double(f.gravity), GDCBOL2), 4.0) + absorb * std::pow(T1, 4.) * geom,
                            0.25);
        } else if (mu > -r1) {
            double x0 = -mu / r1;
            geom = sqr(r1 / r) * r1 * (std::sqrt(1. - x0 * x0) * (2 + x0 * x0) / 3 - x0 * (PI / 2 - std::asin(x0))) / PI;
            temp = std::pow(std::pow(t2 * std::pow(double(f.gravity), GDCBOL2), 4.0) + absorb * std::pow(T1, 4.) * geom,
                            0.25);
        } else {
            geom = 0.;
            temp = t2 * std::pow(double(f.gravity), GDCBOL2);
        }
        f.flux = (float)(f.area * planck(wave, temp));
        if (m.mirror) f.flux = (float)(f.flux + f.area * geom * planck(wave, T1));
    }
}

// --------------------------------------------------------------- flat grid
// Lcurve::FlatGrid (lcurve.h:260-272) and its build (comp_light_flat.cpp:26-65)
struct Flat {
    std::vector<double> dx, dy, dz, px, py, pz, flux, in1, out1;
    size_t n = 0, n0 = 0, n1 = 0;
};

void flat_build(const std::vector<Face> &pts, Flat &g) {
    g.n = pts.size();
    size_t c0 = 0, c1 = 0;
    for (const Face &p : pts) {
        c0 += (p.necl == 0);
        c1 += (p.necl == 1);
    }
    g.n0 = c0;
    g.n1 = c0 + c1;
    g.dx.resize(g.n); g.dy.resize(g.n); g.dz.resize(g.n);
    g.px.resize(g.n); g.py.resize(g.n); g.pz.resize(g.n);
    g.flux.resize(g.n);
    g.in1.resize(g.n1 - g.n0);
    g.out1.resize(g.n1 - g.n0);
    size_t i0 = 0, i1 = g.n0;
    for (const Face &p : pts) {
        size_t i = p.necl == 0 ? i0++ : i1++;
        g.dx[i] = p.dirn.x; g.dy[i] = p.dirn.y; g.dz[i] = p.dirn.z;
        g.px[i] = p.posn.x; g.py[i] = p.posn.y; g.pz[i] = p.posn.z;
        g.flux[i] = p.flux;
        if (p.necl == 1) {
            g.in1[i - g.n0] = p.ein;
            g.out1[i - g.n0] = p.eout;
        }
    }
}

lc_flat_grid flat_view(const Flat &g) {
    lc_flat_grid v;
    std::memset(&v, 0, sizeof(v));
    v.n = g.n; v.n0 = g.n0; v.n1 = g.n1;
    v.dx = g.dx.data(); v.dy = g.dy.data(); v.dz = g.dz.data();
    v.px = g.px.data(); v.py = g.py.data(); v.pz = g.pz.data();
    v.flux = g.flux.data(); v.in1 = g.in1.data(); v.out1 = g.out1.data();
    return v;
}

// ---------------------------------------------------------- flux elements
// comp_light_flat.cpp:76-95
struct LdcC {
    double c1, c2, c3, c4, base, mucrit;
    int type;
};
LdcC ldc_unpack(const lc_ldc &l) {
    LdcC L = {l.c1, l.c2, l.c3, l.c4, 1.0 - (l.c1 + l.c2 + l.c3 + l.c4), l.mucrit, l.type};
    return L;
}
inline double imu_bl(double mu, const LdcC &L) {
    double m = std::min(mu, 1.0);
    double im;
    if (L.type == LC_LDC_POLY) {
        double om = 1.0 - m;
        im = 1.0 - om * (L.c1 + om * (L.c2 + om * (L.c3 + om * L.c4)));
    } else {
        double msq = std::sqrt(std::max(m, 0.0));
        im = L.base + msq * (L.c1 + msq * (L.c2 + msq * (L.c3 + msq * L.c4)));
    }
    return static_cast<double>(mu > 0.0) * im;
}

struct ElemArgs {
    int star;
    bool beam_on, glens;
    double beam, spin, VFAC, XCOFM, rlens1;
    LdcC L;
};

// s1_elem / s2_elem (comp_light_flat.cpp:110-168)
inline double elem(const lc_flat_grid &g, size_t i, double ex, double ey, double ez, const ElemArgs &a) {
    double mu = ex * g.dx[i] + ey * g.dy[i] + ez * g.dz[i];
    double magn = 1.0;
    if (a.star == SECONDARY && a.glens) {
        double sx = g.px[i], sy = g.py[i], sz = g.pz[i];
        double d = -(sx * ex + sy * ey + sz * ez);
        double qx = sx + d * ex, qy = sy + d * ey, qz = sz + d * ez;
        double p = std::sqrt(qx * qx + qy * qy + qz * qz);
        double ph = 0.5 * p;
        double rd = a.rlens1 * d;
        double pd = (ph * ph > 25.0 * rd) ? p + rd / p : ph + std::sqrt(ph * ph + rd);
        double mg = pd * pd / (pd - ph) / ph / 4.0;
        magn = d > 0.0 ? mg : 1.0;
    }
    double val;
    if (!a.beam_on) {
        if (a.star == PRIMARY)
            val = mu * g.flux[i] * imu_bl(mu, a.L);
        else
            val = mu * magn * g.flux[i] * imu_bl(mu, a.L);
    } else {
        double vx = -a.VFAC * a.spin * g.py[i];
        double vy = a.star == PRIMARY ? a.VFAC * (a.spin * g.px[i] - a.XCOFM)
                                      : a.VFAC * (a.spin * (g.px[i] - 1.0) + 1.0 - a.XCOFM);
        double vr = -(ex * vx + ey * vy);
        double vn = g.dx[i] * vx + g.dy[i] * vy;
        double mud = mu - mu * vr - vn;
        if (a.star == PRIMARY)
            val = mu * g.flux[i] * (1.0 - a.beam * vr) * imu_bl(mud, a.L);
        else
            val = mu * magn * g.flux[i] * (1.0 - a.beam * vr) * imu_bl(mud, a.L);
    }
    if (a.star == SECONDARY && a.glens) return mu > a.L.mucrit ? val : 0.0;
    return static_cast<double>(mu > a.L.mucrit) * val;
}

// comp_light_flat.cpp:171-173
inline bool ecl_range(double phin, double in, double out) { return (phin >= in && phin <= out) || phin <= out - 1.0; }

// sum_star{1,2}_multi_t (comp_light_flat.cpp:241-372)
void flat_sum_multi(const lc_flat_grid &g, const lc_phase_batch &pb, const ElemArgs &a, double *out) {
    const size_t FACE_BLOCK = 1024;
    const size_t n0 = g.n0, n1 = g.n1, n = g.n;
    const long nent = (long)pb.n;
    const long nblk = (long)((n + FACE_BLOCK - 1) / FACE_BLOCK);
    std::fill(out, out + nent, 0.0);
    std::vector<double> blk((size_t)nblk * (size_t)nent);
#pragma omp parallel for schedule(dynamic)
    for (long b = 0; b < nblk; b++) {
        double *brow = blk.data() + (size_t)b * (size_t)nent;
        const size_t lo = b * FACE_BLOCK;
        const size_t hi = std::min(n, lo + FACE_BLOCK);
        const size_t a1 = std::min(hi, n0);
        const size_t b0 = std::max(lo, n0);
        const size_t b1 = std::min(hi, n1);
        const size_t c0 = std::max(lo, n1);
        std::fill(brow, brow + nent, 0.0);
        for (size_t i = lo; i < a1; i++)
            for (long k = 0; k < nent; k++) brow[k] += elem(g, i, pb.ex[k], pb.ey[k], pb.ez[k], a);
        for (size_t i = b0; i < b1; i++) {
            const double ein = g.in1[i - n0], eout = g.out1[i - n0];
            for (long k = 0; k < nent; k++) {
                double v = elem(g, i, pb.ex[k], pb.ey[k], pb.ez[k], a);
                brow[k] += ecl_range(pb.phin[k], ein, eout) ? 0.0 : v;
            }
        }
        for (size_t i = c0; i < hi; i++) {
            for (long k = 0; k < nent; k++) {
                bool ecl = false;
                for (int j = g.moff[i - n1]; j < g.moff[i - n1 + 1]; j++)
                    if (ecl_range(pb.phin[k], g.min_[j], g.mout_[j])) {
                        ecl = true;
                        break;
                    }
                if (!ecl) brow[k] += elem(g, i, pb.ex[k], pb.ey[k], pb.ez[k], a);
            }
        }
    }
    for (long b = 0; b < nblk; b++) {
        const double *brow = blk.data() + (size_t)b * (size_t)nent;
        for (long k = 0; k < nent; k++) out[k] += brow[k];
    }
}

// single-phase flat sum (comp_light_flat.cpp:175-232), faces ascending
double flat_sum_one(const lc_flat_grid &g, double ex, double ey, double ez, double phin, const ElemArgs &a) {
    double s = 0.0;
    for (size_t i = 0; i < g.n0; i++) s += elem(g, i, ex, ey, ez, a);
    for (size_t i = g.n0; i < g.n1; i++) {
        double v = elem(g, i, ex, ey, ez, a);
        s += ecl_range(phin, g.in1[i - g.n0], g.out1[i - g.n0]) ? 0.0 : v;
    }
    return s;
}

// ------------------------------------------------------------- Point path
// Lcurve::LDC::imu (lcurve.h:216-231)
inline double ldc_imu(double mu, const lc_ldc &l) {
    if (mu <= 0) return 0.;
    mu = std::min(mu, 1.);
    double ommu = 1. - mu, im = 1.;
    if (l.type == LC_LDC_POLY) {
        im -= ommu * (l.c1 + ommu * (l.c2 + ommu * (l.c3 + ommu * l.c4)));
    } else {
        im -= l.c1 + l.c2 + l.c3 + l.c4;
        double msq = std::sqrt(mu);
        im += msq * (l.c1 + msq * (l.c2 + msq * (l.c3 + msq * l.c4)));
    }
    return im;
}

// Point::visible (lcurve.h:81-89)
inline bool visible(const Face &f, double phase) {
    double phi = phase - std::floor(phase);
    if (f.necl && ((phi >= f.ein && phi <= f.eout) || phi <= f.eout - 1.0)) return false;
    return true;
}

// star1_elem / star2_elem (comp_light.cpp:21-92)
double point_elem(const Face &pt, const V3 &earth, double phi, const lc_ldc &ldc, int star, double beam,
                  double spin, double VFAC, double XCOFM, bool glens1, double rlens1) {
    if (!visible(pt, phi)) return 0.0;
    double mu = dot(earth, pt.dirn);
    if (!(mu > ldc.mucrit)) return 0.0;
    double magn = 1.0;
    if (star == SECONDARY && glens1) {
        V3 s = pt.posn;
        double d = -dot(s, earth);
        if (d > 0.0) {
            double p = len(add(s, scl(d, earth)));
            double ph = 0.5 * p;
            double rd = rlens1 * d;
            double pd = (ph * ph > 25.0 * rd) ? p + rd / p : ph + std::sqrt(ph * ph + rd);
            magn = pd * pd / (pd - ph) / ph / 4.0;
        }
    }
    if (beam == 0.0) {
        if (star == PRIMARY) return mu * pt.flux * ldc_imu(mu, ldc);
        return mu * magn * pt.flux * ldc_imu(mu, ldc);
    }
    double vx = -VFAC * spin * pt.posn.y;
    double vy = star == PRIMARY ? VFAC * (spin * pt.posn.x - XCOFM) : VFAC * (spin * (pt.posn.x - 1.0) + 1.0 - XCOFM);
    double vr = -(earth.x * vx + earth.y * vy);
    double vn = pt.dirn.x * vx + pt.dirn.y * vy;
    double mud = mu - mu * vr - vn;
    if (star == PRIMARY) return mu * pt.flux * (1.0 - beam * vr) * ldc_imu(mud, ldc);
    return mu * magn * pt.flux * (1.0 - beam * vr) * ldc_imu(mud, ldc);
}

// Ginterp (lcurve.h:92-148)
struct Ginterp {
    double phase1, phase2, scale11, scale12, scale21, scale22;
    double scale1(double phase) const {
        double pnorm = phase - std::floor(phase);
        if (pnorm <= phase1 || pnorm >= 1. - phase1) return 1.;
        return (scale11 * (1. - phase1 - pnorm) + scale12 * (pnorm - phase1)) / (1 - 2. * phase1);
    }
    double scale2(double phase) const {
        double pnorm = phase - std::floor(phase);
        if (pnorm >= phase2 && pnorm <= 1. - phase2) return 1.;
        if (pnorm < 0.5) return (scale22 * (phase2 - pnorm) + scale21 * (pnorm + phase2)) / (2. * phase2);
        return (scale21 * (1. + phase2 - pnorm) + scale22 * (pnorm - 1. + phase2)) / (2. * phase2);
    }
    int type(double phase) const {
        double pnorm = phase - std::floor(phase);
        if (pnorm <= phase1 || pnorm >= 1. - phase1) return 1;
        if ((pnorm > phase1 && pnorm < phase2) || (pnorm > 1. - phase2 && pnorm < 1. - phase1)) return 2;
        return 3;
    }
};

// comp_star1 / comp_star2 with ndiv = 1, expose = 0 (comp_light.cpp:221-315):
// the Point-path sum at one phase, faces ascending, scaled by gint.
double comp_star(int star, double iangle, const lc_ldc &ldc, double phase, double q, double beam, float vscale,
                 bool glens1, double rlens1, const Ginterp &gint, const std::vector<Face> &sf,
                 const std::vector<Face> &sc) {
    const double XCOFM = q / (1.0 + q);
    const double VFAC = vscale / (C_SI / 1.e3);
    double phi = phase, wgt = 1.0;
    double ri = deg2rad(iangle);
    double cosi = std::cos(ri), sini = std::sin(ri);
    V3 earth = set_earth(cosi, sini, phi);
    double s = 0.0, part;
    if (star == PRIMARY) {
        const std::vector<Face> &st = (gint.type(phi) == 1) ? sf : sc;
        for (const Face &pt : st) s += point_elem(pt, earth, phi, ldc, PRIMARY, beam, 1.0, VFAC, XCOFM, false, 0.);
        part = wgt * gint.scale1(phi) * s;
    } else {
        const std::vector<Face> &st = (gint.type(phi) == 3) ? sf : sc;
        for (const Face &pt : st)
            s += point_elem(pt, earth, phi, ldc, SECONDARY, beam, 1.0, VFAC, XCOFM, glens1, rlens1);
        part = wgt * gint.scale2(phi) * s;
    }
    double sum = 0.0;
    sum += part;
    return sum / std::max(1, 0);
}

// ------------------------------------------------------------ diagnostics
// comp_gravity.cpp: flux-weighted gravity (8192-face blocks)
double comp_gravity(const lc_model &m, const std::vector<Face> &st, int star) {
    double gm = std::pow(1000. * m.velocity_scale.value, 3.0) * m.tperiod * DAY / TWOPI;
    double a = std::pow(gm / sqr(TWOPI / DAY / m.tperiod), 1. / 3.);
    double gscale = 100 * gm / sqr(a), gref;
    double r1, r2;
    get_r1r2(m, r1, r2);
    double rl2 = 1. - xl12(m.q.value, m.spin2.value);
    if (m.r2.value < 0) r2 = rl2;
    if (star == PRIMARY) {
        double rl1 = xl12(m.q.value, m.spin1.value); // the reference uses xl12 here
        if (m.roche1) {
            const double ACC = m.delta_phase / 10.;
            double rref1, pref1, rad, ffac1 = r1 / rl1;
            ref_sphere(m.q.value, PRIMARY, m.spin1.value, ffac1, rref1, pref1);
            V3 posn, dvec;
            face(m.q.value, PRIMARY, m.spin1.value, v3(-1., 0., 0.), rref1, pref1, ACC, posn, dvec, rad, gref);
            gref *= gscale;
        } else {
            gref = gscale / (1 + m.q.value) / sqr(r1);
        }
    } else {
        if (m.roche2) {
            const double ACC = m.delta_phase / 10.;
            double rref2, pref2, rad, ffac2 = r2 / rl2;
            ref_sphere(m.q.value, SECONDARY, m.spin2.value, ffac2, rref2, pref2);
            V3 posn, dvec;
            face(m.q.value, SECONDARY, m.spin2.value, v3(1., 0., 0.), rref2, pref2, ACC, posn, dvec, rad, gref);
            gref *= gscale;
        } else {
            gref = gscale * m.q.value / (1 + m.q.value) / sqr(r2);
        }
    }
    const size_t BLK = 8192, nn = st.size(), nblk = (nn + BLK - 1) / BLK;
    double sumfg = 0., sumf = 0.;
    for (size_t b = 0; b < nblk; b++) {
        double bfg = 0., bf = 0.;
        for (size_t i = b * BLK, e = std::min(nn, b * BLK + BLK); i < e; i++) {
            bfg += st[i].flux * st[i].gravity;
            bf += st[i].flux;
        }
        sumfg += bfg;
        sumf += bf;
    }
    if ((gref > 0) & (sumfg > 0) & (sumf > 0)) return std::log10(gref * sumfg / sumf);
    return 0.;
}

// comp_radius.cpp:16-42 / 51-76
double comp_radius(const std::vector<Face> &st, int star) {
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    const size_t BLK = 8192, nn = st.size(), nblk = (nn + BLK - 1) / BLK;
    double sumsa = 0., sumvol = 0.;
    for (size_t b = 0; b < nblk; b++) {
        double bsa = 0., bvol = 0.;
        for (size_t i = b * BLK, e = std::min(nn, b * BLK + BLK); i < e; i++) {
            V3 vec = sub(st[i].posn, cofm);
            double r = len(vec);
            double rcosa = dot(st[i].dirn, vec);
            bsa += st[i].area * rcosa / std::pow(r, 3.0);
            bvol += st[i].area * rcosa;
        }
        sumsa += bsa;
        sumvol += bvol;
    }
    return std::pow(sumvol / sumsa, 1. / 3.);
}

// ------------------------------------------------------- light_curve_comp
// re_scale (re_scale.cpp:16-43)
double re_scale(const lc_datum *data, size_t n, double *fit, double &chisq, double &wnok) {
    double wgt, sdy = 0., syy = 0.;
    wnok = 0.;
    for (size_t i = 0; i < n; i++) {
        if (data[i].weight > 0.) {
            wgt = data[i].weight / sqr(data[i].ferr);
            sdy += wgt * data[i].flux * fit[i];
            syy += wgt * fit[i] * fit[i];
            wnok += data[i].weight;
        }
    }
    double scale;
    if (wnok > 0.0) {
        scale = sdy / syy;
        for (size_t i = 0; i < n; i++) fit[i] = fit[i] * scale;
        chisq = 0.;
        for (size_t i = 0; i < n; i++)
            if (data[i].weight > 0.0) chisq += data[i].weight * sqr((data[i].flux - fit[i]) / data[i].ferr);
    } else {
        scale = 1.;
        chisq = 0.;
    }
    return scale;
}

struct Batch {
    std::vector<double> ex, ey, ez, phin, w;
    std::vector<int> idx;
    void push(double a, double b, double c, double p, double ww, int i) {
        ex.push_back(a); ey.push_back(b); ez.push_back(c);
        phin.push_back(p); w.push_back(ww); idx.push_back(i);
    }
    lc_phase_batch view() const {
        lc_phase_batch v;
        v.n = idx.size();
        v.ex = ex.data(); v.ey = ey.data(); v.ez = ez.data(); v.phin = phin.data();
        return v;
    }
};

lc_ldc make_ldc(const lc_model &m, int star) {
    lc_ldc l;
    std::memset(&l, 0, sizeof(l));
    if (star == PRIMARY) {
        l.c1 = m.ldc1_1.value; l.c2 = m.ldc1_2.value; l.c3 = m.ldc1_3.value; l.c4 = m.ldc1_4.value;
        l.mucrit = m.mucrit1; l.type = m.limb1;
    } else {
        l.c1 = m.ldc2_1.value; l.c2 = m.ldc2_2.value; l.c3 = m.ldc2_3.value; l.c4 = m.ldc2_4.value;
        l.mucrit = m.mucrit2; l.type = m.limb2;
    }
    return l;
}

// Weighted linear least squares through a singular value decomposition
// whose singular values below TOL = 1e-5 of the largest are set to zero
// (Subs::svdfit, svdfit.cpp:22-76; svbksb then skips them).  The
// decomposition here is the one-sided Jacobi method (Hestenes 1958): pairs
// of columns of the weighted design matrix are rotated until orthogonal;
// the column norms are then the singular values and the accumulated
// rotations the right singular vectors, so a = sum over kept j of
// v_j (col_j . b) / sigma_j^2.  cols[j][k] = design value of row k scaled by
// 1/sigma_k, b[k] = y_k / sigma_k.
void svd_lsq(std::vector<std::vector<double>> cols, const std::vector<double> &b, std::vector<double> &a) {
    const size_t nc = cols.size(), m = b.size();
    std::vector<std::vector<double>> V(nc, std::vector<double>(nc, 0.));
    for (size_t j = 0; j < nc; j++) V[j][j] = 1.;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.;
        for (size_t p = 0; p < nc; p++)
            for (size_t q = p + 1; q < nc; q++) {
                double al = 0., be = 0., ga = 0.;
                for (size_t k = 0; k < m; k++) {
                    al += cols[p][k] * cols[p][k];
                    be += cols[q][k] * cols[q][k];
                    ga += cols[p][k] * cols[q][k];
                }
                if (al == 0. || be == 0. || ga == 0.) continue;
                off = std::max(off, std::fabs(ga) / std::sqrt(al * be));
                const double zeta = (be - al) / (2. * ga);
                const double t = (zeta >= 0. ? 1. : -1.) / (std::fabs(zeta) + std::sqrt(1. + zeta * zeta));
                const double c = 1. / std::sqrt(1. + t * t), sn = c * t;
                for (size_t k = 0; k < m; k++) {
                    const double x = cols[p][k], y = cols[q][k];
                    cols[p][k] = c * x - sn * y;
                    cols[q][k] = sn * x + c * y;
                }
                for (size_t k = 0; k < nc; k++) {
                    const double x = V[k][p], y = V[k][q];
                    V[k][p] = c * x - sn * y;
                    V[k][q] = sn * x + c * y;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<double> sig(nc);
    double wmax = 0.;
    for (size_t j = 0; j < nc; j++) {
        double ss = 0.;
        for (size_t k = 0; k < m; k++) ss += cols[j][k] * cols[j][k];
        sig[j] = std::sqrt(ss);
        wmax = std::max(wmax, sig[j]);
    }
    const double thresh = 1e-5 * wmax;
    a.assign(nc, 0.);
    for (size_t j = 0; j < nc; j++) {
        if (!(sig[j] >= thresh) || sig[j] == 0.) continue;
        double ub = 0.;
        for (size_t k = 0; k < m; k++) ub += cols[j][k] * b[k];
        const double f = ub / (sig[j] * sig[j]);
        for (size_t i = 0; i < nc; i++) a[i] += V[i][j] * f;
    }
}

// Lcurve::light_curve_comp (light_curve_comp.cpp:119-599), no disc, no
// bright spot.  With iscale the two stars are separate components scaled
// by svd_lsq (light_curve_comp.cpp:378-416, 511-573): the disc, edge and
// spot columns are zero, so their factors come out 0, and star 2 is a
// component only when t2 > 0.  Each star's curve is summed here in the
// order of the batched path (fine sub-phases, then coarse), not per point
// as the reference's single-phase kernels do: the same sums up to rounding.
void light_curve_comp(const lc_model &mdl, const lc_datum *data, size_t n, const lc_eval_opts &o, double *calc,
                      lc_scalars &io, double *comp1 = nullptr, double *comp2 = nullptr) {
    if (mdl.add_disc || mdl.add_spot)
        throw OracleError(LC_E_UNSUPPORTED, "disc and bright spot are not part of this path");
    if (n == 0) throw OracleError(LC_E_ARG, "light_curve_comp: empty data");
    double r1, r2;
    get_r1r2(mdl, r1, r2);
    double rl2 = 1. - xl12(mdl.q.value, mdl.spin2.value);
    if (r2 < 0)
        r2 = rl2;
    else if (r2 > rl2)
        throw OracleError(LC_E_ROCHE_OVERFLOW, "light_curve_comp: the secondary star is larger than its Roche lobe!");
    const lc_ldc ldc1 = make_ldc(mdl, PRIMARY), ldc2 = make_ldc(mdl, SECONDARY);

    double rlens1 = 0.;
    if (mdl.glens1) {
        double gm = std::pow(1000. * mdl.velocity_scale.value, 3.0) * mdl.tperiod * DAY / TWOPI;
        double a = std::pow(gm / (TWOPI / DAY / mdl.tperiod * TWOPI / DAY / mdl.tperiod), 1. / 3.);
        rlens1 = 4. * gm / (1. + mdl.q.value) / a / (C_SI * C_SI);
    }

    std::vector<Face> star1f, star2f, star1c_own, star2c_own;
    set_star_grid(mdl, PRIMARY, true, star1f);
    set_star_grid(mdl, SECONDARY, true, star2f);
    const bool copy1 = (mdl.nlat1f == mdl.nlat1c);
    const bool copy2 = (mdl.nlat2f == mdl.nlat2c) && (!mdl.npole || r1 >= r2 || (mdl.nlatfill == 0 && mdl.nlngfill == 0));
    if (!copy1) set_star_grid(mdl, PRIMARY, false, star1c_own);
    if (!copy2) set_star_grid(mdl, SECONDARY, false, star2c_own);
    std::vector<Face> &star1c = copy1 ? star1f : star1c_own;
    std::vector<Face> &star2c = copy2 ? star2f : star2c_own;
    if (o.info) {
        std::fprintf(stderr, "Number of points for star 1 (fine) = %zu\n", star1f.size());
        std::fprintf(stderr, "Number of points for star 2 (fine) = %zu\n", star2f.size());
        std::fprintf(stderr, "Number of points for star 1 (coarse) = %zu\n", star1c.size());
        std::fprintf(stderr, "Number of points for star 2 (coarse) = %zu\n", star2c.size());
    }

    set_star_continuum(mdl, star1f, star2f);
    if (!copy1 || !copy2) set_star_continuum(mdl, star1c, star2c);

    Ginterp gint;
    gint.phase1 = mdl.phase1;
    gint.phase2 = mdl.phase2;
    gint.scale11 = gint.scale12 = gint.scale21 = gint.scale22 = 1.;
    const float vscale = (float)mdl.velocity_scale.value;
    const double q = mdl.q.value, iangle = mdl.iangle.value;
    if (mdl.nlat1c != mdl.nlat1f) {
        double ff = comp_star(PRIMARY, iangle, ldc1, 0.9999999999 * mdl.phase1, q, mdl.beam_factor1.value, vscale,
                              false, 0., gint, star1f, star1c);
        double fc = comp_star(PRIMARY, iangle, ldc1, 1.0000000001 * mdl.phase1, q, mdl.beam_factor1.value, vscale,
                              false, 0., gint, star1f, star1c);
        gint.scale11 = ff / fc;
        ff = comp_star(PRIMARY, iangle, ldc1, 1. - 0.9999999999 * mdl.phase1, q, mdl.beam_factor1.value, vscale, false,
                       0., gint, star1f, star1c);
        fc = comp_star(PRIMARY, iangle, ldc1, 1. - 1.0000000001 * mdl.phase1, q, mdl.beam_factor1.value, vscale, false,
                       0., gint, star1f, star1c);
        gint.scale12 = ff / fc;
    }
    if (!copy2) {
        const bool gl = mdl.glens1 != 0;
        double ff = comp_star(SECONDARY, iangle, ldc2, 1 - 1.0000000001 * mdl.phase2, q, mdl.beam_factor2.value, vscale,
                              gl, rlens1, gint, star2f, star2c);
        double fc = comp_star(SECONDARY, iangle, ldc2, 1 - 0.9999999999 * mdl.phase2, q, mdl.beam_factor2.value, vscale,
                              gl, rlens1, gint, star2f, star2c);
        gint.scale21 = ff / fc;
        ff = comp_star(SECONDARY, iangle, ldc2, 1.0000000001 * mdl.phase2, q, mdl.beam_factor2.value, vscale, gl,
                       rlens1, gint, star2f, star2c);
        fc = comp_star(SECONDARY, iangle, ldc2, 0.9999999999 * mdl.phase2, q, mdl.beam_factor2.value, vscale, gl,
                       rlens1, gint, star2f, star2c);
        gint.scale22 = ff / fc;
    }

    Flat fg1f, fg2f, fg1c_own, fg2c_own;
    flat_build(star1f, fg1f);
    flat_build(star2f, fg2f);
    if (!copy1) flat_build(star1c, fg1c_own);
    if (!copy2) flat_build(star2c, fg2c_own);
    const Flat &fg1c = copy1 ? fg1f : fg1c_own;
    const Flat &fg2c = copy2 ? fg2f : fg2c_own;

    double xmin = data[0].time, xmax = data[0].time;
    for (size_t np = 1; np < n; np++) {
        xmin = data[np].time > xmin ? xmin : data[np].time;
        xmax = data[np].time < xmax ? xmax : data[np].time;
    }
    double middle = (xmin + xmax) / 2., range = (xmax - xmin) / 2.;
    std::vector<double> phases(n), slfacs(n);
    const double t0 = mdl.t0.value, period = mdl.period.value, pdot = mdl.pdot.value, deltat = mdl.deltat.value;
    for (size_t np = 0; np < n; np++) {
        double phase = (data[np].time - t0) / period;
        for (int it = 0; it < 4; it++)
            phase -= (t0 + phase * (period + pdot * phase) - data[np].time) / (period + 2. * pdot * phase);
        phase += deltat / period / 2. * (std::cos(2. * PI * phase) - 1.);
        phases[np] = phase;
        double frac = (data[np].time - middle) / range;
        slfacs[np] = 1. + frac * (mdl.slope.value + frac * (mdl.quad.value + frac * mdl.cube.value));
    }

    const double ri = deg2rad(iangle);
    const double cosi = std::cos(ri), sini = std::sin(ri);
    const float vs = vscale;
    const double XCOFM = q / (1.0 + q);
    const double VFAC = vs / (C_SI / 1.e3);
    Batch pb1f, pb1c, pb2f, pb2c;
    for (size_t np = 0; np < n; np++) {
        double phase = phases[np], slfac = slfacs[np];
        double expose = data[np].expose / period;
        int ndiv = data[np].ndiv;
        double norm = std::max(1, ndiv - 1);
        for (int nd = 0; nd < ndiv; ++nd) {
            double phi, wgt;
            if (ndiv == 1) {
                phi = phase;
                wgt = 1.0;
            } else {
                phi = phase + expose * (nd - (ndiv - 1) / 2.0) / (ndiv - 1);
                wgt = (nd == 0 || nd == ndiv - 1) ? 0.5 : 1.0;
            }
            V3 earth = set_earth(cosi, sini, phi);
            double phin = phi - std::floor(phi);
            double w1 = slfac * wgt * gint.scale1(phi) / norm;
            double w2 = slfac * wgt * gint.scale2(phi) / norm;
            Batch &b1 = copy1 ? pb1f : ((gint.type(phi) == 1) ? pb1f : pb1c);
            Batch &b2 = copy2 ? pb2f : ((gint.type(phi) == 3) ? pb2f : pb2c);
            b1.push(earth.x, earth.y, earth.z, phin, w1, (int)np);
            b2.push(earth.x, earth.y, earth.z, phin, w2, (int)np);
        }
    }

    // iscale components come from comp_star1_flat / comp_star2_flat, which
    // use spin = 1 in the beaming term (comp_light_flat.cpp:565-568, 596-599;
    // light_curve_comp.cpp:389-413); the plain path uses the stars' spins
    // (light_curve_comp.cpp:462-475)
    ElemArgs a1 = {PRIMARY, mdl.beam_factor1.value != 0.0, false, mdl.beam_factor1.value,
                   mdl.iscale ? 1.0 : mdl.spin1.value, VFAC, XCOFM, 0., ldc_unpack(ldc1)};
    ElemArgs a2 = {SECONDARY, mdl.beam_factor2.value != 0.0, mdl.glens1 != 0, mdl.beam_factor2.value,
                   mdl.iscale ? 1.0 : mdl.spin2.value, VFAC, XCOFM, rlens1, ldc_unpack(ldc2)};
    std::vector<double> o1f(pb1f.idx.size()), o1c(pb1c.idx.size()), o2f(pb2f.idx.size()), o2c(pb2c.idx.size());
    if (!o1f.empty()) flat_sum_multi(flat_view(fg1f), pb1f.view(), a1, o1f.data());
    if (!o1c.empty()) flat_sum_multi(flat_view(fg1c), pb1c.view(), a1, o1c.data());
    if (!o2f.empty()) flat_sum_multi(flat_view(fg2f), pb2f.view(), a2, o2f.data());
    if (!o2c.empty()) flat_sum_multi(flat_view(fg2c), pb2c.view(), a2, o2c.data());

    std::vector<double> f1, f2;  // iscale: each star's curve
    if (mdl.iscale) {
        f1.assign(n, 0.);
        f2.assign(n, 0.);
        for (size_t k = 0; k < pb1f.idx.size(); k++) f1[pb1f.idx[k]] += pb1f.w[k] * o1f[k];
        for (size_t k = 0; k < pb1c.idx.size(); k++) f1[pb1c.idx[k]] += pb1c.w[k] * o1c[k];
        for (size_t k = 0; k < pb2f.idx.size(); k++) f2[pb2f.idx[k]] += pb2f.w[k] * o2f[k];
        for (size_t k = 0; k < pb2c.idx.size(); k++) f2[pb2c.idx[k]] += pb2c.w[k] * o2c[k];
        if (comp1) std::copy(f1.begin(), f1.end(), comp1);
        if (comp2) std::copy(f2.begin(), f2.end(), comp2);
    } else {
        for (size_t np = 0; np < n; np++) calc[np] = 0.;
        for (size_t k = 0; k < pb1f.idx.size(); k++) calc[pb1f.idx[k]] += pb1f.w[k] * o1f[k];
        for (size_t k = 0; k < pb1c.idx.size(); k++) calc[pb1c.idx[k]] += pb1c.w[k] * o1c[k];
        for (size_t k = 0; k < pb2f.idx.size(); k++) calc[pb2f.idx[k]] += pb2f.w[k] * o2f[k];
        for (size_t k = 0; k < pb2c.idx.size(); k++) calc[pb2c.idx[k]] += pb2c.w[k] * o2c[k];
        for (size_t np = 0; np < n; np++) calc[np] += mdl.third.value;
    }

    // wdwarf: comp_star1_flat at phase 0.5 (light_curve_comp.cpp:504-508;
    // comp_light_flat.cpp:543-571), beaming with spin = 1
    double wdwarf = 0.0;
    if (o.diagnostics) {
        double phi = 0.5;
        V3 earth = set_earth(cosi, sini, phi);
        double phin = phi - std::floor(phi);
        const Flat &st1 = (gint.type(phi) == 1) ? fg1f : fg1c;
        ElemArgs aw = a1;
        aw.spin = 1.0;
        double s = 1.0 * gint.scale1(phi) * flat_sum_one(flat_view(st1), earth.x, earth.y, earth.z, phin, aw);
        wdwarf = s / std::max(1, 0);
    }

    const bool has2 = mdl.t2.value > 0.;
    auto chi_all = [&] {
        io.wnok = 0.;
        io.chisq = 0.;
        for (size_t i = 0; i < n; i++) {
            if (data[i].weight > 0.) {
                io.wnok += data[i].weight;
                io.chisq += data[i].weight * sqr((data[i].flux - calc[i]) / data[i].ferr);
            }
        }
    };
    if (mdl.iscale) {
        if (o.scale) {
            // rows of points with weight > 0 and sigma = ferr / sqrt(weight) > 0
            std::vector<std::vector<double>> cols(has2 ? 5 : 4);
            std::vector<double> b;
            for (size_t np = 0; np < n; np++) {
                if (!(data[np].weight > 0.)) continue;
                const double z = data[np].ferr / std::sqrt(data[np].weight);
                if (!(z > 0.)) continue;
                const double is = 1. / z;
                cols[0].push_back(is * f1[np]);
                for (int c = 1; c < 4; c++) cols[c].push_back(0.);
                if (has2) cols[4].push_back(is * f2[np]);
                b.push_back(is * data[np].flux);
            }
            std::vector<double> a;
            svd_lsq(cols, b, a);
            a.resize(5, 0.);
            for (int k = 0; k < 5; k++) io.sfac[k] = a[k];
            wdwarf *= io.sfac[0];
        } else {
            wdwarf *= io.sfac[0];
        }
        for (size_t np = 0; np < n; np++) {
            calc[np] = io.sfac[0] * f1[np] + io.sfac[1] * 0. + io.sfac[2] * 0. + io.sfac[3] * 0.;
            if (has2) calc[np] += io.sfac[4] * f2[np];
        }
        if (o.scale || o.rdata) chi_all();
    } else if (o.scale) {
        double ssfac = re_scale(data, n, calc, io.chisq, io.wnok);
        wdwarf *= ssfac;
        for (int k = 0; k < 5; k++) io.sfac[k] = ssfac;
    } else {
        wdwarf *= io.sfac[0];
        for (size_t np = 0; np < n; np++) calc[np] *= io.sfac[0];
        if (o.rdata) {
            io.wnok = 0.;
            io.chisq = 0.;
            for (size_t i = 0; i < n; i++) {
                if (data[i].weight > 0.) {
                    io.wnok += data[i].weight;
                    io.chisq += data[i].weight * sqr((data[i].flux - calc[i]) / data[i].ferr);
                }
            }
        }
    }
    io.wdwarf = wdwarf;
    if (o.diagnostics) {
        io.logg1 = comp_gravity(mdl, star1f, PRIMARY);
        io.logg2 = comp_gravity(mdl, star2f, SECONDARY);
        io.rv1 = comp_radius(star1f, PRIMARY);
        io.rv2 = comp_radius(star2f, SECONDARY);
    } else {
        io.logg1 = io.logg2 = io.rv1 = io.rv2 = 0.0;
    }
}

// Lcurve::light_curve_comp_fast (light_curve_comp_fast.cpp:6-84)
void light_curve_comp_fast(const lc_model &mdl, const lc_datum *data, size_t n_data, const lc_eval_opts &o,
                           double *calc, lc_scalars &io) {
    const size_t M = (size_t)o.max_model_points;
    if (!(n_data > M)) {
        light_curve_comp(mdl, data, n_data, o, calc, io);
        return;
    }
    if (o.max_model_points < 2) throw OracleError(LC_E_ARG, "light_curve_comp_fast: max_model_points < 2");
    std::vector<lc_datum> model_data;
    std::vector<size_t> sample;
    for (int i = 0; i < o.max_model_points; ++i) {
        size_t idx = (size_t)i * (n_data - 1) / (size_t)(o.max_model_points - 1);
        model_data.push_back(data[idx]);
        sample.push_back(idx);
    }
    std::vector<double> model_calc(model_data.size());
    light_curve_comp(mdl, model_data.data(), model_data.size(), o, model_calc.data(), io);
    for (size_t i = 0; i < n_data; ++i) {
        auto it = std::lower_bound(sample.begin(), sample.end(), i);
        if (it == sample.begin()) {
            calc[i] = model_calc[0];
        } else if (it == sample.end()) {
            calc[i] = model_calc.back();
        } else {
            size_t idx2 = (size_t)(it - sample.begin());
            size_t idx1 = idx2 - 1;
            size_t i1 = sample[idx1], i2 = sample[idx2];
            double t = (data[i].time - data[i1].time) / (data[i2].time - data[i1].time);
            calc[i] = model_calc[idx1] * (1.0 - t) + model_calc[idx2] * t;
        }
    }
    if (o.rdata) {
        io.wnok = 0.;
        io.chisq = 0.;
        for (size_t i = 0; i < n_data; ++i) {
            if (data[i].weight > 0.) {
                io.wnok += data[i].weight;
                io.chisq += data[i].weight * sqr((data[i].flux - calc[i]) / data[i].ferr);
            }
        }
    }
}

// ---------------------------------------------------------- SoA transfer
void faces_to_soa(const std::vector<Face> &v, lc_star_faces *out) {
    for (size_t i = 0; i < v.size(); i++) {
        const Face &f = v[i];
        if (out->px) { out->px[i] = f.posn.x; out->py[i] = f.posn.y; out->pz[i] = f.posn.z; }
        if (out->dx) { out->dx[i] = f.dirn.x; out->dy[i] = f.dirn.y; out->dz[i] = f.dirn.z; }
        if (out->area) out->area[i] = f.area;
        if (out->gravity) out->gravity[i] = f.gravity;
        if (out->flux) out->flux[i] = f.flux;
        if (out->necl) out->necl[i] = f.necl;
        if (out->ingress) out->ingress[i] = f.ein;
        if (out->egress) out->egress[i] = f.eout;
    }
}

void soa_to_faces(const lc_star_faces *in, std::vector<Face> &v) {
    v.resize(in->n);
    for (size_t i = 0; i < in->n; i++) {
        Face &f = v[i];
        f.posn = v3(in->px[i], in->py[i], in->pz[i]);
        f.dirn = v3(in->dx[i], in->dy[i], in->dz[i]);
        f.area = in->area[i];
        f.gravity = in->gravity[i];
        f.flux = in->flux ? in->flux[i] : 0.f;
        f.necl = in->necl ? in->necl[i] : 0;
        f.ein = in->ingress ? in->ingress[i] : 0.;
        f.eout = in->egress ? in->egress[i] : 0.;
    }
}

template <class Fn> int guarded(Fn fn) {
    try {
        fn();
        return LC_OK;
    } catch (const OracleError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return LC_E_PARAM;
    }
}

} // namespace

// ====================================================================== C API
extern "C" {

const char *orc_last_error(void) { return g_last_error.c_str(); }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

double orc_xl1(double q, int *err) {
    double r = 0.;
    int e = guarded([&] { r = xl1(q); });
    if (err) *err = e;
    return r;
}
double orc_xl11(double q, double spin, int *err) {
    double r = 0.;
    int e = guarded([&] { r = xl11(q, spin); });
    if (err) *err = e;
    return r;
}
double orc_xl12(double q, double spin, int *err) {
    double r = 0.;
    int e = guarded([&] { r = xl12(q, spin); });
    if (err) *err = e;
    return r;
}

int orc_ref_sphere(double q, int star, double spin, double ffac, double *rref, double *pref) {
    return guarded([&] { ref_sphere(q, star, spin, ffac, *rref, *pref); });
}

double orc_rpot1(double q, double spin, const double *p) { return rpot1(q, spin, v3(p[0], p[1], p[2])); }
double orc_rpot2(double q, double spin, const double *p) { return rpot2(q, spin, v3(p[0], p[1], p[2])); }
void orc_drpot1(double q, double spin, const double *p, double *out) {
    V3 d = drpot1(q, spin, v3(p[0], p[1], p[2]));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void orc_drpot2(double q, double spin, const double *p, double *out) {
    V3 d = drpot2(q, spin, v3(p[0], p[1], p[2]));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

int orc_face(double q, int star, double spin, const double *dirn, double rref, double pref, double acc, double *pvec,
             double *dvec, double *r, double *g) {
    return guarded([&] {
        V3 pv, dv;
        face(q, star, spin, v3(dirn[0], dirn[1], dirn[2]), rref, pref, acc, pv, dv, *r, *g);
        pvec[0] = pv.x; pvec[1] = pv.y; pvec[2] = pv.z;
        dvec[0] = dv.x; dvec[1] = dv.y; dvec[2] = dv.z;
    });
}

void orc_face_chop_batch(double q, int star, double spin, const double *dx, const double *dy, const double *dz, int n,
                         double rref, double pref, double acc, double *rad, unsigned char *fail) {
    (void)dz;
    face_chop_batch(q, star, spin, dx, dy, n, rref, pref, acc, rad, fail);
}

int orc_sphere_eclipse(double cosi, double sini, const double *r, const double *c, double rsphere, double *phi1,
                       double *phi2, double *lam1, double *lam2) {
    return sphere_eclipse(cosi, sini, v3(r[0], r[1], r[2]), v3(c[0], c[1], c[2]), rsphere, *phi1, *phi2, *lam1, *lam2)
               ? 1
               : 0;
}

int orc_ingress_egress(double q, int star, double spin, double ffac, double iangle, double delta, const double *r,
                       double rref, double pref, double *ingress, double *egress) {
    (void)ffac;
    bool res = false;
    int e = guarded([&] {
        res = ingress_egress(q, star, spin, iangle, delta, v3(r[0], r[1], r[2]), *ingress, *egress, rref, pref);
    });
    if (e) return -e;
    return res ? 1 : 0;
}

int orc_fblink(double q, int star, double spin, double acc, const double *earth, const double *p, double rref,
               double pref) {
    bool res = false;
    int e = guarded([&] {
        res = fblink(q, star, spin, acc, v3(earth[0], earth[1], earth[2]), v3(p[0], p[1], p[2]), rref, pref);
    });
    if (e) return -e;
    return res ? 1 : 0;
}

int orc_numface(int nlat, int infill, double thelo, double thehi, int nlatfill, int nlngfill) {
    return numface(nlat, infill != 0, thelo, thehi, nlatfill, nlngfill);
}

double orc_planck(double wave, double temp) { return planck(wave, temp); }
double orc_dlpdlt(double wave, double temp) { return dlpdlt(wave, temp); }

int orc_star_grid_size(const lc_model *m, int star, int fine, size_t *n) {
    return guarded([&] { *n = (size_t)grid_setup(*m, star, fine != 0, false).nface; });
}

int orc_set_star_grid(const lc_model *m, int star, int fine, lc_star_faces *out) {
    return guarded([&] {
        std::vector<Face> v;
        set_star_grid(*m, star, fine != 0, v);
        if (v.size() != out->n) throw OracleError(LC_E_ARG, "orc_set_star_grid: output size mismatch");
        faces_to_soa(v, out);
    });
}

int orc_set_star_continuum(const lc_model *m, lc_star_faces *star1, lc_star_faces *star2) {
    return guarded([&] {
        std::vector<Face> s1, s2;
        soa_to_faces(star1, s1);
        soa_to_faces(star2, s2);
        set_star_continuum(*m, s1, s2);
        for (size_t i = 0; i < s1.size(); i++) star1->flux[i] = s1[i].flux;
        for (size_t i = 0; i < s2.size(); i++) star2->flux[i] = s2[i].flux;
    });
}

int orc_flat_sum_multi(int star, const lc_flat_grid *g, const lc_phase_batch *pb, const lc_ldc *ldc, double beam,
                       double spin, double VFAC, double XCOFM, int glens1, double rlens1, double *out) {
    return guarded([&] {
        ElemArgs a = {star, beam != 0.0, star == SECONDARY && glens1 != 0, beam, spin, VFAC, XCOFM, rlens1,
                      ldc_unpack(*ldc)};
        flat_sum_multi(*g, *pb, a, out);
    });
}

double orc_point_sum(int star, const lc_star_faces *f, const lc_ldc *ldc, double iangle, double phase, double q,
                     double beam, double vscale, int glens1, double rlens1) {
    std::vector<Face> v;
    soa_to_faces(f, v);
    const double XCOFM = q / (1.0 + q);
    const double VFAC = (float)vscale / (C_SI / 1.e3);
    double ri = deg2rad(iangle);
    V3 earth = set_earth(std::cos(ri), std::sin(ri), phase);
    double s = 0.0;
    for (const Face &pt : v) s += point_elem(pt, earth, phase, *ldc, star, beam, 1.0, VFAC, XCOFM, glens1 != 0, rlens1);
    return s;
}

double orc_re_scale(const lc_datum *data, size_t n, double *fit, double *chisq, double *wnok) {
    return re_scale(data, n, fit, *chisq, *wnok);
}

int orc_light_curve_comp(const lc_model *m, const lc_datum *data, size_t n, const lc_eval_opts *opts, double *calc,
                         lc_scalars *io) {
    return guarded([&] {
        if (opts->max_model_points > 0)
            light_curve_comp_fast(*m, data, n, *opts, calc, *io);
        else
            light_curve_comp(*m, data, n, *opts, calc, *io);
    });
}

int orc_star_curves(const lc_model *m, const lc_datum *data, size_t n, double *f1, double *f2) {
    return guarded([&] {
        lc_model mi = *m;
        mi.iscale = 1;
        lc_eval_opts o{};
        lc_scalars io{};
        for (double &f : io.sfac) f = 1.;
        std::vector<double> calc(n);
        light_curve_comp(mi, data, n, o, calc.data(), io, f1, f2);
    });
}

} // extern "C"
