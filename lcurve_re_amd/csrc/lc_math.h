// lc_math.h -- Roche geometry and eclipse search shared by host and device
// code of the engine (FP64).  Compiled with -ffp-contract=off so every
// expression rounds as written.  Functions return lc_status-style codes
// instead of throwing; the reference file:line each one follows is cited.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define LC_HD __host__ __device__ __forceinline__

namespace lcx {

// constants.h:15-96 (H and K are the reference's non-CODATA values)
constexpr double C_SI = 2.99792458e8;
constexpr double H_SI = 6.6262e-34;
constexpr double K_SI = 1.3806e-23;
constexpr double PI = 3.14159265358979323846264;
constexpr double TWOPI = 2. * PI;
constexpr float DAY = 86400.f;
constexpr double M_PI_ = 3.14159265358979323846; // new_subs.h:199 uses M_PI

enum { PRIMARY = 0, SECONDARY = 1 };

// status codes (mirror lc_status of lcurve_amd.h)
enum {
    ST_OK = 0,
    ST_ROCHE_OVERFLOW = 1,
    ST_FACE = 2,
    ST_ECLIPSE = 3,
    ST_PARAM = 4,
    ST_DEVICE = 5,
    ST_UNSUPPORTED = 6,
    ST_ARG = 7,
    ST_LENSING = 8
};

struct V3 {
    double x, y, z;
};
LC_HD V3 v3(double x, double y, double z) {
    V3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
LC_HD double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LC_HD double sqrlen(const V3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
LC_HD double len(const V3 &a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
LC_HD V3 add(const V3 &a, const V3 &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
LC_HD V3 sub(const V3 &a, const V3 &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
LC_HD V3 scl(double c, const V3 &a) { return v3(c * a.x, c * a.y, c * a.z); }
LC_HD double sqr(double a) { return a * a; }
LC_HD double deg2rad(double d) { return M_PI_ * d / 180.; }
LC_HD double rad2deg(double r) { return 180. * r / M_PI_; }
LC_HD double dmin(double a, double b) { return b < a ? b : a; }  // std::min(a, b)
LC_HD double dmax(double a, double b) { return a < b ? b : a; }  // std::max(a, b)
LC_HD double sign_of(double a, double b) { return b >= 0. ? fabs(a) : -fabs(a); }

// ------------------------------------------------------------ potentials
// Synchronous-or-not Roche potential about star 1 (rpot.cpp:37-48)
LC_HD double rpot1(double q, double spin, const V3 &p) {
    double mu = q / (1 + q);
    double comp = 1. - mu;
    double x2y2 = sqr(p.x) + sqr(p.y);
    double r1sq = x2y2 + sqr(p.z);
    double r1 = sqrt(r1sq);
    double r2 = sqrt(r1sq + 1. - 2. * p.x);
    return (-comp / r1 - mu / r2 - sqr(spin) * x2y2 / 2. + mu * p.x);
}
// ... about star 2 (rpot.cpp:58-86)
LC_HD double rpot2(double q, double spin, const V3 &p) {
    double mu = q / (1 + q);
    double comp = 1. - mu;
    double x2y2 = p.x * p.x + p.y * p.y;
    double r1sq = x2y2 + p.z * p.z;
    double r1 = sqrt(r1sq);
    double r2 = sqrt(r1sq + 1.0 - 2.0 * p.x);
    double spin_sq = spin * spin;
    return -comp / r1 - mu / r2 - spin_sq * (0.5 + 0.5 * x2y2 - p.x) - comp * p.x;
}
LC_HD double rpot_star(int star, double q, double spin, const V3 &p) {
    return star == PRIMARY ? rpot1(q, spin, p) : rpot2(q, spin, p);
}
// gradients (drpot.cpp:40-85)
LC_HD V3 drpot_star(int star, double q, double spin, const V3 &p) {
    double r1sq = sqrlen(p);
    double r1 = sqrt(r1sq);
    double r2sq = r1sq + 1. - 2. * p.x;
    double r2 = sqrt(r2sq);
    double mu = q / (1 + q);
    double mu1 = mu / r2 / r2sq;
    double comp = (1. - mu) / r1 / r1sq;
    double ssq = sqr(spin);
    double gx = star == PRIMARY ? comp * p.x + mu1 * (p.x - 1.) - ssq * p.x + mu
                                : comp * p.x + mu1 * (p.x - 1.) - ssq * (p.x - 1.) + mu - 1.;
    return v3(gx, comp * p.y + mu1 * p.y - ssq * p.y, comp * p.z + mu1 * p.z);
}

// A potential evaluation that keeps its distances, so that the gradient at
// the same point needs two products or divisions more instead of a second
// evaluation (dbrent asks for f(u) and then, unless it stops, for df(u)).
// Host code: the operations of rpot1/rpot2 and drpot_star term for term
// ((1-mu)/r1 and mu/r2 are the common subexpressions), so the values are
// bit-identical to calling the two functions.  Device code (the line
// searches of the eclipse kernels): 1/r1 and 1/r2 from reciprocal square
// roots (within about an ulp), a third of the instructions of a square root
// and a division each; the searches compare these values with one another
// and with the surface potential, so results move only where a comparison
// is within an ulp.
struct PotPoint {
    V3 r;
    double d1, d2;  // host: r1^2, r2^2; device: 1/r1, 1/r2
    double c1, m2;  // (1-mu)/r1, mu/r2
};
LC_HD double rpot_keep(int star, double q, double spin, const V3 &p, PotPoint &k) {
    double mu = q / (1 + q);
    double comp = 1. - mu;
    double x2y2 = p.x * p.x + p.y * p.y;
    double r1sq = x2y2 + p.z * p.z;
    double r2sq = r1sq + 1. - 2. * p.x;
    k.r = p;
#ifdef __HIP_DEVICE_COMPILE__
    k.d1 = rsqrt(r1sq);
    k.d2 = rsqrt(r2sq);
    k.c1 = comp * k.d1;
    k.m2 = mu * k.d2;
#else
    double r1 = sqrt(r1sq);
    double r2 = sqrt(r2sq);
    k.d1 = r1sq;
    k.d2 = r2sq;
    k.c1 = comp / r1;
    k.m2 = mu / r2;
#endif
    if (star == PRIMARY) return (-k.c1 - k.m2 - sqr(spin) * x2y2 / 2. + mu * p.x);
    return -k.c1 - k.m2 - spin * spin * (0.5 + 0.5 * x2y2 - p.x) - comp * p.x;
}
LC_HD V3 drpot_kept(int star, double q, double spin, const PotPoint &k) {
    const V3 &p = k.r;
    double mu = q / (1 + q);
#ifdef __HIP_DEVICE_COMPILE__
    double mu1 = k.m2 * (k.d2 * k.d2);
    double comp = k.c1 * (k.d1 * k.d1);
#else
    double mu1 = k.m2 / k.d2;
    double comp = k.c1 / k.d1;
#endif
    double ssq = sqr(spin);
    double gx = star == PRIMARY ? comp * p.x + mu1 * (p.x - 1.) - ssq * p.x + mu
                                : comp * p.x + mu1 * (p.x - 1.) - ssq * (p.x - 1.) + mu - 1.;
    return v3(gx, comp * p.y + mu1 * p.y - ssq * p.y, comp * p.z + mu1 * p.z);
}

// Earth vector at an orbital phase (set_earth.cpp:39-46)
LC_HD V3 set_earth(double cosi, double sini, double phase) {
    double rp = TWOPI * phase;
    double s, c;
#ifdef __HIP_DEVICE_COMPILE__
    sincos(rp, &s, &c);
#else
    s = sin(rp);
    c = cos(rp);
#endif
    return v3(sini * c, -sini * s, cosi);
}

LC_HD double rpot_val(double q, int star, double spin, const V3 &earth, const V3 &p, double lam) {
    return rpot_star(star, q, spin, add(p, scl(lam, earth)));  // rpot_val.cpp:21-30
}
LC_HD void rpot_grad(double q, int star, double spin, const V3 &earth, const V3 &p, double lam, double &dphi,
                     double &dlam) {  // rpot_grad.cpp:25-40
    V3 d = drpot_star(star, q, spin, add(p, scl(lam, earth)));
    V3 ed = v3(earth.y, -earth.x, 0.);
    dphi = TWOPI * lam * dot(d, ed);
    dlam = dot(d, earth);
}

// ------------------------------------------------------------ L1 points
// Newton iteration on the L1 quintic (roche.cpp:8-80, xl1.cpp:42-76).
// kind 0: xl1(q); 1: xl11(q, spin); 2: xl12(q, spin).  Returns < 0 on
// failure.
LC_HD double lagrange_l1(int kind, double q, double spin, int *st) {
    *st = ST_OK;
    if (q <= 0.) {
        *st = ST_PARAM;
        return 0.;
    }
    double ssq = spin * spin, mu = q / (1. + q);
    double a1 = -1. + mu, a2 = 2. - 2. * mu, a3, a4, a5, a6;
    if (kind == 0) {
        a3 = -1. + mu; a4 = 1. + 2. * mu; a5 = -2. - mu; a6 = 1.;
    } else if (kind == 1) {
        a3 = -1. + mu; a4 = ssq + 2. * mu; a5 = -2. * ssq - mu; a6 = ssq;
    } else {
        a3 = -ssq + mu; a4 = 3. * ssq + 2. * mu - 2.; a5 = 1. - mu - 3. * ssq; a6 = ssq;
    }
    double d1 = 1. * a2, d2 = 2. * a3, d3 = 3. * a4, d4 = 4. * a5, d5 = 5. * a6;
    int n = 0;
    double xold = 0., x = 1. / (1. + q);
    while (n < 1000 && fabs(x - xold) > 1.e-12 * fabs(x)) {
        xold = x;
        double f = x * (x * (x * (x * (x * a6 + a5) + a4) + a3) + a2) + a1;
        double df = x * (x * (x * (x * d5 + d4) + d3) + d2) + d1;
        x -= f / df;
        n++;
    }
    if (n == 1000) *st = ST_PARAM;
    return x;
}

// ref_sphere.cpp:30-41
LC_HD int ref_sphere(double q, int star, double spin, double ffac, double &rref, double &pref) {
    int st;
    if (star == PRIMARY) {
        double tref = lagrange_l1(1, q, spin, &st);
        rref = dmin(1.001 * ffac, 1.0) * tref;
        pref = rpot1(q, spin, v3(ffac * tref, 0., 0.));
    } else {
        double tref = 1. - lagrange_l1(2, q, spin, &st);
        rref = dmin(1.001 * ffac, 1.0) * tref;
        pref = rpot2(q, spin, v3(1. - ffac * tref, 0., 0.));
    }
    return st;
}

// ----------------------------------------------------------- face solves
// Potential along the ray cofm + r*d.  `own_sqrt` selects the scalar
// Roche::face form (distance to the own centre as a square root,
// face.cpp:98-113) or the ring-solve form that uses r itself
// (face.cpp:137-154).
struct RayPot {
    double mu, comp, spin_sq, cx, dx, dy, dz;
    bool primary, own_sqrt;
    LC_HD double operator()(double r) const {
        if (own_sqrt) {
            double px = cx + r * dx, py = r * dy, pz = r * dz;
            double x2y2 = px * px + py * py;
            double r1sq = x2y2 + pz * pz;
            double r1 = sqrt(r1sq);
            double r2 = sqrt(r1sq + 1. - 2. * px);
            if (primary) return -comp / r1 - mu / r2 - spin_sq * x2y2 / 2. + mu * px;
            return -comp / r1 - mu / r2 - spin_sq * (0.5 + 0.5 * x2y2 - px) - comp * px;
        }
        double rdx = r * dx;
        double px = cx + rdx;
        double py = r * dy;
        double x2y2 = px * px + py * py;
        double rr = r * r;
        double down = r;
#ifdef __HIP_DEVICE_COMPILE__
        // device: the companion term from a reciprocal square root (within
        // about an ulp; a chop decision moves only where the potential is
        // within an ulp of pref, and then the radius by at most acc)
        const double icomp = rsqrt(rr + 1. + (primary ? -2. : 2.) * rdx);
        if (primary) return -comp / down - mu * icomp - spin_sq * x2y2 / 2. + mu * px;
        return -comp * icomp - mu / down - spin_sq * (0.5 + 0.5 * x2y2 - px) - comp * px;
#else
        double dcomp = sqrt(rr + 1. + (primary ? -2. : 2.) * rdx);
        if (primary) return -comp / down - mu / dcomp - spin_sq * x2y2 / 2. + mu * px;
        return -comp / dcomp - mu / down - spin_sq * (0.5 + 0.5 * x2y2 - px) - comp * px;
#endif
    }
};

LC_HD RayPot make_raypot(double q, int star, double spin, const V3 &d, bool own_sqrt) {
    RayPot rp;
    rp.mu = q / (1 + q);
    rp.comp = 1. - rp.mu;
    rp.spin_sq = spin * spin;
    rp.cx = star == PRIMARY ? 0. : 1.;
    rp.dx = d.x;
    rp.dy = d.y;
    rp.dz = d.z;
    rp.primary = star == PRIMARY;
    rp.own_sqrt = own_sqrt;
    return rp;
}

// Radius along a ray where the potential reaches pref: reference check,
// up to 30 halvings, binary chop to acc (face.cpp:163-265 / 71-161).
// Ring form (own_sqrt=false): returns 1/2/3 on the three failures of
// face_chop_batch.  Scalar form: returns ST_FACE on any failure, including
// a chop count of exactly 100 (face.cpp:264).
LC_HD int solve_radius(const RayPot &rp, double rref, double pref, double acc, double &rad) {
    if (rp(rref) < pref) return rp.own_sqrt ? ST_FACE : 1;
    double r1 = rref / 2, r2 = rref, tref = pref + 1.;
    for (int it = 0; it < 30 && tref > pref; it++) {
        r1 = r2 / 2;
        tref = rp(r1);
        if (tref > pref) r2 = r1;
    }
    if (tref > pref) return rp.own_sqrt ? ST_FACE : 2;
    int nchop = 0;
    while (r2 - r1 > acc && nchop < 100) {
        double r = (r1 + r2) / 2.;
        if (rp(r) < pref)
            r1 = r;
        else
            r2 = r;
        nchop++;
    }
    rad = (r1 + r2) / 2.;
    if (rp.own_sqrt) return nchop == 100 ? ST_FACE : ST_OK;
    return (r2 - r1 > acc) ? 3 : ST_OK;
}

// Scalar Roche::face: position, unit normal, |grad| (face.cpp:215-272)
LC_HD int face_scalar(double q, int star, double spin, const V3 &dirn, double rref, double pref, double acc, V3 &pvec,
                      V3 &dvec, double &r, double &g) {
    RayPot rp = make_raypot(q, star, spin, dirn, true);
    int st = solve_radius(rp, rref, pref, acc, r);
    if (st) return st;
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    pvec = add(cofm, scl(r, dirn));
    dvec = drpot_star(star, q, spin, pvec);
    g = len(dvec);
    dvec = v3(dvec.x / g, dvec.y / g, dvec.z / g);
    return ST_OK;
}

// ------------------------------------------------------- sphere eclipses
// Phase/multiplier window of the eclipse of point r by a sphere
// (sphere_eclipse.cpp:33-75)
// Whether sphere_eclipse reports an eclipse, without the window: the same
// operations up to its two early returns.
LC_HD bool sphere_eclipse_test(double cosi, double sini, const V3 &r, const V3 &c, double rs) {
    const V3 d = sub(r, c);
    const double PDIST = sqrt(sqr(d.x) + sqr(d.y));
    const double BQUAD = d.z * cosi - PDIST * sini;
    if (BQUAD >= 0.) return false;
    const double CQUAD = sqrlen(d) - sqr(rs);
    return !(sqr(BQUAD) - CQUAD <= 0.);  // as written there, NaN included
}
LC_HD bool sphere_eclipse(double cosi, double sini, const V3 &r, const V3 &c, double rs, double &phi1, double &phi2,
                          double &lam1, double &lam2) {
    const V3 d = sub(r, c);
    const double PDIST = sqrt(sqr(d.x) + sqr(d.y));
    const double BQUAD = d.z * cosi - PDIST * sini;
    if (BQUAD >= 0.) return false;
    const double CQUAD = sqrlen(d) - sqr(rs);
    double fac = sqr(BQUAD) - CQUAD;
    if (fac <= 0.) return false;
    fac = sqrt(fac);
    lam2 = -BQUAD + fac;
    lam1 = dmax(0., CQUAD / lam2);
    if (CQUAD < 0.) {
        phi1 = 0.;
        phi2 = 1.;
    } else {
        double delta = acos((cosi * d.z + sqrt(CQUAD)) / (sini * PDIST));
        double phi = atan2(d.y, -d.x);
        phi1 = (phi - delta) / TWOPI;
        phi1 -= floor(phi1);
        phi2 = phi1 + 2. * delta / TWOPI;
    }
    return true;
}

// Multiplier window along a given line of sight (sphere_eclipse.cpp:94-116)
LC_HD bool sphere_cut(const V3 &earth, const V3 &r, const V3 &c, double rs, double &lam1, double &lam2) {
    const V3 d = sub(r, c);
    const double BQUAD = dot(earth, d);
    if (BQUAD >= 0.) return false;
    const double CQUAD = sqrlen(d) - sqr(rs);
    double fac = sqr(BQUAD) - CQUAD;
    if (fac <= 0.) return false;
    fac = sqrt(fac);
    lam2 = -BQUAD + fac;
    lam1 = dmax(0., CQUAD / lam2);
    return true;
}

// ----------------------------------------------------------------- dbrent
// Derivative-assisted Brent minimisation with early exit once the value
// drops below fref (dbrent.cpp:19-126).  Returns ST_ECLIPSE after 100
// iterations.  F supplies f(x, keep) and df(keep): the derivative at the
// point of the last value, from what that evaluation kept.  dbrent_at
// starts from a value fx = f(bx) (and its kept distances) that the caller
// has already computed; dbrent evaluates it first.
template <class F>
LC_HD int dbrent_at(double ax, double bx, double cx, const F &fn, double acc, double fref, double fx,
                    typename F::Keep &kp, double &xmin, double &fmin) {
    double a = (ax < cx ? ax : cx);
    double b = (ax > cx ? ax : cx);
    double x = bx, w = bx, v = bx;
    double fw = fx, fv = fx;
    if (fx < fref) {
        xmin = x;
        fmin = fx;
        return ST_OK;
    }
    double dx = fn.df(kp), dw = dx, dv = dx;
    double e = 0., d = 0.;
    for (int iter = 0; iter < 100; iter++) {
        double xm = 0.5 * (a + b);
        double tol1 = acc, tol2 = 2. * tol1;
        if (fabs(x - xm) <= (tol2 - 0.5 * (b - a))) {
            xmin = x;
            fmin = fx;
            return ST_OK;
        }
        if (fabs(e) > tol1) {
            double d1 = 2. * (b - a), d2 = d1;
            if (dw != dx) d1 = (w - x) * dx / (dx - dw);
            if (dv != dx) d2 = (v - x) * dx / (dx - dv);
            double u1 = x + d1, u2 = x + d2;
            bool ok1 = (a - u1) * (u1 - b) > 0. && dx * d1 <= 0.;
            bool ok2 = (a - u2) * (u2 - b) > 0. && dx * d2 <= 0.;
            double olde = e;
            e = d;
            if (ok1 || ok2) {
                if (ok1 && ok2)
                    d = (fabs(d1) < fabs(d2) ? d1 : d2);
                else
                    d = ok1 ? d1 : d2;
                if (fabs(d) <= fabs(0.5 * olde)) {
                    double u = x + d;
                    if (u - a < tol2 || b - u < tol2) d = sign_of(tol1, xm - x);
                } else {
                    e = (dx >= 0. ? a - x : b - x);
                    d = 0.5 * e;
                }
            } else {
                e = (dx >= 0. ? a - x : b - x);
                d = 0.5 * e;
            }
        } else {
            e = (dx >= 0. ? a - x : b - x);
            d = 0.5 * e;
        }
        // the full step or, below tol1, a step of tol1; one evaluation site
        // for both, so that lanes taking either branch share it
        const bool full = fabs(d) >= tol1;
        const double u = full ? x + d : x + sign_of(tol1, d);
        const double fu = fn.f(u, kp);
        if (fu < fref) {
            xmin = u;
            fmin = fu;
            return ST_OK;
        }
        if (!full && fu > fx) {
            xmin = x;
            fmin = fx;
            return ST_OK;
        }
        double du = fn.df(kp);
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
    return ST_ECLIPSE;
}
template <class F>
LC_HD int dbrent(double ax, double bx, double cx, const F &fn, double acc, double fref, double &xmin, double &fmin) {
    typename F::Keep kp;
    const double fx = fn.f(bx, kp);
    return dbrent_at(ax, bx, cx, fn, acc, fref, fx, kp, xmin, fmin);
}

// Potential and directional derivative along a line in (phase, lambda)
// space (roche.h:258-324, Rpot / Drpot)
struct PhaseLamLine {
    double q, spin, cosi, sini;
    V3 p;
    double phi, dphi, lam, dlam;
    int star;
    // the two values, sharing the earth vector and the distances
    struct Keep {
        PotPoint pt;
        V3 earth;
        double lam;
    };
#ifdef __HIP_DEVICE_COMPILE__
    // device: every value of the search from the kept pair (one arithmetic)
    LC_HD double f(double x) const {
        Keep k;
        return f(x, k);
    }
    LC_HD double df(double x) const {
        Keep k;
        f(x, k);
        return df(k);
    }
#else
    LC_HD double f(double x) const {
        return rpot_val(q, star, spin, set_earth(cosi, sini, phi + dphi * x), p, lam + dlam * x);
    }
    LC_HD double df(double x) const {
        double dp, dl;
        rpot_grad(q, star, spin, set_earth(cosi, sini, phi + dphi * x), p, lam + dlam * x, dp, dl);
        return (dp * dphi + dl * dlam);
    }
#endif
    LC_HD double f(double x, Keep &k) const {
        k.earth = set_earth(cosi, sini, phi + dphi * x);
        k.lam = lam + dlam * x;
        return rpot_keep(star, q, spin, add(p, scl(k.lam, k.earth)), k.pt);
    }
    LC_HD double df(const Keep &k) const {
        V3 d = drpot_kept(star, q, spin, k.pt);
        V3 ed = v3(k.earth.y, -k.earth.x, 0.);
        double dp = TWOPI * k.lam * dot(d, ed), dl = dot(d, k.earth);
        return (dp * dphi + dl * dlam);
    }
};

// Potential along the line of sight (roche.h:328-365, Rlpot / Dlrpot)
struct SightLine {
    double q, spin;
    V3 earth, p;
    int star;
    using Keep = PotPoint;
#ifdef __HIP_DEVICE_COMPILE__
    LC_HD double f(double x) const {
        Keep k;
        return f(x, k);
    }
#else
    LC_HD double f(double x) const { return rpot_val(q, star, spin, earth, p, x); }
    LC_HD double df(double x) const {
        double dp, dl;
        rpot_grad(q, star, spin, earth, p, x, dp, dl);
        return dl;
    }
#endif
    LC_HD double f(double x, Keep &k) const { return rpot_keep(star, q, spin, add(p, scl(x, earth)), k); }
    LC_HD double df(const Keep &k) const { return dot(drpot_kept(star, q, spin, k), earth); }
};

// Line minimisation in (phase, lambda) space with boundary handling
// (pot_min.cpp:127-420).  Returns a status; results through references.
LC_HD int linmin(double q, int star, double spin, double cosi, double sini, const V3 &p, double &phi, double &lam,
                 double dphi, double dlam, double phi1, double phi2, double lam1, double lam2, double pref, double acc,
                 double &pmin, bool &jammed) {
    PhaseLamLine fn;
    fn.q = q; fn.spin = spin; fn.cosi = cosi; fn.sini = sini; fn.p = p;
    fn.phi = phi; fn.dphi = dphi; fn.lam = lam; fn.dlam = dlam; fn.star = star;

    double xmax = 1.e30, xend;
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
    const double xa = 0., fa = fn.f(xa);
    double xb = 1.e-8 * xmax, fb = fn.f(xb);
    int nten = 0;
    while ((fb >= fa || xa == xb) && nten < 7) {
        nten++;
        xb *= 10.;
        fb = fn.f(xb);
    }
    if (fb <= pref) {
        phi += dphi * xb;
        lam += dlam * xb;
        pmin = fb;
        return ST_OK;
    }
    if (fb >= fa) {
        pmin = fa;
        return ST_OK;
    }
    bool bracketted = false;
    double xc = 0., fc = 0.;
    const double xbold = xb, fbold = fb;
    xmax -= xb;
    for (int n = 1; n <= 5; n++) {
        xc = xbold + xmax * n / 5;
        fc = fn.f(xc);
        if (fc <= pref) {
            phi += dphi * xc;
            lam += dlam * xc;
            pmin = fc;
            return ST_OK;
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
        double dc = fn.df(xc);
        if (dc > 0.) {
            xb = xbold;
            double db = fn.df(xb);
            if (!(db < 0.)) return ST_ECLIPSE;  // "failed to bracket minimum, error 1"
            double xm = (xb + xc) / 2.;
            while ((xc - xb) > 1.e-8 * xc) {
                xm = (xb + xc) / 2.;
                if (fn.df(xm) > 0.)
                    xc = xm;
                else
                    xb = xm;
            }
            double fm = fn.f(xm);
            if (fm <= pref) {
                phi += dphi * xm;
                lam += dlam * xm;
                pmin = fm;
                return ST_OK;
            }
            if (!(fm < fc && fm < fbold)) return ST_ECLIPSE;  // "error 3"
            xb = xm;
        } else {
            // stuck against a boundary: slide along it
            jammed = true;
            phi += dphi * xmax;
            lam += dlam * xmax;
            xmax = 1.;
            rpot_grad(q, star, spin, set_earth(cosi, sini, phi), p, lam, dphi, dlam);
            if (nbound == 1 || nbound == 2) {
                phi = nbound == 1 ? phi1 : phi2;
                dphi = 0.;
                dlam = dlam > 0. ? lam1 - lam : lam2 - lam;
            } else if (nbound == 3 || nbound == 4) {
                lam = nbound == 3 ? lam1 : lam2;
                dlam = 0.;
                dphi = dphi > 0. ? phi1 - phi : phi2 - phi;
            }
            fn.phi = phi;
            fn.dphi = dphi;
            fn.lam = lam;
            fn.dlam = dlam;
            nten = 0;
            xb = 1.e-6;
            fb = fn.f(xb);
            while (fb >= fa && nten < 7) {
                xb *= 10.;
                nten++;
                fb = fn.f(xb);
            }
            if (fb <= pref) {
                phi += dphi * xb;
                lam += dlam * xb;
                pmin = fb;
                return ST_OK;
            }
            if (fb >= fa) {
                pmin = fa;
                return ST_OK;
            }
            bracketted = false;
            for (int n = 1; n <= 5; n++) {
                xc = xmax * n / 5;
                fc = fn.f(xc);
                if (fc <= pref) {
                    phi += dphi * xc;
                    lam += dlam * xc;
                    pmin = fc;
                    return ST_OK;
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
                if (fn.df(xc) > 0.) return ST_ECLIPSE;  // "error 2"
                phi += dphi * xmax;
                lam += dlam * xmax;
                pmin = fc;
                return ST_OK;
            }
        }
    }
    double xacc = acc / sqrt(sqr(TWOPI * dphi) + sqr(dlam));
    double xmin, fmin;
    int st = dbrent(xa, xb, xc, fn, xacc, pref, xmin, fmin);
    if (st) return st;
    pmin = fmin;
    phi += dphi * xmin;
    lam += dlam * xmin;
    return ST_OK;
}

// pot_min's first point, the middle of the region, and its potential
// (pot_min.cpp:45-49); the search ends there when it is <= pref.
LC_HD double pot_min_start(double q, double cosi, double sini, int star, double spin, const V3 &p, double phi1,
                           double phi2, double lam1, double lam2, double &phi, double &lam, V3 &earth, V3 &r) {
    phi = (phi1 + phi2) / 2.;
    lam = (lam1 + lam2) / 2.;
    earth = set_earth(cosi, sini, phi);
    r = add(p, scl(lam, earth));
    return rpot_star(star, q, spin, r);
}

// Conjugate-gradient search for a potential below pref over the region
// (phi1..phi2, lam1..lam2) (pot_min.cpp:38-100).  *hit = eclipse found.
LC_HD int pot_min(double q, double cosi, double sini, int star, double spin, const V3 &p, double phi1, double phi2,
                  double lam1, double lam2, double rref, double pref, double acc, double &phi, double &lam, bool &hit) {
    double dphi, dlam;
    V3 earth, r;
    double pot = pot_min_start(q, cosi, sini, star, spin, p, phi1, phi2, lam1, lam2, phi, lam, earth, r);
    {
        V3 d = drpot_star(star, q, spin, r);
        V3 ed = v3(earth.y, -earth.x, 0.);
        dphi = TWOPI * lam * dot(d, ed);
        dlam = dot(d, earth);
    }
    hit = true;
    if (pot <= pref) return ST_OK;
    double gdphi = -dphi, gdlam = -dlam;
    double hdphi = gdphi, hdlam = gdlam;
    dphi = gdphi;
    dlam = gdlam;
    const double DELPHI = q / (1. + q) * sqr(acc / rref) / 2.;
    bool jammed = false;
    for (int its = 0; its < 200; its++) {
        double pmin;
        int st = linmin(q, star, spin, cosi, sini, p, phi, lam, dphi, dlam, phi1, phi2, lam1, lam2, pref, acc, pmin,
                        jammed);
        if (st) return st;
        if (pmin <= pref) return ST_OK;
        hit = false;
        if (jammed || fabs(pmin - pot) < DELPHI) return ST_OK;
        pot = pmin;
        rpot_grad(q, star, spin, set_earth(cosi, sini, phi), p, lam, dphi, dlam);
        double gg = gdphi * gdphi + gdlam * gdlam;
        if (gg == 0.) return ST_OK;
        double dgg = (dphi + gdphi) * dphi + (dlam + gdlam) * dlam;
        double gam = dgg / gg;
        gdphi = -dphi;
        gdlam = -dlam;
        dphi = hdphi = gdphi + gam * hdphi;
        dlam = hdlam = gdlam + gam * hdlam;
        hit = true;
    }
    return ST_ECLIPSE;  // "too many iterations"
}

// Is p hidden by the Roche-distorted star along `earth`? (fblink.cpp:34-104)
LC_HD int fblink(double q, int star, double spin, double acc, const V3 &earth, const V3 &p, double rref, double pref,
                 bool &hidden) {
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    double lam1, lam2;
    hidden = false;
    if (!sphere_cut(earth, p, cofm, rref, lam1, lam2)) return ST_OK;
    if (lam1 == 0.) {
        hidden = true;
        return ST_OK;
    }
    SightLine fn;
    fn.q = q; fn.spin = spin; fn.earth = earth; fn.p = p; fn.star = star;
    int nstep = 1;
    double step = lam2 - lam1;
    double f1 = 0., f2 = 0., flam = 1., lam = lam1;
    SightLine::Keep klam;  // distances at lam, handed to dbrent
    while (step > acc) {
        lam = lam1 + step / 2.;
        for (int n = 0; n < nstep; n++) {
            flam = fn.f(lam, klam);
            if (flam <= pref) {
                hidden = true;
                return ST_OK;
            }
            if (nstep == 1) {
                f1 = fn.f(lam1);
                f2 = fn.f(lam2);
            }
            if (flam < f1 && flam < f2) break;
            lam += step;
        }
        if (flam < f1 && flam < f2) break;
        step /= 2.;
        nstep *= 2;
    }
    if (flam < f1 && flam < f2) {
        double xmin, fmin;
        // the reference's dbrent evaluates f(lam) again first: same value
        int st = dbrent_at(lam1, lam, lam2, fn, acc, pref, flam, klam, xmin, fmin);
        if (st) return st;
        hidden = fmin < pref;
    }
    return ST_OK;
}

// Ingress/egress phases of point r eclipsed by a Roche-distorted star
// (ingress_egress.cpp:35-97).  *ecl = eclipsed.
LC_HD int ingress_egress(double q, int star, double spin, double cosi, double sini, double delta, const V3 &r,
                         double rref, double pref, double &ingress, double &egress, bool &ecl) {
    const V3 cofm = star == PRIMARY ? v3(0., 0., 0.) : v3(1., 0., 0.);
    double phi1, phi2, lam1, lam2, phi, lam;
    ecl = false;
    if (!sphere_eclipse(cosi, sini, r, cofm, rref, phi1, phi2, lam1, lam2)) return ST_OK;
    const double ACC = 2. * sqrt(2. * TWOPI * (lam2 - lam1) * delta);
    bool hit;
    int st = pot_min(q, cosi, sini, star, spin, r, phi1, phi2, lam1, lam2, rref, pref, ACC, phi, lam, hit);
    if (st) return st;
    if (!hit) return ST_OK;
    double pin = phi, pout = phi1;
    while (fabs(pin - pout) > delta) {
        double pmid = (pin + pout) / 2.;
        bool h;
        st = fblink(q, star, spin, ACC, set_earth(cosi, sini, pmid), r, rref, pref, h);
        if (st) return st;
        if (h)
            pin = pmid;
        else
            pout = pmid;
    }
    ingress = (pin + pout) / 2.;
    ingress = ingress - floor(ingress);
    pin = phi;
    pout = phi2;
    while (fabs(pin - pout) > delta) {
        double pmid = (pin + pout) / 2.;
        bool h;
        st = fblink(q, star, spin, ACC, set_earth(cosi, sini, pmid), r, rref, pref, h);
        if (st) return st;
        if (h)
            pin = pmid;
        else
            pout = pmid;
    }
    egress = (pin + pout) / 2.;
    egress = egress - floor(egress);
    if (egress < ingress) egress++;
    ecl = true;
    return ST_OK;
}

// Black body (new_subs.cpp:125-136) and d ln B / d ln T (158-164)
LC_HD double planck(double wave, double temp) {
    const double FAC1 = 2.e27 * H_SI * C_SI;
    const double FAC2 = 1.e9 * H_SI * C_SI / K_SI;
    double efac = FAC2 / (wave * temp);
    if (efac > 40.) return FAC1 * exp(-efac) / (wave * sqr(wave));
    return FAC1 / (exp(efac) - 1.) / (wave * sqr(wave));
}
LC_HD double dlpdlt(double wave, double temp) {
    const double FAC2 = 1.e9 * H_SI * C_SI / K_SI;
    double efac = FAC2 / (wave * temp);
    return efac / (1. - exp(-efac));
}

}  // namespace lcx
