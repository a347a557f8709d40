/*
 * ref_harness.cpp -- extern "C" wrappers around functions of the compiled
 * reference (TEST INFRASTRUCTURE ONLY).  Linked with the reference's own
 * translation units into oracle/_ref/libref.so by oracle/Makefile; used by
 * tests/golden/make_golden.py to write golden vectors and by the tests that
 * check the oracle against the reference where the library is built.
 *
 * Lcurve::Model cannot be constructed here (model.cpp is not built, see the
 * Makefile), so only functions that take explicit arguments are wrapped:
 * Roche geometry, add_faces, the flux sums, the Point-path sums, re_scale.
 */
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "lcurve.h"
#include "../lroche_base/roche.h"
#include "../new_subs.h"

namespace {
thread_local std::string g_err;

template <class F> int guard(F f) {
    try {
        f();
        return 0;
    } catch (const Roche::Roche_Error &e) {
        g_err = e.what();
        return 2;
    } catch (const Lcurve::Lcurve_Error &e) {
        g_err = e.what();
        return 3;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 4;
    }
}

Roche::STAR star_of(int s) { return s == 0 ? Roche::PRIMARY : Roche::SECONDARY; }
} // namespace

Subs::xy<double, double> envelope(double iangle, double lambda, double r1);

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

double ref_xl1(double q) { return Roche::xl1(q); }
double ref_xl11(double q, double spin) { return Roche::xl11(q, spin); }
double ref_xl12(double q, double spin) { return Roche::xl12(q, spin); }

void ref_ref_sphere(double q, int star, double spin, double ffac, double *rref, double *pref) {
    Roche::ref_sphere(q, star_of(star), spin, ffac, *rref, *pref);
}

double ref_rpot1(double q, double spin, const double *p) { return Roche::rpot1(q, spin, Subs::Vec3(p[0], p[1], p[2])); }
double ref_rpot2(double q, double spin, const double *p) { return Roche::rpot2(q, spin, Subs::Vec3(p[0], p[1], p[2])); }

int ref_face(double q, int star, double spin, const double *dirn, double rref, double pref, double acc, double *pvec,
             double *dvec, double *r, double *g) {
    return guard([&] {
        Subs::Vec3 pv, dv;
        Roche::face(q, star_of(star), spin, Subs::Vec3(dirn[0], dirn[1], dirn[2]), rref, pref, acc, pv, dv, *r, *g);
        pvec[0] = pv.x(); pvec[1] = pv.y(); pvec[2] = pv.z();
        dvec[0] = dv.x(); dvec[1] = dv.y(); dvec[2] = dv.z();
    });
}

void ref_face_chop_batch(double q, int star, double spin, const double *dx, const double *dy, const double *dz, int n,
                         double rref, double pref, double acc, double *rad, unsigned char *fail) {
    Roche::face_chop_batch(q, star_of(star), spin, dx, dy, dz, n, rref, pref, acc, rad, fail);
}

int ref_sphere_eclipse(double cosi, double sini, const double *r, const double *c, double rs, double *phi1,
                       double *phi2, double *lam1, double *lam2) {
    return Roche::sphere_eclipse(cosi, sini, Subs::Vec3(r[0], r[1], r[2]), Subs::Vec3(c[0], c[1], c[2]), rs, *phi1,
                                 *phi2, *lam1, *lam2)
               ? 1
               : 0;
}

int ref_ingress_egress(double q, int star, double spin, double ffac, double iangle, double delta, const double *r,
                       double rref, double pref, double *ingress, double *egress) {
    bool res = false;
    int e = guard([&] {
        res = Roche::ingress_egress(q, star_of(star), spin, ffac, iangle, delta, Subs::Vec3(r[0], r[1], r[2]),
                                    *ingress, *egress, rref, pref);
    });
    if (e) return -e;
    return res ? 1 : 0;
}

int ref_fblink(double q, int star, double spin, double acc, const double *earth, const double *p, double rref,
               double pref) {
    bool res = false;
    int e = guard([&] {
        res = Roche::fblink(q, star_of(star), spin, acc, Subs::Vec3(earth[0], earth[1], earth[2]),
                            Subs::Vec3(p[0], p[1], p[2]), rref, pref);
    });
    if (e) return -e;
    return res ? 1 : 0;
}

int ref_numface(int nlat, int infill, double thelo, double thehi, int nlatfill, int nlngfill) {
    return Lcurve::numface(nlat, infill != 0, thelo, thehi, nlatfill, nlngfill);
}

void ref_envelope(double iangle_rad, double lambda, double r1, double *x, double *y) {
    Subs::xy<double, double> e = envelope(iangle_rad, lambda, r1);
    *x = e.x;
    *y = e.y;
}

double ref_planck(double wave, double temp) { return Subs::planck(wave, temp); }
double ref_dlpdlt(double wave, double temp) { return Subs::dlpdlt(wave, temp); }

/* Lcurve::add_faces over one theta segment; nface_total is the size of the
 * grid (numface), *nface the running offset (in/out).  Outputs are SoA
 * arrays of nface_total entries; necl holds the eclipse count (0/1). */
int ref_add_faces(int nface_total, int *nface, double tlo, double thi, double dtheta, int nlatfill, int nlngfill,
                  int npole, int which_star, double q, double iangle, double r1, double r2, double rref1, double rref2,
                  int roche1, int roche2, double spin1, double spin2, int eclipse, double gref, double pref1,
                  double pref2, double ffac1, double ffac2, double delta, double *px, double *py, double *pz,
                  double *dx, double *dy, double *dz, float *area, float *gravity, int *necl, double *ingress,
                  double *egress) {
    return guard([&] {
        std::vector<Lcurve::Point> star(nface_total);
        int nf = *nface;
        Lcurve::add_faces(star, nf, tlo, thi, dtheta, nlatfill, nlngfill, npole != 0, star_of(which_star), q, iangle,
                          r1, r2, rref1, rref2, roche1 != 0, roche2 != 0, spin1, spin2, eclipse != 0, gref, pref1,
                          pref2, ffac1, ffac2, delta);
        for (int i = *nface; i < nf && i < nface_total; i++) {
            const Lcurve::Point &p = star[i];
            px[i] = p.posn.x(); py[i] = p.posn.y(); pz[i] = p.posn.z();
            dx[i] = p.dirn.x(); dy[i] = p.dirn.y(); dz[i] = p.dirn.z();
            area[i] = p.area;
            gravity[i] = p.gravity;
            necl[i] = (int)p.eclipse.size();
            ingress[i] = p.eclipse.empty() ? 0. : p.eclipse[0].first;
            egress[i] = p.eclipse.empty() ? 0. : p.eclipse[0].second;
        }
        *nface = nf;
    });
}

static void fill_flat(Lcurve::FlatGrid &g, size_t n, size_t n0, size_t n1, const double *dx, const double *dy,
                      const double *dz, const double *px, const double *py, const double *pz, const double *flux,
                      const double *in1, const double *out1) {
    g.n = n; g.n0 = n0; g.n1 = n1;
    g.dx.assign(dx, dx + n); g.dy.assign(dy, dy + n); g.dz.assign(dz, dz + n);
    g.px.assign(px, px + n); g.py.assign(py, py + n); g.pz.assign(pz, pz + n);
    g.flux.assign(flux, flux + n);
    g.in1.assign(in1, in1 + (n1 - n0));
    g.out1.assign(out1, out1 + (n1 - n0));
    g.moff.assign(1, 0);
    g.min_.clear();
    g.mout_.clear();
}

/* Lcurve::flat_sum_star{1,2}_multi on a FlatGrid without ragged faces. */
void ref_flat_sum_multi(int star, size_t n, size_t n0, size_t n1, const double *dx, const double *dy,
                        const double *dz, const double *px, const double *py, const double *pz, const double *flux,
                        const double *in1, const double *out1, size_t np, const double *ex, const double *ey,
                        const double *ez, const double *phin, double c1, double c2, double c3, double c4,
                        double mucrit, int ltype, double beam, double spin, double VFAC, double XCOFM, int glens1,
                        double rlens1, double *out) {
    Lcurve::FlatGrid g;
    fill_flat(g, n, n0, n1, dx, dy, dz, px, py, pz, flux, in1, out1);
    Lcurve::PhaseBatch pb;
    for (size_t k = 0; k < np; k++) pb.push(ex[k], ey[k], ez[k], phin[k], 1.0, (int)k);
    Lcurve::LDC ldc(c1, c2, c3, c4, mucrit, ltype == 0 ? Lcurve::LDC::POLY : Lcurve::LDC::CLARET);
    if (star == 0)
        Lcurve::flat_sum_star1_multi(g, pb, ldc, beam, spin, VFAC, XCOFM, out);
    else
        Lcurve::flat_sum_star2_multi(g, pb, ldc, beam, spin, VFAC, XCOFM, glens1 != 0, rlens1, out);
}

static void fill_points(std::vector<Lcurve::Point> &v, size_t n, const double *px, const double *py,
                        const double *pz, const double *dx, const double *dy, const double *dz, const float *flux,
                        const int *necl, const double *ingress, const double *egress) {
    v.resize(n);
    for (size_t i = 0; i < n; i++) {
        Lcurve::Point::etype ecl;
        if (necl[i]) ecl.push_back(std::make_pair(ingress[i], egress[i]));
        v[i] = Lcurve::Point(Subs::Vec3(px[i], py[i], pz[i]), Subs::Vec3(dx[i], dy[i], dz[i]), 0., 1., ecl);
        v[i].flux = flux[i];
    }
}

/* Lcurve::comp_star1 / comp_star2 (Point path) with one grid serving as fine
 * and coarse and the given Ginterp. */
double ref_comp_star(int star, size_t n, const double *px, const double *py, const double *pz, const double *dx,
                     const double *dy, const double *dz, const float *flux, const int *necl, const double *ingress,
                     const double *egress, double c1, double c2, double c3, double c4, double mucrit, int ltype,
                     double iangle, double phase, double expose, int ndiv, double q, double beam, float vscale,
                     int glens1, double rlens1, double phase1, double phase2) {
    std::vector<Lcurve::Point> v;
    fill_points(v, n, px, py, pz, dx, dy, dz, flux, necl, ingress, egress);
    Lcurve::LDC ldc(c1, c2, c3, c4, mucrit, ltype == 0 ? Lcurve::LDC::POLY : Lcurve::LDC::CLARET);
    Lcurve::Ginterp gint{};
    gint.phase1 = phase1;
    gint.phase2 = phase2;
    gint.scale11 = gint.scale12 = gint.scale21 = gint.scale22 = 1.;
    if (star == 0) return Lcurve::comp_star1(iangle, ldc, phase, expose, ndiv, q, beam, vscale, gint, v, v);
    return Lcurve::comp_star2(iangle, ldc, phase, expose, ndiv, q, beam, vscale, glens1 != 0, rlens1, gint, v, v);
}

/* Lcurve::re_scale */
double ref_re_scale(size_t n, const double *time, const double *flux, const double *ferr, const double *weight,
                    double *fit, double *chisq, double *wnok) {
    Lcurve::Data d(n);
    for (size_t i = 0; i < n; i++) {
        d[i].time = time[i];
        d[i].expose = 0.;
        d[i].flux = flux[i];
        d[i].ferr = ferr[i];
        d[i].weight = weight[i];
        d[i].ndiv = 1;
    }
    std::vector<double> f(fit, fit + n);
    double s = Lcurve::re_scale(d, f, *chisq, *wnok);
    std::memcpy(fit, f.data(), n * sizeof(double));
    return s;
}

/* Lcurve::comp_radius1 / comp_radius2 */
double ref_comp_radius(int star, size_t n, const double *px, const double *py, const double *pz, const double *dx,
                       const double *dy, const double *dz, const float *area) {
    std::vector<Lcurve::Point> v(n);
    for (size_t i = 0; i < n; i++)
        v[i] = Lcurve::Point(Subs::Vec3(px[i], py[i], pz[i]), Subs::Vec3(dx[i], dy[i], dz[i]), area[i], 1.,
                             Lcurve::Point::etype());
    return star == 0 ? Lcurve::comp_radius1(v) : Lcurve::comp_radius2(v);
}

} // extern "C"
