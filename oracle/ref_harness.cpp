/*
 * ref_harness.cpp -- extern "C" wrappers around functions of the compiled
 * reference (TEST INFRASTRUCTURE ONLY).  Linked with the reference's own
 * translation units into oracle/_ref/libref.so by oracle/Makefile; used by
 * tests/golden/make_golden.py to write golden vectors and by the tests that
 * check the oracle against the reference where the library is built.
 *
 * The reference's JSON constructor of Lcurve::Model lives in model.cpp,
 * which is not built (see the Makefile).  Model itself is a plain struct of
 * Pparam and scalar members: trivially copyable with a trivial destructor,
 * so it is an implicit-lifetime type and storage obtained from malloc holds
 * a Model object (C++20 [intro.object]/10-13; the static_asserts below check
 * the two properties).  ref_model_new fills such an object member by member
 * from an lc_model, the field-for-field layout that this repository's own
 * parser produces, so the functions taking a Model -- set_star_grid,
 * set_star_continuum, light_curve_comp(_fast), comp_gravity -- run here as
 * compiled from the reference sources.  The other wrappers take explicit
 * arguments: Roche geometry, add_faces, the flux sums, the Point-path sums,
 * re_scale, comp_radius.
 */
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <type_traits>
#include <vector>

#include "../include/lcurve_amd.h"

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

/* The same on a FlatGrid with ragged faces [n1, n): face n1 + r has the
 * ranges min_[moff[r] .. moff[r+1]) / mout_[...] (comp_light_flat.cpp:26-65). */
void ref_flat_sum_multi_ragged(int star, size_t n, size_t n0, size_t n1, const double *dx, const double *dy,
                               const double *dz, const double *px, const double *py, const double *pz,
                               const double *flux, const double *in1, const double *out1, const int *moff,
                               const double *min_, const double *mout_, size_t np, const double *ex,
                               const double *ey, const double *ez, const double *phin, double c1, double c2,
                               double c3, double c4, double mucrit, int ltype, double beam, double spin, double VFAC,
                               double XCOFM, int glens1, double rlens1, double *out) {
    Lcurve::FlatGrid g;
    fill_flat(g, n, n0, n1, dx, dy, dz, px, py, pz, flux, in1, out1);
    g.moff.assign(moff, moff + (n - n1 + 1));
    g.min_.assign(min_, min_ + moff[n - n1]);
    g.mout_.assign(mout_, mout_ + moff[n - n1]);
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

/* ---------------------------------------------------------------- Model */
static_assert(std::is_trivially_copyable<Lcurve::Model>::value, "Model must be trivially copyable");
static_assert(std::is_trivially_destructible<Lcurve::Model>::value, "Model must be trivially destructible");

#define REF_PPARAMS(X)                                                                                            \
    X(q) X(iangle) X(r1) X(r2) X(cphi3) X(cphi4) X(spin1) X(spin2) X(t1) X(t2) X(ldc1_1) X(ldc1_2) X(ldc1_3)      \
    X(ldc1_4) X(ldc2_1) X(ldc2_2) X(ldc2_3) X(ldc2_4) X(velocity_scale) X(beam_factor1) X(beam_factor2) X(t0)    \
    X(period) X(pdot) X(deltat) X(gravity_dark1) X(gravity_dark2) X(absorb) X(slope) X(quad) X(cube) X(third)    \
    X(rdisc1) X(rdisc2) X(height_disc) X(beta_disc) X(temp_disc) X(texp_disc) X(lin_limb_disc)                   \
    X(quad_limb_disc) X(temp_edge) X(absorb_edge) X(radius_spot) X(length_spot) X(height_spot) X(expon_spot)     \
    X(epow_spot) X(angle_spot) X(yaw_spot) X(temp_spot) X(tilt_spot) X(cfrac_spot) X(stsp11_long)                \
    X(stsp11_lat) X(stsp11_fwhm) X(stsp11_tcen) X(stsp12_long) X(stsp12_lat) X(stsp12_fwhm) X(stsp12_tcen)       \
    X(stsp13_long) X(stsp13_lat) X(stsp13_fwhm) X(stsp13_tcen) X(stsp21_long) X(stsp21_lat) X(stsp21_fwhm)       \
    X(stsp21_tcen) X(stsp22_long) X(stsp22_lat) X(stsp22_fwhm) X(stsp22_tcen) X(uesp_long1) X(uesp_long2)        \
    X(uesp_lathw) X(uesp_taper) X(uesp_temp)

#define REF_NUMBERS(X)                                                                                            \
    X(delta_phase) X(nlat1f) X(nlat2f) X(nlat1c) X(nlat2c) X(nlatfill) X(nlngfill) X(lfudge) X(llo) X(lhi)       \
    X(phase1) X(phase2) X(wavelength) X(tperiod) X(mucrit1) X(mucrit2) X(nrad) X(nspot)

#define REF_FLAGS(X)                                                                                              \
    X(npole) X(roche1) X(roche2) X(eclipse1) X(eclipse2) X(glens1) X(use_radii) X(gdark_bolom1) X(gdark_bolom2)  \
    X(mirror) X(add_disc) X(opaque) X(add_spot) X(iscale)

static Lcurve::Model *model_of(const lc_model *pod) {
    void *mem = std::malloc(sizeof(Lcurve::Model));
    if (!mem) throw std::bad_alloc();
    std::memset(mem, 0, sizeof(Lcurve::Model));
    Lcurve::Model *m = static_cast<Lcurve::Model *>(mem);
#define SET_PP(n)                                                                                                 \
    m->n.value = pod->n.value;                                                                                    \
    m->n.range = pod->n.range;                                                                                    \
    m->n.dstep = pod->n.dstep;                                                                                    \
    m->n.vary = pod->n.vary != 0;                                                                                 \
    m->n.defined = pod->n.defined != 0;
    REF_PPARAMS(SET_PP)
#undef SET_PP
#define SET_NUM(n) m->n = pod->n;
    REF_NUMBERS(SET_NUM)
#undef SET_NUM
#define SET_FLAG(n) m->n = pod->n != 0;
    REF_FLAGS(SET_FLAG)
#undef SET_FLAG
    m->limb1 = pod->limb1 == LC_LDC_CLARET ? Lcurve::LDC::CLARET : Lcurve::LDC::POLY;
    m->limb2 = pod->limb2 == LC_LDC_CLARET ? Lcurve::LDC::CLARET : Lcurve::LDC::POLY;
    return m;
}

struct ModelHolder {
    Lcurve::Model *m;
    explicit ModelHolder(const lc_model *pod) : m(model_of(pod)) {}
    ~ModelHolder() { std::free(m); }
    const Lcurve::Model &operator*() const { return *m; }
};

static void faces_to_points(const lc_star_faces *f, std::vector<Lcurve::Point> &v) {
    v.resize(f->n);
    for (size_t i = 0; i < f->n; i++) {
        Lcurve::Point::etype ecl;
        if (f->necl && f->necl[i]) ecl.push_back(std::make_pair(f->ingress[i], f->egress[i]));
        v[i] = Lcurve::Point(Subs::Vec3(f->px[i], f->py[i], f->pz[i]), Subs::Vec3(f->dx[i], f->dy[i], f->dz[i]),
                             f->area[i], f->gravity[i], ecl);
        v[i].flux = f->flux ? f->flux[i] : 0.f;
    }
}

/* Lcurve::set_star_grid (set_star_grid.cpp:49-523) for one star and grid;
 * *n in: capacity of the out arrays, out: faces written. */
int ref_set_star_grid(const lc_model *pod, int star, int fine, lc_star_faces *out) {
    return guard([&] {
        ModelHolder m(pod);
        std::vector<Lcurve::Point> v;
        Lcurve::set_star_grid(*m, star_of(star), fine != 0, v);
        if (v.size() > out->n) throw std::runtime_error("ref_set_star_grid: output arrays too small");
        out->n = v.size();
        for (size_t i = 0; i < v.size(); i++) {
            const Lcurve::Point &p = v[i];
            out->px[i] = p.posn.x(); out->py[i] = p.posn.y(); out->pz[i] = p.posn.z();
            out->dx[i] = p.dirn.x(); out->dy[i] = p.dirn.y(); out->dz[i] = p.dirn.z();
            out->area[i] = p.area;
            out->gravity[i] = p.gravity;
            out->flux[i] = p.flux;
            out->necl[i] = p.eclipse.empty() ? 0 : 1;
            out->ingress[i] = p.eclipse.empty() ? 0. : p.eclipse[0].first;
            out->egress[i] = p.eclipse.empty() ? 0. : p.eclipse[0].second;
        }
    });
}

/* Lcurve::set_star_continuum (set_star_continuum.cpp:26-297): fluxes of
 * the faces given (posn, dirn, area, gravity) written to star?->flux. */
int ref_set_star_continuum(const lc_model *pod, lc_star_faces *star1, lc_star_faces *star2) {
    return guard([&] {
        ModelHolder m(pod);
        std::vector<Lcurve::Point> v1, v2;
        faces_to_points(star1, v1);
        faces_to_points(star2, v2);
        Lcurve::set_star_continuum(*m, v1, v2);
        for (size_t i = 0; i < v1.size(); i++) star1->flux[i] = v1[i].flux;
        for (size_t i = 0; i < v2.size(); i++) star2->flux[i] = v2[i].flux;
    });
}

/* Lcurve::comp_gravity1 / comp_gravity2 (comp_gravity.cpp) on the faces
 * given (area, gravity, flux). */
double ref_comp_gravity(const lc_model *pod, int star, const lc_star_faces *f) {
    double out = 0.;
    if (guard([&] {
            ModelHolder m(pod);
            std::vector<Lcurve::Point> v;
            faces_to_points(f, v);
            out = star == 0 ? Lcurve::comp_gravity1(*m, v) : Lcurve::comp_gravity2(*m, v);
        }) != 0)
        return -1.0e300;
    return out;
}

/* Lcurve::light_curve_comp (light_curve_comp.cpp:119-599), or
 * light_curve_comp_fast (light_curve_comp_fast.cpp:6-84) when
 * o->max_model_points > 0.  io->sfac is read when o->scale is 0. */
int ref_light_curve_comp(const lc_model *pod, const lc_datum *data, size_t n, const lc_eval_opts *o, double *calc,
                         lc_scalars *io) {
    return guard([&] {
        ModelHolder m(pod);
        Lcurve::Data d(n);
        for (size_t i = 0; i < n; i++) {
            d[i].time = data[i].time;
            d[i].expose = data[i].expose;
            d[i].flux = data[i].flux;
            d[i].ferr = data[i].ferr;
            d[i].weight = data[i].weight;
            d[i].ndiv = data[i].ndiv;
        }
        std::vector<double> sfac(io->sfac, io->sfac + 5), c;
        double wdwarf = 0., chisq = 0., wnok = 0., logg1 = 0., logg2 = 0., rv1 = 0., rv2 = 0.;
        if (o->max_model_points > 0)
            Lcurve::light_curve_comp_fast(*m, d, o->scale != 0, o->rdata != 0, false, sfac, c, wdwarf, chisq, wnok,
                                          logg1, logg2, rv1, rv2, o->max_model_points, o->diagnostics != 0);
        else
            Lcurve::light_curve_comp(*m, d, o->scale != 0, o->rdata != 0, false, sfac, c, wdwarf, chisq, wnok,
                                     logg1, logg2, rv1, rv2, o->diagnostics != 0);
        for (size_t i = 0; i < n && i < c.size(); i++) calc[i] = c[i];
        for (size_t k = 0; k < 5 && k < sfac.size(); k++) io->sfac[k] = sfac[k];
        io->wdwarf = wdwarf;
        io->chisq = chisq;
        io->wnok = wnok;
        io->logg1 = logg1;
        io->logg2 = logg2;
        io->rv1 = rv1;
        io->rv2 = rv2;
    });
}

} // extern "C"
