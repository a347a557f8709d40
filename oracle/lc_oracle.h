/*
 * lc_oracle.h -- FP64 CPU/OpenMP restatement of the lcurve_re light-curve
 * path.  TEST INFRASTRUCTURE ONLY: the correctness oracle that tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use as the
 * checker.  The product (lcurve_re_amd/) never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative
 * to the reference repository root).  Arithmetic follows the reference
 * expression by expression; the library is compiled with -ffp-contract=off
 * so no FMA contraction changes the rounding.
 */
#ifndef LC_ORACLE_H
#define LC_ORACLE_H

#include "../include/lcurve_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *orc_last_error(void);

/* src/lroche_base/xl1.cpp:42-76, roche.cpp:8-80.  *err = lc_status. */
double orc_xl1(double q, int *err);
double orc_xl11(double q, double spin, int *err);
double orc_xl12(double q, double spin, int *err);

/* src/lroche_base/ref_sphere.cpp:30-41 */
int orc_ref_sphere(double q, int star, double spin, double ffac,
                   double *rref, double *pref);

/* src/lroche_base/rpot.cpp:37-86, drpot.cpp:40-85 */
double orc_rpot1(double q, double spin, const double *p);
double orc_rpot2(double q, double spin, const double *p);
void orc_drpot1(double q, double spin, const double *p, double *out);
void orc_drpot2(double q, double spin, const double *p, double *out);

/* src/lroche_base/face.cpp:163-220 (scalar face) */
int orc_face(double q, int star, double spin, const double *dirn, double rref,
             double pref, double acc, double *pvec, double *dvec, double *r,
             double *g);

/* src/lroche_base/face.cpp:71-161 (lock-step batched radius solve) */
void orc_face_chop_batch(double q, int star, double spin, const double *dx,
                         const double *dy, const double *dz, int n,
                         double rref, double pref, double acc, double *rad,
                         unsigned char *fail);

/* src/lroche_base/sphere_eclipse.cpp:33-75 */
int orc_sphere_eclipse(double cosi, double sini, const double *r,
                       const double *c, double rsphere, double *phi1,
                       double *phi2, double *lam1, double *lam2);

/* src/lroche_base/ingress_egress.cpp:35-97.  Returns 1 eclipsed, 0 not,
 * -(lc_status) on a search failure. */
int orc_ingress_egress(double q, int star, double spin, double ffac,
                       double iangle, double delta, const double *r,
                       double rref, double pref, double *ingress,
                       double *egress);

/* src/lroche_base/fblink.cpp:34-104.  1 eclipsed, 0 not, <0 error. */
int orc_fblink(double q, int star, double spin, double acc,
               const double *earth, const double *p, double rref,
               double pref);

/* src/lcurve_base/numface.cpp:14-53 */
int orc_numface(int nlat, int infill, double thelo, double thehi,
                int nlatfill, int nlngfill);

/* src/new_subs.cpp:125-164 */
double orc_planck(double wave, double temp);
double orc_dlpdlt(double wave, double temp);

/* Face count of set_star_grid (set_star_grid.cpp:49-162). */
int orc_star_grid_size(const lc_model *m, int star, int fine, size_t *n);

/* Lcurve::set_star_grid (set_star_grid.cpp:49-523).  out arrays hold
 * orc_star_grid_size() entries; flux is set to 0. */
int orc_set_star_grid(const lc_model *m, int star, int fine,
                      lc_star_faces *out);

/* Lcurve::set_star_continuum (set_star_continuum.cpp:26-297) */
int orc_set_star_continuum(const lc_model *m, lc_star_faces *star1,
                           lc_star_faces *star2);

/* Lcurve::flat_sum_star{1,2}_multi (comp_light_flat.cpp:241-465) */
int orc_flat_sum_multi(int star, const lc_flat_grid *g,
                       const lc_phase_batch *pb, const lc_ldc *ldc,
                       double beam, double spin, double VFAC, double XCOFM,
                       int glens1, double rlens1, double *out);

/* Raw Point-path sum of comp_star1 / comp_star2 at one phase with spin = 1
 * for the beaming term (comp_light.cpp:21-92, 221-315): faces ascending,
 * no Ginterp scale.  The Ginterp scale factors are ratios of these sums. */
double orc_point_sum(int star, const lc_star_faces *f, const lc_ldc *ldc,
                     double iangle, double phase, double q, double beam,
                     double vscale, int glens1, double rlens1);

/* Lcurve::re_scale (re_scale.cpp:16-43) */
double orc_re_scale(const lc_datum *data, size_t n, double *fit, double *chisq, double *wnok);

/* Lcurve::light_curve_comp (light_curve_comp.cpp:119-599) or
 * light_curve_comp_fast (light_curve_comp_fast.cpp:6-84) when
 * opts->max_model_points > 0.  Returns lc_status. */
int orc_light_curve_comp(const lc_model *m, const lc_datum *data, size_t n,
                         const lc_eval_opts *opts, double *calc,
                         lc_scalars *io);

/* The two stars' curves that the iscale path scales separately (each with
 * the slope factor and exposure weights, no third light). */
int orc_star_curves(const lc_model *m, const lc_datum *data, size_t n,
                    double *f1, double *f2);

/* Number of OpenMP threads the oracle uses (0 = runtime default). */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
