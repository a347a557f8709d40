/*
 * lcurve_amd.h -- C ABI of the MI355X light-curve engine.
 *
 * Plain C types only (no C++ or HIP types in any signature).  Every entry
 * point returns an lc_status; on failure lc_last_error() holds the message.
 *
 * The reference (Fabmat1/lcurve_re) has no FFI: its light-curve path is
 * called from C++ through Lcurve::light_curve_comp / light_curve_comp_fast
 * and the Lcurve::Model JSON/parameter API.  Each entry point below names
 * the reference interface it replaces; include/lcurve_re/lcurve.h is the
 * C++ drop-in layer built on top of this ABI.
 */
#ifndef LCURVE_AMD_H
#define LCURVE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum lc_status {
    LC_OK = 0,
    LC_E_ROCHE_OVERFLOW = 1, /* star larger than its Roche lobe
                                (light_curve_comp.cpp:135-138, set_star_grid.cpp:64-79) */
    LC_E_FACE = 2,           /* surface solve failed (face.cpp:229-265)              */
    LC_E_ECLIPSE_SEARCH = 3, /* pot_min/linmin/dbrent/fblink failed
                                (pot_min.cpp:98,271,291,401; dbrent.cpp:649)         */
    LC_E_PARAM = 4,          /* malformed or illegal parameter (model.cpp:10-360)     */
    LC_E_DEVICE = 5,         /* HIP runtime error                                     */
    LC_E_UNSUPPORTED = 6,    /* an option outside this engine (none at present)        */
    LC_E_ARG = 7,            /* invalid argument (null pointer, bad size)             */
    LC_E_LENSING = 8         /* lensing configuration rejected (set_star_grid.cpp:80-100) */
} lc_status;

/* Precision of the per-face flux arithmetic. */
enum {
    LC_PREC_MIXED = 0, /* FP32 per face, FP64 per-phase accumulation (default) */
    LC_PREC_FP64 = 1   /* FP64 per face in the reference's block order         */
};

enum { LC_STAR1 = 0, LC_STAR2 = 1 };      /* Roche::PRIMARY / SECONDARY (roche.h:13) */
enum { LC_LDC_POLY = 0, LC_LDC_CLARET = 1 }; /* LDC::LDCtype (lcurve.h:204)        */

/* ------------------------------------------------------------ parameters */
/* One physical parameter: Lcurve::Pparam (lcurve.h:159-195). */
typedef struct lc_pparam {
    double value, range, dstep;
    int32_t vary, defined;
#ifdef __cplusplus
    /* Pparam's implicit conversion to its value (lcurve.h:179) */
    operator double() const { return value; }
#endif
} lc_pparam;

/* Lcurve::Model (lcurve.h:317-421), field for field. */
typedef struct lc_model {
    lc_pparam q, iangle, r1, r2, cphi3, cphi4, spin1, spin2;
    lc_pparam t1, t2;
    lc_pparam ldc1_1, ldc1_2, ldc1_3, ldc1_4;
    lc_pparam ldc2_1, ldc2_2, ldc2_3, ldc2_4;
    lc_pparam velocity_scale, beam_factor1, beam_factor2;
    lc_pparam t0, period, pdot, deltat;
    lc_pparam gravity_dark1, gravity_dark2;
    lc_pparam absorb, slope, quad, cube, third;
    lc_pparam rdisc1, rdisc2, height_disc, beta_disc;
    lc_pparam temp_disc, texp_disc, lin_limb_disc, quad_limb_disc;
    lc_pparam temp_edge, absorb_edge;
    lc_pparam radius_spot, length_spot, height_spot;
    lc_pparam expon_spot, epow_spot, angle_spot, yaw_spot;
    lc_pparam temp_spot, tilt_spot, cfrac_spot;
    lc_pparam stsp11_long, stsp11_lat, stsp11_fwhm, stsp11_tcen;
    lc_pparam stsp12_long, stsp12_lat, stsp12_fwhm, stsp12_tcen;
    lc_pparam stsp13_long, stsp13_lat, stsp13_fwhm, stsp13_tcen;
    lc_pparam stsp21_long, stsp21_lat, stsp21_fwhm, stsp21_tcen;
    lc_pparam stsp22_long, stsp22_lat, stsp22_fwhm, stsp22_tcen;
    lc_pparam uesp_long1, uesp_long2, uesp_lathw, uesp_taper, uesp_temp;
    /* computational parameters */
    double delta_phase, lfudge, llo, lhi, phase1, phase2, wavelength;
    double tperiod, mucrit1, mucrit2;
    int32_t nlat1f, nlat2f, nlat1c, nlat2c, npole, nlatfill, nlngfill;
    int32_t roche1, roche2, eclipse1, eclipse2, glens1, use_radii;
    int32_t gdark_bolom1, gdark_bolom2, limb1, limb2, mirror;
    int32_t add_disc, nrad, opaque, add_spot, nspot, iscale;
    int32_t reserved_;
} lc_model;

/* Lcurve::Datum (lcurve.h:427-445); same field order and padding. */
typedef struct lc_datum {
    double time, expose, flux, ferr, weight;
    int32_t ndiv;
    int32_t reserved_;
} lc_datum;

/* Scalar outputs of light_curve_comp (lcurve.h:669-674).  sfac is in/out:
 * with scale=0, sfac[0] is applied to the curve (light_curve_comp.cpp:562-573). */
typedef struct lc_scalars {
    double sfac[5];
    double wdwarf, chisq, wnok, logg1, logg2, rv1, rv2;
} lc_scalars;

typedef struct lc_eval_opts {
    int32_t scale;            /* fit the scale factor (re_scale)                  */
    int32_t rdata;            /* real data: compute chi**2                         */
    int32_t info;             /* print grid sizes to stderr                        */
    int32_t diagnostics;      /* wdwarf/logg/rv (light_curve_comp.cpp:504-596)     */
    int32_t max_model_points; /* 0: light_curve_comp; >0: light_curve_comp_fast   */
    int32_t precision;        /* LC_PREC_MIXED or LC_PREC_FP64                     */
} lc_eval_opts;

/* Limb darkening: Lcurve::LDC (lcurve.h:202-251). */
typedef struct lc_ldc {
    double c1, c2, c3, c4, mucrit;
    int32_t type; /* LC_LDC_POLY / LC_LDC_CLARET */
    int32_t reserved_;
} lc_ldc;

/* Structure-of-arrays star grid: Lcurve::FlatGrid (lcurve.h:260-272).
 * Faces [0,n0) are never eclipsed, [n0,n1) have one range in1/out1
 * (indexed i-n0), [n1,n) have ragged ranges moff/min_/mout_. */
typedef struct lc_flat_grid {
    size_t n, n0, n1;
    const double *dx, *dy, *dz, *px, *py, *pz, *flux;
    const double *in1, *out1;
    const int32_t *moff;
    const double *min_, *mout_;
} lc_flat_grid;

/* Sub-phases of one sum: Lcurve::PhaseBatch (lcurve.h:279-295), minus the
 * scatter fields (w, idx) that the sum does not read. */
typedef struct lc_phase_batch {
    size_t n;
    const double *ex, *ey, *ez, *phin;
} lc_phase_batch;

/* One star grid as built by set_star_grid (+ set_star_continuum):
 * a vector<Lcurve::Point> (lcurve.h:49-90) in SoA form.  necl[i] is 0 or 1
 * (a star can only be eclipsed by its companion). */
typedef struct lc_star_faces {
    size_t n;
    double *px, *py, *pz, *dx, *dy, *dz;
    float *area, *gravity, *flux;
    int32_t *necl;
    double *ingress, *egress;
} lc_star_faces;

/* Engine timing counters (hipEvent based, on the engine's stream). */
typedef struct lc_profile {
    double flux_ms_total;   /* summed duration of flux-sum kernel launches      */
    double grid_ms_total;   /* face geometry kernels                            */
    double eclipse_ms_total;/* eclipse search kernels                           */
    double cont_ms_total;   /* continuum kernels                                */
    double other_ms_total;  /* phases, reductions, scatter, chi**2              */
    int64_t flux_launches;
    int64_t evaluations;    /* light curves evaluated                            */
    double flux_elements;   /* algorithmic face x sub-phase elements summed      */
    double flux_flops;      /* algorithmic FLOP of those elements                */
    double flux_bytes;      /* algorithmic HBM bytes of the flux kernels         */
    double host_plan_ms_total; /* host planning + table building (wall)          */
    double call_ms_total;   /* wall time inside the evaluation entry points      */
    double eclipse_candidates; /* faces sent to the FP64 eclipse search          */
    double eclipse_hits;    /* of those, faces found eclipsed                    */
    /* per flux-kernel variant v = ldc_type | beaming<<1 | lensing<<2 | mucrit<<3 */
    double flux_var_ms[16];     /* summed launch durations of variant v          */
    double flux_var_flops[16];  /* algorithmic FLOP of variant v                  */
    int64_t flux_var_launches[16];
    double flux_var_evaluated[16]; /* elements actually evaluated (after culling) */
    int64_t grids_computed; /* star grids built by the face + eclipse kernels     */
    int64_t grids_reused;   /* star grids copied (same geometry earlier in the
                               batch, or kept from an earlier call)              */
    /* batched calls whose kernel sequence was replayed as a captured graph:
     * their time is in graph_ms_total only (no per-stage split); the
     * stage and flux-variant fields above count the other calls */
    int64_t graph_launches;
    int64_t graph_captures;
    double graph_ms_total;
    /* flux tile shape of the last batched call (sub-phases per lane, face chunk) */
    int64_t last_rphase;
    int64_t last_chunk;
    /* star grids whose fluxes were copied from a grid with the same faces
     * and continuum inputs instead of recomputed (ContKey reuse,
     * light_curve_comp.cpp:43-52, 205-268) */
    int64_t continua_reused;
    /* of the 2 x eclipse_hits ingress/egress searches, the ones bisected:
     * the others were ended ahead of the bisection because their face is
     * not seen over their first bracket (calls with early stops only) */
    double eclipse_bisected;
} lc_profile;

typedef struct lc_ctx lc_ctx;

/* ------------------------------------------------------------ context */
/* No reference counterpart: the reference keeps thread_local state
 * (light_curve_comp.cpp:161, comp_light_cuda.cu:462-466).  One ctx per host
 * thread; a ctx is not thread-safe. */
lc_status lc_ctx_create(int device, lc_ctx **out);
void lc_ctx_destroy(lc_ctx *ctx);
const char *lc_last_error(const lc_ctx *ctx);
lc_status lc_ctx_get_profile(const lc_ctx *ctx, lc_profile *out);
lc_status lc_ctx_reset_profile(lc_ctx *ctx);
/* Stage reuse across calls (the reference's per-thread grid cache,
 * light_curve_comp.cpp:34-117, 156-207): keep up to max_faces faces of
 * recently built star grids on the device and copy, instead of rebuilding,
 * any grid whose geometry inputs match exactly.  0 turns reuse across calls
 * off (grids equal within one batch are still built once).  Default 4M
 * faces, or 0 when LCURVE_NO_GRID_CACHE is set. */
lc_status lc_ctx_set_stage_cache(lc_ctx *ctx, size_t max_faces);

/* No reference counterpart.  on != 0 (the default; LCURVE_AMD_GRAPHS=0 at
 * creation turns it off): a batched call whose launch shape (sizes, device
 * buffers, tile shape) equals the previous call's replays that call's
 * kernels as one captured hipGraph; the second call of a shape captures it.
 * Uploads, clears and downloads stay ordinary stream operations, so the
 * graph holds kernel launches only.  Results are the same bits either way. */
lc_status lc_ctx_set_graphs(lc_ctx *ctx, int on);
/* No reference counterpart.  Autotuner of the flux kernel's tile shape:
 * runs the given workload with every (sub-phases per lane R in {1, 2, 4})
 * x (face chunk in {64, 128, ..., 8192}; chunks whose partial-sum buffer
 * would pass 4 GiB are skipped) and keeps for this ctx the pair with the
 * shortest time of the flux sums and the merge of their chunk partials
 * (reported through best_rphase / best_chunk when not NULL); a pair must
 * beat the current shape by 2%.  The default is chosen per call: R = 2
 * when a grid has beaming, else 1, and the largest chunk of 2048, 1024,
 * ..., 64 that still makes 2048 flux tasks; when it wins, the shape it
 * resolved to on this workload is kept and reported.  FP64 mode always
 * uses 1024-face chunks. */
lc_status lc_ctx_autotune(lc_ctx *ctx, const lc_model *models, size_t m, const lc_datum *data, size_t n,
                          const lc_eval_opts *opts, int32_t *best_rphase, int32_t *best_chunk);
/* Library version string; usable without a GPU. */
const char *lc_version(void);
/* Message for the last failure of a ctx-free call on this thread. */
const char *lc_thread_error(void);
/* sizeof of a struct of this header by name ("lc_model", ...), 0 if
 * unknown: lets a foreign-language binding verify its layouts. */
size_t lc_sizeof(const char *struct_name);

/* ------------------------------------------------------ model API (host) */
/* Lcurve::Model::Model(const json&) (model.cpp:10-360): parses a JSON
 * document with a "model_parameters" object of string values. */
lc_status lc_model_from_json(const char *json_text, size_t len, lc_model *out);
/* Model::nvary (model.cpp:363) */
int lc_model_nvary(const lc_model *m);
/* Model::set_param (model.cpp:463) */
lc_status lc_model_set_param(lc_model *m, const double *vpar, size_t n);
/* Model::get_param (model.cpp:1160) */
lc_status lc_model_get_param(const lc_model *m, double *vpar, size_t n);
/* Model::get_range (model.cpp:1412) / get_dstep (model.cpp:1539) */
lc_status lc_model_get_range(const lc_model *m, double *v, size_t n);
lc_status lc_model_get_dstep(const lc_model *m, double *v, size_t n);
/* Model::get_limit (model.cpp:1286): lo[i], hi[i] */
lc_status lc_model_get_limit(const lc_model *m, double *lo, double *hi, size_t n);
/* Model::is_not_legal (model.cpp:788): returns 1 if illegal, 0 if legal */
int lc_model_is_not_legal(const lc_model *m, const double *vpar, size_t n);
/* Model::is_not_legal over nvec parameter vectors of n values each (rows n
   apart): out[k] = 1 if vector k is illegal.  No counterpart in the
   reference, which tests one vector per call; for samplers screening many
   proposals at once. */
lc_status lc_model_is_not_legal_batch(const lc_model *m, const double *vpar, size_t nvec, size_t n, int *out);
/* Model::get_name (model.cpp:562): writes NUL-terminated name */
lc_status lc_model_get_name(const lc_model *m, int i, char *buf, size_t buflen);
/* Lcurve::numface via set_star_grid's face count (set_star_grid.cpp:110-162) */
lc_status lc_star_grid_size(const lc_model *m, int star, int fine, size_t *nface);

/* ------------------------------------------------------- light curves */
/* Lcurve::light_curve_comp (lcurve.h:669-674; light_curve_comp.cpp:119-599),
 * or light_curve_comp_fast (lcurve.h:474-483) when opts->max_model_points > 0.
 * calc has n entries; io->sfac is read when opts->scale == 0. */
lc_status lc_light_curve_comp(lc_ctx *ctx, const lc_model *model,
                              const lc_datum *data, size_t n,
                              const lc_eval_opts *opts, double *calc,
                              lc_scalars *io);

/* Batched form: m parameter vectors against one data set, one launch
 * sequence.  calc is m x n (row per model), io has m entries, status[i] is
 * the lc_status of instance i.  Returns LC_OK if the batch ran (instances
 * may still fail individually), or the error that stopped it. */
lc_status lc_light_curve_comp_batch(lc_ctx *ctx, const lc_model *models,
                                    size_t m, const lc_datum *data, size_t n,
                                    const lc_eval_opts *opts, double *calc,
                                    lc_scalars *io, int32_t *status);

/* No counterpart in the reference (its calc is a host vector).  The batched
 * form with the curves left in GPU memory, for callers that use them there
 * (a sampler or a training set on the device): dev_calc is device memory
 * of ctx's GPU, m x n doubles; rows of failed instances are left as they
 * were.  Scalars and statuses come back in host memory as above, and the
 * curves are written when the call returns.  light_curve_comp_fast sampling
 * (max_model_points below n) is refused (its interpolation runs on the
 * host). */
lc_status lc_light_curve_comp_batch_dev(lc_ctx *ctx, const lc_model *models,
                                        size_t m, const lc_datum *data, size_t n,
                                        const lc_eval_opts *opts, double *dev_calc,
                                        lc_scalars *io, int32_t *status);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): flat_sum_star1_multi (star 0) / flat_sum_star2_multi
 * (star 1) as lc_flat_sum_star{1,2}_multi compute them, bit for bit, but on
 * DEVICE arrays: every pointer in *g and *pb, and out (pb->n doubles), are
 * device addresses on ctx's GPU, written and synchronised by the caller
 * before the call (g and pb themselves are host structs).  The work memory
 * (flux records, task table, partial sums) is one static device array of
 * 64 MiB per GPU (needs about 64 B per face + 8 B per 1024 faces per
 * sub-phase); larger inputs return LC_E_ARG with the sizes.  The call
 * allocates nothing, runs on ctx's stream and returns when out is written.
 * Grids with ragged faces (n1 < n) return LC_E_UNSUPPORTED here; the host-
 * array entry points take them.  Calls on one GPU share that array: calls that overlap on one GPU (other
 * threads, other contexts) race. */
lc_status lc_flat_sum_multi_noalloc(lc_ctx *ctx, int star, const lc_flat_grid *g, const lc_phase_batch *pb,
                                    const lc_ldc *ldc, double beam, double spin, double VFAC, double XCOFM,
                                    int glens1, double rlens1, int precision, double *out);

/* ------------------------------------------------ stage-level entry points */
/* Lcurve::flat_sum_star1_multi (lcurve.h:299-301; comp_light_flat.cpp:413) */
lc_status lc_flat_sum_star1_multi(lc_ctx *ctx, const lc_flat_grid *g,
                                  const lc_phase_batch *pb, const lc_ldc *ldc,
                                  double beam, double spin, double VFAC,
                                  double XCOFM, int precision, double *out);
/* Lcurve::flat_sum_star2_multi (lcurve.h:304-307; comp_light_flat.cpp:434) */
lc_status lc_flat_sum_star2_multi(lc_ctx *ctx, const lc_flat_grid *g,
                                  const lc_phase_batch *pb, const lc_ldc *ldc,
                                  double beam, double spin, double VFAC,
                                  double XCOFM, int glens1, double rlens1,
                                  int precision, double *out);
/* Lcurve::set_star_grid (lcurve.h:528; set_star_grid.cpp:49-523): fills out
 * (arrays of lc_star_grid_size() entries, caller-owned host memory). */
lc_status lc_set_star_grid(lc_ctx *ctx, const lc_model *m, int star, int fine,
                           lc_star_faces *out);
/* Lcurve::set_star_continuum (lcurve.h:550-552; set_star_continuum.cpp:26-297):
 * sets star1->flux and star2->flux from the other fields. */
lc_status lc_set_star_continuum(lc_ctx *ctx, const lc_model *m,
                                lc_star_faces *star1, lc_star_faces *star2);

/* ------------------------------------------------ learned surrogate (MLP) */
/* No reference counterpart: a BF16 multilayer perceptron (SiLU hidden
 * layers, linear output) trained on light curves from this engine, for the
 * delayed-acceptance screen of an MCMC sampler (the sampler's cheap first
 * stage; every accepted step is judged by lc_light_curve_comp_batch).
 * widths = {inputs, hidden..., outputs}; parameters are laid out layer by
 * layer as W_l (fan_in x fan_out, row-major) followed by b_l.  Host arrays
 * are row-major float32; a lc_mlp is not thread-safe. */
typedef struct lc_mlp lc_mlp;
lc_status lc_mlp_create(int device, const int32_t *widths, int nwidth, int max_batch, lc_mlp **out);
void lc_mlp_destroy(lc_mlp *mlp);
const char *lc_mlp_last_error(const lc_mlp *mlp);
size_t lc_mlp_num_params(const lc_mlp *mlp);
lc_status lc_mlp_set_params(lc_mlp *mlp, const float *params, size_t n);  /* also resets Adam */
lc_status lc_mlp_get_params(lc_mlp *mlp, float *params, size_t n);
/* y (batch x outputs) = network(x (batch x inputs)) */
lc_status lc_mlp_forward(lc_mlp *mlp, const float *x, size_t batch, float *y);
/* gradient of the mean squared error over batch <= max_batch rows */
lc_status lc_mlp_grad(lc_mlp *mlp, const float *x, const float *target, size_t batch, float *grad, double *loss);
/* device-resident training set: allocate nrows, then load row ranges */
lc_status lc_mlp_set_dataset(lc_mlp *mlp, size_t nrows);
lc_status lc_mlp_load_dataset(lc_mlp *mlp, size_t row0, const float *x, const float *target, size_t count);
/* steps of Adam (0.9, 0.999, 1e-8) on minibatches drawn from the data set */
lc_status lc_mlp_train(lc_mlp *mlp, int steps, int batch, float lr, uint64_t seed, double *last_loss);
/* mean squared error of the network over data-set rows [row0, row0+count) */
lc_status lc_mlp_eval_dataset(lc_mlp *mlp, size_t row0, size_t count, double *mse);
/* Optional FP32 linear skip S (inputs x outputs, row-major) added to the
 * output of lc_mlp_forward / lc_mlp_chisq: y = network(x) + x S.  Training
 * (lc_mlp_grad, lc_mlp_train, lc_mlp_eval_dataset) leaves it out, so the
 * network learns what the skip does not explain.  S = NULL removes it. */
lc_status lc_mlp_set_skip(lc_mlp *mlp, const float *S, size_t n);
/* Likelihood of predicted curves (outputs = standardised curve at phases
 * k / outputs: v = y * scale + mu[k] is log F when log_output, else F)
 * against data: weight = w / ferr^2, zero for ignored points.  lc_mlp_chisq
 * then gives, for each of m input rows and its t0, chi**2 with the optimal
 * scale factor, the model flux at a data time being the linear
 * interpolation of F at phase (t - t0) / period. */
lc_status lc_mlp_set_likelihood(lc_mlp *mlp, const double *time, const double *flux, const double *weight, size_t n,
                                const float *mu, float scale, double period, int log_output);
lc_status lc_mlp_chisq(lc_mlp *mlp, const float *x, const double *t0, size_t m, double *chisq);
/* GEMM FLOP issued so far and wall time spent in lc_mlp_train */
lc_status lc_mlp_stats(const lc_mlp *mlp, double *gemm_flop, double *train_ms);

#ifdef __cplusplus
}
#endif
#endif /* LCURVE_AMD_H */
