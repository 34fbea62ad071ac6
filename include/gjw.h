/*
 * gjw.h -- C ABI of libgjw: MI355X (gfx950) radar FDTD forward solve and
 * adjoint gradient for gerjoii's src/w path.
 *
 * The entry points replace the numerics behind these MATLAB functions
 * (reference: diegozain/gerjoii, paths relative to its root):
 *   [finite_,gerjoii_] = w_fwd(geome_,parame_,finite_,gerjoii_)   src/w/w_fwd.m:1
 *   [finite_,gerjoii_] = w_fwd_(geome_,parame_,finite_,gerjoii_,is) src/w/w_fwd_.m:1
 *   gerjoii_ = w_solve(geome_,parame_,finite_,gerjoii_)             src/w/w_solve.m:1
 *   [g_w,a_max] = w_grad(u_w,geome_,finite_,gerjoii_,parame_)      src/w/w_grad.m:1
 *   [e,d_e,d_oe] = w_bozdag(d_o,d,obj_fnc)                         src/w/w_bozdag.m:1
 *   Ew = w_E(d_o,d,Nw,obj_fnc)                                     src/w/w_E.m:1
 *   v = differentiate_cube(v,dt)                  src/shared/w-dc/differentiate_cube.m:1
 * A MEX gateway (or ctypes / torch) fills a gjw_grid from geome_/parame_/
 * finite_, passes gerjoii_.w.{s,M,wvlet_} as the shot geometry, and reads
 * traces and gradients back.
 *
 * Conventions
 *  - Arrays are MATLAB column-major (depth fastest).  Model arrays are m x n
 *    (rows = depth, parame_.w.epsilon / .sigma); traces of one shot are
 *    nt x nr (gerjoii_.w.d_2d); shots are concatenated in shot order.
 *  - Indices are MATLAB 1-based: a source is [iz, ix] on the padded Ez grid
 *    (gerjoii_.w.s, src/w/w_load.m:31-32); a receiver is the linear index Mw
 *    into the (nx+1) x (ny+1) Ez grid (src/w/w_M.m:20).
 *  - Every pointer argument may be host memory or device memory of the
 *    context's GPU; the library tells them apart itself.
 *  - Every function returns GJW_OK (0) or a negative code; gjw_last_error()
 *    returns a per-thread message describing the last failure.
 *  - One context per process and GPU; calls on one context are serialised by
 *    the caller (MATLAB calls MEX on one thread).
 */
#ifndef GJW_H
#define GJW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GJW_OK 0
#define GJW_ERR_ARG -1     /* invalid argument (size, index, option)     */
#define GJW_ERR_HIP -2     /* HIP runtime error                          */
#define GJW_ERR_STATE -3   /* call out of order (e.g. gradient before forward) */
#define GJW_ERR_MEM -4     /* device memory exhausted                    */

typedef struct gjw_ctx gjw_ctx;
typedef struct gjw_survey gjw_survey;
typedef struct gjw_batch gjw_batch;

/* Geometry and physics of one padded grid.  Field names follow the MATLAB
 * structs (src/w/w_geom.m:78-172, src/joint/wdc_geom.m:43-125). */
typedef struct {
  int32_t m;         /* geome_.m: model rows (depth, no PML, no air)      */
  int32_t n;         /* geome_.n: model columns                           */
  int32_t pml;       /* finite_.w.pml_w == finite_.w.pml_d                */
  int32_t air;       /* parame_.w.air: air rows above the model           */
  int32_t nt;        /* geome_.w.nt                                       */
  int32_t pml_order; /* finite_.w.pml_order (2)                           */
  double dx;         /* geome_.dx == geome_.dy  [m]                       */
  double dt;         /* geome_.w.dt [s]                                   */
  double eps0;       /* parame_.w.eps_0                                   */
  double mu0;        /* parame_.w.mu_0                                    */
  double R0;         /* finite_.w.R_0 (1e-8)                              */
} gjw_grid;

/* ---- context ---------------------------------------------------------- */
int gjw_create(int32_t device, gjw_ctx **out);
int gjw_destroy(gjw_ctx *ctx);
const char *gjw_last_error(void);
/* Library version string. */
const char *gjw_version(void);

/* ---- survey: one grid + its model + its shots -------------------------- */
int gjw_survey_create(gjw_ctx *ctx, const gjw_grid *grid, gjw_survey **out);
int gjw_survey_destroy(gjw_survey *s);

/* Model (parame_.w.epsilon, parame_.w.sigma), m x n doubles.  Pads with air
 * and PML as src/w/w_exp2pml.m:17-37 and builds the coefficients of
 * src/w/w_cpml.m:44-185 and src/w/w_coeff.m:28-55 in FP64, stored as FP32. */
int gjw_set_model(gjw_survey *s, const double *eps_r, const double *sigma);

/* Shot geometry.  src: 2 x ns int32 ([iz; ix] per shot, 1-based);
 * wavelets: nt x ns float (gerjoii_.w.wvlet_ per shot);
 * rec_ptr: ns+1 int64 offsets into rec_lin; rec_lin: 1-based Mw per shot. */
int gjw_set_shots(gjw_survey *s, int32_t ns, const int32_t *src, const float *wavelets,
                  const int64_t *rec_ptr, const int32_t *rec_lin);

/* Options (gjw_set_option keys). */
#define GJW_OPT_PATH 1       /* GJW_PATH_BATCHED (default) | GJW_PATH_PLAIN        */
#define GJW_OPT_CKPT_EVERY 2 /* checkpoint interval in steps; 0 = about 2 sqrt(nt),
                                in whole launches of the forward and adjoint depths */
#define GJW_OPT_TILE_L 3     /* columns marched per wavefront; 0 = autotuned        */
#define GJW_OPT_SNAPSHOT 4   /* GJW_SNAP_F32 (default) | GJW_SNAP_BF16 | GJW_SNAP_F16 */
#define GJW_OPT_KTIMING 5    /* n > 0: time one launch of each step kernel per n
                                time steps it covers, with HIP events */
#define GJW_OPT_FUSE 6       /* time steps fused per forward launch on the batched path:
                                0 (default: 2, 4 or 8, picked by the autotuner), 1, 2, 4 or
                                8; the adjoint sweep fuses as many, at most 4 */
#define GJW_OPT_MISFIT 8     /* misfit of gjw_fwi_gradient: GJW_MISFIT_* (batched path) */
#define GJW_OPT_ADJ_FUSE 9   /* time steps fused per adjoint launch on the batched path:
                                0 (default: six), 4, 6, or 8 (with an eight-step forward) */
#define GJW_OPT_GRAPH 7      /* 1: gjw_fwi_gradient (batched path) replays its whole launch
                                sequence as one captured hipGraph (captured on the second call
                                and again whenever a buffer or option it uses changes) */
#define GJW_OPT_SHOT_GRADS 10 /* 1: gjw_gradient / gjw_fwi_gradient / gjw_batch_fwi_gradient
                                return one gradient per shot, ns*m*n values each (shot-major),
                                as w_update_e.m:98 keeps g_e per source before its
                                post-processing; 0 (default): the sum over shots */
#define GJW_OPT_ADJ_TILE 11  /* columns per wavefront of the 4-, 6- and 8-step adjoint on the
                                batched path; 0 (default): tuned in the first gradient sweep
                                (the widest of the candidates within 1.5% of the fastest).
                                A fixed width makes gradients repeat bit for bit across runs */
#define GJW_OPT_FRAMES 12    /* element type of the Ez frames the recompute sweep hands to the
                                adjoint on the batched path: GJW_FRAMES_F32 (default) or
                                GJW_FRAMES_BF16 (half the frame traffic; the adjoint then sums
                                its frame correlation by parts so that no difference of rounded
                                frames enters the gradient).  BF16 frames apply to single-survey
                                gradients with FP32 snapshots, an even nt and checkpoint
                                interval and at least two fused forward steps; other gradients
                                keep FP32 frames (gjw_query GJW_Q_FRAMES tells which was used) */
#define GJW_FRAMES_F32 0
#define GJW_FRAMES_BF16 1
#define GJW_OPT_ROW_PAIRS 13 /* 1 (default): the 8-step forward of the batched path marches the
                                rows between the PML, receiver and source rows with two rows
                                per wavefront lane (112 owned rows per wave instead of 48), in
                                strips of interior columns, except with FP32 frames; 0: one
                                row per lane everywhere.  The update is the same, row by row */
#define GJW_PATH_BATCHED 0   /* all shots per launch, checkpoint + recompute        */
#define GJW_PATH_PLAIN 1     /* one shot at a time, full wavefield stored (as w_solve.m) */
#define GJW_SNAP_F32 0
#define GJW_SNAP_BF16 1
#define GJW_SNAP_F16 2
int gjw_set_option(gjw_survey *s, int32_t key, int64_t value);

/* Forward solve of every shot (src/w/w_solve.m:80-118).
 * traces: sum_s nt*nr_s floats (gerjoii_.w.d_2d per shot, concatenated);
 * jz: nt x ns floats or NULL (gerjoii_.w.Jy, the field at the source).
 * keep != 0 keeps the forward state (checkpoints, or the stored wavefield on
 * the plain path) for a following gjw_gradient. */
int gjw_forward(gjw_survey *s, float *traces, float *jz, int32_t keep);

/* Misfit and adjoint source of the last forward (src/w/w_bozdag.m:13-55,
 * src/w/w_E.m:7-16, instant_.m:14-20) against observed traces d_obs (same
 * layout as traces).  kind: the reference's obj_FNC
 *   GJW_MISFIT_L2        'L2'        e = d - d_o,                J = 0.5|d - d_o|^2
 *   GJW_MISFIT_ENV_L2    'env_L2'    envelopes,                  J = 0.5|A(d) - A(d_o)|^2
 *   GJW_MISFIT_PHASE_L2  'phase_L2'  instantaneous phases,       J = 0.5|phi(d) - phi(d_o)|^2
 *   GJW_MISFIT_ENV_L22   'env_L22'   squared envelopes,          J = 0.5|A(d)^2 - A(d_o)^2|^2
 *   GJW_MISFIT_ENV_LOG   'env_LOG'   log envelopes,              J = 0.5|log A(d) - log A(d_o)|^2
 * (Hilbert transforms per trace by FFT of length nt in FP32, as MATLAB's
 * hilbert).  e_out (optional) receives the adjoint source; J (optional) the
 * misfit in FP64.  The adjoint source stays on the device for gjw_gradient. */
#define GJW_MISFIT_L2 0
#define GJW_MISFIT_ENV_L2 1
#define GJW_MISFIT_PHASE_L2 2
#define GJW_MISFIT_ENV_L22 3
#define GJW_MISFIT_ENV_LOG 4
#define GJW_MISFIT_FREQSTP 5 /* line-search objectives only: L2 after an extra band filter */
int gjw_misfit(gjw_survey *s, const float *d_obs, int32_t kind, float *e_out, double *J);

/* [e, ~, ~] = w_bozdag(d_o, d, obj_fnc) and J = w_E(...) (src/w/w_bozdag.m:1,
 * src/w/w_E.m:1) for synthetic traces d given by the caller instead of the
 * last forward (same layout; they replace the survey's traces). */
int gjw_bozdag(gjw_survey *s, const float *d_obs, const float *d, int32_t kind, float *e_out,
               double *J);

/* [e_, d_e, d_oe] = w_bozdag(d_o, d, obj_fnc) with all three outputs, and
 * J = w_E(d_o, d, I, obj_fnc) on d_e / d_oe (src/w/w_bozdag.m:1, src/w/w_E.m:1,
 * src/shared/w-dc/instant_.m:1), for ntr traces of nt samples (trace-major,
 * MATLAB's nt x nr) with no survey.  d_e / d_oe are d and d_o for L2, the
 * envelopes (env_*) or instantaneous phases (phase_L2) otherwise.  Any output
 * may be NULL. */
int gjw_w_bozdag(gjw_ctx *ctx, const float *d_o, const float *d, int32_t ntr, int32_t nt,
                 int32_t kind, float *e, float *d_e, float *d_oe, double *J);

/* Gradient of the misfit w.r.t. the model (m x n doubles, FP64 accumulation
 * over time and shots).
 * mode GJW_GRAD_EXACT: the exact discrete adjoint of the forward solve with
 *   adjoint source adj (traces layout; NULL = the one left by gjw_misfit).
 *   Strip (PML) coefficients are held fixed, i.e. the gradient covers every
 *   model cell whose Ez is updated by the inner Yee formula.
 * mode GJW_GRAD_REF: the reference's w_grad.m semantics (negated sigma,
 *   residual written into Ez at the receivers, correlation with u for sigma
 *   and with d_t u for eps); plain path only; a_max (m x n float, optional)
 *   receives |a| of the last step (w_grad.m:110-112).
 * g_w (optional, nt x ns doubles, EXACT only): dJ/d(wavelet). */
#define GJW_GRAD_EXACT 0
#define GJW_GRAD_REF 1
int gjw_gradient(gjw_survey *s, const float *adj, int32_t mode, double *g_eps, double *g_sig,
                 double *g_w, float *a_max);

/* Forward + L2 misfit + exact gradient in one call, everything on the device
 * (what w_update_e.m / w_update_s.m need per iteration, for all shots). */
int gjw_fwi_gradient(gjw_survey *s, const float *d_obs, double *g_eps, double *g_sig, double *J);

/* Band filter of the adjoint source that gjw_fwi_gradient applies before the
 * adjoint sweep (src/w/w_update_e.m:69-70: filt_gauss(e, dt, f_, f__, 8)).
 * f_high <= f_low turns it off (the default). */
int gjw_set_adj_filter(gjw_survey *s, double f_low, double f_high, int32_t steep);

/* Band filter of ntr traces of nt samples in place, trace-major (MATLAB's
 * nt x nr array): Tukey taper (r = 0.1), zero padding to the next power of
 * two, super-Gaussian band exp(-((f-mid)^steep)/((df/2) width^(steep-1))) on
 * the signed frequencies, real part of the inverse FFT
 * (src/shared/w-dc/filt_gauss.m:1-47; dt in s, frequencies in Hz). */
int gjw_filt_gauss(gjw_ctx *ctx, float *d, int32_t ntr, int32_t nt, double dt, double f_low,
                   double f_high, int32_t steep);

/* Smooth mute of a linear arrival (src/w/w_lmute.m:13-41): d and d_o (ntr
 * traces of nt samples, trace-major) are multiplied in place by
 * 1/(1+exp(5 (t_mute + (rx - rx[0])/v_mute - t))) and e = d - d_o (e may be
 * NULL).  t: nt times [s]; rx: ntr receiver x [m]; v_mute [m/ns], t_mute [ns]
 * as in gerjoii_.w.v_mute / t_mute. */
int gjw_lmute(gjw_ctx *ctx, float *d, float *d_o, float *e, int32_t ntr, int32_t nt,
              const double *t, const double *rx, double v_mute, double t_mute);

/* Time derivative of an m x n x nt cube in place, second order
 * (src/shared/w-dc/differentiate_cube.m:23-51). */
int gjw_differentiate_cube(gjw_ctx *ctx, float *v, int32_t nz, int32_t nx, int32_t nt, double dt);

/* Several surveys (own grid, model, shots, nt and misfit each; at most 8,
 * one context) whose FWI gradients run in shared launches: every time-step
 * launch covers the tiles of all surveys still active.  Replaces running
 * w_update_*.m once per survey; no counterpart in the reference. */
int gjw_batch_create(gjw_ctx *ctx, int32_t n, gjw_survey *const *surveys, gjw_batch **out);
int gjw_batch_destroy(gjw_batch *b);
/* gjw_fwi_gradient of every survey of the batch: d_obs[k], g_eps[k], g_sig[k]
 * as for survey k (g_* entries may be NULL), J[k] its misfit.  The surveys
 * share one checkpoint interval (the largest GJW_OPT_CKPT_EVERY, else
 * about 2 sqrt(max nt)). */
int gjw_batch_fwi_gradient(gjw_batch *b, const float *const *d_obs, double *const *g_eps,
                           double *const *g_sig, double *J);
/* Objective of every survey's current model, as w_objs_e.m:22-63 evaluates
 * it for one line-search step (src/w/w_stepparabola_e.m:20-31 calls it per
 * step): forward without wavefield storage for all surveys in shared
 * launches, the survey's band filter (gjw_set_adj_filter) on the synthetic
 * and the observed traces, J[k] = 0.5 |d - d_o|^2 / |d_o|^2 (L2 misfit).
 * Line searches batch their perturbed models as surveys of one batch. */
int gjw_batch_objective(gjw_batch *b, const float *const *d_obs, double *J);

/* The line-search objective of w_objs_e.m:22-63 in full, for every survey's
 * current model (forward in shared launches, no wavefield; the survey's band
 * filter on both trace sets, as w_objs_e.m:34-37):
 * - nobj = 0 (no gerjoii_.w.obj_FNCs): J[k] = w_E(d_o, d, I, e_kind) / |d_o|^2
 *   (w_objs_e.m:61-62; w_objs_s.m:37-38 for a conductivity step);
 * - nobj > 0 (gerjoii_.w.obj_FNCs, w_objs_e.m:41-60): for each objective q in
 *   order, kind GJW_MISFIT_FREQSTP first band-filters both trace sets again
 *   with (fstp_lo, fstp_hi) (they stay filtered for the objectives after it);
 *   [~, d', d_o'] = w_bozdag(d_o, d, kinds[q]) (the traces for L2 / FREQstp,
 *   envelopes or phases otherwise); J[k] += weights[q] * w_E(d_o', d', I,
 *   e_kind) / |d_o'|^2 (gerjoii_.w.g_e_weights).
 * e_kind is gerjoii_.w.obj_FNC, whose w_E transform (squares for env_L22,
 * logarithms for env_LOG) applies to every term, as in the reference. */
int gjw_batch_objective_multi(gjw_batch *b, const float *const *d_obs, int32_t nobj,
                              const int32_t *kinds, const double *weights, int32_t e_kind,
                              double fstp_lo, double fstp_hi, double *J);

/* Forward of every survey's current model in shared launches, no wavefield
 * or checkpoints: traces[k] (survey k's traces layout, host or device)
 * receives its traces.  Replaces the forward runs w_fwd_(..., 0) that
 * w_pica_e_.m:24 and w_pica_s_.m:24 make for one trial model each. */
int gjw_batch_forward(gjw_batch *b, float *const *traces);

/* Device time and cell-updates of the last gjw_batch_fwi_gradient or
 * gjw_batch_objective. */
int gjw_batch_last_timing(gjw_batch *b, double *ms, double *cells);

/* Device-time of the last gjw_forward / gjw_gradient / gjw_fwi_gradient
 * call and the number of cell-updates it executed. */
int gjw_last_timing(gjw_survey *s, double *ms, double *cell_updates);

/* Gradient post-processing, on `count` FP64 images of m x n (column-major,
 * image c at offset c*m*n; host or device memory).  The reference runs it per
 * source inside its parfor (src/w/w_update_e.m:108-145, w_update_es.m:105-180);
 * here every source's image goes through one launch sequence. */

/* Source taper (src/shared/w-dc/gaussi.m:10-16, applied at
 * src/w/w_update_e.m:112-118): image c is multiplied in place by
 * normali(1 - exp(-(((x - X[s_x-1])/L)^2 + ((z - Y[s_z-1])/L)^2))), L = 2*lo[c].
 * X: n column coordinates (geome_.X), Y: m depth coordinates (geome_.Y), s:
 * (iz, ix) per image, 1-based model indices, lo: gaussi's lo argument (w_update_e
 * passes twice the wavelength at the source).  X, Y, s and lo are host arrays. */
int gjw_gaussi(gjw_ctx *ctx, double *g, int32_t m, int32_t n, int32_t count, const double *X,
               const double *Y, const int32_t *s, const double *lo);

/* k-space Gaussian filter (src/shared/w-dc/image_gaussian.m:21-61): zero
 * padding to 2^nextpow2 in both directions, fft2, the cake exp(-r^2)
 * (GJW_LOW_PASS) or 1 - exp(-r^2) (GJW_HI_PASS) normalised to its maximum,
 * r^2 = ((kx - n_/2)/a)^2 + ((kz - m_/2)/b)^2 with a = n_ ax/2 sqrt(2),
 * b = m_ az/2 sqrt(2), fftshifted; real part of ifft2, trimmed.  in may equal
 * out; m, n >= 2. */
#define GJW_LOW_PASS 0
#define GJW_HI_PASS 1
int gjw_image_gaussian(gjw_ctx *ctx, const double *in, double *out, int32_t m, int32_t n,
                       int32_t count, double ax, double az, int32_t type);

/* normali (src/shared/w-dc/normali.m:1-7) of each image in place: a / max|a|
 * (NaN ignored by the maximum, as MATLAB's max; an all-zero image is left
 * alone).  nan_guard != 0 then zeroes every image holding a NaN
 * (src/w/w_update_e.m:133-137); a shape mask (m x n, or NULL) then multiplies
 * every image (src/w/w_update_e.m:141-143). */
int gjw_normali(gjw_ctx *ctx, double *a, int32_t m, int32_t n, int32_t count, int32_t nan_guard,
                const double *shape);

/* gjw_normali for images and mask already in device memory, allocating
 * nothing: no counterpart in the reference; for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Its partial maxima live in a static device scratch array of
 * 768 KiB per GPU (count * P <= 65536 partials, P = min(256,
 * ceil(m n / 4096)): up to 256 images of a million cells or more per call;
 * larger calls fail with GJW_ERR_ARG), resolved once per device by
 * gjw_create.  Host pointers fail with GJW_ERR_ARG.  Calls that overlap on
 * one GPU share the scratch and race. */
int gjw_normali_noalloc(gjw_ctx *ctx, double *a, int32_t m, int32_t n, int32_t count,
                        int32_t nan_guard, const double *shape);

/* Shot reduction and update (src/w/w_update_e.m:177, src/w/w_update_e_.m:38-41,
 * src/w/w_update_s_.m:38, src/w/w_image_e.m:50-51, src/w/w_image_s.m:38-39):
 * dparam = sum_c(-steps[c] * g_c) / count; with neg_rule != 0 all zero when
 * sum(steps) < 0 (the permittivity reduction's rule); with a model (may be
 * NULL), model .*= exp(model .* dparam).  steps: count host doubles. */
int gjw_stack_update(gjw_ctx *ctx, const double *g, int32_t m, int32_t n, int32_t count,
                     const double *steps, int32_t neg_rule, double *dparam, double *model);

/* Average device duration (ms) and count of the sampled step kernels, and the
 * number of launches of that kernel, while GJW_OPT_KTIMING is on.  which: 0 = one-step forward, 1 = one-step adjoint,
 * 2 = two-step forward, 3 = two-step adjoint, 4 = four-step forward,
 * 5 = four-step adjoint, 6 = eight-step forward (forward sweep), 7 = eight-step
 * adjoint, 8 = six-step adjoint, 9 = eight-step forward that also stores the
 * Ez frames (recompute sweep).  Has no
 * counterpart in the reference (measurement only). */
int gjw_kernel_stats(gjw_survey *s, int32_t which, double *avg_ms, int64_t *count,
                     int64_t *launches, int32_t reset);

/* Columns per wavefront chosen by the autotuner for the step kernel `which`
 * (numbered as in gjw_kernel_stats; 0 before the first forward, and for the
 * adjoint kernels before the first gradient).  Has no counterpart in the
 * reference (measurement only). */
int gjw_tile(gjw_survey *s, int32_t which, int32_t *L);

/* Effective settings of a survey: GJW_Q_CKPT_INTERVAL, the checkpoint interval
 * the batched path uses (the option, else about 2 sqrt(nt) in whole launches,
 * else the least-memory interval; 0 on the plain path); GJW_Q_FWD_DEPTH and
 * GJW_Q_ADJ_DEPTH, the forward and adjoint steps per launch (0 before the
 * first forward has tuned them); GJW_Q_FRAMES, the element type (GJW_FRAMES_*)
 * of the frames of the last batched gradient.  Has no counterpart in the
 * reference (the MATLAB code stores the whole wavefield and steps one at a
 * time). */
#define GJW_Q_CKPT_INTERVAL 1
#define GJW_Q_FWD_DEPTH 2
#define GJW_Q_ADJ_DEPTH 3
#define GJW_Q_FRAMES 4 /* GJW_FRAMES_* of the last batched gradient's frames */
#define GJW_Q_ROW_PAIRS 5 /* two-row tiles per strip of the last 8-step forward launch
                             (GJW_OPT_ROW_PAIRS; 0: none) */
int gjw_query(gjw_survey *s, int32_t what, int64_t *value);

#ifdef __cplusplus
}
#endif
#endif /* GJW_H */
