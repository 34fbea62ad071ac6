/*
 * gjw_mex.c -- MATLAB MEX gateway over libgjw's C ABI (include/gjw.h).
 *
 * gerjoii's src/w entry points keep their names and signatures and become
 * thin .m wrappers around this one MEX function (INTEGRATION.md shows them):
 *
 *   [d_2d, Jy]          = gjw_mex('fwd', geome_, parame_, finite_, gerjoii_)
 *                         w_exp2pml + w_cpml + w_coeff + w_solve (src/w/w_fwd.m:17-25)
 *   [g, a_max]          = gjw_mex('grad', mode, e_2d, dtu)           (src/w/w_grad.m:1)
 *   [e, d_e, d_oe, J]   = gjw_mex('bozdag', d_o, d, obj_fnc)         (src/w/w_bozdag.m:1, w_E.m)
 *   d_filt              = gjw_mex('filt_gauss', d, dt, f_low, f_high, steep)
 *                                                            (src/shared/w-dc/filt_gauss.m:1)
 *   [d, d_o, e]         = gjw_mex('lmute', d, d_o, t, rx, v_mute, t_mute) (src/w/w_lmute.m:1)
 *   g                   = gjw_mex('gaussi', g, x, z, s, lo)     (src/shared/w-dc/gaussi.m:1)
 *   img                 = gjw_mex('image_gaussian', img, a, b, type)
 *                                                       (src/shared/w-dc/image_gaussian.m:1)
 *   a                   = gjw_mex('normali', a)                 (src/shared/w-dc/normali.m:1)
 *
 * Conventions (as MEX requires): MATLAB owns the inputs; outputs are created
 * with mxCreate*; every library failure becomes mexErrMsgIdAndTxt with
 * gjw_last_error()'s text.  Traces are nt x nr column-major, which is the
 * library's trace-major layout, so they pass without a transpose; models are
 * m x n column-major doubles (depth fastest), which the library takes as is.
 * 1-based indices (s, M) pass unchanged: the library converts them.
 *
 * State: one context per MATLAB process (parfor workers are processes, each
 * loads the MEX and gets its own), created on the first call that needs the
 * GPU and freed by mexAtExit; one survey per padded-grid shape, on the plain
 * path (the wavefield of the last 'fwd' is kept on the GPU for 'grad', as
 * w_solve.m keeps u_2d).
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "gjw.h"
#include "mex.h"

static gjw_ctx *ctx = NULL;
static gjw_survey *sv = NULL;
static gjw_grid last;
static int have_fwd = 0;

static void cleanup(void) {
  gjw_survey_destroy(sv);
  gjw_destroy(ctx);
  sv = NULL;
  ctx = NULL;
  have_fwd = 0;
}

static void chk(int rc) {
  if (rc) mexErrMsgIdAndTxt("gjw:call", "%s", gjw_last_error());
}

static void need(int nrhs, int want, const char *op) {
  if (nrhs != want) mexErrMsgIdAndTxt("gjw:nrhs", "gjw_mex('%s', ...) takes %d inputs, got %d", op,
                                      want - 1, nrhs - 1);
}

static void ensure_ctx(void) {
  if (!ctx) {
    chk(gjw_create(0, &ctx));
    mexAtExit(cleanup);
  }
}

static const mxArray *field(const mxArray *s, const char *a, const char *b) {
  const mxArray *x = mxIsStruct(s) ? mxGetField(s, 0, a) : NULL;
  if (x && b) x = mxIsStruct(x) ? mxGetField(x, 0, b) : NULL;
  if (!x) mexErrMsgIdAndTxt("gjw:field", "missing field %s%s%s", a, b ? "." : "", b ? b : "");
  return x;
}

static double fscalar(const mxArray *s, const char *a, const char *b) {
  return mxGetScalar(field(s, a, b));
}

/* a numeric array of any of the classes gerjoii passes for indices and
 * samples, converted element by element */
static double num_at(const mxArray *x, size_t k) {
  const void *p = mxGetData(x);
  switch (mxGetClassID(x)) {
    case mxDOUBLE_CLASS: return ((const double *)p)[k];
    case mxSINGLE_CLASS: return ((const float *)p)[k];
    case mxINT32_CLASS: return ((const int32_t *)p)[k];
    case mxUINT32_CLASS: return ((const uint32_t *)p)[k];
    default: mexErrMsgIdAndTxt("gjw:class", "numeric input expected (double, single, int32, uint32)");
  }
  return 0;
}

static float *to_f32(const mxArray *x) {
  const size_t n = mxGetNumberOfElements(x);
  float *o = (float *)mxMalloc((n ? n : 1) * sizeof(float));
  for (size_t k = 0; k < n; ++k) o[k] = (float)num_at(x, k);
  return o;
}

static const double *doubles(const mxArray *x, const char *what) {
  if (!mxIsDouble(x)) mexErrMsgIdAndTxt("gjw:class", "%s must be double", what);
  return mxGetDoubles(x);
}

static void get_str(const mxArray *x, char *buf, size_t n, const char *what) {
  if (!mxIsChar(x) || mxGetString(x, buf, n))
    mexErrMsgIdAndTxt("gjw:arg", "%s must be a character array shorter than %d", what, (int)n);
}

/* 'fwd': geome_, parame_, finite_, gerjoii_ -> d_2d (nt x nr single), Jy (nt x 1 single) */
static void op_fwd(int nlhs, mxArray *plhs[], const mxArray *prhs[]) {
  const mxArray *geo = prhs[1], *par = prhs[2], *fin = prhs[3], *ger = prhs[4];
  const mxArray *eps = field(par, "w", "epsilon"), *sig = field(par, "w", "sigma");
  const mxArray *air = mxGetField(field(par, "w", NULL), 0, "air");
  if (fscalar(fin, "w", "pml_w") != fscalar(fin, "w", "pml_d"))
    mexErrMsgIdAndTxt("gjw:arg", "pml_w must equal pml_d (w_exp2pml.m pads columns with pml_w)");
  gjw_grid g;
  memset(&g, 0, sizeof g);
  g.m = (int32_t)mxGetM(eps);
  g.n = (int32_t)mxGetN(eps);
  if (mxGetM(sig) != mxGetM(eps) || mxGetN(sig) != mxGetN(eps))
    mexErrMsgIdAndTxt("gjw:arg", "parame_.w.sigma and parame_.w.epsilon differ in size");
  g.pml = (int32_t)fscalar(fin, "w", "pml_w");
  g.air = air ? (int32_t)mxGetScalar(air) : 0;
  g.nt = (int32_t)fscalar(geo, "w", "nt");
  g.dt = fscalar(geo, "w", "dt");
  g.dx = fscalar(geo, "dx", NULL);
  g.eps0 = fscalar(par, "w", "eps_0");
  g.mu0 = fscalar(par, "w", "mu_0");
  g.R0 = fscalar(fin, "w", "R_0");
  g.pml_order = (int32_t)fscalar(fin, "w", "pml_order");
  ensure_ctx();
  if (!sv || memcmp(&g, &last, sizeof g)) {
    gjw_survey_destroy(sv);
    sv = NULL;
    have_fwd = 0;
    chk(gjw_survey_create(ctx, &g, &sv));
    last = g;
    chk(gjw_set_option(sv, GJW_OPT_PATH, GJW_PATH_PLAIN)); /* w_solve keeps u whole */
  }
  chk(gjw_set_model(sv, doubles(eps, "parame_.w.epsilon"), doubles(sig, "parame_.w.sigma")));
  const mxArray *s = field(ger, "w", "s"), *M = field(ger, "w", "M"), *wv = field(ger, "w", "wvlet_");
  if (mxGetNumberOfElements(s) != 2) mexErrMsgIdAndTxt("gjw:arg", "gerjoii_.w.s must be [iz ix]");
  if (mxGetNumberOfElements(wv) != (size_t)g.nt)
    mexErrMsgIdAndTxt("gjw:arg", "gerjoii_.w.wvlet_ must have nt = %d samples", g.nt);
  const int32_t src[2] = {(int32_t)num_at(s, 0), (int32_t)num_at(s, 1)};
  const size_t nr = mxGetNumberOfElements(M);
  int64_t rp[2] = {0, (int64_t)nr};
  int32_t *mw = (int32_t *)mxMalloc((nr ? nr : 1) * sizeof(int32_t));
  for (size_t r = 0; r < nr; ++r) mw[r] = (int32_t)num_at(M, r);
  float *w = to_f32(wv);
  chk(gjw_set_shots(sv, 1, src, w, rp, mw));
  plhs[0] = mxCreateNumericMatrix((mwSize)g.nt, (mwSize)nr, mxSINGLE_CLASS, mxREAL);
  mxArray *jy = mxCreateNumericMatrix((mwSize)g.nt, 1, mxSINGLE_CLASS, mxREAL);
  chk(gjw_forward(sv, mxGetSingles(plhs[0]), mxGetSingles(jy), 1));
  have_fwd = 1;
  if (nlhs > 1)
    plhs[1] = jy;
  else
    mxDestroyArray(jy);
  mxFree(mw);
  mxFree(w);
}

/* 'grad': mode ('REF' | 'EXACT'), e_2d (nt x nr), dtu -> g (m x n double),
 * a_max (m x n single; [] for EXACT).  dtu = 1: the permittivity gradient
 * (w_grad of differentiate_cube(u), w_update_e.m:83-98); 0: the conductivity
 * gradient (w_grad of u, w_update_s.m:77). */
static void op_grad(int nlhs, mxArray *plhs[], const mxArray *prhs[]) {
  char mode[8];
  get_str(prhs[1], mode, sizeof mode, "mode");
  const int ref = !strcmp(mode, "REF");
  if (!ref && strcmp(mode, "EXACT")) mexErrMsgIdAndTxt("gjw:arg", "mode must be 'REF' or 'EXACT'");
  if (!sv || !have_fwd) mexErrMsgIdAndTxt("gjw:state", "call gjw_mex('fwd', ...) first");
  const mxArray *e = prhs[2];
  if (mxGetM(e) != (size_t)last.nt)
    mexErrMsgIdAndTxt("gjw:arg", "e_2d must have nt = %d rows", last.nt);
  const int dtu = mxGetScalar(prhs[3]) != 0;
  float *adj = to_f32(e);
  mxArray *ge = mxCreateDoubleMatrix((mwSize)last.m, (mwSize)last.n, mxREAL);
  mxArray *gs = mxCreateDoubleMatrix((mwSize)last.m, (mwSize)last.n, mxREAL);
  mxArray *am = ref ? mxCreateNumericMatrix((mwSize)last.m, (mwSize)last.n, mxSINGLE_CLASS, mxREAL)
                    : mxCreateNumericMatrix(0, 0, mxSINGLE_CLASS, mxREAL);
  chk(gjw_gradient(sv, adj, ref ? GJW_GRAD_REF : GJW_GRAD_EXACT, mxGetDoubles(ge),
                   mxGetDoubles(gs), NULL, ref ? mxGetSingles(am) : NULL));
  plhs[0] = dtu ? ge : gs;
  mxDestroyArray(dtu ? gs : ge);
  if (nlhs > 1)
    plhs[1] = am;
  else
    mxDestroyArray(am);
  mxFree(adj);
}

static int32_t misfit_kind(const mxArray *x) {
  char obj[16];
  get_str(x, obj, sizeof obj, "obj_fnc");
  if (!strcmp(obj, "L2") || !strcmp(obj, "FREQstp")) return GJW_MISFIT_L2;
  if (!strcmp(obj, "env_L2")) return GJW_MISFIT_ENV_L2;
  if (!strcmp(obj, "phase_L2")) return GJW_MISFIT_PHASE_L2;
  if (!strcmp(obj, "env_L22")) return GJW_MISFIT_ENV_L22;
  if (!strcmp(obj, "env_LOG")) return GJW_MISFIT_ENV_LOG;
  mexErrMsgIdAndTxt("gjw:arg", "unknown obj_fnc %s", obj);
  return -1;
}

/* 'bozdag': d_o, d (nt x nr), obj_fnc -> e, d_e, d_oe (nt x nr single), J */
static void op_bozdag(int nlhs, mxArray *plhs[], const mxArray *prhs[]) {
  const mxArray *dob = prhs[1], *d = prhs[2];
  const size_t nt = mxGetM(d), nr = mxGetN(d);
  if (mxGetM(dob) != nt || mxGetN(dob) != nr) mexErrMsgIdAndTxt("gjw:arg", "d_o and d differ in size");
  const int32_t kind = misfit_kind(prhs[3]);
  ensure_ctx();
  float *o = to_f32(dob), *x = to_f32(d);
  mxArray *out[3];
  for (int k = 0; k < 3; ++k) out[k] = mxCreateNumericMatrix(nt, nr, mxSINGLE_CLASS, mxREAL);
  mxArray *J = mxCreateDoubleMatrix(1, 1, mxREAL);
  chk(gjw_w_bozdag(ctx, o, x, (int32_t)nr, (int32_t)nt, kind, mxGetSingles(out[0]),
                   mxGetSingles(out[1]), mxGetSingles(out[2]), mxGetDoubles(J)));
  for (int k = 0; k < 3; ++k) {
    if (k < (nlhs > 0 ? nlhs : 1))
      plhs[k] = out[k];
    else
      mxDestroyArray(out[k]);
  }
  if (nlhs > 3)
    plhs[3] = J;
  else
    mxDestroyArray(J);
  mxFree(o);
  mxFree(x);
}

/* 'filt_gauss': d (nt x nr single), dt, f_low, f_high, steep -> filtered copy */
static void op_filt_gauss(mxArray *plhs[], const mxArray *prhs[]) {
  if (!mxIsSingle(prhs[1])) mexErrMsgIdAndTxt("gjw:class", "d must be single");
  const size_t nt = mxGetM(prhs[1]), nr = mxGetN(prhs[1]);
  ensure_ctx();
  plhs[0] = mxDuplicateArray(prhs[1]);
  chk(gjw_filt_gauss(ctx, mxGetSingles(plhs[0]), (int32_t)nr, (int32_t)nt, mxGetScalar(prhs[2]),
                     mxGetScalar(prhs[3]), mxGetScalar(prhs[4]), (int32_t)mxGetScalar(prhs[5])));
}

/* 'lmute': d, d_o (nt x nr single), t (nt), rx (nr), v_mute [m/ns], t_mute [ns]
 * -> muted d, muted d_o, e = d - d_o */
static void op_lmute(int nlhs, mxArray *plhs[], const mxArray *prhs[]) {
  if (!mxIsSingle(prhs[1]) || !mxIsSingle(prhs[2]))
    mexErrMsgIdAndTxt("gjw:class", "d and d_o must be single");
  const size_t nt = mxGetM(prhs[1]), nr = mxGetN(prhs[1]);
  if (mxGetM(prhs[2]) != nt || mxGetN(prhs[2]) != nr)
    mexErrMsgIdAndTxt("gjw:arg", "d and d_o differ in size");
  if (mxGetNumberOfElements(prhs[3]) != nt || mxGetNumberOfElements(prhs[4]) != nr)
    mexErrMsgIdAndTxt("gjw:arg", "t must have nt and rx nr elements");
  ensure_ctx();
  mxArray *out[3] = {mxDuplicateArray(prhs[1]), mxDuplicateArray(prhs[2]),
                     mxCreateNumericMatrix(nt, nr, mxSINGLE_CLASS, mxREAL)};
  chk(gjw_lmute(ctx, mxGetSingles(out[0]), mxGetSingles(out[1]), mxGetSingles(out[2]),
                (int32_t)nr, (int32_t)nt, doubles(prhs[3], "t"), doubles(prhs[4], "rx"),
                mxGetScalar(prhs[5]), mxGetScalar(prhs[6])));
  for (int k = 0; k < 3; ++k) {
    if (k < (nlhs > 0 ? nlhs : 1))
      plhs[k] = out[k];
    else
      mxDestroyArray(out[k]);
  }
}

/* 'gaussi': g (m x n double), x (n), z (m), s = [iz ix], lo -> tapered copy */
static void op_gaussi(mxArray *plhs[], const mxArray *prhs[]) {
  const size_t m = mxGetM(prhs[1]), n = mxGetN(prhs[1]);
  if (mxGetNumberOfElements(prhs[2]) != n || mxGetNumberOfElements(prhs[3]) != m)
    mexErrMsgIdAndTxt("gjw:arg", "x must have n and z m elements");
  if (mxGetNumberOfElements(prhs[4]) != 2) mexErrMsgIdAndTxt("gjw:arg", "s must be [iz ix]");
  const int32_t s[2] = {(int32_t)num_at(prhs[4], 0), (int32_t)num_at(prhs[4], 1)};
  const double lo = mxGetScalar(prhs[5]);
  ensure_ctx();
  plhs[0] = mxDuplicateArray(prhs[1]);
  chk(gjw_gaussi(ctx, (double *)doubles(plhs[0], "g"), (int32_t)m, (int32_t)n, 1,
                 doubles(prhs[2], "x"), doubles(prhs[3], "z"), s, &lo));
}

/* 'image_gaussian': img (m x n double), a, b, 'LOW_PASS' | 'HI_PASS' */
static void op_image_gaussian(mxArray *plhs[], const mxArray *prhs[]) {
  char typ[16];
  get_str(prhs[4], typ, sizeof typ, "filt_TYPE");
  const int hi = !strcmp(typ, "HI_PASS");
  if (!hi && strcmp(typ, "LOW_PASS"))
    mexErrMsgIdAndTxt("gjw:arg", "filt_TYPE must be 'LOW_PASS' or 'HI_PASS'");
  const size_t m = mxGetM(prhs[1]), n = mxGetN(prhs[1]);
  ensure_ctx();
  plhs[0] = mxCreateDoubleMatrix(m, n, mxREAL);
  chk(gjw_image_gaussian(ctx, doubles(prhs[1], "image_xy"), mxGetDoubles(plhs[0]), (int32_t)m,
                         (int32_t)n, 1, mxGetScalar(prhs[2]), mxGetScalar(prhs[3]),
                         hi ? GJW_HI_PASS : GJW_LOW_PASS));
}

/* 'normali': a (m x n double) -> a / max|a| (a zero maximum taken as 1) */
static void op_normali(mxArray *plhs[], const mxArray *prhs[]) {
  const size_t m = mxGetM(prhs[1]), n = mxGetN(prhs[1]);
  ensure_ctx();
  plhs[0] = mxDuplicateArray(prhs[1]);
  chk(gjw_normali(ctx, (double *)doubles(plhs[0], "a"), (int32_t)m, (int32_t)n, 1, 0, NULL));
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  char op[24];
  if (nrhs < 1) mexErrMsgIdAndTxt("gjw:nrhs", "gjw_mex(op, ...): missing op");
  get_str(prhs[0], op, sizeof op, "op");
  if (!strcmp(op, "fwd")) {
    need(nrhs, 5, op);
    op_fwd(nlhs, plhs, prhs);
  } else if (!strcmp(op, "grad")) {
    need(nrhs, 4, op);
    op_grad(nlhs, plhs, prhs);
  } else if (!strcmp(op, "bozdag")) {
    need(nrhs, 4, op);
    op_bozdag(nlhs, plhs, prhs);
  } else if (!strcmp(op, "filt_gauss")) {
    need(nrhs, 6, op);
    op_filt_gauss(plhs, prhs);
  } else if (!strcmp(op, "lmute")) {
    need(nrhs, 7, op);
    op_lmute(nlhs, plhs, prhs);
  } else if (!strcmp(op, "gaussi")) {
    need(nrhs, 6, op);
    op_gaussi(plhs, prhs);
  } else if (!strcmp(op, "image_gaussian")) {
    need(nrhs, 5, op);
    op_image_gaussian(plhs, prhs);
  } else if (!strcmp(op, "normali")) {
    need(nrhs, 2, op);
    op_normali(plhs, prhs);
  } else {
    mexErrMsgIdAndTxt("gjw:op", "unknown op %s", op);
  }
}
