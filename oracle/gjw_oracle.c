/*
 * gjw_oracle.c -- CPU oracle for the radar FDTD hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see gjw_oracle_impl.h).  Built by oracle/Makefile
 * into oracle/libgjw_oracle.so; loaded by oracle/oracle.py, which only tests/,
 * smoke() and bench.py's cpu_baseline leg import.
 *
 * Restates, for REAL = double and REAL = float:
 *   w_exp2pml.m:9-37   material padding           -> orc_exp2pml_{d,f}
 *   w_cpml.m:44-185    split-field PML coefficients \
 *   w_coeff.m:28-55    inner coefficients           -> orc_coeff_export_{d,f}
 *   w_solve.m:80-268   forward leapfrog             -> orc_forward_{d,f}
 *   w_grad.m:20-115    reference gradient (REF)     -> orc_grad_ref_{d,f}
 *   differentiate_cube.m:23-51                      -> orc_differentiate_cube_{d,f}
 * and the exact discrete adjoint of w_solve.m       -> orc_grad_exact_{d,f}.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* padded-grid dimensions, w_geom.m:135-172 (1-based, as in MATLAB) */
typedef struct {
  int m, n, P, air;          /* model rows (depth), cols, PML width, air rows */
  int NX, NY;                /* finite_.w.nx, finite_.w.ny                   */
  int pis, pie, pjs, pje;    /* w_geom.m:161-164                             */
} orc_dims;

typedef struct {
  double dx, dt, eps0, mu0, R0;
  int order, nt;
} orc_phys;

int orc_dims_make(int m, int n, int P, int air, orc_dims *g) {
  if (m < 1 || n < 1 || P < 1 || air < 0) return -1;
  g->m = m; g->n = n; g->P = P; g->air = air;
  g->NY = n + 2 * P;           /* w_geom.m:135 (pml_d pads columns)  */
  g->NX = m + 2 * P + air;     /* w_geom.m:136 (pml_w pads depth)    */
  g->pis = P + 1;
  g->pie = g->NX - P + 1;
  g->pjs = P + 1;
  g->pje = g->NY - P + 1;
  return 0;
}

/* 16-bit checkpoint snapshots with per-snapshot scaling (mode 1 = BF16,
 * 2 = FP16), the format of the device checkpoints: a plane is scaled by 2^sh,
 * sh from its max |v| (largest value just below 2^15 for FP16, in [1, 2) for
 * BF16), rounded to nearest even at 8 (BF16) or 11 (FP16) significant bits
 * with the format's subnormal step (2^-133, 2^-24), and scaled back.  Only
 * the rounding changes a value: the power-of-two scaling is exact. */
static int snap_shift(double amax, int mode) {
  if (!(amax > 0)) return 0;
  int e;
  frexp(amax, &e); /* amax = f * 2^e, 0.5 <= f < 1 */
  return mode == 2 ? 15 - e : 1 - e;
}

static double snap_round(double v, int sh, int mode) {
  const double x = ldexp(v, sh);
  if (x == 0 || !isfinite(x)) return v;
  int e;
  frexp(fabs(x), &e); /* |x| in [2^(e-1), 2^e) */
  const int bits = mode == 2 ? 11 : 8, qmin = mode == 2 ? -24 : -133;
  int qe = (e - 1) - (bits - 1);
  if (qe < qmin) qe = qmin;
  const double q = ldexp(1.0, qe);
  return ldexp(nearbyint(x / q) * q, -sh);
}

/* exported for the tests of the rounding itself */
int orc_snap_shift(double amax, int mode) { return snap_shift(amax, mode); }
double orc_snap_round(double v, int sh, int mode) { return snap_round(v, sh, mode); }

#define REAL double
#define SFX _d
#include "gjw_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX _f
#include "gjw_oracle_impl.h"
#undef REAL
#undef SFX

/* Chain rule of w_exp2pml.m: the padded eps copies the nearest model (or air)
 * value into the PML; padded sigma is zero outside the model.  Sums the
 * padded-grid gradient back onto the m x n model. */
int orc_fold_pad(const orc_dims *g, const double *g_eps_pad,
                 const double *g_sig_pad, double *g_eps, double *g_sig) {
  const int NX = g->NX, NY = g->NY, P = g->P, m = g->m, n = g->n, air = g->air;
  if (g_eps) memset(g_eps, 0, sizeof(double) * m * n);
  if (g_sig) memset(g_sig, 0, sizeof(double) * m * n);
  for (int j = 1; j <= NY; ++j) {
    int jm = j - P;
    if (jm < 1) jm = 1;
    if (jm > n) jm = n;
    for (int i = 1; i <= NX; ++i) {
      int r = i - P;
      if (r < 1) r = 1;
      if (r > air + m) r = air + m;
      size_t kp = (size_t)(i - 1) + (size_t)(j - 1) * NX;
      if (r > air) {
        size_t km = (size_t)(r - air - 1) + (size_t)(jm - 1) * m;
        if (g_eps && g_eps_pad) g_eps[km] += g_eps_pad[kp];
        int in_model = (i - P >= air + 1 && i - P <= air + m && j - P >= 1 && j - P <= n);
        if (in_model && g_sig && g_sig_pad) g_sig[km] += g_sig_pad[kp];
      }
    }
  }
  return 0;
}

/* CPU baseline: float instance, OpenMP over shots (one shot per thread),
 * forward + exact adjoint per shot.  e_all is nt x nr per shot; the
 * per-shot padded gradients are summed in shot order afterwards.  Returns
 * the number of threads used through *threads. */
int orc_survey_grad_f(const orc_dims *g, const orc_phys *ph, const float *eps,
                      const float *sig, int ns, const int *src /* 2 x ns */,
                      const float *wv /* nt x ns */, int nr, const int *Mw /* nr x ns */,
                      const float *e_all, int K, double *g_eps_pad,
                      double *g_sig_pad, int *threads) {
  const size_t npad = (size_t)g->NX * g->NY;
  double *ge = (double *)calloc((size_t)ns * npad, sizeof(double));
  double *gs = (double *)calloc((size_t)ns * npad, sizeof(double));
  if (!ge || !gs) return -1;
  int fail = 0, nthr = 1;
#pragma omp parallel
  {
#ifdef _OPENMP
#pragma omp single
    nthr = omp_get_num_threads();
#endif
#pragma omp for schedule(dynamic, 1)
    for (int s = 0; s < ns; ++s) {
      int rc = orc_grad_exact_f(g, ph, eps, sig, src[2 * s], src[2 * s + 1],
                                wv + (size_t)s * ph->nt, nr, Mw + (size_t)s * nr,
                                e_all + (size_t)s * ph->nt * nr, K, 0,
                                ge + (size_t)s * npad, gs + (size_t)s * npad, NULL, 0);
      if (rc) fail = rc;
    }
  }
  memset(g_eps_pad, 0, npad * sizeof(double));
  memset(g_sig_pad, 0, npad * sizeof(double));
  for (int s = 0; s < ns; ++s)
    for (size_t k = 0; k < npad; ++k) {
      g_eps_pad[k] += ge[(size_t)s * npad + k];
      g_sig_pad[k] += gs[(size_t)s * npad + k];
    }
  free(ge);
  free(gs);
  if (threads) *threads = nthr;
  return fail;
}

/* forward only, float, OpenMP over shots (CPU baseline for the trace path) */
int orc_survey_forward_f(const orc_dims *g, const orc_phys *ph, const float *eps,
                         const float *sig, int ns, const int *src,
                         const float *wv, int nr, const int *Mw, float *d_all,
                         int *threads) {
  int fail = 0, nthr = 1;
#pragma omp parallel
  {
#ifdef _OPENMP
#pragma omp single
    nthr = omp_get_num_threads();
#endif
#pragma omp for schedule(dynamic, 1)
    for (int s = 0; s < ns; ++s) {
      int rc = orc_forward_f(g, ph, eps, sig, src[2 * s], src[2 * s + 1],
                             wv + (size_t)s * ph->nt, nr, Mw + (size_t)s * nr,
                             d_all + (size_t)s * ph->nt * nr, NULL, NULL);
      if (rc) fail = rc;
    }
  }
  if (threads) *threads = nthr;
  return fail;
}
