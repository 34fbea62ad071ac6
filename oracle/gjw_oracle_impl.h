/*
 * gjw_oracle_impl.h -- body of the CPU oracle, instantiated once per REAL.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the reference's
 * radar FDTD forward solve and gradient (MATLAB, src/w/).  It is used by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker, never as the thing measured or shipped.  Nothing under gerjoii_amd/
 * links, loads or calls it.
 *
 * Included by gjw_oracle.c with REAL = double (suffix _d) and REAL = float
 * (suffix _f).  The float instance is the "singles" solver that
 * src/w/w_fwd.m:7-15 describes; the double instance is the FP64 oracle that
 * the same comment calls the singles->doubles switch.
 *
 * Index conventions follow MATLAB: 1-based, column-major, depth (i) fastest.
 *   Ez (NX+1) x (NY+1), Hx (NX+1) x NY, Hy NX x (NY+1), eps/sig NX x NY.
 */

/* Column loops of one step run on OpenMP threads when the grid is large: each
 * iteration writes its own cells with the same arithmetic as the serial loop,
 * so results are bit-identical for any thread count.  Inside an enclosing
 * parallel region (one shot per thread) the loops stay serial. */
#ifndef ORC_PAR_DEFINED
#define ORC_PAR_DEFINED
#define ORC_PAR_CELLS 65536
#define OMP_COLS _Pragma("omp parallel for schedule(static) if (par)")
static void orc_par_copy(void *dst, const void *src, size_t bytes, int par) {
  const size_t chunk = (size_t)1 << 20;
  const long nb = (long)((bytes + chunk - 1) / chunk);
#pragma omp parallel for schedule(static) if (par)
  for (long b = 0; b < nb; ++b) {
    const size_t o = (size_t)b * chunk, n = o + chunk <= bytes ? chunk : bytes - o;
    memcpy((char *)dst + o, (const char *)src + o, n);
  }
}
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)
#define TY(name) CAT(name, SFX)

/* ------------------------------------------------------------------------- */
/* coefficient set: w_cpml.m + w_coeff.m                                      */
/* ------------------------------------------------------------------------- */
typedef struct {
  REAL *Ceze, *Cezhy, *Cezhx;                /* NX x NY (w_coeff.m:28-30)   */
  REAL Chxh, Chxez, Chyh, Chyez, Cezj;       /* mu_r=1, sigma_m=0 scalars   */
  /* x strips, P x (NY-1) unless noted (w_cpml.m:44-104)                     */
  REAL *Chyh_xn, *Chyez_xn, *Cezxe_xn, *Cezxhy_xn;
  REAL Cezye_xn, Cezyhx_xn;
  REAL *Chyh_xp, *Chyez_xp, *Cezxe_xp, *Cezxhy_xp;
  REAL Cezye_xp;
  REAL *Cezyhx_xp;                            /* P x (NY-1-2P)               */
  /* y strips (w_cpml.m:108-185)                                              */
  REAL *Chxh_yn, *Chxez_yn;                   /* (NX-1) x P                  */
  REAL Cezxe_yn;
  REAL *Cezxhy_yn;                            /* (NX-1-2P) x P               */
  REAL *Cezye_yn, *Cezyhx_yn;                 /* (NX-1) x P                  */
  REAL *Chxh_yp, *Chxez_yp;
  REAL Cezxe_yp;
  REAL *Cezxhy_yp;
  REAL *Cezye_yp, *Cezyhx_yp;
  /* PML conductivity profiles before division by eps (for the chain rule)  */
  double *spex_xn, *spmx_xn, *spex_xp, *spmx_xp; /* length P, by strip row  */
  double *spey_yn, *spmy_yn, *spey_yp, *spmy_yp; /* length P, by strip col  */
} TY(orc_coeffs);

static void FN(coeffs_free)(TY(orc_coeffs) * c) {
  REAL **arrs[] = {&c->Ceze,     &c->Cezhy,    &c->Cezhx,     &c->Chyh_xn,
                   &c->Chyez_xn, &c->Cezxe_xn, &c->Cezxhy_xn, &c->Chyh_xp,
                   &c->Chyez_xp, &c->Cezxe_xp, &c->Cezxhy_xp, &c->Cezyhx_xp,
                   &c->Chxh_yn,  &c->Chxez_yn, &c->Cezxhy_yn, &c->Cezye_yn,
                   &c->Cezyhx_yn, &c->Chxh_yp, &c->Chxez_yp,  &c->Cezxhy_yp,
                   &c->Cezye_yp, &c->Cezyhx_yp};
  for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); ++k) {
    free(*arrs[k]);
    *arrs[k] = NULL;
  }
  double **d[] = {&c->spex_xn, &c->spmx_xn, &c->spex_xp, &c->spmx_xp,
                  &c->spey_yn, &c->spmy_yn, &c->spey_yp, &c->spmy_yp};
  for (size_t k = 0; k < sizeof(d) / sizeof(d[0]); ++k) {
    free(*d[k]);
    *d[k] = NULL;
  }
}

/* eps/sig are the padded NX x NY materials (w_exp2pml.m output).
 * sig_sign = +1 for the forward (w_fwd.m:20-22), -1 for the REF gradient's
 * negated conductivity (w_grad.m:24-27). */
static int FN(coeffs_build)(const orc_dims *g, const orc_phys *ph,
                            const REAL *eps, const REAL *sig, int si, int sj,
                            double sig_sign, TY(orc_coeffs) * c) {
  memset(c, 0, sizeof(*c));
  const int NX = g->NX, NY = g->NY, P = g->P;
  const int pis = g->pis, pie = g->pie, pjs = g->pjs, pje = g->pje;
  const double e0 = ph->eps0, u0 = ph->mu0, dt = ph->dt, dx = ph->dx,
               dy = ph->dx;
  const double cl = 1.0 / sqrt(u0 * e0);
  const int ord = ph->order;
#define EPS(i, j) ((double)eps[((i)-1) + (size_t)((j)-1) * NX])
#define SIG(i, j) (sig_sign * (double)sig[((i)-1) + (size_t)((j)-1) * NX])
#define ALLOC(p, cnt)                                      \
  do {                                                     \
    (p) = (REAL *)calloc((size_t)(cnt) > 0 ? (cnt) : 1, sizeof(REAL)); \
    if (!(p)) return -1;                                   \
  } while (0)
#define ALLOCD(p, cnt)                                        \
  do {                                                        \
    (p) = (double *)calloc((size_t)(cnt) > 0 ? (cnt) : 1, sizeof(double)); \
    if (!(p)) return -1;                                      \
  } while (0)
  /* ---- w_coeff.m:28-55 ---- */
  ALLOC(c->Ceze, (size_t)NX * NY);
  ALLOC(c->Cezhy, (size_t)NX * NY);
  ALLOC(c->Cezhx, (size_t)NX * NY);
  for (int j = 1; j <= NY; ++j)
    for (int i = 1; i <= NX; ++i) {
      double ee = EPS(i, j), ss = SIG(i, j);
      double den = 2.0 * ee * e0 + dt * ss;
      size_t k = (size_t)(i - 1) + (size_t)(j - 1) * NX;
      c->Ceze[k] = (REAL)((2.0 * ee * e0 - dt * ss) / den);
      c->Cezhy[k] = (REAL)((2.0 * dt / dx) / den);
      c->Cezhx[k] = (REAL)(-(2.0 * dt / dy) / den);
    }
  c->Chxh = (REAL)((2.0 * u0 - dt * 0.0) / (2.0 * u0 + dt * 0.0));
  c->Chxez = (REAL)(-(2.0 * dt / dy) / (2.0 * u0 + dt * 0.0));
  c->Chyh = (REAL)((2.0 * u0 - dt * 0.0) / (2.0 * u0 + dt * 0.0));
  c->Chyez = (REAL)((2.0 * dt / dx) / (2.0 * u0 + dt * 0.0));
  if (si >= 1 && si <= NX && sj >= 1 && sj <= NY)
    c->Cezj = (REAL)(-(2.0 * dt) / (2.0 * EPS(si, sj) * e0 + dt * SIG(si, sj)));
  else
    c->Cezj = 0;

  /* ---- w_cpml.m:44-64, xn strip (vacuum) ---- */
  const double smx = -(ord + 1) * e0 * cl * log(ph->R0) / (2.0 * dx * P);
  const double smy = -(ord + 1) * e0 * cl * log(ph->R0) / (2.0 * dy * P);
  ALLOCD(c->spex_xn, P); ALLOCD(c->spmx_xn, P);
  ALLOCD(c->spex_xp, P); ALLOCD(c->spmx_xp, P);
  ALLOCD(c->spey_yn, P); ALLOCD(c->spmy_yn, P);
  ALLOCD(c->spey_yp, P); ALLOCD(c->spmy_yp, P);
  for (int a = 1; a <= P; ++a) {
    double re = ((P - a + 1) - 0.75) / P, rm = ((P - a + 1) - 0.25) / P;
    c->spex_xn[a - 1] = smx * pow(re, ord);
    c->spmx_xn[a - 1] = (u0 / e0) * smx * pow(rm, ord);
    re = (a - 0.75) / P; rm = (a - 0.25) / P;
    c->spex_xp[a - 1] = smx * pow(re, ord);
    c->spmx_xp[a - 1] = (u0 / e0) * smx * pow(rm, ord);
  }
  for (int b = 1; b <= P; ++b) {
    double re = ((P - b + 1) - 0.75) / P, rm = ((P - b + 1) - 0.25) / P;
    c->spey_yn[b - 1] = smy * pow(re, ord);
    c->spmy_yn[b - 1] = (u0 / e0) * smy * pow(rm, ord);
    re = (b - 0.75) / P; rm = (b - 0.25) / P;
    c->spey_yp[b - 1] = smy * pow(re, ord);
    c->spmy_yp[b - 1] = (u0 / e0) * smy * pow(rm, ord);
  }
  const int nym1 = NY - 1, nxm1 = NX - 1;
  const int wxy = NY - 1 - 2 * P; /* Ezy_xn/xp columns   */
  const int hyx = NX - 1 - 2 * P; /* Ezx_yn/yp rows      */
  ALLOC(c->Chyh_xn, (size_t)P * nym1);
  ALLOC(c->Chyez_xn, (size_t)P * nym1);
  ALLOC(c->Cezxe_xn, (size_t)P * nym1);
  ALLOC(c->Cezxhy_xn, (size_t)P * nym1);
  for (int b = 1; b <= nym1; ++b)
    for (int a = 1; a <= P; ++a) {
      size_t k = (size_t)(a - 1) + (size_t)(b - 1) * P;
      double sm = c->spmx_xn[a - 1], se = c->spex_xn[a - 1];
      c->Chyh_xn[k] = (REAL)((2 * u0 - dt * sm) / (2 * u0 + dt * sm));
      c->Chyez_xn[k] = (REAL)((2 * dt / dx) / (2 * u0 + dt * sm));
      c->Cezxe_xn[k] = (REAL)((2 * e0 - dt * se) / (2 * e0 + dt * se));
      c->Cezxhy_xn[k] = (REAL)((2 * dt / dx) / (2 * e0 + dt * se));
    }
  c->Cezye_xn = (REAL)1;
  c->Cezyhx_xn = (REAL)(-dt / (dy * e0));
  /* ---- w_cpml.m:68-104, xp strip ---- */
  ALLOC(c->Chyh_xp, (size_t)P * nym1);
  ALLOC(c->Chyez_xp, (size_t)P * nym1);
  ALLOC(c->Cezxe_xp, (size_t)P * nym1);
  ALLOC(c->Cezxhy_xp, (size_t)P * nym1);
  ALLOC(c->Cezyhx_xp, (size_t)P * (wxy > 0 ? wxy : 1));
  for (int b = 1; b <= nym1; ++b)
    for (int a = 1; a <= P; ++a) {
      size_t k = (size_t)(a - 1) + (size_t)(b - 1) * P;
      double ee = EPS(pie - 1 + a, b + 1);
      double sm = c->spmx_xp[a - 1] / ee, se = c->spex_xp[a - 1];
      c->Chyh_xp[k] = (REAL)((2 * u0 - dt * sm) / (2 * u0 + dt * sm));
      c->Chyez_xp[k] = (REAL)((2 * dt / dx) / (2 * u0 + dt * sm));
      c->Cezxe_xp[k] =
          (REAL)((2 * e0 * ee - dt * se) / (2 * e0 * ee + dt * se));
      c->Cezxhy_xp[k] = (REAL)((2 * dt / dx) / (2 * e0 * ee + dt * se));
    }
  for (int b = 1; b <= wxy; ++b)
    for (int a = 1; a <= P; ++a)
      c->Cezyhx_xp[(size_t)(a - 1) + (size_t)(b - 1) * P] =
          (REAL)(-dt / (dy * e0 * EPS(pie - 1 + a, pjs + b)));
  c->Cezye_xp = (REAL)1;
  /* ---- w_cpml.m:108-144, yn strip ---- */
  ALLOC(c->Chxh_yn, (size_t)nxm1 * P);
  ALLOC(c->Chxez_yn, (size_t)nxm1 * P);
  ALLOC(c->Cezye_yn, (size_t)nxm1 * P);
  ALLOC(c->Cezyhx_yn, (size_t)nxm1 * P);
  ALLOC(c->Cezxhy_yn, (size_t)(hyx > 0 ? hyx : 1) * P);
  ALLOC(c->Chxh_yp, (size_t)nxm1 * P);
  ALLOC(c->Chxez_yp, (size_t)nxm1 * P);
  ALLOC(c->Cezye_yp, (size_t)nxm1 * P);
  ALLOC(c->Cezyhx_yp, (size_t)nxm1 * P);
  ALLOC(c->Cezxhy_yp, (size_t)(hyx > 0 ? hyx : 1) * P);
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= nxm1; ++a) {
      size_t k = (size_t)(a - 1) + (size_t)(b - 1) * nxm1;
      /* yn: Hx at eps(2:nx, 1:pjs-1), Ezy at eps(2:nx, 2:pjs) */
      double eh = EPS(a + 1, b), ee = EPS(a + 1, b + 1);
      double sm = c->spmy_yn[b - 1] / eh, se = c->spey_yn[b - 1];
      c->Chxh_yn[k] = (REAL)((2 * u0 - dt * sm) / (2 * u0 + dt * sm));
      c->Chxez_yn[k] = (REAL)(-(2 * dt / dy) / (2 * u0 + dt * sm));
      c->Cezye_yn[k] =
          (REAL)((2 * e0 * ee - dt * se) / (2 * e0 * ee + dt * se));
      c->Cezyhx_yn[k] = (REAL)(-(2 * dt / dy) / (2 * e0 * ee + dt * se));
      /* yp: Hx at eps(2:end, pje:ny), Ezy at eps(2:nx, pje:ny) */
      eh = EPS(a + 1, pje - 1 + b);
      ee = EPS(a + 1, pje - 1 + b);
      sm = c->spmy_yp[b - 1] / eh;
      se = c->spey_yp[b - 1];
      c->Chxh_yp[k] = (REAL)((2 * u0 - dt * sm) / (2 * u0 + dt * sm));
      c->Chxez_yp[k] = (REAL)(-(2 * dt / dy) / (2 * u0 + dt * sm));
      c->Cezye_yp[k] =
          (REAL)((2 * e0 * ee - dt * se) / (2 * e0 * ee + dt * se));
      c->Cezyhx_yp[k] = (REAL)(-(2 * dt / dy) / (2 * e0 * ee + dt * se));
    }
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= hyx; ++a) {
      size_t k = (size_t)(a - 1) + (size_t)(b - 1) * hyx;
      c->Cezxhy_yn[k] = (REAL)(dt / (dx * e0 * EPS(pis + a, b + 1)));
      c->Cezxhy_yp[k] = (REAL)(dt / (dx * e0 * EPS(pis + a, pje - 1 + b)));
    }
  c->Cezxe_yn = (REAL)1;
  c->Cezxe_yp = (REAL)1;
  (void)pje;
#undef EPS
#undef SIG
#undef ALLOC
#undef ALLOCD
  return 0;
}

/* ------------------------------------------------------------------------- */
/* field state: Ez, Hx, Hy and the 8 split-field arrays (w_solve.m:61-72)    */
/* ------------------------------------------------------------------------- */
typedef struct {
  REAL *buf; /* one allocation, so a snapshot is a memcpy                   */
  size_t len;
  REAL *Ez, *Hx, *Hy;
  REAL *Ezx_xn, *Ezy_xn, *Ezx_xp, *Ezy_xp, *Ezx_yn, *Ezy_yn, *Ezx_yp, *Ezy_yp;
} TY(orc_state);

static int FN(state_alloc)(const orc_dims *g, TY(orc_state) * s) {
  const size_t NX = g->NX, NY = g->NY, P = g->P;
  size_t sz[11] = {(NX + 1) * (NY + 1), (NX + 1) * NY, NX * (NY + 1),
                   P * (NY - 1), P * (NY - 1 - 2 * P), P * (NY - 1),
                   P * (NY - 1 - 2 * P), (NX - 1 - 2 * P) * P, (NX - 1) * P,
                   (NX - 1 - 2 * P) * P, (NX - 1) * P};
  size_t tot = 0;
  for (int k = 0; k < 11; ++k) tot += sz[k];
  s->buf = (REAL *)calloc(tot, sizeof(REAL));
  if (!s->buf) return -1;
  s->len = tot;
  REAL **p[11] = {&s->Ez,     &s->Hx,     &s->Hy,     &s->Ezx_xn,
                  &s->Ezy_xn, &s->Ezx_xp, &s->Ezy_xp, &s->Ezx_yn,
                  &s->Ezy_yn, &s->Ezx_yp, &s->Ezy_yp};
  size_t off = 0;
  for (int k = 0; k < 11; ++k) {
    *p[k] = s->buf + off;
    off += sz[k];
  }
  return 0;
}

static void FN(state_rebase)(const TY(orc_state) * src, REAL *buf,
                             TY(orc_state) * dst) {
  /* view of another buffer with the same offsets as src */
  *dst = *src;
  dst->buf = buf;
  ptrdiff_t d = buf - src->buf;
  dst->Ez += d; dst->Hx += d; dst->Hy += d;
  dst->Ezx_xn += d; dst->Ezy_xn += d; dst->Ezx_xp += d; dst->Ezy_xp += d;
  dst->Ezx_yn += d; dst->Ezy_yn += d; dst->Ezx_yp += d; dst->Ezy_yp += d;
}

/* ------------------------------------------------------------------------- */
/* one leapfrog step, literal to w_solve.m:85-94 and its local functions     */
/* ------------------------------------------------------------------------- */
static void FN(step_literal)(const orc_dims *g, const TY(orc_coeffs) * c,
                             TY(orc_state) * s, int si, int sj, REAL wsrc,
                             int with_source) {
  const int NX = g->NX, NY = g->NY, P = g->P;
  const int pis = g->pis, pie = g->pie, pjs = g->pjs, pje = g->pje;
  const size_t le = (size_t)NX + 1; /* Ez and Hx leading dimension */
  const size_t lh = (size_t)NX;     /* Hy leading dimension        */
  REAL *Ez = s->Ez, *Hx = s->Hx, *Hy = s->Hy;
  const int par = (size_t)(NX + 1) * (NY + 1) >= ORC_PAR_CELLS && !omp_in_parallel();
#define EZ(i, j) Ez[((i)-1) + (size_t)((j)-1) * le]
#define HX(i, j) Hx[((i)-1) + (size_t)((j)-1) * le]
#define HY(i, j) Hy[((i)-1) + (size_t)((j)-1) * lh]
#define A2(A, a, b, ld) (A)[((a)-1) + (size_t)((b)-1) * (size_t)(ld)]
#define CC(A, i, j) A2(c->A, i, j, NX)
  /* new_H, w_solve.m:136-148 */
  OMP_COLS
  for (int j = pjs; j <= pje - 1; ++j)
    for (int i = 1; i <= NX + 1; ++i)
      HX(i, j) = c->Chxh * HX(i, j) + c->Chxez * (EZ(i, j + 1) - EZ(i, j));
  OMP_COLS
  for (int j = 1; j <= NY + 1; ++j)
    for (int i = pis; i <= pie - 1; ++i)
      HY(i, j) = c->Chyh * HY(i, j) + c->Chyez * (EZ(i + 1, j) - EZ(i, j));
  /* new_H_pml, w_solve.m:154-187 */
  OMP_COLS
  for (int b = 1; b <= NY - 1; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = a, j = b + 1;
      HY(i, j) = A2(c->Chyh_xn, a, b, P) * HY(i, j) +
                 A2(c->Chyez_xn, a, b, P) * (EZ(i + 1, j) - EZ(i, j));
    }
  OMP_COLS
  for (int b = 1; b <= NY - 1; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = pie - 1 + a, j = b + 1;
      HY(i, j) = A2(c->Chyh_xp, a, b, P) * HY(i, j) +
                 A2(c->Chyez_xp, a, b, P) * (EZ(i + 1, j) - EZ(i, j));
    }
  OMP_COLS
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= NX - 1; ++a) {
      int i = a + 1, j = b;
      HX(i, j) = A2(c->Chxh_yn, a, b, NX - 1) * HX(i, j) +
                 A2(c->Chxez_yn, a, b, NX - 1) * (EZ(i, j + 1) - EZ(i, j));
    }
  OMP_COLS
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= NX - 1; ++a) {
      int i = a + 1, j = pje - 1 + b;
      HX(i, j) = A2(c->Chxh_yp, a, b, NX - 1) * HX(i, j) +
                 A2(c->Chxez_yp, a, b, NX - 1) * (EZ(i, j + 1) - EZ(i, j));
    }
  /* new_Ez, w_solve.m:193-207 */
  OMP_COLS
  for (int j = pjs + 1; j <= pje - 1; ++j)
    for (int i = pis + 1; i <= pie - 1; ++i)
      EZ(i, j) = CC(Ceze, i, j) * EZ(i, j) +
                 CC(Cezhy, i, j) * (HY(i, j) - HY(i - 1, j)) +
                 CC(Cezhx, i, j) * (HX(i, j) - HX(i, j - 1));
  /* source, w_solve.m:92 */
  if (with_source) EZ(si, sj) = EZ(si, sj) + c->Cezj * wsrc;
  /* new_Ez_pml, w_solve.m:215-268 */
  const int wxy = NY - 1 - 2 * P, hyx = NX - 1 - 2 * P;
  OMP_COLS
  for (int b = 1; b <= NY - 1; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = a + 1, j = b + 1;
      A2(s->Ezx_xn, a, b, P) =
          A2(c->Cezxe_xn, a, b, P) * A2(s->Ezx_xn, a, b, P) +
          A2(c->Cezxhy_xn, a, b, P) * (HY(i, j) - HY(i - 1, j));
    }
  OMP_COLS
  for (int b = 1; b <= wxy; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = a + 1, j = pjs + b;
      A2(s->Ezy_xn, a, b, P) = c->Cezye_xn * A2(s->Ezy_xn, a, b, P) +
                               c->Cezyhx_xn * (HX(i, j) - HX(i, j - 1));
    }
  OMP_COLS
  for (int b = 1; b <= NY - 1; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = pie - 1 + a, j = b + 1;
      A2(s->Ezx_xp, a, b, P) =
          A2(c->Cezxe_xp, a, b, P) * A2(s->Ezx_xp, a, b, P) +
          A2(c->Cezxhy_xp, a, b, P) * (HY(i, j) - HY(i - 1, j));
    }
  OMP_COLS
  for (int b = 1; b <= wxy; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = pie - 1 + a, j = pjs + b;
      A2(s->Ezy_xp, a, b, P) =
          c->Cezye_xp * A2(s->Ezy_xp, a, b, P) +
          A2(c->Cezyhx_xp, a, b, P) * (HX(i, j) - HX(i, j - 1));
    }
  OMP_COLS
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= hyx; ++a) {
      int i = pis + a, j = b + 1;
      A2(s->Ezx_yn, a, b, hyx) =
          c->Cezxe_yn * A2(s->Ezx_yn, a, b, hyx) +
          A2(c->Cezxhy_yn, a, b, hyx) * (HY(i, j) - HY(i - 1, j));
    }
  OMP_COLS
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= NX - 1; ++a) {
      int i = a + 1, j = b + 1;
      A2(s->Ezy_yn, a, b, NX - 1) =
          A2(c->Cezye_yn, a, b, NX - 1) * A2(s->Ezy_yn, a, b, NX - 1) +
          A2(c->Cezyhx_yn, a, b, NX - 1) * (HX(i, j) - HX(i, j - 1));
    }
  OMP_COLS
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= hyx; ++a) {
      int i = pis + a, j = pje - 1 + b;
      A2(s->Ezx_yp, a, b, hyx) =
          c->Cezxe_yp * A2(s->Ezx_yp, a, b, hyx) +
          A2(c->Cezxhy_yp, a, b, hyx) * (HY(i, j) - HY(i - 1, j));
    }
  OMP_COLS
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= NX - 1; ++a) {
      int i = a + 1, j = pje - 1 + b;
      A2(s->Ezy_yp, a, b, NX - 1) =
          A2(c->Cezye_yp, a, b, NX - 1) * A2(s->Ezy_yp, a, b, NX - 1) +
          A2(c->Cezyhx_yp, a, b, NX - 1) * (HX(i, j) - HX(i, j - 1));
    }
  /* Ez = Ezx + Ezy, eight regions (w_solve.m:260-267), in that order */
  OMP_COLS
  for (int j = 2; j <= pjs; ++j)
    for (int i = 2; i <= pis; ++i)
      EZ(i, j) = A2(s->Ezx_xn, i - 1, j - 1, P) +
                 A2(s->Ezy_yn, i - 1, j - 1, NX - 1);
  OMP_COLS
  for (int j = pje; j <= NY; ++j)
    for (int i = 2; i <= pis; ++i)
      EZ(i, j) = A2(s->Ezx_xn, i - 1, j - 1, P) +
                 A2(s->Ezy_yp, i - 1, j - pje + 1, NX - 1);
  OMP_COLS
  for (int j = pje; j <= NY; ++j)
    for (int i = pie; i <= NX; ++i)
      EZ(i, j) = A2(s->Ezx_xp, i - pie + 1, j - 1, P) +
                 A2(s->Ezy_yp, i - 1, j - pje + 1, NX - 1);
  OMP_COLS
  for (int j = 2; j <= pjs; ++j)
    for (int i = pie; i <= NX; ++i)
      EZ(i, j) = A2(s->Ezx_xp, i - pie + 1, j - 1, P) +
                 A2(s->Ezy_yn, i - 1, j - 1, NX - 1);
  OMP_COLS
  for (int j = 2; j <= pjs; ++j)
    for (int i = pis + 1; i <= pie - 1; ++i)
      EZ(i, j) = A2(s->Ezx_yn, i - pis, j - 1, hyx) +
                 A2(s->Ezy_yn, i - 1, j - 1, NX - 1);
  OMP_COLS
  for (int j = pje; j <= NY; ++j)
    for (int i = pis + 1; i <= pie - 1; ++i)
      EZ(i, j) = A2(s->Ezx_yp, i - pis, j - pje + 1, hyx) +
                 A2(s->Ezy_yp, i - 1, j - pje + 1, NX - 1);
  OMP_COLS
  for (int j = pjs + 1; j <= pje - 1; ++j)
    for (int i = 2; i <= pis; ++i)
      EZ(i, j) = A2(s->Ezx_xn, i - 1, j - 1, P) + A2(s->Ezy_xn, i - 1, j - pjs, P);
  OMP_COLS
  for (int j = pjs + 1; j <= pje - 1; ++j)
    for (int i = pie; i <= NX; ++i)
      EZ(i, j) = A2(s->Ezx_xp, i - pie + 1, j - 1, P) +
                 A2(s->Ezy_xp, i - pie + 1, j - pjs, P);
#undef EZ
#undef HX
#undef HY
#undef A2
#undef CC
}

/* ------------------------------------------------------------------------- */
/* public: w_exp2pml.m:17-37                                                  */
/* ------------------------------------------------------------------------- */
int FN(orc_exp2pml)(int m, int n, int P, int air, const double *eps_model,
                    const double *sig_model, REAL *eps_rz, REAL *sig_ez) {
  const int NX = m + 2 * P + air, NY = n + 2 * P;
  for (int j = 1; j <= NY; ++j) {
    int jm = j - P; /* model column, clamped = side replication */
    if (jm < 1) jm = 1;
    if (jm > n) jm = n;
    int in_model_col = (j - P >= 1 && j - P <= n);
    for (int i = 1; i <= NX; ++i) {
      /* rows of [air; model] before up/down padding */
      int r = i - P;
      if (r < 1) r = 1; /* "up" replicates the first row of [air; model] */
      if (r > air + m) r = air + m; /* "down" replicates the last row      */
      double e, s;
      if (r <= air) {
        e = 1.0;
        s = 0.0;
      } else {
        e = eps_model[(size_t)(r - air - 1) + (size_t)(jm - 1) * m];
        s = sig_model[(size_t)(r - air - 1) + (size_t)(jm - 1) * m];
      }
      /* sigma is zero-padded on every side (w_exp2pml.m:23,32) */
      int in_model_row = (i - P >= 1 && i - P <= air + m);
      if (!(in_model_col && in_model_row)) s = 0.0;
      size_t k = (size_t)(i - 1) + (size_t)(j - 1) * NX;
      eps_rz[k] = (REAL)e;
      sig_ez[k] = (REAL)s;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* public: coefficient export for closed-form checks                         */
/* which: 0 Ceze, 1 Cezhy, 2 Cezhx (NX x NY); 3 Chyh_xn .. (see oracle.py)   */
/* ------------------------------------------------------------------------- */
int FN(orc_coeff_export)(const orc_dims *g, const orc_phys *ph, const REAL *eps,
                         const REAL *sig, int si, int sj, int which,
                         REAL *out, double *scal) {
  TY(orc_coeffs) c;
  if (FN(coeffs_build)(g, ph, eps, sig, si, sj, 1.0, &c)) return -1;
  const size_t NX = g->NX, NY = g->NY, P = g->P;
  const size_t nxy = NX * NY, xs = P * (NY - 1), xw = P * (NY - 1 - 2 * P),
               yn = (NX - 1) * P, yh = (NX - 1 - 2 * P) * P;
  struct {
    REAL *p;
    size_t n;
  } tab[] = {{c.Ceze, nxy},      {c.Cezhy, nxy},     {c.Cezhx, nxy},
             {c.Chyh_xn, xs},    {c.Chyez_xn, xs},   {c.Cezxe_xn, xs},
             {c.Cezxhy_xn, xs},  {c.Chyh_xp, xs},    {c.Chyez_xp, xs},
             {c.Cezxe_xp, xs},   {c.Cezxhy_xp, xs},  {c.Cezyhx_xp, xw},
             {c.Chxh_yn, yn},    {c.Chxez_yn, yn},   {c.Cezxhy_yn, yh},
             {c.Cezye_yn, yn},   {c.Cezyhx_yn, yn},  {c.Chxh_yp, yn},
             {c.Chxez_yp, yn},   {c.Cezxhy_yp, yh},  {c.Cezye_yp, yn},
             {c.Cezyhx_yp, yn}};
  int ntab = (int)(sizeof(tab) / sizeof(tab[0]));
  if (which >= 0 && which < ntab && out)
    memcpy(out, tab[which].p, tab[which].n * sizeof(REAL));
  if (scal) {
    scal[0] = c.Chxh; scal[1] = c.Chxez; scal[2] = c.Chyh; scal[3] = c.Chyez;
    scal[4] = c.Cezj; scal[5] = c.Cezye_xn; scal[6] = c.Cezyhx_xn;
    scal[7] = c.Cezye_xp; scal[8] = c.Cezxe_yn; scal[9] = c.Cezxe_yp;
  }
  FN(coeffs_free)(&c);
  return (which >= ntab) ? -2 : 0;
}

/* ------------------------------------------------------------------------- */
/* public: w_solve.m (forward).  d is nt x nr (column-major, d[it + r*nt]);  */
/* Jz nt; u (optional) m x n x nt domain wavefield (w_solve.m:102).          */
/* ------------------------------------------------------------------------- */
int FN(orc_forward)(const orc_dims *g, const orc_phys *ph, const REAL *eps,
                    const REAL *sig, int si, int sj, const REAL *wvlet, int nr,
                    const int *Mw, REAL *d, REAL *Jz, REAL *u) {
  TY(orc_coeffs) c;
  TY(orc_state) s;
  if (FN(coeffs_build)(g, ph, eps, sig, si, sj, 1.0, &c)) return -1;
  if (FN(state_alloc)(g, &s)) return -1;
  const int nt = ph->nt;
  const size_t le = (size_t)g->NX + 1;
  const int di0 = g->pis + g->air, dj0 = g->pjs;
  for (int it = 1; it <= nt; ++it) {
    FN(step_literal)(g, &c, &s, si, sj, wvlet[it - 1], 1);
    for (int r = 0; r < nr; ++r) d[(size_t)(it - 1) + (size_t)r * nt] = s.Ez[Mw[r] - 1];
    if (u) {
      REAL *uk = u + (size_t)(it - 1) * g->m * g->n;
      for (int b = 0; b < g->n; ++b)
        for (int a = 0; a < g->m; ++a)
          uk[a + (size_t)b * g->m] = s.Ez[(size_t)(di0 + a - 1) + (size_t)(dj0 + b - 1) * le];
    }
    if (Jz) Jz[it - 1] = s.Ez[(size_t)(si - 1) + (size_t)(sj - 1) * le];
  }
  free(s.buf);
  FN(coeffs_free)(&c);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* public: differentiate_cube.m:23-51, in place on an m x n x nt cube        */
/* ------------------------------------------------------------------------- */
int FN(orc_differentiate_cube)(REAL *v, int nz, int nx, int nt, double dt) {
  const double a = -1.5 / dt, b = 2 / dt, c = -0.5 / dt, d = 2 / dt;
  const size_t fr = (size_t)nz * nx;
  if (nt < 3) return -1;
  REAL *v_ = (REAL *)malloc(fr * sizeof(REAL));
  REAL *v__ = (REAL *)calloc(fr, sizeof(REAL));
  if (!v_ || !v__) return -1;
#define V(k) (v + (size_t)((k)-1) * fr)
  memcpy(v_, V(1), fr * sizeof(REAL));
  for (size_t p = 0; p < fr; ++p)
    V(1)[p] = (REAL)(a * V(1)[p] + b * V(2)[p] + c * V(3)[p]);
  for (int it = 2; it <= nt - 2; it += 2) {
    for (size_t p = 0; p < fr; ++p) {
      v__[p] = V(it)[p];
      V(it)[p] = (REAL)(c * v_[p] - c * V(it + 1)[p]);
      v_[p] = V(it + 1)[p];
      V(it + 1)[p] = (REAL)(c * v__[p] - c * V(it + 2)[p]);
    }
  }
  if (nt % 2 == 1) {
    for (size_t p = 0; p < fr; ++p) {
      v__[p] = V(nt - 1)[p];
      V(nt - 1)[p] = (REAL)(c * v_[p] - c * V(nt)[p]);
      V(nt)[p] = (REAL)(-c * v_[p] - d * v__[p] - a * V(nt)[p]);
    }
  } else {
    for (size_t p = 0; p < fr; ++p)
      V(nt)[p] = (REAL)(-c * v__[p] - d * v_[p] - a * V(nt)[p]);
  }
#undef V
  free(v_);
  free(v__);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* public: w_grad.m (the reference's gradient, "REF").                        */
/* u is m x n x nt (u, or d_t u for the permittivity gradient); e nt x nr.   */
/* g (m x n) = dt * sum_it -u(:,:,nt-it) .* a ; amax = |a| of the last step. */
/* ------------------------------------------------------------------------- */
int FN(orc_grad_ref)(const orc_dims *g, const orc_phys *ph, const REAL *eps,
                     const REAL *sig, int si, int sj, int nr, const int *Mw,
                     const REAL *e, const REAL *u, REAL *gw, REAL *amax) {
  TY(orc_coeffs) c;
  TY(orc_state) s;
  if (FN(coeffs_build)(g, ph, eps, sig, si, sj, -1.0, &c)) return -1;
  if (FN(state_alloc)(g, &s)) return -1;
  const int nt = ph->nt, m = g->m, n = g->n;
  const size_t le = (size_t)g->NX + 1, fr = (size_t)m * n;
  const int di0 = g->pis + g->air, dj0 = g->pjs;
  for (size_t p = 0; p < fr; ++p) gw[p] = 0;
  for (int it = 0; it <= nt - 1; ++it) {
    for (int r = 0; r < nr; ++r) /* hard injection, w_grad.m:72 */
      s.Ez[Mw[r] - 1] = e[(size_t)(nt - it - 1) + (size_t)r * nt];
    FN(step_literal)(g, &c, &s, si, sj, 0, 0);
    const REAL *uk = u + (size_t)(nt - it - 1) * fr;
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < m; ++a) {
        REAL aw = s.Ez[(size_t)(di0 + a - 1) + (size_t)(dj0 + b - 1) * le];
        gw[a + (size_t)b * m] = gw[a + (size_t)b * m] - uk[a + (size_t)b * m] * aw;
        if (it == nt - 1) amax[a + (size_t)b * m] = (REAL)fabs((double)aw);
      }
  }
  for (size_t p = 0; p < fr; ++p) gw[p] = (REAL)(ph->dt * gw[p]);
  free(s.buf);
  FN(coeffs_free)(&c);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Exact discrete adjoint ("EXACT").                                          */
/*                                                                            */
/* J(eps, sig) is any functional of the traces d_n(r) with dJ/dd_n(r) = e_n(r)*/
/* (for the L2 misfit of w_bozdag.m:13-15, e = d - d_o).  The adjoint below is */
/* the transpose of step_literal, written per cell from coefficient maps     */
/* that are read out of the reference-shaped arrays.  The gradient is taken  */
/* with respect to every coefficient slot and chained through the closed-form */
/* coefficient formulas of w_coeff.m / w_cpml.m to the padded NX x NY eps and */
/* sigma; orc_fold_pad then applies the w_exp2pml.m replication chain rule.   */
/* ------------------------------------------------------------------------- */
typedef struct {
  /* Hx cells (NX+1) x NY, Hy cells NX x (NY+1) */
  REAL *AX, *BX, *AY, *BY;
  double *dAX, *dBX, *dAY, *dBY; /* d/d eps at the same index (strips only) */
  /* Ez cells (NX+1) x (NY+1) */
  unsigned char *kind; /* 0 never written, 1 inner, 2 split-field PML */
  REAL *CE, *CH, *CHX, *CXE, *CXH, *CYE, *CYH;
  double *dCE_e, *dCH_e, *dCHX_e, *dCE_s, *dCH_s, *dCHX_s;
  double *dCXE, *dCXH, *dCYE, *dCYH; /* d/d eps */
  int *xid, *xoff, *yid, *yoff; /* which reference aux array holds Ezx/Ezy  */
  double dCJ_e, dCJ_s;
} TY(orc_cellmap);

static void FN(cellmap_free)(TY(orc_cellmap) * q) {
  free(q->AX); free(q->BX); free(q->AY); free(q->BY);
  free(q->dAX); free(q->dBX); free(q->dAY); free(q->dBY);
  free(q->kind); free(q->CE); free(q->CH); free(q->CHX); free(q->CXE);
  free(q->CXH); free(q->CYE); free(q->CYH);
  free(q->dCE_e); free(q->dCH_e); free(q->dCHX_e); free(q->dCE_s);
  free(q->dCH_s); free(q->dCHX_s); free(q->dCXE); free(q->dCXH);
  free(q->dCYE); free(q->dCYH); free(q->xid); free(q->xoff); free(q->yid);
  free(q->yoff);
}

static int FN(cellmap_build)(const orc_dims *g, const orc_phys *ph,
                             const TY(orc_coeffs) * c, const REAL *eps,
                             const REAL *sig, int si, int sj,
                             TY(orc_cellmap) * q) {
  memset(q, 0, sizeof(*q));
  const int NX = g->NX, NY = g->NY, P = g->P;
  const int pis = g->pis, pie = g->pie, pjs = g->pjs, pje = g->pje;
  const size_t le = (size_t)NX + 1, lh = (size_t)NX;
  const size_t nE = le * (NY + 1), nHx = le * NY, nHy = lh * (NY + 1);
  const double e0 = ph->eps0, u0 = ph->mu0, dt = ph->dt, dx = ph->dx, dy = ph->dx;
#define CALLOC(p, cnt, T)                          \
  do {                                             \
    (p) = (T *)calloc((cnt), sizeof(T));           \
    if (!(p)) return -1;                           \
  } while (0)
  CALLOC(q->AX, nHx, REAL); CALLOC(q->BX, nHx, REAL);
  CALLOC(q->AY, nHy, REAL); CALLOC(q->BY, nHy, REAL);
  CALLOC(q->dAX, nHx, double); CALLOC(q->dBX, nHx, double);
  CALLOC(q->dAY, nHy, double); CALLOC(q->dBY, nHy, double);
  CALLOC(q->kind, nE, unsigned char);
  CALLOC(q->CE, nE, REAL); CALLOC(q->CH, nE, REAL); CALLOC(q->CHX, nE, REAL);
  CALLOC(q->CXE, nE, REAL); CALLOC(q->CXH, nE, REAL);
  CALLOC(q->CYE, nE, REAL); CALLOC(q->CYH, nE, REAL);
  CALLOC(q->dCE_e, nE, double); CALLOC(q->dCH_e, nE, double);
  CALLOC(q->dCHX_e, nE, double); CALLOC(q->dCE_s, nE, double);
  CALLOC(q->dCH_s, nE, double); CALLOC(q->dCHX_s, nE, double);
  CALLOC(q->dCXE, nE, double); CALLOC(q->dCXH, nE, double);
  CALLOC(q->dCYE, nE, double); CALLOC(q->dCYH, nE, double);
  CALLOC(q->xid, nE, int); CALLOC(q->xoff, nE, int);
  CALLOC(q->yid, nE, int); CALLOC(q->yoff, nE, int);
#undef CALLOC
#define EPS(i, j) ((double)eps[((i)-1) + (size_t)((j)-1) * NX])
#define SIG(i, j) ((double)sig[((i)-1) + (size_t)((j)-1) * NX])
#define IE(i, j) ((size_t)((i)-1) + (size_t)((j)-1) * le)
#define IHX(i, j) ((size_t)((i)-1) + (size_t)((j)-1) * le)
#define IHY(i, j) ((size_t)((i)-1) + (size_t)((j)-1) * lh)
#define A2(A, a, b, ld) (A)[((a)-1) + (size_t)((b)-1) * (size_t)(ld)]
  /* default: identity H cells */
  for (size_t k = 0; k < nHx; ++k) q->AX[k] = 1;
  for (size_t k = 0; k < nHy; ++k) q->AY[k] = 1;
  /* inner H (w_solve.m:144-147) */
  for (int j = pjs; j <= pje - 1; ++j)
    for (int i = 1; i <= NX + 1; ++i) {
      q->AX[IHX(i, j)] = c->Chxh;
      q->BX[IHX(i, j)] = c->Chxez;
    }
  for (int j = 1; j <= NY + 1; ++j)
    for (int i = pis; i <= pie - 1; ++i) {
      q->AY[IHY(i, j)] = c->Chyh;
      q->BY[IHY(i, j)] = c->Chyez;
    }
  /* H strips (w_solve.m:170-186) and their eps derivatives.
   * A = (2u0 - dt s)/(2u0 + dt s), B = +-(2dt/h)/(2u0 + dt s), s = sm/eps. */
  for (int b = 1; b <= NY - 1; ++b)
    for (int a = 1; a <= P; ++a) {
      int i = a, j = b + 1;
      q->AY[IHY(i, j)] = A2(c->Chyh_xn, a, b, P);
      q->BY[IHY(i, j)] = A2(c->Chyez_xn, a, b, P);
      i = pie - 1 + a;
      q->AY[IHY(i, j)] = A2(c->Chyh_xp, a, b, P);
      q->BY[IHY(i, j)] = A2(c->Chyez_xp, a, b, P);
      double ee = EPS(i, j), s = c->spmx_xp[a - 1] / ee, dm = 2 * u0 + dt * s;
      double dsde = -s / ee;
      q->dAY[IHY(i, j)] = (-4 * u0 * dt / (dm * dm)) * dsde;
      q->dBY[IHY(i, j)] = (-(2 * dt / dx) * dt / (dm * dm)) * dsde;
    }
  for (int b = 1; b <= P; ++b)
    for (int a = 1; a <= NX - 1; ++a) {
      int i = a + 1, j = b;
      q->AX[IHX(i, j)] = A2(c->Chxh_yn, a, b, NX - 1);
      q->BX[IHX(i, j)] = A2(c->Chxez_yn, a, b, NX - 1);
      double ee = EPS(i, j), s = c->spmy_yn[b - 1] / ee, dm = 2 * u0 + dt * s;
      double dsde = -s / ee;
      q->dAX[IHX(i, j)] = (-4 * u0 * dt / (dm * dm)) * dsde;
      q->dBX[IHX(i, j)] = ((2 * dt / dy) * dt / (dm * dm)) * dsde;
      j = pje - 1 + b;
      q->AX[IHX(i, j)] = A2(c->Chxh_yp, a, b, NX - 1);
      q->BX[IHX(i, j)] = A2(c->Chxez_yp, a, b, NX - 1);
      ee = EPS(i, j); s = c->spmy_yp[b - 1] / ee; dm = 2 * u0 + dt * s;
      dsde = -s / ee;
      q->dAX[IHX(i, j)] = (-4 * u0 * dt / (dm * dm)) * dsde;
      q->dBX[IHX(i, j)] = ((2 * dt / dy) * dt / (dm * dm)) * dsde;
    }
  /* Ez cells */
  for (int j = 1; j <= NY + 1; ++j)
    for (int i = 1; i <= NX + 1; ++i) {
      size_t k = IE(i, j);
      if (i == 1 || j == 1 || i == NX + 1 || j == NY + 1) {
        q->kind[k] = 0;
        continue;
      }
      int xrow = (i <= pis) || (i >= pie);
      int ycol = (j <= pjs) || (j >= pje);
      if (!xrow && !ycol) {
        q->kind[k] = 1;
        double ee = EPS(i, j), ss = SIG(i, j), D = 2 * ee * e0 + dt * ss;
        q->CE[k] = A2(c->Ceze, i, j, NX);
        q->CH[k] = A2(c->Cezhy, i, j, NX);
        q->CHX[k] = A2(c->Cezhx, i, j, NX);
        q->dCE_e[k] = 4 * e0 * dt * ss / (D * D);
        q->dCE_s[k] = -4 * e0 * ee * dt / (D * D);
        q->dCH_e[k] = -(2 * dt / dx) * 2 * e0 / (D * D);
        q->dCH_s[k] = -(2 * dt / dx) * dt / (D * D);
        q->dCHX_e[k] = (2 * dt / dy) * 2 * e0 / (D * D);
        q->dCHX_s[k] = (2 * dt / dy) * dt / (D * D);
        continue;
      }
      q->kind[k] = 2;
      double ee = EPS(i, j);
      /* Ezx slot: x strip by row, else y strip (w_solve.m:260-267) */
      if (i <= pis) {
        int a = i - 1, b = j - 1;
        q->xid[k] = 0; q->xoff[k] = (a - 1) + (b - 1) * P;
        q->CXE[k] = A2(c->Cezxe_xn, a, b, P);
        q->CXH[k] = A2(c->Cezxhy_xn, a, b, P);
      } else if (i >= pie) {
        int a = i - pie + 1, b = j - 1;
        q->xid[k] = 1; q->xoff[k] = (a - 1) + (b - 1) * P;
        q->CXE[k] = A2(c->Cezxe_xp, a, b, P);
        q->CXH[k] = A2(c->Cezxhy_xp, a, b, P);
        double se = c->spex_xp[a - 1], D = 2 * e0 * ee + dt * se;
        q->dCXE[k] = 4 * e0 * dt * se / (D * D);
        q->dCXH[k] = -(2 * dt / dx) * 2 * e0 / (D * D);
      } else if (j <= pjs) {
        int a = i - pis, b = j - 1, hyx = NX - 1 - 2 * P;
        q->xid[k] = 2; q->xoff[k] = (a - 1) + (b - 1) * hyx;
        q->CXE[k] = c->Cezxe_yn;
        q->CXH[k] = A2(c->Cezxhy_yn, a, b, hyx);
        q->dCXH[k] = -dt / (dx * e0 * ee * ee);
      } else {
        int a = i - pis, b = j - pje + 1, hyx = NX - 1 - 2 * P;
        q->xid[k] = 3; q->xoff[k] = (a - 1) + (b - 1) * hyx;
        q->CXE[k] = c->Cezxe_yp;
        q->CXH[k] = A2(c->Cezxhy_yp, a, b, hyx);
        q->dCXH[k] = -dt / (dx * e0 * ee * ee);
      }
      /* Ezy slot: y strip by column, else x strip */
      if (j <= pjs) {
        int a = i - 1, b = j - 1;
        q->yid[k] = 6; q->yoff[k] = (a - 1) + (b - 1) * (NX - 1);
        q->CYE[k] = A2(c->Cezye_yn, a, b, NX - 1);
        q->CYH[k] = A2(c->Cezyhx_yn, a, b, NX - 1);
        double se = c->spey_yn[b - 1], D = 2 * e0 * ee + dt * se;
        q->dCYE[k] = 4 * e0 * dt * se / (D * D);
        q->dCYH[k] = (2 * dt / dy) * 2 * e0 / (D * D);
      } else if (j >= pje) {
        int a = i - 1, b = j - pje + 1;
        q->yid[k] = 7; q->yoff[k] = (a - 1) + (b - 1) * (NX - 1);
        q->CYE[k] = A2(c->Cezye_yp, a, b, NX - 1);
        q->CYH[k] = A2(c->Cezyhx_yp, a, b, NX - 1);
        double se = c->spey_yp[b - 1], D = 2 * e0 * ee + dt * se;
        q->dCYE[k] = 4 * e0 * dt * se / (D * D);
        q->dCYH[k] = (2 * dt / dy) * 2 * e0 / (D * D);
      } else if (i <= pis) {
        int a = i - 1, b = j - pjs;
        q->yid[k] = 4; q->yoff[k] = (a - 1) + (b - 1) * P;
        q->CYE[k] = c->Cezye_xn;
        q->CYH[k] = c->Cezyhx_xn;
      } else {
        int a = i - pie + 1, b = j - pjs;
        q->yid[k] = 5; q->yoff[k] = (a - 1) + (b - 1) * P;
        q->CYE[k] = c->Cezye_xp;
        q->CYH[k] = A2(c->Cezyhx_xp, a, b, P);
        q->dCYH[k] = dt / (dy * e0 * ee * ee);
      }
    }
  if (si >= 1 && si <= NX && sj >= 1 && sj <= NY) {
    double ee = EPS(si, sj), ss = SIG(si, sj), D = 2 * ee * e0 + dt * ss;
    q->dCJ_e = 2 * dt * 2 * e0 / (D * D);
    q->dCJ_s = 2 * dt * dt / (D * D);
  }
#undef EPS
#undef SIG
#undef A2
  (void)u0;
  return 0;
}

static REAL *FN(aux_ptr)(TY(orc_state) * s, int id) {
  switch (id) {
    case 0: return s->Ezx_xn;
    case 1: return s->Ezx_xp;
    case 2: return s->Ezx_yn;
    case 3: return s->Ezx_yp;
    case 4: return s->Ezy_xn;
    case 5: return s->Ezy_xp;
    case 6: return s->Ezy_yn;
    default: return s->Ezy_yp;
  }
}

/* Round a checkpoint state as a 16-bit snapshot (snap_round), with one
 * scale per block: a grid column of Ez, Hx or Hy, or 256 consecutive
 * split-field cells (column-major cell order) of the Ezx / Ezy values the
 * assembly reads (the other aux entries never reach Ez). */
static void FN(snap_state)(const orc_dims *g, const TY(orc_cellmap) * q,
                           TY(orc_state) * s, int mode) {
  const int NX = g->NX, NY = g->NY;
  const size_t nE = (size_t)(NX + 1) * (NY + 1);
  REAL *pl[3] = {s->Ez, s->Hx, s->Hy};
  const size_t rows[3] = {(size_t)NX + 1, (size_t)NX + 1, (size_t)NX};
  const int cols[3] = {NY + 1, NY, NY + 1};
  for (int p = 0; p < 3; ++p)
    for (int j = 0; j < cols[p]; ++j) {
      REAL *c = pl[p] + (size_t)j * rows[p];
      double amax = 0;
      for (size_t i = 0; i < rows[p]; ++i) amax = fmax(amax, fabs((double)c[i]));
      const int sh = snap_shift(amax, mode);
      for (size_t i = 0; i < rows[p]; ++i) c[i] = (REAL)snap_round((double)c[i], sh, mode);
    }
  REAL **cell = (REAL **)malloc(nE * sizeof(REAL *));
  if (!cell) return;
  for (int w = 0; w < 2; ++w) {
    size_t n = 0;
    for (size_t k = 0; k < nE; ++k)
      if (q->kind[k] == 2)
        cell[n++] = FN(aux_ptr)(s, w ? q->yid[k] : q->xid[k]) + (w ? q->yoff[k] : q->xoff[k]);
    for (size_t b0 = 0; b0 < n; b0 += 256) {
      const size_t b1 = b0 + 256 < n ? b0 + 256 : n;
      double amax = 0;
      for (size_t t = b0; t < b1; ++t) amax = fmax(amax, fabs((double)*cell[t]));
      const int sh = snap_shift(amax, mode);
      for (size_t t = b0; t < b1; ++t) *cell[t] = (REAL)snap_round((double)*cell[t], sh, mode);
    }
  }
  free(cell);
}

/* g_eps_pad, g_sig_pad: NX x NY; g_w: nt (dJ/dw, the source-wavelet
 * gradient); any may be NULL.  pml_terms = 0 drops the strip-coefficient
 * eps terms (the strips then act as a fixed absorber). K = recompute
 * segment length (checkpoint interval).  snap = 1 / 2 stores the checkpoints
 * as BF16 / FP16 snapshots (snap_state), 0 keeps them exact. */
int FN(orc_grad_exact)(const orc_dims *g, const orc_phys *ph, const REAL *eps,
                       const REAL *sig, int si, int sj, const REAL *wvlet,
                       int nr, const int *Mw, const REAL *e, int K,
                       int pml_terms, double *g_eps_pad, double *g_sig_pad,
                       double *g_w, int snap) {
  const int NX = g->NX, NY = g->NY, nt = ph->nt;
  const size_t le = (size_t)NX + 1, lh = (size_t)NX;
  const size_t nE = le * (NY + 1), nHx = le * NY, nHy = lh * (NY + 1);
  if (si <= 1 || sj <= 1 || si > NX || sj > NY) return -3; /* ring source */
  if (K < 1) K = 1;
  const int par = (size_t)(NX + 1) * (NY + 1) >= ORC_PAR_CELLS && !omp_in_parallel();
  TY(orc_coeffs) c;
  TY(orc_cellmap) q;
  TY(orc_state) s0;
  if (FN(coeffs_build)(g, ph, eps, sig, si, sj, 1.0, &c)) return -1;
  if (FN(cellmap_build)(g, ph, &c, eps, sig, si, sj, &q)) return -1;
  if (FN(state_alloc)(g, &s0)) return -1;
  const size_t SL = s0.len;
  const int nck = (nt + K - 1) / K; /* checkpoints at steps 0, K, 2K, ... */
  REAL *ck = (REAL *)calloc((size_t)nck * SL, sizeof(REAL));
  REAL *seg = (REAL *)calloc((size_t)(K + 1) * SL, sizeof(REAL));
  double *mE = (double *)calloc(nE, sizeof(double));
  double *nE_ = (double *)calloc(nE, sizeof(double));
  double *mX = (double *)calloc(nE, sizeof(double));
  double *mY = (double *)calloc(nE, sizeof(double));
  double *WX = (double *)calloc(nE, sizeof(double));
  double *WY = (double *)calloc(nE, sizeof(double));
  double *nHx_ = (double *)calloc(nHx, sizeof(double));
  double *nHy_ = (double *)calloc(nHy, sizeof(double));
  if (!ck || !seg || !mE || !nE_ || !mX || !mY || !WX || !WY || !nHx_ || !nHy_)
    return -1;
  if (g_eps_pad) memset(g_eps_pad, 0, sizeof(double) * NX * NY);
  if (g_sig_pad) memset(g_sig_pad, 0, sizeof(double) * NX * NY);
  if (g_w) memset(g_w, 0, sizeof(double) * nt);
  /* forward sweep, keeping checkpoints */
  for (int it = 0; it < nt; ++it) {
    if (it % K == 0) {
      orc_par_copy(ck + (size_t)(it / K) * SL, s0.buf, SL * sizeof(REAL), par);
      if (snap) {
        TY(orc_state) cv;
        FN(state_rebase)(&s0, ck + (size_t)(it / K) * SL, &cv);
        FN(snap_state)(g, &q, &cv, snap);
      }
    }
    FN(step_literal)(g, &c, &s0, si, sj, wvlet[it], 1);
  }
  const int sk = (int)IE(si, sj);
#define GE(i, j) g_eps_pad[((i)-1) + (size_t)((j)-1) * NX]
#define GS(i, j) g_sig_pad[((i)-1) + (size_t)((j)-1) * NX]
  for (int cidx = nck - 1; cidx >= 0; --cidx) {
    const int k0 = cidx * K, k1 = (k0 + K < nt) ? k0 + K : nt;
    /* recompute states k0..k1 into seg */
    orc_par_copy(seg, ck + (size_t)cidx * SL, SL * sizeof(REAL), par);
    for (int k = k0 + 1; k <= k1; ++k) {
      orc_par_copy(seg + (size_t)(k - k0) * SL, seg + (size_t)(k - 1 - k0) * SL,
                   SL * sizeof(REAL), par);
      TY(orc_state) sv;
      FN(state_rebase)(&s0, seg + (size_t)(k - k0) * SL, &sv);
      FN(step_literal)(g, &c, &sv, si, sj, wvlet[k - 1], 1);
    }
    for (int nstep = k1; nstep > k0; --nstep) {
      TY(orc_state) Sn, Sp; /* state after step nstep, and before it */
      FN(state_rebase)(&s0, seg + (size_t)(nstep - k0) * SL, &Sn);
      FN(state_rebase)(&s0, seg + (size_t)(nstep - 1 - k0) * SL, &Sp);
      const REAL *Ezp = Sp.Ez, *Hxn = Sn.Hx, *Hyn = Sn.Hy, *Hxp = Sp.Hx,
                 *Hyp = Sp.Hy;
      /* 1. data adjoint at step nstep */
      for (int r = 0; r < nr; ++r)
        mE[Mw[r] - 1] += (double)e[(size_t)(nstep - 1) + (size_t)r * nt];
      /* 2. source term (source is added to an inner cell only; a source in
       *    the split-field strip is overwritten, w_solve.m:92-94) */
      if (q.kind[sk] == 1) {
        if (g_w) g_w[nstep - 1] = (double)c.Cezj * mE[sk];
        double gcj = mE[sk] * (double)wvlet[nstep - 1];
        if (g_eps_pad) GE(si, sj) += gcj * q.dCJ_e;
        if (g_sig_pad) GS(si, sj) += gcj * q.dCJ_s;
      }
      /* 3. E* pass with the E-coefficient gradient */
      OMP_COLS
      for (int j = 1; j <= NY + 1; ++j)
        for (int i = 1; i <= NX + 1; ++i) {
          size_t k = IE(i, j);
          double me = mE[k];
          if (q.kind[k] == 1) {
            double dhy = (double)Hyn[IHY(i, j)] - (double)Hyn[IHY(i - 1, j)];
            double dhx = (double)Hxn[IHX(i, j)] - (double)Hxn[IHX(i, j - 1)];
            double gce = me * (double)Ezp[k], gch = me * dhy, ghx = me * dhx;
            if (g_eps_pad)
              GE(i, j) += gce * q.dCE_e[k] + gch * q.dCH_e[k] + ghx * q.dCHX_e[k];
            if (g_sig_pad)
              GS(i, j) += gce * q.dCE_s[k] + gch * q.dCH_s[k] + ghx * q.dCHX_s[k];
            WY[k] = (double)q.CH[k] * me;
            WX[k] = (double)q.CHX[k] * me;
            nE_[k] = (double)q.CE[k] * me;
          } else if (q.kind[k] == 2) {
            double lx = mX[k] + me, ly = mY[k] + me;
            if (g_eps_pad && pml_terms) {
              double xo = (double)FN(aux_ptr)(&Sp, q.xid[k])[q.xoff[k]];
              double yo = (double)FN(aux_ptr)(&Sp, q.yid[k])[q.yoff[k]];
              double dhy = (double)Hyn[IHY(i, j)] - (double)Hyn[IHY(i - 1, j)];
              double dhx = (double)Hxn[IHX(i, j)] - (double)Hxn[IHX(i, j - 1)];
              GE(i, j) += lx * xo * q.dCXE[k] + lx * dhy * q.dCXH[k] +
                          ly * yo * q.dCYE[k] + ly * dhx * q.dCYH[k];
            }
            WY[k] = (double)q.CXH[k] * lx;
            WX[k] = (double)q.CYH[k] * ly;
            mX[k] = (double)q.CXE[k] * lx;
            mY[k] = (double)q.CYE[k] * ly;
            nE_[k] = 0.0;
          } else {
            WY[k] = 0.0;
            WX[k] = 0.0;
            nE_[k] = me;
          }
        }
      /* 4. H* pass: lambda_H' = A * lambda_H'(next) + curl^T W */
      OMP_COLS
      for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX + 1; ++i) {
          size_t k = IHX(i, j);
          double w0 = WX[IE(i, j)], w1 = WX[IE(i, j + 1)];
          nHx_[k] = (double)q.AX[k] * nHx_[k] + w0 - w1;
          if (g_eps_pad && pml_terms && (q.dAX[k] != 0 || q.dBX[k] != 0)) {
            double dez = (double)Ezp[IE(i, j + 1)] - (double)Ezp[IE(i, j)];
            GE(i, j) += nHx_[k] * ((double)Hxp[k] * q.dAX[k] + dez * q.dBX[k]);
          }
        }
      OMP_COLS
      for (int j = 1; j <= NY + 1; ++j)
        for (int i = 1; i <= NX; ++i) {
          size_t k = IHY(i, j);
          double w0 = WY[IE(i, j)], w1 = WY[IE(i + 1, j)];
          nHy_[k] = (double)q.AY[k] * nHy_[k] + w0 - w1;
          if (g_eps_pad && pml_terms && (q.dAY[k] != 0 || q.dBY[k] != 0)) {
            double dez = (double)Ezp[IE(i + 1, j)] - (double)Ezp[IE(i, j)];
            GE(i, j) += nHy_[k] * ((double)Hyp[k] * q.dAY[k] + dez * q.dBY[k]);
          }
        }
      /* 5. E-final: mE(n-1) = (CE mE | 0 | mE) + grad^T (B lambda_H') */
      OMP_COLS
      for (int j = 1; j <= NY + 1; ++j)
        for (int i = 1; i <= NX + 1; ++i) {
          size_t k = IE(i, j);
          double v = nE_[k];
          if (i >= 2) v += (double)q.BY[IHY(i - 1, j)] * nHy_[IHY(i - 1, j)];
          if (i <= NX) v -= (double)q.BY[IHY(i, j)] * nHy_[IHY(i, j)];
          if (j >= 2) v += (double)q.BX[IHX(i, j - 1)] * nHx_[IHX(i, j - 1)];
          if (j <= NY) v -= (double)q.BX[IHX(i, j)] * nHx_[IHX(i, j)];
          mE[k] = v;
        }
    }
  }
#undef GE
#undef GS
  free(ck); free(seg); free(mE); free(nE_); free(mX); free(mY); free(WX);
  free(WY); free(nHx_); free(nHy_); free(s0.buf);
  FN(cellmap_free)(&q);
  FN(coeffs_free)(&c);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* per-cell forward from the coefficient maps; must agree with the literal   */
/* solver bit for bit (the GPU kernels use the same per-cell form)           */
/* ------------------------------------------------------------------------- */
int FN(orc_forward_percell)(const orc_dims *g, const orc_phys *ph,
                            const REAL *eps, const REAL *sig, int si, int sj,
                            const REAL *wvlet, int nr, const int *Mw, REAL *d) {
  const int NX = g->NX, NY = g->NY, nt = ph->nt;
  const size_t le = (size_t)NX + 1, lh = (size_t)NX;
  const size_t nE = le * (NY + 1), nHx = le * NY, nHy = lh * (NY + 1);
  TY(orc_coeffs) c;
  TY(orc_cellmap) q;
  if (FN(coeffs_build)(g, ph, eps, sig, si, sj, 1.0, &c)) return -1;
  if (FN(cellmap_build)(g, ph, &c, eps, sig, si, sj, &q)) return -1;
  REAL *Ez = (REAL *)calloc(nE, sizeof(REAL)), *Hx = (REAL *)calloc(nHx, sizeof(REAL)),
       *Hy = (REAL *)calloc(nHy, sizeof(REAL)), *X = (REAL *)calloc(nE, sizeof(REAL)),
       *Y = (REAL *)calloc(nE, sizeof(REAL));
  if (!Ez || !Hx || !Hy || !X || !Y) return -1;
  for (int it = 1; it <= nt; ++it) {
    for (int j = 1; j <= NY; ++j)
      for (int i = 1; i <= NX + 1; ++i) {
        size_t k = IHX(i, j);
        Hx[k] = q.AX[k] * Hx[k] + q.BX[k] * (Ez[IE(i, j + 1)] - Ez[IE(i, j)]);
      }
    for (int j = 1; j <= NY + 1; ++j)
      for (int i = 1; i <= NX; ++i) {
        size_t k = IHY(i, j);
        Hy[k] = q.AY[k] * Hy[k] + q.BY[k] * (Ez[IE(i + 1, j)] - Ez[IE(i, j)]);
      }
    for (int j = 1; j <= NY + 1; ++j)
      for (int i = 1; i <= NX + 1; ++i) {
        size_t k = IE(i, j);
        if (q.kind[k] == 1) {
          Ez[k] = q.CE[k] * Ez[k] + q.CH[k] * (Hy[IHY(i, j)] - Hy[IHY(i - 1, j)]) +
                  q.CHX[k] * (Hx[IHX(i, j)] - Hx[IHX(i, j - 1)]);
          if (i == si && j == sj) Ez[k] = Ez[k] + c.Cezj * wvlet[it - 1];
        } else if (q.kind[k] == 2) {
          X[k] = q.CXE[k] * X[k] + q.CXH[k] * (Hy[IHY(i, j)] - Hy[IHY(i - 1, j)]);
          Y[k] = q.CYE[k] * Y[k] + q.CYH[k] * (Hx[IHX(i, j)] - Hx[IHX(i, j - 1)]);
          Ez[k] = X[k] + Y[k];
        }
      }
    for (int r = 0; r < nr; ++r) d[(size_t)(it - 1) + (size_t)r * nt] = Ez[Mw[r] - 1];
  }
  free(Ez); free(Hx); free(Hy); free(X); free(Y);
  FN(cellmap_free)(&q);
  FN(coeffs_free)(&c);
  return 0;
}

#undef IE
#undef IHX
#undef IHY
#undef FN
#undef TY
