// gjw_api.hip -- host driver and C ABI of libgjw (declared in include/gjw.h).
//
// Owns the per-survey device memory, builds the coefficient arrays from the
// model (src/w/w_exp2pml.m, w_cpml.m, w_coeff.m restated in FP64 on the host),
// and sequences the kernels of gjw_kernels.hip:
//   forward:  nt x k_fwd_step (all shots per launch), checkpoint every K steps
//   gradient: per segment (last to first) restore checkpoint, recompute K
//             steps storing Ez frames, K x k_adj_step, flush FP32 partial
//             sums of the correlation into the FP64 gradient.
// The plain path runs the same kernels one shot at a time and stores the
// whole domain wavefield, as src/w/w_solve.m:76,102 does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hipfft/hipfft.h>

#include "../../include/gjw.h"
#include "gjw_internal.h"

namespace gjw {
// kernels (gjw_kernels.hip)
__global__ void k_fwd_step(DevGrid, DevShots, int, const float*, const float*, const float*,
                           float*, float*, float*, float*, float*, int, float*, float*,
                           long long);
__global__ void k_adj_step(DevGrid, DevShots, int, const float*, const float*, const float*,
                           const float*, const float*, float*, float*, float*, float*, float*,
                           int, const float*, const float*, const float*, long long, float*,
                           long long, float*);
#define GJW_ADJN_ARGS DevGrid, DevShots, int, const float*, const float*, const float*, \
    const float*, const float*, float*, float*, float*, float*, float*, int, const float*, \
    const float*, long long, float*, long long, float*
// FR: frames FP32 (1) or BF16 (2)
template <int FR>
__global__ void k_adj_step2(GJW_ADJN_ARGS);
template <int FR>
__global__ void k_adj_step4(GJW_ADJN_ARGS);
template <int FR>
__global__ void k_adj_step8(GJW_ADJN_ARGS);
template <int FR>
__global__ void k_adj_step6(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step2<1>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step2<2>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step4<1>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step4<2>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step6<1>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step6<2>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step8<1>(GJW_ADJN_ARGS);
extern template __global__ void k_adj_step8<2>(GJW_ADJN_ARGS);
#undef GJW_ADJN_ARGS
#define GJW_FWDN_ARGS DevGrid, DevShots, int, const float*, const float*, const float*, \
    const float*, const float*, float*, float*, float*, float*, float*, int, float*, float*, long long
// FR: no frames (0), FP32 frames (1), BF16 frames (2)
template <int FR>
__global__ void k_fwd_step2(GJW_FWDN_ARGS);
template <int FR>
__global__ void k_fwd_step4(GJW_FWDN_ARGS);
template <int FR>
__global__ void k_fwd_step8(GJW_FWDN_ARGS, RowPairs);
extern template __global__ void k_fwd_step2<0>(GJW_FWDN_ARGS);
extern template __global__ void k_fwd_step2<1>(GJW_FWDN_ARGS);
extern template __global__ void k_fwd_step2<2>(GJW_FWDN_ARGS);
extern template __global__ void k_fwd_step4<0>(GJW_FWDN_ARGS);
extern template __global__ void k_fwd_step4<1>(GJW_FWDN_ARGS);
extern template __global__ void k_fwd_step4<2>(GJW_FWDN_ARGS);
extern template __global__ void k_fwd_step8<0>(GJW_FWDN_ARGS, RowPairs);
extern template __global__ void k_fwd_step8<1>(GJW_FWDN_ARGS, RowPairs);
extern template __global__ void k_fwd_step8<2>(GJW_FWDN_ARGS, RowPairs);
#undef GJW_FWDN_ARGS
__global__ void k_fwd_step_multi(const MultiTable, MultiArgs);
template <int NS, bool FRAMES>
__global__ void k_fwd_stepN_multi(const MultiTable, MultiArgs);
extern template __global__ void k_fwd_stepN_multi<2, false>(const MultiTable, MultiArgs);
extern template __global__ void k_fwd_stepN_multi<2, true>(const MultiTable, MultiArgs);
extern template __global__ void k_fwd_stepN_multi<4, false>(const MultiTable, MultiArgs);
extern template __global__ void k_fwd_stepN_multi<4, true>(const MultiTable, MultiArgs);
template <bool FRAMES>
__global__ void k_fwd_step8_multi(const MultiTable, MultiArgs);
extern template __global__ void k_fwd_step8_multi<false>(const MultiTable, MultiArgs);
extern template __global__ void k_fwd_step8_multi<true>(const MultiTable, MultiArgs);
__global__ void k_adj_step_multi(const MultiTable, MultiArgs);
template <int NS>
__global__ void k_adj_stepN_multi(const MultiTable, MultiArgs);
extern template __global__ void k_adj_stepN_multi<2>(const MultiTable, MultiArgs);
extern template __global__ void k_adj_stepN_multi<4>(const MultiTable, MultiArgs);
extern template __global__ void k_adj_stepN_multi<6>(const MultiTable, MultiArgs);
__global__ void k_fill(float*, long long, float);
__global__ void k_fill_d(double*, long long, double);
__global__ void k_inject_last(DevGrid, int, int, int, const int*, const int*, const int*,
                              const float*, int, float*);
__global__ void k_gather(const float*, long long, const int*, int, int, float*);
__global__ void k_pack16(const float*, long long, const int*, long long, int, int, int*,
                         unsigned short*);
__global__ void k_unpack16(const unsigned short*, long long, int, const int*, int, const int*,
                           long long, float*);
__global__ void k_scatter(const float*, long long, const int*, int, int, float*);
__global__ void k_domain_frame(DevGrid, int, const float*, float*, long long, int);
__global__ void k_flush(int, int, int, int, int, float*, long long, const double*, const double*,
                        double*, double*, double, long long);
__global__ void k_ref_overwrite(DevGrid, int, int, int, const int*, const int*, const float*, int,
                                float*);
__global__ void k_ref_corr(DevGrid, const float*, const float*, int, int, double, float*, float*);
__global__ void k_resid_l2(const float*, const float*, float*, long long, double*);
__global__ void k_analytic_spec(const float2*, float2*, int, long long);
__global__ void k_obs_instant(const float2*, long long, int, int, float*);
__global__ void k_misfit_stage1(const float*, const float2*, const float*, long long, int, int,
                                float*, float*, double*);
__global__ void k_misfit_stage2(const float*, const float2*, long long, int, int, float*);
__global__ void k_reduce_d(const double*, int, double, double*);
__global__ void k_dcube(float*, long long, int, double);
__global__ void k_tukey_pad(const float*, int, int, long long, double, float*);
__global__ void k_sumsq(const float*, long long, double*);
__global__ void k_obj_terms(const float*, const float*, long long, int, double*, double*);
__global__ void k_gauss_band(float2*, int, long long, double, double, double, int);
__global__ void k_trim(const float*, int, int, long long, float*);
__global__ void k_lmute(float*, float*, float*, int, long long, const double*, const double*, double,
                        double);
__global__ void k_gaussi_norm(const double*, int, const double*, int, const int*, const double*,
                              double*);
__global__ void k_gaussi_apply(double*, int, int, long long, const double*, const double*,
                               const int*, const double*, const double*);
__global__ void k_img_pad(const double*, int, int, int, int, long long, double*);
__global__ void k_img_kfilter(double2*, int, int, long long, double, double, int);
__global__ void k_img_trim(const double*, int, int, int, int, long long, double*);
__global__ void k_absmax_part(const double*, long long, double*, int*);
__global__ void k_normali(double*, long long, int, const double*, const int*, int,
                          const double*);
__global__ void k_stack_update(const double*, long long, int, const double*, int, double*,
                               double*);
}  // namespace gjw

using namespace gjw;

// --------------------------------------------------------------------------
// errors
// --------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? GJW_ERR_MEM : GJW_ERR_HIP, "%s: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

#define TRY(expr)            \
  do {                       \
    int r_ = (expr);         \
    if (r_ != GJW_OK) return r_; \
  } while (0)

// --------------------------------------------------------------------------
// objects
// --------------------------------------------------------------------------
struct GaussFilter;  // band filter of traces (filt_gauss.m), defined with the FFT code
static void filter_free(GaussFilter* f);
struct PostWork;  // gradient post-processing: 2-D FFT plans and workspace
static void post_free(PostWork* w);

struct gjw_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  GaussFilter* flt = nullptr;  // gjw_filt_gauss's plans and buffers
  PostWork* post = nullptr;    // plans and buffers of the gradient post-processing
  gjw_survey* mws = nullptr;   // gjw_w_bozdag's trace workspace (FFT plans, buffers)
};

template <class T>
struct DBuf {  // device buffer, grows on demand
  T* p = nullptr;
  size_t n = 0;
  int ensure(size_t cnt) {
    if (cnt <= n && p) return GJW_OK;
    if (p) {
      HIPCHK(hipFree(p));
      p = nullptr;
      n = 0;
    }
    if (cnt == 0) cnt = 1;
    HIPCHK(hipMalloc(&p, cnt * sizeof(T)));
    n = cnt;
    return GJW_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// a captured launch sequence and the addresses / parameters it bakes in
struct GraphCache {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  std::vector<uintptr_t> key;
  void reset() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
    key.clear();
  }
};

// step kernels timed by GJW_OPT_KTIMING (numbered as in gjw_kernel_stats)
constexpr int kKernKinds = 10;

struct gjw_survey {
  gjw_ctx* ctx = nullptr;
  gjw_grid gr{};
  int NX = 0, NY = 0, P = 0, LD = 0, di0 = 0, dj0 = 0;
  long long FS = 0;
  bool have_model = false, have_shots = false;
  // options
  int path = GJW_PATH_BATCHED, ck_every = 0, tile_L = 0, snap = GJW_SNAP_F32, fuse = 0;
  int ck_batch = 0;  // checkpoint interval imposed by gjw_batch_fwi_gradient
  int tuned_L = 0, tuned_L1 = 0, tuned_L4 = 0, tuned_L8 = 0;  // 2-, 1-, 4-, 8-step forward
  int tuned_La2 = 0, tuned_La4 = 0, tuned_La6 = 0, tuned_La8 = 0;  // 2-, 4-, 6-, 8-step adjoint
  int adj_fuse = 0;  // adjoint steps per launch (GJW_OPT_ADJ_FUSE; 0: six)
  int adj_tile = 0;  // adjoint tile width (GJW_OPT_ADJ_TILE; 0: tuned in context)
  int frames16 = 0;  // GJW_OPT_FRAMES: BF16 recompute frames asked for
  int fr16 = 0;      // the gradient in progress keeps its frames in BF16 (frames_bf16)
  int fr16_used = 0; // ... and the last one did (GJW_Q_FRAMES)
  int row_pairs = 1; // GJW_OPT_ROW_PAIRS: two-row interior tiles in the 8-step forward
  int pairs_last = 0; // two-row tiles of the last 8-step forward launch (GJW_Q_ROW_PAIRS)
  // adjoint-source band filter of gjw_fwi_gradient (w_update_e.m:69-70); off
  // while flt_hi <= flt_lo
  double flt_lo = 0, flt_hi = 0;
  int flt_steep = 8;
  GaussFilter* flt = nullptr;
  std::vector<hipEvent_t> ex_ev;                // adjoint tile-width trials of the first sweep
  int fuse_eff = 0;  // forward steps per launch in use (fuse, or the autotuner's choice)
  // model, padded (NX x NY, column-major, FP64)
  std::vector<double> pe, ps;
  float bx0 = 0.f, by0 = 0.f;
  DBuf<float> coef;  // cc (2), cxe, cxh, cye, cyh, hax, hbx, hay, hby, uc0 (4), uc1 (4), uc2
  DBuf<float> coef_ref;  // ce, ch with sigma negated (w_grad.m:26-27)
  DBuf<float> zero_sc;   // per-shot zero source coefficients
  DBuf<double> fe, fs;
  std::vector<int> pml_idx;
  DBuf<int> d_pml;
  // shots
  int ns = 0, ntr = 0, rrow_min = 1 << 30, rrow_max = -1;
  std::vector<int> si, sj, tr_shot, tr_row, tr_col;
  std::vector<long long> rec_ptr;
  std::vector<float> sc;
  DBuf<int> d_si, d_sj, d_cptr, d_crow, d_cid, d_trs, d_trr, d_trc, d_rp;
  int rec_multi = 0;  // some cell holds several receivers of one shot
  DBuf<float> d_sc, d_wav;
  // state
  DBuf<float> fld;   // forward: E0,Hx0,Hy0,E1,Hx1,Hy1,X,Y  (8 * ns * FS)
  DBuf<float> adjf;  // adjoint: M0,Nx0,Ny0,X0,Y0,M1,Nx1,Ny1,X1,Y1 (10 * ns * FS)
  DBuf<float> tr, adj, obs;
  DBuf<float> jzb;   // source recorder [shot][nt] (gjw_forward with jz)
  bool want_jz = false;
  DBuf<float> ck;    // checkpoints (FP32, or 16-bit planes + shifts; see ck_stride)
  DBuf<float> frames;
  DBuf<float> acc;
  DBuf<double> ge, gsg, part, Jd, gw;
  int ck_K = 0, ck_n = 0, kept = 0;  // kept: 1 batched checkpoints, 2 plain wavefield
  int cur = 0;                       // which forward buffer holds the current state
  bool have_adj = false;
  double last_ms = 0, last_cells = 0;
  // sampled per-kernel timing (GJW_OPT_KTIMING): event pairs around every
  // ksample-th launch of k_fwd_step / k_adj_step, resolved after each call
  int ksample = 0;
  // Hilbert-based misfits: FFT plans over all traces and their buffers
  int misfit_kind = GJW_MISFIT_L2;  // misfit of gjw_fwi_gradient (GJW_OPT_MISFIT)
  int shot_grads = 0;               // 1: one gradient per shot (GJW_OPT_SHOT_GRADS)
  hipfftHandle fft_r2c = 0, fft_c2c = 0;
  int fft_nt = 0, fft_ntr = 0;
  DBuf<float2> fhalf, fana;
  DBuf<float> fin, fov, fP;
  // GJW_OPT_GRAPH: the captured batched FWI gradient and the buffers it uses
  int graph_on = 0, graph_calls = 0;
  GraphCache g_all, g_fwd, g_adj;  // whole gradient (L2); forward / adjoint sweeps
  std::vector<hipEvent_t> kev_pool;
  std::vector<std::pair<int, int>> kev_used;  // (kernel id, index of start event)
  // k_fwd_step, k_adj_step, k_fwd_step2, k_adj_step2, k_fwd_step4, k_adj_step4, k_fwd_step8
  // (forward sweep), k_adj_step8, k_adj_step6, k_fwd_step8 with frames (recompute sweep)
  double kt_ms[kKernKinds] = {};
  long long kt_n[kKernKinds] = {};
  long long kl_n[kKernKinds] = {};  // all launches while timing is on
};

static inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

static int host_or_dev_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return GJW_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st));
  return GJW_OK;
}

// --------------------------------------------------------------------------
// context
// --------------------------------------------------------------------------
extern "C" const char* gjw_last_error(void) { return g_err.c_str(); }
extern "C" const char* gjw_version(void) { return "gjw 0.1 (gfx950)"; }

// Device scratch of gjw_normali_noalloc: the per-block partial maxima and NaN
// counts of k_absmax_part, kNoAllocParts of each (768 KiB), in a file-scope
// device array so that the entry point allocates nothing.  Each GPU holds its
// own copy until the process ends; host code uses the address that
// hipGetSymbolAddress returns on that device (noalloc_scratch), never the
// array's name.  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU (several host threads or streams) race.
constexpr int kNoAllocParts = 65536;
static __device__ __align__(256) unsigned char scratch[(size_t)kNoAllocParts * (8 + 4)];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::once_flag g_scratch_once[kMaxDevices];
static int g_scratch_rc[kMaxDevices] = {};

// Resolve this device's copy of the scratch array, once per device, with the
// device current (gjw_create calls it).
static int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(GJW_ERR_ARG, "device %d beyond the scratch table (%d devices)", device, kMaxDevices);
  std::call_once(g_scratch_once[device], [device] {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) {
      g_scratch_ptr[device] = p;
    } else {
      (void)hipGetLastError();
      g_scratch_rc[device] = (int)e;
    }
  });
  if (!g_scratch_ptr[device])
    return fail(GJW_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                hipGetErrorString((hipError_t)g_scratch_rc[device]));
  return GJW_OK;
}

extern "C" int gjw_create(int32_t device, gjw_ctx** out) {
  if (!out) return fail(GJW_ERR_ARG, "gjw_create: out is NULL");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(GJW_ERR_ARG, "gjw_create: device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  TRY(noalloc_scratch_init(device));
  gjw_ctx* c = new gjw_ctx;
  c->device = device;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  *out = c;
  return GJW_OK;
}

extern "C" int gjw_destroy(gjw_ctx* c) {
  if (!c) return GJW_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  filter_free(c->flt);
  post_free(c->post);
  if (c->mws) gjw_survey_destroy(c->mws);
  delete c;
  return GJW_OK;
}

// --------------------------------------------------------------------------
// survey
// --------------------------------------------------------------------------
extern "C" int gjw_survey_create(gjw_ctx* c, const gjw_grid* g, gjw_survey** out) {
  if (!c || !g || !out) return fail(GJW_ERR_ARG, "gjw_survey_create: NULL argument");
  if (g->m < 1 || g->n < 1 || g->pml < 1 || g->air < 0 || g->nt < 3)
    return fail(GJW_ERR_ARG, "gjw_survey_create: bad sizes m=%d n=%d pml=%d air=%d nt=%d", g->m,
                g->n, g->pml, g->air, g->nt);
  if (!(g->dx > 0) || !(g->dt > 0) || !(g->eps0 > 0) || !(g->mu0 > 0) || !(g->R0 > 0) ||
      !(g->R0 < 1) || g->pml_order < 0)
    return fail(GJW_ERR_ARG, "gjw_survey_create: bad physics (dx, dt, eps0, mu0, R0, order)");
  gjw_survey* s = new gjw_survey;
  s->ctx = c;
  s->gr = *g;
  s->P = g->pml;
  s->NX = g->m + 2 * g->pml + g->air;  // w_geom.m:136
  s->NY = g->n + 2 * g->pml;           // w_geom.m:135
  s->LD = (int)cdiv(s->NX + 1, 64) * 64;
  s->FS = (long long)s->LD * (s->NY + 1);
  // the multi-step kernels address the 19 FS-float coefficient block (and
  // every field) with 32-bit byte offsets below 2^31
  if (19LL * 4 * s->FS >= (1LL << 31)) {
    const long long FS = s->FS;
    delete s;
    return fail(GJW_ERR_ARG,
                "gjw_survey_create: padded grid of %lld cells per field exceeds the %lld-cell limit",
                FS, ((1LL << 31) - 1) / 76);
  }
  s->di0 = s->P + g->air;  // pis + air, 0-based (w_geom.m:168)
  s->dj0 = s->P;           // pjs, 0-based (w_geom.m:169)
  *out = s;
  return GJW_OK;
}

extern "C" int gjw_survey_destroy(gjw_survey* s) {
  if (!s) return GJW_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  for (auto e : s->kev_pool) (void)hipEventDestroy(e);
  for (auto e : s->ex_ev) (void)hipEventDestroy(e);
  filter_free(s->flt);
  if (s->fft_r2c) (void)hipfftDestroy(s->fft_r2c);
  if (s->fft_c2c) (void)hipfftDestroy(s->fft_c2c);
  s->fhalf.release(); s->fana.release(); s->fin.release(); s->fov.release(); s->fP.release();
  s->g_all.reset();
  s->g_fwd.reset();
  s->g_adj.reset();
  s->coef.release(); s->coef_ref.release(); s->zero_sc.release();
  s->fe.release(); s->fs.release(); s->d_pml.release();
  s->d_si.release(); s->d_sj.release(); s->d_cptr.release(); s->d_crow.release();
  s->d_cid.release(); s->d_trs.release(); s->d_trr.release(); s->d_trc.release();
  s->d_rp.release();
  s->d_sc.release(); s->d_wav.release(); s->fld.release(); s->adjf.release();
  s->tr.release(); s->adj.release(); s->obs.release(); s->ck.release(); s->jzb.release();
  s->frames.release(); s->acc.release(); s->ge.release(); s->gsg.release();
  s->part.release(); s->Jd.release(); s->gw.release();
  delete s;
  return GJW_OK;
}

extern "C" int gjw_set_option(gjw_survey* s, int32_t key, int64_t v) {
  if (!s) return fail(GJW_ERR_ARG, "gjw_set_option: NULL survey");
  switch (key) {
    case GJW_OPT_PATH:
      if (v != GJW_PATH_BATCHED && v != GJW_PATH_PLAIN)
        return fail(GJW_ERR_ARG, "GJW_OPT_PATH: bad value %lld", (long long)v);
      s->path = (int)v;
      break;
    case GJW_OPT_CKPT_EVERY:
      if (v < 0 || v > s->gr.nt) return fail(GJW_ERR_ARG, "GJW_OPT_CKPT_EVERY out of range");
      s->ck_every = (int)v;
      break;
    case GJW_OPT_TILE_L:
      if (v < 0 || v > 4096) return fail(GJW_ERR_ARG, "GJW_OPT_TILE_L out of range");
      s->tile_L = (int)v;
      s->tuned_L = s->tuned_L1 = s->tuned_L4 = s->tuned_L8 = 0;
      s->tuned_La2 = s->tuned_La4 = s->tuned_La6 = s->tuned_La8 = 0;
      s->fuse_eff = 0;
      break;
    case GJW_OPT_MISFIT:
      if (v < GJW_MISFIT_L2 || v > GJW_MISFIT_ENV_LOG)
        return fail(GJW_ERR_ARG, "GJW_OPT_MISFIT: unknown misfit %lld", (long long)v);
      s->misfit_kind = (int)v;
      return GJW_OK;
    case GJW_OPT_ADJ_FUSE:
      if (v != 0 && v != 4 && v != 6 && v != 8)
        return fail(GJW_ERR_ARG, "GJW_OPT_ADJ_FUSE must be 0 (six), 4, 6 or 8");
      s->adj_fuse = (int)v;
      return GJW_OK;
    case GJW_OPT_GRAPH:
      if (v != 0 && v != 1) return fail(GJW_ERR_ARG, "GJW_OPT_GRAPH must be 0 or 1");
      s->graph_on = (int)v;
      s->graph_calls = 0;
      return GJW_OK;
    case GJW_OPT_FUSE:
      if (v != 0 && v != 1 && v != 2 && v != 4 && v != 8)
        return fail(GJW_ERR_ARG, "GJW_OPT_FUSE must be 0 (autotuned), 1, 2, 4 or 8");
      s->fuse = (int)v;
      s->tuned_L = 0;
      s->tuned_La2 = s->tuned_La4 = s->tuned_La6 = s->tuned_La8 = 0;
      s->fuse_eff = 0;
      break;
    case GJW_OPT_ADJ_TILE:
      if (v < 0 || v > 4096) return fail(GJW_ERR_ARG, "GJW_OPT_ADJ_TILE out of range");
      s->adj_tile = (int)v;
      s->tuned_La2 = s->tuned_La4 = s->tuned_La6 = s->tuned_La8 = 0;
      s->graph_calls = 0;
      return GJW_OK;
    case GJW_OPT_FRAMES:
      if (v != GJW_FRAMES_F32 && v != GJW_FRAMES_BF16)
        return fail(GJW_ERR_ARG, "GJW_OPT_FRAMES: GJW_FRAMES_F32 or GJW_FRAMES_BF16");
      s->frames16 = (int)v;
      break;
    case GJW_OPT_ROW_PAIRS:
      if (v != 0 && v != 1) return fail(GJW_ERR_ARG, "GJW_OPT_ROW_PAIRS must be 0 or 1");
      s->row_pairs = (int)v;
      s->tuned_L8 = 0;  // the 8-step forward's width is tuned with its tiling
      s->graph_calls = 0;
      break;
    case GJW_OPT_SHOT_GRADS:
      if (v != 0 && v != 1) return fail(GJW_ERR_ARG, "GJW_OPT_SHOT_GRADS must be 0 or 1");
      s->shot_grads = (int)v;
      return GJW_OK;
    case GJW_OPT_KTIMING:
      if (v < 0) return fail(GJW_ERR_ARG, "GJW_OPT_KTIMING must be >= 0");
      s->ksample = (int)v;
      return GJW_OK;
    case GJW_OPT_SNAPSHOT:
      if (v != GJW_SNAP_F32 && v != GJW_SNAP_BF16 && v != GJW_SNAP_F16)
        return fail(GJW_ERR_ARG, "GJW_OPT_SNAPSHOT: F32, BF16 or F16");
      if (v != GJW_SNAP_F32 && s->LD > 2048)
        return fail(GJW_ERR_ARG, "GJW_OPT_SNAPSHOT: 16-bit snapshots need NX+1 <= 2048 rows");
      s->snap = (int)v;
      break;
    default:
      return fail(GJW_ERR_ARG, "gjw_set_option: unknown key %d", key);
  }
  s->kept = 0;
  return GJW_OK;
}

// w_exp2pml.m:17-37 (air rows, edge replication of eps into the PML, zero
// sigma outside the model), then w_cpml.m / w_coeff.m in FP64.
extern "C" int gjw_set_model(gjw_survey* s, const double* eps_r, const double* sigma) {
  if (!s || !eps_r || !sigma) return fail(GJW_ERR_ARG, "gjw_set_model: NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  const int m = s->gr.m, n = s->gr.n, P = s->P, air = s->gr.air, NX = s->NX, NY = s->NY;
  std::vector<double> em((size_t)m * n), sm((size_t)m * n);
  // inputs may live on the device
  HIPCHK(hipMemcpy(em.data(), eps_r, em.size() * sizeof(double), hipMemcpyDefault));
  HIPCHK(hipMemcpy(sm.data(), sigma, sm.size() * sizeof(double), hipMemcpyDefault));
  for (size_t k = 0; k < em.size(); ++k)
    if (!(em[k] > 0) || !std::isfinite(em[k]) || !std::isfinite(sm[k]) || sm[k] < 0)
      return fail(GJW_ERR_ARG, "gjw_set_model: eps must be > 0 and sigma >= 0 (cell %zu)", k);
  s->pe.assign((size_t)NX * NY, 0.0);
  s->ps.assign((size_t)NX * NY, 0.0);
  for (int j = 0; j < NY; ++j) {
    const int jm = std::min(std::max(j - P, 0), n - 1);
    const bool mcol = (j - P >= 0 && j - P < n);
    for (int i = 0; i < NX; ++i) {
      const int r = std::min(std::max(i - P, 0), air + m - 1);  // row of [air; model]
      double e = 1.0, sg = 0.0;
      if (r >= air) {
        e = em[(size_t)(r - air) + (size_t)jm * m];
        sg = sm[(size_t)(r - air) + (size_t)jm * m];
      }
      const bool mrow = (i - P >= 0 && i - P < air + m);
      if (!(mcol && mrow)) sg = 0.0;
      s->pe[(size_t)i + (size_t)j * NX] = e;
      s->ps[(size_t)i + (size_t)j * NX] = sg;
    }
  }
  const double e0 = s->gr.eps0, u0 = s->gr.mu0, dt = s->gr.dt, dx = s->gr.dx, dy = dx;
  const int ord = s->gr.pml_order;
  const double cl = 1.0 / std::sqrt(u0 * e0);
  const double smax = -(ord + 1) * e0 * cl * std::log(s->gr.R0) / (2.0 * dx * P);  // w_cpml.m:47
  const long long FS = s->FS, LD = s->LD;
  std::vector<float> h((size_t)kCoefFloats * FS, 0.f);
  // (Ceze, Cezhy) interleaved per cell in the first 2*FS floats
  float *cc = h.data(), *cxe = cc + 2 * FS, *cxh = cxe + FS, *cye = cxh + FS,
        *cyh = cye + FS, *hax = cyh + FS, *hbx = hax + FS, *hay = hbx + FS, *hby = hay + FS;
  auto EPS = [&](int i, int j) { return s->pe[(size_t)i + (size_t)j * NX]; };
  auto SIG = [&](int i, int j) { return s->ps[(size_t)i + (size_t)j * NX]; };
  // profile of strip cell a (1 = outermost on the n-sides), electric / magnetic
  auto prof_n = [&](int a, double off) { return std::pow(((P - a + 1) - off) / P, ord); };
  auto prof_p = [&](int a, double off) { return std::pow((a - off) / P, ord); };
  s->pml_idx.clear();
  for (int j = 0; j <= NY; ++j)
    for (int i = 0; i <= NX; ++i) {
      const long long k = (long long)j * LD + i;
      const int kd = ekind(NX, NY, P, i, j);
      if (kd == KIND_INNER) {
        const double D = 2.0 * EPS(i, j) * e0 + dt * SIG(i, j);
        cc[2 * k] = (float)((2.0 * EPS(i, j) * e0 - dt * SIG(i, j)) / D);
        cc[2 * k + 1] = (float)((2.0 * dt / dx) / D);
      } else if (kd == KIND_PML) {
        s->pml_idx.push_back((int)k);
        const double ee = EPS(i, j);
        // Ezx (w_cpml.m:60-61, 93-96, 133-134, 174-175)
        if (i <= P) {
          const double se = smax * prof_n(i, 0.75);
          cxe[k] = (float)((2 * e0 - dt * se) / (2 * e0 + dt * se));
          cxh[k] = (float)((2 * dt / dx) / (2 * e0 + dt * se));
        } else if (i >= NX - P) {
          const double se = smax * prof_p(i - NX + P + 1, 0.75);
          cxe[k] = (float)((2 * e0 * ee - dt * se) / (2 * e0 * ee + dt * se));
          cxh[k] = (float)((2 * dt / dx) / (2 * e0 * ee + dt * se));
        } else {
          cxe[k] = 1.f;
          cxh[k] = (float)(dt / (dx * e0 * ee));
        }
        // Ezy (w_cpml.m:63-64, 103-104, 141-144, 182-185)
        if (j <= P) {
          const double se = smax * (dx / dy) * prof_n(j, 0.75);
          cye[k] = (float)((2 * e0 * ee - dt * se) / (2 * e0 * ee + dt * se));
          cyh[k] = (float)(-(2 * dt / dy) / (2 * e0 * ee + dt * se));
        } else if (j >= NY - P) {
          const double se = smax * (dx / dy) * prof_p(j - NY + P + 1, 0.75);
          cye[k] = (float)((2 * e0 * ee - dt * se) / (2 * e0 * ee + dt * se));
          cyh[k] = (float)(-(2 * dt / dy) / (2 * e0 * ee + dt * se));
        } else if (i <= P) {
          cye[k] = 1.f;
          cyh[k] = (float)(-dt / (dy * e0));
        } else {
          cye[k] = 1.f;
          cyh[k] = (float)(-dt / (dy * e0 * ee));
        }
      }
      // Hy strips: rows 0..P-1 (xn, vacuum) and NX-P..NX-1 (xp), columns 1..NY-1
      if (i <= NX - 1 && j >= 1 && j <= NY - 1 && (i <= P - 1 || i >= NX - P)) {
        double sm_;
        if (i <= P - 1)
          sm_ = (u0 / e0) * smax * prof_n(i + 1, 0.25);
        else
          sm_ = (u0 / e0) * smax * prof_p(i - NX + P + 1, 0.25) / EPS(i, j);
        hay[k] = (float)((2 * u0 - dt * sm_) / (2 * u0 + dt * sm_));
        hby[k] = (float)((2 * dt / dx) / (2 * u0 + dt * sm_));
      }
      // Hx strips: columns 0..P-1 (yn) and NY-P..NY-1 (yp), rows 1..NX-1
      if (j <= NY - 1 && i >= 1 && i <= NX - 1 && (j <= P - 1 || j >= NY - P)) {
        double sm_;
        if (j <= P - 1)
          sm_ = (u0 / e0) * smax * (dx / dy) * prof_n(j + 1, 0.25) / EPS(i, j);
        else
          sm_ = (u0 / e0) * smax * (dx / dy) * prof_p(j - NY + P + 1, 0.25) / EPS(i, j);
        hax[k] = (float)((2 * u0 - dt * sm_) / (2 * u0 + dt * sm_));
        hbx[k] = (float)(-(2 * dt / dy) / (2 * u0 + dt * sm_));
      }
    }
  s->bx0 = (float)(-(2 * dt / dy) / (2 * u0));  // Chxez
  s->by0 = (float)((2 * dt / dx) / (2 * u0));   // Chyez
  // Per-cell table of every update coefficient, for the kernels' edge and
  // strip columns: uc0 = (ce, cxe, cxh, cye), uc1 = (cyh, AX, BX, AY), uc2 = BY
  // with  Ez' = ce Ez + (cxe X + cxh dHy) + (cye Y + cyh dHx)  and the split
  // parts stored at PML cells only (inner: ce = Ceze, cxh = -cyh = Cezhy,
  // cxe = cye = 0; ring: ce = 1, rest 0), Hx' = AX Hx + BX dEz/dj,
  // Hy' = AY Hy + BY dEz/di ((1, 0) where a component is not updated).
  {
    float *uc0 = h.data() + 10 * FS, *uc1 = uc0 + 4 * FS, *uc2 = uc1 + 4 * FS;
    for (int j = 0; j <= NY; ++j)
      for (int i = 0; i <= NX; ++i) {
        const long long k = (long long)j * LD + i;
        float ce = 1.f, a = 0.f, b = 0.f, c = 0.f, d = 0.f;
        const int kd = ekind(NX, NY, P, i, j);
        if (kd == KIND_INNER) {
          ce = cc[2 * k];
          b = cc[2 * k + 1];
          d = -cc[2 * k + 1];
        } else if (kd == KIND_PML) {
          ce = 0.f;
          a = cxe[k]; b = cxh[k]; c = cye[k]; d = cyh[k];
        }
        float ax = 1.f, bx = 0.f, ay = 1.f, by = 0.f;
        if (j <= NY - 1) {
          if (j >= P && j <= NY - P - 1) {
            bx = s->bx0;
          } else if (i >= 1 && i <= NX - 1) {
            ax = hax[k];
            bx = hbx[k];
          }
        }
        if (i <= NX - 1) {
          if (i >= P && i <= NX - P - 1) {
            by = s->by0;
          } else if (j >= 1 && j <= NY - 1) {
            ay = hay[k];
            by = hby[k];
          }
        }
        uc0[4 * k] = ce; uc0[4 * k + 1] = a; uc0[4 * k + 2] = b; uc0[4 * k + 3] = c;
        uc1[4 * k] = d; uc1[4 * k + 1] = ax; uc1[4 * k + 2] = bx; uc1[4 * k + 3] = ay;
        uc2[k] = by;
      }
  }
  // gradient scale factors on the domain: d/d eps = -2 eps0 / D, d/d sigma = -dt / D
  std::vector<double> hfe((size_t)m * n, 0.0), hfs((size_t)m * n, 0.0);
  for (int b = 0; b < n; ++b)
    for (int a = 0; a < m; ++a) {
      const int i = s->di0 + a, j = s->dj0 + b;
      if (ekind(NX, NY, P, i, j) != KIND_INNER) continue;
      const double D = 2.0 * EPS(i, j) * e0 + dt * SIG(i, j);
      hfe[(size_t)a + (size_t)b * m] = -2.0 * e0 / D;
      hfs[(size_t)a + (size_t)b * m] = -dt / D;
    }
  hipStream_t st = s->ctx->stream;
  TRY(s->coef.ensure(h.size()));
  HIPCHK(hipMemcpyAsync(s->coef.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st));
  TRY(s->fe.ensure(hfe.size()));
  TRY(s->fs.ensure(hfs.size()));
  HIPCHK(hipMemcpyAsync(s->fe.p, hfe.data(), hfe.size() * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->fs.p, hfs.data(), hfs.size() * 8, hipMemcpyHostToDevice, st));
  TRY(s->d_pml.ensure(s->pml_idx.size()));
  HIPCHK(hipMemcpyAsync(s->d_pml.p, s->pml_idx.data(), s->pml_idx.size() * sizeof(int),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  s->have_model = true;
  s->kept = 0;
  if (s->have_shots) {  // source coefficients depend on the model
    for (int q = 0; q < s->ns; ++q) {
      const int i = s->si[q], j = s->sj[q];
      s->sc[q] = ekind(NX, NY, P, i, j) == KIND_INNER
                     ? (float)(-(2.0 * dt) / (2.0 * EPS(i, j) * e0 + dt * SIG(i, j)))
                     : 0.f;
    }
    HIPCHK(hipMemcpy(s->d_sc.p, s->sc.data(), s->ns * sizeof(float), hipMemcpyHostToDevice));
  }
  return GJW_OK;
}

extern "C" int gjw_set_shots(gjw_survey* s, int32_t ns, const int32_t* src, const float* wavelets,
                             const int64_t* rec_ptr, const int32_t* rec_lin) {
  if (!s || !src || !wavelets || !rec_ptr || (ns > 0 && !rec_lin))
    return fail(GJW_ERR_ARG, "gjw_set_shots: NULL argument");
  if (ns < 1) return fail(GJW_ERR_ARG, "gjw_set_shots: ns must be >= 1");
  HIPCHK(hipSetDevice(s->ctx->device));
  const int NX = s->NX, NY = s->NY, P = s->P, nt = s->gr.nt;
  std::vector<int32_t> hs(2 * (size_t)ns);
  std::vector<int64_t> hp(ns + 1);
  HIPCHK(hipMemcpy(hs.data(), src, hs.size() * 4, hipMemcpyDefault));
  HIPCHK(hipMemcpy(hp.data(), rec_ptr, hp.size() * 8, hipMemcpyDefault));
  if (hp[0] != 0) return fail(GJW_ERR_ARG, "gjw_set_shots: rec_ptr[0] must be 0");
  for (int q = 0; q < ns; ++q)
    if (hp[q + 1] < hp[q]) return fail(GJW_ERR_ARG, "gjw_set_shots: rec_ptr not monotone");
  const long long ntr = hp[ns];
  if (ntr > (1LL << 30)) return fail(GJW_ERR_ARG, "gjw_set_shots: too many receivers");
  std::vector<int32_t> hl(std::max<long long>(ntr, 1));
  if (ntr) HIPCHK(hipMemcpy(hl.data(), rec_lin, ntr * 4, hipMemcpyDefault));
  s->ns = ns;
  s->ntr = (int)ntr;
  s->si.resize(ns);
  s->sj.resize(ns);
  s->sc.assign(ns, 0.f);
  for (int q = 0; q < ns; ++q) {
    const int i = hs[2 * q] - 1, j = hs[2 * q + 1] - 1;  // [iz, ix] 1-based
    if (i < 1 || i > NX - 1 || j < 1 || j > NY - 1)
      return fail(GJW_ERR_ARG,
                  "gjw_set_shots: source %d at [%d,%d] must lie in rows 2..%d, cols 2..%d",
                  q + 1, i + 1, j + 1, NX, NY);
    s->si[q] = i;
    s->sj[q] = j;
  }
  s->rec_ptr.assign(hp.begin(), hp.end());
  s->rrow_min = 1 << 30;
  s->rrow_max = -1;
  s->tr_shot.resize(ntr);
  s->tr_row.resize(ntr);
  s->tr_col.resize(ntr);
  const long long NE = (long long)(NX + 1) * (NY + 1);
  for (int q = 0; q < ns; ++q)
    for (long long t = hp[q]; t < hp[q + 1]; ++t) {
      const long long lin = (long long)hl[t] - 1;  // w_M.m:20 sub2ind([nx+1,ny+1],rz,rx)
      if (lin < 0 || lin >= NE)
        return fail(GJW_ERR_ARG, "gjw_set_shots: receiver index %d out of range 1..%lld",
                    hl[t], NE);
      s->tr_shot[t] = q;
      s->tr_row[t] = (int)(lin % (NX + 1));
      s->tr_col[t] = (int)(lin / (NX + 1));
      s->rrow_min = std::min(s->rrow_min, s->tr_row[t]);
      s->rrow_max = std::max(s->rrow_max, s->tr_row[t]);
    }
  // column CSR per shot
  std::vector<int> cptr((size_t)ns * (NY + 2), 0), crow(ntr), cid(ntr);
  {
    std::vector<long long> order(ntr);
    for (long long t = 0; t < ntr; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) {
      if (s->tr_shot[a] != s->tr_shot[b]) return s->tr_shot[a] < s->tr_shot[b];
      return s->tr_col[a] < s->tr_col[b];
    });
    std::vector<int> cnt((size_t)ns * (NY + 1), 0);
    for (long long t = 0; t < ntr; ++t) cnt[(size_t)s->tr_shot[t] * (NY + 1) + s->tr_col[t]]++;
    int run = 0;
    for (int q = 0; q < ns; ++q) {
      for (int j = 0; j <= NY; ++j) {
        cptr[(size_t)q * (NY + 2) + j] = run;
        run += cnt[(size_t)q * (NY + 1) + j];
      }
      cptr[(size_t)q * (NY + 2) + NY + 1] = run;
    }
    for (long long t = 0; t < ntr; ++t) {
      crow[t] = s->tr_row[order[t]];
      cid[t] = (int)order[t];
    }
  }
  hipStream_t st = s->ctx->stream;
  TRY(s->d_si.ensure(ns));
  TRY(s->d_sj.ensure(ns));
  TRY(s->d_sc.ensure(ns));
  TRY(s->d_wav.ensure((size_t)ns * nt));
  TRY(s->d_cptr.ensure(cptr.size()));
  TRY(s->d_crow.ensure(ntr));
  TRY(s->d_cid.ensure(ntr));
  TRY(s->d_trs.ensure(ntr));
  TRY(s->d_trr.ensure(ntr));
  TRY(s->d_trc.ensure(ntr));
  TRY(s->d_rp.ensure(ns + 1));
  {
    std::vector<int> rp(ns + 1);
    for (int q = 0; q <= ns; ++q) rp[q] = (int)s->rec_ptr[q];
    HIPCHK(hipMemcpy(s->d_rp.p, rp.data(), (ns + 1) * 4, hipMemcpyHostToDevice));
    // a cell with several receivers of one shot: the kernels' per-lane
    // receiver lookup keeps one trace per cell, so such surveys use the
    // per-receiver loop
    std::vector<long long> cells(ntr);
    for (long long t = 0; t < ntr; ++t)
      cells[t] = ((long long)s->tr_shot[t] * (s->NY + 1) + s->tr_col[t]) * (s->NX + 1) + s->tr_row[t];
    std::sort(cells.begin(), cells.end());
    s->rec_multi = std::adjacent_find(cells.begin(), cells.end()) != cells.end();
  }
  HIPCHK(hipMemcpy(s->d_si.p, s->si.data(), ns * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->d_sj.p, s->sj.data(), ns * 4, hipMemcpyHostToDevice));
  TRY(host_or_dev_copy(s->d_wav.p, wavelets, (size_t)ns * nt * 4, st));
  HIPCHK(hipMemcpy(s->d_cptr.p, cptr.data(), cptr.size() * 4, hipMemcpyHostToDevice));
  if (ntr) {
    HIPCHK(hipMemcpy(s->d_crow.p, crow.data(), ntr * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->d_cid.p, cid.data(), ntr * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->d_trs.p, s->tr_shot.data(), ntr * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->d_trr.p, s->tr_row.data(), ntr * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->d_trc.p, s->tr_col.data(), ntr * 4, hipMemcpyHostToDevice));
  }
  HIPCHK(hipStreamSynchronize(st));
  s->have_shots = true;
  s->kept = 0;
  s->have_adj = false;
  if (s->have_model) {
    for (int q = 0; q < ns; ++q) {
      const int i = s->si[q], j = s->sj[q];
      const double ee = s->pe[(size_t)i + (size_t)j * NX], sg = s->ps[(size_t)i + (size_t)j * NX];
      // w_coeff.m:43-44; a source inside the strips is overwritten (w_solve.m:92-94)
      s->sc[q] = ekind(NX, NY, P, i, j) == KIND_INNER
                     ? (float)(-(2.0 * s->gr.dt) / (2.0 * ee * s->gr.eps0 + s->gr.dt * sg))
                     : 0.f;
    }
  }
  HIPCHK(hipMemcpy(s->d_sc.p, s->sc.data(), ns * 4, hipMemcpyHostToDevice));
  return GJW_OK;
}

// --------------------------------------------------------------------------
// launch helpers
// --------------------------------------------------------------------------
// FP64 gradient values per parameter: m*n, or m*n per shot (shot-major) with
// GJW_OPT_SHOT_GRADS; the stride is the flush's distance between shots.
static long long grad_len(const gjw_survey* s) {
  return (long long)s->gr.m * s->gr.n * (s->shot_grads ? s->ns : 1);
}
static long long grad_stride(const gjw_survey* s) {
  return s->shot_grads ? (long long)s->gr.m * s->gr.n : 0;
}

static DevGrid dev_grid(const gjw_survey* s) {
  DevGrid g;
  g.NX = s->NX; g.NY = s->NY; g.P = s->P; g.LD = s->LD; g.FS = s->FS;
  g.bx0 = s->bx0; g.by0 = s->by0;
  const float* c = s->coef.p;
  const long long FS = s->FS;
  g.cc = reinterpret_cast<const float2*>(c); g.cxe = c + 2 * FS; g.cxh = c + 3 * FS; g.cye = c + 4 * FS;
  g.cyh = c + 5 * FS; g.hax = c + 6 * FS; g.hbx = c + 7 * FS; g.hay = c + 8 * FS;
  g.hby = c + 9 * FS;
  g.uc0 = reinterpret_cast<const float4*>(c + 10 * FS);
  g.uc1 = reinterpret_cast<const float4*>(c + 14 * FS);
  g.uc2 = c + 18 * FS;
  g.coef = c;
  g.di0 = s->di0; g.dj0 = s->dj0; g.m = s->gr.m; g.n = s->gr.n;
  return g;
}

// shots [q0, q0+cnt) of the survey; jz: record the source cell (forward sweeps
// of gjw_forward that asked for it)
static DevShots dev_shots(const gjw_survey* s, int q0, bool jz = false) {
  DevShots d;
  d.nt = s->gr.nt;
  d.si = s->d_si.p + q0;
  d.sj = s->d_sj.p + q0;
  d.sc = s->d_sc.p + q0;
  d.wav = s->d_wav.p + (long long)q0 * s->gr.nt;
  d.cptr = s->d_cptr.p + (long long)q0 * (s->NY + 2);
  d.crow = s->d_crow.p;
  d.cid = s->d_cid.p;
  d.rmin = s->rrow_min;
  d.rmax = s->rrow_max;
  d.jz = jz ? s->jzb.p + (long long)q0 * s->gr.nt : nullptr;
  d.rp = s->d_rp.p + q0;
  d.rec_multi = s->rec_multi;
  return d;
}

static int launch_fill(float* p, long long n, float v, hipStream_t st) {
  if (n <= 0) return GJW_OK;
  const int blocks = (int)std::min<long long>(cdiv(n, 256), 65536);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, v);
  HIPCHK(hipGetLastError());
  return GJW_OK;
}

static int launch_fill_d(double* p, long long n, double v, hipStream_t st) {
  if (n <= 0) return GJW_OK;
  const int blocks = (int)std::min<long long>(cdiv(n, 256), 65536);
  hipLaunchKernelGGL(k_fill_d, dim3(blocks), dim3(256), 0, st, p, n, v);
  HIPCHK(hipGetLastError());
  return GJW_OK;
}

// 1-D launch; the kernels decode (row tile, column strip, shot) themselves
static inline dim3 step_grid(const gjw_survey* s, int L, int nshots) {
  return dim3((unsigned)(cdiv(s->NX + 1, 64) * cdiv(cdiv(s->NY + 1, L), 4) * nshots));
}

struct FwdBufs {
  float *E[2], *Hx[2], *Hy[2], *X[2], *Y[2];
};

static FwdBufs fwd_bufs(gjw_survey* s, int nsh) {
  FwdBufs b;
  const long long sz = (long long)nsh * s->FS;
  float* p = s->fld.p;
  b.E[0] = p; b.Hx[0] = p + sz; b.Hy[0] = p + 2 * sz;
  b.E[1] = p + 3 * sz; b.Hx[1] = p + 4 * sz; b.Hy[1] = p + 5 * sz;
  b.X[0] = p + 6 * sz; b.Y[0] = p + 7 * sz; b.X[1] = p + 8 * sz; b.Y[1] = p + 9 * sz;
  return b;
}

struct AdjBufs {
  float *M[2], *Nx[2], *Ny[2], *X[2], *Y[2];
};

static AdjBufs adj_bufs(gjw_survey* s, int nsh) {
  AdjBufs b;
  const long long sz = (long long)nsh * s->FS;
  float* p = s->adjf.p;
  for (int t = 0; t < 2; ++t) {
    b.M[t] = p + (5 * t + 0) * sz;
    b.Nx[t] = p + (5 * t + 1) * sz;
    b.Ny[t] = p + (5 * t + 2) * sz;
    b.X[t] = p + (5 * t + 3) * sz;
    b.Y[t] = p + (5 * t + 4) * sz;
  }
  return b;
}

static int kev_begin(gjw_survey* s, int which, int it, int* slot) {
  *slot = -1;
  if (s->ksample <= 0) return GJW_OK;
  (void)it;
  // time one launch of this kernel in every ksample steps it covers
  const int steps = which == 8 ? 6 : which >= 6 ? 8 : which >= 4 ? 4 : which >= 2 ? 2 : 1;  // 9: 8
  const long long every = std::max(1, s->ksample / steps);
  if (s->kl_n[which]++ % every != 0) return GJW_OK;
  const size_t need = 2 * (s->kev_used.size() + 1);
  while (s->kev_pool.size() < need) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    s->kev_pool.push_back(e);
  }
  *slot = (int)(2 * s->kev_used.size());
  s->kev_used.push_back({which, *slot});
  HIPCHK(hipEventRecord(s->kev_pool[*slot], s->ctx->stream));
  return GJW_OK;
}

static int kev_end(gjw_survey* s, int slot) {
  if (slot >= 0) HIPCHK(hipEventRecord(s->kev_pool[slot + 1], s->ctx->stream));
  return GJW_OK;
}

// fold the sampled event pairs of the last call into the per-kernel averages
static int kev_resolve(gjw_survey* s) {
  if (s->kev_used.empty()) return GJW_OK;
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  for (auto& u : s->kev_used) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, s->kev_pool[u.second], s->kev_pool[u.second + 1]));
    s->kt_ms[u.first] += ms;
    s->kt_n[u.first] += 1;
  }
  s->kev_used.clear();
  return GJW_OK;
}

static int fwd_launch(gjw_survey* s, const DevGrid& g, const DevShots& sh, int nsh, int L,
                      FwdBufs& b, int cur, int it, float* tr, float* frame, long long fstride,
                      int xc = 0) {
  int slot;
  TRY(kev_begin(s, 0, it, &slot));
  hipLaunchKernelGGL(k_fwd_step, step_grid(s, L, nsh), dim3(64, 4), 0, s->ctx->stream, g, sh, L,
                     b.E[cur], b.Hx[cur], b.Hy[cur], b.E[cur ^ 1], b.Hx[cur ^ 1], b.Hy[cur ^ 1],
                     b.X[xc], b.Y[xc], it, tr, frame, fstride);
  HIPCHK(hipGetLastError());
  return kev_end(s, slot);
}

// NS steps per launch; grid of (64 - 2 NS)-row overlapped tiles
// (a block stacks 4 row tiles of one strip)
// wpb: waves per block (4, or 2: each block-row's stack of four row tiles
// as two blocks, see vtile_of_block)
static inline dim3 stepN_grid(const gjw_survey* s, int NS, int L, int nshots, int wpb = 4) {
  return dim3((unsigned)(cdiv(cdiv(s->NX + 1, 64 - 2 * NS), 4) * cdiv(s->NY + 1, L) * nshots *
                         (4 / wpb)));
}

// Two-row interior tiles of the 8-step forward (fwd8_pairs) for strips of L
// columns.  One-row tiles stay where a wave's rows reach a PML row, a
// receiver or a source (the top and bottom tiles of row_tile's layout); the
// rows between go to waves of two rows per lane, in the strips whose march
// runs through interior columns only.  n = 0 (every tile one row per lane)
// when no such rows or strips exist.
static RowPairs row_pairs_plan(const gjw_survey* s, int L) {
  RowPairs rp{};
  if (!s->row_pairs) return rp;
  constexpr int NS = 8, R = 64 - 2 * NS, R2 = 128 - 2 * NS;
  const int NX = s->NX, NY = s->NY, P = s->P;
  const int T = (int)cdiv(NX + 1, R);
  if (T < 3) return rp;
  auto first = [&](int t) { return t == T - 1 ? std::max(NX + 1 - R, 0) : t * R; };
  int smax = -1;
  for (int q = 0; q < s->ns; ++q) smax = std::max(smax, s->si[q]);
  // top: the first tile whose wave rows (from its first owned row - NS) are
  // inner rows below every receiver and source row
  int t0 = 1;
  while (t0 < T && !(first(t0) - NS >= P && first(t0) > s->rrow_max && first(t0) - NS > smax))
    ++t0;
  // bottom: the one-row tiles tb .. T-1 hold the rows from where a pair
  // wave's needed rows (up to its last owned row + NS - 1) would reach the PML
  int tb = T - 1;
  while (tb > t0 && first(tb) - 1 + NS - 1 > NX - P - 1) --tb;
  if (tb <= t0) return rp;
  const int lo = first(t0), hi = first(tb) - 1;
  const int n = (int)cdiv(hi - lo + 1, R2);
  // frames are stored for rows of the stored domain only
  if (hi < lo || t0 + n > tb || lo < s->di0 || hi > s->di0 + s->gr.m - 1) return rp;
  // strips: level 0 marches columns jb-NS+1 .. je+NS-2 and level NS-1 starts
  // NS-1 columns behind; all of them interior
  int slo = 0;
  while (slo * L - 2 * (NS - 1) < P + 1) ++slo;
  int shi = (int)cdiv(NY + 1, L) - 1;
  while (shi >= slo && std::min(shi * L + L, NY + 1) + NS - 2 > NY - P - 1) --shi;
  if (shi < slo) return rp;
  rp.t0 = t0;
  rp.n = n;
  rp.tb = tb;
  rp.lo = lo;
  rp.hi = hi;
  rp.slo = slo;
  rp.shi = shi;
  return rp;
}

// NS = 2 or 4 fused forward steps; fr1 (optional) receives the NS domain
// frames Ez(S_{it+1}) .. Ez(S_{it+NS}) at fr1 + l * LD * n
static int fwdN_launch(gjw_survey* s, int NS, const DevGrid& g, const DevShots& sh, int nsh, int L,
                       FwdBufs& b, int cur, int xc, int it, float* tr, float* fr1,
                       long long fstride) {
  int slot;
  TRY(kev_begin(s, NS == 8 ? (fr1 ? 9 : 6) : NS == 4 ? 4 : 2, it, &slot));
  // the forward sweep's 8-step kernel runs its stacks as blocks of two waves
  const int wpb = (NS == 8 && !fr1) ? 2 : 4;
  const dim3 grid = stepN_grid(s, NS, L, nsh, wpb), block(64, wpb);
  hipStream_t st = s->ctx->stream;
#define GJW_FWDN_CALL(K)                                                                     \
  hipLaunchKernelGGL(K, grid, block, 0, st, g, sh, L, b.E[cur], b.Hx[cur], b.Hy[cur], b.X[xc], \
                     b.Y[xc], b.E[cur ^ 1], b.Hx[cur ^ 1], b.Hy[cur ^ 1], b.X[xc ^ 1],        \
                     b.Y[xc ^ 1], it, tr, fr1, fstride)
#define GJW_FWD8_CALL(K, RP)                                                                 \
  hipLaunchKernelGGL(K, grid, block, 0, st, g, sh, L, b.E[cur], b.Hx[cur], b.Hy[cur], b.X[xc], \
                     b.Y[xc], b.E[cur ^ 1], b.Hx[cur ^ 1], b.Hy[cur ^ 1], b.X[xc ^ 1],        \
                     b.Y[xc ^ 1], it, tr, fr1, fstride, RP)
  const int FR = fr1 ? (s->fr16 ? 2 : 1) : 0;
  if (NS == 8) {
    // FP32 frames keep one row per lane everywhere
    const RowPairs rp = FR == 1 ? RowPairs{} : row_pairs_plan(s, L);
    s->pairs_last = rp.n;
    if (FR == 2)
      GJW_FWD8_CALL(k_fwd_step8<2>, rp);
    else if (FR == 1)
      GJW_FWD8_CALL(k_fwd_step8<1>, rp);
    else
      GJW_FWD8_CALL(k_fwd_step8<0>, rp);
  } else if (NS == 4) {
    if (FR == 2)
      GJW_FWDN_CALL(k_fwd_step4<2>);
    else if (FR == 1)
      GJW_FWDN_CALL(k_fwd_step4<1>);
    else
      GJW_FWDN_CALL(k_fwd_step4<0>);
  } else {
    if (FR == 2)
      GJW_FWDN_CALL(k_fwd_step2<2>);
    else if (FR == 1)
      GJW_FWDN_CALL(k_fwd_step2<1>);
    else
      GJW_FWDN_CALL(k_fwd_step2<0>);
  }
#undef GJW_FWDN_CALL
#undef GJW_FWD8_CALL
  HIPCHK(hipGetLastError());
  return kev_end(s, slot);
}

// Advance the batched forward state from step it by up to `limit` steps (8,
// 4, 2 or 1 per launch, at most the depth in use); frame (optional) receives
// Ez(S_{it+1}), Ez(S_{it+2}), .. at frame, frame + fstep, ..  Returns the
// steps taken.
static int fwd_advance(gjw_survey* s, const DevGrid& g, const DevShots& sh, int nsh, FwdBufs& b,
                       int& cur, int& xc, int it, int limit, float* tr, float* frame,
                       long long fstep, long long fstride, int* taken) {
  (void)fstep;  // frame slots are LD * n apart, the kernels know it
  const int fz = s->fuse_eff > 0 ? s->fuse_eff : 2;
  const int NS = (fz >= 8 && limit >= 8)   ? 8
                 : (fz >= 4 && limit >= 4) ? 4
                 : (fz >= 2 && limit >= 2) ? 2
                                           : 1;
  if (NS > 1) {
    const int L = NS == 8 ? s->tuned_L8 : NS == 4 ? s->tuned_L4 : s->tuned_L;
    TRY(fwdN_launch(s, NS, g, sh, nsh, L, b, cur, xc, it, tr, frame, fstride));
    xc ^= 1;
  } else {
    if (frame && s->fr16) return fail(GJW_ERR_STATE, "BF16 frames reached a one-step forward launch");
    TRY(fwd_launch(s, g, sh, nsh, s->tuned_L1, b, cur, it, tr, frame, fstride, xc));
  }
  cur ^= 1;
  *taken = NS;
  return GJW_OK;
}

static int adj_launch(gjw_survey* s, const DevGrid& g, const DevShots& sh, int nsh, int L,
                      AdjBufs& b, int cur, int it, const float* adj, const float* F1,
                      const float* F0, long long fstride, float* acc, long long astride,
                      float* gw) {
  int slot;
  TRY(kev_begin(s, 1, it, &slot));
  hipLaunchKernelGGL(k_adj_step, step_grid(s, L, nsh), dim3(64, 4), 0, s->ctx->stream, g, sh, L,
                     b.M[cur], b.Nx[cur], b.Ny[cur], b.X[cur], b.Y[cur], b.M[cur ^ 1],
                     b.Nx[cur ^ 1], b.Ny[cur ^ 1], b.X[cur ^ 1], b.Y[cur ^ 1], it, adj, F1, F0,
                     fstride, acc, astride, gw);
  HIPCHK(hipGetLastError());
  return kev_end(s, slot);
}

// transposes forward steps it, .., it-NS+1 (NS = 2 or 4, it-NS+1 >= 0) in
// one launch; Fa is the frame Ez(S_{it+1}), the earlier ones precede it
static int adjN_launch(gjw_survey* s, int NS, const DevGrid& g, const DevShots& sh, int nsh,
                       int L, AdjBufs& b, int cur, int it, const float* adj, const float* Fa,
                       long long fstride, float* acc, long long astride, float* gw) {
  int slot;
  TRY(kev_begin(s, NS == 8 ? 7 : NS == 6 ? 8 : NS == 4 ? 5 : 3, it, &slot));
  const dim3 grid = stepN_grid(s, NS, L, nsh), block(64, 4);
#define GJW_ADJN_CALL(K)                                                                   \
  hipLaunchKernelGGL(K, grid, block, 0, s->ctx->stream, g, sh, L, b.M[cur], b.Nx[cur],    \
                     b.Ny[cur], b.X[cur], b.Y[cur], b.M[cur ^ 1], b.Nx[cur ^ 1],          \
                     b.Ny[cur ^ 1], b.X[cur ^ 1], b.Y[cur ^ 1], it, adj, Fa, fstride, acc, \
                     astride, gw)
  if (s->fr16) {
    if (NS == 8)
      GJW_ADJN_CALL(k_adj_step8<2>);
    else if (NS == 6)
      GJW_ADJN_CALL(k_adj_step6<2>);
    else if (NS == 4)
      GJW_ADJN_CALL(k_adj_step4<2>);
    else
      GJW_ADJN_CALL(k_adj_step2<2>);
  } else {
    if (NS == 8)
      GJW_ADJN_CALL(k_adj_step8<1>);
    else if (NS == 6)
      GJW_ADJN_CALL(k_adj_step6<1>);
    else if (NS == 4)
      GJW_ADJN_CALL(k_adj_step4<1>);
    else
      GJW_ADJN_CALL(k_adj_step2<1>);
  }
#undef GJW_ADJN_CALL
  HIPCHK(hipGetLastError());
  return kev_end(s, slot);
}

static int check_ready(gjw_survey* s, const char* who) {
  if (!s) return fail(GJW_ERR_ARG, "%s: NULL survey", who);
  if (!s->have_model) return fail(GJW_ERR_STATE, "%s: gjw_set_model has not been called", who);
  if (!s->have_shots) return fail(GJW_ERR_STATE, "%s: gjw_set_shots has not been called", who);
  HIPCHK(hipSetDevice(s->ctx->device));
  return GJW_OK;
}

// Tile width L (columns per wavefront) of the one-, two-, four- and
// eight-step forward kernels, and the forward fusion depth: time a few
// launches of each candidate on zeroed batch state and keep the fastest;
// with GJW_OPT_FUSE = 0 the depth is the deepest whose launch beats two
// launches of half its depth.
static int autotune_L(gjw_survey* s, int nsh) {
  if (s->tile_L > 0) {
    s->tuned_L = s->tuned_L1 = s->tuned_L4 = s->tuned_L8 = s->tile_L;
    s->fuse_eff = s->fuse > 0 ? s->fuse : 2;
    return GJW_OK;
  }
  if (s->tuned_L > 0 && s->tuned_L1 > 0 && s->tuned_L4 > 0 && s->tuned_L8 > 0 && s->fuse_eff > 0)
    return GJW_OK;
  hipStream_t st = s->ctx->stream;
  DevGrid g = dev_grid(s);
  DevShots sh = dev_shots(s, 0);
  FwdBufs b = fwd_bufs(s, nsh);
  const int ks = s->ksample;
  s->ksample = 0;
  const int reps = 4;
  float best_of[9] = {};
  for (int NS : {1, 2, 4, 8}) {
    const int c1[] = {8, 16, 32, 64}, c2[] = {12, 16, 24, 32}, c4[] = {16, 24, 32, 48},
              c8[] = {48, 64, 96, 128};
    const int* cand = NS == 1 ? c1 : NS == 2 ? c2 : NS == 4 ? c4 : c8;
    const int ncand = NS == 8 ? (int)(sizeof(c8) / sizeof(int)) : 4;
    float best = 1e30f, tms[8] = {};
    int bestL = 16;
    for (int ci = 0; ci < ncand; ++ci) {
      const int L = cand[ci];
      int cur = 0, xc = 0;
      for (int r = -1; r < reps; ++r) {
        if (r == 0) HIPCHK(hipEventRecord(s->ctx->ev0, st));
        if (NS == 1)
          TRY(fwd_launch(s, g, sh, nsh, L, b, cur, 0, nullptr, nullptr, 0, xc));
        else
          TRY(fwdN_launch(s, NS, g, sh, nsh, L, b, cur, xc, 0, nullptr, nullptr, 0));
        cur ^= 1;
        if (NS > 1) xc ^= 1;
      }
      HIPCHK(hipEventRecord(s->ctx->ev1, st));
      HIPCHK(hipEventSynchronize(s->ctx->ev1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1));
      tms[ci] = ms;
      best = std::min(best, ms);
    }
    // the widest candidate within 1.5% of the fastest (candidates ascend):
    // near-equal widths would otherwise flip between runs on timing noise
    for (int ci = 0; ci < ncand; ++ci)
      if (tms[ci] <= 1.015f * best) bestL = cand[ci];
    (NS == 1 ? s->tuned_L1 : NS == 2 ? s->tuned_L : NS == 4 ? s->tuned_L4 : s->tuned_L8) = bestL;
    best_of[NS] = best;
  }
  int fz = 2;
  if (best_of[4] < 2.f * best_of[2]) fz = best_of[8] < 2.f * best_of[4] ? 8 : 4;
  s->fuse_eff = s->fuse > 0 ? s->fuse : fz;
  s->ksample = ks;
  return GJW_OK;
}

static int ensure_fwd_state(gjw_survey* s, int nsh) {
  TRY(s->fld.ensure((size_t)10 * nsh * s->FS));
  return GJW_OK;
}

static int zero_fwd_state(gjw_survey* s, int nsh) {
  return launch_fill(s->fld.p, (long long)10 * nsh * s->FS, 0.f, s->ctx->stream);
}

// Checkpoint layout, one slot per checkpoint:
//   FP32:   [E, Hx, Hy: ns*FS floats each][X, Y: ns*npml floats each]
//   16-bit: the same five planes as BF16/FP16 values, each plane padded to
//           256 bytes, then per plane the power-of-two shifts (int32) of its
//           scale blocks (one grid column of E/Hx/Hy; 256 split-field cells)
static inline long long al256(long long b) { return (b + 255) / 256 * 256; }

static long long ck_plane_cnt(const gjw_survey* s, int p) {  // values per shot
  return p < 3 ? s->FS : (long long)s->pml_idx.size();
}
static int ck_plane_B(const gjw_survey* s, int p) { return p < 3 ? s->LD : 256; }
static long long ck_plane_nb(const gjw_survey* s, int p) {  // scale blocks per shot
  return cdiv(ck_plane_cnt(s, p), ck_plane_B(s, p));
}

// slot size in floats (the buffer is typed float)
static long long ck_stride(const gjw_survey* s) {
  if (s->snap == GJW_SNAP_F32)
    return 3LL * s->ns * s->FS + 2LL * s->ns * (long long)s->pml_idx.size();
  long long bytes = 0;
  for (int p = 0; p < 5; ++p)
    bytes += al256(2LL * s->ns * ck_plane_cnt(s, p)) + al256(4LL * s->ns * ck_plane_nb(s, p));
  return bytes / 4;
}

// 16-bit slot: plane pointers and their shift tables
static void ck16_layout(const gjw_survey* s, int c, unsigned short* pl[5], int* sh[5]) {
  char* base = reinterpret_cast<char*>(s->ck.p + (long long)c * ck_stride(s));
  for (int p = 0; p < 5; ++p) {
    pl[p] = reinterpret_cast<unsigned short*>(base);
    base += al256(2LL * s->ns * ck_plane_cnt(s, p));
    sh[p] = reinterpret_cast<int*>(base);
    base += al256(4LL * s->ns * ck_plane_nb(s, p));
  }
}

static int ck_save(gjw_survey* s, FwdBufs& b, int cur, int xc, int c) {
  hipStream_t st = s->ctx->stream;
  const long long sz = (long long)s->ns * s->FS, npml = (long long)s->pml_idx.size();
  if (s->snap != GJW_SNAP_F32) {
    unsigned short* pl[5];
    int* sh[5];
    ck16_layout(s, c, pl, sh);
    const float* src[5] = {b.E[cur], b.Hx[cur], b.Hy[cur], b.X[xc], b.Y[xc]};
    for (int p = 0; p < 5; ++p) {
      const long long cnt = ck_plane_cnt(s, p);
      if (cnt == 0) continue;
      hipLaunchKernelGGL(k_pack16, dim3((unsigned)ck_plane_nb(s, p), (unsigned)s->ns), dim3(256), 0,
                         st, src[p], s->FS, p < 3 ? nullptr : s->d_pml.p, cnt, ck_plane_B(s, p),
                         s->snap, sh[p], pl[p]);
      HIPCHK(hipGetLastError());
    }
    return GJW_OK;
  }
  float* dst = s->ck.p + (long long)c * ck_stride(s);
  HIPCHK(hipMemcpyAsync(dst, b.E[cur], sz * 4, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(dst + sz, b.Hx[cur], sz * 4, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(dst + 2 * sz, b.Hy[cur], sz * 4, hipMemcpyDeviceToDevice, st));
  const long long nn = npml * s->ns;
  if (nn > 0) {
    hipLaunchKernelGGL(k_gather, dim3((unsigned)cdiv(nn, 256)), dim3(256), 0, st, b.X[xc], s->FS,
                       s->d_pml.p, (int)npml, s->ns, dst + 3 * sz);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)cdiv(nn, 256)), dim3(256), 0, st, b.Y[xc], s->FS,
                       s->d_pml.p, (int)npml, s->ns, dst + 3 * sz + nn);
    HIPCHK(hipGetLastError());
  }
  return GJW_OK;
}

static int ck_restore(gjw_survey* s, FwdBufs& b, int cur, int xc, int c) {
  hipStream_t st = s->ctx->stream;
  const long long sz = (long long)s->ns * s->FS, npml = (long long)s->pml_idx.size();
  if (s->snap != GJW_SNAP_F32) {
    unsigned short* pl[5];
    int* sh[5];
    ck16_layout(s, c, pl, sh);
    float* dst[5] = {b.E[cur], b.Hx[cur], b.Hy[cur], b.X[xc], b.Y[xc]};
    for (int p = 0; p < 5; ++p) {
      const long long cnt = ck_plane_cnt(s, p);
      if (cnt == 0) continue;
      hipLaunchKernelGGL(k_unpack16, dim3((unsigned)ck_plane_nb(s, p), (unsigned)s->ns), dim3(256),
                         0, st, pl[p], cnt, ck_plane_B(s, p), sh[p], s->snap,
                         p < 3 ? nullptr : s->d_pml.p, s->FS, dst[p]);
      HIPCHK(hipGetLastError());
    }
    return GJW_OK;
  }
  const float* src = s->ck.p + (long long)c * ck_stride(s);
  HIPCHK(hipMemcpyAsync(b.E[cur], src, sz * 4, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(b.Hx[cur], src + sz, sz * 4, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(b.Hy[cur], src + 2 * sz, sz * 4, hipMemcpyDeviceToDevice, st));
  const long long nn = npml * s->ns;
  if (nn > 0) {
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)cdiv(nn, 256)), dim3(256), 0, st,
                       src + 3 * sz, s->FS, s->d_pml.p, (int)npml, s->ns, b.X[xc]);
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)cdiv(nn, 256)), dim3(256), 0, st,
                       src + 3 * sz + nn, s->FS, s->d_pml.p, (int)npml, s->ns, b.Y[xc]);
    HIPCHK(hipGetLastError());
  }
  return GJW_OK;
}

// adjoint steps per launch: six by default (GJW_OPT_ADJ_FUSE: 4, 6, or 8 when
// the forward fuses at least as many), at most the forward's depth when it
// is shallower than four
static int adj_depth(const gjw_survey* s) {
  if (s->fuse_eff > 0 && s->fuse_eff < 4) return s->fuse_eff;
  if (s->adj_fuse == 8) return s->fuse_eff >= 8 ? 8 : 6;
  return s->adj_fuse == 4 ? 4 : 6;
}

// Depths of the shared launches of a survey batch.  A survey's own autotuner
// judges a depth by the survey alone, which a small survey cannot fill; the
// batch fills the GPU together, so it takes the deepest kernels (8-step
// forward, 6-step adjoint) unless an option asks a survey for less.
static int batch_fwd_depth(const gjw_survey* s) { return s->fuse > 0 ? s->fuse : 8; }
static int batch_adj_depth(const gjw_survey* s) {
  if (s->fuse > 0 && s->fuse < 4) return s->fuse;
  return s->adj_fuse == 4 ? 4 : 6;
}

static long long ck_stride(const gjw_survey* s);

// bytes per frame element: BF16 frames when asked for with FP32 snapshots
// (the checkpoint intervals the library picks are even; frames_bf16 checks
// the rest)
static int frame_esize(const gjw_survey* s) {
  return s->frames16 && s->snap == GJW_SNAP_F32 ? 2 : 4;
}
// BF16 frames for a gradient at interval K: every segment then has an even
// number of steps, which the fused kernels cover without one-step launches
// (the one-step kernels keep FP32 frames)
static bool frames_bf16(const gjw_survey* s, int K) {
  return frame_esize(s) == 2 && K % 2 == 0 && s->gr.nt % 2 == 0 && s->fuse_eff >= 2;
}

static int ck_interval(const gjw_survey* s) {
  if (s->ck_batch > 0) return s->ck_batch;  // common interval of a survey batch
  if (s->ck_every > 0) return s->ck_every;
  // about 2 sqrt(nt) steps, rounded up to whole launches of the forward and
  // the adjoint depths.  A checkpoint holds about three frames' worth of
  // state, so frames + checkpoints take least memory near 1.8 sqrt(nt); the
  // longer segments also halve the per-segment work (restore, first frame,
  // flush): 1.2% faster than sqrt(nt) at nt = 8000.
  const int nt = s->gr.nt;
  const int qf = s->fuse_eff >= 8 ? 8 : 4, qa = adj_depth(s);
  const int q = qf * qa / std::gcd(qf, qa);
  const int k = std::max(1, (int)std::ceil(2.0 * std::sqrt((double)nt)));
  int K = std::min(nt, (k + q - 1) / q * q);
  // A survey whose checkpoints and frames would pass half of the device's
  // memory at that interval takes the interval that needs the least memory:
  // ceil(nt/K) checkpoints of c bytes plus K+1 frames of f bytes.
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return K;
  }
  const double c = 4.0 * (double)ck_stride(s),
               f = (double)frame_esize(s) * s->ns * (double)s->LD * s->gr.n;
  auto bytes = [&](int kk) { return std::ceil((double)nt / kk) * c + (kk + 1) * f; };
  if (bytes(K) <= 0.5 * (double)tot) return K;
  const int k0 = std::max(1, (int)std::lround(std::sqrt(nt * c / f)));
  int best = K;
  for (int kk = std::max(q, k0 / q * q); kk <= std::min(nt, (k0 / q + 1) * q); kk += q)
    if (bytes(kk) < bytes(best)) best = kk;
  return best;
}

// --------------------------------------------------------------------------
// forward
// --------------------------------------------------------------------------
static int forward_batched(gjw_survey* s, int keep) {
  const int ns = s->ns, nt = s->gr.nt;
  hipStream_t st = s->ctx->stream;
  TRY(ensure_fwd_state(s, ns));
  TRY(s->tr.ensure(std::max<long long>((long long)s->ntr * nt, 1)));
  TRY(zero_fwd_state(s, ns));
  TRY(autotune_L(s, ns));
  TRY(zero_fwd_state(s, ns));
  if (keep) {
    s->ck_K = ck_interval(s);
    s->ck_n = (int)cdiv(nt, s->ck_K);
    TRY(s->ck.ensure((size_t)s->ck_n * ck_stride(s)));
  }
  DevGrid g = dev_grid(s);
  DevShots sh = dev_shots(s, 0, s->want_jz);
  FwdBufs b = fwd_bufs(s, ns);
  int cur = 0, xc = 0;
  HIPCHK(hipEventRecord(s->ctx->ev0, st));
  for (int it = 0; it < nt;) {
    if (keep && it % s->ck_K == 0) TRY(ck_save(s, b, cur, xc, it / s->ck_K));
    // never step across a checkpoint boundary
    int limit = nt - it;
    if (keep) limit = std::min(limit, s->ck_K - it % s->ck_K);
    int taken = 0;
    TRY(fwd_advance(s, g, sh, ns, b, cur, xc, it, limit, s->tr.p, nullptr, 0, 0, &taken));
    it += taken;
  }
  HIPCHK(hipEventRecord(s->ctx->ev1, st));
  s->cur = cur;
  s->kept = keep ? 1 : 0;
  s->last_cells = (double)(s->NX + 1) * (s->NY + 1) * nt * ns;
  return GJW_OK;
}

// plain: one shot at a time, whole domain wavefield kept (w_solve.m:76,102).
// frames (nt+1 domain frames of the shot) may be null when nothing is kept.
static int forward_plain_shot(gjw_survey* s, int q, float* fr) {
  const int nt = s->gr.nt;
  const long long FD = (long long)s->LD * s->gr.n;  // one domain frame, field pitch
  const int L = s->tile_L > 0 ? s->tile_L : 16;
  DevGrid g = dev_grid(s);
  FwdBufs b = fwd_bufs(s, 1);
  DevShots sh = dev_shots(s, q, s->want_jz);
  TRY(zero_fwd_state(s, 1));
  if (fr) TRY(launch_fill(fr, FD, 0.f, s->ctx->stream));  // frame 0 = Ez(S_0) = 0
  int cur = 0;
  for (int it = 0; it < nt; ++it) {
    TRY(fwd_launch(s, g, sh, 1, L, b, cur, it, s->tr.p, fr ? fr + (long long)(it + 1) * FD : nullptr,
                   0));
    cur ^= 1;
  }
  return GJW_OK;
}

static int forward_plain(gjw_survey* s, int keep) {
  const int ns = s->ns, nt = s->gr.nt;
  const long long FD = (long long)s->LD * s->gr.n;
  TRY(ensure_fwd_state(s, 1));
  TRY(s->tr.ensure(std::max<long long>((long long)s->ntr * nt, 1)));
  if (keep) TRY(s->frames.ensure((size_t)ns * (nt + 1) * FD));
  HIPCHK(hipEventRecord(s->ctx->ev0, s->ctx->stream));
  for (int q = 0; q < ns; ++q)
    TRY(forward_plain_shot(s, q, keep ? s->frames.p + (long long)q * (nt + 1) * FD : nullptr));
  HIPCHK(hipEventRecord(s->ctx->ev1, s->ctx->stream));
  s->kept = keep ? 2 : 0;
  s->last_cells = (double)(s->NX + 1) * (s->NY + 1) * nt * ns;
  return GJW_OK;
}

static int finish_timing(gjw_survey* s) {
  TRY(kev_resolve(s));
  HIPCHK(hipEventSynchronize(s->ctx->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1));
  s->last_ms = ms;
  return GJW_OK;
}

extern "C" int gjw_forward(gjw_survey* s, float* traces, float* jz, int32_t keep) {
  TRY(check_ready(s, "gjw_forward"));
  s->have_adj = false;
  s->want_jz = jz != nullptr;
  if (jz) TRY(s->jzb.ensure((size_t)s->ns * s->gr.nt));
  const int rc = s->path == GJW_PATH_PLAIN ? forward_plain(s, keep) : forward_batched(s, keep);
  s->want_jz = false;
  TRY(rc);
  TRY(finish_timing(s));
  if (traces) {
    HIPCHK(hipMemcpyAsync(traces, s->tr.p, (size_t)s->ntr * s->gr.nt * 4, hipMemcpyDefault,
                          s->ctx->stream));
  }
  if (jz)
    HIPCHK(hipMemcpyAsync(jz, s->jzb.p, (size_t)s->ns * s->gr.nt * 4, hipMemcpyDefault,
                          s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return GJW_OK;
}

// --------------------------------------------------------------------------
// misfit
// --------------------------------------------------------------------------
#define FFTCHK(x)                                                             \
  do {                                                                        \
    hipfftResult r_ = (x);                                                    \
    if (r_ != HIPFFT_SUCCESS) return fail(GJW_ERR_HIP, "hipfft %d at %s:%d", (int)r_, __FILE__, __LINE__); \
  } while (0)

// --------------------------------------------------------------------------
// band filter of traces (filt_gauss.m:16-47): Tukey taper (r = 0.1), zero
// padding to the next power of two N, one-sided super-Gaussian band in the
// frequency domain, real part of the inverse transform, trim to nt.  Traces
// are filtered in place, trace-major.
// --------------------------------------------------------------------------
struct GaussFilter {
  hipfftHandle r2c = 0, c2r = 0;
  int N = 0, ntr = 0;
  DBuf<float> pad;
  DBuf<float2> spec;
  void release() {
    if (r2c) (void)hipfftDestroy(r2c);
    if (c2r) (void)hipfftDestroy(c2r);
    r2c = c2r = 0;
    N = ntr = 0;
    pad.release();
    spec.release();
  }
  int run(hipStream_t st, float* d, int ntr_, int nt, double dt, double f_low, double f_high,
          int steep) {
    if (ntr_ <= 0) return GJW_OK;
    int n2 = 1;
    while (n2 < nt) n2 *= 2;  // nextpow2
    if (!r2c || N != n2 || ntr != ntr_) {
      if (r2c) FFTCHK(hipfftDestroy(r2c));
      if (c2r) FFTCHK(hipfftDestroy(c2r));
      r2c = c2r = 0;
      int n1 = n2;
      FFTCHK(hipfftPlanMany(&r2c, 1, &n1, nullptr, 1, n2, nullptr, 1, n2 / 2 + 1, HIPFFT_R2C,
                            ntr_));
      FFTCHK(hipfftPlanMany(&c2r, 1, &n1, nullptr, 1, n2 / 2 + 1, nullptr, 1, n2, HIPFFT_C2R,
                            ntr_));
      N = n2;
      ntr = ntr_;
    }
    TRY(pad.ensure((size_t)ntr_ * N));
    TRY(spec.ensure((size_t)ntr_ * (N / 2 + 1)));
    FFTCHK(hipfftSetStream(r2c, st));
    FFTCHK(hipfftSetStream(c2r, st));
    const long long np = (long long)ntr_ * N;
    const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(np, 256), 1), 8192);
    hipLaunchKernelGGL(k_tukey_pad, dim3(blocks), dim3(256), 0, st, d, nt, N, (long long)ntr_, 0.1,
                       pad.p);
    HIPCHK(hipGetLastError());
    FFTCHK(hipfftExecR2C(r2c, reinterpret_cast<hipfftReal*>(pad.p),
                         reinterpret_cast<hipfftComplex*>(spec.p)));
    const double df = 1.0 / dt / N, width = f_high - f_low, mid = f_low + 0.5 * width;
    hipLaunchKernelGGL(k_gauss_band, dim3(blocks), dim3(256), 0, st, spec.p, N, (long long)ntr_,
                       df, mid, width, steep);
    HIPCHK(hipGetLastError());
    FFTCHK(hipfftExecC2R(c2r, reinterpret_cast<hipfftComplex*>(spec.p),
                         reinterpret_cast<hipfftReal*>(pad.p)));
    hipLaunchKernelGGL(k_trim, dim3(blocks), dim3(256), 0, st, pad.p, N, nt, (long long)ntr_, d);
    HIPCHK(hipGetLastError());
    return GJW_OK;
  }
};

static void filter_free(GaussFilter* f) {
  if (!f) return;
  f->release();
  delete f;
}

// --------------------------------------------------------------------------
// k-space filter of images (image_gaussian.m:21-61): zero padding to powers
// of two M x N, batched 2-D D2Z, the Gaussian cake of lines 37-50 on the
// one-sided spectrum, Z2D, trim.  FP64 throughout, as MATLAB's fft2.
// --------------------------------------------------------------------------
struct PostWork {
  hipfftHandle d2z = 0, z2d = 0;
  int M = 0, N = 0, count = 0;
  DBuf<double> pad, dsmall;
  DBuf<double2> spec;
  DBuf<int> ismall;
  void release() {
    if (d2z) (void)hipfftDestroy(d2z);
    if (z2d) (void)hipfftDestroy(z2d);
    d2z = z2d = 0;
    M = N = count = 0;
    pad.release();
    spec.release();
    dsmall.release();
    ismall.release();
  }
  int kfilter(hipStream_t st, const double* in, double* out, int m, int n, int cnt, double ax,
              double az, int type) {
    int M2 = 1, N2 = 1;
    while (M2 < m) M2 *= 2;  // 2^nextpow2
    while (N2 < n) N2 *= 2;
    if (!d2z || M != M2 || N != N2 || count != cnt) {
      if (d2z) FFTCHK(hipfftDestroy(d2z));
      if (z2d) FFTCHK(hipfftDestroy(z2d));
      d2z = z2d = 0;
      int dims[2] = {N2, M2};  // slowest first: columns, then depth
      const int nr = M2 * N2, nc = N2 * (M2 / 2 + 1);
      FFTCHK(hipfftPlanMany(&d2z, 2, dims, nullptr, 1, nr, nullptr, 1, nc, HIPFFT_D2Z, cnt));
      FFTCHK(hipfftPlanMany(&z2d, 2, dims, nullptr, 1, nc, nullptr, 1, nr, HIPFFT_Z2D, cnt));
      M = M2;
      N = N2;
      count = cnt;
    }
    TRY(pad.ensure((size_t)cnt * M * N));
    TRY(spec.ensure((size_t)cnt * N * (M / 2 + 1)));
    FFTCHK(hipfftSetStream(d2z, st));
    FFTCHK(hipfftSetStream(z2d, st));
    auto blocks = [](long long tot) {
      return dim3((unsigned)std::min<long long>(std::max<long long>(cdiv(tot, 256), 1), 8192));
    };
    hipLaunchKernelGGL(k_img_pad, blocks((long long)cnt * M * N), dim3(256), 0, st, in, m, n, M, N,
                       (long long)cnt, pad.p);
    HIPCHK(hipGetLastError());
    FFTCHK(hipfftExecD2Z(d2z, reinterpret_cast<hipfftDoubleReal*>(pad.p),
                         reinterpret_cast<hipfftDoubleComplex*>(spec.p)));
    // image_gaussian.m:37: a = n_ (ax/2) sqrt(2), b = m_ (az/2) sqrt(2)
    const double a = N * (ax / 2) * std::sqrt(2.0), b = M * (az / 2) * std::sqrt(2.0);
    hipLaunchKernelGGL(k_img_kfilter, blocks((long long)cnt * N * (M / 2 + 1)), dim3(256), 0, st,
                       spec.p, M, N, (long long)cnt, a, b, type);
    HIPCHK(hipGetLastError());
    FFTCHK(hipfftExecZ2D(z2d, reinterpret_cast<hipfftDoubleComplex*>(spec.p),
                         reinterpret_cast<hipfftDoubleReal*>(pad.p)));
    hipLaunchKernelGGL(k_img_trim, blocks((long long)cnt * m * n), dim3(256), 0, st, pad.p, M, N,
                       m, n, (long long)cnt, out);
    HIPCHK(hipGetLastError());
    return GJW_OK;
  }
};

static void post_free(PostWork* w) {
  if (!w) return;
  w->release();
  delete w;
}

static bool filter_on(const gjw_survey* s) { return s->flt_hi > s->flt_lo; }

// the survey's band filter on cnt traces at p (nt samples each)
static int trace_filter(gjw_survey* s, float* p, long long cnt) {
  if (!filter_on(s) || cnt <= 0) return GJW_OK;
  if (!s->flt) s->flt = new GaussFilter();
  return s->flt->run(s->ctx->stream, p, (int)cnt, s->gr.nt, s->gr.dt, s->flt_lo, s->flt_hi,
                     s->flt_steep);
}

// band filter of the adjoint sources of traces [t0, t0+cnt) (w_update_e.m:69-70)
static int adj_filter(gjw_survey* s, long long t0, long long cnt) {
  return trace_filter(s, s->adj.p + t0 * s->gr.nt, cnt);
}

// analytic signal (unnormalised: x nt) of the ntr traces in s->fin, into s->fana
static int hilbert_dev(gjw_survey* s) {
  const int nt = s->gr.nt, ntr = (int)s->ntr;
  hipStream_t st = s->ctx->stream;
  if (!s->fft_r2c || s->fft_nt != nt || s->fft_ntr != ntr) {
    if (s->fft_r2c) FFTCHK(hipfftDestroy(s->fft_r2c));
    if (s->fft_c2c) FFTCHK(hipfftDestroy(s->fft_c2c));
    s->fft_r2c = s->fft_c2c = 0;
    int n1 = nt;
    FFTCHK(hipfftPlanMany(&s->fft_r2c, 1, &n1, nullptr, 1, nt, nullptr, 1, nt / 2 + 1,
                          HIPFFT_R2C, ntr));
    FFTCHK(hipfftPlanMany(&s->fft_c2c, 1, &n1, nullptr, 1, nt, nullptr, 1, nt, HIPFFT_C2C, ntr));
    s->fft_nt = nt;
    s->fft_ntr = ntr;
  }
  FFTCHK(hipfftSetStream(s->fft_r2c, st));
  FFTCHK(hipfftSetStream(s->fft_c2c, st));
  const long long n = (long long)ntr * nt;
  FFTCHK(hipfftExecR2C(s->fft_r2c, reinterpret_cast<hipfftReal*>(s->fin.p),
                       reinterpret_cast<hipfftComplex*>(s->fhalf.p)));
  const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(n, 256), 1), 8192);
  hipLaunchKernelGGL(k_analytic_spec, dim3(blocks), dim3(256), 0, st, s->fhalf.p, s->fana.p, nt,
                     (long long)ntr);
  HIPCHK(hipGetLastError());
  FFTCHK(hipfftExecC2C(s->fft_c2c, reinterpret_cast<hipfftComplex*>(s->fana.p),
                       reinterpret_cast<hipfftComplex*>(s->fana.p), HIPFFT_BACKWARD));
  return GJW_OK;
}

// Adjoint source and misfit of the last forward against s->obs (w_bozdag.m,
// w_E.m): L2 directly; the envelope / phase kinds through three Hilbert
// transforms (observed, synthetic, and the adjoint-source term).
static int misfit_enqueue(gjw_survey* s, int32_t kind) {
  hipStream_t st = s->ctx->stream;
  const int nt = s->gr.nt;
  const long long n = (long long)s->ntr * nt;
  if (kind == GJW_MISFIT_L2) {
    const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(n, 256), 1), 1024);
    hipLaunchKernelGGL(k_resid_l2, dim3(blocks), dim3(256), 0, st, s->tr.p, s->obs.p, s->adj.p, n,
                       s->part.p);
    hipLaunchKernelGGL(k_reduce_d, dim3(1), dim3(256), 0, st, s->part.p, blocks, 0.5, s->Jd.p);
    HIPCHK(hipGetLastError());
    return GJW_OK;
  }
  if (n == 0) return fail(GJW_ERR_STATE, "misfit: no traces");
  TRY(s->fin.ensure(n));
  TRY(s->fov.ensure(n));
  TRY(s->fP.ensure(n));
  TRY(s->fana.ensure(n));
  TRY(s->fhalf.ensure((size_t)s->ntr * (nt / 2 + 1)));
  const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(n, 256), 1), 1024);
  // observed envelope or phase
  HIPCHK(hipMemcpyAsync(s->fin.p, s->obs.p, n * 4, hipMemcpyDeviceToDevice, st));
  TRY(hilbert_dev(s));
  hipLaunchKernelGGL(k_obs_instant, dim3(blocks), dim3(256), 0, st, s->fana.p, n, nt, kind,
                     s->fov.p);
  // synthetic: misfit terms, q and P
  HIPCHK(hipMemcpyAsync(s->fin.p, s->tr.p, n * 4, hipMemcpyDeviceToDevice, st));
  TRY(hilbert_dev(s));
  hipLaunchKernelGGL(k_misfit_stage1, dim3(blocks), dim3(256), 0, st, s->tr.p, s->fana.p,
                     s->fov.p, n, nt, kind, s->fin.p, s->fP.p, s->part.p);
  hipLaunchKernelGGL(k_reduce_d, dim3(1), dim3(256), 0, st, s->part.p, blocks, 1.0, s->Jd.p);
  // e = P -+ imag(hilbert(q))
  TRY(hilbert_dev(s));
  hipLaunchKernelGGL(k_misfit_stage2, dim3(blocks), dim3(256), 0, st, s->fP.p, s->fana.p, n, nt,
                     kind, s->adj.p);
  HIPCHK(hipGetLastError());
  return GJW_OK;
}

static int misfit_dev(gjw_survey* s, const float* d_obs, int32_t kind) {
  if (kind < GJW_MISFIT_L2 || kind > GJW_MISFIT_ENV_LOG)
    return fail(GJW_ERR_ARG, "gjw_misfit: unknown kind %d", kind);
  hipStream_t st = s->ctx->stream;
  const long long n = (long long)s->ntr * s->gr.nt;
  TRY(s->adj.ensure(std::max<long long>(n, 1)));
  TRY(s->obs.ensure(std::max<long long>(n, 1)));
  TRY(s->part.ensure(1024));
  TRY(s->Jd.ensure(1));
  TRY(host_or_dev_copy(s->obs.p, d_obs, n * 4, st));
  TRY(misfit_enqueue(s, kind));
  s->have_adj = true;
  return GJW_OK;
}

extern "C" int gjw_bozdag(gjw_survey* s, const float* d_obs, const float* d, int32_t kind,
                          float* e_out, double* J) {
  TRY(check_ready(s, "gjw_bozdag"));
  if (!d_obs || !d) return fail(GJW_ERR_ARG, "gjw_bozdag: NULL traces");
  const long long n = (long long)s->ntr * s->gr.nt;
  TRY(s->tr.ensure(std::max<long long>(n, 1)));
  TRY(host_or_dev_copy(s->tr.p, d, n * 4, s->ctx->stream));
  return gjw_misfit(s, d_obs, kind, e_out, J);
}

extern "C" int gjw_misfit(gjw_survey* s, const float* d_obs, int32_t kind, float* e_out,
                          double* J) {
  TRY(check_ready(s, "gjw_misfit"));
  if (!d_obs) return fail(GJW_ERR_ARG, "gjw_misfit: d_obs is NULL");
  if (!s->tr.p) return fail(GJW_ERR_STATE, "gjw_misfit: run gjw_forward first");
  TRY(misfit_dev(s, d_obs, kind));
  hipStream_t st = s->ctx->stream;
  if (e_out)
    HIPCHK(hipMemcpyAsync(e_out, s->adj.p, (size_t)s->ntr * s->gr.nt * 4, hipMemcpyDefault, st));
  if (J) HIPCHK(hipMemcpyAsync(J, s->Jd.p, 8, hipMemcpyDefault, st));
  HIPCHK(hipStreamSynchronize(st));
  return GJW_OK;
}

// [e, d_e, d_oe] = w_bozdag(d_o, d, kind) and J = w_E for any ntr x nt traces,
// with no survey: the context keeps a trace workspace (plans, buffers) that
// the misfit kernels of the surveys use in the same way.
extern "C" int gjw_w_bozdag(gjw_ctx* c, const float* d_o, const float* d, int32_t ntr,
                            int32_t nt, int32_t kind, float* e, float* d_e, float* d_oe,
                            double* J) {
  if (!c || !d_o || !d) return fail(GJW_ERR_ARG, "gjw_w_bozdag: NULL argument");
  if (ntr < 1 || nt < 3) return fail(GJW_ERR_ARG, "gjw_w_bozdag: bad sizes ntr=%d nt=%d", ntr, nt);
  if (kind < GJW_MISFIT_L2 || kind > GJW_MISFIT_ENV_LOG)
    return fail(GJW_ERR_ARG, "gjw_w_bozdag: unknown kind %d", kind);
  HIPCHK(hipSetDevice(c->device));
  if (!c->mws) {
    c->mws = new gjw_survey;
    c->mws->ctx = c;
  }
  gjw_survey* w = c->mws;
  w->gr.nt = nt;
  w->ntr = ntr;
  hipStream_t st = c->stream;
  const long long n = (long long)ntr * nt;
  TRY(w->tr.ensure(n));
  TRY(w->obs.ensure(n));
  TRY(w->adj.ensure(n));
  TRY(w->part.ensure(1024));
  TRY(w->Jd.ensure(1));
  TRY(host_or_dev_copy(w->obs.p, d_o, n * 4, st));
  TRY(host_or_dev_copy(w->tr.p, d, n * 4, st));
  TRY(misfit_enqueue(w, kind));
  if (e) HIPCHK(hipMemcpyAsync(e, w->adj.p, n * 4, hipMemcpyDefault, st));
  if (J) HIPCHK(hipMemcpyAsync(J, w->Jd.p, 8, hipMemcpyDefault, st));
  if (kind == GJW_MISFIT_L2) {  // w_bozdag.m:14-15: d_e = d, d_oe = d_o
    if (d_e) HIPCHK(hipMemcpyAsync(d_e, w->tr.p, n * 4, hipMemcpyDefault, st));
    if (d_oe) HIPCHK(hipMemcpyAsync(d_oe, w->obs.p, n * 4, hipMemcpyDefault, st));
  } else {
    // the misfit left the envelope (phase for phase_L2) of d_o in fov
    if (d_oe) HIPCHK(hipMemcpyAsync(d_oe, w->fov.p, n * 4, hipMemcpyDefault, st));
    if (d_e) {
      HIPCHK(hipMemcpyAsync(w->fin.p, w->tr.p, n * 4, hipMemcpyDeviceToDevice, st));
      TRY(hilbert_dev(w));
      const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(n, 256), 1), 1024);
      hipLaunchKernelGGL(k_obs_instant, dim3(blocks), dim3(256), 0, st, w->fana.p, n, nt, kind,
                         w->fov.p);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(d_e, w->fov.p, n * 4, hipMemcpyDefault, st));
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  return GJW_OK;
}

// --------------------------------------------------------------------------
// gradient
// --------------------------------------------------------------------------
static int adj_init(gjw_survey* s, int nsh, int q0, const float* adj) {
  hipStream_t st = s->ctx->stream;
  TRY(s->adjf.ensure((size_t)10 * nsh * s->FS));
  TRY(launch_fill(s->adjf.p, (long long)10 * nsh * s->FS, 0.f, st));
  // residual of the last step at the receivers of shots [q0, q0+nsh)
  const long long t0 = s->rec_ptr[q0], t1 = s->rec_ptr[q0 + nsh];
  if (t1 > t0) {
    DevGrid g = dev_grid(s);
    AdjBufs b = adj_bufs(s, nsh);
    hipLaunchKernelGGL(k_inject_last, dim3((unsigned)cdiv(t1 - t0, 256)), dim3(256), 0, st, g,
                       s->gr.nt, q0, (int)(t1 - t0), s->d_trs.p + t0, s->d_trr.p + t0,
                       s->d_trc.p + t0, adj + t0 * s->gr.nt, s->gr.nt - 1, b.M[0]);
    HIPCHK(hipGetLastError());
  }
  return GJW_OK;
}

// candidate tile widths of the deepest adjoint kernel, tried in turn by the
// first gradient sweep (kAdjRounds launches each, the first round a warm-up)
constexpr int kAdjCand4[] = {12, 16, 24, 32, 48}, kAdjCand6[] = {32, 48, 64, 96, 128};
constexpr int kAdjNCand = 5, kAdjRounds = 3;

static int grad_batched_run(gjw_survey* s, const float* adj, float* gw);
static int grad_batched(gjw_survey* s, const float* adj, float* gw) {
  s->fr16 = frames_bf16(s, s->ck_K) ? 1 : 0;
  const int rc = grad_batched_run(s, adj, gw);
  s->fr16_used = s->fr16;
  s->fr16 = 0;
  return rc;
}
static int grad_batched_run(gjw_survey* s, const float* adj, float* gw) {
  const int ns = s->ns, nt = s->gr.nt, K = s->ck_K;
  const long long per = (long long)s->gr.m * s->gr.n, FD = (long long)s->LD * s->gr.n;
  hipStream_t st = s->ctx->stream;
  // FP32 frames: slot q of the segment at q * FD; BF16 pairs: word plane
  // (q + 1) / 2 holds slots q and q + 1 for odd q, plane 0 slot 0 (the
  // kernels' layout).  The frames of one shot are fstride floats apart.
  const bool f16 = s->fr16 != 0;
  const long long planes = f16 ? K / 2 + 1 : K + 1;
  TRY(s->frames.ensure((size_t)ns * planes * FD));
  TRY(s->acc.ensure((size_t)2 * ns * FD));
  TRY(launch_fill(s->acc.p, 2LL * ns * FD, 0.f, st));
  TRY(adj_init(s, ns, 0, adj));
  DevGrid g = dev_grid(s);
  DevShots sh = dev_shots(s, 0);
  FwdBufs fb = fwd_bufs(s, ns);
  AdjBufs ab = adj_bufs(s, ns);
  const long long fstride = planes * FD;  // frames per shot
  int acur = 0;
  const int dmax = adj_depth(s);
  int& tuned_deep = dmax == 8 ? s->tuned_La8 : dmax == 6 ? s->tuned_La6 : s->tuned_La4;
  // Adjoint tile width, tuned in context: the first (uncaptured) sweep gives
  // its first launches of the deepest kernel the candidate widths in turn,
  // timed with events; later sweeps use the fastest.
  if (s->tile_L > 0) s->tuned_La2 = s->tuned_La4 = s->tuned_La6 = s->tuned_La8 = s->tile_L;
  if (s->adj_tile > 0) s->tuned_La2 = s->tuned_La4 = s->tuned_La6 = s->tuned_La8 = s->adj_tile;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(st, &cap));
  const bool explore = tuned_deep <= 0 && cap == hipStreamCaptureStatusNone;
  const int nx = explore ? kAdjRounds * kAdjNCand : 0;
  int ex = 0;  // trial launches issued
  if (explore)
    while ((int)s->ex_ev.size() < 2 * nx) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      s->ex_ev.push_back(e);
    }
  const int La_fallback = s->tuned_L4 > 0 ? s->tuned_L4 : 16;
  for (int c = s->ck_n - 1; c >= 0; --c) {
    const int k0 = c * K, k1 = std::min(k0 + K, nt);
    int cur = 0, xc = 0;
    TRY(ck_restore(s, fb, cur, xc, c));
    hipLaunchKernelGGL(k_domain_frame, dim3((unsigned)cdiv(per * ns, 256)), dim3(256), 0, st, g, ns,
                       fb.E[cur], s->frames.p, fstride, s->fr16);
    HIPCHK(hipGetLastError());
    for (int it = k0; it < k1;) {
      int taken = 0;
      // frames of steps it+1 ..: slot it+1-k0, or (BF16, it - k0 even) the
      // plane of the pair (it+1-k0, it+2-k0)
      float* fr = s->frames.p + (f16 ? (long long)((it - k0) / 2 + 1) : (long long)(it + 1 - k0)) * FD;
      TRY(fwd_advance(s, g, sh, ns, fb, cur, xc, it, k1 - it, nullptr, fr, FD, fstride, &taken));
      it += taken;
    }
    for (int it = k1 - 1; it >= k0;) {
      const float* F = s->frames.p + (long long)(it - k0) * FD;  // Ez(S_it) (FP32 frames)
      if (f16 && (it - k0) % 2 == 0)
        return fail(GJW_ERR_STATE, "BF16 frames: an adjoint pass starts at an even slot");
      // the frame of S_{it+1}: slot it+1-k0, or (BF16) the plane holding it
      const float* Fa = f16 ? s->frames.p + (long long)((it - k0 + 2) / 2) * FD : F + FD;
      const int NS = (dmax >= 6 && it - (dmax - 1) >= k0)       ? dmax
                     : (s->fuse_eff >= 4 && it - 3 >= k0) ? 4
                     : (s->fuse_eff >= 2 && it - 1 >= k0) ? 2
                                                          : 1;
      if (NS > 1) {
        int L = NS == 8   ? s->tuned_La8
                : NS == 6 ? s->tuned_La6
                : NS == 4 ? s->tuned_La4
                          : s->tuned_La2;
        const bool trial = NS == dmax && ex < nx;
        if (trial) L = (dmax >= 6 ? kAdjCand6 : kAdjCand4)[ex % kAdjNCand];
        if (L <= 0) L = La_fallback;
        if (trial) HIPCHK(hipEventRecord(s->ex_ev[2 * ex], st));
        TRY(adjN_launch(s, NS, g, sh, ns, L, ab, acur, it, adj, Fa, fstride, s->acc.p, FD, gw));
        if (trial) HIPCHK(hipEventRecord(s->ex_ev[2 * ex++ + 1], st));
        it -= NS;
      } else {
        if (s->fr16) return fail(GJW_ERR_STATE, "BF16 frames reached a one-step adjoint launch");
        TRY(adj_launch(s, g, sh, ns, s->tuned_L1, ab, acur, it, adj, F + FD, F, fstride,
                       s->acc.p, FD, gw));
        it -= 1;
      }
      acur ^= 1;
    }
    hipLaunchKernelGGL(k_flush, dim3((unsigned)cdiv(per, 256)), dim3(256), 0, st, s->gr.m,
                       s->gr.n, s->LD, s->di0, ns, s->acc.p, FD, s->fe.p, s->fs.p, s->ge.p, s->gsg.p, 1.0,
                       grad_stride(s));
    HIPCHK(hipGetLastError());
  }
  if (explore) {
    // Mean time of each candidate over the trial rounds after the first
    // (warm-up) one; the widest candidate within 1.5% of the fastest wins
    // (near-equal widths would otherwise flip between runs on timing noise;
    // the widest re-reads the fewest strip-edge columns).
    int bestL = La_fallback;
    if (ex >= 2 * kAdjNCand) {
      HIPCHK(hipEventSynchronize(s->ex_ev[2 * ex - 1]));
      float mean[kAdjNCand];
      float best = 1e30f;
      for (int c = 0; c < kAdjNCand; ++c) {
        float tot = 0;
        int n = 0;
        for (int k = kAdjNCand + c; k < ex; k += kAdjNCand) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, s->ex_ev[2 * k], s->ex_ev[2 * k + 1]));
          tot += ms;
          ++n;
        }
        mean[c] = n > 0 ? tot / n : 1e30f;
        best = std::min(best, mean[c]);
      }
      for (int c = 0; c < kAdjNCand; ++c)  // candidates ascend: the last one within wins
        if (mean[c] <= 1.015f * best) bestL = (dmax >= 6 ? kAdjCand6 : kAdjCand4)[c];
    }
    if (dmax >= 6)
      tuned_deep = bestL;
    else
      s->tuned_La2 = s->tuned_La4 = bestL;
  }
  return GJW_OK;
}

static int grad_plain_shot(gjw_survey* s, int q, const float* adj, const float* u, float* gw) {
  const int nt = s->gr.nt;
  const long long per = (long long)s->gr.m * s->gr.n, FD = (long long)s->LD * s->gr.n;
  hipStream_t st = s->ctx->stream;
  const int L = s->tile_L > 0 ? s->tile_L : 16;
  const int K = ck_interval(s);
  DevGrid g = dev_grid(s);
  DevShots sh = dev_shots(s, q);
  TRY(adj_init(s, 1, q, adj));
  AdjBufs ab = adj_bufs(s, 1);
  int acur = 0;
  for (int it = nt - 1; it >= 0; --it) {
    TRY(adj_launch(s, g, sh, 1, L, ab, acur, it, adj, u + (long long)(it + 1) * FD,
                   u + (long long)it * FD, 0, s->acc.p, 0, gw ? gw + (long long)q * nt : nullptr));
    acur ^= 1;
    if (it % K == 0) {
      hipLaunchKernelGGL(k_flush, dim3((unsigned)cdiv(per, 256)), dim3(256), 0, st, s->gr.m,
                         s->gr.n, s->LD, s->di0, 1, s->acc.p, FD, s->fe.p, s->fs.p,
                         s->ge.p + q * grad_stride(s), s->gsg.p + q * grad_stride(s), 1.0, 0LL);
      HIPCHK(hipGetLastError());
    }
  }
  return GJW_OK;
}

static int grad_plain_exact(gjw_survey* s, const float* adj, float* gw) {
  const long long FD = (long long)s->LD * s->gr.n;
  TRY(s->acc.ensure((size_t)2 * FD));
  TRY(launch_fill(s->acc.p, 2LL * FD, 0.f, s->ctx->stream));
  for (int q = 0; q < s->ns; ++q)
    TRY(grad_plain_shot(s, q, adj, s->frames.p + (long long)q * (s->gr.nt + 1) * FD, gw));
  return GJW_OK;
}

// plain FWI gradient: per shot, forward storing the whole domain wavefield,
// L2 residual of that shot, adjoint; J summed over shots at the end.
static int fwi_plain(gjw_survey* s, const float* d_obs) {
  if (s->misfit_kind != GJW_MISFIT_L2)
    return fail(GJW_ERR_ARG, "gjw_fwi_gradient: the plain path computes the L2 misfit only");
  const int ns = s->ns, nt = s->gr.nt;
  const long long n = (long long)s->ntr * nt;
  const long long FD = (long long)s->LD * s->gr.n;
  hipStream_t st = s->ctx->stream;
  const int B = 64;  // residual blocks per shot
  TRY(ensure_fwd_state(s, 1));
  TRY(s->tr.ensure(std::max<long long>(n, 1)));
  TRY(s->adj.ensure(std::max<long long>(n, 1)));
  TRY(s->obs.ensure(std::max<long long>(n, 1)));
  TRY(s->part.ensure((size_t)ns * B));
  TRY(s->Jd.ensure(1));
  TRY(s->frames.ensure((size_t)(nt + 1) * FD));
  TRY(s->acc.ensure((size_t)2 * FD));
  TRY(s->ge.ensure(grad_len(s)));
  TRY(s->gsg.ensure(grad_len(s)));
  TRY(host_or_dev_copy(s->obs.p, d_obs, n * 4, st));
  TRY(launch_fill(s->acc.p, 2LL * FD, 0.f, st));
  TRY(launch_fill_d(s->ge.p, grad_len(s), 0.0, st));
  TRY(launch_fill_d(s->gsg.p, grad_len(s), 0.0, st));
  TRY(launch_fill_d(s->part.p, (long long)ns * B, 0.0, st));
  for (int q = 0; q < ns; ++q) {
    TRY(forward_plain_shot(s, q, s->frames.p));
    const long long t0 = s->rec_ptr[q] * nt, cnt = (s->rec_ptr[q + 1] - s->rec_ptr[q]) * nt;
    if (cnt > 0) {
      hipLaunchKernelGGL(k_resid_l2, dim3(B), dim3(256), 0, st, s->tr.p + t0, s->obs.p + t0,
                         s->adj.p + t0, cnt, s->part.p + (long long)q * B);
      HIPCHK(hipGetLastError());
    }
    TRY(adj_filter(s, s->rec_ptr[q], s->rec_ptr[q + 1] - s->rec_ptr[q]));
    TRY(grad_plain_shot(s, q, s->adj.p, s->frames.p, nullptr));
  }
  hipLaunchKernelGGL(k_reduce_d, dim3(1), dim3(256), 0, st, s->part.p, ns * B, 0.5, s->Jd.p);
  HIPCHK(hipGetLastError());
  s->have_adj = true;
  s->kept = 0;
  return GJW_OK;
}

// Reference gradient (w_grad.m:20-115) on the plain path's stored wavefield:
// per shot, a forward-type sweep with negated conductivity in the inner
// coefficients, the residual written into Ez at the receivers before each
// step, no source, and the correlation g -= u(nt-it) .* a, g *= dt.
static int grad_ref(gjw_survey* s, const float* adj, float* amax) {
  const int ns = s->ns, nt = s->gr.nt, NX = s->NX, NY = s->NY, P = s->P;
  const long long per = (long long)s->gr.m * s->gr.n, FS = s->FS, FD = (long long)s->LD * s->gr.n;
  hipStream_t st = s->ctx->stream;
  const int L = s->tile_L > 0 ? s->tile_L : 16;
  const int K = ck_interval(s);
  if (!s->coef_ref.p) {  // inner Ceze / Cezhy with -sigma
    std::vector<float> h((size_t)2 * FS, 0.f);
    const double e0 = s->gr.eps0, dt = s->gr.dt, dx = s->gr.dx;
    for (int j = 0; j <= NY; ++j)
      for (int i = 0; i <= NX; ++i) {
        if (ekind(NX, NY, P, i, j) != KIND_INNER) continue;
        const double ee = s->pe[(size_t)i + (size_t)j * NX], sg = -s->ps[(size_t)i + (size_t)j * NX];
        const double D = 2.0 * ee * e0 + dt * sg;
        h[2 * ((size_t)j * s->LD + i)] = (float)((2.0 * ee * e0 - dt * sg) / D);
        h[2 * ((size_t)j * s->LD + i) + 1] = (float)((2.0 * dt / dx) / D);
      }
    TRY(s->coef_ref.ensure(h.size()));
    HIPCHK(hipMemcpy(s->coef_ref.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  }
  TRY(s->zero_sc.ensure(ns));
  TRY(launch_fill(s->zero_sc.p, ns, 0.f, st));
  DevGrid g = dev_grid(s);
  g.cc = reinterpret_cast<const float2*>(s->coef_ref.p);
  TRY(s->acc.ensure((size_t)2 * FD));
  TRY(launch_fill(s->acc.p, 2LL * FD, 0.f, st));
  TRY(ensure_fwd_state(s, 1));
  FwdBufs b = fwd_bufs(s, 1);
  for (int q = 0; q < ns; ++q) {
    DevShots sh = dev_shots(s, q);
    sh.sc = s->zero_sc.p + q;
    TRY(zero_fwd_state(s, 1));
    const float* U = s->frames.p + (long long)q * (nt + 1) * FD;
    const int t0 = (int)s->rec_ptr[q], t1 = (int)s->rec_ptr[q + 1];
    int cur = 0;
    for (int it = 0; it < nt; ++it) {
      if (t1 > t0) {
        hipLaunchKernelGGL(k_ref_overwrite, dim3(1), dim3(64), 0, st, g, nt, t0, t1, s->d_trr.p,
                           s->d_trc.p, adj, nt - 1 - it, b.E[cur]);
        HIPCHK(hipGetLastError());
      }
      TRY(fwd_launch(s, g, sh, 1, L, b, cur, it, nullptr, nullptr, 0));
      cur ^= 1;
      hipLaunchKernelGGL(k_ref_corr, dim3((unsigned)cdiv(per, 256)), dim3(256), 0, st, g, b.E[cur],
                         U, nt - it, nt, s->gr.dt, s->acc.p,
                         (amax && it == nt - 1 && q == ns - 1) ? amax : nullptr);
      HIPCHK(hipGetLastError());
      if (it % K == K - 1 || it == nt - 1) {
        hipLaunchKernelGGL(k_flush, dim3((unsigned)cdiv(per, 256)), dim3(256), 0, st, s->gr.m,
                           s->gr.n, s->LD, s->di0, 1, s->acc.p, FD, nullptr, nullptr,
                           s->ge.p + q * grad_stride(s), s->gsg.p + q * grad_stride(s), -s->gr.dt,
                           0LL);
        HIPCHK(hipGetLastError());
      }
    }
  }
  return GJW_OK;
}

static int grad_run(gjw_survey* s, const float* adj_in, int32_t mode, double* g_eps,
                    double* g_sig, double* g_w, float* a_max = nullptr) {
  if (mode != GJW_GRAD_EXACT && mode != GJW_GRAD_REF)
    return fail(GJW_ERR_ARG, "gjw_gradient: unknown mode %d", mode);
  if (!s->kept) return fail(GJW_ERR_STATE, "gjw_gradient: run gjw_forward(keep=1) first");
  if (mode == GJW_GRAD_REF && s->kept != 2)
    return fail(GJW_ERR_STATE,
                "gjw_gradient: REF mode needs the stored wavefield of the plain path "
                "(GJW_OPT_PATH = GJW_PATH_PLAIN, then gjw_forward(keep=1))");
  hipStream_t st = s->ctx->stream;
  const int ns = s->ns, nt = s->gr.nt;
  const long long per = (long long)s->gr.m * s->gr.n, n = (long long)s->ntr * nt;
  const float* adj = nullptr;
  if (adj_in) {
    TRY(s->adj.ensure(std::max<long long>(n, 1)));
    TRY(host_or_dev_copy(s->adj.p, adj_in, n * 4, st));
    adj = s->adj.p;
  } else {
    if (!s->have_adj) return fail(GJW_ERR_STATE, "gjw_gradient: no adjoint source (call gjw_misfit)");
    adj = s->adj.p;
  }
  TRY(s->ge.ensure(grad_len(s)));
  TRY(s->gsg.ensure(grad_len(s)));
  TRY(launch_fill_d(s->ge.p, grad_len(s), 0.0, st));
  TRY(launch_fill_d(s->gsg.p, grad_len(s), 0.0, st));
  float* gwf = nullptr;
  DBuf<float> gwtmp;
  if (g_w) {
    TRY(gwtmp.ensure((size_t)ns * nt));
    TRY(launch_fill(gwtmp.p, (long long)ns * nt, 0.f, st));
    gwf = gwtmp.p;
  }
  HIPCHK(hipEventRecord(s->ctx->ev0, st));
  DBuf<float> amtmp;
  if (a_max) TRY(amtmp.ensure(per));
  int rc = mode == GJW_GRAD_REF ? grad_ref(s, adj, a_max ? amtmp.p : nullptr)
           : s->kept == 1       ? grad_batched(s, adj, gwf)
                                : grad_plain_exact(s, adj, gwf);
  if (rc != GJW_OK) {
    gwtmp.release();
    return rc;
  }
  HIPCHK(hipEventRecord(s->ctx->ev1, st));
  TRY(finish_timing(s));
  s->last_cells = (double)(s->NX + 1) * (s->NY + 1) * nt * ns * (s->kept == 1 ? 2 : 1);
  if (g_eps) HIPCHK(hipMemcpyAsync(g_eps, s->ge.p, grad_len(s) * 8, hipMemcpyDefault, st));
  if (g_sig) HIPCHK(hipMemcpyAsync(g_sig, s->gsg.p, grad_len(s) * 8, hipMemcpyDefault, st));
  if (g_w) {
    std::vector<float> h((size_t)ns * nt);
    HIPCHK(hipMemcpyAsync(h.data(), gwf, h.size() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> hd(h.begin(), h.end());
    HIPCHK(hipMemcpy(g_w, hd.data(), hd.size() * 8, hipMemcpyDefault));
  }
  if (a_max) HIPCHK(hipMemcpyAsync(a_max, amtmp.p, per * 4, hipMemcpyDefault, st));
  HIPCHK(hipStreamSynchronize(st));
  gwtmp.release();
  amtmp.release();
  return GJW_OK;
}

extern "C" int gjw_gradient(gjw_survey* s, const float* adj, int32_t mode, double* g_eps,
                            double* g_sig, double* g_w, float* a_max) {
  TRY(check_ready(s, "gjw_gradient"));
  if (a_max && mode != GJW_GRAD_REF)
    return fail(GJW_ERR_ARG, "gjw_gradient: a_max is only produced in REF mode");
  if (g_w && mode != GJW_GRAD_EXACT)
    return fail(GJW_ERR_ARG, "gjw_gradient: g_w is only produced in EXACT mode");
  return grad_run(s, adj, mode, g_eps, g_sig, g_w, a_max);
}

// Batched FWI gradient in two parts.  fwi_batched_prepare allocates every
// buffer and runs the autotuner (host synchronisation); fwi_batched_enqueue
// then only enqueues kernels and device copies on the survey stream (forward
// sweep with checkpoints, L2 residual and misfit, adjoint sweep with
// recomputation), so it can be captured as one hipGraph.  The observed traces
// are expected in s->obs.
static int fwi_batched_prepare(gjw_survey* s) {
  const int ns = s->ns, nt = s->gr.nt;
  const long long n = (long long)s->ntr * nt;
  const long long FD = (long long)s->LD * s->gr.n;
  TRY(ensure_fwd_state(s, ns));
  TRY(s->tr.ensure(std::max<long long>(n, 1)));
  TRY(zero_fwd_state(s, ns));
  TRY(autotune_L(s, ns));
  s->ck_K = ck_interval(s);
  s->ck_n = (int)cdiv(nt, s->ck_K);
  TRY(s->ck.ensure((size_t)s->ck_n * ck_stride(s)));
  TRY(s->adj.ensure(std::max<long long>(n, 1)));
  TRY(s->obs.ensure(std::max<long long>(n, 1)));
  TRY(s->part.ensure(1024));
  TRY(s->Jd.ensure(1));
  TRY(s->ge.ensure(grad_len(s)));
  TRY(s->gsg.ensure(grad_len(s)));
  TRY(s->frames.ensure((size_t)ns *
                       (frames_bf16(s, s->ck_K) ? s->ck_K / 2 + 1 : s->ck_K + 1) * FD));
  TRY(s->acc.ensure((size_t)2 * ns * FD));
  TRY(s->adjf.ensure((size_t)10 * ns * s->FS));
  return GJW_OK;
}

static int fwi_fwd_enqueue(gjw_survey* s) {
  const int ns = s->ns, nt = s->gr.nt;
  DevGrid g = dev_grid(s);
  DevShots sh = dev_shots(s, 0);
  FwdBufs b = fwd_bufs(s, ns);
  TRY(zero_fwd_state(s, ns));
  int cur = 0, xc = 0;
  for (int it = 0; it < nt;) {
    if (it % s->ck_K == 0) TRY(ck_save(s, b, cur, xc, it / s->ck_K));
    int taken = 0;
    TRY(fwd_advance(s, g, sh, ns, b, cur, xc, it, std::min(nt - it, s->ck_K - it % s->ck_K),
                    s->tr.p, nullptr, 0, 0, &taken));
    it += taken;
  }
  s->cur = cur;
  return GJW_OK;
}

static int fwi_adj_enqueue(gjw_survey* s) {
  hipStream_t st = s->ctx->stream;
  TRY(launch_fill_d(s->ge.p, grad_len(s), 0.0, st));
  TRY(launch_fill_d(s->gsg.p, grad_len(s), 0.0, st));
  return grad_batched(s, s->adj.p, nullptr);
}

static int fwi_batched_enqueue(gjw_survey* s) {
  TRY(fwi_fwd_enqueue(s));
  TRY(misfit_enqueue(s, s->misfit_kind));
  TRY(adj_filter(s, 0, s->ntr));
  return fwi_adj_enqueue(s);
}

// Replay the sequence `enqueue` puts on stream st, capturing it first (or
// again, when the key of addresses and parameters changed).  No timing events
// are recorded while capturing.
template <class F>
static int graph_run(GraphCache& gc, hipStream_t st, const std::vector<uintptr_t>& key,
                     int* ksample, F&& enqueue) {
  if (!gc.x || key != gc.key) {
    gc.reset();
    const int ks = *ksample;
    *ksample = 0;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = enqueue();
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &gr);
    *ksample = ks;
    if (rc != GJW_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    HIPCHK(ec);
    gc.g = gr;
    HIPCHK(hipGraphInstantiate(&gc.x, gc.g, nullptr, nullptr, 0));
    gc.key = key;
  }
  HIPCHK(hipGraphLaunch(gc.x, st));
  return GJW_OK;
}

// every device address and launch parameter the enqueued sequence bakes in
static std::vector<uintptr_t> fwi_graph_key(const gjw_survey* s) {
  auto u = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
  auto fbits = [](float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (uintptr_t)b;
  };
  auto dbits = [](double f) {
    uint64_t b;
    std::memcpy(&b, &f, 8);
    return (uintptr_t)b;
  };
  return {u(s->fld.p),    u(s->tr.p),    u(s->ck.p),     u(s->adj.p),   u(s->obs.p),
          u(s->part.p),   u(s->Jd.p),    u(s->ge.p),     u(s->gsg.p),   u(s->frames.p),
          u(s->acc.p),    u(s->adjf.p),  u(s->coef.p),   u(s->fe.p),    u(s->fs.p),
          u(s->d_pml.p),  u(s->d_si.p),  u(s->d_sj.p),   u(s->d_sc.p),  u(s->d_wav.p),
          u(s->d_cptr.p), u(s->d_crow.p), u(s->d_cid.p), u(s->d_trs.p), u(s->d_trr.p),
          u(s->d_trc.p),  (uintptr_t)s->frames16, (uintptr_t)s->row_pairs, (uintptr_t)s->ns, (uintptr_t)s->ntr, (uintptr_t)s->shot_grads,
          (uintptr_t)s->gr.nt, (uintptr_t)s->tuned_L, (uintptr_t)s->tuned_L1,
          (uintptr_t)s->tuned_L4, (uintptr_t)s->tuned_L8, (uintptr_t)s->fuse_eff,
          (uintptr_t)s->tuned_La2,
          (uintptr_t)s->tuned_La4, (uintptr_t)s->tuned_La6, (uintptr_t)s->tuned_La8,
          (uintptr_t)s->adj_fuse,
          (uintptr_t)s->ck_K, (uintptr_t)s->fuse, (uintptr_t)s->snap,
          (uintptr_t)s->rrow_min, (uintptr_t)s->rrow_max, (uintptr_t)s->pml_idx.size(),
          dbits(s->flt_lo), dbits(s->flt_hi), (uintptr_t)s->flt_steep,
          fbits(s->bx0), fbits(s->by0)};
}

// L2: the whole gradient is one graph.  Hilbert-based misfits run their FFTs
// eagerly between a forward-sweep graph and an adjoint-sweep graph.
static int fwi_graph_run(gjw_survey* s) {
  hipStream_t st = s->ctx->stream;
  std::vector<uintptr_t> key = fwi_graph_key(s);
  if (s->misfit_kind == GJW_MISFIT_L2 && !filter_on(s))
    return graph_run(s->g_all, st, key, &s->ksample, [&] { return fwi_batched_enqueue(s); });
  // FFT misfits and the adjoint-source filter run eagerly between the sweeps
  TRY(graph_run(s->g_fwd, st, key, &s->ksample, [&] { return fwi_fwd_enqueue(s); }));
  TRY(misfit_enqueue(s, s->misfit_kind));
  TRY(adj_filter(s, 0, s->ntr));
  return graph_run(s->g_adj, st, key, &s->ksample, [&] { return fwi_adj_enqueue(s); });
}

extern "C" int gjw_fwi_gradient(gjw_survey* s, const float* d_obs, double* g_eps, double* g_sig,
                                double* J) {
  TRY(check_ready(s, "gjw_fwi_gradient"));
  if (!d_obs) return fail(GJW_ERR_ARG, "gjw_fwi_gradient: d_obs is NULL");
  hipStream_t st = s->ctx->stream;
  hipEvent_t e0;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventRecord(e0, st));
  int rc;
  if (s->path == GJW_PATH_PLAIN) {
    rc = fwi_plain(s, d_obs);
    if (rc == GJW_OK) {
      if (hipEventRecord(s->ctx->ev1, st) != hipSuccess) rc = fail(GJW_ERR_HIP, "event record");
      if (rc == GJW_OK) rc = kev_resolve(s);
      if (rc == GJW_OK && g_eps)
        rc = host_or_dev_copy(g_eps, s->ge.p, grad_len(s) * 8, st);
      if (rc == GJW_OK && g_sig)
        rc = host_or_dev_copy(g_sig, s->gsg.p, grad_len(s) * 8, st);
    }
  } else {
    rc = fwi_batched_prepare(s);
    if (rc == GJW_OK) rc = host_or_dev_copy(s->obs.p, d_obs, (long long)s->ntr * s->gr.nt * 4, st);
    // the first call runs eagerly (and settles the autotuner); with
    // GJW_OPT_GRAPH later calls replay the captured sequence
    const bool use_graph = s->graph_on && s->graph_calls++ > 0;
    if (rc == GJW_OK) rc = use_graph ? fwi_graph_run(s) : fwi_batched_enqueue(s);
    if (rc == GJW_OK && hipEventRecord(s->ctx->ev1, st) != hipSuccess)
      rc = fail(GJW_ERR_HIP, "event record");
    if (rc == GJW_OK) rc = kev_resolve(s);
    if (rc == GJW_OK) {
      s->kept = 1;
      s->have_adj = true;
      if (g_eps) rc = host_or_dev_copy(g_eps, s->ge.p, grad_len(s) * 8, st);
      if (rc == GJW_OK && g_sig) rc = host_or_dev_copy(g_sig, s->gsg.p, grad_len(s) * 8, st);
    }
  }
  if (rc != GJW_OK) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, s->ctx->ev1));
  HIPCHK(hipEventDestroy(e0));
  s->last_ms = ms;
  s->last_cells = (double)(s->NX + 1) * (s->NY + 1) * s->gr.nt * s->ns *
                  (s->path == GJW_PATH_PLAIN ? 2 : 3);
  if (J) HIPCHK(hipMemcpy(J, s->Jd.p, 8, hipMemcpyDefault));
  return GJW_OK;
}

// --------------------------------------------------------------------------
// several surveys in shared launches
// --------------------------------------------------------------------------
// A batch runs the FWI gradients of up to kMaxSurveys surveys (each with its
// own grid, model, shots, trace length and misfit) with one launch per time
// step (or pair of steps) covering the tiles of every survey still active.
// All surveys start at step 0 and share one checkpoint interval K; a launch
// fuses two steps only where no survey starts or ends between them.
struct gjw_batch {
  gjw_ctx* ctx = nullptr;
  std::vector<gjw_survey*> sv;
  DBuf<DevSurvey> dsv;  // the table on the device (its address keys the captured graphs)
  MultiTable tab{};     // the table the kernels take by value
  int K = 0, T = 0, fuse = 2, afuse = 2, calls = 0;  // afuse: adjoint steps per launch
  int L8 = 0, La6 = 0;  // strip widths of the shared 8-step forward / 6-step adjoint launches
  bool keep = true;  // forward sweep saves checkpoints (gradients); not for objectives
  GraphCache g_all, g_fwd, g_adj;
  double last_ms = 0, last_cells = 0;
};

extern "C" int gjw_batch_create(gjw_ctx* ctx, int32_t n, gjw_survey* const* surveys,
                                gjw_batch** out) {
  if (!ctx || !surveys || !out) return fail(GJW_ERR_ARG, "gjw_batch_create: NULL argument");
  if (n < 1 || n > kMaxSurveys)
    return fail(GJW_ERR_ARG, "gjw_batch_create: 1 to %d surveys", kMaxSurveys);
  for (int k = 0; k < n; ++k) {
    if (!surveys[k] || surveys[k]->ctx != ctx)
      return fail(GJW_ERR_ARG, "gjw_batch_create: survey %d is NULL or of another context", k);
    for (int q = 0; q < k; ++q)
      if (surveys[q] == surveys[k]) return fail(GJW_ERR_ARG, "gjw_batch_create: survey repeated");
  }
  gjw_batch* b = new gjw_batch;
  b->ctx = ctx;
  b->sv.assign(surveys, surveys + n);
  *out = b;
  return GJW_OK;
}

extern "C" int gjw_batch_destroy(gjw_batch* b) {
  if (!b) return GJW_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  b->dsv.release();
  b->g_all.reset();
  b->g_fwd.reset();
  b->g_adj.reset();
  delete b;
  return GJW_OK;
}

static DevSurvey dev_survey(gjw_survey* s) {
  DevSurvey d;
  d.g = dev_grid(s);
  d.s = dev_shots(s, 0);
  d.ns = s->ns;
  d.fsz = (long long)s->ns * s->FS;
  d.fld = s->fld.p;
  d.adjf = s->adjf.p;
  d.tr = s->tr.p;
  d.adj = s->adj.p;
  d.frames = s->frames.p;
  d.FD = (long long)s->LD * s->gr.n;
  d.fstride = (long long)(s->ck_K + 1) * d.FD;
  d.acc = s->acc.p;
  return d;
}

// one multi-survey launch over surveys ids (steps: 1 or 2; adj: adjoint kernels)
static int batch_launch(gjw_batch* b, const std::vector<int>& ids, int steps, bool adj,
                        bool frames, const int* it, const int* cur, const int* xc,
                        const int* fslot, int rec) {
  MultiArgs a{};
  a.n = (int)ids.size();
  a.rec = rec;
  int blk = 0;
  const int rows = steps == 1 ? 64 : 64 - 2 * steps;
  for (int k = 0; k < a.n; ++k) {
    gjw_survey* s = b->sv[ids[k]];
    // (the adjoint widths are tuned by single-survey gradient sweeps)
    const int La6 = s->tuned_La6 > 0 ? s->tuned_La6 : s->tuned_L8 > 0 ? s->tuned_L8 : 64;
    const int La4 = s->tuned_La4 > 0 ? s->tuned_La4 : s->tuned_L4;
    const int La2 = s->tuned_La2 > 0 ? s->tuned_La2 : s->tuned_L;
    const int L = steps == 8   ? (b->L8 > 0 ? b->L8 : s->tuned_L8)
                  : steps == 6 ? (b->La6 > 0 ? b->La6 : La6)
                  : steps == 4 ? (adj ? La4 : s->tuned_L4)
                  : steps == 2 ? (adj ? La2 : s->tuned_L)
                               : s->tuned_L1;
    a.blk0[k] = blk;
    a.sid[k] = ids[k];
    a.it[k] = it[ids[k]];
    a.L[k] = L;
    a.cur[k] = cur[ids[k]];
    a.xc[k] = xc ? xc[ids[k]] : 0;
    a.fslot[k] = fslot ? fslot[ids[k]] : -1;
    const long long nb = steps == 1 ? cdiv(s->NX + 1, rows) * cdiv(cdiv(s->NY + 1, L), 4) * s->ns
                                    : cdiv(cdiv(s->NX + 1, rows), 4) * cdiv(s->NY + 1, L) * s->ns;
    blk += (int)(cdiv(nb, 8) * 8);
  }
  a.blk0[a.n] = blk;
  if (blk == 0) return GJW_OK;
  hipStream_t st = b->ctx->stream;
  const dim3 grid((unsigned)blk), block(64, 4);
  if ((adj && steps == 8) || (!adj && steps == 6))
    return fail(GJW_ERR_STATE, "batch_launch: no %d-step %s kernel", steps,
                adj ? "adjoint" : "forward");
  if (!adj && steps == 1)
    hipLaunchKernelGGL(k_fwd_step_multi, grid, block, 0, st, b->tab, a);
  else if (!adj && steps == 8 && frames)
    hipLaunchKernelGGL((k_fwd_step8_multi<true>), grid, block, 0, st, b->tab, a);
  else if (!adj && steps == 8)
    hipLaunchKernelGGL((k_fwd_step8_multi<false>), grid, block, 0, st, b->tab, a);
  else if (!adj && steps == 4 && frames)
    hipLaunchKernelGGL((k_fwd_stepN_multi<4, true>), grid, block, 0, st, b->tab, a);
  else if (!adj && steps == 4)
    hipLaunchKernelGGL((k_fwd_stepN_multi<4, false>), grid, block, 0, st, b->tab, a);
  else if (!adj && frames)
    hipLaunchKernelGGL((k_fwd_stepN_multi<2, true>), grid, block, 0, st, b->tab, a);
  else if (!adj)
    hipLaunchKernelGGL((k_fwd_stepN_multi<2, false>), grid, block, 0, st, b->tab, a);
  else if (steps == 1)
    hipLaunchKernelGGL(k_adj_step_multi, grid, block, 0, st, b->tab, a);
  else if (steps == 6)
    hipLaunchKernelGGL((k_adj_stepN_multi<6>), grid, block, 0, st, b->tab, a);
  else if (steps == 4)
    hipLaunchKernelGGL((k_adj_stepN_multi<4>), grid, block, 0, st, b->tab, a);
  else
    hipLaunchKernelGGL((k_adj_stepN_multi<2>), grid, block, 0, st, b->tab, a);
  HIPCHK(hipGetLastError());
  return GJW_OK;
}

// Strip width of the shared deep launches, tuned for the batch as a whole:
// a survey's own tuning sees the survey alone, which a small survey cannot
// fill, and picks narrow strips (many blocks) that cost the filled batch
// halo work.  Times the 8-step forward of every survey on its current state
// at each candidate (the state is zeroed before any sweep); the 6-step
// adjoint takes 3/4 of that width (the ratio measured on single surveys).
static int batch_tune(gjw_batch* b) {
  if (b->L8 > 0 || b->fuse < 8) return GJW_OK;
  const int S = (int)b->sv.size();
  bool fixed = true;  // every survey's width set by GJW_OPT_TILE_L: keep those
  for (gjw_survey* s : b->sv) fixed &= s->tile_L > 0;
  if (fixed) return GJW_OK;
  hipStream_t st = b->ctx->stream;
  std::vector<int> all(S);
  int it[kMaxSurveys] = {}, cur[kMaxSurveys] = {}, xc[kMaxSurveys] = {};
  for (int k = 0; k < S; ++k) all[k] = k;
  float best = 1e30f;
  int bestL = 0;
  for (int L : {48, 64, 96, 128}) {
    b->L8 = L;
    for (int r = -1; r < 3; ++r) {
      if (r == 0) HIPCHK(hipEventRecord(b->ctx->ev0, st));
      TRY(batch_launch(b, all, 8, false, false, it, cur, xc, nullptr, 0));
      for (int k = 0; k < S; ++k) cur[k] ^= 1, xc[k] ^= 1;
    }
    HIPCHK(hipEventRecord(b->ctx->ev1, st));
    HIPCHK(hipEventSynchronize(b->ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
    if (ms < best) {
      best = ms;
      bestL = L;
    }
  }
  b->L8 = bestL;
  b->La6 = std::max(32, bestL * 3 / 4 / 8 * 8);
  return GJW_OK;
}

// forward sweep of every survey with checkpoints (shared launches)
static int batch_fwd_enqueue(gjw_batch* b) {
  const int S = (int)b->sv.size(), T = b->T, K = b->K;
  int nt[kMaxSurveys], cur[kMaxSurveys], xc[kMaxSurveys], itv[kMaxSurveys];
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    nt[k] = s->gr.nt;
    TRY(zero_fwd_state(s, s->ns));
    cur[k] = xc[k] = 0;
  }
  for (int tau = 0; tau < T;) {
    std::vector<int> act;
    int room = std::min(b->fuse, K - tau % K);  // steps this launch may take
    for (int k = 0; k < S; ++k)
      if (tau < nt[k]) {
        act.push_back(k);
        gjw_survey* s = b->sv[k];
        if (b->keep && tau % K == 0) {
          FwdBufs fb = fwd_bufs(s, s->ns);
          TRY(ck_save(s, fb, cur[k], xc[k], tau / K));
        }
        room = std::min(room, nt[k] - tau);
        itv[k] = tau;
      }
    const int steps = room >= 8 ? 8 : room >= 4 ? 4 : room >= 2 ? 2 : 1;
    TRY(batch_launch(b, act, steps, false, false, itv, cur, xc, nullptr, 1));
    for (int k : act) {
      cur[k] ^= 1;
      if (steps > 1) xc[k] ^= 1;
    }
    tau += steps;
  }
  for (int k = 0; k < S; ++k) b->sv[k]->cur = cur[k];
  return GJW_OK;
}

// l2: the surveys whose misfit and adjoint source go into a captured graph
// (L2 without the FFT filter), else the others
static int batch_misfit_enqueue(gjw_batch* b, bool l2) {
  for (gjw_survey* s : b->sv)
    if ((s->misfit_kind == GJW_MISFIT_L2 && !filter_on(s)) == l2) {
      TRY(misfit_enqueue(s, s->misfit_kind));
      TRY(adj_filter(s, 0, s->ntr));
    }
  return GJW_OK;
}

// adjoint sweep of every survey, segment by segment, recomputing the forward
static int batch_adj_enqueue(gjw_batch* b) {
  const int S = (int)b->sv.size(), T = b->T, K = b->K;
  hipStream_t st = b->ctx->stream;
  int nt[kMaxSurveys], cur[kMaxSurveys], xc[kMaxSurveys], itv[kMaxSurveys], fsl[kMaxSurveys],
      acur[kMaxSurveys];
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    const long long FD = (long long)s->LD * s->gr.n;
    nt[k] = s->gr.nt;
    acur[k] = 0;
    TRY(launch_fill_d(s->ge.p, grad_len(s), 0.0, st));
    TRY(launch_fill_d(s->gsg.p, grad_len(s), 0.0, st));
    TRY(launch_fill(s->acc.p, 2LL * s->ns * FD, 0.f, st));
    TRY(adj_init(s, s->ns, 0, s->adj.p));
  }
  const int nseg = (int)cdiv(T, K);
  for (int c = nseg - 1; c >= 0; --c) {
    const int k0 = c * K;
    int k1[kMaxSurveys], kmax = k0;
    std::vector<int> seg;
    for (int k = 0; k < S; ++k) {
      k1[k] = std::min(k0 + K, nt[k]);
      if (k0 >= nt[k]) continue;
      seg.push_back(k);
      kmax = std::max(kmax, k1[k]);
      gjw_survey* s = b->sv[k];
      FwdBufs fb = fwd_bufs(s, s->ns);
      cur[k] = xc[k] = 0;
      TRY(ck_restore(s, fb, 0, 0, c));
      const long long per = (long long)s->gr.m * s->gr.n;
      hipLaunchKernelGGL(k_domain_frame, dim3((unsigned)cdiv(per * s->ns, 256)), dim3(256), 0, st,
                         dev_grid(s), s->ns, fb.E[0], s->frames.p,
                         (long long)(K + 1) * s->LD * s->gr.n, 0);
      HIPCHK(hipGetLastError());
    }
    for (int tau = k0; tau < kmax;) {
      std::vector<int> act;
      int room = b->fuse;
      for (int k : seg)
        if (tau < k1[k]) {
          act.push_back(k);
          room = std::min(room, k1[k] - tau);
          itv[k] = tau;
          fsl[k] = tau + 1 - k0;
        }
      const int steps = room >= 8 ? 8 : room >= 4 ? 4 : room >= 2 ? 2 : 1;
      TRY(batch_launch(b, act, steps, false, true, itv, cur, xc, fsl, 0));
      for (int k : act) {
        cur[k] ^= 1;
        if (steps > 1) xc[k] ^= 1;
      }
      tau += steps;
    }
    for (int tau = kmax - 1; tau >= k0;) {
      std::vector<int> act;
      // steps this launch may take: none below k0, and no survey may start
      // (take its first adjoint step) inside the launch
      int room = std::min(b->afuse, tau - k0 + 1);
      for (int k : seg)
        if (tau <= k1[k] - 1) {
          act.push_back(k);
          itv[k] = tau;
          fsl[k] = tau - k0;
        } else {
          room = std::min(room, tau - (k1[k] - 1));
        }
      const int steps = room >= 6 ? 6 : room >= 4 ? 4 : room >= 2 ? 2 : 1;
      TRY(batch_launch(b, act, steps, true, false, itv, acur, nullptr, fsl, 0));
      for (int k : act) acur[k] ^= 1;
      tau -= steps;
    }
    for (int k : seg) {
      gjw_survey* s = b->sv[k];
      const long long per = (long long)s->gr.m * s->gr.n, FD = (long long)s->LD * s->gr.n;
      hipLaunchKernelGGL(k_flush, dim3((unsigned)cdiv(per, 256)), dim3(256), 0, st, s->gr.m,
                         s->gr.n, s->LD, s->di0, s->ns, s->acc.p, FD, s->fe.p, s->fs.p, s->ge.p,
                         s->gsg.p, 1.0, grad_stride(s));
      HIPCHK(hipGetLastError());
    }
  }
  return GJW_OK;
}

// Checkpoint interval of a batch that needs the least memory, when the
// interval K would put checkpoints and frames of all surveys over half of
// the device: ceil(nt_s / K) checkpoints of c_s bytes plus K + 1 frames of f_s
// bytes per survey, minimised over multiples of q near sqrt(sum nt c / sum f).
static int batch_mem_interval(const gjw_batch* b, int K, int q, int T) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return K;
  }
  double ntc = 0, fsum = 0;
  for (const gjw_survey* s : b->sv) {
    ntc += (double)s->gr.nt * 4.0 * (double)ck_stride(s);
    fsum += 4.0 * s->ns * (double)s->LD * s->gr.n;
  }
  auto bytes = [&](int kk) {
    double t = 0;
    for (const gjw_survey* s : b->sv)
      t += std::ceil((double)s->gr.nt / kk) * 4.0 * (double)ck_stride(s) +
           (kk + 1) * 4.0 * s->ns * (double)s->LD * s->gr.n;
    return t;
  };
  if (bytes(K) <= 0.5 * (double)tot) return K;
  const int k0 = std::max(1, (int)std::lround(std::sqrt(ntc / std::max(fsum, 1.0))));
  int best = K;
  for (int kk = std::max(q, k0 / q * q); kk <= std::min(T, (k0 / q + 1) * q); kk += q)
    if (bytes(kk) < bytes(best)) best = kk;
  return best;
}

extern "C" int gjw_batch_fwi_gradient(gjw_batch* b, const float* const* d_obs,
                                      double* const* g_eps, double* const* g_sig, double* J) {
  if (!b || !d_obs) return fail(GJW_ERR_ARG, "gjw_batch_fwi_gradient: NULL argument");
  const int S = (int)b->sv.size();
  HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  int T = 0, fuse = 2, graph = 1, all_l2 = 1, K = 0;
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    TRY(check_ready(s, "gjw_batch_fwi_gradient"));
    if (!d_obs[k]) return fail(GJW_ERR_ARG, "gjw_batch_fwi_gradient: d_obs[%d] is NULL", k);
    if (s->path != GJW_PATH_BATCHED)
      return fail(GJW_ERR_ARG, "gjw_batch_fwi_gradient: survey %d is on the plain path", k);
    T = std::max(T, s->gr.nt);
    graph &= s->graph_on;
    all_l2 &= s->misfit_kind == GJW_MISFIT_L2 && !filter_on(s);
    K = std::max(K, s->ck_every);
  }
  const bool user_K = K > 0;
  if (!user_K) {
    // about 2 sqrt(T) in whole launches of both depths, or, when the
    // checkpoints and frames of all surveys would pass half of the device,
    // the interval that needs the least memory (as ck_interval for one survey)
    int f0 = 8, a0 = 6;
    for (int k = 0; k < S; ++k) {
      f0 = std::min(f0, batch_fwd_depth(b->sv[k]));
      a0 = std::min(a0, batch_adj_depth(b->sv[k]));
    }
    const int q = f0 * a0 / std::gcd(f0, a0);
    K = std::max(1, (int)std::ceil(2.0 * std::sqrt((double)T)));
    K = std::min(T, (K + q - 1) / q * q);
    K = batch_mem_interval(b, K, q, T);
  }
  b->T = T;
  // forward and adjoint depths: the shallowest any survey asks for (the
  // shared adjoint kernels go up to six steps)
  int afuse = 6;
  fuse = 8;
  auto prepare_all = [&](int Kc) {
    for (int k = 0; k < S; ++k) {
      gjw_survey* s = b->sv[k];
      s->ck_batch = Kc;
      const int rc = fwi_batched_prepare(s);
      s->ck_batch = 0;
      TRY(rc);
      fuse = std::min(fuse, batch_fwd_depth(s));
      afuse = std::min(afuse, batch_adj_depth(s));
    }
    return (int)GJW_OK;
  };
  TRY(prepare_all(K));
  if (!user_K) {  // about 2 sqrt(T), in whole launches of both depths (as ck_interval)
    const int q = fuse * afuse / std::gcd(fuse, afuse);
    const int K2 = std::min(T, (K + q - 1) / q * q);
    if (K2 != K) {
      K = K2;
      TRY(prepare_all(K));  // the autotuner's results are kept; buffers for the new K
    }
  }
  b->K = K;
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    s->ck_K = K;
    s->ck_n = (int)cdiv(s->gr.nt, K);
    TRY(host_or_dev_copy(s->obs.p, d_obs[k], (long long)s->ntr * s->gr.nt * 4, st));
  }
  std::vector<DevSurvey> hs(S);
  b->fuse = fuse;
  b->afuse = afuse;
  for (int k = 0; k < S; ++k) hs[k] = dev_survey(b->sv[k]);
  TRY(b->dsv.ensure(S));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(b->dsv.p, hs.data(), S * sizeof(DevSurvey), hipMemcpyHostToDevice));
  for (int k = 0; k < S; ++k) b->tab.sv[k] = hs[k];
  TRY(batch_tune(b));
  HIPCHK(hipEventRecord(b->ctx->ev0, st));
  // the first call runs eagerly; with GJW_OPT_GRAPH on every survey, later
  // calls replay captured sweeps (one graph when every misfit is L2)
  const bool use_graph = graph && b->calls++ > 0;
  if (!use_graph) {
    TRY(batch_fwd_enqueue(b));
    TRY(batch_misfit_enqueue(b, true));
    TRY(batch_misfit_enqueue(b, false));
    TRY(batch_adj_enqueue(b));
  } else {
    std::vector<uintptr_t> key = {reinterpret_cast<uintptr_t>(b->dsv.p), (uintptr_t)K,
                                  (uintptr_t)T, (uintptr_t)fuse, (uintptr_t)afuse,
                                  (uintptr_t)b->L8, (uintptr_t)b->La6};
    for (gjw_survey* s : b->sv) {
      const std::vector<uintptr_t> ks = fwi_graph_key(s);
      key.insert(key.end(), ks.begin(), ks.end());
      key.push_back((uintptr_t)s->misfit_kind);
    }
    int ks0 = 0;
    if (all_l2) {
      TRY(graph_run(b->g_all, st, key, &ks0, [&] {
        TRY(batch_fwd_enqueue(b));
        TRY(batch_misfit_enqueue(b, true));
        return batch_adj_enqueue(b);
      }));
    } else {
      TRY(graph_run(b->g_fwd, st, key, &ks0, [&] {
        TRY(batch_fwd_enqueue(b));
        return batch_misfit_enqueue(b, true);
      }));
      TRY(batch_misfit_enqueue(b, false));
      TRY(graph_run(b->g_adj, st, key, &ks0, [&] { return batch_adj_enqueue(b); }));
    }
  }
  HIPCHK(hipEventRecord(b->ctx->ev1, st));
  double cells = 0;
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    s->kept = 0;
    s->have_adj = true;
    if (g_eps && g_eps[k]) TRY(host_or_dev_copy(g_eps[k], s->ge.p, grad_len(s) * 8, st));
    if (g_sig && g_sig[k]) TRY(host_or_dev_copy(g_sig[k], s->gsg.p, grad_len(s) * 8, st));
    if (J) HIPCHK(hipMemcpyAsync(J + k, s->Jd.p, 8, hipMemcpyDefault, st));
    cells += 3.0 * (s->NX + 1) * (s->NY + 1) * (double)s->gr.nt * s->ns;
  }
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
  b->last_ms = ms;
  b->last_cells = cells;
  return GJW_OK;
}

// Objective of each survey's current model (w_objs_e.m:22-63 with one L2
// objective): forward without checkpoints for all surveys in shared launches,
// the survey's band filter on the synthetic and the observed traces, then
// J = 0.5 |d - d_o|^2 / |d_o|^2 in FP64.
// Forward sweep of every survey of the batch in shared launches, without
// checkpoints or wavefield (the line-search forwards w_fwd_(..., 0) of
// w_objs_e.m / w_pica_e_.m): traces in each survey's s->tr; d_obs[k] (may be
// NULL) copied to s->obs.  Returns the cell-updates in *cells.
static int batch_forward_only(gjw_batch* b, const float* const* d_obs, const char* who,
                              double* cells) {
  const int S = (int)b->sv.size();
  HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  int T = 0, fuse = 8;
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    TRY(check_ready(s, who));
    if (d_obs && !d_obs[k]) return fail(GJW_ERR_ARG, "%s: d_obs[%d] is NULL", who, k);
    if (s->path != GJW_PATH_BATCHED)
      return fail(GJW_ERR_ARG, "%s: survey %d is on the plain path", who, k);
    T = std::max(T, s->gr.nt);
  }
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    const long long n = (long long)s->ntr * s->gr.nt;
    TRY(ensure_fwd_state(s, s->ns));
    TRY(s->tr.ensure(std::max<long long>(n, 1)));
    TRY(s->obs.ensure(std::max<long long>(n, 1)));
    TRY(s->adj.ensure(std::max<long long>(n, 1)));
    TRY(s->part.ensure(2048));
    TRY(s->Jd.ensure(2));
    TRY(zero_fwd_state(s, s->ns));
    TRY(autotune_L(s, s->ns));
    fuse = std::min(fuse, batch_fwd_depth(s));
    if (d_obs) TRY(host_or_dev_copy(s->obs.p, d_obs[k], n * 4, st));
  }
  std::vector<DevSurvey> hs(S);
  for (int k = 0; k < S; ++k) hs[k] = dev_survey(b->sv[k]);
  TRY(b->dsv.ensure(S));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(b->dsv.p, hs.data(), S * sizeof(DevSurvey), hipMemcpyHostToDevice));
  for (int k = 0; k < S; ++k) b->tab.sv[k] = hs[k];
  b->T = T;
  b->K = T;  // one segment: no checkpoints
  b->fuse = fuse;
  b->keep = false;
  TRY(batch_tune(b));
  HIPCHK(hipEventRecord(b->ctx->ev0, st));
  const int rc = batch_fwd_enqueue(b);
  b->keep = true;
  TRY(rc);
  *cells = 0;
  for (int k = 0; k < S; ++k) {
    gjw_survey* s = b->sv[k];
    *cells += (double)(s->NX + 1) * (s->NY + 1) * (double)s->gr.nt * s->ns;
    s->kept = 0;
  }
  return GJW_OK;
}

static int batch_finish_timing(gjw_batch* b, double cells) {
  hipStream_t st = b->ctx->stream;
  HIPCHK(hipEventRecord(b->ctx->ev1, st));
  HIPCHK(hipEventSynchronize(b->ctx->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
  b->last_ms = ms;
  b->last_cells = cells;
  return GJW_OK;
}

// 0.5 sum (T(d) - T(o))^2 and sum o^2 over n samples into Jd[0], Jd[1]
// (T: w_E.m's transform of kind e_kind)
static int obj_terms(gjw_survey* s, const float* d, const float* o, long long n, int e_kind,
                     double out[2]) {
  hipStream_t st = s->ctx->stream;
  const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(n, 256), 1), 1024);
  hipLaunchKernelGGL(k_obj_terms, dim3(blocks), dim3(256), 0, st, d, o, n, e_kind, s->part.p,
                     s->part.p + 1024);
  hipLaunchKernelGGL(k_reduce_d, dim3(1), dim3(256), 0, st, s->part.p, blocks, 1.0, s->Jd.p);
  hipLaunchKernelGGL(k_reduce_d, dim3(1), dim3(256), 0, st, s->part.p + 1024, blocks, 1.0,
                     s->Jd.p + 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->Jd.p, 16, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return GJW_OK;
}

// w_bozdag.m's second and third outputs of the survey's working traces:
// envelopes (env_*) or instantaneous phases (phase_L2) of s->tr into s->fP
// and of s->obs into s->fov
static int instants(gjw_survey* s, int kind) {
  hipStream_t st = s->ctx->stream;
  const int nt = s->gr.nt;
  const long long n = (long long)s->ntr * nt;
  TRY(s->fin.ensure(n));
  TRY(s->fov.ensure(n));
  TRY(s->fP.ensure(n));
  TRY(s->fana.ensure(n));
  TRY(s->fhalf.ensure((size_t)s->ntr * (nt / 2 + 1)));
  const int blocks = (int)std::min<long long>(std::max<long long>(cdiv(n, 256), 1), 1024);
  float* dst[2] = {s->fov.p, s->fP.p};
  const float* src[2] = {s->obs.p, s->tr.p};
  for (int w = 0; w < 2; ++w) {
    HIPCHK(hipMemcpyAsync(s->fin.p, src[w], n * 4, hipMemcpyDeviceToDevice, st));
    TRY(hilbert_dev(s));
    hipLaunchKernelGGL(k_obs_instant, dim3(blocks), dim3(256), 0, st, s->fana.p, n, nt, kind,
                       dst[w]);
    HIPCHK(hipGetLastError());
  }
  return GJW_OK;
}

extern "C" int gjw_batch_objective_multi(gjw_batch* b, const float* const* d_obs, int32_t nobj,
                                         const int32_t* kinds, const double* weights,
                                         int32_t e_kind, double fstp_lo, double fstp_hi,
                                         double* J) {
  const char* who = "gjw_batch_objective_multi";
  if (!b || !d_obs || !J) return fail(GJW_ERR_ARG, "%s: NULL argument", who);
  if (nobj < 0 || (nobj > 0 && (!kinds || !weights)))
    return fail(GJW_ERR_ARG, "%s: nobj kinds and weights expected", who);
  if (e_kind < GJW_MISFIT_L2 || e_kind > GJW_MISFIT_ENV_LOG)
    return fail(GJW_ERR_ARG, "%s: unknown e_kind %d", who, e_kind);
  for (int q = 0; q < nobj; ++q) {
    if (kinds[q] < GJW_MISFIT_L2 || kinds[q] > GJW_MISFIT_FREQSTP)
      return fail(GJW_ERR_ARG, "%s: unknown objective kind %d", who, kinds[q]);
    if (kinds[q] == GJW_MISFIT_FREQSTP && !(fstp_hi > fstp_lo))
      return fail(GJW_ERR_ARG, "%s: FREQstp needs fstp_hi > fstp_lo", who);
  }
  double cells = 0;
  TRY(batch_forward_only(b, d_obs, who, &cells));
  for (size_t k = 0; k < b->sv.size(); ++k) {
    gjw_survey* s = b->sv[k];
    const long long n = (long long)s->ntr * s->gr.nt;
    // w_objs_e.m:34-37: the band filter of both trace sets
    TRY(trace_filter(s, s->tr.p, s->ntr));
    TRY(trace_filter(s, s->obs.p, s->ntr));
    double t[2];
    if (nobj == 0) {  // w_objs_e.m:61-62: one objective, w_E of the traces
      TRY(obj_terms(s, s->tr.p, s->obs.p, n, e_kind, t));
      J[k] = t[1] > 0 ? t[0] / t[1] : 0.0;
      continue;
    }
    // w_objs_e.m:41-60: weighted objectives; FREQstp filters both trace sets
    // again (and keeps them filtered for the objectives after it)
    double acc = 0;
    for (int q = 0; q < nobj; ++q) {
      const int kind = kinds[q];
      if (kind == GJW_MISFIT_FREQSTP) {
        if (!s->flt) s->flt = new GaussFilter();
        TRY(s->flt->run(s->ctx->stream, s->tr.p, (int)s->ntr, s->gr.nt, s->gr.dt, fstp_lo,
                        fstp_hi, 8));
        TRY(s->flt->run(s->ctx->stream, s->obs.p, (int)s->ntr, s->gr.nt, s->gr.dt, fstp_lo,
                        fstp_hi, 8));
      }
      if (kind == GJW_MISFIT_L2 || kind == GJW_MISFIT_FREQSTP) {
        TRY(obj_terms(s, s->tr.p, s->obs.p, n, e_kind, t));
      } else {
        TRY(instants(s, kind));
        TRY(obj_terms(s, s->fP.p, s->fov.p, n, e_kind, t));
      }
      acc += weights[q] * (t[1] > 0 ? t[0] / t[1] : 0.0);
    }
    J[k] = acc;
  }
  return batch_finish_timing(b, cells);
}

extern "C" int gjw_batch_objective(gjw_batch* b, const float* const* d_obs, double* J) {
  return gjw_batch_objective_multi(b, d_obs, 0, nullptr, nullptr, GJW_MISFIT_L2, 0, 0, J);
}

extern "C" int gjw_batch_forward(gjw_batch* b, float* const* traces) {
  const char* who = "gjw_batch_forward";
  if (!b || !traces) return fail(GJW_ERR_ARG, "%s: NULL argument", who);
  for (size_t k = 0; k < b->sv.size(); ++k)
    if (!traces[k]) return fail(GJW_ERR_ARG, "%s: traces[%d] is NULL", who, (int)k);
  double cells = 0;
  TRY(batch_forward_only(b, nullptr, who, &cells));
  TRY(batch_finish_timing(b, cells));
  hipStream_t st = b->ctx->stream;
  for (size_t k = 0; k < b->sv.size(); ++k) {
    const gjw_survey* s = b->sv[k];
    HIPCHK(hipMemcpyAsync(traces[k], s->tr.p, (size_t)s->ntr * s->gr.nt * 4, hipMemcpyDefault, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return GJW_OK;
}

extern "C" int gjw_batch_last_timing(gjw_batch* b, double* ms, double* cells) {
  if (!b) return fail(GJW_ERR_ARG, "gjw_batch_last_timing: NULL batch");
  if (ms) *ms = b->last_ms;
  if (cells) *cells = b->last_cells;
  return GJW_OK;
}

extern "C" int gjw_kernel_stats(gjw_survey* s, int32_t which, double* avg_ms, int64_t* count,
                                int64_t* launches, int32_t reset) {
  if (!s || which < 0 || which >= kKernKinds)
    return fail(GJW_ERR_ARG, "gjw_kernel_stats: bad argument");
  TRY(kev_resolve(s));
  if (avg_ms) *avg_ms = s->kt_n[which] ? s->kt_ms[which] / s->kt_n[which] : 0.0;
  if (count) *count = s->kt_n[which];
  if (launches) *launches = s->kl_n[which];
  if (reset) {
    s->kt_ms[which] = 0;
    s->kt_n[which] = 0;
    s->kl_n[which] = 0;
  }
  return GJW_OK;
}

extern "C" int gjw_tile(gjw_survey* s, int32_t which, int32_t* L) {
  if (!s || !L || which < 0 || which >= kKernKinds) return fail(GJW_ERR_ARG, "gjw_tile: bad argument");
  const int t[kKernKinds] = {s->tuned_L1, s->tuned_L1, s->tuned_L,  s->tuned_La2,
                             s->tuned_L4, s->tuned_La4, s->tuned_L8, s->tuned_La8,
                             s->tuned_La6, s->tuned_L8};
  *L = t[which];
  return GJW_OK;
}

extern "C" int gjw_query(gjw_survey* s, int32_t what, int64_t* value) {
  if (!s || !value) return fail(GJW_ERR_ARG, "gjw_query: NULL argument");
  switch (what) {
    case GJW_Q_CKPT_INTERVAL:
      if (!s->have_model || !s->have_shots)
        return fail(GJW_ERR_STATE, "gjw_query: set the model and the shots first");
      *value = s->path == GJW_PATH_PLAIN ? 0 : ck_interval(s);
      return GJW_OK;
    case GJW_Q_FWD_DEPTH:
      *value = s->path == GJW_PATH_PLAIN ? 1 : s->fuse_eff;
      return GJW_OK;
    case GJW_Q_ADJ_DEPTH:
      *value = s->path == GJW_PATH_PLAIN ? 1 : s->fuse_eff > 0 ? adj_depth(s) : 0;
      return GJW_OK;
    case GJW_Q_FRAMES:
      *value = s->fr16_used ? GJW_FRAMES_BF16 : GJW_FRAMES_F32;
      return GJW_OK;
    case GJW_Q_ROW_PAIRS:
      *value = s->pairs_last;
      return GJW_OK;
    default:
      return fail(GJW_ERR_ARG, "gjw_query: unknown query %d", what);
  }
}

extern "C" int gjw_last_timing(gjw_survey* s, double* ms, double* cells) {
  if (!s) return fail(GJW_ERR_ARG, "gjw_last_timing: NULL survey");
  if (ms) *ms = s->last_ms;
  if (cells) *cells = s->last_cells;
  return GJW_OK;
}

extern "C" int gjw_set_adj_filter(gjw_survey* s, double f_low, double f_high, int32_t steep) {
  if (!s) return fail(GJW_ERR_ARG, "gjw_set_adj_filter: NULL survey");
  if (f_high > f_low && (steep < 1 || steep > 64))
    return fail(GJW_ERR_ARG, "gjw_set_adj_filter: steep must be in 1..64");
  s->flt_lo = f_low;
  s->flt_hi = f_high;
  s->flt_steep = steep;
  return GJW_OK;
}

// device copy of a host or device array for an in-place operation: *dp is v
// when v is device memory, else a fresh device copy (freed by stage_end)
static int stage_begin(gjw_ctx* c, void* v, size_t bytes, void** dp, bool* own) {
  hipPointerAttribute_t at;
  const hipError_t pe = hipPointerGetAttributes(&at, v);
  const bool on_dev = pe == hipSuccess && at.type == hipMemoryTypeDevice;
  (void)hipGetLastError();
  *own = !on_dev;
  *dp = v;
  if (on_dev) return GJW_OK;
  HIPCHK(hipMalloc(dp, std::max<size_t>(bytes, 1)));
  HIPCHK(hipMemcpyAsync(*dp, v, bytes, hipMemcpyDefault, c->stream));
  return GJW_OK;
}
static int stage_end(gjw_ctx* c, void* v, void* dp, size_t bytes, bool own, bool back) {
  if (own && back) HIPCHK(hipMemcpyAsync(v, dp, bytes, hipMemcpyDefault, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (own) HIPCHK(hipFree(dp));
  return GJW_OK;
}

extern "C" int gjw_filt_gauss(gjw_ctx* c, float* d, int32_t ntr, int32_t nt, double dt,
                              double f_low, double f_high, int32_t steep) {
  if (!c || (!d && ntr > 0)) return fail(GJW_ERR_ARG, "gjw_filt_gauss: NULL argument");
  if (ntr < 0 || nt < 2 || !(dt > 0) || !(f_high > f_low) || steep < 1 || steep > 64)
    return fail(GJW_ERR_ARG, "gjw_filt_gauss: bad arguments");
  if (ntr == 0) return GJW_OK;
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)ntr * nt * 4;
  void* dp;
  bool own;
  TRY(stage_begin(c, d, bytes, &dp, &own));
  if (!c->flt) c->flt = new GaussFilter();
  const int rc = c->flt->run(c->stream, static_cast<float*>(dp), ntr, nt, dt, f_low, f_high, steep);
  TRY(stage_end(c, d, dp, bytes, own, rc == GJW_OK));
  return rc;
}

extern "C" int gjw_lmute(gjw_ctx* c, float* d, float* d_o, float* e, int32_t ntr, int32_t nt,
                         const double* t, const double* rx, double v_mute, double t_mute) {
  if (!c || !d || !d_o || !t || !rx) return fail(GJW_ERR_ARG, "gjw_lmute: NULL argument");
  if (ntr < 1 || nt < 1 || !(v_mute != 0)) return fail(GJW_ERR_ARG, "gjw_lmute: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)ntr * nt * 4;
  void *dd, *dob, *de = nullptr, *dt_, *drx;
  bool od, oo, oe = false, ot, orx;
  TRY(stage_begin(c, d, bytes, &dd, &od));
  TRY(stage_begin(c, d_o, bytes, &dob, &oo));
  if (e) TRY(stage_begin(c, e, bytes, &de, &oe));
  TRY(stage_begin(c, const_cast<double*>(t), (size_t)nt * 8, &dt_, &ot));
  TRY(stage_begin(c, const_cast<double*>(rx), (size_t)ntr * 8, &drx, &orx));
  const long long n = (long long)ntr * nt;
  const int blocks = (int)std::min<long long>(cdiv(n, 256), 8192);
  // m/ns -> m/s, ns -> s (w_lmute.m:22-23)
  hipLaunchKernelGGL(k_lmute, dim3(blocks), dim3(256), 0, c->stream, static_cast<float*>(dd),
                     static_cast<float*>(dob), static_cast<float*>(de), nt, (long long)ntr,
                     static_cast<const double*>(dt_), static_cast<const double*>(drx),
                     v_mute * 1e9, t_mute * 1e-9);
  const bool ok = hipGetLastError() == hipSuccess;
  TRY(stage_end(c, d, dd, bytes, od, ok));
  TRY(stage_end(c, d_o, dob, bytes, oo, ok));
  if (e) TRY(stage_end(c, e, de, bytes, oe, ok));
  TRY(stage_end(c, const_cast<double*>(t), dt_, (size_t)nt * 8, ot, false));
  TRY(stage_end(c, const_cast<double*>(rx), drx, (size_t)ntr * 8, orx, false));
  if (!ok) return fail(GJW_ERR_HIP, "gjw_lmute: launch failed");
  return GJW_OK;
}

extern "C" int gjw_differentiate_cube(gjw_ctx* c, float* v, int32_t nz, int32_t nx, int32_t nt,
                                      double dt) {
  if (!c || !v) return fail(GJW_ERR_ARG, "gjw_differentiate_cube: NULL argument");
  if (nz < 1 || nx < 1 || nt < 3) return fail(GJW_ERR_ARG, "gjw_differentiate_cube: bad sizes");
  HIPCHK(hipSetDevice(c->device));
  hipPointerAttribute_t at;
  const long long fr = (long long)nz * nx, n = fr * nt;
  hipError_t pe = hipPointerGetAttributes(&at, v);
  const bool on_dev = (pe == hipSuccess && at.type == hipMemoryTypeDevice);
  (void)hipGetLastError();
  float* d = v;
  if (!on_dev) HIPCHK(hipMalloc(&d, n * 4));
  if (!on_dev) HIPCHK(hipMemcpyAsync(d, v, n * 4, hipMemcpyDefault, c->stream));
  hipLaunchKernelGGL(k_dcube, dim3((unsigned)cdiv(fr, 256)), dim3(256), 0, c->stream, d, fr, nt,
                     dt);
  HIPCHK(hipGetLastError());
  if (!on_dev) {
    HIPCHK(hipMemcpyAsync(v, d, n * 4, hipMemcpyDefault, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(d));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return GJW_OK;
}

// --------------------------------------------------------------------------
// gradient post-processing (w_update_e.m:108-145, w_update_e_.m:28-42,
// w_image_e.m:50-51) on batches of FP64 m x n images
// --------------------------------------------------------------------------
static PostWork* post_work(gjw_ctx* c) {
  if (!c->post) c->post = new PostWork();
  return c->post;
}

static dim3 grid_for(long long tot) {
  return dim3((unsigned)std::min<long long>(std::max<long long>(cdiv(tot, 256), 1), 8192));
}

extern "C" int gjw_gaussi(gjw_ctx* c, double* g, int32_t m, int32_t n, int32_t count,
                          const double* X, const double* Y, const int32_t* s, const double* lo) {
  if (!c || (count > 0 && (!g || !X || !Y || !s || !lo)))
    return fail(GJW_ERR_ARG, "gjw_gaussi: NULL argument");
  if (m < 1 || n < 1 || count < 0) return fail(GJW_ERR_ARG, "gjw_gaussi: bad sizes");
  if (count == 0) return GJW_OK;
  for (int k = 0; k < count; ++k) {
    if (s[2 * k] < 1 || s[2 * k] > m || s[2 * k + 1] < 1 || s[2 * k + 1] > n)
      return fail(GJW_ERR_ARG, "gjw_gaussi: source %d at (%d, %d) outside the %d x %d model", k,
                  s[2 * k], s[2 * k + 1], m, n);
    if (!(lo[k] > 0) || !std::isfinite(lo[k]))
      return fail(GJW_ERR_ARG, "gjw_gaussi: lo[%d] must be positive and finite", k);
  }
  HIPCHK(hipSetDevice(c->device));
  PostWork* w = post_work(c);
  // small host arrays -> device: X (n), Y (m), lo (count), gmax (count); s
  std::vector<double> hv((size_t)n + m + count);
  std::copy(X, X + n, hv.begin());
  std::copy(Y, Y + m, hv.begin() + n);
  std::copy(lo, lo + count, hv.begin() + n + m);
  TRY(w->dsmall.ensure(hv.size() + count));
  TRY(w->ismall.ensure((size_t)2 * count));
  HIPCHK(hipMemcpy(w->dsmall.p, hv.data(), hv.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(w->ismall.p, s, (size_t)2 * count * 4, hipMemcpyHostToDevice));
  const double *dX = w->dsmall.p, *dY = dX + n, *dlo = dY + m;
  double* gmax = w->dsmall.p + hv.size();
  const size_t bytes = (size_t)count * m * n * 8;
  void* dp;
  bool own;
  TRY(stage_begin(c, g, bytes, &dp, &own));
  hipLaunchKernelGGL(k_gaussi_norm, dim3(count), dim3(256), 0, c->stream, dX, n, dY, m,
                     w->ismall.p, dlo, gmax);
  bool ok = hipGetLastError() == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_gaussi_apply, grid_for((long long)count * m * n), dim3(256), 0, c->stream,
                       static_cast<double*>(dp), m, n, (long long)count, dX, dY, w->ismall.p, dlo,
                       gmax);
    ok = hipGetLastError() == hipSuccess;
  }
  TRY(stage_end(c, g, dp, bytes, own, ok));
  if (!ok) return fail(GJW_ERR_HIP, "gjw_gaussi: launch failed");
  return GJW_OK;
}

extern "C" int gjw_image_gaussian(gjw_ctx* c, const double* in, double* out, int32_t m,
                                  int32_t n, int32_t count, double ax, double az, int32_t type) {
  if (!c || (count > 0 && (!in || !out)))
    return fail(GJW_ERR_ARG, "gjw_image_gaussian: NULL argument");
  if (m < 2 || n < 2 || count < 0)
    return fail(GJW_ERR_ARG, "gjw_image_gaussian: images must be at least 2 x 2");
  if (type != GJW_LOW_PASS && type != GJW_HI_PASS)
    return fail(GJW_ERR_ARG, "gjw_image_gaussian: type must be GJW_LOW_PASS or GJW_HI_PASS");
  if (!(ax > 0) || !(az > 0) || !std::isfinite(ax) || !std::isfinite(az))
    return fail(GJW_ERR_ARG, "gjw_image_gaussian: ax and az must be positive and finite");
  if (count == 0) return GJW_OK;
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)count * m * n * 8;
  void *di, *dout;
  bool oi, oo;
  TRY(stage_begin(c, const_cast<double*>(in), bytes, &di, &oi));
  if (out == in) {
    dout = di;
    oo = false;
  } else {
    TRY(stage_begin(c, out, bytes, &dout, &oo));
  }
  const int rc = post_work(c)->kfilter(c->stream, static_cast<const double*>(di),
                                       static_cast<double*>(dout), m, n, count, ax, az, type);
  if (out == in) {
    TRY(stage_end(c, out, di, bytes, oi, rc == GJW_OK));
  } else {
    TRY(stage_end(c, const_cast<double*>(in), di, bytes, oi, false));
    TRY(stage_end(c, out, dout, bytes, oo, rc == GJW_OK));
  }
  return rc;
}

extern "C" int gjw_normali(gjw_ctx* c, double* a, int32_t m, int32_t n, int32_t count,
                           int32_t nan_guard, const double* shape) {
  if (!c || (count > 0 && !a)) return fail(GJW_ERR_ARG, "gjw_normali: NULL argument");
  if (m < 1 || n < 1 || count < 0) return fail(GJW_ERR_ARG, "gjw_normali: bad sizes");
  if (count == 0) return GJW_OK;
  if (count > 65535) return fail(GJW_ERR_ARG, "gjw_normali: at most 65535 images per call");
  HIPCHK(hipSetDevice(c->device));
  PostWork* w = post_work(c);
  const long long per = (long long)m * n;
  const int P = (int)std::min<long long>(256, std::max<long long>(1, cdiv(per, 256 * 16)));
  TRY(w->dsmall.ensure((size_t)count * P));
  TRY(w->ismall.ensure((size_t)count * P));
  const size_t bytes = (size_t)per * count * 8;
  void *dp, *dsh = nullptr;
  bool own, osh = false;
  TRY(stage_begin(c, a, bytes, &dp, &own));
  if (shape) TRY(stage_begin(c, const_cast<double*>(shape), (size_t)per * 8, &dsh, &osh));
  hipLaunchKernelGGL(k_absmax_part, dim3(P, count), dim3(256), 0, c->stream,
                     static_cast<const double*>(dp), per, w->dsmall.p, w->ismall.p);
  bool ok = hipGetLastError() == hipSuccess;
  if (ok) {
    const unsigned bx = (unsigned)std::min<long long>(cdiv(per, 256), 1024);
    hipLaunchKernelGGL(k_normali, dim3(bx, count), dim3(256), 0, c->stream,
                       static_cast<double*>(dp), per, P, w->dsmall.p, w->ismall.p, nan_guard ? 1 : 0,
                       static_cast<const double*>(dsh));
    ok = hipGetLastError() == hipSuccess;
  }
  if (shape) TRY(stage_end(c, const_cast<double*>(shape), dsh, (size_t)per * 8, osh, false));
  TRY(stage_end(c, a, dp, bytes, own, ok));
  if (!ok) return fail(GJW_ERR_HIP, "gjw_normali: launch failed");
  return GJW_OK;
}

// gjw_normali for device-resident images, allocating nothing: the partials
// go to this device's scratch array (every slot k_normali reads is written by
// k_absmax_part of the same call, so nothing needs clearing first).
extern "C" int gjw_normali_noalloc(gjw_ctx* c, double* a, int32_t m, int32_t n, int32_t count,
                                   int32_t nan_guard, const double* shape) {
  if (!c || (count > 0 && !a)) return fail(GJW_ERR_ARG, "gjw_normali_noalloc: NULL argument");
  if (m < 1 || n < 1 || count < 0) return fail(GJW_ERR_ARG, "gjw_normali_noalloc: bad sizes");
  if (count == 0) return GJW_OK;
  const long long per = (long long)m * n;
  const int P = (int)std::min<long long>(256, std::max<long long>(1, cdiv(per, 256 * 16)));
  if ((long long)count * P > kNoAllocParts || count > 65535)
    return fail(GJW_ERR_ARG,
                "gjw_normali_noalloc: %d images of %lld cells need %lld partials; the scratch "
                "holds %d (at most %d images of 1M cells or more)",
                count, per, (long long)count * P, kNoAllocParts, kNoAllocParts / 256);
  auto on_device = [](const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return at.type == hipMemoryTypeDevice;
  };
  if (!on_device(a) || (shape && !on_device(shape)))
    return fail(GJW_ERR_ARG, "gjw_normali_noalloc: the images and the mask must be device memory");
  HIPCHK(hipSetDevice(c->device));
  TRY(noalloc_scratch_init(c->device));
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[c->device]);
  double* pmax = reinterpret_cast<double*>(base);
  int* pnan = reinterpret_cast<int*>(base + (size_t)kNoAllocParts * 8);
  hipLaunchKernelGGL(k_absmax_part, dim3(P, count), dim3(256), 0, c->stream, a, per, pmax, pnan);
  HIPCHK(hipGetLastError());
  const unsigned bx = (unsigned)std::min<long long>(cdiv(per, 256), 1024);
  hipLaunchKernelGGL(k_normali, dim3(bx, count), dim3(256), 0, c->stream, a, per, P, pmax, pnan,
                     nan_guard ? 1 : 0, shape);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GJW_OK;
}

extern "C" int gjw_stack_update(gjw_ctx* c, const double* g, int32_t m, int32_t n, int32_t count,
                                const double* steps, int32_t neg_rule, double* dparam,
                                double* model) {
  if (!c || !g || !steps || !dparam) return fail(GJW_ERR_ARG, "gjw_stack_update: NULL argument");
  if (m < 1 || n < 1 || count < 1) return fail(GJW_ERR_ARG, "gjw_stack_update: bad sizes");
  HIPCHK(hipSetDevice(c->device));
  PostWork* w = post_work(c);
  // w_update_e_.m:40-41: the update is dropped when the steps sum below zero
  // (the conductivity and joint reductions have no such rule)
  double ssum = 0.0;
  for (int k = 0; k < count; ++k) ssum += steps[k];
  const int zero = (neg_rule && ssum < 0) ? 1 : 0;
  TRY(w->dsmall.ensure((size_t)count));
  HIPCHK(hipMemcpy(w->dsmall.p, steps, (size_t)count * 8, hipMemcpyHostToDevice));
  const long long per = (long long)m * n;
  void *dg, *dd, *dm = nullptr;
  bool og, od, om = false;
  TRY(stage_begin(c, const_cast<double*>(g), (size_t)per * count * 8, &dg, &og));
  TRY(stage_begin(c, dparam, (size_t)per * 8, &dd, &od));
  if (model) TRY(stage_begin(c, model, (size_t)per * 8, &dm, &om));
  hipLaunchKernelGGL(k_stack_update, grid_for(per), dim3(256), 0, c->stream,
                     static_cast<const double*>(dg), per, count, w->dsmall.p, zero,
                     static_cast<double*>(dd), static_cast<double*>(dm));
  const bool ok = hipGetLastError() == hipSuccess;
  TRY(stage_end(c, const_cast<double*>(g), dg, (size_t)per * count * 8, og, false));
  TRY(stage_end(c, dparam, dd, (size_t)per * 8, od, ok));
  if (model) TRY(stage_end(c, model, dm, (size_t)per * 8, om, ok));
  if (!ok) return fail(GJW_ERR_HIP, "gjw_stack_update: launch failed");
  return GJW_OK;
}
