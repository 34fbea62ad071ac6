// gjw_kernels.hip -- HIP kernels (gfx950) for the radar TMz Yee FDTD path.
//
// Hot kernels:
//   k_fwd_step  one fused leapfrog step (H then Ez, split-field PML, source,
//               receiver sampling, optional domain-frame store) for every shot
//               of a survey in one launch.  Restates w_solve.m:85-104.
//   k_adj_step  one step of the exact discrete adjoint of k_fwd_step, fused
//               with the permittivity / conductivity correlation.
// The reference solver runs the same step as four MATLAB slice statements
// (new_H, new_H_pml, new_Ez, new_Ez_pml; w_solve.m:136-268).
//
// Both hot kernels march a 64-lane wavefront down one column strip: lane l
// owns depth row i0+l, the loop runs over columns j.  Vertical neighbours come
// from the adjacent lane (cross-lane shuffle), horizontal neighbours from the
// previous / next loop iteration, so every field value is read from HBM once
// per step (plus one halo row per 64 and one halo column per strip).
// Ping-pong buffers (E0,H0 -> E1,H1) make the single fused pass race-free.
#include "gjw_internal.h"

namespace gjw {

__device__ __forceinline__ float shfl_dn1(float v) { return __shfl_down(v, 1, 64); }
__device__ __forceinline__ float shfl_up1(float v) { return __shfl_up(v, 1, 64); }

// --------------------------------------------------------------------------
// per-cell coefficient selection (w_solve.m:144-147 inner, :170-186 strips)
// --------------------------------------------------------------------------
__device__ __forceinline__ void hy_coef(const DevGrid& g, int i, int j, float& a, float& b) {
  if (i >= g.P && i <= g.NX - g.P - 1) {
    a = 1.f;
    b = g.by0;
  } else if (i <= g.NX - 1 && j >= 1 && j <= g.NY - 1) {
    const long long k = (long long)j * g.LD + i;
    a = g.hay[k];
    b = g.hby[k];
  } else {
    a = 1.f;
    b = 0.f;
  }
}

__device__ __forceinline__ void hx_coef(const DevGrid& g, int i, int j, float& a, float& b) {
  if (j >= g.P && j <= g.NY - g.P - 1) {
    a = 1.f;
    b = g.bx0;
  } else if (j <= g.NY - 1 && i >= 1 && i <= g.NX - 1) {
    const long long k = (long long)j * g.LD + i;
    a = g.hax[k];
    b = g.hbx[k];
  } else {
    a = 1.f;
    b = 0.f;
  }
}

__device__ __forceinline__ float hy_new(const DevGrid& g, int i, int j, float hy, float e_i,
                                        float e_ip1) {
  float a, b;
  hy_coef(g, i, j, a, b);
  return a * hy + b * (e_ip1 - e_i);
}

__device__ __forceinline__ float hx_new(const DevGrid& g, int i, int j, float hx, float e_j,
                                        float e_jp1) {
  float a, b;
  hx_coef(g, i, j, a, b);
  return a * hx + b * (e_jp1 - e_j);
}

// ==========================================================================
// forward: fused step
// ==========================================================================
__global__ void __launch_bounds__(256)
k_fwd_step(DevGrid g, DevShots s, int L, const float* __restrict__ E0,
           const float* __restrict__ Hx0, const float* __restrict__ Hy0, float* __restrict__ E1,
           float* __restrict__ Hx1, float* __restrict__ Hy1, float* __restrict__ X,
           float* __restrict__ Y, int it, float* __restrict__ tr, float* __restrict__ frame,
           long long fstride) {
  const int lane = threadIdx.x;
  const int shot = blockIdx.z;
  const int i0 = blockIdx.x * 64;
  const int jb = (blockIdx.y * blockDim.y + threadIdx.y) * L;
  const int NX = g.NX, NY = g.NY, P = g.P;
  const long long LD = g.LD;
  if (jb > NY) return;
  const int je = min(jb + L, NY + 1);
  const int i = i0 + lane;
  const bool ok = i <= NX;
  const long long so = (long long)shot * g.FS;
  E0 += so; Hx0 += so; Hy0 += so; E1 += so; Hx1 += so; Hy1 += so; X += so; Y += so;
  const int si = s.si[shot], sj = s.sj[shot];
  const float wsrc = s.sc[shot] * s.wav[(long long)shot * s.nt + it];
  const int* cp = s.cptr + (long long)shot * (NY + 2);
  float* fr = frame ? frame + (long long)shot * fstride : nullptr;
  const bool fastrows = (i0 >= P + 1) && (i0 + 63 <= NX - P - 1);

  float e_c = ok ? E0[(long long)jb * LD + i] : 0.f;
  float hx_prev = 0.f;  // Hx'(i, j-1)
  if (jb >= 1 && ok) {
    const long long km = (long long)(jb - 1) * LD + i;
    hx_prev = hx_new(g, i, jb - 1, Hx0[km], E0[km], e_c);
  }
  for (int j = jb; j < je; ++j) {
    const long long k = (long long)j * LD + i;
    const float e_n = (ok && j < NY) ? E0[k + LD] : 0.f;
    const float hx = (ok && j < NY) ? Hx0[k] : 0.f;
    const float hy = (ok && i < NX) ? Hy0[k] : 0.f;
    float e_dn = shfl_dn1(e_c);
    if (lane == 63) e_dn = (i + 1 <= NX) ? E0[k + 1] : 0.f;
    float hyn, hxn, en;
    const bool fast = fastrows && j >= P + 1 && j <= NY - P - 1;
    if (fast) {
      // interior: mu_r = 1, sigma_m = 0 => Chxh = Chyh = 1 (w_coeff.m:34-40)
      hyn = hy + g.by0 * (e_dn - e_c);
      hxn = hx + g.bx0 * (e_n - e_c);
      float hy_up = shfl_up1(hyn);
      if (lane == 0) hy_up = Hy0[k - 1] + g.by0 * (e_c - E0[k - 1]);
      en = g.ce[k] * e_c + g.ch[k] * ((hyn - hy_up) - (hxn - hx_prev));
    } else {
      hyn = (ok && i < NX) ? hy_new(g, i, j, hy, e_c, e_dn) : 0.f;
      hxn = (ok && j < NY) ? hx_new(g, i, j, hx, e_c, e_n) : 0.f;
      float hy_up = shfl_up1(hyn);
      if (lane == 0) hy_up = (ok && i >= 1) ? hy_new(g, i - 1, j, Hy0[k - 1], E0[k - 1], e_c) : 0.f;
      en = 0.f;
      if (ok) {
        const int kd = ekind(NX, NY, P, i, j);
        const float dhy = hyn - hy_up, dhx = hxn - hx_prev;
        if (kd == KIND_INNER) {
          en = g.ce[k] * e_c + g.ch[k] * (dhy - dhx);
        } else if (kd == KIND_PML) {
          // split field Ez = Ezx + Ezy (w_solve.m:232-267)
          const float x = g.cxe[k] * X[k] + g.cxh[k] * dhy;
          const float y = g.cye[k] * Y[k] + g.cyh[k] * dhx;
          X[k] = x;
          Y[k] = y;
          en = x + y;
        }
      }
    }
    // soft source (w_solve.m:92); sc is 0 when the strip overwrites the cell
    if (j == sj && i == si) en += wsrc;
    if (ok) {
      E1[k] = en;
      if (j < NY) Hx1[k] = hxn;
      if (i < NX) Hy1[k] = hyn;
      if (tr) {  // receivers of this column (w_solve.m:100)
        const int qa = cp[j], qb = cp[j + 1];
        for (int q = qa; q < qb; ++q)
          if (s.crow[q] == i) tr[(long long)s.cid[q] * s.nt + it] = en;
      }
      if (fr) {
        const int a = i - g.di0, b = j - g.dj0;
        if (a >= 0 && a < g.m && b >= 0 && b < g.n) fr[(long long)b * g.m + a] = en;
      }
    }
    e_c = e_n;
    hx_prev = hxn;
  }
}

// ==========================================================================
// adjoint: one step of the transpose of k_fwd_step
// ==========================================================================
// Adjoint state per shot: M (adjoint of Ez), Nx/Ny (adjoint of the updated
// Hx/Hy, before their decay factor), X/Y (adjoint of Ezx/Ezy).  W is the
// curl weight an Ez cell hands to its H neighbours:
//   inner: Wy = Cezhy*M, Wx = -Cezhy*M;   PML: Wy = cxh*(X+M), Wx = cyh*(Y+M).
__device__ __forceinline__ void w_of(const DevGrid& g, int i, int j, long long k, float me,
                                     const float* __restrict__ X0,
                                     const float* __restrict__ Y0, float& wy, float& wx) {
  const int kd = ekind(g.NX, g.NY, g.P, i, j);
  if (kd == KIND_INNER) {
    const float c = g.ch[k];
    wy = c * me;
    wx = -c * me;
  } else if (kd == KIND_PML) {
    wy = g.cxh[k] * (X0[k] + me);
    wx = g.cyh[k] * (Y0[k] + me);
  } else {
    wy = 0.f;
    wx = 0.f;
  }
}

// it: forward step being transposed (0-based).  M0 holds dJ/dEz(S_{it+1})
// with the data residual of step it already added; M1 receives dJ/dEz(S_it)
// with the residual of step it-1 added.  F1/F0 are the stored forward frames
// Ez(S_{it+1}), Ez(S_it) on the domain (F0 may be null for S_0 = 0).
__global__ void __launch_bounds__(256)
k_adj_step(DevGrid g, DevShots s, int L, const float* __restrict__ M0,
           const float* __restrict__ Nx0, const float* __restrict__ Ny0,
           const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M1,
           float* __restrict__ Nx1, float* __restrict__ Ny1, float* __restrict__ X1,
           float* __restrict__ Y1, int it, const float* __restrict__ adj,
           const float* __restrict__ F1, const float* __restrict__ F0, long long fstride,
           float* __restrict__ acc, long long astride, float* __restrict__ gw) {
  const int lane = threadIdx.x;
  const int shot = blockIdx.z;
  const int i0 = blockIdx.x * 64;
  const int jb = (blockIdx.y * blockDim.y + threadIdx.y) * L;
  const int NX = g.NX, NY = g.NY, P = g.P;
  const long long LD = g.LD;
  if (jb > NY) return;
  const int je = min(jb + L, NY + 1);
  const int i = i0 + lane;
  const bool ok = i <= NX;
  const long long so = (long long)shot * g.FS;
  M0 += so; Nx0 += so; Ny0 += so; X0 += so; Y0 += so;
  M1 += so; Nx1 += so; Ny1 += so; X1 += so; Y1 += so;
  const int si = s.si[shot], sj = s.sj[shot];
  const float sc = s.sc[shot];
  const int* cp = s.cptr + (long long)shot * (NY + 2);
  const float* f1 = F1 ? F1 + (long long)shot * fstride : nullptr;
  const float* f0 = F0 ? F0 + (long long)shot * fstride : nullptr;
  float2* ac = acc ? reinterpret_cast<float2*>(acc) + (long long)shot * astride : nullptr;
  const bool fastrows = (i0 >= P + 2) && (i0 + 64 <= NX - P - 1);

  float me_c = 0.f, wy_c = 0.f, wx_c = 0.f;
  if (ok) {
    const long long k = (long long)jb * LD + i;
    me_c = M0[k];
    w_of(g, i, jb, k, me_c, X0, Y0, wy_c, wx_c);
  }
  float px_prev = 0.f;  // BX(i,j-1) * Nx'(i,j-1)
  if (jb >= 1 && ok) {
    const long long km = (long long)(jb - 1) * LD + i;
    float wy_m, wx_m, a, b;
    w_of(g, i, jb - 1, km, M0[km], X0, Y0, wy_m, wx_m);
    hx_coef(g, i, jb - 1, a, b);
    px_prev = b * (a * Nx0[km] + wx_m - wx_c);
  }
  for (int j = jb; j < je; ++j) {
    const long long k = (long long)j * LD + i;
    const bool fast = fastrows && j >= P + 2 && j <= NY - P - 2;
    float me_n = 0.f, wy_n = 0.f, wx_n = 0.f;
    float nyv = 0.f, nxv = 0.f, py, px, py_up, mn;
    if (fast) {
      // every cell touched here is inner: W = (c M, -c M), A = 1, B = (by0, bx0)
      me_n = M0[k + LD];
      const float cn = g.ch[k + LD];
      wy_n = cn * me_n;
      wx_n = -cn * me_n;
      float wy_dn = shfl_dn1(wy_c);
      if (lane == 63) wy_dn = g.ch[k + 1] * M0[k + 1];
      nyv = Ny0[k] + wy_c - wy_dn;
      nxv = Nx0[k] + wx_c - wx_n;
      py = g.by0 * nyv;
      px = g.bx0 * nxv;
      py_up = shfl_up1(py);
      if (lane == 0) py_up = g.by0 * (Ny0[k - 1] + g.ch[k - 1] * M0[k - 1] - wy_c);
      mn = g.ce[k] * me_c + py_up - py + px_prev - px;
    } else {
      if (ok && j < NY) {
        me_n = M0[k + LD];
        w_of(g, i, j + 1, k + LD, me_n, X0, Y0, wy_n, wx_n);
      }
      float wy_dn = shfl_dn1(wy_c);
      if (lane == 63) {
        wy_dn = 0.f;
        if (i + 1 <= NX) {
          float a_, b_;
          w_of(g, i + 1, j, k + 1, M0[k + 1], X0, Y0, a_, b_);
          wy_dn = a_;
        }
      }
      py = 0.f;
      px = 0.f;
      if (ok && i < NX) {
        float a, b;
        hy_coef(g, i, j, a, b);
        nyv = a * Ny0[k] + wy_c - wy_dn;
        py = b * nyv;
      }
      if (ok && j < NY) {
        float a, b;
        hx_coef(g, i, j, a, b);
        nxv = a * Nx0[k] + wx_c - wx_n;
        px = b * nxv;
      }
      py_up = shfl_up1(py);
      if (lane == 0) {
        py_up = 0.f;
        if (ok && i >= 1) {
          float wyu, wxu, a, b;
          w_of(g, i - 1, j, k - 1, M0[k - 1], X0, Y0, wyu, wxu);
          hy_coef(g, i - 1, j, a, b);
          py_up = b * (a * Ny0[k - 1] + wyu - wy_c);
        }
      }
      mn = 0.f;
      if (ok) {
        const int kd = ekind(NX, NY, P, i, j);
        const float base = (kd == KIND_INNER) ? g.ce[k] * me_c : (kd == KIND_PML ? 0.f : me_c);
        mn = base + py_up - py + px_prev - px;
        if (kd == KIND_PML) {
          X1[k] = g.cxe[k] * (X0[k] + me_c);
          Y1[k] = g.cye[k] * (Y0[k] + me_c);
        }
      }
    }
    if (ok) {
      if (adj && it >= 1) {  // residual of step it-1 enters dJ/dEz(S_it)
        const int qa = cp[j], qb = cp[j + 1];
        for (int q = qa; q < qb; ++q)
          if (s.crow[q] == i) mn += adj[(long long)s.cid[q] * s.nt + (it - 1)];
      }
      if (ac && ekind(NX, NY, P, i, j) == KIND_INNER) {
        const int a = i - g.di0, b = j - g.dj0;
        if (a >= 0 && a < g.m && b >= 0 && b < g.n) {
          const long long fk = (long long)b * g.m + a;
          const float v1 = f1[fk], v0 = f0 ? f0[fk] : 0.f;
          float2 t = ac[fk];
          t.x += me_c * (v1 - v0);  // -> eps gradient
          t.y += me_c * (v1 + v0);  // -> sigma gradient
          ac[fk] = t;
        }
      }
      if (gw && i == si && j == sj) gw[(long long)shot * s.nt + it] = sc * me_c;
      M1[k] = mn;
      if (i < NX) Ny1[k] = nyv;
      if (j < NY) Nx1[k] = nxv;
    }
    me_c = me_n;
    wy_c = wy_n;
    wx_c = wx_n;
    px_prev = px;
  }
}

// ==========================================================================
// small kernels
// ==========================================================================
__global__ void k_fill(float* p, long long n, float v) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    p[t] = v;
}

__global__ void k_fill_d(double* p, long long n, double v) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    p[t] = v;
}

// dJ/dEz(S_nt) = residual of the last step at every receiver (duplicates add)
// Traces [t0, t0+ntr) belong to shots q0.. which occupy buffer slots 0..
__global__ void k_inject_last(DevGrid g, int nt, int q0, int ntr, const int* __restrict__ tr_shot,
                              const int* __restrict__ tr_row, const int* __restrict__ tr_col,
                              const float* __restrict__ adj, int it, float* __restrict__ M) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntr) return;
  const int sh = tr_shot[t] - q0;
  const long long k = (long long)sh * g.FS + (long long)tr_col[t] * g.LD + tr_row[t];
  atomicAdd(&M[k], adj[(long long)t * nt + it]);
}

// Gather/scatter of the split-field aux state through a list of PML cells.
__global__ void k_gather(const float* __restrict__ src, long long FS, const int* __restrict__ idx,
                         int n, int ns, float* __restrict__ dst) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)n * ns) return;
  const int sh = (int)(t / n), q = (int)(t % n);
  dst[t] = src[(long long)sh * FS + idx[q]];
}

__global__ void k_scatter(const float* __restrict__ src, long long FS, const int* __restrict__ idx,
                          int n, int ns, float* __restrict__ dst) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)n * ns) return;
  const int sh = (int)(t / n), q = (int)(t % n);
  dst[(long long)sh * FS + idx[q]] = src[t];
}

// Ez(S) on the domain -> frame (used for the frame of a restored checkpoint)
__global__ void k_domain_frame(DevGrid g, int ns, const float* __restrict__ E,
                               float* __restrict__ frame, long long fstride) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long per = (long long)g.m * g.n;
  if (t >= per * ns) return;
  const int sh = (int)(t / per);
  const long long r = t % per;
  const int b = (int)(r / g.m), a = (int)(r % g.m);
  frame[(long long)sh * fstride + r] =
      E[(long long)sh * g.FS + (long long)(b + g.dj0) * g.LD + (a + g.di0)];
}

// acc (float2 per domain cell per shot) -> FP64 gradient, summed over shots
// in shot order; acc is cleared for the next segment.
__global__ void k_flush(int m, int n, int ns, float* __restrict__ acc, long long astride,
                        const double* __restrict__ fe, const double* __restrict__ fs,
                        double* __restrict__ ge, double* __restrict__ gs) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long per = (long long)m * n;
  if (t >= per) return;
  float2* a2 = reinterpret_cast<float2*>(acc);
  double se = 0.0, ss = 0.0;
  for (int sh = 0; sh < ns; ++sh) {
    float2 v = a2[(long long)sh * astride + t];
    se += (double)v.x;
    ss += (double)v.y;
    a2[(long long)sh * astride + t] = make_float2(0.f, 0.f);
  }
  ge[t] += fe[t] * se;
  gs[t] += fs[t] * ss;
}

// L2 residual e = d - d_o (w_bozdag.m:13-15) and per-block partial sums of
// e^2 in FP64 (w_E.m:15), finished by k_reduce_d.
__global__ void k_resid_l2(const float* __restrict__ d, const float* __restrict__ dob,
                           float* __restrict__ e, long long n, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const float v = d[t] - dob[t];
    e[t] = v;
    acc += (double)v * (double)v;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_reduce_d(const double* __restrict__ part, int n, double scale,
                           double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) acc += part[t];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = scale * sh[0];
}

// In-place second-order time derivative of an nz*nx x nt cube, one thread per
// spatial cell marching in time (differentiate_cube.m:23-51): one-sided
// three-point formulas at both ends, central differences inside.
__global__ void k_dcube(float* __restrict__ v, long long fr, int nt, double dt) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= fr) return;
  const double a = -1.5 / dt, b = 2.0 / dt, c = -0.5 / dt, d = 2.0 / dt;
  double u_km2 = 0.0, u_km1 = v[p], u_k = v[fr + p];
  v[p] = (float)(a * u_km1 + b * u_k + c * v[2 * fr + p]);
  for (int k = 1; k <= nt - 2; ++k) {
    const double u_kp1 = v[(long long)(k + 1) * fr + p];
    v[(long long)k * fr + p] = (float)(c * u_km1 - c * u_kp1);
    u_km2 = u_km1;
    u_km1 = u_k;
    u_k = u_kp1;
  }
  v[(long long)(nt - 1) * fr + p] = (float)(-c * u_km2 - d * u_km1 - a * u_k);
}

}  // namespace gjw
