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

// Neighbour-lane reads through DPP wave shifts (a VALU modifier, no LDS
// round trip): shfl_dn1 gives lane l the value of lane l+1 (wave_shl:1),
// shfl_up1 the value of lane l-1 (wave_shr:1).  The lane without a source
// (63, resp. 0) reads 0 (bound_ctrl); callers overwrite it or treat it as
// halo.  With a zero old value the compiler folds the shift into the
// instruction that consumes it instead of issuing a separate move.
__device__ __forceinline__ float shfl_dn1(float v) {
  const int x = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float shfl_up1(float v) {
  const int x = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, true));
}

// Raw buffer stores: a lane that has nothing to write passes an offset past
// the end of the buffer, which the hardware drops.  The store stays a single
// unconditional instruction, so the compiler can count the vector-memory
// operations in flight exactly (a store under a branch makes it wait for all
// of them, including loads issued for the next chunk).
constexpr unsigned kOOB = 0x7ffffff0u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(float* p, long long nfloats) {
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)(nfloats * 4), 0x00020000);
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, unsigned off_floats,
                                       bool on) {
  // the builtin takes the 32 data bits as an integer: pass the bits, not a converted value
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(on ? off_floats * 4u : kOOB),
                                        0, 0);
}

// Ez frames of the recompute sweep: FP32 (frame kind FR = 1), or BF16 pairs
// (FR = 2): the BF16 values (upper half of the FP32 bits, rounded to nearest
// even) of two consecutive steps in one 32-bit word, the earlier step in the
// low half.  Word plane q of a segment holds the frames of its slots 2q-1
// (low) and 2q (high); slot 0, the segment's first frame, is the high half
// of plane 0.  One store instruction per two frames halves the recompute
// sweep's frame stores, the instruction count that bounds it.
// Both conversions are one v_cvt_pk_bf16_f32 (round to nearest even), so a
// frame written alone (slot 0) and in a pair round alike.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned frame_pair_bits(float lo, float hi) {
  const f32x2_t v = {lo, hi};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ unsigned short frame_bf16(float v) {
  return __builtin_bit_cast(unsigned short, (__bf16)v);
}

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

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2), so launch-order neighbours land on different
// L2s.  The launch is 1-D; block b is remapped to logical tile
// t = (b mod 8) * (B/8) + b/8, which gives each XCD a contiguous run of
// logical tiles.  Logical tiles enumerate, fastest first: the shots of a
// group of kShotGroup shots, row tiles, column strips, shot groups.  The
// shots of a group run the same (row tile, strip) back to back on one XCD, so
// they share its coefficient lines in L2; the rows and columns a wave reads
// as halo were fetched by a wave on the same XCD a few tiles earlier.
constexpr int kShotGroup = 8;  // measured best against 4 and 16

struct TileId {
  int rowtile, strip, shot;
  int wy;  // multi-step kernels: the wave's place in its block-row's stack of four row tiles
};

// Logical tile t of a grid of gx row tiles, gy strip groups and ns shots.
__device__ __forceinline__ TileId tile_decode(unsigned t, unsigned gx, unsigned gy, unsigned ns) {
  const unsigned span = (unsigned)kShotGroup * gx * gy;  // tiles per full shot group
  const unsigned grp = t / span, rem = t - grp * span;
  const unsigned gc = min((unsigned)kShotGroup, ns - grp * (unsigned)kShotGroup);
  const unsigned rest = rem / gc;
  TileId id;
  id.shot = (int)(grp * (unsigned)kShotGroup + rem % gc);
  id.rowtile = (int)(rest % gx);
  id.strip = (int)(rest / gx);
  id.wy = (int)threadIdx.y;
  return id;
}

__device__ __forceinline__ unsigned tiles_x(int NX, int rows) {
  return (unsigned)((NX + 1 + rows - 1) / rows);
}
__device__ __forceinline__ unsigned tiles_y(int NY, int L) {
  return (unsigned)(((NY + 1 + L - 1) / L + 3) / 4);
}

__device__ __forceinline__ TileId tile_of_block(int NX, int NY, int L, int rows = 64) {
  const unsigned B = gridDim.x, b = blockIdx.x;
  const unsigned per = B / 8u;
  const unsigned t = (b < per * 8u) ? (b % 8u) * per + b / 8u : b;
  const unsigned gx = tiles_x(NX, rows), gy = tiles_y(NY, L);
  return tile_decode(t, gx, gy, B / (gx * gy));
}

// Multi-step kernels stack a block's 4 waves vertically: block (rowtile,
// strip) covers row tiles 4*rowtile .. 4*rowtile+3 of one column strip, so
// the block reads about 4 x 256 contiguous bytes of each column and field.
__device__ __forceinline__ unsigned vtiles_x(int NX, int rows) {
  return (unsigned)(((NX + 1 + rows - 1) / rows + 3) / 4);
}
__device__ __forceinline__ unsigned vtiles_y(int NY, int L) {
  return (unsigned)((NY + 1 + L - 1) / L);
}
// Logical tile t of the stacked layout.  With slow_first the boundary
// block-rows (0 and gx-1, whose waves take the slower row-class path through
// the PML strips) come first within each shot group: the slow blocks start in
// the first rounds and the uniform interior blocks fill in after them, so no
// slow block is left for the end of the launch (the 8-step forward: 1.49 ms
// against 1.56 ms).  The adjoint keeps the plain order, which measured
// faster for it (1.84 ms against 1.87 ms).  Shots stay fastest, as in
// tile_decode.
__device__ __forceinline__ TileId vtile_decode(unsigned t, unsigned gx, unsigned gy, unsigned ns,
                                              bool slow_first) {
  // Each XCD takes a contiguous eighth of the logical tiles; the reordering
  // keeps the XCDs balanced only when that eighth is whole shot groups (else
  // some XCDs would get only the slow block-rows).
  const unsigned groups = (ns + (unsigned)kShotGroup - 1) / (unsigned)kShotGroup;
  if (!slow_first || gx < 3 || groups % 8u != 0) return tile_decode(t, gx, gy, ns);
  const unsigned span = (unsigned)kShotGroup * gx * gy;
  const unsigned grp = t / span, rem = t - grp * span;
  const unsigned gc = min((unsigned)kShotGroup, ns - grp * (unsigned)kShotGroup);
  const unsigned nb = 2u * gy * gc;  // tiles of the boundary block-rows
  TileId id;
  id.wy = (int)threadIdx.y;
  if (rem < nb) {
    const unsigned r = rem / gc;
    id.shot = (int)(grp * (unsigned)kShotGroup + rem % gc);
    id.rowtile = (r & 1u) ? (int)gx - 1 : 0;
    id.strip = (int)(r >> 1);
  } else {
    const unsigned q = rem - nb, r = q / gc;
    id.shot = (int)(grp * (unsigned)kShotGroup + q % gc);
    id.rowtile = 1 + (int)(r % (gx - 2));
    id.strip = (int)(r / (gx - 2));
  }
  return id;
}

// A block-row's stack of four row tiles may also run as q = 4 / blockDim.y
// blocks of blockDim.y waves (the forward sweep's 8-step kernel takes two
// blocks of two: 1.017 against 1.071 ms per launch, where the other step
// kernels ran 1.5-2% slower that way).  Physical block b takes sub-stack
// k mod q of logical block k / q, with k its place in its XCD's run (b / 8):
// the q blocks of one stack stay on one XCD and its L2.
__device__ __forceinline__ TileId vtile_of_block(int NX, int NY, int L, int rows,
                                                bool slow_first) {
  const unsigned q = 4u / blockDim.y;
  const unsigned b = blockIdx.x, B = gridDim.x / q;  // logical blocks
  const unsigned per = B / 8u;
  unsigned t, sub;
  if (b < per * 8u * q) {
    const unsigned k = b / 8u;
    t = (b % 8u) * per + k / q;
    sub = k % q;
  } else {
    t = b / q;
    sub = b % q;
  }
  const unsigned gx = vtiles_x(NX, rows), gy = vtiles_y(NY, L);
  TileId id = vtile_decode(t, gx, gy, B / (gx * gy), slow_first);
  id.wy = (int)(sub * blockDim.y + threadIdx.y);
  return id;
}

// Row tile t of the multi-step kernels (R = 64 - 2 NS owned rows per wave):
// tile t owns rows t*R .. t*R+R-1, except that the last tile is aligned to
// the bottom row (it owns NX+1-R .. NX) and the tile before it stops where
// the last one begins.  The wave's 64 lanes hold rows r0 = (first owned row)
// - NS .. r0+63.  With the bottom alignment a wave whose rows reach into the
// bottom PML strip owns only rows near it: for a 40-cell PML and NS = 8 the
// top and the bottom strip take one wave each.
struct RowTile {
  int r0, own_hi;
};
__device__ __forceinline__ RowTile row_tile(int t, int NX, int R, int NS) {
  const int T = (NX + 1 + R - 1) / R;
  RowTile rt;
  if (t >= T) {  // no owned row
    rt.r0 = t * R - NS;
    rt.own_hi = -1;
    return rt;
  }
  const int lo = t == T - 1 ? max(NX + 1 - R, 0) : t * R;
  rt.r0 = lo - NS;
  rt.own_hi = t == T - 1 ? NX : t + 1 == T - 1 ? NX - R : (t + 1) * R - 1;
  return rt;
}

// Several surveys in one launch: surveys[a.sid[k]] owns blocks
// [a.blk0[k], a.blk0[k+1]) (a multiple of 8 each, so the XCD-aware order holds
// per survey); its own step, tile width, buffer parities and frame slot come
// with it.  Blocks past a survey's tiles exit.
__device__ __forceinline__ bool multi_tile(const MultiArgs& a, const MultiTable& tab,
                                           int rows, int& k, TileId& tid, bool vert = false,
                                           bool slow_first = false) {
  const int b = (int)blockIdx.x;
  k = 0;
  while (k + 1 < a.n && b >= a.blk0[k + 1]) ++k;
  const DevSurvey& S = tab.sv[a.sid[k]];
  const unsigned B = (unsigned)(a.blk0[k + 1] - a.blk0[k]), bl = (unsigned)(b - a.blk0[k]);
  const unsigned per = B / 8u;
  const unsigned t = (bl % 8u) * per + bl / 8u;
  const unsigned gx = vert ? vtiles_x(S.g.NX, rows) : tiles_x(S.g.NX, rows);
  const unsigned gy = vert ? vtiles_y(S.g.NY, a.L[k]) : tiles_y(S.g.NY, a.L[k]);
  if (t >= gx * gy * (unsigned)S.ns) return false;
  tid = vert ? vtile_decode(t, gx, gy, (unsigned)S.ns, slow_first)
             : tile_decode(t, gx, gy, (unsigned)S.ns);
  return true;
}

// ==========================================================================
// forward: fused step
// ==========================================================================
// Columns per software-pipelined chunk: the loads of U columns are issued
// together before any of them is used, so each wavefront keeps ~5U 256-byte
// requests in flight (one column alone leaves the kernel latency-bound).
constexpr int U = 4;

// One column of the general (strip-aware) update; used for every column the
// unrolled interior chunk does not cover.
__device__ __forceinline__ void fwd_column(const DevGrid& g, int lane, int i, int j, bool ok,
                                           const float* __restrict__ E0,
                                           const float* __restrict__ Hx0,
                                           const float* __restrict__ Hy0, float* __restrict__ X,
                                           float* __restrict__ Y, float& e_c, float& hx_prev,
                                           float& en, float& hxn, float& hyn) {
  const int NX = g.NX, NY = g.NY, P = g.P;
  const long long LD = g.LD;
  const long long k = (long long)j * LD + i;
  const float e_n = (ok && j < NY) ? E0[k + LD] : 0.f;
  const float hx = (ok && j < NY) ? Hx0[k] : 0.f;
  const float hy = (ok && i < NX) ? Hy0[k] : 0.f;
  float e_dn = shfl_dn1(e_c);
  if (lane == 63) e_dn = (i + 1 <= NX) ? E0[k + 1] : 0.f;
  hyn = (ok && i < NX) ? hy_new(g, i, j, hy, e_c, e_dn) : 0.f;
  hxn = (ok && j < NY) ? hx_new(g, i, j, hx, e_c, e_n) : 0.f;
  float hy_up = shfl_up1(hyn);
  if (lane == 0) hy_up = (ok && i >= 1) ? hy_new(g, i - 1, j, Hy0[k - 1], E0[k - 1], e_c) : 0.f;
  en = 0.f;
  if (ok) {
    const int kd = ekind(NX, NY, P, i, j);
    const float dhy = hyn - hy_up, dhx = hxn - hx_prev;
    if (kd == KIND_INNER) {
      const float2 c = g.cc[k];
      en = c.x * e_c + c.y * (dhy - dhx);
    } else if (kd == KIND_PML) {
      // split field Ez = Ezx + Ezy (w_solve.m:232-267)
      const float x = g.cxe[k] * X[k] + g.cxh[k] * dhy;
      const float y = g.cye[k] * Y[k] + g.cyh[k] * dhx;
      X[k] = x;
      Y[k] = y;
      en = x + y;
    }
  }
  e_c = e_n;
}

__device__ __forceinline__ void fwd1_body(const DevGrid& g, const DevShots& s, int L, TileId tid, const float* __restrict__ E0,
           const float* __restrict__ Hx0, const float* __restrict__ Hy0, float* __restrict__ E1,
           float* __restrict__ Hx1, float* __restrict__ Hy1, float* __restrict__ X,
           float* __restrict__ Y, int it, float* __restrict__ tr, float* __restrict__ frame,
           long long fstride) {
  const int lane = threadIdx.x;
  const int NX = g.NX, NY = g.NY, P = g.P;
  const int shot = tid.shot;
  const int i0 = tid.rowtile * 64;
  // threadIdx.y is constant across a 64-lane wavefront: make the strip start
  // (and everything derived from it) scalar so loop control stays uniform
  const int jb = __builtin_amdgcn_readfirstlane((tid.strip * blockDim.y + threadIdx.y) * L);
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
  const float bx0 = g.bx0, by0 = g.by0;
  const bool has_rec = tr && i0 <= s.rmax && i0 + 63 >= s.rmin;

  float e_c = ok ? E0[(long long)jb * LD + i] : 0.f;
  float hx_prev = 0.f;  // Hx'(i, j-1)
  if (jb >= 1 && ok) {
    const long long km = (long long)(jb - 1) * LD + i;
    hx_prev = hx_new(g, i, jb - 1, Hx0[km], E0[km], e_c);
  }
  // per-column epilogue: source, stores, receivers, domain frame
  auto emit = [&](int j, float en, float hxn, float hyn) {
    if (j == sj && i == si) {
      en += wsrc;                                       // soft source (w_solve.m:92)
      if (s.jz) s.jz[(long long)shot * s.nt + it] = en;  // source recorder (w_solve.m:104)
    }
    if (ok) {
      const long long k = (long long)j * LD + i;
      E1[k] = en;
      if (j < NY) Hx1[k] = hxn;
      if (i < NX) Hy1[k] = hyn;
      if (has_rec) {  // receivers of this column (w_solve.m:100)
        const int qa = cp[j], qb = cp[j + 1];
        for (int q = qa; q < qb; ++q)
          if (s.crow[q] == i) tr[(long long)s.cid[q] * s.nt + it] = en;
      }
      if (fr) {  // domain frames keep the field pitch: cell (i, j) at (j - dj0) * LD + i
        const int a = i - g.di0, b = j - g.dj0;
        if (a >= 0 && a < g.m && b >= 0 && b < g.n) fr[(long long)b * LD + i] = en;
      }
    }
  };
  int j = jb;
  while (j < je) {
    if (fastrows && j >= P + 1 && j + U - 1 <= NY - P - 1 && j + U <= je) {
      // interior chunk: mu_r = 1, sigma_m = 0 => Chxh = Chyh = 1 (w_coeff.m:34-40)
      const long long k = (long long)j * LD + i;
      float en_[U], hx_[U], hy_[U], h63[U], hu0[U], eu0[U];
      float2 c_[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long kk = k + u * LD;
        en_[u] = E0[kk + LD];
        hx_[u] = Hx0[kk];
        hy_[u] = Hy0[kk];
        c_[u] = g.cc[kk];
      }
      if (lane == 63) {
#pragma unroll
        for (int u = 0; u < U; ++u) h63[u] = E0[k + u * LD + 1];
      }
      if (lane == 0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          hu0[u] = Hy0[k + u * LD - 1];
          eu0[u] = E0[k + u * LD - 1];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        float e_dn = shfl_dn1(e_c);
        if (lane == 63) e_dn = h63[u];
        const float hyn = hy_[u] + by0 * (e_dn - e_c);
        const float hxn = hx_[u] + bx0 * (en_[u] - e_c);
        float hy_up = shfl_up1(hyn);
        if (lane == 0) hy_up = hu0[u] + by0 * (e_c - eu0[u]);
        const float en = c_[u].x * e_c + c_[u].y * ((hyn - hy_up) - (hxn - hx_prev));
        emit(j + u, en, hxn, hyn);
        e_c = en_[u];
        hx_prev = hxn;
      }
      j += U;
    } else {
      float en, hxn, hyn;
      fwd_column(g, lane, i, j, ok, E0, Hx0, Hy0, X, Y, e_c, hx_prev, en, hxn, hyn);
      emit(j, en, hxn, hyn);
      hx_prev = hxn;
      ++j;
    }
  }
}

__global__ void __launch_bounds__(256)
k_fwd_step(DevGrid g, DevShots s, int L, const float* __restrict__ E0,
           const float* __restrict__ Hx0, const float* __restrict__ Hy0, float* __restrict__ E1,
           float* __restrict__ Hx1, float* __restrict__ Hy1, float* __restrict__ X,
           float* __restrict__ Y, int it, float* __restrict__ tr, float* __restrict__ frame,
           long long fstride) {
  fwd1_body(g, s, L, tile_of_block(g.NX, g.NY, L), E0, Hx0, Hy0, E1, Hx1, Hy1, X, Y, it, tr, frame, fstride);
}

__global__ void __launch_bounds__(256)
k_fwd_step_multi(const MultiTable tab, MultiArgs a) {
  int k;
  TileId tid;
  if (!multi_tile(a, tab, 64, k, tid)) return;
  const DevSurvey& S = tab.sv[a.sid[k]];
  const int c = a.cur[k], x = a.xc[k];
  float* f = S.fld;
  const long long z = S.fsz;
  float* fr = a.fslot[k] >= 0 ? S.frames + (long long)a.fslot[k] * S.FD : nullptr;
  fwd1_body(S.g, S.s, a.L[k], tid, f + 3 * c * z, f + (3 * c + 1) * z, f + (3 * c + 2) * z,
            f + 3 * (c ^ 1) * z, f + (3 * (c ^ 1) + 1) * z, f + (3 * (c ^ 1) + 2) * z,
            f + (6 + 2 * x) * z, f + (7 + 2 * x) * z, a.it[k], a.rec ? S.tr : nullptr, fr,
            S.fstride);
}

// ==========================================================================
// forward: two fused steps per pass (temporal blocking)
// ==========================================================================
// One launch advances every shot from step n to n+2.  A wavefront's 64 lanes
// hold rows r0..r0+63 with r0 = 60*t - 2: lanes 2..61 own their rows, lanes
// 0-1 and 62-63 recompute the neighbours' rows that the second step needs,
// so no lane loads a halo of its own.  Along the march, step 2 runs one
// column behind step 1 (column j for step 1, j-1 for step 2), which is all
// the second step needs in that direction.  Each field value is read and
// written once per two steps.  Split-field aux X/Y ping-pong (X0 -> X2) so
// that halo lanes can read the start-of-pass values of cells other waves own.

// Masked raw buffer load: a lane with on == false reads past the end of the
// buffer and gets 0 without touching memory (no branch, see bstore).
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned off_floats, bool on) {
  return __uint_as_float(
      __builtin_amdgcn_raw_buffer_load_b32(r, (int)(on ? off_floats * 4u : kOOB), 0, 0));
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 bload2(__amdgpu_buffer_rsrc_t r, unsigned off_floats, bool on) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(on ? off_floats * 4u : kOOB), 0, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ __forceinline__ void bstore2(float2 v, __amdgpu_buffer_rsrc_t r, unsigned off_floats,
                                        bool on) {
  u32x2 w;
  w.x = __float_as_uint(v.x);
  w.y = __float_as_uint(v.y);
  __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)(on ? off_floats * 4u : kOOB), 0, 0);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 bload4(__amdgpu_buffer_rsrc_t r, unsigned off_floats, bool on) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(on ? off_floats * 4u : kOOB), 0, 0);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                     __uint_as_float(v.w));
}

// Loads from a survey's coefficient block through one descriptor (DevGrid::coef):
// soff is the array's byte offset in the block.  It is wave-uniform and rides
// in the instruction's scalar offset, so the per-lane offset and the
// out-of-range mask stay those of the array itself; one descriptor instead of
// one per array keeps the kernels' scalar registers from spilling.
__device__ __forceinline__ float bloadk(__amdgpu_buffer_rsrc_t r, unsigned soff,
                                        unsigned off_floats, bool on) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
      r, (int)(on ? off_floats * 4u : kOOB), (int)soff, 0));
}
__device__ __forceinline__ float2 bload2k(__amdgpu_buffer_rsrc_t r, unsigned soff,
                                          unsigned off_floats, bool on) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(on ? off_floats * 4u : kOOB),
                                                       (int)soff, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ __forceinline__ float4 bload4k(__amdgpu_buffer_rsrc_t r, unsigned soff,
                                          unsigned off_floats, bool on) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(on ? off_floats * 4u : kOOB),
                                                        (int)soff, 0);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                     __uint_as_float(v.w));
}
__device__ __forceinline__ void bstorek(float v, __amdgpu_buffer_rsrc_t r, unsigned soff,
                                        unsigned off_floats, bool on) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(on ? off_floats * 4u : kOOB),
                                        (int)soff, 0);
}
// byte offsets of the arrays in the coefficient block (units of FS floats: see DevGrid)
struct CoefOff {
  unsigned cc, cxe, cxh, cyh, hay, hby, u0, u1, u2;
};
__device__ __forceinline__ CoefOff coef_off(const DevGrid& g) {
  const unsigned f = (unsigned)g.FS * 4u;
  return CoefOff{0u, 2u * f, 3u * f, 5u * f, 8u * f, 9u * f, 10u * f, 14u * f, 18u * f};
}

// Waves per SIMD the multi-step kernels are compiled for (the compiler keeps
// their registers within that occupancy; the kernels are bandwidth-bound).
#ifndef GJW_FWD_WAVES
#define GJW_FWD_WAVES 4
#endif
#ifndef GJW_ADJ_WAVES
#define GJW_ADJ_WAVES 4
#endif
#ifndef GJW_ADJM_WAVES  // multi-survey adjoint (carries the survey table too)
#define GJW_ADJM_WAVES 3
#endif
#ifndef GJW_ADJ6_WAVES
#define GJW_ADJ6_WAVES 3
#endif
#ifndef GJW_ADJ8_WAVES
#define GJW_ADJ8_WAVES 2
#endif
#ifndef GJW_FWD8_WAVES
#define GJW_FWD8_WAVES 3
#endif

// FR: the recompute sweep stores Ez frames of the domain (1: FP32, 2: BF16).
//
// Column handling.  Step-1 columns jb-1 .. je are computed (step 2 runs one
// column behind; the first two produce only what the strip's first output
// column needs).  Interior columns (both steps' columns inside
// [P+1, NY-P-1]) go through U-column chunks whose memory operations are all
// unconditional raw buffer loads/stores with per-lane masks, so the compiler
// waits only for what each chunk needs:
//   * waves whose 64 rows are all inner rows take the scalar-coefficient path;
//   * the other waves (top and bottom PML rows) take a row-class path: every
//     lane knows whether its row is inner, in the top strip (coefficients
//     depend on the row only and sit in registers) or in the bottom strip
//     (coefficients depend on eps(i, j) and are loaded per cell).
// The left/right PML columns, the grid's edge columns and the chunks that
// straddle them use a per-cell coefficient table (univ_chunk).
// NS fused forward steps per pass (NS = 2 or 4).  A wavefront's 64 lanes
// hold rows r0 .. r0+63 with r0 = (64 - 2 NS) t - NS; lanes NS .. 63-NS own
// their rows, the outer lanes recompute the neighbours' rows the later steps
// need.  Level l (step it+l) runs l columns behind level 0 along the march,
// so every level reads its inputs from registers: the field values of the
// previous level's current column and of its own previous column.  Only the
// last level stores the fields; every level records receivers, the source
// recorder and (FR, recompute sweep) its domain frame, slot l at
// frame1 + l * LD * n.
template <int NS, int FR>
__device__ __forceinline__ void fwdN_body(const DevGrid& g, const DevShots& s, int L, TileId tid,
                                          const float* __restrict__ E0,
                                          const float* __restrict__ Hx0,
                                          const float* __restrict__ Hy0,
                                          const float* __restrict__ X0,
                                          const float* __restrict__ Y0, float* __restrict__ E2,
                                          float* __restrict__ Hx2, float* __restrict__ Hy2,
                                          float* __restrict__ X2, float* __restrict__ Y2, int it,
                                          float* __restrict__ tr, float* __restrict__ frame1,
                                          long long fstride) {
  constexpr int R = 64 - 2 * NS;
  constexpr int UC = NS >= 8 ? 1 : NS == 4 ? 2 : U;  // columns per strip / row-class chunk
  constexpr int UI = NS >= 8 ? 2 : U;                // columns per interior chunk
  const int lane = threadIdx.x;
  const int NX = g.NX, NY = g.NY, P = g.P;
  const long long LD = g.LD;
  const int shot = tid.shot;
  // threadIdx.y is constant across a 64-lane wavefront: make the tile origin
  // (and everything derived from it) scalar so loop control stays uniform
  const RowTile rt = row_tile((int)(tid.rowtile * 4 + tid.wy), NX, R, NS);
  const int r0 = __builtin_amdgcn_readfirstlane(rt.r0);
  const int own_hi = __builtin_amdgcn_readfirstlane(rt.own_hi);
  const int jb = __builtin_amdgcn_readfirstlane(tid.strip * L);
  if (r0 + NS > NX) return;  // no owned row
  const int je = min(jb + L, NY + 1);
  const int i = r0 + lane;
  const bool okrow = i >= 0 && i <= NX;
  const bool own = okrow && lane >= NS && lane <= 63 - NS && i <= own_hi;
  const long long so = (long long)shot * g.FS;
  E0 += so; Hx0 += so; Hy0 += so; X0 += so; Y0 += so;
  E2 += so; Hx2 += so; Hy2 += so; X2 += so; Y2 += so;
  const int si = s.si[shot], sj = s.sj[shot];
  float ws[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) ws[l] = s.sc[shot] * s.wav[(long long)shot * s.nt + it + l];
  const int* cp = s.cptr + (long long)shot * (NY + 2);
  const bool has_rec = tr && r0 + NS <= s.rmax && own_hi >= s.rmin;
  // Every row the owned rows depend on is an interior row: level l's values
  // are needed on rows within NS-1-l of the owned ones, so the rows of lanes 0
  // and 63 (and, in a tile that stops early, the rows past own_hi+NS-1) are
  // never needed; the interior update may run there whatever class they are.
  const bool inner_rows = (r0 + 1 >= P + 1) && (own_hi + NS - 1 <= NX - P - 1);
  const bool dom_row = i - g.di0 >= 0 && i - g.di0 < g.m;
  auto dom_col = [&](int c) { return c - g.dj0 >= 0 && c - g.dj0 < g.n; };
  const float bx0 = g.bx0, by0 = g.by0;
  const long long ic = okrow ? i : 0;  // address row for loads of valid lanes
  const long long FD = (long long)LD * g.n;
  // source recorder (w_solve.m:104): Ez at the source cell after each step
  const __amdgpu_buffer_rsrc_t rJZ =
      rsrc_of(s.jz ? s.jz + (long long)shot * s.nt : nullptr, s.jz ? s.nt : 0);
  // domain frames of the NS steps (slot l at l * FD)
  // domain frames of the NS steps: slot l at l * FD (FR = 1), or the pairs of
  // slots (2p, 2p+1) in word plane p (FR = 2)
  // (FR = 2: the descriptor starts NS * LD words early, so that a level's
  // column lag, -l * LD, rides in the scalar offset with the plane)
  const __amdgpu_buffer_rsrc_t rFR =
      FR == 2 ? rsrc_of(frame1 + (long long)shot * fstride - NS * LD, NS / 2 * FD + NS * LD)
              : rsrc_of(FR ? frame1 + (long long)shot * fstride : nullptr, FR ? NS * FD : 0);
  // FR = 2: the frame of each even level is held a column (fh), until the
  // odd level after it reaches that column and stores the pair in one 32-bit
  // store; the even level's value of this column (fhn) replaces it after
  // that store.  Every chunk kind runs every level of every column, so the
  // held values carry across the kinds.
  float fh[NS / 2 > 0 ? NS / 2 : 1], fhn[NS / 2 > 0 ? NS / 2 : 1];
#pragma unroll
  for (int m = 0; m < NS / 2; ++m) fh[m] = fhn[m] = 0.f;
  // level l's frame at column c - l: v0 = (c - dj0) * LD + row, the lanes'
  // offset of level 0's column; plane l / 2 and the column lag in the scalar
  // offset
  auto fsoff = [&](int l) { return (unsigned)(l / 2 * FD + (NS - l) * LD) * 4u; };
  auto frame_pair = [&](int l, float e, unsigned v0, bool on) {
    if (l % 2 == 0) {
      fhn[l / 2] = e;
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(frame_pair_bits(fh[l / 2], e), rFR,
                                            (int)(on ? v0 * 4u : kOOB), (int)fsoff(l), 0);
      fh[l / 2] = fhn[l / 2];
    }
  };

  const __amdgpu_buffer_rsrc_t rE0 = rsrc_of(const_cast<float*>(E0), g.FS),
                               rHx0 = rsrc_of(const_cast<float*>(Hx0), g.FS),
                               rHy0 = rsrc_of(const_cast<float*>(Hy0), g.FS),
                               rX0 = rsrc_of(const_cast<float*>(X0), g.FS),
                               rY0 = rsrc_of(const_cast<float*>(Y0), g.FS),
                               rE2 = rsrc_of(E2, g.FS), rHx2 = rsrc_of(Hx2, g.FS),
                               rHy2 = rsrc_of(Hy2, g.FS), rX2 = rsrc_of(X2, g.FS),
                               rY2 = rsrc_of(Y2, g.FS),
                               rK = rsrc_of(const_cast<float*>(g.coef), 19 * g.FS);
  const CoefOff ko = coef_off(g);

  // row classes (valid in interior columns)
  const bool e_inner = okrow && i >= P + 1 && i <= NX - P - 1;
  const bool e_xn = okrow && i >= 1 && i <= P;
  const bool e_xp = okrow && i >= NX - P && i <= NX - 1;
  const bool e_pml = e_xn || e_xp;
  const bool h_inner = okrow && i >= P && i <= NX - P - 1;
  const bool h_xn = okrow && i >= 0 && i <= P - 1;
  const bool h_xp = okrow && i >= NX - P && i <= NX - 1;
  const bool h_any = h_inner || h_xn || h_xp;  // Hy exists (i <= NX-1)
  const bool notring = okrow && i >= 1 && i <= NX - 1;
  // split-field (PML) Ez cell in column c
  auto pml_cell = [&](int c) {
    return notring && c >= 1 && c <= NY - 1 && (e_pml || c <= P || c >= NY - P);
  };
  // top-strip coefficients depend on the row only: read them once
  // (from the coefficient table at column P+1, loaded without a branch)
  const bool rows_on = okrow && !inner_rows;
  const unsigned kr = (unsigned)((P + 1) * LD + ic);
  const float4 kr0 = bload4k(rK, ko.u0, 4 * kr, rows_on), kr1 = bload4k(rK, ko.u1, 4 * kr, rows_on);
  const float kr2 = bloadk(rK, ko.u2, kr, rows_on);
  const float cxe_r = rows_on ? kr0.y : 1.f, cxh_r = rows_on ? kr0.z : 0.f;
  const float cyh_r = rows_on ? kr1.x : 0.f;
  // Hy coefficients of rows whose Hy update does not depend on the column
  const float hay_row = rows_on ? kr1.w : 1.f;
  const float hby_row = rows_on ? kr2 : by0;

  // receiver samples of column j (0 <= j <= NY)
  auto record = [&](int j, float e, int step) {
    const int qa = cp[j], qb = cp[j + 1];
    for (int q = qa; q < qb; ++q)
      if (s.crow[q] == i) tr[(long long)s.cid[q] * s.nt + step] = e;
  };
  // Receivers: level 0 looks up the trace of the lane's cell in its column
  // (-1 if none; one receiver per cell) and the lookup travels with the
  // column to the later levels (rid_l[l]: level l's column); the samples are
  // written through the shot's traces.  Surveys with several receivers in one
  // cell use the per-receiver loop instead.
  const int t_lo = has_rec ? s.rp[shot] : 0;
  const __amdgpu_buffer_rsrc_t rTR =
      rsrc_of(has_rec ? tr + (long long)t_lo * s.nt : nullptr,
              has_rec ? (long long)(s.rp[shot + 1] - t_lo) * s.nt : 0);
  auto rec_lookup = [&](int j) {
    int r = -1;
    if (has_rec && j >= 0 && j <= NY) {
      const int qa = cp[j], qb = cp[j + 1];
      for (int q = qa; q < qb; ++q)
        if (s.crow[q] == i) r = s.cid[q];
    }
    return r;
  };
  int rid_l[NS];  // trace of the lane's cell in the column of level l
#pragma unroll
  for (int l = 0; l < NS; ++l) rid_l[l] = -1;
  // per-level outputs that leave the kernel: receivers, source recorder, frame
  auto level_out = [&](int l, int cl, float e) {
    const bool ol = own && cl >= jb && cl < je;
    if (has_rec) {
      if (s.rec_multi) {
        if (ol) record(cl, e, it + l);
      } else {
        bstore(e, rTR, (unsigned)((rid_l[l] - t_lo) * s.nt + it + l), ol && rid_l[l] >= 0);
      }
    }
    bstore(e, rJZ, (unsigned)(it + l), ol && i == si && cl == sj);
    if (FR == 1)
      bstore(e, rFR, (unsigned)(l * FD + (cl - g.dj0) * LD + ic), ol && dom_row && dom_col(cl));
    if (FR == 2)
      frame_pair(l, e, (unsigned)((cl + l - g.dj0) * LD + ic), ol && dom_row && dom_col(cl));
  };
  // Levels whose column c-l is one of the strip's own columns [jb, je), as
  // bit l of a wave-uniform mask (one range per column instead of two
  // compares per level).
  auto level_mask = [&](int c) {
    const int lo = max(0, c - je + 1), hi = min(NS - 1, c - jb);
    return hi >= lo ? (2u << hi) - (1u << lo) : 0u;
  };
  // level_out for the interior chunks: their columns all lie inside the
  // stored domain's columns (P+1 .. NY-P-1 within dj0 .. dj0+n-1), so only the
  // strip's own columns and the lane's row decide a frame store
  const bool own_dom = own && dom_row;
  auto level_out_in = [&](int l, int cl, float e, unsigned lv) {
    const bool on = (lv >> l) & 1u;
    const bool ol = own && on;
    if (has_rec) {
      if (s.rec_multi) {
        if (ol) record(cl, e, it + l);
      } else {
        bstore(e, rTR, (unsigned)((rid_l[l] - t_lo) * s.nt + it + l), ol && rid_l[l] >= 0);
      }
    }
    bstore(e, rJZ, (unsigned)(it + l), ol && i == si && cl == sj);
    if (FR == 1)
      bstorek(e, rFR, (unsigned)(l * (FD - LD)) * 4u, (unsigned)((cl + l - g.dj0) * LD + ic),
              own_dom && on);
    if (FR == 2) frame_pair(l, e, (unsigned)((cl + l - g.dj0) * LD + ic), own_dom && on);
  };

  // ---- state.  Level 0 steps column c from the loaded fields; level l >= 1
  // steps column c-l from level l-1's outputs (its previous column: *_in[l];
  // its current column: this iteration's output) with the coefficients of
  // column c-l carried along (*_l[l]).  hxp[l]: Hx of level l at its previous
  // column.  Level 0 starts at column jb-NS+1, which needs Hx^{n+1}(jb-NS).
  const int j1 = jb - (NS - 1);
  float e0_c, e_in[NS], hx_in[NS], hy_in[NS], x_in[NS], y_in[NS], hxp[NS];
  float2 cc_l[NS];
  float cxe_l[NS], cxh_l[NS], cyh_l[NS], hay_l[NS], hby_l[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    e_in[l] = hx_in[l] = hy_in[l] = x_in[l] = y_in[l] = hxp[l] = 0.f;
    cc_l[l] = make_float2(0.f, 0.f);
    cxe_l[l] = 1.f;
    cxh_l[l] = cyh_l[l] = 0.f;
    hay_l[l] = 1.f;
    hby_l[l] = 0.f;
  }
  {
    const unsigned ka = (unsigned)(j1 * LD + ic), kb = ka - (unsigned)LD;
    const bool ina = okrow && j1 >= 0 && j1 <= NY, inb = okrow && j1 - 1 >= 0 && j1 - 1 <= NY;
    e0_c = bload(rE0, ka, ina);
    const float eb = bload(rE0, kb, inb), hxb = bload(rHx0, kb, inb && j1 - 1 < NY);
    const float4 kbb = bload4k(rK, ko.u1, 4 * kb, inb);  // (cyh, AX, BX, AY) of column j1-1
    hxp[0] = kbb.y * hxb + kbb.z * (e0_c - eb);
  }
  // hand level l's outputs to level l+1 and shift the coefficient carries
  auto shift = [&](const float* eo, const float* hxo, const float* hyo, const float* xo,
                   const float* yo) {
#pragma unroll
    for (int l = NS - 1; l >= 1; --l) {
      e_in[l] = eo[l - 1];
      hx_in[l] = hxo[l - 1];
      hy_in[l] = hyo[l - 1];
      x_in[l] = xo[l - 1];
      y_in[l] = yo[l - 1];
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) hxp[l] = hxo[l];
  };

  // every level's columns of a chunk of w columns at jj are interior
  auto chunkable = [&](int jj, int w) { return jj - (NS - 1) >= P + 1 && jj + w - 1 <= NY - P - 1; };

  // ---- chunk of any columns (left/right strips, grid edges, strip ends):
  // every coefficient per cell from the table, Ez' = ce Ez + X' + Y'
  auto univ_chunk = [&](int jj) {
    float en_[UC], hx_[UC], hy_[UC], x_[UC], y_[UC], kc_[UC];
    float4 ka_[UC], kb_[UC];
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const int c = jj + u;
      const bool cin = okrow && c >= 0 && c <= NY;
      const unsigned kk = (unsigned)(c * LD + ic);
      en_[u] = bload(rE0, kk + (unsigned)LD, okrow && c + 1 >= 0 && c + 1 <= NY);
      hx_[u] = bload(rHx0, kk, cin && c < NY);
      hy_[u] = bload(rHy0, kk, cin && i < NX);
      const bool pm = pml_cell(c);
      x_[u] = bload(rX0, kk, pm);
      y_[u] = bload(rY0, kk, pm);
      ka_[u] = bload4k(rK, ko.u0, 4 * kk, cin);
      kb_[u] = bload4k(rK, ko.u1, 4 * kk, cin);
      kc_[u] = bloadk(rK, ko.u2, kk, cin);
    }
    // table coefficients of the later levels' first columns
    float4 ka_l[NS], kb_l[NS];
    float kc_l[NS];
#pragma unroll
    for (int l = 1; l < NS; ++l) {
      const int c = jj - l;
      const bool cin = okrow && c >= 0 && c <= NY;
      const unsigned kk = (unsigned)(c * LD + ic);
      ka_l[l] = bload4k(rK, ko.u0, 4 * kk, cin);
      kb_l[l] = bload4k(rK, ko.u1, 4 * kk, cin);
      kc_l[l] = bloadk(rK, ko.u2, kk, cin);
    }
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const int c = jj + u;
      float eo[NS], hxo[NS], hyo[NS], xo[NS], yo[NS];
      if (has_rec) rid_l[0] = rec_lookup(c);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const int cl = c - l;
        const float ec = l == 0 ? e0_c : e_in[l], en = l == 0 ? en_[u] : eo[l - 1];
        const float hx = l == 0 ? hx_[u] : hx_in[l], hy = l == 0 ? hy_[u] : hy_in[l];
        const float x = l == 0 ? x_[u] : x_in[l], y = l == 0 ? y_[u] : y_in[l];
        const float4 ka = l == 0 ? ka_[u] : ka_l[l], kb = l == 0 ? kb_[u] : kb_l[l];
        const float kc = l == 0 ? kc_[u] : kc_l[l];
        // (ce, cxe, cxh, cye) = ka, (cyh, AX, BX, AY) = kb, BY = kc
        const float e_dn = shfl_dn1(ec);
        const float hy1 = kb.w * hy + kc * (e_dn - ec);
        const float hx1 = kb.y * hx + kb.z * (en - ec);
        const float dhy = hy1 - shfl_up1(hy1), dhx = hx1 - hxp[l];
        const float x1 = ka.y * x + ka.z * dhy;
        const float y1 = ka.w * y + kb.x * dhx;
        float e1 = ka.x * ec + x1 + y1;
        e1 += (cl == sj && i == si) ? ws[l] : 0.f;
        eo[l] = e1;
        hxo[l] = hx1;
        hyo[l] = hy1;
        xo[l] = x1;
        yo[l] = y1;
        level_out(l, cl, e1);
        if (l == NS - 1) {
          const bool ol = own && cl >= jb && cl < je;
          const bool pm = pml_cell(cl);
          const unsigned k2 = (unsigned)(cl * LD + ic);
          bstore(e1, rE2, k2, ol);
          bstore(hx1, rHx2, k2, ol && cl < NY);
          bstore(hy1, rHy2, k2, ol && h_any);
          bstore(x1, rX2, k2, ol && pm);
          bstore(y1, rY2, k2, ol && pm);
        }
      }
      e0_c = en_[u];
      shift(eo, hxo, hyo, xo, yo);
      if (has_rec) {
#pragma unroll
        for (int l = NS - 1; l >= 1; --l) rid_l[l] = rid_l[l - 1];
      }
#pragma unroll
      for (int l = NS - 1; l >= 1; --l) {
        ka_l[l] = l == 1 ? ka_[u] : ka_l[l - 1];
        kb_l[l] = l == 1 ? kb_[u] : kb_l[l - 1];
        kc_l[l] = l == 1 ? kc_[u] : kc_l[l - 1];
      }
    }
    // coefficients of the later levels' next columns, in the interior chunks' form
#pragma unroll
    for (int l = 1; l < NS; ++l) {
      cc_l[l] = make_float2(ka_l[l].x, ka_l[l].z);
      cxe_l[l] = ka_l[l].y;
      cxh_l[l] = ka_l[l].z;
      cyh_l[l] = kb_l[l].x;
      hay_l[l] = kb_l[l].w;
      hby_l[l] = kc_l[l];
    }
  };

  // Loads of one interior chunk (all rows inner).  The interior run is
  // software-pipelined: the next chunk's loads are issued before the current
  // chunk is computed, so a wave always has a chunk of requests in flight.
  // Every load is issued unconditionally (past the end of the run it reads
  // out of bounds, which returns 0 without a memory access) so the compiler
  // can count outstanding requests exactly.
  struct InnerBuf {
    float en[UI], hx[UI], hy[UI];
    float2 c[UI];
  };
  auto inner_load = [&](int jj, InnerBuf& b, bool on) {
#pragma unroll
    for (int u = 0; u < UI; ++u) {
      const unsigned kk = (unsigned)((jj + u) * LD + i);
      b.en[u] = bload(rE0, kk + (unsigned)LD, on);
      b.hx[u] = bload(rHx0, kk, on);
      b.hy[u] = bload(rHy0, kk, on);
      b.c[u] = bload2k(rK, ko.cc, 2 * kk, on);
    }
  };
  // ---- interior chunk, all rows inner: scalar H coefficients, (Ceze, Cezhy) loads.
  // SRC: the wave's rows hold the source row (injection and source recorder);
  // the other waves carry neither.
  auto inner_compute = [&](int jj, const InnerBuf& b, auto SRC) {
    constexpr bool src = decltype(SRC)::value;
#pragma unroll
    for (int u = 0; u < UI; ++u) {
      const int c = jj + u;
      float eo[NS], hxo[NS], hyo[NS], zo[NS];
      if (src) rid_l[0] = rec_lookup(c);
      // per-level stores (frames, receivers, source recorder) test the level mask
      const unsigned lv = (FR || src) ? level_mask(c) : 0u;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const int cl = c - l;
        const float ec = l == 0 ? e0_c : e_in[l], en = l == 0 ? b.en[u] : eo[l - 1];
        const float hx = l == 0 ? b.hx[u] : hx_in[l], hy = l == 0 ? b.hy[u] : hy_in[l];
        const float2 cf = l == 0 ? b.c[u] : cc_l[l];
        const float e_dn = shfl_dn1(ec);
        const float hy1 = hy + by0 * (e_dn - ec);
        const float hx1 = hx + bx0 * (en - ec);
        const float hyu = shfl_up1(hy1);
        float e1 = cf.x * ec + cf.y * ((hy1 - hyu) - (hx1 - hxp[l]));
        if (src) e1 += (cl == sj && i == si) ? ws[l] : 0.f;
        eo[l] = e1;
        hxo[l] = hx1;
        hyo[l] = hy1;
        zo[l] = 0.f;  // no split-field cells in these rows
        const bool on = (lv >> l) & 1u;
        if (src) {
          bstore(e1, rJZ, (unsigned)(it + l), own && on && i == si && cl == sj);
          bstore(e1, rTR, (unsigned)((rid_l[l] - t_lo) * s.nt + it + l), own && on && rid_l[l] >= 0);
        }
        // interior columns are columns of the stored domain (see level_out_in)
        // level l's frame: plane l, column c-l; the per-level part l (FD - LD)
        // rides in the scalar offset, the lanes share one offset per column
        if (FR == 1)
          bstorek(e1, rFR, (unsigned)(l * (FD - LD)) * 4u, (unsigned)((c - g.dj0) * LD + i),
                  own_dom && on);
        if (FR == 2) frame_pair(l, e1, (unsigned)((c - g.dj0) * LD + i), own_dom && on);
        if (l == NS - 1) {
          const bool ol = own && cl >= jb && cl < je;
          const unsigned k2 = (unsigned)(cl * LD + i);
          bstore(e1, rE2, k2, ol);
          bstore(hx1, rHx2, k2, ol);
          bstore(hy1, rHy2, k2, ol);
        }
      }
      e0_c = b.en[u];
      shift(eo, hxo, hyo, zo, zo);
#pragma unroll
      for (int l = NS - 1; l >= 1; --l) cc_l[l] = l == 1 ? b.c[u] : cc_l[l - 1];
      if (src) {
#pragma unroll
        for (int l = NS - 1; l >= 1; --l) rid_l[l] = rid_l[l - 1];
      }
    }
  };

  // ---- interior chunk, rows of every class (top / bottom PML rows present)
  auto class_chunk = [&](int jj) {
    float en_[UC], hx_[UC], hy_[UC], x_[UC], y_[UC], cxe_[UC], cxh_[UC], cyh_[UC], hay_[UC], hby_[UC];
    float2 c_[UC];
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const unsigned kk = (unsigned)((jj + u) * LD + ic);
      en_[u] = bload(rE0, kk + (unsigned)LD, okrow);
      hx_[u] = bload(rHx0, kk, okrow);
      hy_[u] = bload(rHy0, kk, h_any);
      c_[u] = bload2k(rK, ko.cc, 2 * kk, e_inner);
      x_[u] = bload(rX0, kk, e_pml);
      y_[u] = bload(rY0, kk, e_pml);
      cxe_[u] = e_xp ? bloadk(rK, ko.cxe, kk, e_xp) : cxe_r;
      cxh_[u] = e_xp ? bloadk(rK, ko.cxh, kk, e_xp) : cxh_r;
      cyh_[u] = e_xp ? bloadk(rK, ko.cyh, kk, e_xp) : cyh_r;
      hay_[u] = h_xp ? bloadk(rK, ko.hay, kk, h_xp) : hay_row;
      hby_[u] = h_xp ? bloadk(rK, ko.hby, kk, h_xp) : hby_row;
    }
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const int c = jj + u;
      float eo[NS], hxo[NS], hyo[NS], xo[NS], yo[NS];
      if (has_rec) rid_l[0] = rec_lookup(c);
      const unsigned lv = level_mask(c);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const int cl = c - l;
        const float ec = l == 0 ? e0_c : e_in[l], en = l == 0 ? en_[u] : eo[l - 1];
        const float hx = l == 0 ? hx_[u] : hx_in[l], hy = l == 0 ? hy_[u] : hy_in[l];
        const float x = l == 0 ? x_[u] : x_in[l], y = l == 0 ? y_[u] : y_in[l];
        const float2 cf = l == 0 ? c_[u] : cc_l[l];
        const float cxe = l == 0 ? cxe_[u] : cxe_l[l], cxh = l == 0 ? cxh_[u] : cxh_l[l];
        const float cyh = l == 0 ? cyh_[u] : cyh_l[l];
        const float hay = l == 0 ? hay_[u] : hay_l[l], hby = l == 0 ? hby_[u] : hby_l[l];
        const float e_dn = shfl_dn1(ec);
        const float hy1 = h_any ? hay * hy + hby * (e_dn - ec) : 0.f;
        const float hx1 = okrow ? hx + bx0 * (en - ec) : 0.f;
        const float hyu = shfl_up1(hy1);
        const float dhy = hy1 - hyu, dhx = hx1 - hxp[l];
        const float x1 = cxe * x + cxh * dhy;
        const float y1 = y + cyh * dhx;
        // ring rows (0, NX) are never updated: they keep their value
        float e1 = e_inner ? cf.x * ec + cf.y * (dhy - dhx) : (e_pml ? x1 + y1 : ec);
        e1 += (cl == sj && i == si) ? ws[l] : 0.f;
        eo[l] = e1;
        hxo[l] = hx1;
        hyo[l] = hy1;
        xo[l] = x1;
        yo[l] = y1;
        level_out_in(l, cl, e1, lv);
        if (l == NS - 1) {
          const bool ol = own && ((lv >> l) & 1u);
          const unsigned k2 = (unsigned)(cl * LD + ic);
          bstore(e1, rE2, k2, ol);
          bstore(hx1, rHx2, k2, ol);
          bstore(hy1, rHy2, k2, ol && h_any);
          bstore(x1, rX2, k2, ol && e_pml);
          bstore(y1, rY2, k2, ol && e_pml);
        }
      }
      e0_c = en_[u];
      shift(eo, hxo, hyo, xo, yo);
      if (has_rec) {
#pragma unroll
        for (int l = NS - 1; l >= 1; --l) rid_l[l] = rid_l[l - 1];
      }
#pragma unroll
      for (int l = NS - 1; l >= 1; --l) {
        cc_l[l] = l == 1 ? c_[u] : cc_l[l - 1];
        cxe_l[l] = l == 1 ? cxe_[u] : cxe_l[l - 1];
        cxh_l[l] = l == 1 ? cxh_[u] : cxh_l[l - 1];
        cyh_l[l] = l == 1 ? cyh_[u] : cyh_l[l - 1];
        hay_l[l] = l == 1 ? hay_[u] : hay_l[l - 1];
        hby_l[l] = l == 1 ? hby_[u] : hby_l[l - 1];
      }
    }
  };

  // Level 0 at columns j .. j+U-1 per chunk, from jb-NS+1 to je+NS-2 (the
  // last level's last column is je-1).  Chunk starts split into a
  // strip/edge prefix, an interior run and a suffix.
  // The two wave kinds get separate copies of the loop, so that each keeps
  // only its own coefficient carries live.
  const int jlast = je + NS - 2;
  // Interior tiles (all rows inner, no receivers, no source row, every column
  // of the march in interior chunks) run a loop of their own and return: in
  // it only the interior chunk's descriptors and coefficients are live.
  if (inner_rows && !has_rec && !(si >= r0 && si <= r0 + 63) && chunkable(j1, UI) &&
      (jlast - j1 + 1) % UI == 0 && chunkable(jlast - UI + 1, UI)) {
    int j = j1;
    InnerBuf A, B;
    inner_load(j, A, true);
    while (j <= jlast) {
      inner_load(j + UI, B, j + UI <= jlast);
      inner_compute(j, A, std::false_type{});
      j += UI;
      if (j > jlast) break;
      inner_load(j + UI, A, j + UI <= jlast);
      inner_compute(j, B, std::false_type{});
      j += UI;
    }
    return;
  }
  if (inner_rows && (!has_rec || !s.rec_multi)) {
    int j = j1;
    for (; j <= jlast && !chunkable(j, UI); j += UC) univ_chunk(j);
    int jz = j;  // end of the interior run
    while (jz <= jlast && chunkable(jz, UI)) jz += UI;
    auto run = [&](auto SRC) {
      InnerBuf A, B;
      inner_load(j, A, j < jz);
      while (j < jz) {
        inner_load(j + UI, B, j + UI < jz);
        inner_compute(j, A, SRC);
        j += UI;
        if (j >= jz) break;
        inner_load(j + UI, A, j + UI < jz);
        inner_compute(j, B, SRC);
        j += UI;
      }
    };
    if (has_rec || (si >= r0 && si <= r0 + 63))
      run(std::true_type{});
    else
      run(std::false_type{});
    for (; j <= jlast; j += UC) univ_chunk(j);
  } else {
    int j = j1;
    for (; j <= jlast && !chunkable(j, UC); j += UC) univ_chunk(j);
    for (; j <= jlast && chunkable(j, UC); j += UC) class_chunk(j);
    for (; j <= jlast; j += UC) univ_chunk(j);
  }
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_FWD_WAVES)
k_fwd_step2(DevGrid g, DevShots s, int L, const float* __restrict__ E0,
            const float* __restrict__ Hx0, const float* __restrict__ Hy0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ E2,
            float* __restrict__ Hx2, float* __restrict__ Hy2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, float* __restrict__ tr, float* __restrict__ frame1,
            long long fstride) {
  fwdN_body<2, FR>(g, s, L, vtile_of_block(g.NX, g.NY, L, 60, true), E0, Hx0, Hy0, X0, Y0, E2, Hx2,
                       Hy2, X2, Y2, it, tr, frame1, fstride);
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_FWD_WAVES)
k_fwd_step4(DevGrid g, DevShots s, int L, const float* __restrict__ E0,
            const float* __restrict__ Hx0, const float* __restrict__ Hy0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ E2,
            float* __restrict__ Hx2, float* __restrict__ Hy2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, float* __restrict__ tr, float* __restrict__ frame1,
            long long fstride) {
  fwdN_body<4, FR>(g, s, L, vtile_of_block(g.NX, g.NY, L, 56, true), E0, Hx0, Hy0, X0, Y0, E2, Hx2,
                       Hy2, X2, Y2, it, tr, frame1, fstride);
}

// Interior row tiles of the 8-step forward with two rows per lane (FR = 0,
// or 2: BF16 frame pairs).  Lane k holds rows r0 + 2k (a) and r0 + 2k + 1
// (b): one 64-bit load or store moves a field's two rows, and one DPP shift
// per level serves both rows' vertical neighbours (row a's Hy above is lane
// k-1's row b, row b's Ez below is lane k+1's row a).  The wave covers 128
// rows with the same 2 x 8 rows of trapezoid halo as a one-row wave's 64, so
// it owns 112 rows instead of 48.  The host gives these tiles only inner
// rows of the stored domain, no receiver or source row within their 128 rows,
// and strips whose march runs through interior columns only (RowPairs); the
// update is the interior chunk's of fwdN_body, row by row.  ODD: the tile
// owns an odd number of rows, so one lane owns row a only and stores it
// alone.
template <int FR, bool ODD>
__device__ __forceinline__ void fwd8_pairs(const DevGrid& g, int shot, int strip, int q,
                                           const RowPairs& rp, int L,
                                           const float* __restrict__ E0,
                                           const float* __restrict__ Hx0,
                                           const float* __restrict__ Hy0, float* __restrict__ E2,
                                           float* __restrict__ Hx2, float* __restrict__ Hy2,
                                           float* __restrict__ frame1, long long fstride) {
  constexpr int NS = 8, R2 = 128 - 2 * NS;
  const int lane = threadIdx.x;
  const long long LD = g.LD;
  const int own_lo = __builtin_amdgcn_readfirstlane(rp.lo + q * R2);
  const int own_hi = __builtin_amdgcn_readfirstlane(min(rp.lo + q * R2 + R2 - 1, rp.hi));
  const int jb = __builtin_amdgcn_readfirstlane(strip * L);
  const int je = min(jb + L, g.NY + 1);
  const int ia = own_lo - NS + 2 * lane;  // row a; row b = ia + 1
  const bool own_a = ia >= own_lo && ia <= own_hi;
  const bool own_b = ia + 1 <= own_hi && ia + 1 >= own_lo;
  const bool own2 = own_a && own_b, own1 = ODD && own_a && !own_b;
  const long long so = (long long)shot * g.FS;
  const __amdgpu_buffer_rsrc_t rE0 = rsrc_of(const_cast<float*>(E0 + so), g.FS),
                               rHx0 = rsrc_of(const_cast<float*>(Hx0 + so), g.FS),
                               rHy0 = rsrc_of(const_cast<float*>(Hy0 + so), g.FS),
                               rE2 = rsrc_of(E2 + so, g.FS), rHx2 = rsrc_of(Hx2 + so, g.FS),
                               rHy2 = rsrc_of(Hy2 + so, g.FS),
                               rK = rsrc_of(const_cast<float*>(g.coef), 19 * g.FS);
  const CoefOff ko = coef_off(g);
  const float bx0 = g.bx0, by0 = g.by0;
  const long long FD = (long long)LD * g.n;
  // frames as in fwdN_body (FR = 2): the descriptor starts NS * LD words
  // early, plane l / 2 and the column lag of level l ride in the scalar offset
  const __amdgpu_buffer_rsrc_t rFR =
      rsrc_of(FR == 2 ? frame1 + (long long)shot * fstride - NS * LD : nullptr,
              FR == 2 ? NS / 2 * FD + NS * LD : 0);
  auto fsoff = [&](int l) { return (int)((unsigned)(l / 2 * FD + (NS - l) * LD) * 4u); };
  auto level_mask = [&](int c) {
    const int lo = max(0, c - je + 1), hi = min(NS - 1, c - jb);
    return hi >= lo ? (2u << hi) - (1u << lo) : 0u;
  };

  // state of fwdN_body's interior chunks, for rows (a, b): level l >= 1's
  // inputs of its previous column, hxp[l] its Hx of that column, cc_l[l] the
  // (Ceze, Cezhy) of rows a and b of its column; fh: even levels' values held
  // for the frame pairs
  const int j1 = jb - (NS - 1);
  float2 e0_c, e_in[NS], hx_in[NS], hy_in[NS], hxp[NS], fh[NS / 2];
  float4 cc_l[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    e_in[l] = hx_in[l] = hy_in[l] = hxp[l] = make_float2(0.f, 0.f);
    cc_l[l] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int m = 0; m < NS / 2; ++m) fh[m] = make_float2(0.f, 0.f);
  {
    const unsigned ka = (unsigned)(j1 * LD + ia), kb = ka - (unsigned)LD;
    e0_c = bload2(rE0, ka, true);
    const float2 eb = bload2(rE0, kb, true), hxb = bload2(rHx0, kb, true);
    // (cyh, AX, BX, AY) of column j1-1
    const float4 ua = bload4k(rK, ko.u1, 4 * kb, true), ub = bload4k(rK, ko.u1, 4 * (kb + 1), true);
    hxp[0].x = ua.y * hxb.x + ua.z * (e0_c.x - eb.x);
    hxp[0].y = ub.y * hxb.y + ub.z * (e0_c.y - eb.y);
  }
  struct Buf {
    float2 en, hx, hy;
    float4 c;  // (Ceze, Cezhy) of row a, then of row b
  };
  auto load = [&](int c, Buf& b, bool on) {
    const unsigned kk = (unsigned)(c * LD + ia);
    b.en = bload2(rE0, kk + (unsigned)LD, on);
    b.hx = bload2(rHx0, kk, on);
    b.hy = bload2(rHy0, kk, on);
    b.c = bload4k(rK, ko.cc, 2 * kk, on);
  };
  auto compute = [&](int c, const Buf& b) {
    const unsigned lv = FR ? level_mask(c) : 0u;
    const unsigned v0 = (unsigned)((c - g.dj0) * LD + ia);
    float2 eo[NS], hxo[NS], hyo[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const int cl = c - l;
      const float2 ec = l == 0 ? e0_c : e_in[l], en = l == 0 ? b.en : eo[l - 1];
      const float2 hx = l == 0 ? b.hx : hx_in[l], hy = l == 0 ? b.hy : hy_in[l];
      const float4 cf = l == 0 ? b.c : cc_l[l];
      const float e_dn = shfl_dn1(ec.x);  // Ez of row b + 1 (lane k+1's row a)
      float2 hy1, hx1, e1;
      hy1.x = hy.x + by0 * (ec.y - ec.x);
      hy1.y = hy.y + by0 * (e_dn - ec.y);
      hx1.x = hx.x + bx0 * (en.x - ec.x);
      hx1.y = hx.y + bx0 * (en.y - ec.y);
      const float hyu = shfl_up1(hy1.y);  // Hy of row a - 1 (lane k-1's row b)
      e1.x = cf.x * ec.x + cf.y * ((hy1.x - hyu) - (hx1.x - hxp[l].x));
      e1.y = cf.z * ec.y + cf.w * ((hy1.y - hy1.x) - (hx1.y - hxp[l].y));
      eo[l] = e1;
      hxo[l] = hx1;
      hyo[l] = hy1;
      if (FR == 2 && l % 2 == 1) {
        // the pair (level l-1, level l) of column c - l: level l-1 reached
        // it one column ago
        const bool on = (lv >> l) & 1u;
        u32x2 w;
        w.x = frame_pair_bits(fh[l / 2].x, e1.x);
        w.y = frame_pair_bits(fh[l / 2].y, e1.y);
        __builtin_amdgcn_raw_buffer_store_b64(w, rFR, (int)(own2 && on ? v0 * 4u : kOOB),
                                              fsoff(l), 0);
        if (ODD)
          __builtin_amdgcn_raw_buffer_store_b32(w.x, rFR, (int)(own1 && on ? v0 * 4u : kOOB),
                                                fsoff(l), 0);
        fh[l / 2] = eo[l - 1];
      }
      if (l == NS - 1) {
        const bool ol = cl >= jb && cl < je;
        const unsigned k2 = (unsigned)(cl * LD + ia);
        bstore2(e1, rE2, k2, own2 && ol);
        bstore2(hx1, rHx2, k2, own2 && ol);
        bstore2(hy1, rHy2, k2, own2 && ol);
        if (ODD) {
          bstore(e1.x, rE2, k2, own1 && ol);
          bstore(hx1.x, rHx2, k2, own1 && ol);
          bstore(hy1.x, rHy2, k2, own1 && ol);
        }
      }
    }
    e0_c = b.en;
#pragma unroll
    for (int l = NS - 1; l >= 1; --l) {
      e_in[l] = eo[l - 1];
      hx_in[l] = hxo[l - 1];
      hy_in[l] = hyo[l - 1];
      cc_l[l] = l == 1 ? b.c : cc_l[l - 1];
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) hxp[l] = hxo[l];
  };
  // level 0 at columns j1 .. je+NS-2, one column per chunk, the next
  // column's loads in flight while this one is computed
  const int jlast = je + NS - 2;
  int j = j1;
  Buf A, B;
  load(j, A, true);
  while (j <= jlast) {
    load(j + 1, B, j + 1 <= jlast);
    compute(j, A);
    if (++j > jlast) break;
    load(j + 1, A, j + 1 <= jlast);
    compute(j, B);
    ++j;
  }
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_FWD8_WAVES)
k_fwd_step8(DevGrid g, DevShots s, int L, const float* __restrict__ E0,
            const float* __restrict__ Hx0, const float* __restrict__ Hy0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ E2,
            float* __restrict__ Hx2, float* __restrict__ Hy2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, float* __restrict__ tr, float* __restrict__ frame1,
            long long fstride, RowPairs rp) {
  const TileId tid = vtile_of_block(g.NX, g.NY, L, 48, true);
  if (FR != 1 && rp.n > 0 && tid.strip >= rp.slo && tid.strip <= rp.shi) {
    const int t = tid.rowtile * 4 + tid.wy;
    if (t >= rp.t0 && t < rp.tb) {
      const int q = t - rp.t0;
      if (q < rp.n) {
        const int lo = rp.lo + q * 112, hi = min(lo + 111, rp.hi);
        if ((hi - lo + 1) & 1)
          fwd8_pairs<FR, true>(g, tid.shot, tid.strip, q, rp, L, E0, Hx0, Hy0, E2, Hx2, Hy2,
                               frame1, fstride);
        else
          fwd8_pairs<FR, false>(g, tid.shot, tid.strip, q, rp, L, E0, Hx0, Hy0, E2, Hx2, Hy2,
                                frame1, fstride);
      }
      return;
    }
  }
  fwdN_body<8, FR>(g, s, L, tid, E0, Hx0, Hy0, X0, Y0, E2, Hx2, Hy2, X2, Y2, it, tr, frame1,
                   fstride);
}

template <int NS, bool FRAMES>
__device__ __forceinline__ void fwd_multi_body(const MultiTable& tab, const MultiArgs& a) {
  int k;
  TileId tid;
  if (!multi_tile(a, tab, 64 - 2 * NS, k, tid, true, true)) return;
  const DevSurvey& S = tab.sv[a.sid[k]];
  const int c = a.cur[k], x = a.xc[k];
  float* f = S.fld;
  const long long z = S.fsz;
  float* fr1 = FRAMES ? S.frames + (long long)a.fslot[k] * S.FD : nullptr;
  fwdN_body<NS, FRAMES ? 1 : 0>(S.g, S.s, a.L[k], tid, f + 3 * c * z, f + (3 * c + 1) * z,
                        f + (3 * c + 2) * z, f + (6 + 2 * x) * z, f + (7 + 2 * x) * z,
                        f + 3 * (c ^ 1) * z, f + (3 * (c ^ 1) + 1) * z, f + (3 * (c ^ 1) + 2) * z,
                        f + (6 + 2 * (x ^ 1)) * z, f + (7 + 2 * (x ^ 1)) * z, a.it[k],
                        a.rec ? S.tr : nullptr, fr1, S.fstride);
}

template <int NS, bool FRAMES>
__global__ void __launch_bounds__(256, GJW_FWD_WAVES)
k_fwd_stepN_multi(const MultiTable tab, MultiArgs a) {
  fwd_multi_body<NS, FRAMES>(tab, a);
}

// eight steps over several surveys, with the 8-step forward's register budget
template <bool FRAMES>
__global__ void __launch_bounds__(256, GJW_FWD8_WAVES)
k_fwd_step8_multi(const MultiTable tab, MultiArgs a) {
  fwd_multi_body<8, FRAMES>(tab, a);
}
template __global__ void k_fwd_step8_multi<false>(const MultiTable, MultiArgs);
template __global__ void k_fwd_step8_multi<true>(const MultiTable, MultiArgs);

#define GJW_FWDN_ARGS DevGrid, DevShots, int, const float*, const float*, const float*, \
    const float*, const float*, float*, float*, float*, float*, float*, int, float*, float*, long long
template __global__ void k_fwd_step2<0>(GJW_FWDN_ARGS);
template __global__ void k_fwd_step2<1>(GJW_FWDN_ARGS);
template __global__ void k_fwd_step2<2>(GJW_FWDN_ARGS);
template __global__ void k_fwd_step4<0>(GJW_FWDN_ARGS);
template __global__ void k_fwd_step4<1>(GJW_FWDN_ARGS);
template __global__ void k_fwd_step4<2>(GJW_FWDN_ARGS);
template __global__ void k_fwd_step8<0>(GJW_FWDN_ARGS, RowPairs);
template __global__ void k_fwd_step8<1>(GJW_FWDN_ARGS, RowPairs);
template __global__ void k_fwd_step8<2>(GJW_FWDN_ARGS, RowPairs);
#undef GJW_FWDN_ARGS
template __global__ void k_fwd_stepN_multi<2, false>(const MultiTable, MultiArgs);
template __global__ void k_fwd_stepN_multi<2, true>(const MultiTable, MultiArgs);
template __global__ void k_fwd_stepN_multi<4, false>(const MultiTable, MultiArgs);
template __global__ void k_fwd_stepN_multi<4, true>(const MultiTable, MultiArgs);

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
    const float c = g.cc[k].y;
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
__device__ __forceinline__ void adj1_body(const DevGrid& g, const DevShots& s, int L, TileId tid, const float* __restrict__ M0,
           const float* __restrict__ Nx0, const float* __restrict__ Ny0,
           const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M1,
           float* __restrict__ Nx1, float* __restrict__ Ny1, float* __restrict__ X1,
           float* __restrict__ Y1, int it, const float* __restrict__ adj,
           const float* __restrict__ F1, const float* __restrict__ F0, long long fstride,
           float* __restrict__ acc, long long astride, float* __restrict__ gw) {
  const int lane = threadIdx.x;
  const int NX = g.NX, NY = g.NY, P = g.P;
  const int shot = tid.shot;
  const int i0 = tid.rowtile * 64;
  // threadIdx.y is constant across a 64-lane wavefront: make the strip start
  // (and everything derived from it) scalar so loop control stays uniform
  const int jb = __builtin_amdgcn_readfirstlane((tid.strip * blockDim.y + threadIdx.y) * L);
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
  const float bx0 = g.bx0, by0 = g.by0;
  const int ai = i - g.di0;  // domain row of this lane
  const bool arow = ai >= 0 && ai < g.m;
  const bool has_rec = adj && it >= 1 && i0 <= s.rmax && i0 + 63 >= s.rmin;

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
  // per-column epilogue: residual injection, correlation, wavelet gradient, stores
  // pre: the correlation operands (v1, v0, t) were loaded by the caller when
  // pre_ok; otherwise they are loaded here.
  auto emit = [&](int j, float mn, float me, float nxv, float nyv, bool inner, bool pre_ok,
                  float v1, float v0, float2 t) {
    if (!ok) return;
    const long long k = (long long)j * LD + i;
    if (has_rec) {  // residual of step it-1 enters dJ/dEz(S_it)
      const int qa = cp[j], qb = cp[j + 1];
      for (int q = qa; q < qb; ++q)
        if (s.crow[q] == i) mn += adj[(long long)s.cid[q] * s.nt + (it - 1)];
    }
    if (ac && inner && arow) {
      const int b = j - g.dj0;
      if (b >= 0 && b < g.n) {
        const long long fk = (long long)b * LD + i;
        if (!pre_ok) {
          v1 = f1[fk];
          v0 = f0 ? f0[fk] : 0.f;
          t = ac[fk];
        }
        t.x += me * (v1 - v0);  // -> eps gradient
        t.y += me * (v1 + v0);  // -> sigma gradient
        ac[fk] = t;
      }
    }
    if (gw && i == si && j == sj) gw[(long long)shot * s.nt + it] = sc * me;
    M1[k] = mn;
    if (i < NX) Ny1[k] = nyv;
    if (j < NY) Nx1[k] = nxv;
  };
  int j = jb;
  while (j < je) {
    if (fastrows && j >= P + 2 && j + U - 1 <= NY - P - 2 && j + U <= je) {
      // interior chunk: W = (c M, -c M), A = 1, B = (by0, bx0) for every cell touched
      const long long k = (long long)j * LD + i;
      float mn_[U], nx_[U], ny_[U], m63[U], c63[U], nu0[U], mu0[U], cu0[U];
      float2 cn_[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long kk = k + u * LD;
        mn_[u] = M0[kk + LD];
        cn_[u] = g.cc[kk + LD];
        nx_[u] = Nx0[kk];
        ny_[u] = Ny0[kk];
      }
      if (lane == 63) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          m63[u] = M0[k + u * LD + 1];
          c63[u] = g.cc[k + u * LD + 1].y;
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          nu0[u] = Ny0[k + u * LD - 1];
          mu0[u] = M0[k + u * LD - 1];
          cu0[u] = g.cc[k + u * LD - 1].y;
        }
      }
      // correlation operands of the chunk's columns (all inside the domain)
      float v1_[U], v0_[U];
      float2 t_[U];
      const bool pre = ac && arow;
      if (pre) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const long long fk = (long long)(j + u - g.dj0) * LD + i;
          v1_[u] = f1[fk];
          v0_[u] = f0 ? f0[fk] : 0.f;
          t_[u] = ac[fk];
        }
      }
      float ce_c = g.cc[k].x;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float wy_n = cn_[u].y * mn_[u];
        const float wx_n = -wy_n;
        float wy_dn = shfl_dn1(wy_c);
        if (lane == 63) wy_dn = c63[u] * m63[u];
        const float nyv = ny_[u] + wy_c - wy_dn;
        const float nxv = nx_[u] + wx_c - wx_n;
        const float py = by0 * nyv, px = bx0 * nxv;
        float py_up = shfl_up1(py);
        if (lane == 0) py_up = by0 * (nu0[u] + cu0[u] * mu0[u] - wy_c);
        const float mn = ce_c * me_c + py_up - py + px_prev - px;
        emit(j + u, mn, me_c, nxv, nyv, true, pre, pre ? v1_[u] : 0.f, pre ? v0_[u] : 0.f,
             pre ? t_[u] : make_float2(0.f, 0.f));
        me_c = mn_[u];
        wy_c = wy_n;
        wx_c = wx_n;
        ce_c = cn_[u].x;
        px_prev = px;
      }
      j += U;
    } else {
      const long long k = (long long)j * LD + i;
      float me_n = 0.f, wy_n = 0.f, wx_n = 0.f;
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
      float nyv = 0.f, nxv = 0.f, py = 0.f, px = 0.f;
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
      float py_up = shfl_up1(py);
      if (lane == 0) {
        py_up = 0.f;
        if (ok && i >= 1) {
          float wyu, wxu, a, b;
          w_of(g, i - 1, j, k - 1, M0[k - 1], X0, Y0, wyu, wxu);
          hy_coef(g, i - 1, j, a, b);
          py_up = b * (a * Ny0[k - 1] + wyu - wy_c);
        }
      }
      float mn = 0.f;
      bool inner = false;
      if (ok) {
        const int kd = ekind(NX, NY, P, i, j);
        inner = kd == KIND_INNER;
        const float base = inner ? g.cc[k].x * me_c : (kd == KIND_PML ? 0.f : me_c);
        mn = base + py_up - py + px_prev - px;
        if (kd == KIND_PML) {
          X1[k] = g.cxe[k] * (X0[k] + me_c);
          Y1[k] = g.cye[k] * (Y0[k] + me_c);
        }
      }
      emit(j, mn, me_c, nxv, nyv, inner, false, 0.f, 0.f, make_float2(0.f, 0.f));
      me_c = me_n;
      wy_c = wy_n;
      wx_c = wx_n;
      px_prev = px;
      ++j;
    }
  }
}

__global__ void __launch_bounds__(256)
k_adj_step(DevGrid g, DevShots s, int L, const float* __restrict__ M0,
           const float* __restrict__ Nx0, const float* __restrict__ Ny0,
           const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M1,
           float* __restrict__ Nx1, float* __restrict__ Ny1, float* __restrict__ X1,
           float* __restrict__ Y1, int it, const float* __restrict__ adj,
           const float* __restrict__ F1, const float* __restrict__ F0, long long fstride,
           float* __restrict__ acc, long long astride, float* __restrict__ gw) {
  adj1_body(g, s, L, tile_of_block(g.NX, g.NY, L), M0, Nx0, Ny0, X0, Y0, M1, Nx1, Ny1, X1, Y1, it, adj, F1, F0, fstride, acc, astride, gw);
}

__global__ void __launch_bounds__(256)
k_adj_step_multi(const MultiTable tab, MultiArgs a) {
  int k;
  TileId tid;
  if (!multi_tile(a, tab, 64, k, tid)) return;
  const DevSurvey& S = tab.sv[a.sid[k]];
  const int c = a.cur[k], d = c ^ 1;
  float* m = S.adjf;
  const long long z = S.fsz;
  const float* F0 = S.frames + (long long)a.fslot[k] * S.FD;
  adj1_body(S.g, S.s, a.L[k], tid, m + 5 * c * z, m + (5 * c + 1) * z, m + (5 * c + 2) * z,
            m + (5 * c + 3) * z, m + (5 * c + 4) * z, m + 5 * d * z, m + (5 * d + 1) * z,
            m + (5 * d + 2) * z, m + (5 * d + 3) * z, m + (5 * d + 4) * z, a.it[k], S.adj,
            F0 + S.FD, F0, S.fstride, S.acc, S.FD, nullptr);
}

// ==========================================================================
// adjoint: two fused steps per pass (temporal blocking)
// ==========================================================================
// One launch transposes forward steps it and it-1 for every shot, with the
// tiling of k_fwd_step2: 64-lane waves over rows r0 = 60*t - 2 .. r0+63, lanes
// 2..61 own their rows, step 2 one column behind step 1.  The correlation
// partials of a cell are read and written once per pass:
//   acc.x += me1 (F[it+1] - F[it]) + me2 (F[it] - F[it-1])
//   acc.y += me1 (F[it+1] + F[it]) + me2 (F[it] + F[it-1])
// with me1 = M(S_{it+1}) (input) and me2 = M(S_it) (after step 1 and its
// residual injection).  Requires it >= 1.


// NS fused adjoint steps per pass (NS = 2 or 4), the transpose of NS
// forward steps it, it-1, .., it-NS+1 (it >= NS-1), with the tiling of the
// forward: rows r0 = (64 - 2 NS) t - NS .. r0+63, lanes NS .. 63-NS own their
// rows, level l (forward step it-l) runs l columns behind level 0.  M0 holds
// dJ/dEz(S_{it+1}) with the residual of step it added; M2 receives
// dJ/dEz(S_{it-NS+1}) with the residual of step it-NS added (when >= 0).
// Fa: the frame Ez(S_{it+1}); the frames of S_{it}, .., S_{it-NS+1} lie at
// Fa - FD, .., Fa - NS FD (FD = LD * n).  The correlation of a cell is
// accumulated once per pass, at the last level's column, from each level's
// input M (kept for the levels' column lag) and the NS+1 frames:
//   acc.x += sum_l M_l (F_l - F_{l+1}),  acc.y += sum_l M_l (F_l + F_{l+1}).
// FR = 2: BF16 frames.  Then acc.x is summed by parts,
//   acc.x += M_0 F_0 + sum_{l>=1} (M_l - M_{l-1}) F_l - M_{NS-1} F_NS,
// the same sum in exact arithmetic: the difference is taken of M (FP32, in
// registers) instead of the rounded frames, whose consecutive values differ
// by a few percent (their difference would carry the frames' 2^-9 rounding
// amplified by that factor).  A pass's last frame is the next pass's first,
// the same stored value, so the boundary terms cancel across passes.
template <int NS, int FR>
__device__ __forceinline__ void adjN_body(const DevGrid& g, const DevShots& s, int L, TileId tid,
                                          const float* __restrict__ M0,
                                          const float* __restrict__ Nx0,
                                          const float* __restrict__ Ny0,
                                          const float* __restrict__ X0,
                                          const float* __restrict__ Y0, float* __restrict__ M2,
                                          float* __restrict__ Nx2, float* __restrict__ Ny2,
                                          float* __restrict__ X2, float* __restrict__ Y2, int it,
                                          const float* __restrict__ adj,
                                          const float* __restrict__ Fa, long long fstride,
                                          float* __restrict__ acc, long long astride,
                                          float* __restrict__ gw) {
  constexpr int R = 64 - 2 * NS;
  constexpr int UA = NS >= 6 ? 1 : NS == 4 ? 2 : 4;  // columns per chunk
  const int lane = threadIdx.x;
  const int NX = g.NX, NY = g.NY, P = g.P;
  const long long LD = g.LD;
  const int shot = tid.shot;
  const RowTile rt = row_tile((int)(tid.rowtile * 4 + tid.wy), NX, R, NS);
  const int r0 = __builtin_amdgcn_readfirstlane(rt.r0);
  const int own_hi = __builtin_amdgcn_readfirstlane(rt.own_hi);
  const int jb = __builtin_amdgcn_readfirstlane(tid.strip * L);
  if (r0 + NS > NX) return;  // no owned row
  const int je = min(jb + L, NY + 1);
  const int i = r0 + lane;
  const bool okrow = i >= 0 && i <= NX;
  const bool own = okrow && lane >= NS && lane <= 63 - NS && i <= own_hi;
  const long long so = (long long)shot * g.FS;
  M0 += so; Nx0 += so; Ny0 += so; X0 += so; Y0 += so;
  M2 += so; Nx2 += so; Ny2 += so; X2 += so; Y2 += so;
  const int si = s.si[shot], sj = s.sj[shot];
  const float sc = s.sc[shot];
  const int* cp = s.cptr + (long long)shot * (NY + 2);
  const long long FD = (long long)LD * g.n;
  float* ac = acc ? acc + 2 * (long long)shot * astride : nullptr;
  // residual injection reaches halo lanes too: later levels read their values
  const bool has_rec = adj && r0 <= s.rmax && r0 + 63 >= s.rmin;
  // the wavelet gradient is recorded only when it is asked for, by the waves
  // whose rows hold the source row; the other waves issue no stores for it
  const bool src_rows = gw && si >= r0 && si <= r0 + 63;
  // the rows the owned rows depend on are interior rows (see fwdN_body)
  const bool inner_rows = (r0 + 1 >= P + 1) && (own_hi + NS - 1 <= NX - P - 1);
  const bool dom_row = i - g.di0 >= 0 && i - g.di0 < g.m;
  auto dom_col = [&](int c) { return c - g.dj0 >= 0 && c - g.dj0 < g.n; };
  const float bx0 = g.bx0, by0 = g.by0;
  const long long ic = okrow ? i : 0;

  const __amdgpu_buffer_rsrc_t rM0 = rsrc_of(const_cast<float*>(M0), g.FS),
                               rNx0 = rsrc_of(const_cast<float*>(Nx0), g.FS),
                               rNy0 = rsrc_of(const_cast<float*>(Ny0), g.FS),
                               rX0 = rsrc_of(const_cast<float*>(X0), g.FS),
                               rY0 = rsrc_of(const_cast<float*>(Y0), g.FS),
                               rM2 = rsrc_of(M2, g.FS), rNx2 = rsrc_of(Nx2, g.FS),
                               rNy2 = rsrc_of(Ny2, g.FS), rX2 = rsrc_of(X2, g.FS),
                               rY2 = rsrc_of(Y2, g.FS),
                               rK = rsrc_of(const_cast<float*>(g.coef), 19 * g.FS),
                               // frames of S_{it-NS+1} .. S_{it+1}: slot k at k * FD
                               // FR = 1: frames of S_{it-NS+1} .. S_{it+1}, slot k at k * FD;
                               // FR = 2: the NS/2 + 1 word planes that hold them
                               rFR = FR == 2 ? rsrc_of(const_cast<float*>(Fa - NS / 2 * FD + (long long)shot * fstride),
                                                       (NS / 2 + 1) * FD)
                                             : rsrc_of(const_cast<float*>(Fa - NS * FD + (long long)shot * fstride),
                                                       (NS + 1) * FD),
                               rAC = rsrc_of(ac, ac ? 2 * FD : 0),
                               rGW = rsrc_of(gw ? gw + (long long)shot * s.nt : nullptr,
                                             gw ? s.nt : 0);
  const CoefOff ko = coef_off(g);

  // row classes (valid in interior columns)
  const bool e_inner = okrow && i >= P + 1 && i <= NX - P - 1;
  const bool e_xn = okrow && i >= 1 && i <= P;
  const bool e_xp = okrow && i >= NX - P && i <= NX - 1;
  const bool e_pml = e_xn || e_xp;
  const bool h_inner = okrow && i >= P && i <= NX - P - 1;
  const bool h_xn = okrow && i >= 0 && i <= P - 1;
  const bool h_xp = okrow && i >= NX - P && i <= NX - 1;
  const bool h_any = h_inner || h_xn || h_xp;
  const bool notring = okrow && i >= 1 && i <= NX - 1;
  auto pml_cell = [&](int c) {
    return notring && c >= 1 && c <= NY - 1 && (e_pml || c <= P || c >= NY - P);
  };
  // (from the coefficient table at column P+1, loaded without a branch)
  const bool rows_on = okrow && !inner_rows;
  const unsigned kr = (unsigned)((P + 1) * LD + ic);
  const float4 kr0 = bload4k(rK, ko.u0, 4 * kr, rows_on), kr1 = bload4k(rK, ko.u1, 4 * kr, rows_on);
  const float kr2 = bloadk(rK, ko.u2, kr, rows_on);
  const float cxe_r = rows_on ? kr0.y : 1.f, cxh_r = rows_on ? kr0.z : 0.f;
  const float cyh_r = rows_on ? kr1.x : 0.f;
  // Hy coefficients of rows whose Hy update does not depend on the column
  const float hay_row = rows_on ? kr1.w : 1.f;
  const float hby_row = rows_on ? kr2 : by0;

  // data residual of forward step `step` at the receivers of column c
  auto inject = [&](int c, float& mn, int step) {
    if (has_rec && step >= 0 && c >= 0 && c <= NY) {
      const int qa = cp[c], qb = cp[c + 1];
      for (int q = qa; q < qb; ++q)
        if (s.crow[q] == i) mn += adj[(long long)s.cid[q] * s.nt + step];
    }
  };
  // the interior path's receivers: the trace of the lane's cell in column c
  // (-1 if none; one receiver per cell), read through the shot's residuals
  const int t_lo = has_rec ? s.rp[shot] : 0;
  const __amdgpu_buffer_rsrc_t rAD =
      rsrc_of(has_rec ? const_cast<float*>(adj) + (long long)t_lo * s.nt : nullptr,
              has_rec ? (long long)(s.rp[shot + 1] - t_lo) * s.nt : 0);
  auto rec_lookup = [&](int c) {
    int r = -1;
    if (has_rec && c >= 0 && c <= NY) {
      const int qa = cp[c], qb = cp[c + 1];
      for (int q = qa; q < qb; ++q)
        if (s.crow[q] == i) r = s.cid[q];
    }
    return r;
  };
  int rid_h[NS];  // rid_h[k]: trace of the lane's cell k columns before the chunk (k >= 1)
#pragma unroll
  for (int k = 0; k < NS; ++k) rid_h[k] = -1;

  // ---- state.  Level l works on column c-l (c: level 0's column): its
  // input M, W = (wy, wx), split aux (x, y) and nu = (ny, nx) of that column
  // (level 0: M, W, x, y from the previous column's look-ahead, nu loaded;
  // level l >= 1: level l-1's outputs of the previous column), px of its
  // previous column, and the coefficients of its column:
  //   ce, cy (inner), cxe, cxh, cyh (row classes), AY/BY (hay, hby), or the
  //   table entries ka, kb, kc (strip and edge columns).
  // mh[l][k]: level l's input M k+1 columns ago, for the correlation at the
  // last level's column.
  const int j1 = jb - (NS - 1);
  float me_in[NS], wy_in[NS], wx_in[NS], x_in[NS], y_in[NS], ny_in[NS], nx_in[NS], pxp[NS];
  float ce_l[NS], cy_l[NS], cxe_l[NS], cxh_l[NS], cyh_l[NS], hay_l[NS], hby_l[NS];
  float mh[NS][NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    me_in[l] = wy_in[l] = wx_in[l] = x_in[l] = y_in[l] = ny_in[l] = nx_in[l] = pxp[l] = 0.f;
    ce_l[l] = cy_l[l] = 0.f;
    cxe_l[l] = 1.f;
    cxh_l[l] = cyh_l[l] = 0.f;
    hay_l[l] = 1.f;
    hby_l[l] = 0.f;
#pragma unroll
    for (int k = 0; k < NS; ++k) mh[l][k] = 0.f;
  }
  {  // level 0 at column j1, px of level 0 at j1-1 (branch-free loads)
    const unsigned ka = (unsigned)(j1 * LD + ic), kb = ka - (unsigned)LD;
    const bool ina = okrow && j1 >= 0 && j1 <= NY, inb = okrow && j1 - 1 >= 0 && j1 - 1 <= NY;
    const bool pa = pml_cell(j1), pb = pml_cell(j1 - 1);
    me_in[0] = bload(rM0, ka, ina);
    x_in[0] = bload(rX0, ka, pa);
    y_in[0] = bload(rY0, ka, pa);
    const float4 a0 = bload4k(rK, ko.u0, 4 * ka, ina), a1 = bload4k(rK, ko.u1, 4 * ka, ina);
    const float mb = bload(rM0, kb, inb), yb = bload(rY0, kb, pb);
    const float nxb = bload(rNx0, kb, inb && j1 - 1 < NY);
    const float4 b1 = bload4k(rK, ko.u1, 4 * kb, inb);  // (cyh, AX, BX, AY) of column j1-1
    wy_in[0] = a0.z * (x_in[0] + me_in[0]);
    wx_in[0] = a1.x * (y_in[0] + me_in[0]);
    pxp[0] = b1.z * (b1.y * nxb + b1.x * (yb + mb) - wx_in[0]);
    ce_l[0] = a0.x;
    cy_l[0] = a0.z;
    cxe_l[0] = a0.y;
    cxh_l[0] = a0.z;
    cyh_l[0] = a1.x;
  }

  // every level's column and level 0's look-ahead column are interior
  auto chunkable = [&](int jj) { return jj - (NS - 1) >= P + 1 && jj + UA <= NY - P - 1; };
  // the last level's column cz (owned, in the domain) correlates
  auto corr_on = [&](int cz) { return own && dom_row && cz >= jb && cz < je && dom_col(cz); };

  // Hand level l's outputs to level l+1, keep the M history, store the last
  // level's fields, correlate.  mo, wyo, wxo, xo, yo, nyo, nxo, pxo: outputs
  // of each level this column; f[k]: frames S_{it+1-k} at column cz.
  // SPLIT: the chunk may hold split-field cells (else X, Y are not stored);
  // SRC: the wave's rows may hold the source (recorder of the source gradient).
  auto finish_column = [&](int c, const float* mo, const float* wyo, const float* wxo,
                           const float* xo, const float* yo, const float* nyo, const float* nxo,
                           const float* pxo, const float* f, float2 t, auto SPLIT, auto SRC) {
    const int cz = c - (NS - 1);
    const bool oz = own && cz >= jb && cz < je;
    const unsigned kz = (unsigned)(cz * LD + ic);
    bstore(mo[NS - 1], rM2, kz, oz);
    bstore(nxo[NS - 1], rNx2, kz, oz && cz < NY);
    bstore(nyo[NS - 1], rNy2, kz, oz && h_any);
    if (decltype(SPLIT)::value) {
      const bool pz = pml_cell(cz);
      bstore(xo[NS - 1], rX2, kz, oz && pz);
      bstore(yo[NS - 1], rY2, kz, oz && pz);
    }
    // each level's input M at column cz
    float mz[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) mz[l] = l == NS - 1 ? me_in[l] : mh[l][NS - 2 - l];
    const bool inner_z = e_inner && cz >= P + 1 && cz <= NY - P - 1;
    if constexpr (FR == 2) {
      t.x = t.x + mz[0] * f[0];
#pragma unroll
      for (int l = 1; l < NS; ++l) t.x = t.x + (mz[l] - mz[l - 1]) * f[l];
      t.x = t.x - mz[NS - 1] * f[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) t.y = t.y + mz[l] * (f[l] + f[l + 1]);
    } else {
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        t.x = t.x + mz[l] * (f[l] - f[l + 1]);
        t.y = t.y + mz[l] * (f[l] + f[l + 1]);
      }
    }
    bstore2(t, rAC, 2 * (unsigned)((cz - g.dj0) * LD + ic), ac && inner_z && corr_on(cz));
    if (decltype(SRC)::value && src_rows) {  // wave-uniform: a scalar branch
      const bool at_src = oz && i == si && cz == sj;
#pragma unroll
      for (int l = 0; l < NS; ++l) bstore(sc * mz[l], rGW, (unsigned)(it - l), at_src);
    }
    // histories, then level l+1 inputs from level l outputs
#pragma unroll
    for (int l = 0; l < NS - 1; ++l) {
#pragma unroll
      for (int k = NS - 2 - l; k >= 1; --k) mh[l][k] = mh[l][k - 1];
      mh[l][0] = me_in[l];
    }
#pragma unroll
    for (int l = NS - 1; l >= 1; --l) {
      me_in[l] = mo[l - 1];
      wy_in[l] = wyo[l - 1];
      wx_in[l] = wxo[l - 1];
      x_in[l] = xo[l - 1];
      y_in[l] = yo[l - 1];
      ny_in[l] = nyo[l - 1];
      nx_in[l] = nxo[l - 1];
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) pxp[l] = pxo[l];
    (void)c;
  };
  // correlation operands of column cz: frames S_{it+1} .. S_{it-NS+1}, partials
  // (the frames differ by whole planes: that part rides in the scalar offset,
  // so the NS+1 loads share one per-lane offset)
  auto load_corr = [&](int cz, float* f, float2& t, bool on) {
    const unsigned fk = (unsigned)((cz - g.dj0) * LD + ic);
    if constexpr (FR == 2) {
      // it - k0 is odd in every pass: S_{it+1-2m} is the high half of plane
      // m back from Fa, S_{it-2m} the low half
#pragma unroll
      for (int m = 0; m <= NS / 2; ++m) {
        const unsigned w =
            __float_as_uint(bloadk(rFR, (unsigned)((NS / 2 - m) * FD) * 4u, fk, on));
        f[2 * m] = __uint_as_float(w & 0xFFFF0000u);
        if (2 * m + 1 <= NS) f[2 * m + 1] = __uint_as_float(w << 16);
      }
    } else {
#pragma unroll
      for (int k = 0; k <= NS; ++k) f[k] = bloadk(rFR, (unsigned)((NS - k) * FD) * 4u, fk, on);
    }
    t = bload2(rAC, 2 * fk, on);
  };

  // ---- chunk of any columns (strips, grid edges, strip ends): the table
  auto univ_chunk = [&](int jj) {
    float mL[UA], xL[UA], yL[UA], kcL[UA], nx_[UA], ny_[UA], f_[UA][NS + 1];
    float4 kaL[UA], kbL[UA];
    float2 t_[UA];
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int c = jj + u, cn = c + 1;
      const unsigned kk = (unsigned)(c * LD + ic), kn = kk + (unsigned)LD;
      const bool nin = okrow && cn >= 0 && cn <= NY;
      mL[u] = bload(rM0, kn, nin);
      const bool pmn = pml_cell(cn);
      xL[u] = bload(rX0, kn, pmn);
      yL[u] = bload(rY0, kn, pmn);
      kaL[u] = bload4k(rK, ko.u0, 4 * kn, nin);
      kbL[u] = bload4k(rK, ko.u1, 4 * kn, nin);
      kcL[u] = bloadk(rK, ko.u2, kn, nin);
      nx_[u] = bload(rNx0, kk, okrow && c >= 0 && c < NY);
      ny_[u] = bload(rNy0, kk, okrow && c >= 0 && c <= NY && i < NX);
      const int cz = c - (NS - 1);
      load_corr(cz, f_[u], t_[u], ac && e_inner && cz >= P + 1 && cz <= NY - P - 1 && corr_on(cz));
    }
    // table entries of every level's first column (level 0: column jj)
    float4 ka_l[NS], kb_l[NS];
    float kc_l[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const int c = jj - l;
      const bool cin = okrow && c >= 0 && c <= NY;
      const unsigned kk = (unsigned)(c * LD + ic);
      ka_l[l] = bload4k(rK, ko.u0, 4 * kk, cin);
      kb_l[l] = bload4k(rK, ko.u1, 4 * kk, cin);
      kc_l[l] = bloadk(rK, ko.u2, kk, cin);
    }
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int c = jj + u;
      // W of level 0's look-ahead column c+1: (ce, cxe, cxh, cye) = ka, (cyh, AX, BX, AY) = kb
      const float wy0n = kaL[u].z * (xL[u] + mL[u]);
      const float wx0n = kbL[u].x * (yL[u] + mL[u]);
      float mo[NS], wyo[NS], wxo[NS], xo[NS], yo[NS], nyo[NS], nxo[NS], pxo[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const int cl = c - l;
        const float4 ka = ka_l[l], kb = kb_l[l];
        const float kc = kc_l[l];
        const float me = me_in[l], ny = l == 0 ? ny_[u] : ny_in[l], nx = l == 0 ? nx_[u] : nx_in[l];
        const float wxn = l == 0 ? wx0n : wxo[l - 1];
        const float nyv = kb.w * ny + wy_in[l] - shfl_dn1(wy_in[l]);
        const float nxv = kb.y * nx + wx_in[l] - wxn;
        const float py = kc * nyv, px = kb.z * nxv;
        float mn = ka.x * me + shfl_up1(py) - py + pxp[l] - px;
        const float x1 = ka.y * (x_in[l] + me);
        const float y1 = ka.w * (y_in[l] + me);
        inject(cl, mn, it - l - 1);
        mo[l] = mn;
        wyo[l] = ka.z * (x1 + mn);
        wxo[l] = kb.x * (y1 + mn);
        xo[l] = x1;
        yo[l] = y1;
        nyo[l] = nyv;
        nxo[l] = nxv;
        pxo[l] = px;
      }
      finish_column(c, mo, wyo, wxo, xo, yo, nyo, nxo, pxo, f_[u], t_[u], std::true_type{},
                    std::true_type{});
      me_in[0] = mL[u];
      wy_in[0] = wy0n;
      wx_in[0] = wx0n;
      x_in[0] = xL[u];
      y_in[0] = yL[u];
#pragma unroll
      for (int l = NS - 1; l >= 1; --l) {
        ka_l[l] = ka_l[l - 1];
        kb_l[l] = kb_l[l - 1];
        kc_l[l] = kc_l[l - 1];
      }
      ka_l[0] = kaL[u];
      kb_l[0] = kbL[u];
      kc_l[0] = kcL[u];
    }
    // coefficients of every level's next column, in the interior chunks' form
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      ce_l[l] = ka_l[l].x;
      cy_l[l] = ka_l[l].z;
      cxe_l[l] = ka_l[l].y;
      cxh_l[l] = ka_l[l].z;
      cyh_l[l] = kb_l[l].x;
      hay_l[l] = kb_l[l].w;
      hby_l[l] = kc_l[l];
    }
  };

  // ---- interior chunk, all rows inner, no receivers
  // REC: the wave's rows hold receivers (their residuals are injected)
  auto inner_chunk = [&](int jj, auto SRC, auto REC) {
    constexpr bool rec = decltype(REC)::value;
    float mL[UA], nx_[UA], ny_[UA], f_[UA][NS + 1];
    float2 cL[UA], t_[UA];
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const unsigned kk = (unsigned)((jj + u) * LD + i);
      mL[u] = bload(rM0, kk + (unsigned)LD, true);
      cL[u] = bload2k(rK, ko.cc, 2 * (kk + (unsigned)LD), true);
      nx_[u] = bload(rNx0, kk, true);
      ny_[u] = bload(rNy0, kk, true);
      const int cz = jj + u - (NS - 1);
      load_corr(cz, f_[u], t_[u], ac && corr_on(cz));
    }
    // residuals injected by level l at column jj+u-l, loaded with the chunk
    float ra[UA][NS];
    int rc[UA];
    if (rec) {
#pragma unroll
      for (int u = 0; u < UA; ++u) rc[u] = rec_lookup(jj + u);
#pragma unroll
      for (int u = 0; u < UA; ++u)
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const int r = u - l >= 0 ? rc[u - l] : rid_h[l - u];
          const int step = it - l - 1;
          ra[u][l] = bload(rAD, (unsigned)((r - t_lo) * s.nt + step), r >= 0 && step >= 0);
        }
    }
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int c = jj + u;
      const float wy0n = cL[u].y * mL[u];
      const float wx0n = -cL[u].y * mL[u];
      float mo[NS], wyo[NS], wxo[NS], zo[NS], nyo[NS], nxo[NS], pxo[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const float me = me_in[l], ny = l == 0 ? ny_[u] : ny_in[l], nx = l == 0 ? nx_[u] : nx_in[l];
        const float wxn = l == 0 ? wx0n : wxo[l - 1];
        const float nyv = ny + wy_in[l] - shfl_dn1(wy_in[l]);
        const float nxv = nx + wx_in[l] - wxn;
        const float py = by0 * nyv, px = bx0 * nxv;
        float mn = ce_l[l] * me + shfl_up1(py) - py + pxp[l] - px;
        if (rec) mn += ra[u][l];
        mo[l] = mn;
        wyo[l] = cy_l[l] * mn;
        wxo[l] = -cy_l[l] * mn;
        zo[l] = 0.f;  // no split-field cells in these rows
        nyo[l] = nyv;
        nxo[l] = nxv;
        pxo[l] = px;
      }
      finish_column(c, mo, wyo, wxo, zo, zo, nyo, nxo, pxo, f_[u], t_[u], std::false_type{}, SRC);
      me_in[0] = mL[u];
      wy_in[0] = wy0n;
      wx_in[0] = wx0n;
      x_in[0] = y_in[0] = 0.f;
#pragma unroll
      for (int l = NS - 1; l >= 1; --l) {
        ce_l[l] = ce_l[l - 1];
        cy_l[l] = cy_l[l - 1];
      }
      ce_l[0] = cL[u].x;
      cy_l[0] = cL[u].y;
    }
    if (rec) {  // receiver history for the next chunk
#pragma unroll
      for (int k = NS - 1; k >= 1; --k) rid_h[k] = k <= UA ? rc[UA - k] : rid_h[k - UA];
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cxe_l[l] = 1.f;
      cxh_l[l] = cyh_l[l] = 0.f;
      hay_l[l] = 1.f;
      hby_l[l] = by0;
    }
  };

  // ---- interior chunk, rows of every class (top / bottom PML rows, receivers)
  auto class_chunk = [&](int jj, auto SRC) {
    float mL[UA], xL[UA], yL[UA], cxeL[UA], cxhL[UA], cyhL[UA], nx_[UA], ny_[UA], hayA[UA],
        hbyA[UA], f_[UA][NS + 1];
    float2 cL[UA], t_[UA];
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const unsigned kk = (unsigned)((jj + u) * LD + ic), kn = kk + (unsigned)LD;
      mL[u] = bload(rM0, kn, okrow);
      cL[u] = bload2k(rK, ko.cc, 2 * kn, e_inner);
      xL[u] = bload(rX0, kn, e_pml);
      yL[u] = bload(rY0, kn, e_pml);
      cxeL[u] = e_xp ? bloadk(rK, ko.cxe, kn, e_xp) : cxe_r;
      cxhL[u] = e_xp ? bloadk(rK, ko.cxh, kn, e_xp) : cxh_r;
      cyhL[u] = e_xp ? bloadk(rK, ko.cyh, kn, e_xp) : cyh_r;
      nx_[u] = bload(rNx0, kk, okrow);
      ny_[u] = bload(rNy0, kk, h_any);
      hayA[u] = h_xp ? bloadk(rK, ko.hay, kk, h_xp) : hay_row;
      hbyA[u] = h_xp ? bloadk(rK, ko.hby, kk, h_xp) : hby_row;
      const int cz = jj + u - (NS - 1);
      load_corr(cz, f_[u], t_[u], ac && e_inner && corr_on(cz));
    }
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int c = jj + u;
      const float wy0n = e_inner ? cL[u].y * mL[u] : (e_pml ? cxhL[u] * (xL[u] + mL[u]) : 0.f);
      const float wx0n = e_inner ? -cL[u].y * mL[u] : (e_pml ? cyhL[u] * (yL[u] + mL[u]) : 0.f);
      float mo[NS], wyo[NS], wxo[NS], xo[NS], yo[NS], nyo[NS], nxo[NS], pxo[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const int cl = c - l;
        const float me = me_in[l], ny = l == 0 ? ny_[u] : ny_in[l], nx = l == 0 ? nx_[u] : nx_in[l];
        const float hay = l == 0 ? hayA[u] : hay_l[l], hby = l == 0 ? hbyA[u] : hby_l[l];
        const float wxn = l == 0 ? wx0n : wxo[l - 1];
        const float wdn = shfl_dn1(wy_in[l]);
        const float nyv = h_any ? hay * ny + wy_in[l] - wdn : 0.f;
        const float nxv = okrow ? nx + wx_in[l] - wxn : 0.f;
        const float py = hby * nyv, px = bx0 * nxv;
        const float pup = shfl_up1(py);
        const float base = e_inner ? ce_l[l] * me : (e_pml ? 0.f : me);
        float mn = base + pup - py + pxp[l] - px;
        const float x1 = e_pml ? cxe_l[l] * (x_in[l] + me) : 0.f;
        const float y1 = e_pml ? y_in[l] + me : 0.f;  // cye = 1 off the left/right strips
        inject(cl, mn, it - l - 1);
        mo[l] = mn;
        wyo[l] = e_inner ? cy_l[l] * mn : (e_pml ? cxh_l[l] * (x1 + mn) : 0.f);
        wxo[l] = e_inner ? -cy_l[l] * mn : (e_pml ? cyh_l[l] * (y1 + mn) : 0.f);
        xo[l] = x1;
        yo[l] = y1;
        nyo[l] = nyv;
        nxo[l] = nxv;
        pxo[l] = px;
      }
      finish_column(c, mo, wyo, wxo, xo, yo, nyo, nxo, pxo, f_[u], t_[u], std::true_type{}, SRC);
      me_in[0] = mL[u];
      wy_in[0] = wy0n;
      wx_in[0] = wx0n;
      x_in[0] = xL[u];
      y_in[0] = yL[u];
#pragma unroll
      for (int l = NS - 1; l >= 1; --l) {
        ce_l[l] = ce_l[l - 1];
        cy_l[l] = cy_l[l - 1];
        cxe_l[l] = cxe_l[l - 1];
        cxh_l[l] = cxh_l[l - 1];
        cyh_l[l] = cyh_l[l - 1];
        hay_l[l] = l == 1 ? hayA[u] : hay_l[l - 1];
        hby_l[l] = l == 1 ? hbyA[u] : hby_l[l - 1];
      }
      ce_l[0] = cL[u].x;
      cy_l[0] = cL[u].y;
      cxe_l[0] = cxeL[u];
      cxh_l[0] = cxhL[u];
      cyh_l[0] = cyhL[u];
    }
  };

  // level 0 at columns jb-NS+1 .. je+NS-2 (the last level's last column is je-1)
  const bool use_inner = inner_rows && (!has_rec || !s.rec_multi);
  const int jlast = je + NS - 2;
  // Interior tiles (all rows inner, no receivers, no source, every column of
  // the march interior) run a loop of their own and return: in it only the
  // interior chunk's descriptors and coefficients are live, so the scalar
  // registers of the strip and row-class paths do not crowd it.
  if (use_inner && !has_rec && !src_rows && chunkable(j1) &&
      chunkable(j1 + (jlast - j1) / UA * UA)) {
    // two chunks per trip: the level carries alternate between two sets of
    // registers instead of being moved at every column
    int jj = j1;
    for (; jj + UA <= jlast; jj += 2 * UA) {
      inner_chunk(jj, std::false_type{}, std::false_type{});
      inner_chunk(jj + UA, std::false_type{}, std::false_type{});
    }
    if (jj <= jlast) inner_chunk(jj, std::false_type{}, std::false_type{});
    return;
  }
  int j = j1;
  for (; j <= jlast && !chunkable(j); j += UA) univ_chunk(j);
  if (use_inner && has_rec) {
    // levels >= 1 start on columns j-1 .. j-NS+1, passed by the strip chunks
    if (j <= jlast && chunkable(j)) {
#pragma unroll
      for (int k = 1; k < NS; ++k) rid_h[k] = rec_lookup(j - k);
    }
    for (; j <= jlast && chunkable(j); j += UA) inner_chunk(j, std::true_type{}, std::true_type{});
  }
  for (; j <= jlast && chunkable(j); j += UA) {
    if (use_inner)
      inner_chunk(j, std::true_type{}, std::false_type{});
    else
      class_chunk(j, std::true_type{});
  }
  for (; j <= jlast; j += UA) univ_chunk(j);
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_ADJ_WAVES)
k_adj_step2(DevGrid g, DevShots s, int L, const float* __restrict__ M0,
            const float* __restrict__ Nx0, const float* __restrict__ Ny0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M2,
            float* __restrict__ Nx2, float* __restrict__ Ny2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, const float* __restrict__ adj,
            const float* __restrict__ Fa, long long fstride, float* __restrict__ acc,
            long long astride, float* __restrict__ gw) {
  adjN_body<2, FR>(g, s, L, vtile_of_block(g.NX, g.NY, L, 60, false), M0, Nx0, Ny0, X0, Y0, M2, Nx2, Ny2, X2,
               Y2, it, adj, Fa, fstride, acc, astride, gw);
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_ADJ_WAVES)
k_adj_step4(DevGrid g, DevShots s, int L, const float* __restrict__ M0,
            const float* __restrict__ Nx0, const float* __restrict__ Ny0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M2,
            float* __restrict__ Nx2, float* __restrict__ Ny2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, const float* __restrict__ adj,
            const float* __restrict__ Fa, long long fstride, float* __restrict__ acc,
            long long astride, float* __restrict__ gw) {
  adjN_body<4, FR>(g, s, L, vtile_of_block(g.NX, g.NY, L, 56, false), M0, Nx0, Ny0, X0, Y0, M2, Nx2, Ny2, X2,
               Y2, it, adj, Fa, fstride, acc, astride, gw);
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_ADJ6_WAVES)
k_adj_step6(DevGrid g, DevShots s, int L, const float* __restrict__ M0,
            const float* __restrict__ Nx0, const float* __restrict__ Ny0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M2,
            float* __restrict__ Nx2, float* __restrict__ Ny2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, const float* __restrict__ adj,
            const float* __restrict__ Fa, long long fstride, float* __restrict__ acc,
            long long astride, float* __restrict__ gw) {
  adjN_body<6, FR>(g, s, L, vtile_of_block(g.NX, g.NY, L, 52, false), M0, Nx0, Ny0, X0, Y0, M2, Nx2, Ny2, X2,
               Y2, it, adj, Fa, fstride, acc, astride, gw);
}

template <int FR>
__global__ void __launch_bounds__(256, GJW_ADJ8_WAVES)
k_adj_step8(DevGrid g, DevShots s, int L, const float* __restrict__ M0,
            const float* __restrict__ Nx0, const float* __restrict__ Ny0,
            const float* __restrict__ X0, const float* __restrict__ Y0, float* __restrict__ M2,
            float* __restrict__ Nx2, float* __restrict__ Ny2, float* __restrict__ X2,
            float* __restrict__ Y2, int it, const float* __restrict__ adj,
            const float* __restrict__ Fa, long long fstride, float* __restrict__ acc,
            long long astride, float* __restrict__ gw) {
  adjN_body<8, FR>(g, s, L, vtile_of_block(g.NX, g.NY, L, 48, false), M0, Nx0, Ny0, X0, Y0, M2, Nx2, Ny2, X2,
               Y2, it, adj, Fa, fstride, acc, astride, gw);
}

template <int NS>
__global__ void __launch_bounds__(256, GJW_ADJM_WAVES)
k_adj_stepN_multi(const MultiTable tab, MultiArgs a) {
  int k;
  TileId tid;
  if (!multi_tile(a, tab, 64 - 2 * NS, k, tid, true)) return;
  const DevSurvey& S = tab.sv[a.sid[k]];
  const int c = a.cur[k], d = c ^ 1;
  float* m = S.adjf;
  const long long z = S.fsz;
  const float* Fa = S.frames + (long long)(a.fslot[k] + 1) * S.FD;
  adjN_body<NS, 1>(S.g, S.s, a.L[k], tid, m + 5 * c * z, m + (5 * c + 1) * z, m + (5 * c + 2) * z,
                m + (5 * c + 3) * z, m + (5 * c + 4) * z, m + 5 * d * z, m + (5 * d + 1) * z,
                m + (5 * d + 2) * z, m + (5 * d + 3) * z, m + (5 * d + 4) * z, a.it[k], S.adj, Fa,
                S.fstride, S.acc, S.FD, nullptr);
}
#define GJW_ADJN_ARGS DevGrid, DevShots, int, const float*, const float*, const float*, \
    const float*, const float*, float*, float*, float*, float*, float*, int, const float*, \
    const float*, long long, float*, long long, float*
template __global__ void k_adj_step2<1>(GJW_ADJN_ARGS);
template __global__ void k_adj_step2<2>(GJW_ADJN_ARGS);
template __global__ void k_adj_step4<1>(GJW_ADJN_ARGS);
template __global__ void k_adj_step4<2>(GJW_ADJN_ARGS);
template __global__ void k_adj_step6<1>(GJW_ADJN_ARGS);
template __global__ void k_adj_step6<2>(GJW_ADJN_ARGS);
template __global__ void k_adj_step8<1>(GJW_ADJN_ARGS);
template __global__ void k_adj_step8<2>(GJW_ADJN_ARGS);
#undef GJW_ADJN_ARGS
template __global__ void k_adj_stepN_multi<2>(const MultiTable, MultiArgs);
template __global__ void k_adj_stepN_multi<4>(const MultiTable, MultiArgs);
template __global__ void k_adj_stepN_multi<6>(const MultiTable, MultiArgs);

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

// --------------------------------------------------------------------------
// 16-bit checkpoint snapshots with per-block scaling
// --------------------------------------------------------------------------
// A snapshot plane is one field of every shot: either the whole pitched field
// (idx == null, cnt = FS values per shot, scale blocks of B = LD values: one
// grid column) or its split-field cells (idx = PML cell list, cnt entries per
// shot, blocks of B = 256 list entries).  Each block is scaled by a power of
// two 2^sh chosen from its max |v| so that its largest value lands just below
// 2^15 (FP16) or in [1, 2) (BF16); power-of-two scaling is exact, so the only
// rounding is the round-to-nearest-even conversion.  One workgroup per block.
constexpr int kSnapMaxPer = 8;  // values per thread: B <= 256 * kSnapMaxPer

__device__ __forceinline__ int snap_shift(float a, int mode) {
  if (!(a > 0.f)) return 0;
  int e;
  frexpf(a, &e);  // a = f * 2^e, 0.5 <= f < 1
  return mode == 2 ? 15 - e : 1 - e;
}

__device__ __forceinline__ unsigned short to16(float v, int mode) {
  if (mode == 2) return __builtin_bit_cast(unsigned short, (_Float16)v);  // RNE
  unsigned u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);  // bf16 round to nearest even
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float from16(unsigned short q, int mode) {
  if (mode == 2) return (float)__builtin_bit_cast(_Float16, q);
  return __uint_as_float((unsigned)q << 16);
}

// grid (blocks per shot, shots), 256 threads; mode 1 = BF16, 2 = FP16;
// shifts[shot * nb + block] receives sh
__global__ void __launch_bounds__(256)
k_pack16(const float* __restrict__ src, long long FS, const int* __restrict__ idx, long long cnt,
         int B, int mode, int* __restrict__ shifts, unsigned short* __restrict__ dst) {
  const int shot = blockIdx.y, blk = blockIdx.x;
  const long long t0 = (long long)blk * B;
  const int nb = gridDim.x;
  const float* p = src + (long long)shot * FS;
  float v[kSnapMaxPer];
  float m = 0.f;
#pragma unroll
  for (int r = 0; r < kSnapMaxPer; ++r) {
    const int o = r * 256 + (int)threadIdx.x;
    const long long t = t0 + o;
    v[r] = (o < B && t < cnt) ? p[idx ? idx[t] : t] : 0.f;
    m = fmaxf(m, fabsf(v[r]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
  const int sh = snap_shift(m, mode);
  if (threadIdx.x == 0) shifts[(long long)shot * nb + blk] = sh;
  unsigned short* q = dst + (long long)shot * cnt;
#pragma unroll
  for (int r = 0; r < kSnapMaxPer; ++r) {
    const int o = r * 256 + (int)threadIdx.x;
    const long long t = t0 + o;
    if (o < B && t < cnt) q[t] = to16(ldexpf(v[r], sh), mode);
  }
}

__global__ void __launch_bounds__(256)
k_unpack16(const unsigned short* __restrict__ src, long long cnt, int B,
           const int* __restrict__ shifts, int mode, const int* __restrict__ idx, long long FS,
           float* __restrict__ dst) {
  const int shot = blockIdx.y, blk = blockIdx.x;
  const long long t0 = (long long)blk * B;
  const int sh = shifts[(long long)shot * gridDim.x + blk];
  const unsigned short* q = src + (long long)shot * cnt;
  float* p = dst + (long long)shot * FS;
  for (int o = threadIdx.x; o < B; o += 256) {
    const long long t = t0 + o;
    if (t < cnt) p[idx ? idx[t] : t] = ldexpf(from16(q[t], mode), -sh);
  }
}

// --------------------------------------------------------------------------
// Hilbert-based misfits (w_bozdag.m:13-55, instant_.m:14-20, w_E.m:7-16)
// --------------------------------------------------------------------------
// Traces are rows of nt samples.  MATLAB's hilbert(x) along time: FFT of
// length nt, weights h = [1, 2 .. 2, (1 if nt even), 0 ..], inverse FFT.  The
// forward real-to-complex FFT gives the nt/2+1 non-negative frequencies; this
// kernel writes the full weighted spectrum for the complex inverse FFT (the
// inverse is unnormalised: the analytic signal is the result / nt).
__global__ void k_analytic_spec(const float2* __restrict__ half, float2* __restrict__ full, int nt,
                                long long ntr) {
  const int nh = nt / 2 + 1;
  const long long n = ntr * nt;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long r = t / nt;
    const int k = (int)(t - r * nt);
    float2 v = make_float2(0.f, 0.f);
    if (k < nh) {
      const float2 x = half[r * nh + k];
      const float w = (k == 0 || (nt % 2 == 0 && k == nt / 2)) ? 1.f : 2.f;
      v = make_float2(w * x.x, w * x.y);
    }
    full[t] = v;
  }
}

// envelope (kind 1, 3, 4) or instantaneous phase (kind 2) of the observed
// traces: ov = |A| or angle(A), A = analytic signal (ana / nt)
__global__ void k_obs_instant(const float2* __restrict__ ana, long long n, int nt, int kind,
                             float* __restrict__ ov) {
  const float inv = 1.f / (float)nt;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const float re = ana[t].x * inv, im = ana[t].y * inv;
    ov[t] = kind == 2 ? atan2f(im, re) : sqrtf(re * re + im * im);
  }
}

// Stage 1 per sample of the synthetic traces d with analytic signal A:
// the misfit term (FP64 block partials of 0.5 * r^2), q (the signal whose
// Hilbert transform enters the adjoint source) and P (the rest):
//   env_L2   r = de - oe,        f = r / de,       q = f dH,       P = d f
//   phase_L2 r = ph - oph,       f = r / de^2,     q = f d,        P = -f
//   env_L22  r = de^2 - oe^2,    f = r,            q = f dH,       P = d f
//   env_LOG  r = log de - log oe, f = log(oe / de), q = f dH / de^2, P = -f d / de^2
__global__ void k_misfit_stage1(const float* __restrict__ d, const float2* __restrict__ ana,
                                const float* __restrict__ ov, long long n, int nt, int kind,
                                float* __restrict__ q, float* __restrict__ P,
                                double* __restrict__ part) {
  const float inv = 1.f / (float)nt;
  double acc = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const float re = ana[t].x * inv, dH = ana[t].y * inv, x = d[t], o = ov[t];
    const float de = sqrtf(re * re + dH * dH);
    float f, qq, pp;
    double r;
    if (kind == 1) {
      f = (de - o) / de;
      qq = f * dH;
      pp = x * f;
      r = (double)de - (double)o;
    } else if (kind == 2) {
      const float ph = atan2f(dH, re);
      f = (ph - o) / (de * de);
      qq = f * x;
      pp = -f;
      r = (double)ph - (double)o;
    } else if (kind == 3) {
      f = de * de - o * o;
      qq = f * dH;
      pp = x * f;
      r = (double)de * de - (double)o * o;
    } else {
      f = logf(o / de);
      qq = f * (dH / (de * de));
      pp = -f * (x / (de * de));
      r = log((double)de) - log((double)o);
    }
    q[t] = qq;
    P[t] = pp;
    acc += 0.5 * r * r;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double ps[4];
  if ((threadIdx.x & 63) == 0) ps[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (ps[0] + ps[1]) + (ps[2] + ps[3]);
}

// Stage 2: e = P -+ imag(hilbert(q)); NaN -> 0 for env_L2 and env_LOG
// (w_bozdag.m:25,54)
__global__ void k_misfit_stage2(const float* __restrict__ P, const float2* __restrict__ anaq,
                                long long n, int nt, int kind, float* __restrict__ e) {
  const float inv = 1.f / (float)nt;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const float h = anaq[t].y * inv;
    float v = (kind == 1 || kind == 3) ? P[t] - h : P[t] + h;
    if ((kind == 1 || kind == 4) && v != v) v = 0.f;
    e[t] = v;
  }
}

// Ez(S) on the domain -> frame (used for the frame of a restored checkpoint)
// (bf16: frames of kind 2, rounded as the recompute sweep's frame stores)
__global__ void k_domain_frame(DevGrid g, int ns, const float* __restrict__ E,
                               float* __restrict__ frame, long long fstride, int bf16) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long per = (long long)g.m * g.n;
  if (t >= per * ns) return;
  const int sh = (int)(t / per);
  const long long r = t % per;
  const int b = (int)(r / g.m), a = (int)(r % g.m);
  const long long o = (long long)sh * fstride + (long long)b * g.LD + (a + g.di0);
  const float v = E[(long long)sh * g.FS + (long long)(b + g.dj0) * g.LD + (a + g.di0)];
  if (bf16)  // slot 0: the high half of word plane 0
    reinterpret_cast<unsigned short*>(frame)[2 * o + 1] = frame_bf16(v);
  else
    frame[o] = v;
}

// acc (float2 per domain cell per shot) -> FP64 gradient, summed over shots
// in shot order, or with gstride != 0 kept per shot (shot sh at
// ge + sh*gstride, as w_update_e.m keeps g_e per source before its
// post-processing); acc is cleared for the next segment.
__global__ void k_flush(int m, int n, int LD, int di0, int ns, float* __restrict__ acc,
                        long long astride, const double* __restrict__ fe,
                        const double* __restrict__ fs, double* __restrict__ ge,
                        double* __restrict__ gs, double scale, long long gstride) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long per = (long long)m * n;
  if (t >= per) return;
  const long long ka = (t / m) * LD + (t % m) + di0;  // partials use the field pitch
  float2* a2 = reinterpret_cast<float2*>(acc);
  const double ce = fe ? fe[t] : scale, cs = fs ? fs[t] : scale;
  double se = 0.0, ss = 0.0;
  for (int sh = 0; sh < ns; ++sh) {
    float2 v = a2[(long long)sh * astride + ka];
    if (gstride) {  // one gradient per shot
      ge[sh * gstride + t] += ce * (double)v.x;
      gs[sh * gstride + t] += cs * (double)v.y;
    } else {
      se += (double)v.x;
      ss += (double)v.y;
    }
    a2[(long long)sh * astride + ka] = make_float2(0.f, 0.f);
  }
  if (!gstride) {
    ge[t] += ce * se;
    gs[t] += cs * ss;
  }
}

// L2 residual e = d - d_o (w_bozdag.m:13-15) and per-block partial sums of
// e^2 in FP64 (w_E.m:15), finished by k_reduce_d.
// per-block partial sums of x^2 in FP64 (the energy of the observed traces)
__global__ void k_sumsq(const float* __restrict__ x, long long n, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    acc += (double)x[t] * (double)x[t];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Line-search objective terms (w_E.m:7-16 and the normalisation of
// w_objs_e.m:56-57) of n samples d (synthetic) and o (observed), each taken
// through w_E's transform T of kind e_kind (3, env_L22: squares; 4, env_LOG:
// logarithms; else the values): per block, 0.5 sum (T(d) - T(o))^2 into
// pa[block] and sum o^2 into pb[block], FP64.
__global__ void k_obj_terms(const float* __restrict__ d, const float* __restrict__ o, long long n,
                            int e_kind, double* __restrict__ pa, double* __restrict__ pb) {
  double a = 0.0, b = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    double x = d[t], y = o[t];
    b += y * y;
    if (e_kind == 3) {
      x = x * x;
      y = y * y;
    } else if (e_kind == 4) {
      x = log(x);
      y = log(y);
    }
    const double r = x - y;
    a += 0.5 * r * r;
  }
  for (int s = 32; s > 0; s >>= 1) {
    a += __shfl_xor(a, s);
    b += __shfl_xor(b, s);
  }
  __shared__ double ps[2][4];
  if ((threadIdx.x & 63) == 0) {
    ps[0][threadIdx.x >> 6] = a;
    ps[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = (ps[0][0] + ps[0][1]) + (ps[0][2] + ps[0][3]);
    pb[blockIdx.x] = (ps[1][0] + ps[1][1]) + (ps[1][2] + ps[1][3]);
  }
}

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

// Reference gradient (w_grad.m:72): the residual of step it_r overwrites Ez
// at every receiver of the shot before the step.  One thread walks the
// receivers in order, so a repeated receiver keeps the last value, as
// MATLAB's Ez(Mw) = e assignment does.
__global__ void k_ref_overwrite(DevGrid g, int nt, int t0, int t1, const int* __restrict__ tr_row,
                                const int* __restrict__ tr_col, const float* __restrict__ adj,
                                int it_r, float* __restrict__ E) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int t = t0; t < t1; ++t)
    E[(long long)tr_col[t] * g.LD + tr_row[t]] = adj[(long long)t * nt + it_r];
}

// Reference correlation (w_grad.m:86-87) for one step: with a = Ez on the
// domain after the step and u the stored forward frame k (1-based step),
//   acc.x += (d_t u)_k * a   (differentiate_cube.m:23-51 stencil)
//   acc.y += u_k * a
// U points at frame 0 (Ez(S_0)); frames 1..nt are the forward steps.
__global__ void k_ref_corr(DevGrid g, const float* __restrict__ E, const float* __restrict__ U,
                           int k, int nt, double dt, float* __restrict__ acc,
                           float* __restrict__ amax) {
  const long long per = (long long)g.m * g.n;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per) return;
  const int b = (int)(t / g.m), a = (int)(t % g.m);
  const long long FD = (long long)g.LD * g.n, ka = (long long)b * g.LD + a + g.di0;
  const float av = E[(long long)(b + g.dj0) * g.LD + (a + g.di0)];
  auto u = [&](int f) { return (double)U[(long long)f * FD + ka]; };
  double du;
  if (k == 1)
    du = (-1.5 * u(1) + 2.0 * u(2) - 0.5 * u(3)) / dt;
  else if (k == nt)
    du = (1.5 * u(nt) - 2.0 * u(nt - 1) + 0.5 * u(nt - 2)) / dt;
  else
    du = (u(k + 1) - u(k - 1)) / (2.0 * dt);
  float2* a2 = reinterpret_cast<float2*>(acc);
  float2 v = a2[ka];
  v.x += (float)du * av;
  v.y += (float)u(k) * av;
  a2[ka] = v;
  if (amax) amax[t] = fabsf(av);
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

// ==========================================================================
// adjoint-source conditioning: band filter (filt_gauss.m) and linear mute
// (w_lmute.m).  Traces are trace-major, trace r at d + r*nt (the layout of a
// MATLAB nt x nr array).
// ==========================================================================
// Tukey window of n samples with taper fraction r at sample k (tukeywin.m:
// raised-cosine flanks over floor(r/2*(n-1))+1 samples at each end)
__device__ __forceinline__ double tukey_w(int k, int n, double r) {
  if (n <= 1 || r <= 0.0) return 1.0;
  const double per = 0.5 * r, x = (double)k / (double)(n - 1);
  const int tl = (int)floor(per * (double)(n - 1));  // last 0-based sample of the left flank
  if (k <= tl) return 0.5 * (1.0 + cos(M_PI / per * (x - per)));
  if (k >= n - 1 - tl) return 0.5 * (1.0 + cos(M_PI / per * (x - 1.0 + per)));
  return 1.0;
}

// pad[r][k] = d[r][k] * tukey(k) for k < nt, 0 up to N (filt_gauss.m:19-24)
__global__ void k_tukey_pad(const float* __restrict__ d, int nt, int N, long long ntr, double r,
                            float* __restrict__ pad) {
  const long long n = ntr * N;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const long long tr = p / N;
    const int k = (int)(p - tr * N);
    pad[p] = k < nt ? (float)((double)d[tr * nt + k] * tukey_w(k, nt, r)) : 0.f;
  }
}

// integer power of a double (MATLAB's x.^steep for an integer steep)
__device__ __forceinline__ double ipow(double x, int e) {
  double y = 1.0;
  for (int i = 0; i < e; ++i) y *= x;
  return y;
}

// Band weight g(f) = exp(-((f - mid)^steep) / ((df/2) * width^(steep-1)))
// (filt_gauss.m:32-35) on the frequencies of the FFT order (bin k: k*df for
// k < N/2, (k-N)*df above).  The filtered trace is the real part of the
// inverse transform, which is the inverse of the Hermitian part of the
// product: bin k of the one-sided spectrum takes (g(f_k) + g(-f_k)) / 2
// (bins 0 and N/2 their own weight), and the 1/N of the inverse.
__global__ void k_gauss_band(float2* __restrict__ spec, int N, long long ntr, double df,
                             double mid, double width, int steep) {
  const int nb = N / 2 + 1;
  const long long n = ntr * nb;
  const double den = (0.5 * df) * ipow(width, steep - 1);
  auto g = [&](double f) { return exp(-ipow(f - mid, steep) / den); };
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(p % nb);
    double w;
    if (k == 0)
      w = g(0.0);
    else if (2 * k == N)
      w = g(-0.5 * N * df);
    else
      w = 0.5 * (g(k * df) + g(-k * df));
    w /= (double)N;
    const float2 z = spec[p];
    spec[p] = make_float2((float)(z.x * w), (float)(z.y * w));
  }
}

// d[r][k] = pad[r][k], k < nt (filt_gauss.m:44)
__global__ void k_trim(const float* __restrict__ pad, int N, int nt, long long ntr,
                       float* __restrict__ d) {
  const long long n = ntr * nt;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const long long tr = p / nt;
    d[p] = pad[tr * N + (p - tr * nt)];
  }
}

// Sigmoid mute of the linear arrival t0 + x_r/v (w_lmute.m:27-41):
// m(t, r) = 1 / (1 + exp(5 (t0 + (x_r - x_0)/v - t))); d and d_o are muted in
// place and e = d - d_o.  Times in seconds, v in m/s.
__global__ void k_lmute(float* __restrict__ d, float* __restrict__ d_o, float* __restrict__ e,
                        int nt, long long ntr, const double* __restrict__ t,
                        const double* __restrict__ rx, double v, double t0) {
  const long long n = ntr * nt;
  const double x0 = rx[0];
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const long long tr = p / nt;
    const int k = (int)(p - tr * nt);
    const double to = t0 + (rx[tr] - x0) / v;
    const double m = 1.0 / (1.0 + exp(5.0 * (to - t[k])));
    const float a = (float)(d[p] * m), b = (float)(d_o[p] * m);
    d[p] = a;
    d_o[p] = b;
    if (e) e[p] = a - b;
  }
}


// --------------------------------------------------------------------------
// gradient post-processing (w_update_e.m:108-145, w_update_e_.m:38-41,
// w_image_e.m:50-51).  Images are FP64, m x n column-major (depth fastest),
// image c at offset c*m*n.
// --------------------------------------------------------------------------

// Normaliser of the source taper of image c: the maximum over the grid of
// 1 - exp(-(((x_j - x_s)/L)^2 + ((z_i - z_s)/L)^2)), L = 2*lo (gaussi.m:12-16).
// The exponent is a sum of a term in j and a term in i and the map is
// increasing, so the maximum sits at the largest term of each; 0 becomes 1
// (normali.m:3-5).  One block per image.
__global__ void k_gaussi_norm(const double* __restrict__ X, int n, const double* __restrict__ Y,
                              int m, const int* __restrict__ s, const double* __restrict__ lo,
                              double* __restrict__ gmax) {
  __shared__ double sx[256], sz[256];
  const int c = blockIdx.x;
  const double L = 2.0 * lo[c];
  const double xs = X[s[2 * c + 1] - 1], zs = Y[s[2 * c] - 1];
  double ax = 0.0, az = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double a = (X[j] - xs) / L;
    ax = fmax(ax, a * a);
  }
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double b = (Y[i] - zs) / L;
    az = fmax(az, b * b);
  }
  sx[threadIdx.x] = ax;
  sz[threadIdx.x] = az;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sx[threadIdx.x] = fmax(sx[threadIdx.x], sx[threadIdx.x + w]);
      sz[threadIdx.x] = fmax(sz[threadIdx.x], sz[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double g = 1.0 - exp(-(sx[0] + sz[0]));
    gmax[c] = g == 0.0 ? 1.0 : g;
  }
}

// g_c(i,j) *= (1 - exp(-(((x_j - x_s)/L)^2 + ((z_i - z_s)/L)^2))) / gmax_c
// (gaussi.m:10-16 times the gradient, w_update_e.m:118)
__global__ void k_gaussi_apply(double* __restrict__ g, int m, int n, long long count,
                               const double* __restrict__ X, const double* __restrict__ Y,
                               const int* __restrict__ s, const double* __restrict__ lo,
                               const double* __restrict__ gmax) {
  const long long per = (long long)m * n, tot = per * count;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tot;
       p += (long long)gridDim.x * blockDim.x) {
    const long long c = p / per, q = p - c * per;
    const int j = (int)(q / m), i = (int)(q - (long long)j * m);
    const double L = 2.0 * lo[c];
    const double a = (X[j] - X[s[2 * c + 1] - 1]) / L, b = (Y[i] - Y[s[2 * c] - 1]) / L;
    const double t = (1.0 - exp(-(a * a + b * b))) / gmax[c];
    g[p] = t * g[p];
  }
}

// zero-padded copy of m x n images into M x N (powers of two) real planes
// (image_gaussian.m:22-26)
__global__ void k_img_pad(const double* __restrict__ in, int m, int n, int M, int N,
                          long long count, double* __restrict__ pad) {
  const long long per = (long long)M * N, tot = per * count;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tot;
       p += (long long)gridDim.x * blockDim.x) {
    const long long c = p / per, q = p - c * per;
    const int j = (int)(q / M), i = (int)(q - (long long)j * M);
    pad[p] = (i < m && j < n) ? in[c * (long long)m * n + (long long)j * m + i] : 0.0;
  }
}

// k-space filter of image_gaussian.m:37-53 on the one-sided spectrum of the
// padded plane (bins kx = 0..N-1 along x, ky = 0..M/2 along depth).  After
// the fftshift of line 50, bin k carries the filter value of the centred
// grid at (k + N/2) mod N, i.e. the signed offset ((k + N/2) mod N) - N/2.
// The filter is even in both offsets, so the product stays Hermitian and the
// real part of ifft2 is the C2R inverse; 1/(M N) is that inverse's scale.
// type 0: LOW_PASS exp(-r2) (maximum 1); type 1: HI_PASS 1 - exp(-r2),
// whose maximum is at the grid corner (offsets -N/2, -M/2).
__global__ void k_img_kfilter(double2* __restrict__ spec, int M, int N, long long count, double a,
                              double b, int type) {
  const int mb = M / 2 + 1;
  const long long per = (long long)N * mb, tot = per * count;
  const double x0 = 0.5 * N, y0 = 0.5 * M;
  double vmax = 1.0;
  if (type == 1) {
    const double u = (0.0 - x0) / a, v = (0.0 - y0) / b;
    vmax = 1.0 - exp(-(u * u + v * v));
  }
  const double inv = 1.0 / ((double)M * (double)N);
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tot;
       p += (long long)gridDim.x * blockDim.x) {
    const long long q = p % per;
    const int kx = (int)(q / mb), ky = (int)(q - (long long)kx * mb);
    const double u = ((double)((kx + N / 2) % N) - x0) / a;
    const double v = ((double)((ky + M / 2) % M) - y0) / b;
    const double e = exp(-(u * u + v * v));
    const double h = (type == 1 ? 1.0 - e : e) / vmax * inv;
    const double2 z = spec[p];
    spec[p] = make_double2(h * z.x, h * z.y);
  }
}

// out_c(i,j) = pad_c(i,j), i < m, j < n (image_gaussian.m:59)
__global__ void k_img_trim(const double* __restrict__ pad, int M, int N, int m, int n,
                           long long count, double* __restrict__ out) {
  const long long per = (long long)m * n, tot = per * count;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tot;
       p += (long long)gridDim.x * blockDim.x) {
    const long long c = p / per, q = p - c * per;
    const int j = (int)(q / m), i = (int)(q - (long long)j * m);
    out[p] = pad[c * (long long)M * N + (long long)j * M + i];
  }
}

// per-block max |a| (NaN ignored, as MATLAB's max) and NaN count over image
// blockIdx.y; gridDim.x blocks per image
__global__ void k_absmax_part(const double* __restrict__ a, long long per,
                              double* __restrict__ pmax, int* __restrict__ pnan) {
  __shared__ double sm[256];
  __shared__ int sn[256];
  const long long c = blockIdx.y;
  const double* x = a + c * per;
  double mx = 0.0;
  int nn = 0;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < per;
       q += (long long)gridDim.x * blockDim.x) {
    const double v = x[q];
    mx = fmax(mx, fabs(v));
    nn += isnan(v) ? 1 : 0;
  }
  sm[threadIdx.x] = mx;
  sn[threadIdx.x] = nn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
      sn[threadIdx.x] += sn[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[c * gridDim.x + blockIdx.x] = sm[0];
    pnan[c * gridDim.x + blockIdx.x] = sn[0];
  }
}

// a /= max|a| (0 -> 1) per image (normali.m:2-6); with guard, an image that
// holds a NaN becomes zeros (w_update_e.m:133-137); then a .*= shape when a
// mask is given (w_update_e.m:143).  Each block first combines its image's
// P partials (P <= 256).
__global__ void k_normali(double* __restrict__ a, long long per, int P,
                          const double* __restrict__ pmax, const int* __restrict__ pnan,
                          int guard, const double* __restrict__ shape) {
  __shared__ double sm[256];
  __shared__ int sn[256];
  const long long c = blockIdx.y;
  sm[threadIdx.x] = (int)threadIdx.x < P ? pmax[c * P + threadIdx.x] : 0.0;
  sn[threadIdx.x] = (int)threadIdx.x < P ? pnan[c * P + threadIdx.x] : 0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
      sn[threadIdx.x] += sn[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double mx = sm[0] == 0.0 ? 1.0 : sm[0];
  const bool zero = guard && sn[0] > 0;
  double* x = a + c * per;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < per;
       q += (long long)gridDim.x * blockDim.x) {
    const double v = zero ? 0.0 : x[q] / mx;
    x[q] = shape ? v * shape[q] : v;
  }
}

// d(t) = sum_c (-step_c * g_c(t)) / count, in image order (w_update_e.m:177
// and w_update_e_.m:38); zero when the steps sum below zero
// (w_update_e_.m:40-41, flag computed by the caller).  With a model,
// model(t) *= exp(model(t) * d(t)) (w_image_e.m:50-51).
__global__ void k_stack_update(const double* __restrict__ g, long long per, int count,
                               const double* __restrict__ step, int zero, double* __restrict__ d,
                               double* __restrict__ model) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < per;
       t += (long long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    if (!zero) {
      for (int c = 0; c < count; ++c) acc += -step[c] * g[c * per + t];
      acc /= (double)count;
    }
    d[t] = acc;
    if (model) {
      const double e = model[t];
      model[t] = e * exp(e * acc);
    }
  }
}

}  // namespace gjw
