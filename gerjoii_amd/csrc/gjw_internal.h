// gjw_internal.h -- device-side data layout shared by the kernels and the
// host driver of libgjw.  Not part of the public C ABI (include/gjw.h).
//
// Index convention on the device: 0-based, depth i fastest.  Ez lives on
// rows 0..NX and columns 0..NY (MATLAB's (NX+1) x (NY+1), w_solve.m:70-72);
// Hx uses rows 0..NX, columns 0..NY-1; Hy rows 0..NX-1, columns 0..NY.  All
// three share one pitched layout: element (i, j) sits at j*LD + i, with LD a
// multiple of 64 so that a 64-lane wavefront reads one aligned 256-byte row
// segment.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gjw {

// Per-survey constants (shared by every shot of the survey).
struct DevGrid {
  int NX, NY, P, LD;
  long long FS;            // floats per field per shot = LD*(NY+1)
  float bx0, by0;          // inner Chxez, Chyez (w_coeff.m:34-40; Chxh=Chyh=1)
  const float2* cc;        // inner (Ceze, Cezhy) per Ez cell (Cezhx = -Cezhy, dy = dx)
  const float* cxe;        // split-field Ezx/Ezy coefficients at PML Ez cells
  const float* cxh;
  const float* cye;
  const float* cyh;
  const float* hax;        // Hx strip cells (columns 0..P-1, NY-P..NY-1)
  const float* hbx;
  const float* hay;        // Hy strip cells (rows 0..P-1, NX-P..NX-1)
  const float* hby;
  // every coefficient per cell (edge and strip columns of the two-step kernels):
  // uc0 = (ce, cxe, cxh, cye), uc1 = (cyh, AX, BX, AY), uc2 = BY
  const float4* uc0;
  const float4* uc1;
  const float* uc2;
  // the block all of the above live in (19 FS floats: cc at 0, cxe 2 FS,
  // cxh 3 FS, cye 4 FS, cyh 5 FS, hax 6 FS, hbx 7 FS, hay 8 FS, hby 9 FS,
  // uc0 10 FS, uc1 14 FS, uc2 18 FS); the multi-step kernels address every
  // coefficient through one buffer descriptor over it
  const float* coef;
  // stored domain window (w_geom.m:168-172): rows di0..di0+m-1, cols dj0..dj0+n-1
  int di0, dj0, m, n;
};

// Row tiles of the 8-step forward that march two rows per lane (k_fwd_step8,
// fwd8_pairs).  In strips slo..shi, row tiles t0 .. t0+n-1 are such tiles:
// tile t0+q owns rows lo + 112 q .. min(lo + 112 q + 111, hi); row tiles
// t0+n .. tb-1 hold nothing; the others are the one-row tiles of the same
// index.  n = 0: every tile holds one row per lane.
struct RowPairs {
  int t0, n, tb;
  int lo, hi;
  int slo, shi;
};

// Per-shot geometry of one survey.
struct DevShots {
  int nt;
  const int* si;     // source row (0-based Ez index), per shot
  const int* sj;     // source column
  const float* sc;   // Cezj of the shot, 0 when the source cell is overwritten
  const float* wav;  // wavelets [shot][nt]
  const int* cptr;   // receivers by column: [shot][NY+2] offsets into crow/cid
  const int* crow;   // receiver row
  const int* cid;    // global trace index; trace t occupies tr[t*nt .. t*nt+nt)
  int rmin, rmax;    // receiver rows over all shots (waves outside skip the lookup)
  float* jz;         // source recorder Ez(source) per step, [shot][nt], or null
  const int* rp;     // first trace of each shot, [shot], and one past the last
  int rec_multi;     // some cell holds several receivers of one shot
};

// ---- several surveys in one launch (gjw_batch_*) ----
constexpr int kMaxSurveys = 8;

// Static description of one survey of a multi-survey launch (device array).
// Buffer layouts are those of a single survey: forward state fld = E, Hx, Hy
// of parity 0, then of parity 1, then X, Y of parity 0 and 1 (fsz floats
// each); adjoint state adjf = (M, Nx, Ny, X, Y) of parity 0, then 1.
struct DevSurvey {
  DevGrid g;
  DevShots s;
  int ns;              // shots
  long long fsz;       // ns * FS
  float* fld;
  float* adjf;
  float* tr;           // traces, [trace][nt]
  const float* adj;    // adjoint source, traces layout
  float* frames;       // domain frames of the current segment: slot q at q * FD
  long long FD;        // floats per frame of one shot (LD * n)
  long long fstride;   // floats per shot's frame set
  float* acc;          // correlation partials: float2 per frame cell, FD per shot
};

// Per-launch part of a multi-survey launch (kernel argument): survey sid[k]
// owns blocks [blk0[k], blk0[k+1]).
// The surveys of a multi-survey launch, passed by value in the kernel
// arguments: pointers read from the kernel arguments are known to be global
// memory, so the bodies' loads and stores compile to global/buffer
// instructions (pointers read from a device array would take the slower
// flat path).
struct MultiTable {
  DevSurvey sv[kMaxSurveys];
};

struct MultiArgs {
  int n;
  int blk0[kMaxSurveys + 1];
  int sid[kMaxSurveys];
  int it[kMaxSurveys];     // step (forward) / later transposed step (adjoint)
  int L[kMaxSurveys];      // columns per wavefront
  int cur[kMaxSurveys];    // field parity (forward E/H, adjoint state)
  int xc[kMaxSurveys];     // split-field parity (forward)
  int fslot[kMaxSurveys];  // frame slot: of Ez(S_{it+1}) (recompute), of Ez(S_it) (adjoint); -1 none
  int rec;                 // forward: record traces
};

enum { KIND_RING = 0, KIND_INNER = 1, KIND_PML = 2 };

// floats per cell slot of the coefficient buffer (cc 2, split 4, H strips 4, table 9)
constexpr int kCoefFloats = 19;

__host__ __device__ inline int ekind(int NX, int NY, int P, int i, int j) {
  if (i <= 0 || i >= NX || j <= 0 || j >= NY) return KIND_RING;
  if (i <= P || i >= NX - P || j <= P || j >= NY - P) return KIND_PML;
  return KIND_INNER;
}

}  // namespace gjw
