"""BF16 recompute frames (GJW_OPT_FRAMES) against the FP64 oracle, through the
C ABI.

The recompute sweep hands the adjoint one Ez frame per step.  With BF16
frames the adjoint sums its frame correlation by parts, so the gradient
carries the frames' 2^-9 rounding undifferenced:

* 200 x 100, nt 1500, four shots (the small configuration at full size):
  relative L2 <= 1e-3 and cosine >= 0.9999 against the FP64 exact adjoint,
  and within 1e-3 of the FP32-frame gradient;
* 2000 x 600, nt 8000 (the benchmark workload), one shot on the benchmark's
  path: the same bar against the FP64 oracle;
* a white-noise adjoint source, energy up to the Nyquist frequency: a
  characterisation, not a parity case.  The adjoint field is then grid-scale
  noise, its correlation with the smooth forward field is weak, and the
  frames' rounding is a large part of it (1.3e-2 in the eps gradient, as
  with BF16 checkpoint snapshots on such a source); the bar is 3e-2;
* narrow strips (16 and 24 columns) on a grid with PML rows, receiver rows
  and a source, so that many strips start, end and change chunk kind: each
  column of the BF16-frame gradient stays within 1e-2 of the FP32-frame one
  (a frame pair written with a wrong half would put that column's error
  near 1);
* an odd number of steps keeps FP32 frames (the one-step kernels store FP32
  frames only), reported by gjw_query, and then equals the FP32-frame
  gradient bit for bit.
"""
import numpy as np
import pytest

from gerjoii_amd import configs
from gerjoii_amd import _native as N

from test_gpu_scale import cosine, gpu_l2_gradient, observed, oracle_grad, rel_l2, shots, survey

pytestmark = pytest.mark.gpu


def frames_gradient(case, ctx, dobs, frames, adj=None, **kw):
    """Exact gradient with the given frame type; the L2 residual of dobs or
    the adjoint source adj.  Returns (g_eps, g_sig, residual, frames used,
    rec_ptr)."""
    sv = survey(case, ctx, **kw)
    sv.set_option(N.GJW_OPT_FRAMES, frames)
    sv.forward(keep=True)
    if adj is None:
        e, _ = sv.misfit(dobs)
        ge, gs = sv.gradient()
    else:
        e = adj
        ge, gs = sv.gradient(adj)
    used = sv.query(N.GJW_Q_FRAMES)
    rp = sv.rec_ptr
    sv.close()
    return ge, gs, e, used, rp


def test_bf16_frames_small_fp64(oracle_mod, gpu_ctx):
    case = configs.c1()
    dobs = observed(case, gpu_ctx)
    ge, gs, e, used, rp = frames_gradient(case, gpu_ctx, dobs, N.GJW_FRAMES_BF16)
    assert used == N.GJW_FRAMES_BF16
    g32 = frames_gradient(case, gpu_ctx, dobs, N.GJW_FRAMES_F32)
    assert g32[3] == N.GJW_FRAMES_F32
    fe, fs = oracle_grad(oracle_mod, case, e, rp)
    print(f"200x100 BF16 frames vs FP64: eps {rel_l2(ge, fe):.2e} sig {rel_l2(gs, fs):.2e}; "
          f"FP32 frames vs FP64: eps {rel_l2(g32[0], fe):.2e} sig {rel_l2(g32[1], fs):.2e}; "
          f"BF16 vs FP32 frames: eps {rel_l2(ge, g32[0]):.2e} sig {rel_l2(gs, g32[1]):.2e}",
          flush=True)
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999
    assert rel_l2(ge, g32[0]) <= 1e-3 and rel_l2(gs, g32[1]) <= 1e-3


def test_bf16_frames_benchmark_shot_fp64(oracle_mod, gpu_ctx):
    case = shots(configs.c3(), [31])
    dobs = observed(case, gpu_ctx)
    ge, gs, e, used, rp = frames_gradient(case, gpu_ctx, dobs, N.GJW_FRAMES_BF16)
    assert used == N.GJW_FRAMES_BF16
    fe, fs = oracle_grad(oracle_mod, case, e, rp)
    print(f"2000x600 nt 8000 BF16 frames vs FP64: eps {rel_l2(ge, fe):.2e} cos {cosine(ge, fe):.8f}; "
          f"sig {rel_l2(gs, fs):.2e} cos {cosine(gs, fs):.8f}", flush=True)
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999


def test_bf16_frames_white_noise_source(gpu_ctx):
    case = configs.small(nt=600, ns=2)
    sv = survey(case, gpu_ctx)
    d = sv.forward()
    sv.close()
    adj = np.random.default_rng(7).standard_normal(d.shape).astype(np.float32)
    ge, gs, _, used, _ = frames_gradient(case, gpu_ctx, None, N.GJW_FRAMES_BF16, adj=adj)
    g32 = frames_gradient(case, gpu_ctx, None, N.GJW_FRAMES_F32, adj=adj)
    assert used == N.GJW_FRAMES_BF16
    print(f"white-noise source, BF16 vs FP32 frames: eps {rel_l2(ge, g32[0]):.2e} "
          f"sig {rel_l2(gs, g32[1]):.2e}", flush=True)
    assert rel_l2(ge, g32[0]) <= 3e-2 and rel_l2(gs, g32[1]) <= 3e-2
    assert cosine(ge, g32[0]) >= 0.999


def test_bf16_frames_odd_nt_keeps_fp32(gpu_ctx):
    case = configs.small(nt=201, ns=2)
    sv = survey(case, gpu_ctx)
    d = sv.forward()
    sv.close()
    adj = np.random.default_rng(3).standard_normal(d.shape).astype(np.float32)
    ge, gs, _, used, _ = frames_gradient(case, gpu_ctx, None, N.GJW_FRAMES_BF16, adj=adj)
    g32 = frames_gradient(case, gpu_ctx, None, N.GJW_FRAMES_F32, adj=adj)
    assert used == N.GJW_FRAMES_F32
    assert np.array_equal(ge, g32[0]) and np.array_equal(gs, g32[1])


@pytest.mark.parametrize("L", [16, 24])
def test_bf16_frames_narrow_strips_per_column(gpu_ctx, L):
    case = configs.small(nt=400, ns=3)
    dobs = observed(case, gpu_ctx)
    ge, gs, e, used, _ = frames_gradient(case, gpu_ctx, dobs, N.GJW_FRAMES_BF16, L=L, fuse=8)
    g32 = frames_gradient(case, gpu_ctx, dobs, N.GJW_FRAMES_F32, L=L, fuse=8)
    assert used == N.GJW_FRAMES_BF16
    for a, b in ((ge, g32[0]), (gs, g32[1])):
        col = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0),
                                                          1e-30 + 1e-6 * np.linalg.norm(b))
        print(f"L {L}: worst column rel {col.max():.2e}, whole {rel_l2(a, b):.2e}", flush=True)
        assert col.max() <= 1e-2 and rel_l2(a, b) <= 1e-3
