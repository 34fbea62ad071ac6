"""GPU parity at the benchmark configurations' full sizes, through the C ABI.

* 800 x 300, nt 4000, one source: exact gradient of the L2 misfit against the
  FP64 oracle (relative L2 <= 1e-3, cosine >= 0.9999) and central finite
  differences through the FP32 GPU forward on 16 seeded cells;
* 2000 x 600, nt 8000 (the benchmark workload): one shot's gradient on the
  default batched path (8-step forward, 6-step adjoint, default checkpoint
  interval) against the FP64 oracle, same bar;
* 4000 x 1500, nt 12000: two shots with BF16 and FP16 checkpoint snapshots
  against FP32 snapshots, one shot with BF16 snapshots against the FP64
  oracle over the first 1200 steps, and the same survey downsampled 2x
  (2000 x 750, nt 6000, the full duration), one shot with FP32, BF16 and
  FP16 snapshots against the FP64 oracle over all steps;
* the uneven three-survey batch (600 x 200 / 3000 / 24 shots, 1500 x 500 /
  7000 / 48, 3000 x 800 / 10000 / 16; envelope, phase, envelope misfits;
  BF16 snapshots; captured graphs): each survey's gradient from the shared
  launches equals its own one-survey gradient.

The FP64 oracle runs its column loops on OpenMP threads (bit-identical to a
serial run), which keeps each full-size check within about a minute.
"""
import dataclasses
import time

import numpy as np
import pytest

from gerjoii_amd import configs
from gerjoii_amd import survey as S
from gerjoii_amd import _native as N

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def cosine(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def shots(case, idx):
    idx = list(idx)
    return dataclasses.replace(case, src=case.src[idx], wavelets=case.wavelets[idx],
                               receivers=[case.receivers[k] for k in idx])


def survey(case, ctx, snap=None, L=0, K=0, misfit=None, graph=False, fuse=0, adj_fuse=0):
    sv = S.Survey(case.grid, ctx)
    sv.set_shots(case.src, case.wavelets, case.receivers)
    sv.set_model(case.eps0, case.sig0)
    if snap is not None:
        sv.set_option(N.GJW_OPT_SNAPSHOT, snap)
    if L:
        sv.set_option(N.GJW_OPT_TILE_L, L)
    if K:
        sv.set_option(N.GJW_OPT_CKPT_EVERY, K)
    if fuse:
        sv.set_option(N.GJW_OPT_FUSE, fuse)
    if adj_fuse:
        sv.set_option(N.GJW_OPT_ADJ_FUSE, adj_fuse)
    if misfit:
        sv.set_misfit(misfit)
    if graph:
        sv.set_option(N.GJW_OPT_GRAPH, 1)
    return sv


def observed(case, ctx):
    sv = S.Survey(case.grid, ctx)
    sv.set_shots(case.src, case.wavelets, case.receivers)
    sv.set_model(case.eps_true, case.sig_true)
    d = sv.forward()
    sv.close()
    return d


def oracle_grad(O, case, adj, rec_ptr):
    g = case.grid
    og = O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, g.nt, eps0=g.eps0, mu0=g.mu0, R0=g.R0,
                order=g.pml_order)
    ep, sp = O.exp2pml(og, case.eps0, case.sig0)
    ge = np.zeros((og.NX, og.NY))
    gs = np.zeros((og.NX, og.NY))
    for q in range(len(case.src)):
        e = adj[rec_ptr[q]:rec_ptr[q + 1]].T.astype(np.float64)
        a, b, _ = O.grad_exact(og, ep, sp, case.src[q], case.wavelets[q], case.receivers[q], e,
                               pml_terms=False)
        ge += a
        gs += b
    return O.fold_pad(og, ge, gs)


def gpu_l2_gradient(case, ctx, dobs, **kw):
    """Forward with checkpoints, L2 residual, exact gradient; returns
    (g_eps, g_sig, residual, J)."""
    sv = survey(case, ctx, **kw)
    sv.forward(keep=True)
    e, J = sv.misfit(dobs)
    ge, gs = sv.gradient()
    rp = sv.rec_ptr
    sv.close()
    return ge, gs, e, J, rp


def test_gradient_c2_fp64_and_finite_differences(oracle_mod, gpu_ctx):
    case = configs.c2()
    print("C2 gradient: 800 x 300, nt 4000, 1 shot", flush=True)
    dobs = observed(case, gpu_ctx)
    ge, gs, e, J, rp = gpu_l2_gradient(case, gpu_ctx, dobs)
    t0 = time.perf_counter()
    fe, fs = oracle_grad(oracle_mod, case, e, rp)
    print(f"C2 FP64 oracle {time.perf_counter() - t0:.1f} s; eps rel {rel_l2(ge, fe):.2e} "
          f"cos {cosine(ge, fe):.8f}; sig rel {rel_l2(gs, fs):.2e} cos {cosine(gs, fs):.8f}",
          flush=True)
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999
    # central differences of J = 1/2 |d(eps) - d_obs|^2 (FP32 GPU forward, FP64
    # sum) on 16 seeded cells drawn among the ground cells, away from the
    # model's outermost rows and columns, whose |g| is at least a tenth of the
    # maximum (where a 0.5% perturbation moves J well above FP32 noise)
    inner = np.zeros_like(ge, dtype=bool)
    inner[2:-2, 2:-2] = True
    cand = np.flatnonzero(inner & (case.shape > 0) & (np.abs(ge) >= 0.1 * np.abs(ge).max()))
    cells = np.random.default_rng(1002).choice(cand, 16, replace=False)
    sv = survey(case, gpu_ctx)
    dob = dobs.astype(np.float64)
    fd, an = [], []
    for k in cells:
        a, b = np.unravel_index(k, ge.shape)
        h = 5e-3 * case.eps0[a, b]
        vals = []
        for sgn in (1, -1):
            E = case.eps0.copy()
            E[a, b] += sgn * h
            sv.set_model(E, case.sig0)
            d = sv.forward().astype(np.float64)
            vals.append(0.5 * float(((d - dob) ** 2).sum()))
        fd.append((vals[0] - vals[1]) / (2 * h))
        an.append(ge[a, b])
    sv.close()
    fd, an = np.array(fd), np.array(an)
    rel = np.abs(fd - an) / np.abs(an)
    print("C2 FD: per-cell rel err max %.2e median %.2e, rel L2 %.2e" % (rel.max(), np.median(rel),
                                                                       rel_l2(an, fd)), flush=True)
    assert rel.max() <= 0.05 and rel_l2(an, fd) <= 0.02


def test_gradient_c3_shot_fp64(oracle_mod, gpu_ctx):
    """The benchmark workload's grid and trace length, one of its shots, on
    the path the benchmark times."""
    case = shots(configs.c3(), [31])
    print("C3 one shot: 2000 x 600, nt 8000", flush=True)
    dobs = observed(case, gpu_ctx)
    ge, gs, e, J, rp = gpu_l2_gradient(case, gpu_ctx, dobs)
    t0 = time.perf_counter()
    fe, fs = oracle_grad(oracle_mod, case, e, rp)
    print(f"C3 FP64 oracle {time.perf_counter() - t0:.1f} s; eps rel {rel_l2(ge, fe):.2e} "
          f"cos {cosine(ge, fe):.8f}; sig rel {rel_l2(gs, fs):.2e} cos {cosine(gs, fs):.8f}",
          flush=True)
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999


def test_c4_snapshot_precision(gpu_ctx):
    """4000 x 1500, nt 12000, two shots: FWI gradients with BF16 and FP16
    checkpoint snapshots against the FP32-snapshot gradient."""
    case = shots(configs.c4(), [15, 16])
    print("C4 two shots: 4000 x 1500, nt 12000", flush=True)
    dobs = observed(case, gpu_ctx)
    out = {}
    for name, snap in (("fp32", N.GJW_SNAP_F32), ("bf16", N.GJW_SNAP_BF16), ("fp16", N.GJW_SNAP_F16)):
        sv = survey(case, gpu_ctx, snap=snap)
        ge, gs, J = sv.fwi_gradient(dobs)
        sv.close()
        out[name] = (ge, gs, J)
        print(f"C4 {name}: J {J:.6e}", flush=True)
    ref = out["fp32"]
    for name in ("bf16", "fp16"):
        ge, gs, J = out[name]
        print(f"C4 {name} vs fp32: eps rel {rel_l2(ge, ref[0]):.2e} cos {cosine(ge, ref[0]):.10f}; "
              f"sig rel {rel_l2(gs, ref[1]):.2e}; J rel {abs(J - ref[2]) / ref[2]:.1e}", flush=True)
        assert rel_l2(ge, ref[0]) <= 1e-3 and cosine(ge, ref[0]) >= 0.9999
        assert rel_l2(gs, ref[1]) <= 1e-3 and cosine(gs, ref[1]) >= 0.9999
        assert abs(J - ref[2]) <= 1e-6 * ref[2]  # the forward sweep does not read snapshots


@pytest.mark.parametrize("snap", [N.GJW_SNAP_BF16])
def test_c4_shot_fp64_first_steps(oracle_mod, gpu_ctx, snap):
    """4000 x 1500 grid, one shot, the first 1200 of the 12000 steps (17
    checkpoint segments at the default interval): the production path (FP32
    fields, BF16 snapshots, FP64 accumulation) against the FP64 oracle."""
    case = shots(configs.c4(nt=1200), [16])
    print("C4 one shot, first 1200 steps, FP64 oracle", flush=True)
    dobs = observed(case, gpu_ctx)
    ge, gs, e, J, rp = gpu_l2_gradient(case, gpu_ctx, dobs, snap=snap)
    t0 = time.perf_counter()
    fe, fs = oracle_grad(oracle_mod, case, e, rp)
    print(f"C4 FP64 oracle {time.perf_counter() - t0:.1f} s; eps rel {rel_l2(ge, fe):.2e} "
          f"cos {cosine(ge, fe):.8f}; sig rel {rel_l2(gs, fs):.2e} cos {cosine(gs, fs):.8f}",
          flush=True)
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999


def test_c4_downsampled_full_duration_fp64(oracle_mod, gpu_ctx):
    """The same survey downsampled 2x (2000 x 750, dx 0.04 m, nt 6000: the
    full 0.51 us), one shot, all steps: gradients with FP32, BF16 and FP16
    snapshots on the default batched path against one FP64 oracle run."""
    case = shots(configs.c4(coarse=2), [16])
    g = case.grid
    print(f"C4/2 one shot: {g.n} x {g.m}, nt {g.nt}, FP64 oracle", flush=True)
    dobs = observed(case, gpu_ctx)
    runs = {}
    for name, snap in (("fp32", N.GJW_SNAP_F32), ("bf16", N.GJW_SNAP_BF16), ("fp16", N.GJW_SNAP_F16)):
        runs[name] = gpu_l2_gradient(case, gpu_ctx, dobs, snap=snap)
    # the forward sweep reads no snapshot: one residual serves all three
    e, rp = runs["fp32"][2], runs["fp32"][4]
    for name in ("bf16", "fp16"):
        assert np.array_equal(runs[name][2], e)
    t0 = time.perf_counter()
    fe, fs = oracle_grad(oracle_mod, case, e, rp)
    print(f"C4/2 FP64 oracle {time.perf_counter() - t0:.1f} s", flush=True)
    for name, (ge, gs, _, J, _) in runs.items():
        print(f"C4/2 {name} vs FP64: eps rel {rel_l2(ge, fe):.2e} cos {cosine(ge, fe):.8f}; "
              f"sig rel {rel_l2(gs, fs):.2e} cos {cosine(gs, fs):.8f}", flush=True)
        assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
        assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999


def test_c5_uneven_batch_full_size(gpu_ctx):
    """The three surveys at full size with their misfits, BF16 snapshots and
    captured graphs: every survey's gradient and misfit from the shared
    launches equal the survey's own (one common checkpoint interval, strip
    width and fusion depths, so both runs take the same arithmetic), on the
    eager first call and on the graph replay of the second."""
    cases = configs.c5()
    misfits = ("env_L2", "phase_L2", "env_L2")
    K = 216  # 2 sqrt(10000) in whole launches of the 8-step forward and 6-step adjoint
    print("C5 three surveys", flush=True)
    svs, dobs, own = [], [], []
    for case, mf in zip(cases, misfits):
        d = observed(case, gpu_ctx)
        sv = survey(case, gpu_ctx, snap=N.GJW_SNAP_BF16, L=96, K=K, misfit=mf, graph=True,
                    fuse=8, adj_fuse=6)
        own.append(sv.fwi_gradient(d))
        print(f"C5 {case.name} alone: J {own[-1][2]:.6e}", flush=True)
        svs.append(sv)
        dobs.append(d)
    bt = S.Batch(svs)
    for call in range(2):
        ge, gs, J = bt.fwi_gradient(dobs)
        for k, case in enumerate(cases):
            re, rs = rel_l2(ge[k], own[k][0]), rel_l2(gs[k], own[k][1])
            rj = abs(J[k] - own[k][2]) / abs(own[k][2])
            print(f"C5 call {call} {case.name}: eps {re:.1e} sig {rs:.1e} J {rj:.1e}", flush=True)
            assert np.isfinite(ge[k]).all() and np.isfinite(gs[k]).all()
            assert re <= 1e-5 and rs <= 1e-5 and rj <= 1e-9
    bt.close()
    for sv in svs:
        sv.close()
