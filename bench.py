#!/usr/bin/env python3
"""FWI shot-gradients/s on one MI355X: batched vs plain, at <= 1e-3 gradient
error; Gcell-updates/s; roofline of the dominant kernel; CPU baseline.

One step = one full FWI gradient of the survey: forward of every shot (with
checkpoints), L2 misfit against observed traces, exact adjoint with
recomputation, FP64 gradient reduced over time and shots.  Inputs (model,
observed traces, geometry) are resident in HBM before the timed region;
observed traces come from a forward solve of the true model (synthetic data).

Default workload: 2000 x 600 model, 40-cell PML, nt = 8000, 64 shots at 0.3 m
spacing, 400 receivers per shot, FP32 wavefields, FP64 gradient accumulation.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
BYTES_PER_CELL_UPDATE = 32  # single-pass model: Ez, Hx, Hy read+write + 2 coefficient reads (FP32)
# step kernels in the numbering of gjw_kernel_stats / gjw_tile:
# (name, steps per launch, kind of launch for the compulsory-byte model)
KERNELS = [("k_fwd_step", 1, "fwd"), ("k_adj_step", 1, "adj"), ("k_fwd_step2", 2, "fwd"),
           ("k_adj_step2", 2, "adj"), ("k_fwd_step4", 4, "fwd"), ("k_adj_step4", 4, "adj"),
           ("k_fwd_step8", 8, "fwd"), ("k_adj_step8", 8, "adj"), ("k_adj_step6", 6, "adj"),
           ("k_fwd_step8_frames", 8, "fwd_frames")]


def compulsory_bytes(kind, S, g, ns, ntr, fb=4):
    """HBM bytes one S-step launch over all shots must move at least: every
    field read and written once (Ez, Hx, Hy on all padded cells, the split
    fields on the PML cells), the survey's coefficients read once (shared by
    the shots: (Ceze, Cezhy) per cell, the 36-B table per PML cell), plus
    what the launch exists to move: the traces it records (forward sweep),
    the S Ez frames of the domain it stores (recompute sweep), or the S+1
    frames, the adjoint-source samples and the FP32 correlation partials
    (read and written) it uses (adjoint).  fb: bytes per frame element (4
    for FP32 frames, 2 for BF16)."""
    NX, NY = g.nx, g.ny
    cells = (NX + 1) * (NY + 1)
    pml = (NX - 1) * (NY - 1) - (NX - 1 - 2 * g.pml) * (NY - 1 - 2 * g.pml)
    dom = g.m * g.n
    b = (24 * cells + 16 * pml) * ns + 8 * cells + 36 * pml
    if kind == "fwd":
        return b + 4 * S * ntr
    if kind == "fwd_frames":
        return b + fb * S * dom * ns
    return b + fb * (S + 1) * dom * ns + 16 * dom * ns + 4 * S * ntr


def kernel_source_sha16():
    """Hash of the kernel sources, so a PMC summary can be matched to them."""
    import hashlib
    h = hashlib.sha256()
    for f in ("gjw_kernels.hip", "gjw_internal.h"):
        with open(os.path.join(ROOT, "gerjoii_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=8000)
    ap.add_argument("--plain-shots", type=int, default=2)
    ap.add_argument("--cpu-shots", type=int, default=16)
    ap.add_argument("--cpu-nt", type=int, default=480)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-linesearch", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--study", choices=["c4", "c5"], default=None,
                    help="instead of the headline: the precision study at 4000 x 1500 (c4) or the "
                         "uneven three-survey batch (c5), one JSON line")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--fuse", type=int, default=0, help="forward steps per launch: 0 autotuned, 1, 2, 4, 8")
    ap.add_argument("--adj-fuse", type=int, default=0, help="adjoint steps per launch: 0 (six), 4, 6, 8")
    ap.add_argument("--adj-tile", type=int, default=0,
                    help="adjoint tile width (0: tuned in the first gradient sweep)")
    ap.add_argument("--ckpt", type=int, default=0)
    ap.add_argument("--frames", choices=["fp32", "bf16"], default="bf16",
                    help="element type of the recompute sweep's Ez frames (GJW_OPT_FRAMES); "
                         "with bf16 a gradient with FP32 frames is timed beside (frames_fp32)")
    ap.add_argument("--row-pairs", type=int, default=1, choices=[0, 1],
                    help="1: the 8-step forward's interior row tiles hold two rows per lane "
                         "(GJW_OPT_ROW_PAIRS); 0: one row per lane")
    ap.add_argument("--graph", type=int, default=1, help="1: replay the gradient as captured HIP graphs")
    ap.add_argument("--no-sqrt", action="store_true",
                    help="skip the extra gradient timed at a checkpoint interval of about sqrt(nt)")
    ap.add_argument("--probe-copy", action="store_true",
                    help="only run three 1-GiB device copies (calibrates PMC byte counters)")
    args = ap.parse_args()
    if args.probe_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"probe_copy_bytes": a.numel() * 4}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the survey's shots are "
                          "batched inside one device; use --gpus 1"}))
        sys.exit(2)

    import torch
    from gerjoii_amd import configs
    from gerjoii_amd import survey as S
    from gerjoii_amd import _native as N

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.study:
        print(json.dumps(_study_c4() if args.study == "c4" else _study_c5()))
        return
    case = configs.c3(ns=args.shots, nt=args.nt)
    g = case.grid
    cells = g.cells()
    ns = len(case.src)
    log(f"workload C3: model {g.n}x{g.m}, padded Ez {g.ny + 1}x{g.nx + 1} = {cells} cells, "
        f"nt={g.nt}, shots={ns}")

    ctx = S.context(0)
    sv = S.Survey(g, ctx)
    sv.set_shots(case.src, case.wavelets, case.receivers)
    if args.tile:
        sv.set_option(N.GJW_OPT_TILE_L, args.tile)
    if args.fuse:
        sv.set_option(N.GJW_OPT_FUSE, args.fuse)
    if args.adj_fuse:
        sv.set_option(N.GJW_OPT_ADJ_FUSE, args.adj_fuse)
    if args.adj_tile:
        sv.set_option(N.GJW_OPT_ADJ_TILE, args.adj_tile)
    if args.ckpt:
        sv.set_option(N.GJW_OPT_CKPT_EVERY, args.ckpt)
    if args.frames == "bf16":
        sv.set_option(N.GJW_OPT_FRAMES, N.GJW_FRAMES_BF16)
    sv.set_option(N.GJW_OPT_ROW_PAIRS, args.row_pairs)
    if args.graph:
        sv.set_option(N.GJW_OPT_GRAPH, 1)
    # observed data: true model, kept in HBM
    sv.set_model(case.eps_true, case.sig_true)
    d_obs = torch.zeros(sv.traces_shape(), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    sv.forward(out=d_obs)
    log("observed traces ready; tile L =", sv.tile(4))
    sv.set_model(case.eps0, case.sig0)
    g_eps = torch.zeros((g.n, g.m), dtype=torch.float64, device=dev)  # (m, n) column-major
    g_sig = torch.zeros_like(g_eps)

    for w in range(args.warmup):
        sv.fwi_gradient(d_obs, g_eps, g_sig)
        log(f"warmup {w + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = 0.0
    for k in range(args.steps):
        _, _, J = sv.fwi_gradient(d_obs, g_eps, g_sig)
        log(f"step {k + 1}/{args.steps} J={J:.6e}")
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    # kernel durations: one more gradient, launched eagerly (graph replays
    # record no events) with HIP event pairs around one launch in 16 steps of
    # each step kernel, on the survey's stream
    sv.set_option(N.GJW_OPT_GRAPH, 0)
    sv.set_option(N.GJW_OPT_KTIMING, 16)
    for kk in range(len(KERNELS)):
        sv.kernel_stats(kk, reset=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sv.fwi_gradient(d_obs, g_eps, g_sig)
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - t1) * 1e3
    sv.set_option(N.GJW_OPT_KTIMING, 0)
    kstats = [sv.kernel_stats(kk) for kk in range(len(KERNELS))]
    ms_step = dt_s * 1e3 / args.steps
    shot_grads = ns * args.steps / dt_s
    cell_updates = 3.0 * cells * g.nt * ns * args.steps  # forward, recompute, adjoint
    gcells = cell_updates / dt_s / 1e9
    tile_L = {name: sv.tile(w) for w, (name, _, _) in enumerate(KERNELS) if w >= 2}
    K_used = sv.query(N.GJW_Q_CKPT_INTERVAL)
    depths = (sv.query(N.GJW_Q_FWD_DEPTH), sv.query(N.GJW_Q_ADJ_DEPTH))
    frames_used = "bf16" if sv.query(N.GJW_Q_FRAMES) == N.GJW_FRAMES_BF16 else "fp32"
    ntr = sv.ntraces
    # the same gradient at BASELINE's checkpoint interval, about sqrt(nt) steps
    # (rounded up to whole launches of both depths)
    sqrt_run = None
    if not args.no_sqrt and not args.ckpt:
        q = int(np.lcm(*depths))
        Ks = int(np.ceil(np.sqrt(g.nt) / q) * q)
        sv.set_option(N.GJW_OPT_CKPT_EVERY, Ks)
        sv.set_option(N.GJW_OPT_GRAPH, 1 if args.graph else 0)
        sv.fwi_gradient(d_obs, g_eps, g_sig)  # new buffers and (graph) capture
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sv.fwi_gradient(d_obs, g_eps, g_sig)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t1
        sqrt_run = {"checkpoint_interval": Ks, "s": ts, "shot_gradients_per_s": ns / ts,
                    "note": "one gradient after one untimed call, same survey and options"}
        sv.set_option(N.GJW_OPT_CKPT_EVERY, 0)
        log(f"sqrt(nt) interval {Ks}: {ns / ts:.3f} shot-gradients/s")
    # the same gradient with FP32 frames (the library's default)
    f32_run = None
    if frames_used == "bf16":
        sv.set_option(N.GJW_OPT_FRAMES, N.GJW_FRAMES_F32)
        sv.set_option(N.GJW_OPT_GRAPH, 1 if args.graph else 0)
        sv.fwi_gradient(d_obs, g_eps, g_sig)  # new buffers and (graph) capture
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sv.fwi_gradient(d_obs, g_eps, g_sig)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t1
        f32_run = {"s": tf, "shot_gradients_per_s": ns / tf,
                   "frames": "fp32" if sv.query(N.GJW_Q_FRAMES) == N.GJW_FRAMES_F32 else "bf16",
                   "note": "one gradient after one untimed call, same survey, FP32 recompute frames"}
        log(f"FP32 frames: {ns / tf:.3f} shot-gradients/s")
    sv.close()
    del sv
    torch.cuda.empty_cache()

    # Per-kernel roofline.  achieved = compulsory bytes of one launch (all
    # shots; compulsory_bytes) / live average launch duration (HIP events on
    # the survey's stream around one launch in 16 steps).  A fused launch
    # moves each field once per S steps, so these bytes, not S x 32 B per
    # cell, are the ones an S-step launch can be held to: frac <= 1.
    # traffic = measured HBM bytes per launch from the committed rocprofv3
    # PMC summary of the same kernel (profiles/), used only when that summary
    # was taken on the current kernel sources.
    src_sha = kernel_source_sha16()
    kern = {}
    for (name, spl, kind), (avg_ms, samples, launches) in zip(KERNELS, kstats):
        v = {"avg_ms": avg_ms, "samples": samples, "steps_per_launch": spl,
             "launches_per_step": launches,
             "bytes_per_launch_compulsory": compulsory_bytes(kind, spl, g, ns, ntr,
                                                             2 if frames_used == "bf16" else 4)}
        v["achieved_GBs"] = (v["bytes_per_launch_compulsory"] / (avg_ms * 1e-3) / 1e9
                             if avg_ms > 0 else None)
        v["share_of_step"] = avg_ms * launches / ms_eager if ms_eager > 0 else None
        t, note = _traffic_from_profiles(name, src_sha, frames_used)
        v["traffic_bytes"] = t
        if note:
            v["traffic_note"] = note
        v["hbm_GBs"] = (t / (avg_ms * 1e-3) / 1e9) if (t and avg_ms > 0) else None
        v["traffic_over_compulsory"] = t / v["bytes_per_launch_compulsory"] if t else None
        kern[name] = v
    dom = max(kern, key=lambda k: kern[k]["avg_ms"] * kern[k]["launches_per_step"])
    kd = kern[dom]
    roof = {"kernel": dom, "bound": "hbm", "achieved": kd["achieved_GBs"],
            "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": (kd["achieved_GBs"] or 0) / PEAK_HBM_GBS,
            "traffic": kd["traffic_bytes"],
            "hbm_GBs_measured_traffic": kd["hbm_GBs"],
            "hbm_frac_measured_traffic": (kd["hbm_GBs"] or 0) / PEAK_HBM_GBS if kd["hbm_GBs"] else None,
            "bytes_per_launch_compulsory": kd["bytes_per_launch_compulsory"],
            "single_pass_model": {
                "bytes_per_cell_update": BYTES_PER_CELL_UPDATE,
                "GBs_equivalent": cell_updates * BYTES_PER_CELL_UPDATE / dt_s / 1e9,
                "note": "32 B per cell-update as if every step read and wrote HBM; the fused "
                        "kernels do not, so this figure may exceed the peak and is not a "
                        "roofline fraction"},
            "kernel_source_sha16": src_sha,
            "kernels": kern}
    # the whole step's compulsory bytes against its wall time (must stay below the peak)
    step_bytes = sum(v["bytes_per_launch_compulsory"] * v["launches_per_step"] for v in kern.values())
    roof["step_compulsory_GBs"] = step_bytes / (ms_step * 1e-3) / 1e9
    roof["instrumented_gradient_ms"] = ms_eager

    # plain single-shot path on a subset of the same shots
    plain = None
    if not args.no_plain and args.plain_shots > 0:
        q = args.plain_shots
        sp = S.Survey(g, ctx)
        sp.set_shots(case.src[:q], case.wavelets[:q], case.receivers[:q])
        sp.set_path(True)
        sp.set_model(case.eps0, case.sig0)
        rows = int(sp.rec_ptr[-1])
        dsub = d_obs[:rows].contiguous()
        ge_p = torch.zeros_like(g_eps)
        gs_p = torch.zeros_like(g_sig)
        sp.fwi_gradient(dsub, ge_p, gs_p)  # warm
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sp.fwi_gradient(dsub, ge_p, gs_p)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        plain = {"shot_gradients_per_s": q / tp, "shots_timed": q, "s": tp,
                 "gcell_updates_per_s": 2.0 * cells * g.nt * q / tp / 1e9,
                 "path": "one shot per launch, whole domain wavefield stored (no checkpoints)"}
        sp.close()
        del sp
        torch.cuda.empty_cache()
        log(f"plain: {plain['shot_gradients_per_s']:.3f} shot-gradients/s")

    cpu = None if args.no_cpu else _cpu_baseline(case, args)
    cond = _conditioning(ctx, d_obs, g.dt, args.no_cpu)
    lsearch = None if args.no_linesearch else _line_search(ctx, case, d_obs)
    post = _postprocess(ctx, case, args.no_cpu)
    mat = _matio(case, d_obs)
    uneven = None if args.no_batch else _uneven_batch(ctx)

    out = {
        "metric": "FWI shot-gradients/s, batched vs plain, at <=1e-3 grad err; Gcell-updates/s",
        "value": shot_grads,
        "unit": "shot-gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded layered model with topography; observed traces from the true model)",
        "config": {"workload": "batched FWI gradient, 2000x600 model, 40-cell PML, nt=%d, %d shots, "
                               "400 receivers/shot, checkpoint + recompute, FP64 accumulation" % (g.nt, ns),
                   "grid_cells": cells, "nt": g.nt, "shots": ns, "tile_L": tile_L,
                   "checkpoint_interval": K_used,
                   "checkpoint_rule": "about 2 sqrt(nt) in whole launches (sqrt(nt) run beside: "
                                      "checkpoint_sqrt_nt)" if not args.ckpt else "set by --ckpt",
                   "steps_per_launch": {"forward": depths[0], "adjoint": depths[1]},
                   "frames": frames_used,
                   "row_pairs": bool(args.row_pairs),
                   "frames_note": "element type of the Ez frames the recompute sweep hands to the "
                                  "adjoint (GJW_OPT_FRAMES); the BF16-frame gradient of one shot of "
                                  "this workload against the FP64 oracle is checked to <= 1e-3 in "
                                  "tests/test_gpu_frames16.py; FP32 frames timed beside as "
                                  "frames_fp32",
                   "launches_per_gradient": {k: v["launches_per_step"] for k, v in kern.items()
                                             if v["launches_per_step"] > 0},
                   "hip_graph": bool(args.graph),
                   "launch_scheme": "one launch per fused step group for all shots (not one "
                                    "persistent kernel); the sequence replayed as captured "
                                    "HIP graphs" if args.graph else "eager launches",
                   "parallelism": "single GPU, all shots in each launch"},
        "gcell_updates_per_s": gcells,
        "misfit_J": J,
        "plain": plain,
        "batched_vs_plain": (shot_grads / plain["shot_gradients_per_s"]) if plain else None,
        "roofline": roof,
        "checkpoint_sqrt_nt": sqrt_run,
        "frames_fp32": f32_run,
        "cpu_baseline": cpu,
        "adjoint_source_filter": cond,
        "line_search": lsearch,
        "gradient_postprocess": post,
        "mat_io": mat,
        "uneven_batch": uneven,
    }
    print(json.dumps(out))


def _line_search(ctx, case, d_obs, nparabo=3):
    """Objectives of a parabolic line search (w_stepparabola_e.m /
    w_objs_e.m) at the workload: nparabo perturbed models of all shots,
    batched (one Batch of nparabo surveys, shared launches) and one model
    at a time, after a warm-up call each."""
    import torch
    from gerjoii_amd import survey as S
    from gerjoii_amd import linesearch as LS
    rng = np.random.default_rng(5)
    gdir = rng.standard_normal(case.eps0.shape)
    gdir /= np.abs(gdir).max()
    svs = []
    for k in np.linspace(-0.002, 0.002, nparabo):
        sv = S.Survey(case.grid, ctx)
        sv.set_model(LS.perturb_eps(case.eps0, gdir, k), case.sig0)
        sv.set_shots(case.src, case.wavelets, case.receivers)
        svs.append(sv)
    batch = S.Batch(svs)
    batch.objective([d_obs] * nparabo)  # warm-up (allocations, autotuner)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = batch.objective([d_obs] * nparabo)
    tb = time.perf_counter() - t0
    batch.close()
    one = S.Batch(svs[:1])
    one.objective([d_obs])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sv in svs:
        S.Batch([sv]).objective([d_obs])
    ts = time.perf_counter() - t0
    one.close()
    ns = len(case.src)
    for sv in svs:
        sv.close()
    torch.cuda.empty_cache()
    return {"models": nparabo, "shots": ns, "s_batched": tb, "s_one_model_at_a_time": ts,
            "shot_objectives_per_s": nparabo * ns / tb,
            "objectives": [float(x) for x in J],
            "note": "forward only (no wavefield), L2 normalised as w_objs_e.m"}


def _observed(sv, case):
    """Observed traces of the true model, on the device."""
    import torch
    sv.set_model(case.eps_true, case.sig_true)
    d = torch.zeros(sv.traces_shape(), dtype=torch.float32, device="cuda")
    sv.forward(out=d)
    sv.set_model(case.eps0, case.sig0)
    return d


def _study_c4(ns_cmp=16):
    """Precision of 16-bit checkpoints at 4000 x 1500, nt 12000: the FWI
    gradient of the first `ns_cmp` shots with FP32, BF16 and FP16 snapshots
    (FP32 as the reference: at this size FP32 snapshots of all 32 shots do
    not fit one GPU next to the frames), then all 32 shots with BF16
    snapshots timed."""
    import torch
    from gerjoii_amd import configs
    from gerjoii_amd import survey as S
    from gerjoii_amd import _native as N
    ctx = S.context(0)
    case = configs.c4()
    g = case.grid
    sub = dict(src=case.src[:ns_cmp], wavelets=case.wavelets[:ns_cmp], receivers=case.receivers[:ns_cmp])
    grads = {}
    for name, snap in (("fp32", N.GJW_SNAP_F32), ("bf16", N.GJW_SNAP_BF16), ("fp16", N.GJW_SNAP_F16)):
        sv = S.Survey(g, ctx)
        sv.set_shots(sub["src"], sub["wavelets"], sub["receivers"])
        d = _observed(sv, case)
        sv.set_option(N.GJW_OPT_SNAPSHOT, snap)
        t0 = time.perf_counter()
        ge, gs, J = sv.fwi_gradient(d)
        grads[name] = (ge, gs, J, time.perf_counter() - t0)
        log(f"c4 {name}: J={J:.6e} {grads[name][3]:.1f} s")
        sv.close()
        del sv, d
        torch.cuda.empty_cache()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    cos = lambda a, b: float(np.ravel(a) @ np.ravel(b) / (np.linalg.norm(a) * np.linalg.norm(b)))
    ref = grads["fp32"]
    out = {"grid": [g.n, g.m], "nt": g.nt, "shots_compared": ns_cmp}
    for name in ("bf16", "fp16"):
        ge, gs, J, _ = grads[name]
        out[name] = {"eps_rel_l2": rel(ge, ref[0]), "eps_cosine": cos(ge, ref[0]),
                     "sig_rel_l2": rel(gs, ref[1]), "sig_cosine": cos(gs, ref[1]),
                     "J_rel": abs(J - ref[2]) / abs(ref[2])}
    sv = S.Survey(g, ctx)
    sv.set_shots(case.src, case.wavelets, case.receivers)
    d = _observed(sv, case)
    sv.set_option(N.GJW_OPT_SNAPSHOT, N.GJW_SNAP_BF16)
    sv.fwi_gradient(d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sv.fwi_gradient(d)
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    cells = 3.0 * g.cells() * g.nt * len(case.src)
    out["bf16_all_shots"] = {"shots": len(case.src), "s": dt_s,
                             "shot_gradients_per_s": len(case.src) / dt_s,
                             "gcell_updates_per_s": cells / dt_s / 1e9}
    sv.close()
    return {"study": "c4 precision", **out,
            "note": "reference: the FP32-snapshot gradient of the same shots (the library's FP32 path "
                    "meets relative L2 <= 1e-3 against the FP64 oracle in the parity tests)"}


def _study_c5():
    """The uneven three-survey batch at full size: envelope, phase and envelope
    misfits, BF16 checkpoints, captured graphs, per-shot gradients reduced in
    FP64 per survey; shared launches against one survey at a time."""
    import torch
    from gerjoii_amd import configs
    from gerjoii_amd import survey as S
    from gerjoii_amd import _native as N
    ctx = S.context(0)
    cases = configs.c5()
    svs, dobs = [], []
    for case, misfit in zip(cases, ("env_L2", "phase_L2", "env_L2")):
        sv = S.Survey(case.grid, ctx)
        sv.set_shots(case.src, case.wavelets, case.receivers)
        dobs.append(_observed(sv, case))
        sv.set_misfit(misfit)
        sv.set_option(N.GJW_OPT_SNAPSHOT, N.GJW_SNAP_BF16)
        sv.set_option(N.GJW_OPT_GRAPH, 1)
        svs.append(sv)
    cells = sum(3.0 * c.grid.cells() * c.grid.nt * len(c.src) for c in cases)
    for sv, d in zip(svs, dobs):
        sv.fwi_gradient(d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sv, d in zip(svs, dobs):
        sv.fwi_gradient(d)
    torch.cuda.synchronize()
    tq = time.perf_counter() - t0
    bt = S.Batch(svs)
    bt.fwi_gradient(dobs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ge, gs, J = bt.fwi_gradient(dobs)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    finite = all(bool(np.isfinite(a).all()) for a in ge + gs)
    shots = sum(len(c.src) for c in cases)
    bt.close()
    for sv in svs:
        sv.close()
    return {"study": "c5 uneven batch",
            "surveys": [{"grid": [c.grid.n, c.grid.m], "nt": c.grid.nt, "shots": len(c.src)}
                        for c in cases],
            "misfits": ["env_L2", "phase_L2", "env_L2"], "snapshots": "bf16", "graph": True,
            "s_batched": tb, "s_one_survey_at_a_time": tq,
            "shot_gradients_per_s_batched": shots / tb,
            "gcell_updates_per_s_batched": cells / tb / 1e9,
            "gcell_updates_per_s_one_at_a_time": cells / tq / 1e9,
            "J": [float(x) for x in J], "gradients_finite": finite}


def _uneven_batch(ctx, specs=((8, 2000), (8, 1800), (8, 1600), (8, 1400), (6, 1200), (6, 1000))):
    """Surveys of uneven shot count and trace length on the workload's grid
    (each (shots, nt) of `specs`; by default six lines of 6-8 shots, each too
    small to fill the GPU alone): their FWI gradients one survey at a time,
    concurrently (each survey on its own context and stream, one host thread
    each) and in the shared launches of gjw_batch_fwi_gradient; one warm-up
    call each, then one timed call."""
    import torch
    from gerjoii_amd import configs
    from gerjoii_amd import survey as S
    cases = [configs.c3(ns=a, nt=b) for a, b in specs]

    def build(c, cx):
        sv = S.Survey(c.grid, cx)
        sv.set_shots(c.src, c.wavelets, c.receivers)
        return sv, _observed(sv, c)

    def timed(f):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    cells = sum(3.0 * c.grid.cells() * c.grid.nt * len(c.src) for c in cases)
    ctxs = [S.Context(0) for _ in cases]
    own = [build(c, cx) for c, cx in zip(cases, ctxs)]
    svs, dobs = [a for a, _ in own], [b for _, b in own]
    tq = timed(lambda: [sv.fwi_gradient(d) for sv, d in zip(svs, dobs)])
    tc = timed(lambda: S.concurrent_fwi_gradients(svs, dobs))
    for sv in svs:
        sv.close()
    for cx in ctxs:
        cx.close()
    shared = [build(c, ctx) for c in cases]
    # each survey's own first sweep tunes its adjoint width, which the shared
    # launches then use
    for sv, d in shared:
        sv.fwi_gradient(d)
    bt = S.Batch([sv for sv, _ in shared])
    tb = timed(lambda: bt.fwi_gradient([d for _, d in shared]))
    bt.close()
    for sv, _ in shared:
        sv.close()
    del own, svs, dobs, shared
    torch.cuda.empty_cache()
    shots = sum(a for a, _ in specs)
    return {"surveys": [{"shots": a, "nt": b} for a, b in specs], "grid_cells": cases[0].grid.cells(),
            "s_one_survey_at_a_time": tq, "s_concurrent": tc, "s_shared_launches": tb,
            "shot_gradients_per_s_concurrent": shots / tc,
            "gcell_updates_per_s": {"one_at_a_time": cells / tq / 1e9,
                                    "concurrent": cells / tc / 1e9,
                                    "shared_launches": cells / tb / 1e9},
            "note": "concurrent: each survey on its own context (HIP stream), single-survey kernels, "
                    "one host thread per survey; shared launches: eight-step forwards and six-step "
                    "adjoints across surveys, one checkpoint interval"}


def _matio(case, d_obs, lines=8):
    """Radargram .mat I/O (w_natur_u.m writer, w_load.m reader) on a sample
    of the workload's lines: write `lines` line<is>.mat files plus the
    position cells to a temporary folder, read them back as survey inputs,
    and copy the traces to the device."""
    import shutil
    import tempfile
    import torch
    from gerjoii_amd import matio as MIO
    g = case.grid
    lines = min(lines, len(case.src))
    recs = case.receivers[:lines]
    ntr = sum(len(r) for r in recs)
    tr = d_obs[:ntr].cpu().numpy()
    tmp = tempfile.mkdtemp(prefix="gjw_mat_")
    try:
        t0 = time.perf_counter()
        MIO.write_survey(tmp, g, case.src[:lines], recs, tr, case.wavelets[:lines], case.f0)
        tw = time.perf_counter() - t0
        t0 = time.perf_counter()
        src, wav, rr, d = MIO.load_survey(g, tmp, range(1, lines + 1))
        dd = torch.from_numpy(d).cuda()
        torch.cuda.synchronize()
        tl = time.perf_counter() - t0
        ok = bool(np.array_equal(d, tr))
        mb = 4.0 * ntr * g.nt / 1e6
        nbytes = sum(os.path.getsize(os.path.join(tmp, f)) for f in os.listdir(tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    del dd
    return {"lines": lines, "traces": ntr, "nt": g.nt, "trace_MB": mb, "file_MB": nbytes / 1e6,
            "write_MBs": mb / tw, "load_to_device_MBs": mb / tl, "round_trip_exact": ok,
            "note": "MATLAB v5 files with zlib compression (MATLAB's default -v7), scipy.io, "
                    "one host thread; load includes the copy of the traces to the device"}


def _postprocess(ctx, case, no_cpu):
    """Per-shot gradient post-processing and the update (w_update_e.m:108-145,
    w_update_e_.m:38-41, w_image_e.m:50-51) on one FP64 gradient image per
    shot of the workload, resident on the device: time per stage (each entry
    point synchronises on return) over 5 passes after one warm-up, and the
    FP64 numpy oracle chain on a 2-image sample."""
    import torch
    from gerjoii_amd import postproc as PP
    g = case.grid
    m, n, ns = g.m, g.n, len(case.src)
    gen = torch.Generator(device="cuda").manual_seed(5)
    G = torch.randn((ns, n, m), dtype=torch.float64, device="cuda", generator=gen)
    shape = torch.from_numpy(np.ascontiguousarray(case.shape.T, dtype=np.float64)).cuda()
    model = torch.from_numpy(np.ascontiguousarray(case.eps0.T, dtype=np.float64)).cuda()
    X = np.arange(n) * g.dx
    Y = np.arange(m) * g.dx
    s = PP.taper_index(g, case.src)
    lo = PP.source_wavelength(case.eps0, s, case.f0)
    ax = az = g.dx / (1.5 * float(np.median(lo)))  # 1/(1.5 lo) [1/m] times dx, as the project scripts
    steps = np.full(ns, 1e-3)
    stages = {"gaussi": lambda: ctx.gaussi(G, X, Y, s, 2 * lo),
              "image_gaussian": lambda: ctx.image_gaussian(G, ax, az, "LOW_PASS", out=G),
              "normali_guard_mask": lambda: ctx.normali(G, nan_guard=True, shape=shape),
              "stack_update": lambda: ctx.stack_update(G, steps, model=model)}
    for f in stages.values():
        f()
    torch.cuda.synchronize()
    reps = 5
    ms = {}
    for k, f in stages.items():
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        ms[k] = (time.perf_counter() - t0) * 1e3 / reps
    tot = sum(ms.values())
    out = {"images": ns, "m": m, "n": n, "ms_per_stage": ms, "ms_total": tot,
           "images_per_s": ns / (tot * 1e-3),
           "io_GBs_images": 8.0 * m * n * ns * 7 / (tot * 1e-3) / 1e9,
           "note": "FP64 images on the device; io_GBs_images counts 7 passes over the images "
                   "(gaussi R+W, filter R+W, normali R + R+W) against the whole chain's time"}
    if not no_cpu:
        from oracle import oracle as O
        q = min(2, ns)
        h = G[:q].cpu().numpy()
        c0 = time.perf_counter()
        imgs = [O.w_update_post(h[k].T, X, Y, s[k], lo[k], ax, az, case.shape) for k in range(q)]
        O.stack_update(np.stack(imgs, axis=2), steps[:q], model=case.eps0)
        cs = time.perf_counter() - c0
        out["cpu_images_per_s"] = q / cs
        out["cpu_sample"] = "%d images, numpy (FP64), one thread of the host" % q
    del G, shape, model
    torch.cuda.empty_cache()
    return out


def _conditioning(ctx, d_obs, dt, no_cpu):
    """The adjoint-source band filter (filt_gauss.m, gjw_filt_gauss) on all
    traces of the workload: device time per call (HIP events around 5 calls
    after one warm-up) and the FP64 numpy oracle on a 64-trace sample."""
    import torch
    ntr, nt = d_obs.shape
    f_low, f_high = 5e7, 4e8
    d = d_obs.clone()
    ctx.filt_gauss(d, dt, f_low, f_high)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        ctx.filt_gauss(d, dt, f_low, f_high)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out = {"traces": ntr, "nt": nt, "ms_per_call": ms,
           "traces_per_s": ntr / (ms * 1e-3),
           "io_GBs": 8.0 * ntr * nt / (ms * 1e-3) / 1e9,
           "note": "in place on device traces; io_GBs counts one read and one write of the traces"}
    if not no_cpu:
        from oracle import oracle as O
        q = 64
        x = d_obs[:q].cpu().numpy().T.astype(np.float64)
        c0 = time.perf_counter()
        O.filt_gauss(x, dt, f_low, f_high, 8)
        cs = time.perf_counter() - c0
        out["cpu_traces_per_s"] = q / cs
        out["cpu_sample"] = "%d traces, numpy FFT (FP64), one thread of the host" % q
    del d
    torch.cuda.empty_cache()
    return out


def _traffic_from_profiles(kernel, src_sha, frames=None):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    kernel at this workload (profiles/<kernel>_C3_pmc.txt, lines
    'bytes_per_launch: <float>' and 'kernel_source_sha16: <hex>').  Returns
    (bytes or None, note): a summary taken on other kernel sources is not used."""
    p = os.path.join(ROOT, "profiles", f"{kernel}_C3_pmc.txt")
    if not os.path.exists(p):
        return None, None
    val, sha, fr = None, None, "fp32"
    for line in open(p):
        if line.startswith("bytes_per_launch:"):
            try:
                val = float(line.split(":")[1])
            except ValueError:
                val = None
        elif line.startswith("kernel_source_sha16:"):
            sha = line.split(":")[1].strip()
        elif line.startswith("frames:"):
            fr = line.split(":")[1].strip()
    if sha != src_sha:
        return None, f"{os.path.basename(p)} was taken on other kernel sources ({sha})"
    if frames is not None and kernel in ("k_fwd_step8_frames", "k_adj_step6") and fr != frames:
        return None, f"{os.path.basename(p)} was taken with {fr} frames"
    return val, None


def _cpu_baseline(case, args):
    """The FP32 CPU oracle (forward + exact adjoint, one shot per OpenMP
    thread) on a bounded sample of the same workload: the same grid and
    shots, truncated to --cpu-nt steps, extrapolated per step to nt."""
    from oracle import oracle as O
    g = case.grid
    q = min(args.cpu_shots, len(case.src))
    nt_s = min(args.cpu_nt, g.nt)
    og = O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, nt_s)
    ep, sp = O.exp2pml(og, case.eps0, case.sig0)
    rng = np.random.default_rng(0)
    e = rng.standard_normal((q, len(case.receivers[0]), nt_s)).astype(np.float32)
    K = max(1, int(np.ceil(2 * np.sqrt(nt_s))))  # the library's interval, about 2 sqrt(nt)
    t0 = time.perf_counter()
    _, _, thr = O.survey_grad_f32(og, ep, sp, case.src[:q], case.wavelets[:q, :nt_s],
                                  np.array(case.receivers[:q]), e, K)
    dt_s = time.perf_counter() - t0
    per_shot_full = dt_s / q * (g.nt / nt_s)
    cells = g.cells()
    log(f"cpu baseline: {q} shots x {nt_s} steps in {dt_s:.1f} s on {thr} threads")
    cpu_model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": 1.0 / per_shot_full,
            "unit": "shot-gradients/s", "cores": thr, "kind": "port",
            "cpu_model": cpu_model, "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "gcell_updates_per_s": 3.0 * cells * nt_s * q / dt_s / 1e9,
            "sample": f"{q} shots of the workload, first {nt_s} of {g.nt} steps (forward, checkpointed "
                      f"recompute, exact adjoint), float oracle, extrapolated linearly in nt"}


if __name__ == "__main__":
    main()
