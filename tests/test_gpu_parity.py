"""GPU parity: libgjw (HIP, FP32) against the FP64 CPU oracle, through the C ABI.

Tolerances (written per test): traces relative L2 <= 1e-4; gradients relative
L2 <= 1e-3 and cosine >= 0.9999 against the FP64 exact adjoint.
"""
import numpy as np
import pytest

from gerjoii_amd import configs
from gerjoii_amd import survey as S

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def cosine(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def oracle_grid(O, g):
    return O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, g.nt, eps0=g.eps0, mu0=g.mu0, R0=g.R0,
                  order=g.pml_order)


def oracle_forward(O, case, eps, sig):
    g = oracle_grid(O, case.grid)
    ep, sp = O.exp2pml(g, eps, sig)
    out = []
    for q in range(len(case.src)):
        d, _, _ = O.forward(g, ep, sp, case.src[q], case.wavelets[q], case.receivers[q])
        out.append(d.T)  # (nr, nt), traces layout
    return np.concatenate(out, axis=0)


def oracle_grad(O, case, eps, sig, adj, rec_ptr, snap=0, K=None):
    g = oracle_grid(O, case.grid)
    ep, sp = O.exp2pml(g, eps, sig)
    ge = np.zeros((g.NX, g.NY))
    gs = np.zeros((g.NX, g.NY))
    gws = []
    for q in range(len(case.src)):
        e = adj[rec_ptr[q]:rec_ptr[q + 1]].T.astype(np.float64)
        a, b, gw = O.grad_exact(g, ep, sp, case.src[q], case.wavelets[q], case.receivers[q], e,
                                pml_terms=False, snap=snap, K=K)
        ge += a
        gs += b
        gws.append(gw)
    fe, fs = O.fold_pad(g, ge, gs)
    return fe, fs, np.array(gws)


def make_survey(case, ctx, plain=False, ck=0, L=0):
    sv = S.Survey(case.grid, ctx)
    sv.set_model(case.eps0, case.sig0)
    sv.set_shots(case.src, case.wavelets, case.receivers)
    sv.set_path(plain)
    if ck:
        sv.set_option(2, ck)
    if L:
        sv.set_option(3, L)
    return sv


@pytest.mark.parametrize("air", [0, 4])
def test_forward_small(oracle_mod, gpu_ctx, air):
    case = configs.small(air=air)
    print(f"forward small air={air} grid={case.grid.nx}x{case.grid.ny} nt={case.grid.nt}", flush=True)
    sv = make_survey(case, gpu_ctx)
    d = sv.forward()
    ref = oracle_forward(oracle_mod, case, case.eps0, case.sig0)
    err = rel_l2(d, ref)
    print("trace rel L2", err)
    assert np.abs(ref).max() > 0
    assert err <= 1e-4


@pytest.mark.parametrize("L", [8, 16, 64])
def test_forward_tiles_agree(gpu_ctx, L):
    """Every tile width computes the same traces (fast and general paths)."""
    case = configs.small(m=150, n=140, P=12, nt=200)
    base = make_survey(case, gpu_ctx, L=4).forward()
    d = make_survey(case, gpu_ctx, L=L).forward()
    assert rel_l2(d, base) <= 1e-6


def test_forward_plain_equals_batched(gpu_ctx):
    case = configs.small()
    a = make_survey(case, gpu_ctx).forward()
    b = make_survey(case, gpu_ctx, plain=True).forward()
    assert rel_l2(a, b) <= 1e-6


def test_forward_c1(oracle_mod, gpu_ctx):
    case = configs.c1()
    print("forward C1", flush=True)
    sv = make_survey(case, gpu_ctx)
    d = sv.forward()
    ref = oracle_forward(oracle_mod, case, case.eps0, case.sig0)
    err = rel_l2(d, ref)
    print("C1 trace rel L2", err)
    assert err <= 1e-4


def test_forward_c2(oracle_mod, gpu_ctx):
    """The single-shot 800 x 300, nt = 4000 configuration at full size:
    traces against the FP64 oracle."""
    case = configs.c2()
    print("forward C2", flush=True)
    d = make_survey(case, gpu_ctx).forward()
    ref = oracle_forward(oracle_mod, case, case.eps0, case.sig0)
    err = rel_l2(d, ref)
    print("C2 trace rel L2", err)
    assert np.abs(ref).max() > 0 and err <= 1e-4


def test_gradient_c1(oracle_mod, gpu_ctx):
    """The 200 x 100, nt = 1500, 4-shot configuration at full size: the
    batched exact gradient (default depths, tiles and checkpoint interval)
    against the FP64 oracle's exact adjoint."""
    case = configs.c1()
    print("gradient C1", flush=True)
    sv = make_survey(case, gpu_ctx)
    sv.forward(keep=True)
    rng = np.random.default_rng(11)
    adj = rng.standard_normal(sv.traces_shape()).astype(np.float32)
    ge, gs = sv.gradient(adj)
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, adj, sv.rec_ptr)
    print("C1 eps", rel_l2(ge, fe), cosine(ge, fe), "sig", rel_l2(gs, fs), cosine(gs, fs))
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999


def test_c3_full_size_properties(gpu_ctx):
    """The 2000 x 600, nt = 8000 workload at full size (two of its shots; the
    FP64 oracle would take hours): the batched gradient equals the plain
    path's (one step per launch, whole wavefield stored: an independent
    code path) to 1e-4 (measured 2.6e-5 for eps and 4.7e-6 for sigma: FP32
    rounding in two operation orders over 8000 steps), and the gradient is
    linear in the adjoint source:
    doubling it doubles the gradient to 1e-5 (measured 1.8e-6 here; the
    residue is FP32 gradual underflow, test_doubling_is_exact_above_underflow)."""
    import dataclasses
    case = configs.c3()
    case = dataclasses.replace(case, src=case.src[30:32], wavelets=case.wavelets[30:32],
                               receivers=case.receivers[30:32])
    print("C3 two shots", flush=True)
    # a fixed strip width: the first sweep of a tuned survey tries several
    # adjoint widths, and results differ with the tiling in the last bits
    sv = make_survey(case, gpu_ctx, L=96)
    sv.forward(keep=True)
    rng = np.random.default_rng(8)
    adj = rng.standard_normal(sv.traces_shape()).astype(np.float32)
    ge, gs = sv.gradient(adj)
    sv.forward(keep=True)
    ge2, gs2 = sv.gradient(2 * adj)
    print("C3 linearity", rel_l2(ge2, 2 * ge), rel_l2(gs2, 2 * gs), flush=True)
    assert rel_l2(ge2, 2 * ge) <= 1e-5 and rel_l2(gs2, 2 * gs) <= 1e-5
    sp = make_survey(case, gpu_ctx, plain=True)
    sp.forward(keep=True)
    pe, ps = sp.gradient(adj)
    print("C3 batched vs plain", rel_l2(ge, pe), rel_l2(gs, ps), flush=True)
    assert rel_l2(ge, pe) <= 1e-4 and rel_l2(gs, ps) <= 1e-4
    sp.close()
    sv.close()


@pytest.mark.parametrize("plain", [True, False])
def test_doubling_is_exact_above_underflow(gpu_ctx, plain):
    """Power-of-two scaling commutes with every FP32 rounding except in the
    subnormal range (|x| < 2^-126, fixed absolute spacing 2^-149).  The
    adjoint field passes through that range ahead of its numerical front; the
    rounding there is not homogeneous, and over about a thousand steps the
    one-ulp differences it leaves spread through the field.  So, on the
    200 x 100, nt = 1500 case: doubling an adjoint source of unit scale
    leaves a residue at the FP32 rounding level (measured 2.4e-7 to 4.7e-7,
    bar 1e-5), the first differing wavelet-gradient sample appears more
    than 500 backward steps into the sweep (measured after about 1000), and
    with the source scaled by 2^60, which keeps the field clear of the
    subnormal range for the whole sweep, doubling is exact bit for bit.
    The wavelet gradient (the adjoint field at the source cell, in time) is
    what shows where the difference starts."""
    case = configs.c1()
    print(f"doubling C1 plain={plain}", flush=True)
    sv = make_survey(case, gpu_ctx, plain=plain, L=16)
    sv.forward(keep=True)
    rng = np.random.default_rng(11)
    adj = rng.standard_normal(sv.traces_shape()).astype(np.float32)
    nt = case.grid.nt
    e1, s1, w1 = sv.gradient(adj, want_gw=True)
    e2, s2, w2 = sv.gradient(np.float32(2) * adj, want_gw=True)
    bad = np.nonzero(np.any(w2 != 2 * w1, axis=0))[0]
    onset = nt - 1 - int(bad.max()) if bad.size else nt
    print("unit scale", rel_l2(e2, 2 * e1), rel_l2(s2, 2 * s1), "onset", onset, flush=True)
    assert rel_l2(e2, 2 * e1) <= 1e-5 and rel_l2(s2, 2 * s1) <= 1e-5
    assert onset > 500
    big = np.float32(2.0 ** 60) * adj
    e1, s1, w1 = sv.gradient(big, want_gw=True)
    e2, s2, w2 = sv.gradient(np.float32(2) * big, want_gw=True)
    assert np.isfinite(e1).all() and np.abs(e1).max() > 0
    np.testing.assert_array_equal(e2, 2 * e1)
    np.testing.assert_array_equal(s2, 2 * s1)
    np.testing.assert_array_equal(w2, 2 * w1)
    sv.close()


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("air", [0, 3])
def test_gradient_small(oracle_mod, gpu_ctx, plain, air):
    case = configs.small(air=air, nt=240)
    print(f"gradient small plain={plain} air={air}", flush=True)
    sv = make_survey(case, gpu_ctx, plain=plain, ck=7)
    sv.forward(keep=True)
    rng = np.random.default_rng(3)
    adj = rng.standard_normal(sv.traces_shape()).astype(np.float32)
    ge, gs, gw = sv.gradient(adj, want_gw=True)
    fe, fs, fw = oracle_grad(oracle_mod, case, case.eps0, case.sig0, adj, sv.rec_ptr)
    print("eps", rel_l2(ge, fe), cosine(ge, fe), "sig", rel_l2(gs, fs), cosine(gs, fs),
          "w", rel_l2(gw, fw))
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3 and cosine(gs, fs) >= 0.9999
    assert rel_l2(gw, fw) <= 1e-3


def test_dot_product_gpu(gpu_ctx):
    """<F w, e> = <w, F^T e> for the wavelet -> traces map, in FP32 on the GPU."""
    case = configs.small(nt=300)
    sv = make_survey(case, gpu_ctx, ck=17)
    d = sv.forward(keep=True).astype(np.float64)
    rng = np.random.default_rng(5)
    adj = rng.standard_normal(d.shape).astype(np.float32)
    _, _, gw = sv.gradient(adj, want_gw=True)
    lhs = float((d * adj).sum())
    rhs = float((case.wavelets.astype(np.float32).astype(np.float64) * gw).sum())
    print("dot", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


def test_fwi_gradient_l2(oracle_mod, gpu_ctx):
    """Forward + misfit + gradient in one call equals the three-call sequence
    and the oracle's gradient of the L2 misfit."""
    case = configs.small(nt=240, ns=2)
    sv = make_survey(case, gpu_ctx, ck=9)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = svt.forward()
    ge, gs, J = sv.fwi_gradient(dobs)
    d = sv.forward(keep=True)
    e, J2 = sv.misfit(dobs)
    np.testing.assert_allclose(e, d - dobs, rtol=0, atol=1e-6 * np.abs(d).max())
    assert abs(J - J2) <= 1e-9 * abs(J) and abs(J - 0.5 * float((e.astype(np.float64) ** 2).sum())) <= 1e-6 * J
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, e, sv.rec_ptr)
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3


@pytest.mark.parametrize("nt", [240, 241, 256])
def test_filt_gauss_gpu(oracle_mod, gpu_ctx, nt):
    """gjw_filt_gauss (filt_gauss.m) on host and device traces against the
    FP64 oracle; odd, even and power-of-two lengths."""
    import torch
    rng = np.random.default_rng(nt)
    dt, f_low, f_high = 4.2e-11, 1e8, 5e8
    d = rng.standard_normal((5, nt)).astype(np.float32)
    ref = oracle_mod.filt_gauss(d.T.astype(np.float64), dt, f_low, f_high, 8).T
    h = d.copy()
    gpu_ctx.filt_gauss(h, dt, f_low, f_high)
    dd = torch.from_numpy(d).cuda()
    gpu_ctx.filt_gauss(dd, dt, f_low, f_high)
    print("filt_gauss", nt, rel_l2(h, ref), rel_l2(dd.cpu().numpy(), ref))
    assert rel_l2(h, ref) <= 1e-5 and rel_l2(dd.cpu().numpy(), ref) <= 1e-5


def test_lmute_gpu(oracle_mod, gpu_ctx):
    """gjw_lmute (w_lmute.m) against the oracle."""
    rng = np.random.default_rng(3)
    nt, nr = 300, 7
    t = np.arange(nt) * 4e-11
    rx = np.sort(rng.uniform(0, 5, nr))
    d = rng.standard_normal((nr, nt)).astype(np.float32)
    d_o = rng.standard_normal((nr, nt)).astype(np.float32)
    dm, dom, e = oracle_mod.lmute(d.T, d_o.T, t, rx, 0.12, 3.0)
    a, b = d.copy(), d_o.copy()
    eg = gpu_ctx.lmute(a, b, t, rx, 0.12, 3.0)
    for x, y in ((a, dm.T), (b, dom.T), (eg, e.T)):
        assert rel_l2(x, y) <= 1e-6


def test_fwi_gradient_adjoint_filter(oracle_mod, gpu_ctx):
    """gjw_fwi_gradient with the adjoint-source band filter (w_update_e.m:
    69-70): the oracle's gradient for the oracle-filtered L2 residual; graph
    replays equal the eager result bit for bit (the filter runs eagerly
    between the captured sweeps)."""
    import torch
    from gerjoii_amd import _native as N
    case = configs.small(nt=240, ns=2)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = svt.forward()
    f_low, f_high = 1e8, 4e8
    sv = make_survey(case, gpu_ctx, ck=9, L=16)
    sv.set_adj_filter(f_low, f_high)
    ge, gs, J = sv.fwi_gradient(dobs)
    d = sv.forward(keep=True)
    e = (d.astype(np.float64) - dobs.astype(np.float64))
    ef = oracle_mod.filt_gauss(e.T, case.grid.dt, f_low, f_high, 8).T
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, ef, sv.rec_ptr)
    print("filtered", rel_l2(ge, fe), cosine(ge, fe), rel_l2(gs, fs))
    assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
    assert rel_l2(gs, fs) <= 1e-3
    # the filter changes the gradient
    g0 = make_survey(case, gpu_ctx, ck=9, L=16).fwi_gradient(dobs)[0]
    assert rel_l2(g0, ge) > 1e-2
    gr = make_survey(case, gpu_ctx, ck=9, L=16)
    gr.set_adj_filter(f_low, f_high)
    gr.set_option(N.GJW_OPT_GRAPH, 1)
    dt_ = torch.from_numpy(dobs).cuda()
    for rep in range(3):
        a, b, Jr = gr.fwi_gradient(dt_)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a, ge)
        np.testing.assert_array_equal(b, gs)
        assert Jr == J


def _oracle_objective(O, case, eps, dobs, band):
    """w_objs_e.m:22-63 in FP64: forward, band filter of both trace sets,
    0.5 |d - d_o|^2 / |d_o|^2."""
    d = oracle_forward(O, case, eps, case.sig0)
    dt = case.grid.dt
    df = O.filt_gauss(d.T, dt, band[0], band[1], 8)
    do = O.filt_gauss(dobs.T.astype(np.float64), dt, band[0], band[1], 8)
    return 0.5 * float(((df - do) ** 2).sum()) / float((do ** 2).sum())


def test_batch_objective_and_line_search(oracle_mod, gpu_ctx):
    """Batched line-search objectives (gjw_batch_objective, w_objs_e.m) of
    perturbed models against the FP64 oracle, and the parabolic step of
    w_stepparabola_e.m from the device objectives against the step from the
    oracle's objectives at the same trial steps."""
    from gerjoii_amd import linesearch as LS
    case = configs.small(nt=240, ns=2)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = svt.forward()
    band = (1e8, 4e8)
    ge = make_survey(case, gpu_ctx).fwi_gradient(dobs)[0]
    g = ge / np.abs(ge).max()
    ks = [-0.01, 0.0, 0.02]
    svs = []
    for k in ks:
        sv = S.Survey(case.grid, gpu_ctx)
        sv.set_model(LS.perturb_eps(case.eps0, g, k), case.sig0)
        sv.set_shots(case.src, case.wavelets, case.receivers)
        sv.set_adj_filter(*band)
        svs.append(sv)
    J = S.Batch(svs).objective([dobs] * len(ks))
    ref = [_oracle_objective(oracle_mod, case, LS.perturb_eps(case.eps0, g, k), dobs, band)
           for k in ks]
    print("objectives", J, ref)
    for a, b in zip(J, ref):
        assert abs(a - b) <= 1e-4 * abs(b)
    Ee0 = J[1]
    step, ke, Ee = LS.stepparabola_e(gpu_ctx, case.grid, case.src, case.wavelets, case.receivers,
                                     case.eps0, case.sig0, g, dobs, Ee0, k_e=[1e-3, 0.05],
                                     keprct_=0.8, keprct__=0.8, eps_max=30.0, eps_min=0.5,
                                     nparabo=3, band=band)
    Eref = [_oracle_objective(oracle_mod, case, LS.perturb_eps(case.eps0, g, k), dobs, band)
            for k in ke]
    step_ref = LS.parabola_step(ke, Eref)
    print("line search", step, step_ref, ke, Ee, Eref)
    np.testing.assert_allclose(Ee, Eref, rtol=1e-4)
    assert ke.max() - ke.min() > 1e-3  # trial steps wide enough for a curved fit
    assert abs(step - step_ref) <= 1e-2 * max(abs(step_ref), 1e-12)


@pytest.mark.parametrize("fuse", [8, 4])
def test_batch_objective_fused_depths(oracle_mod, gpu_ctx, fuse):
    """gjw_batch_objective with the 8-step multi-survey forward (and the
    4-step one) on two surveys of uneven trace length, so that the shared
    launches also take the 4-, 2- and 1-step tails: each objective against the
    FP64 oracle."""
    cases = [configs.small(nt=240, ns=2), configs.small(nt=203, ns=2, seed=9)]
    band = (1e8, 4e8)
    svs, dobs, ref = [], [], []
    for case in cases:
        svt = make_survey(case, gpu_ctx)
        svt.set_model(case.eps_true, case.sig_true)
        d = svt.forward()
        sv = make_survey(case, gpu_ctx)
        sv.set_option(6, fuse)
        sv.set_adj_filter(*band)
        svs.append(sv)
        dobs.append(d)
        ref.append(_oracle_objective(oracle_mod, case, case.eps0, d, band))
    J = S.Batch(svs).objective(dobs)
    print("fused", fuse, J, ref, flush=True)
    for a, b in zip(J, ref):
        assert abs(a - b) <= 1e-4 * abs(b)


def test_differentiate_cube_gpu(oracle_mod, gpu_ctx):
    rng = np.random.default_rng(2)
    for nt in (7, 8):
        u = np.asfortranarray(rng.standard_normal((5, 4, nt)).astype(np.float32))
        ref = oracle_mod.differentiate_cube(u.astype(np.float64), 0.1)
        v = u.copy(order="F")
        gpu_ctx.differentiate_cube(v, 0.1)
        assert rel_l2(v, ref) <= 1e-6


def test_errors_are_reported(gpu_ctx):
    from gerjoii_amd import _native as N
    case = configs.small()
    sv = S.Survey(case.grid, gpu_ctx)
    with pytest.raises(N.GjwError, match="set_model"):
        sv.forward()
    sv.set_model(case.eps0, case.sig0)
    bad = case.src.copy()
    bad[0] = (1, 5)  # on the never-updated outer ring
    with pytest.raises(N.GjwError, match="source"):
        sv.set_shots(bad, case.wavelets, case.receivers)
    # a grid past the step kernels' 32-bit offsets is refused when the survey
    # is created (6000 x 6000 cells: 38 M padded cells per field)
    import dataclasses
    big = dataclasses.replace(case.grid, m=6000, n=6000)
    with pytest.raises(N.GjwError, match="exceeds"):
        S.Survey(big, gpu_ctx)


@pytest.mark.parametrize("which", ["small", "c1"])
def test_gradient_ref_mode(oracle_mod, gpu_ctx, which):
    """w_grad.m semantics (negated sigma, receivers overwritten, correlation
    with d_t u for eps and u for sigma) against the FP64 restatement: a small
    grid with sigma clipped, and the 200 x 100, nt = 1500, 4-shot case at full
    size, whose backward growth exp(max sigma T / (eps eps0)) is e^2.8 (REF's
    negated-sigma run stays well inside FP32 there)."""
    O = oracle_mod
    if which == "small":
        case = configs.small(nt=200, ns=2)
        case.sig0[:] = np.clip(case.sig0, 0, 5e-3)  # keep the -sigma growth modest
    else:
        case = configs.c1()
    print(f"REF {which}", flush=True)
    sv = make_survey(case, gpu_ctx, plain=True)
    sv.forward(keep=True)
    rng = np.random.default_rng(8)
    adj = rng.standard_normal(sv.traces_shape()).astype(np.float32)
    from gerjoii_amd import _native as N
    import ctypes as C
    ge = np.zeros((case.grid.m, case.grid.n), order="F")
    gs = np.zeros_like(ge)
    am = np.zeros((case.grid.m, case.grid.n), dtype=np.float32, order="F")
    N.check(N.load().gjw_gradient(sv.h, N.ptr(adj), N.GJW_GRAD_REF, N.ptr(ge), N.ptr(gs), None,
                                  N.ptr(am)), "gjw_gradient REF")
    g = oracle_grid(O, case.grid)
    ep, sp = O.exp2pml(g, case.eps0, case.sig0)
    re = np.zeros_like(ge)
    rs = np.zeros_like(ge)
    for q in range(len(case.src)):
        _, _, u = O.forward(g, ep, sp, case.src[q], case.wavelets[q], case.receivers[q], store_u=True)
        e = adj[sv.rec_ptr[q]:sv.rec_ptr[q + 1]].T.astype(np.float64)
        a, amax = O.grad_ref(g, ep, sp, case.src[q], case.receivers[q], e, O.differentiate_cube(u, g.dt))
        b, _ = O.grad_ref(g, ep, sp, case.src[q], case.receivers[q], e, u)
        re += a
        rs += b
    print("REF eps", rel_l2(ge, re), cosine(ge, re), "sig", rel_l2(gs, rs), "amax", rel_l2(am, amax))
    assert rel_l2(ge, re) <= 1e-3 and cosine(ge, re) >= 0.9999
    assert rel_l2(gs, rs) <= 1e-3 and cosine(gs, rs) >= 0.9999
    assert rel_l2(am, amax) <= 1e-3


def test_autograd_function(oracle_mod, gpu_ctx):
    """torch.autograd.Function backward == gjw_gradient == FP64 oracle, and
    central finite differences through the FP32 GPU forward agree on the
    strongest cells within FP32 noise."""
    import torch
    from gerjoii_amd.autograd import traces
    case = configs.small(nt=240, ns=2)
    sv = make_survey(case, gpu_ctx, ck=11)
    sv_true = make_survey(case, gpu_ctx)
    sv_true.set_model(case.eps_true, case.sig_true)
    dobs = torch.tensor(sv_true.forward(), device="cuda")
    eps = torch.tensor(case.eps0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(case.sig0, dtype=torch.float64, requires_grad=True)
    d = traces(eps, sig, sv)
    loss = 0.5 * ((d.double() - dobs.double()) ** 2).sum()
    loss.backward()
    e = (d - dobs).detach().cpu().numpy()
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, e, sv.rec_ptr)
    assert rel_l2(eps.grad.numpy(), fe) <= 1e-3 and rel_l2(sig.grad.numpy(), fs) <= 1e-3
    # finite differences (FP32 forward, FP64 misfit) on the 4 strongest interior cells
    ge = eps.grad.numpy()
    inner = np.zeros_like(ge, dtype=bool)
    inner[2:-2, 2:-2] = True
    idx = np.argsort(np.where(inner, np.abs(ge), 0).ravel())[-4:]
    for k in idx:
        a, b = np.unravel_index(k, ge.shape)
        h = 2e-3 * case.eps0[a, b]
        vals = []
        for sgn in (1, -1):
            E = case.eps0.copy()
            E[a, b] += sgn * h
            sv.set_model(E, case.sig0)
            dd = sv.forward().astype(np.float64)
            vals.append(0.5 * ((dd - dobs.cpu().numpy()) ** 2).sum())
        fd = (vals[0] - vals[1]) / (2 * h)
        print("fd", (a, b), fd, ge[a, b])
        assert abs(fd - ge[a, b]) <= 0.05 * abs(ge[a, b])


def test_autograd_two_nodes_one_survey(oracle_mod, gpu_ctx):
    """Two traces() calls on one survey (two models) before either backward:
    each node's gradient is the gradient of its own model (the first node
    finds the survey moved on and rebuilds its checkpoints), equal to the
    FP64 oracle's to 1e-3, and to a fresh survey's gradient bit for bit."""
    import torch
    from gerjoii_amd.autograd import traces
    case = configs.small(nt=240, ns=2)
    sv = make_survey(case, gpu_ctx, ck=11, L=16)
    rng = np.random.default_rng(21)
    m2 = case.eps0 * (1 + 0.05 * rng.standard_normal(case.eps0.shape))
    e1 = torch.tensor(case.eps0, dtype=torch.float64, requires_grad=True)
    e2 = torch.tensor(m2, dtype=torch.float64, requires_grad=True)
    s1 = torch.tensor(case.sig0, dtype=torch.float64, requires_grad=True)
    s2 = torch.tensor(case.sig0, dtype=torch.float64, requires_grad=True)
    d1 = traces(e1, s1, sv)
    d2 = traces(e2, s2, sv)
    w = torch.tensor(rng.standard_normal(tuple(d1.shape)), dtype=torch.float32, device="cuda")
    (d1 * w).sum().backward()
    (d2 * w).sum().backward()
    adj = w.cpu().numpy()
    for eps, E, S_ in ((e1, case.eps0, s1), (e2, m2, s2)):
        fe, fs, _ = oracle_grad(oracle_mod, case, E, case.sig0, adj, sv.rec_ptr)
        print("two nodes", rel_l2(eps.grad.numpy(), fe), rel_l2(S_.grad.numpy(), fs))
        assert rel_l2(eps.grad.numpy(), fe) <= 1e-3 and rel_l2(S_.grad.numpy(), fs) <= 1e-3
        fresh = make_survey(case, gpu_ctx, ck=11, L=16)
        fresh.set_model(E, case.sig0)
        fresh.forward(keep=True)
        ge, gs = fresh.gradient(adj)
        np.testing.assert_array_equal(eps.grad.numpy(), ge)
        np.testing.assert_array_equal(S_.grad.numpy(), gs)
        fresh.close()
    sv.close()


@pytest.mark.parametrize("fuse,ck,adj_fuse", [(2, 7, 0), (2, 10, 0), (4, 7, 4), (4, 10, 4), (4, 10, 0),
                                              (8, 10, 8), (8, 21, 8), (8, 21, 6), (8, 10, 6)])
def test_two_step_kernel_matches_one_step(gpu_ctx, ck, fuse, adj_fuse):
    """The temporally blocked forward (two, four or eight steps per launch,
    overlapped row tiles) and adjoint (four, six or eight) equal the one-step
    kernels: traces and gradient, odd nt, odd and even checkpoint intervals,
    several tile widths."""
    from gerjoii_amd import _native as N
    case = configs.small(m=130, n=150, P=12, nt=241, ns=2)
    rng = np.random.default_rng(4)
    ref = make_survey(case, gpu_ctx, ck=ck)
    ref.set_option(N.GJW_OPT_FUSE, 1)
    d1 = ref.forward(keep=True)
    adj = rng.standard_normal(d1.shape).astype(np.float32)
    g1 = ref.gradient(adj)
    for L in (5, 16, 64):
        sv = make_survey(case, gpu_ctx, ck=ck, L=L)
        sv.set_option(N.GJW_OPT_FUSE, fuse)
        sv.set_option(N.GJW_OPT_ADJ_FUSE, adj_fuse)
        d2 = sv.forward(keep=True)
        g2 = sv.gradient(adj)
        print("L", L, rel_l2(d2, d1), rel_l2(g2[0], g1[0]), rel_l2(g2[1], g1[1]))
        assert rel_l2(d2, d1) <= 1e-6
        assert rel_l2(g2[0], g1[0]) <= 1e-5 and rel_l2(g2[1], g1[1]) <= 1e-5


def _borehole_case(nt=420):
    """Deep source and a vertical receiver array through the whole model depth,
    so that every row tile's interior update reaches a receiver within nt."""
    case = configs.small(m=170, n=130, P=12, nt=nt, ns=2)
    g = case.grid
    src = np.array([S.src_index(g, 60, 40), S.src_index(g, 110, 90)], dtype=np.int32)
    rz = np.arange(2, g.m - 2, 4)
    rec = S.rec_Mw(g, rz, np.full(rz.size, 70))
    case.src = src
    case.receivers = [rec.copy(), rec.copy()]
    return case


def test_borehole_forward_and_gradient(oracle_mod, gpu_ctx):
    """Interior (fast-path) tiles against the FP64 oracle, one to eight steps per launch."""
    from gerjoii_amd import _native as N
    case = _borehole_case()
    ref = oracle_forward(oracle_mod, case, case.eps0, case.sig0)
    rng = np.random.default_rng(6)
    adj = rng.standard_normal(ref.shape).astype(np.float32)
    rp = np.concatenate([[0], np.cumsum([len(r) for r in case.receivers])])
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, adj, rp)
    for fuse, adj_fuse in ((1, 0), (2, 0), (4, 4), (4, 6), (8, 6), (8, 8)):
        for L in (8, 32):
            sv = make_survey(case, gpu_ctx, ck=13, L=L)
            sv.set_option(N.GJW_OPT_FUSE, fuse)
            sv.set_option(N.GJW_OPT_ADJ_FUSE, adj_fuse)
            d = sv.forward(keep=True)
            ge, gs = sv.gradient(adj)
            print("fuse", fuse, adj_fuse, "L", L, "traces", rel_l2(d, ref), "eps", rel_l2(ge, fe),
                  "sig", rel_l2(gs, fs))
            assert rel_l2(d, ref) <= 1e-4
            assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
            assert rel_l2(gs, fs) <= 1e-3


@pytest.mark.parametrize("mode", [1, 2])
def test_snapshot_checkpoints(oracle_mod, gpu_ctx, mode):
    """Checkpoints stored as BF16 (1) / FP16 (2) snapshots with per-block
    power-of-two scaling: the device gradient equals the oracle's with the same
    snapshot rounding.  Its distance to the exact gradient is the effect of the
    format (measured on this case: FP16 3.4e-3 / 5.9e-4, BF16 2.3e-2 / 4.6e-3
    for eps / sigma), bounded here loosely."""
    from gerjoii_amd import _native as N
    case = configs.small(air=3)
    K = 19
    rng = np.random.default_rng(8)
    sv = make_survey(case, gpu_ctx, ck=K)
    sv.set_option(N.GJW_OPT_SNAPSHOT, mode)
    d = sv.forward(keep=True)
    adj = rng.standard_normal(d.shape).astype(np.float32)
    ge, gs = sv.gradient(adj)
    rp = sv.rec_ptr
    fe_q, fs_q, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, adj, rp, snap=mode, K=K)
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, adj, rp, K=K)
    print("snap", mode, "vs same-rounding oracle", rel_l2(ge, fe_q), rel_l2(gs, fs_q),
          "vs exact", rel_l2(ge, fe), rel_l2(gs, fs), cosine(ge, fe))
    assert rel_l2(ge, fe_q) <= 3e-4 and rel_l2(gs, fs_q) <= 3e-4
    assert rel_l2(ge, fe) > 0
    bound, cos_min = (1e-2, 0.9999) if mode == 2 else (5e-2, 0.999)
    assert rel_l2(ge, fe) <= bound and rel_l2(gs, fs) <= bound and cosine(ge, fe) >= cos_min


def test_source_recorder(oracle_mod, gpu_ctx):
    """gjw_forward's jz output equals the oracle's Jz (w_solve.m:104): the
    field at the source cell after each step, for sources in PML-row tiles and
    in interior tiles, one- to eight-step kernels and the plain path."""
    from gerjoii_amd import _native as N
    for case in (configs.small(air=3), _borehole_case(nt=200)):
        g = oracle_grid(oracle_mod, case.grid)
        ep, sp = oracle_mod.exp2pml(g, case.eps0, case.sig0)
        ref = np.stack([oracle_mod.forward(g, ep, sp, case.src[q], case.wavelets[q],
                                           case.receivers[q])[1] for q in range(len(case.src))])
        for plain, fuse in ((False, 8), (False, 4), (False, 2), (False, 1), (True, 1)):
            sv = make_survey(case, gpu_ctx, plain=plain)
            sv.set_option(N.GJW_OPT_FUSE, fuse)
            d, jz = sv.forward(jz=True)
            print("jz plain", plain, "fuse", fuse, rel_l2(jz, ref))
            assert jz.shape == ref.shape and np.abs(ref).max() > 0
            assert rel_l2(jz, ref) <= 1e-4


@pytest.mark.parametrize("snap,misfit,per_shot", [(0, "L2", False), (1, "L2", False),
                                                  (0, "env_L2", False), (0, "L2", True)])
def test_fwi_gradient_graph_replays(gpu_ctx, snap, misfit, per_shot):
    """GJW_OPT_GRAPH: every replay of the captured batched FWI gradient equals
    the eager result bit for bit, with eager work and synchronisations between
    the replays, and replays follow a change of the model (values, not
    addresses, change)."""
    import torch
    from gerjoii_amd import _native as N
    case = configs.small(nt=150, ns=3)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = torch.from_numpy(svt.forward()).cuda()
    # a fixed tile width: with autotuning, each survey's first sweep tries
    # its own adjoint widths, and results differ with the tiling in the last bits
    eager = make_survey(case, gpu_ctx, ck=7, L=16)
    graph = make_survey(case, gpu_ctx, ck=7, L=16)
    for sv in (eager, graph):
        sv.set_option(N.GJW_OPT_SNAPSHOT, snap)
        sv.set_misfit(misfit)
        sv.set_shot_gradients(per_shot)
    graph.set_option(N.GJW_OPT_GRAPH, 1)
    models = [(case.eps0, case.sig0), (case.eps0 * 1.05, case.sig0 * 0.9)]
    for eps, sig in models:
        eager.set_model(eps, sig)
        graph.set_model(eps, sig)
        ge0, gs0, J0 = eager.fwi_gradient(dobs)
        for rep in range(4):
            ge, gs, J = graph.fwi_gradient(dobs)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(ge, ge0)
            np.testing.assert_array_equal(gs, gs0)
            assert J == J0
            svt.forward()  # eager work on another survey between replays
            torch.cuda.synchronize()
        print("graph replays ok, J", J0)


@pytest.mark.parametrize("kind", ["duplicate_receivers", "tiny_grid"])
def test_receiver_fallback_and_tiny_grids(oracle_mod, gpu_ctx, kind):
    """Two receivers of one shot in one cell (the kernels' per-receiver
    loop instead of the per-lane lookup), and a grid smaller than one row
    tile of the deepest kernels: traces and gradient against the FP64 oracle
    with the eight-step forward and the six- and eight-step adjoints."""
    from gerjoii_amd import _native as N
    if kind == "duplicate_receivers":
        case = configs.small(nt=240, ns=2)
        r = case.receivers[0]
        case.receivers = [np.concatenate([r, r[3:5]]), case.receivers[1]]
    else:
        case = configs.small(m=12, n=30, P=6, nt=200, ns=2, nr=5)
    ref = oracle_forward(oracle_mod, case, case.eps0, case.sig0)
    rng = np.random.default_rng(9)
    adj = rng.standard_normal(ref.shape).astype(np.float32)
    rp = np.concatenate([[0], np.cumsum([len(r) for r in case.receivers])])
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, adj, rp)
    for adj_fuse in (6, 8):
        sv = make_survey(case, gpu_ctx, ck=16, L=8)
        sv.set_option(N.GJW_OPT_FUSE, 8)
        sv.set_option(N.GJW_OPT_ADJ_FUSE, adj_fuse)
        d = sv.forward(keep=True)
        ge, gs = sv.gradient(adj)
        print(kind, adj_fuse, "traces", rel_l2(d, ref), "eps", rel_l2(ge, fe), "sig", rel_l2(gs, fs))
        assert rel_l2(d, ref) <= 1e-4
        assert rel_l2(ge, fe) <= 1e-3 and cosine(ge, fe) >= 0.9999
        assert rel_l2(gs, fs) <= 1e-3


def test_adjoint_width_tuned_in_context(gpu_ctx):
    """The first gradient sweep of an autotuned survey gives its four-step
    adjoint launches the candidate tile widths in turn and keeps the fastest;
    the gradients of that sweep and of later ones agree with a fixed-width
    survey to rounding."""
    import torch
    case = configs.small(m=150, n=260, nt=400, ns=2)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = torch.from_numpy(svt.forward()).cuda()
    ref = make_survey(case, gpu_ctx, L=16).fwi_gradient(dobs)
    sv = make_survey(case, gpu_ctx)
    assert sv.tile(8) == 0  # six-step adjoint, the default depth
    for call in range(2):
        ge, gs, J = sv.fwi_gradient(dobs)
        print("call", call, "adjoint L", sv.tile(8), rel_l2(ge, ref[0]), rel_l2(gs, ref[1]))
        assert sv.tile(8) in (32, 48, 64, 96, 128)
        assert rel_l2(ge, ref[0]) <= 1e-5 and rel_l2(gs, ref[1]) <= 1e-5
        assert abs(J - ref[2]) <= 1e-6 * abs(ref[2])


def test_adjoint_width_pinned(gpu_ctx):
    """GJW_OPT_ADJ_TILE fixes the adjoint's tile width: the survey reports it,
    its gradients repeat bit for bit from the first sweep on, and they agree
    with an autotuned survey's to rounding; a width out of range is refused."""
    import torch
    from gerjoii_amd import _native as N
    case = configs.small(m=150, n=260, nt=400, ns=2)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = torch.from_numpy(svt.forward()).cuda()
    tuned = make_survey(case, gpu_ctx).fwi_gradient(dobs)
    sv = make_survey(case, gpu_ctx)
    sv.set_option(N.GJW_OPT_ADJ_TILE, 64)
    first = sv.fwi_gradient(dobs)
    assert sv.tile(8) == 64
    again = sv.fwi_gradient(dobs)
    np.testing.assert_array_equal(again[0], first[0])
    np.testing.assert_array_equal(again[1], first[1])
    assert rel_l2(first[0], tuned[0]) <= 1e-5 and rel_l2(first[1], tuned[1]) <= 1e-5
    with pytest.raises(RuntimeError):
        sv.set_option(N.GJW_OPT_ADJ_TILE, -1)


@pytest.mark.parametrize("nt", [240, 241])
def test_hilbert_misfits(oracle_mod, gpu_ctx, nt):
    """Envelope / phase misfits on the device (FP32 FFTs) against the oracle's
    w_bozdag.m / w_E.m in FP64, even and odd nt; then the FWI gradient with
    the env_L2 misfit against the oracle's gradient for the oracle's e."""
    case = configs.small(nt=nt, ns=2)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = svt.forward()
    sv = make_survey(case, gpu_ctx, ck=11)
    d = sv.forward(keep=True)
    rp = sv.rec_ptr
    # samples whose envelopes (synthetic and observed) are above 1e-3 of the
    # trace's largest: the phase and log-envelope sources divide by |A(d)|^2 and
    # take log |A|, so below that floor FP32 rounding dominates them
    env = np.abs(oracle_mod.hilbert_analytic(d.T.astype(np.float64))).T
    envo = np.abs(oracle_mod.hilbert_analytic(dobs.T.astype(np.float64))).T
    good = (env >= 1e-3 * env.max(axis=1, keepdims=True)) & \
           (envo >= 1e-3 * envo.max(axis=1, keepdims=True))
    for kind in ("env_L2", "phase_L2", "env_L22", "env_LOG"):
        e, J = sv.misfit(dobs, kind=kind)
        es, Js = [], 0.0
        for q in range(len(case.src)):
            a, b = rp[q], rp[q + 1]
            eq, de, doe = oracle_mod.w_bozdag(dobs[a:b].T, d[a:b].T, kind)
            es.append(eq.T)
            Js += oracle_mod.w_E(doe, de, obj_fnc=kind)
        eref = np.concatenate(es, axis=0)
        eg = rel_l2(e[good], eref[good])
        print(kind, "e rel L2", rel_l2(e, eref), "above the floor", eg, "J", J, Js)
        assert np.isfinite(e).all() and np.abs(eref).max() > 0
        if kind in ("env_L2", "env_L22"):
            assert rel_l2(e, eref) <= 1e-4 and abs(J - Js) <= 1e-5 * abs(Js)
        elif kind == "phase_L2":  # (the phase below the floor also enters J)
            assert eg <= 1e-3 and abs(J - Js) <= 1e-4 * abs(Js)
    # env_LOG (log |A| and 1/|A|^2) on traces with a noise floor, through
    # gjw_bozdag with caller-given synthetic traces
    rng = np.random.default_rng(3)
    dn = (d + 2e-3 * np.abs(d).max() * rng.standard_normal(d.shape)).astype(np.float32)
    on = (dobs + 2e-3 * np.abs(dobs).max() * rng.standard_normal(d.shape)).astype(np.float32)
    for kind in ("env_LOG", "phase_L2", "env_L2"):
        e, J = sv.bozdag(on, dn, kind=kind)
        es, Js = [], 0.0
        for q in range(len(case.src)):
            a, b = rp[q], rp[q + 1]
            eq, de, doe = oracle_mod.w_bozdag(on[a:b].T, dn[a:b].T, kind)
            es.append(eq.T)
            Js += oracle_mod.w_E(doe, de, obj_fnc=kind)
        eref = np.concatenate(es, axis=0)
        print(kind, "noisy traces: e rel L2", rel_l2(e, eref), "J", J, Js)
        assert rel_l2(e, eref) <= 1e-3 and abs(J - Js) <= 1e-5 * abs(Js)
    sv.set_misfit("env_L2")
    ge, gs, J = sv.fwi_gradient(dobs)
    d = sv.forward(keep=True)
    e, _ = sv.misfit(dobs, kind="env_L2")
    fe, fs, _ = oracle_grad(oracle_mod, case, case.eps0, case.sig0, e, rp, K=11)
    assert rel_l2(ge, fe) <= 1e-3 and rel_l2(gs, fs) <= 1e-3 and cosine(ge, fe) >= 0.9999


def _uneven_cases():
    """Three surveys of different grids, trace lengths, shot counts and
    receiver counts (small versions of the uneven-batch workload)."""
    a = configs.small(m=40, n=60, P=8, nt=150, ns=2, air=2)
    b = configs.small(m=70, n=90, P=10, nt=233, ns=3)
    c = configs.small(m=55, n=120, P=9, nt=191, ns=1, air=3)
    return [a, b, c]


@pytest.mark.parametrize("misfit,fuse,K", [("L2", 0, 12), ("env_L2", 0, 12), ("L2", 8, 12),
                                           ("L2", 8, 0), ("L2", 4, 0)])
def test_batch_uneven_surveys(gpu_ctx, misfit, fuse, K):
    """Surveys of uneven grid size and trace length in shared launches give
    each survey's own FWI gradient and misfit (same checkpoint interval):
    with the autotuned depths, and with eight-step forwards and six-step
    adjoints (fuse 8; K 12 splits segments into 8 + 4 and 6 + 6 steps, K 0
    takes the default interval in whole launches of both depths)."""
    from gerjoii_amd import _native as N
    cases = _uneven_cases()
    dobs, single, batch_sv = [], [], []
    for case in cases:
        svt = make_survey(case, gpu_ctx)
        svt.set_model(case.eps_true, case.sig_true)
        dobs.append(svt.forward())
        for lst in (single, batch_sv):
            sv = make_survey(case, gpu_ctx, ck=K)
            sv.set_misfit(misfit)
            if fuse:
                sv.set_option(N.GJW_OPT_FUSE, fuse)
            lst.append(sv)
    ref = [sv.fwi_gradient(d) for sv, d in zip(single, dobs)]
    bt = S.Batch(batch_sv)
    ge, gs, J = bt.fwi_gradient(dobs)
    ms, cells = bt.last_timing()
    print("batch ms", ms, "cells", cells)
    for k in range(len(cases)):
        print("survey", k, rel_l2(ge[k], ref[k][0]), rel_l2(gs[k], ref[k][1]), J[k], ref[k][2])
        assert rel_l2(ge[k], ref[k][0]) <= 1e-5 and rel_l2(gs[k], ref[k][1]) <= 1e-5
        assert abs(J[k] - ref[k][2]) <= 1e-6 * abs(ref[k][2])
    ge2, gs2, J2 = bt.fwi_gradient(dobs)  # a second call reuses the buffers
    for k in range(len(cases)):
        np.testing.assert_array_equal(ge2[k], ge[k])


def test_batch_graph_replays(gpu_ctx):
    """A batch of uneven surveys with mixed misfits, replayed from captured
    forward / adjoint sweep graphs (the Hilbert misfits run between them),
    equals the eager batch bit for bit, call after call."""
    import torch
    from gerjoii_amd import _native as N
    cases = _uneven_cases()
    dobs, eager, graph = [], [], []
    for case, misfit in zip(cases, ("env_L2", "L2", "phase_L2")):
        svt = make_survey(case, gpu_ctx)
        svt.set_model(case.eps_true, case.sig_true)
        dobs.append(svt.forward())
        for lst, on in ((eager, 0), (graph, 1)):
            sv = make_survey(case, gpu_ctx, L=16)  # fixed widths: results depend on the tiling
            sv.set_misfit(misfit)
            sv.set_option(N.GJW_OPT_GRAPH, on)
            lst.append(sv)
    ge0, gs0, J0 = S.Batch(eager).fwi_gradient(dobs)
    bg = S.Batch(graph)
    for rep in range(3):
        ge, gs, J = bg.fwi_gradient(dobs)
        torch.cuda.synchronize()
        for k in range(len(cases)):
            np.testing.assert_array_equal(ge[k], ge0[k])
            np.testing.assert_array_equal(gs[k], gs0[k])
        np.testing.assert_array_equal(J, J0)
