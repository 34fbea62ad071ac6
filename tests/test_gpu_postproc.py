"""GPU parity of the gradient post-processing (gaussi.m, image_gaussian.m,
normali.m, w_update_e_.m's reduction, w_image_e.m's update) and of the
per-shot gradients it starts from, through the C ABI, against the FP64 oracle.

Tolerances (per test): the post-processing is FP64 on both sides, so
elementwise results agree to 1e-13 of the image maximum (exp differs by an
ulp between libraries), normali to the bit (max and division are exact), and
the FFT filter to 1e-12; per-shot gradients sum to the shot-summed gradient
to 1e-12 and meet the gradient bar (relative L2 <= 1e-3, cosine >= 0.9999)
against the oracle shot by shot.
"""
import dataclasses

import numpy as np
import pytest

from gerjoii_amd import configs
from gerjoii_amd import postproc as PP

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _images(rng, m, n, cnt):
    return np.asfortranarray(rng.standard_normal((m, n, cnt)))


def test_gaussi_gpu(oracle_mod, gpu_ctx):
    """gjw_gaussi on host and device images, three sources on a non-square
    grid, against gaussi.m times the image."""
    import torch
    O = oracle_mod
    rng = np.random.default_rng(5)
    m, n, cnt = 37, 53, 3
    X = np.arange(n) * 0.05
    Y = np.arange(m) * 0.05 + 0.3
    s = np.array([[1, 1], [20, 40], [37, 53]], dtype=np.int32)
    lo = np.array([0.2, 0.45, 1.1])
    g = _images(rng, m, n, cnt)
    ref = np.stack([O.gaussi(X, Y, s[k], lo[k]) * g[:, :, k] for k in range(cnt)], axis=2)
    h = g.copy(order="F")
    gpu_ctx.gaussi(h, X, Y, s, lo)
    dev = torch.from_numpy(np.ascontiguousarray(np.transpose(g, (2, 1, 0)))).cuda()
    gpu_ctx.gaussi(dev, X, Y, s, lo)
    dh = np.transpose(dev.cpu().numpy(), (2, 1, 0))
    print("gaussi", rel_max(h, ref), rel_max(dh, ref), flush=True)
    assert rel_max(h, ref) <= 1e-13 and rel_max(dh, ref) <= 1e-13
    for k in range(cnt):  # the taper vanishes on the source cell
        assert h[s[k, 0] - 1, s[k, 1] - 1, k] == 0.0


@pytest.mark.parametrize("m,n", [(37, 53), (64, 32), (2, 2)])
@pytest.mark.parametrize("kind", ["LOW_PASS", "HI_PASS"])
def test_image_gaussian_gpu(oracle_mod, gpu_ctx, m, n, kind):
    """gjw_image_gaussian (batched D2Z/Z2D) against image_gaussian.m's
    complex fft2 / ifft2, padded and power-of-two sizes, host and in-place on
    device."""
    import torch
    rng = np.random.default_rng(m * n)
    cnt, ax, az = 3, 0.08, 0.13
    g = _images(rng, m, n, cnt)
    ref = np.stack([oracle_mod.image_gaussian(g[:, :, k], ax, az, kind) for k in range(cnt)],
                   axis=2)
    out = gpu_ctx.image_gaussian(g, ax, az, kind)
    dev = torch.from_numpy(np.ascontiguousarray(np.transpose(g, (2, 1, 0)))).cuda()
    gpu_ctx.image_gaussian(dev, ax, az, kind, out=dev)
    dh = np.transpose(dev.cpu().numpy(), (2, 1, 0))
    print("image_gaussian", m, n, kind, rel_max(out, ref), rel_max(dh, ref), flush=True)
    assert rel_max(out, ref) <= 1e-12 and rel_max(dh, ref) <= 1e-12


def test_normali_gpu(oracle_mod, gpu_ctx):
    """gjw_normali with the NaN guard and the shape mask, bit for bit: one
    image with a NaN (zeroed), one all-zero image, large images (several
    partial maxima per image)."""
    O = oracle_mod
    rng = np.random.default_rng(9)
    m, n = 300, 210
    g = _images(rng, m, n, 4) * np.array([1.0, 1e-30, 0.0, 1e12])
    g[17, 33, 1] = np.nan
    shape = (rng.random((m, n)) > 0.2).astype(np.float64)
    ref = np.stack([O.normali(g[:, :, k]) for k in range(4)], axis=2)
    ref[:, :, 1] = 0.0  # the NaN guard
    ref = ref * shape[:, :, None]
    out = gpu_ctx.normali(g.copy(order="F"), nan_guard=True, shape=shape)
    np.testing.assert_array_equal(out, ref)
    # without the guard the NaN stays and its image is scaled by the finite max
    plain = gpu_ctx.normali(g.copy(order="F"))
    np.testing.assert_array_equal(plain[:, :, 1], O.normali(g[:, :, 1]))


def test_stack_update_gpu(oracle_mod, gpu_ctx):
    """gjw_stack_update: -sum(step g)/ns, the negative-step-sum rule, and the
    multiplicative model update, against the oracle."""
    O = oracle_mod
    rng = np.random.default_rng(2)
    m, n, cnt = 41, 29, 5
    g = _images(rng, m, n, cnt)
    steps = rng.uniform(0.01, 0.1, cnt)
    eps = np.asfortranarray(rng.uniform(1, 20, (m, n)))
    d_ref, e_ref = O.stack_update(g, steps, model=eps)
    e_dev = eps.copy(order="F")
    d = gpu_ctx.stack_update(g, steps, model=e_dev)
    print("stack", rel_max(d, d_ref), rel_max(e_dev, e_ref), flush=True)
    assert rel_max(d, d_ref) <= 1e-14 and rel_max(e_dev, e_ref) <= 1e-13
    neg = steps.copy()
    neg[0] = -1.0
    np.testing.assert_array_equal(gpu_ctx.stack_update(g, neg), np.zeros((m, n)))
    # the conductivity reduction (w_update_s_.m:38) keeps the mean
    d2 = gpu_ctx.stack_update(g, neg, neg_rule=False)
    assert rel_max(d2, O.stack_update(g, neg, neg_rule=False)) <= 1e-14


@pytest.mark.parametrize("plain", [False, True])
def test_shot_gradients(oracle_mod, gpu_ctx, plain):
    """GJW_OPT_SHOT_GRADS: the per-shot gradients sum to the shot-summed
    gradient and each meets the gradient bar against the oracle's gradient of
    that shot alone."""
    from test_gpu_parity import cosine, make_survey, oracle_grad
    case = configs.small(nt=240)
    sv = make_survey(case, gpu_ctx, plain=plain, ck=7)
    sv.forward(keep=True)
    rng = np.random.default_rng(4)
    adj = rng.standard_normal(sv.traces_shape()).astype(np.float32)
    ge, gs = sv.gradient(adj)
    sv.set_shot_gradients(True)
    sv.forward(keep=True)
    pe, ps = sv.gradient(adj)
    assert pe.shape == (case.grid.m, case.grid.n, sv.ns)
    print("sum of shots", rel_l2(pe.sum(axis=2), ge), rel_l2(ps.sum(axis=2), gs), flush=True)
    assert rel_l2(pe.sum(axis=2), ge) <= 1e-12 and rel_l2(ps.sum(axis=2), gs) <= 1e-12
    for q in range(sv.ns):
        one = dataclasses.replace(case, src=case.src[q:q + 1], wavelets=case.wavelets[q:q + 1],
                                  receivers=case.receivers[q:q + 1])
        lo, hi = sv.rec_ptr[q], sv.rec_ptr[q + 1]
        fe, fs, _ = oracle_grad(oracle_mod, one, case.eps0, case.sig0, adj[lo:hi],
                                np.array([0, hi - lo]))
        print("shot", q, rel_l2(pe[:, :, q], fe), cosine(pe[:, :, q], fe), flush=True)
        assert rel_l2(pe[:, :, q], fe) <= 1e-3 and cosine(pe[:, :, q], fe) >= 0.9999
        assert rel_l2(ps[:, :, q], fs) <= 1e-3


def test_shot_gradients_in_a_batch(gpu_ctx):
    """gjw_batch_fwi_gradient with GJW_OPT_SHOT_GRADS on one of two surveys:
    its per-shot gradients sum to its shot-summed gradient in the same batch,
    and the other survey is unaffected."""
    from gerjoii_amd import survey as S
    from test_gpu_parity import make_survey
    cases = [configs.small(nt=200, ns=3), configs.small(nt=170, ns=2, seed=11)]
    dobs = []
    for case in cases:
        svt = make_survey(case, gpu_ctx)
        svt.set_model(case.eps_true, case.sig_true)
        dobs.append(svt.forward())
    ref = S.Batch([make_survey(c, gpu_ctx, L=16) for c in cases]).fwi_gradient(dobs)
    svs = [make_survey(c, gpu_ctx, L=16) for c in cases]
    svs[0].set_shot_gradients(True)
    got = S.Batch(svs).fwi_gradient(dobs)
    ge_ref, gs_ref = ref[0][0], ref[1][0]
    ge, gs = got[0][0], got[1][0]
    assert ge.shape == (cases[0].grid.m, cases[0].grid.n, 3)
    print("batch shots", rel_l2(ge.sum(axis=2), ge_ref), rel_l2(gs.sum(axis=2), gs_ref), flush=True)
    assert rel_l2(ge.sum(axis=2), ge_ref) <= 1e-12 and rel_l2(gs.sum(axis=2), gs_ref) <= 1e-12
    np.testing.assert_array_equal(got[0][1], ref[0][1])
    np.testing.assert_array_equal(got[1][1], ref[1][1])


def test_joint_update_post(oracle_mod, gpu_ctx):
    """w_update_es.m:105-173 on the device (postproc.w_update_es_post) for
    both gradients of every source, against the oracle, from device per-shot
    gradients."""
    from test_gpu_parity import make_survey
    O = oracle_mod
    case = configs.small(nt=200)
    g = case.grid
    sv = make_survey(case, gpu_ctx)
    sv.set_shot_gradients(True)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    pe, ps, _ = sv.fwi_gradient(svt.forward())
    X = np.arange(g.n) * g.dx
    Y = np.arange(g.m) * g.dx
    s = PP.taper_index(g, case.src)
    lo = PP.source_wavelength(case.eps0, s, case.f0)
    refs = [O.w_update_es_post(pe[:, :, k], ps[:, :, k], X, Y, s[k], lo[k], 0.2, 0.3)
            for k in range(sv.ns)]
    ge, gs = PP.w_update_es_post(gpu_ctx, pe.copy(order="F"), ps.copy(order="F"), g, case.src,
                                 case.eps0, X, Y, case.f0, 0.2, 0.3)
    re = np.stack([r[0] for r in refs], axis=2)
    rs = np.stack([r[1] for r in refs], axis=2)
    print("joint post", rel_max(ge, re), rel_max(gs, rs), flush=True)
    assert rel_max(ge, re) <= 1e-12 and rel_max(gs, rs) <= 1e-12


def test_concurrent_fwi_gradients(gpu_ctx):
    """Surveys on their own contexts, their FWI gradients from concurrent host
    threads: equal to one survey at a time bit for bit; a shared context is
    refused."""
    from gerjoii_amd import survey as S
    from test_gpu_parity import make_survey
    cases = [configs.small(nt=180, ns=2, seed=s) for s in (3, 4, 5)]
    ctxs = [S.Context(0) for _ in cases]
    try:
        svs, dobs = [], []
        for case, cx in zip(cases, ctxs):
            svt = make_survey(case, cx)
            svt.set_model(case.eps_true, case.sig_true)
            dobs.append(svt.forward())
            svs.append(make_survey(case, cx, L=16))
        ref = [sv.fwi_gradient(d) for sv, d in zip(svs, dobs)]
        got = S.concurrent_fwi_gradients(svs, dobs)
        for (a, b, ja), (c, d, jc) in zip(got, ref):
            np.testing.assert_array_equal(a, c)
            np.testing.assert_array_equal(b, d)
            assert ja == jc
        with pytest.raises(ValueError):
            S.concurrent_fwi_gradients([make_survey(cases[0], gpu_ctx), make_survey(cases[1], gpu_ctx)],
                                       dobs[:2])
    finally:
        for cx in ctxs:
            cx.close()


def test_update_pipeline(oracle_mod, gpu_ctx):
    """w_update_e.m:108-145 + w_update_e_.m:38-41 + w_image_e.m:50-51 on the
    device (postproc.w_update_post / w_update_stack) from device per-shot FWI
    gradients, against the oracle chain applied to the same gradients."""
    from test_gpu_parity import make_survey
    O = oracle_mod
    case = configs.small(nt=240)
    g = case.grid
    sv = make_survey(case, gpu_ctx)
    sv.set_shot_gradients(True)
    svt = make_survey(case, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    dobs = svt.forward()
    pe, _, J = sv.fwi_gradient(dobs)
    X = np.arange(g.n) * g.dx
    Y = np.arange(g.m) * g.dx
    ax = az = 0.25
    s = PP.taper_index(g, case.src)
    lo = PP.source_wavelength(case.eps0, s, case.f0)
    ref = np.stack([O.w_update_post(pe[:, :, k], X, Y, s[k], lo[k], ax, az, case.shape)
                    for k in range(sv.ns)], axis=2)
    post = PP.w_update_post(gpu_ctx, pe.copy(order="F"), g, case.src, case.eps0, X, Y, case.f0,
                            ax, az, shape=case.shape)
    print("post", rel_max(post, ref), flush=True)
    assert rel_max(post, ref) <= 1e-12
    steps = np.linspace(0.02, 0.05, sv.ns)
    d_ref, e_ref = O.stack_update(ref, steps, model=case.eps0)
    eps = np.asfortranarray(case.eps0, dtype=np.float64).copy(order="F")
    d = PP.w_update_stack(gpu_ctx, post, steps, model=eps)
    print("update", rel_max(d, d_ref), rel_max(eps, e_ref), flush=True)
    assert rel_max(d, d_ref) <= 1e-12 and rel_max(eps, e_ref) <= 1e-13
