"""The MATLAB-shaped drop-in (gerjoii_amd.w) against the FP64 oracle: the
reference's call sequence w_fwd -> differentiate_cube -> w_grad on struct
dicts, with and without the mute of w_lmute.m, for the L2 and the
Hilbert-based objectives of w_bozdag.m, and w_grad in REF, EXACT and auto
modes.

Tolerances: traces, envelopes and L2 / envelope adjoint sources relative L2
<= 1e-4; instantaneous phases and the phase adjoint source on the samples
whose envelope is above 1e-3 of the gather maximum (below it the phase of
FP32 traces is rounding noise, in the reference as here) <= 1e-3, and the
log-envelope adjoint source where both envelopes are above 1e-2 of their
maxima <= 1e-3; gradients
relative L2 <= 1e-3 and cosine >= 0.9999."""
import warnings

import numpy as np
import pytest

from gerjoii_amd import configs
from gerjoii_amd import w as W

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def cosine(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def oracle_grid(O, g):
    return O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, g.nt, eps0=g.eps0, mu0=g.mu0, R0=g.R0,
                  order=g.pml_order)


def structs(case, q, d_o, obj="L2", mute=False, sig=None):
    """geome_, parame_, finite_, gerjoii_ for shot q, with the field names of
    the reference (w_geom.m, w_load.m, w_lmute.m)."""
    g = case.grid
    nr = len(case.receivers[q])
    geome_ = {"dx": g.dx, "dy": g.dx, "w": {"dt": g.dt, "nt": g.nt}}
    parame_ = {"w": {"epsilon": case.eps0, "sigma": case.sig0 if sig is None else sig,
                     "eps_0": g.eps0, "mu_0": g.mu0, "air": g.air, "t": np.arange(g.nt) * g.dt},
               "natu": {"w": {"d_2d": d_o}}}
    finite_ = {"w": {"pml_w": g.pml, "pml_d": g.pml, "pml_order": g.pml_order, "R_0": g.R0}}
    gerjoii_ = {"w": {"s": case.src[q], "M": case.receivers[q], "wvlet_": case.wavelets[q],
                      "obj_FNC": obj, "MUTE": "yes_MUTE" if mute else "no_MUTE",
                      "v_mute": 0.12, "t_mute": 1.5,
                      "r_real": np.stack([0.1 * np.arange(nr), np.zeros(nr)], axis=1)}}
    return geome_, parame_, finite_, gerjoii_


def oracle_shot(O, case, q, eps, sig):
    og = oracle_grid(O, case.grid)
    ep, sp = O.exp2pml(og, eps, sig)
    d, jz, u = O.forward(og, ep, sp, case.src[q], case.wavelets[q], case.receivers[q], store_u=True)
    return og, ep, sp, d, jz, u


def floor_mask(env, frac=1e-3):
    return env > frac * np.abs(env).max()


def test_w_fwd_l2_traces_recorder_residual(oracle_mod, gpu_ctx):
    O = oracle_mod
    case = configs.small(nt=240, ns=2)
    _, _, _, d_o, _, _ = oracle_shot(O, case, 0, case.eps_true, case.sig_true)
    _, _, _, d, jz, _ = oracle_shot(O, case, 0, case.eps0, case.sig0)
    geome_, parame_, finite_, gerjoii_ = structs(case, 0, d_o.astype(np.float32))
    finite_, gj = W.w_fwd(geome_, parame_, finite_, gerjoii_)
    print("w_fwd L2: d", rel_l2(gj["w"]["d_2d"], d), "Jy", rel_l2(gj["w"]["Jy"], jz), flush=True)
    assert gj["w"]["d_2d"].shape == d.shape and rel_l2(gj["w"]["d_2d"], d) <= 1e-4
    assert rel_l2(gj["w"]["Jy"], jz) <= 1e-4
    assert rel_l2(gj["w"]["e_2d"], d - d_o) <= 1e-4
    assert gj["w"]["obj_FNC"] == "L2"
    # MATLAB value semantics: the caller's structs are not modified
    assert "e_2d" not in gerjoii_["w"] and "u_2d" not in gerjoii_["w"]


@pytest.mark.parametrize("obj", ["L2", "env_L2", "phase_L2", "env_L22", "env_LOG"])
def test_w_fwd_mute_and_objectives(oracle_mod, gpu_ctx, obj):
    """w_fwd.m:29-35: solve, w_lmute when MUTE is 'yes_MUTE', then
    [e_2d, d_2d] = w_bozdag(d_o, d_2d, obj_FNC)."""
    O = oracle_mod
    case = configs.small(nt=240, ns=2)
    g = case.grid
    _, _, _, d_o, _, _ = oracle_shot(O, case, 1, case.eps_true, case.sig_true)
    _, _, _, d, _, _ = oracle_shot(O, case, 1, case.eps0, case.sig0)
    geome_, parame_, finite_, gerjoii_ = structs(case, 1, d_o.astype(np.float32), obj, mute=True)
    _, gj = W.w_fwd(geome_, parame_, finite_, gerjoii_)
    rx = gerjoii_["w"]["r_real"][:, 0]
    dm, dom, _ = O.lmute(d, d_o, np.arange(g.nt) * g.dt, rx, 0.12, 1.5)
    e, de, _ = O.w_bozdag(dom, dm, obj)
    if obj == "phase_L2":
        keep = floor_mask(np.abs(O.hilbert_analytic(dm, axis=0)))
        err_e = rel_l2(gj["w"]["e_2d"][keep], e[keep])
        err_d = rel_l2(gj["w"]["d_2d"][keep], de[keep])
        tol = 1e-3
    elif obj == "env_LOG":
        # log of both envelopes: the floor is 1e-2 of the maximum
        keep = floor_mask(np.abs(O.hilbert_analytic(dm, axis=0)), 1e-2) & floor_mask(
            np.abs(O.hilbert_analytic(dom, axis=0)), 1e-2)
        err_e = rel_l2(gj["w"]["e_2d"][keep], e[keep])
        err_d = rel_l2(gj["w"]["d_2d"], de)
        tol = 1e-3
    else:
        err_e = rel_l2(gj["w"]["e_2d"], e)
        err_d = rel_l2(gj["w"]["d_2d"], de)
        tol = 1e-4
    print(f"w_fwd {obj} muted: e {err_e:.2e} d_2d {err_d:.2e}", flush=True)
    assert err_e <= tol and err_d <= tol
    # the mute acted: early samples of the far receivers are attenuated
    assert np.abs(dm[:20, -1]).max() <= np.abs(d[:20, -1]).max() + 1e-30


def test_w_grad_exact_and_ref_chain(oracle_mod, gpu_ctx):
    """w_update_e.m:83-98 / w_update_s.m:77: w_grad(differentiate_cube(u_2d))
    for the permittivity, w_grad(u_2d) for the conductivity; EXACT against
    the oracle's exact adjoint, REF against its restatement of w_grad.m, and
    'auto' taking REF on this low-loss model."""
    O = oracle_mod
    case = configs.small(nt=200, ns=2)
    case.sig0[:] = np.clip(case.sig0, 0, 5e-3)  # keep the -sigma growth of REF modest
    g = case.grid
    og, ep, sp, d, _, u = oracle_shot(O, case, 0, case.eps0, case.sig0)
    d_o = d * 0.9
    geome_, parame_, finite_, gerjoii_ = structs(case, 0, d_o.astype(np.float32))
    _, gj = W.w_fwd(geome_, parame_, finite_, gerjoii_)
    e = gj["w"]["e_2d"].astype(np.float64)
    u_w = gj["w"]["u_2d"]
    dtu = W.differentiate_cube(u_w, g.dt)
    ge, am = W.w_grad(dtu, geome_, finite_, gj, parame_, mode="EXACT")
    gs, _ = W.w_grad(u_w, geome_, finite_, gj, parame_, mode="EXACT")
    fe, fs, _ = O.grad_exact(og, ep, sp, case.src[0], case.wavelets[0], case.receivers[0], e,
                             pml_terms=False)
    re_, rs_ = O.fold_pad(og, fe, fs)
    print("w_grad EXACT eps", rel_l2(ge, re_), "sig", rel_l2(gs, rs_), flush=True)
    assert am is None
    assert rel_l2(ge, re_) <= 1e-3 and cosine(ge, re_) >= 0.9999
    assert rel_l2(gs, rs_) <= 1e-3 and cosine(gs, rs_) >= 0.9999
    ga, ama = W.w_grad(dtu, geome_, finite_, gj, parame_, mode="REF")
    gb, _ = W.w_grad(u_w, geome_, finite_, gj, parame_, mode="REF")
    ra, amax = O.grad_ref(og, ep, sp, case.src[0], case.receivers[0], e, O.differentiate_cube(u, g.dt))
    rb, _ = O.grad_ref(og, ep, sp, case.src[0], case.receivers[0], e, u)
    print("w_grad REF eps", rel_l2(ga, ra), "sig", rel_l2(gb, rb), "a_max", rel_l2(ama, amax),
          flush=True)
    assert rel_l2(ga, ra) <= 1e-3 and rel_l2(gb, rb) <= 1e-3 and rel_l2(ama, amax) <= 1e-3
    assert W.ref_growth_exponent(parame_, geome_) < W.REF_GROWTH_LIMIT
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        gauto, amauto = W.w_grad(dtu, geome_, finite_, gj, parame_)  # auto -> REF, no warning
    np.testing.assert_array_equal(gauto, ga)
    np.testing.assert_array_equal(amauto, ama)


def test_w_grad_auto_takes_exact_where_ref_overflows(oracle_mod, gpu_ctx):
    """A lossy block (sigma 2 S/m) makes REF's backward run grow as
    e^(sigma T / eps eps0) >> e^88: 'auto' warns and returns the EXACT
    gradient, which stays finite and equals mode='EXACT'."""
    case = configs.small(nt=240, ns=1)
    sig = case.sig0.copy()
    sig[10:20, 20:40] = 2.0
    g = case.grid
    d_o = np.zeros((g.nt, len(case.receivers[0])), dtype=np.float32)
    geome_, parame_, finite_, gerjoii_ = structs(case, 0, d_o, sig=sig)
    assert W.ref_growth_exponent(parame_, geome_) > 88
    _, gj = W.w_fwd(geome_, parame_, finite_, gerjoii_)
    dtu = W.differentiate_cube(gj["w"]["u_2d"], g.dt)
    with pytest.warns(RuntimeWarning, match="EXACT"):
        ga, am = W.w_grad(dtu, geome_, finite_, gj, parame_)
    gx, _ = W.w_grad(dtu, geome_, finite_, gj, parame_, mode="EXACT")
    assert am is None and np.isfinite(ga).all()
    np.testing.assert_array_equal(ga, gx)


@pytest.mark.parametrize("obj", ["L2", "env_L2", "phase_L2", "env_L22", "env_LOG"])
def test_w_bozdag_standalone(oracle_mod, gpu_ctx, obj):
    """w_bozdag(d_o, d, obj) with no survey (gjw_w_bozdag): all three outputs
    against the oracle's w_bozdag.m, on odd-length traces with a DC floor
    (no zero envelope)."""
    O = oracle_mod
    rng = np.random.default_rng(4)
    nt, nr = 301, 5
    t = np.linspace(0, 1, nt)[:, None]
    d = 0.5 + np.sin(50 * t + rng.uniform(0, 3, (1, nr))) * np.exp(-((t - 0.5) / 0.2) ** 2)
    d_o = 0.45 + 0.8 * np.sin(50 * t - 0.4 + rng.uniform(0, 3, (1, nr))) * np.exp(-((t - 0.4) / 0.25) ** 2)
    d, d_o = d.astype(np.float32), d_o.astype(np.float32)
    e, de, doe = W.w_bozdag(d_o, d, obj)
    re, rde, rdoe = O.w_bozdag(d_o.astype(np.float64), d.astype(np.float64), obj)
    print(f"w_bozdag {obj}: e {rel_l2(e, re):.1e} d_e {rel_l2(de, rde):.1e} d_oe {rel_l2(doe, rdoe):.1e}")
    assert e.shape == (nt, nr)
    assert rel_l2(e, re) <= 1e-4 and rel_l2(de, rde) <= 1e-5 and rel_l2(doe, rdoe) <= 1e-5
    if obj in ("L2", "env_L2", "env_L22", "env_LOG"):
        assert abs(W.w_E(doe, de, obj_fnc="L2") - O.w_E(rdoe, rde)) <= 1e-4 * O.w_E(rdoe, rde)
