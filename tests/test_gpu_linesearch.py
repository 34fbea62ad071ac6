"""The line-search step sizes on the GPU against the FP64 oracle (the
reference's src/w/w_stepparabola_e.m, w_stepparabola_s.m, w_objs_e.m,
w_objs_s.m, w_pica_e_.m, w_pica_s_.m).

Tolerances: objectives relative 1e-4 (FP32 traces against FP64); Pica steps
relative 1e-3 (a ratio of inner products of a trace difference, whose FP32
rounding is relative to the traces, not to the difference); parabolic steps
within 1% of the step fitted to the oracle's objectives at the same trial
steps."""
import numpy as np
import pytest

from gerjoii_amd import configs
from gerjoii_amd import linesearch as LS
from gerjoii_amd import survey as S

pytestmark = pytest.mark.gpu

BAND = (1e8, 4e8)


def survey(case, ctx, eps=None, sig=None):
    sv = S.Survey(case.grid, ctx)
    sv.set_model(case.eps0 if eps is None else eps, case.sig0 if sig is None else sig)
    sv.set_shots(case.src, case.wavelets, case.receivers)
    return sv


def oracle_gathers(O, case, eps, sig):
    """FP64 traces of every shot, concatenated as nt x (all receivers)."""
    g = case.grid
    og = O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, g.nt)
    ep, sp = O.exp2pml(og, eps, sig)
    return np.concatenate([O.forward(og, ep, sp, case.src[q], case.wavelets[q],
                                     case.receivers[q])[0] for q in range(len(case.src))], axis=1)


def smooth_field(shape, seed):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), 4)
    return f / np.abs(f).max()


@pytest.mark.parametrize("objectives,obj_fnc", [
    (None, "L2"),
    ([("L2", 0.5), ("env_L2", 0.3), ("FREQstp", 0.2), ("env_L22", 0.4)], "L2"),
    ([("env_L2", 1.0)], "env_L22"),
])
def test_weighted_objectives(oracle_mod, gpu_ctx, objectives, obj_fnc):
    """gjw_batch_objective_multi of two trial models (w_objs_e.m:41-63 with
    gerjoii_.w.obj_FNCs / g_e_weights) against the oracle's composition.
    The instantaneous-phase objective is left out: after the band filter
    the samples before the first arrival are rounding noise in FP32, whose
    phase is arbitrary (in the reference as here)."""
    O = oracle_mod
    case = configs.small(nt=240, ns=2)
    d_obs = survey(case, gpu_ctx, case.eps_true, case.sig_true).forward()
    g = smooth_field(case.eps0.shape, 1)
    models = [LS.perturb_eps(case.eps0, g, k) for k in (-0.01, 0.02)]
    svs = []
    for e in models:
        sv = survey(case, gpu_ctx, e)
        sv.set_adj_filter(*BAND)
        svs.append(sv)
    fstp = (1.5e8, 3e8)
    J = S.Batch(svs).objective([d_obs] * len(svs), objectives=objectives, obj_fnc=obj_fnc,
                               fstp=fstp)
    dt = case.grid.dt
    ref = [O.w_objs_traces(d_obs.T.astype(np.float64), oracle_gathers(O, case, e, case.sig0), dt,
                           BAND, objectives, obj_fnc, fstp) for e in models]
    print("objectives", objectives, obj_fnc, J, ref, flush=True)
    np.testing.assert_allclose(J, ref, rtol=1e-4)


def test_stepparabola_s(oracle_mod, gpu_ctx):
    """w_stepparabola_s.m: nparabo + 2 conductivity trial steps, objectives
    of w_objs_s.m (batched) against the oracle's, and the vertex."""
    O = oracle_mod
    case = configs.small(nt=240, ns=2)
    d_obs = survey(case, gpu_ctx, case.eps_true, case.sig_true).forward()
    g_s = smooth_field(case.sig0.shape, 2)
    sv = survey(case, gpu_ctx)
    sv.set_adj_filter(*BAND)
    Es0 = S.Batch([sv]).objective([d_obs])[0]
    # sig_min = 0: the air cells above the topography have sigma = 0, which
    # any positive floor rejects at every step
    step, ks, Es = LS.stepparabola_s(gpu_ctx, case.grid, case.src, case.wavelets, case.receivers,
                                     case.eps0, case.sig0, g_s, d_obs, Es0, k_s=[1e-2, 1e3],
                                     ksprct_=0.2, ksprct__=0.9, sig_max=0.1, sig_min=0.0,
                                     nparabo=3, band=BAND)
    dt = case.grid.dt
    ref = [O.w_objs_traces(d_obs.T.astype(np.float64),
                           oracle_gathers(O, case, case.eps0, LS.perturb_sig(case.sig0, g_s, k)), dt,
                           BAND) for k in ks]
    p = np.polyfit(ks, ref, 2)
    step_ref = -p[1] / (2 * p[0])
    print("stepparabola_s", step, step_ref, ks, Es, ref, flush=True)
    assert len(ks) == 3 + 2 + 1 and ks[1] > 0 and ks[-1] > ks[1]
    assert np.ptp(ref) > 1e-3 * ref[0]  # trial steps wide enough for a curved fit
    np.testing.assert_allclose(Es, ref, rtol=1e-4)
    assert abs(step - step_ref) <= 1e-2 * abs(step_ref)


@pytest.mark.parametrize("kind", ["e", "s"])
def test_pica_steps(oracle_mod, gpu_ctx, kind):
    """w_pica_e_.m / w_pica_s_.m for each of three shots with its own
    gradient: the device steps (one-shot trial models in shared launches)
    against the oracle's FP64 gathers at the same trial models."""
    O = oracle_mod
    case = configs.small(nt=240, ns=3)
    d_obs = survey(case, gpu_ctx, case.eps_true, case.sig_true).forward()
    sv = survey(case, gpu_ctx)
    d_w = sv.forward()
    e_w = d_w - d_obs
    ns = len(case.src)
    G = np.stack([smooth_field(case.eps0.shape, 10 + q) for q in range(ns)], axis=2)
    if kind == "e":
        steps = LS.pica_e(gpu_ctx, case.grid, case.src, case.wavelets, case.receivers, case.eps0,
                          case.sig0, G, d_w, e_w, k_e=[1e-3, 0.05], eps_max=30.0, eps_min=0.5)
    else:
        steps = LS.pica_s(gpu_ctx, case.grid, case.src, case.wavelets, case.receivers, case.eps0,
                          case.sig0, G, d_w, e_w, k_s=[1e-2, 1e3], sig_max=0.1, sig_min=0.0)
    g = case.grid
    og = O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, g.nt)
    rp = np.concatenate([[0], np.cumsum([len(r) for r in case.receivers])])
    ref = []
    for q in range(ns):
        if kind == "e":
            k = LS.boundstep2(case.eps0, -G[:, :, q], 30.0, 0.5, [1e-3, 0.05])
            eps, sig = case.eps0 * np.exp(-k * case.eps0 * G[:, :, q]), case.sig0
        else:
            k = 0.05 * LS.boundstep2(case.sig0, G[:, :, q], 0.1, 0.0, [1e-2, 1e3])
            eps, sig = case.eps0, case.sig0 * np.exp(k * case.sig0 * G[:, :, q])
        ep, sp = O.exp2pml(og, eps, sig)
        dn = O.forward(og, ep, sp, case.src[q], case.wavelets[q], case.receivers[q])[0]
        e0, s0 = O.exp2pml(og, case.eps0, case.sig0)
        dw = O.forward(og, e0, s0, case.src[q], case.wavelets[q], case.receivers[q])[0]
        ref.append(O.pica_step(dn, dw, e_w[rp[q]:rp[q + 1]].T, k, kind))
    print("pica", kind, steps, ref, flush=True)
    np.testing.assert_allclose(steps, ref, rtol=1e-3)
