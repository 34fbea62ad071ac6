"""Step sizes of the permittivity and conductivity updates: the parabolic line
searches (src/w/w_stepparabola_e.m, w_stepparabola_s.m) and the Pica steps
(w_pica_e_.m, w_pica_s_.m), with their forward solves batched on the device.

The reference evaluates the objective at ``nparabo`` steps one shot and one
step at a time (``w_objs_e.m``, ``w_objs_s.m``: a full forward per shot and
step).  Here the perturbed models are the surveys of one ``Batch`` and every
shot of every step runs in shared launches (``gjw_batch_objective_multi``,
which also forms the weighted objectives of gerjoii_.w.obj_FNCs).  The Pica
steps take one forward per shot at that shot's own trial model; the shots
run as one-shot surveys of shared launches (``gjw_batch_forward``, up to
eight per batch).  The bounds search (``w_boundstep2.m``), the parabola fit
and the Pica ratio of inner products are host arithmetic, as in the
reference.
"""
from __future__ import annotations

import numpy as np

from . import survey as S


def perturb_eps(eps, g_e, ke):
    """w_objs_e.m:22-23: eps * exp(-ke * eps * g_e)."""
    return eps * np.exp(-ke * eps * g_e)


def perturb_sig(sig, g_s, ks):
    """w_objs_s.m:22-23: sigma * exp(-ks * sigma * g_s)."""
    return sig * np.exp(-ks * sig * g_s)


def _first_bad(k, epsi, depsi, eps_max, eps_min, reverse):
    """whileloop_rev / whileloop_fwd of w_boundstep2.m:27-62 (1-based index)."""
    def ok(kk):
        e = epsi * np.exp(kk * epsi * depsi)
        return e.max() <= eps_max and e.min() >= eps_min
    ik = 1
    if reverse:  # advance while out of bounds
        while not ok(k[ik - 1]):
            ik += 1
            if ik > len(k):
                return ik - 1
        return ik
    while ok(k[ik - 1]):  # forward: advance while in bounds, step back once
        ik += 1
        if ik > len(k):
            return ik - 1
    return max(ik - 1, 1)


def boundstep2(epsi, depsi, eps_max, eps_min, k_):
    """w_boundstep2.m:5-23: the largest step k (searched on a reverse
    13-point logspace over [min k_, max k_], then a forward 20-point linspace
    between the first admissible point and its predecessor) for which
    epsi * exp(k * epsi * depsi) stays inside [eps_min, eps_max]."""
    epsi = np.ravel(np.asarray(epsi, dtype=np.float64))
    depsi = np.ravel(np.asarray(depsi, dtype=np.float64))
    k_ = np.atleast_1d(np.asarray(k_, dtype=np.float64))
    k = np.logspace(np.log10(k_.max()), np.log10(k_.min()), 13)
    ik = _first_bad(k, epsi, depsi, eps_max, eps_min, True)
    if ik == 1:
        return k[0]
    k = np.linspace(k[ik - 1], k[ik - 2], 20)
    ik = _first_bad(k, epsi, depsi, eps_max, eps_min, False)
    return k[ik - 1]


def parabola_step(ke, Ee):
    """w_stepparabola_e.m:33-66: least-squares parabola through (ke, Ee) and
    its vertex; 0 when the fit is flat (NaN or zero) or the vertex is at a
    negative step."""
    p = np.polyfit(np.asarray(ke, dtype=np.float64), np.asarray(Ee, dtype=np.float64), 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        step = -p[1] / (2.0 * p[0])
    if not np.isfinite(step) or step == 0 or step < 0:
        return 0.0
    return float(step)


def _objectives(ctx, grid, src, wavelets, receivers, models, d_obs, band, objectives, obj_fnc,
                fstp):
    """Objectives of several (eps, sigma) models of one line of shots, in
    shared launches (at most eight models per batch)."""
    out = []
    for g0 in range(0, len(models), 8):
        surveys = []
        for e, sg in models[g0:g0 + 8]:
            sv = S.Survey(grid, ctx)
            sv.set_model(e, sg)
            sv.set_shots(src, wavelets, receivers)
            if band is not None:
                sv.set_adj_filter(*band)
            surveys.append(sv)
        b = S.Batch(surveys)
        out.extend(b.objective([d_obs] * len(surveys), objectives=objectives, obj_fnc=obj_fnc,
                               fstp=fstp))
        b.close()
        for sv in surveys:
            sv.close()
    return np.array(out)


def objs_e(ctx, grid, src, wavelets, receivers, eps, sig, g_e, d_obs, ke, *, band=None,
           objectives=None, obj_fnc="L2", fstp=None):
    """w_objs_e.m:22-63 at each step of ``ke``: the objective of
    eps * exp(-ke eps g_e), with the band filter, and the weighted sum over
    gerjoii_.w.obj_FNCs when ``objectives`` lists (name, weight) pairs."""
    return _objectives(ctx, grid, src, wavelets, receivers,
                       [(perturb_eps(eps, g_e, k), sig) for k in np.atleast_1d(ke)], d_obs, band,
                       objectives, obj_fnc, fstp)


def objs_s(ctx, grid, src, wavelets, receivers, eps, sig, g_s, d_obs, ks, *, band=None,
           obj_fnc="L2"):
    """w_objs_s.m:22-38 at each step of ``ks``: the objective of
    sigma * exp(-ks sigma g_s) with the band filter (one objective)."""
    return _objectives(ctx, grid, src, wavelets, receivers,
                       [(eps, perturb_sig(sig, g_s, k)) for k in np.atleast_1d(ks)], d_obs, band,
                       None, obj_fnc, None)


def stepparabola_e(ctx, grid, src, wavelets, receivers, eps, sig, g_e, d_obs, Ee0, *, k_e,
                   keprct_, keprct__, eps_max, eps_min, nparabo, band=None, objectives=None,
                   obj_fnc="L2", fstp=None):
    """w_stepparabola_e.m:6-66 for one line of shots: bounds of the step on
    both sides of zero, ``nparabo`` objectives of eps * exp(-ke eps g_e)
    (batched: all shots and steps in shared launches), parabola through
    (0, Ee0) and those, its vertex.  ``band`` = (f_low, f_high) filters the
    traces as w_objs_e.m does; ``objectives`` / ``obj_fnc`` / ``fstp`` as in
    objs_e.  Returns (step, ke, Ee)."""
    ke_ = -1.0 * boundstep2(eps, 1.0 * g_e, eps_max, eps_min, k_e) * keprct_
    ke__ = boundstep2(eps, -g_e, eps_max, eps_min, k_e) * keprct__
    ke = np.linspace(ke_, ke__, nparabo)
    Ee = objs_e(ctx, grid, src, wavelets, receivers, eps, sig, g_e, d_obs, ke, band=band,
                objectives=objectives, obj_fnc=obj_fnc, fstp=fstp)
    ke_all = np.concatenate([[0.0], ke])
    Ee_all = np.concatenate([[Ee0], Ee])
    return parabola_step(ke_all, Ee_all), ke_all, Ee_all


def stepparabola_s(ctx, grid, src, wavelets, receivers, eps, sig, g_s, d_obs, Es0, *, k_s,
                   ksprct_, ksprct__, sig_max, sig_min, nparabo, band=None, obj_fnc="L2"):
    """w_stepparabola_s.m:6-47 for one line of shots: both bounds from
    w_boundstep2 on -g_s (the reference's bwd_fwd = -1 makes the left one
    positive too) scaled by ksprct_ and ksprct__, ``nparabo`` + 2 steps
    between them, the objectives of sigma * exp(-ks sigma g_s) (w_objs_s.m,
    batched), the parabola through (0, Es0) and those.  The vertex is the
    step; 0 when the fit is flat (NaN or zero).  Unlike the permittivity
    search the reference keeps a negative vertex.  Returns (step, ks, Es)."""
    ks_ = boundstep2(sig, -g_s, sig_max, sig_min, k_s) * ksprct_
    ks__ = boundstep2(sig, -g_s, sig_max, sig_min, k_s) * ksprct__
    ks = np.linspace(ks_, ks__, nparabo + 2)
    Es = objs_s(ctx, grid, src, wavelets, receivers, eps, sig, g_s, d_obs, ks, band=band,
                obj_fnc=obj_fnc)
    ks_all = np.concatenate([[0.0], ks])
    Es_all = np.concatenate([[Es0], Es])
    p = np.polyfit(ks_all, Es_all, 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        step = -p[1] / (2.0 * p[0])
    if not np.isfinite(step) or step == 0:
        step = 0.0
    return float(step), ks_all, Es_all


def _pica(ctx, grid, src, wavelets, receivers, models, d_w, e_w, rec_ptr, sign, k):
    """sign * k * (dd . e) / (dd . dd) per shot, dd = sign * (d_new (+/-) d_w)
    (see pica_e / pica_s), with d_new the traces of shot q at models[q]."""
    ns = len(src)
    steps = np.zeros(ns)
    for g0 in range(0, ns, 8):
        qs = list(range(g0, min(ns, g0 + 8)))
        surveys = []
        for q in qs:
            sv = S.Survey(grid, ctx)
            sv.set_model(*models[q])
            sv.set_shots(src[q:q + 1], wavelets[q:q + 1], [receivers[q]])
            surveys.append(sv)
        b = S.Batch(surveys)
        dnew = b.forward()
        b.close()
        for sv in surveys:
            sv.close()
        for q, dq in zip(qs, dnew):
            a, z = int(rec_ptr[q]), int(rec_ptr[q + 1])
            dw = np.asarray(d_w[a:z], dtype=np.float64)
            ew = np.asarray(e_w[a:z], dtype=np.float64).ravel()
            dd = sign(dq.astype(np.float64), dw).ravel()
            steps[q] = k[q] * float(dd @ ew) / float(dd @ dd)
    return steps


def pica_e(ctx, grid, src, wavelets, receivers, eps, sig, g_e, d_w, e_w, *, k_e, eps_max,
           eps_min):
    """w_pica_e_.m:1-30 for every shot of a line, each with its own gradient
    g_e[:, :, q] (the per-source gradients of w_update_e.m): the bound
    k = w_boundstep2(eps, -g_e, eps_max, eps_min, k_e) (backward difference,
    fwd_bwd = -1), one forward at eps * exp(-k eps g_e), dd = -(d_new - d_w),
    step = k (dd . e_w) / (dd . dd).  d_w, e_w: the shots' traces of the
    current model and their residual, (ntraces, nt) in survey order (as
    gerjoii_.w.d_2d / e_2d hold them for the L2 objective)."""
    g_e = np.asarray(g_e, dtype=np.float64)
    ns = len(src)
    if g_e.ndim == 2:
        g_e = np.repeat(g_e[:, :, None], ns, axis=2)
    ks = [boundstep2(eps, -g_e[:, :, q], eps_max, eps_min, k_e) for q in range(ns)]
    models = [(eps * np.exp(-ks[q] * eps * g_e[:, :, q]), sig) for q in range(ns)]
    rec_ptr = np.concatenate([[0], np.cumsum([len(r) for r in receivers])])
    return _pica(ctx, grid, src, wavelets, receivers, models, d_w, e_w, rec_ptr,
                 lambda dn, dw: -(dn - dw), ks)


def pica_s(ctx, grid, src, wavelets, receivers, eps, sig, g_s, d_w, e_w, *, k_s, sig_max,
           sig_min):
    """w_pica_s_.m:1-30 for every shot of a line (w_update_es.m:179 takes
    it for the conductivity): k = 0.05 w_boundstep2(sigma, g_s, sig_max,
    sig_min, k_s) (forward difference, fwd_bwd = +1), one forward at
    sigma * exp(k sigma g_s), dd = d_new + d_w (the reference adds the two
    gathers), step = k (dd . e_w) / (dd . dd)."""
    g_s = np.asarray(g_s, dtype=np.float64)
    ns = len(src)
    if g_s.ndim == 2:
        g_s = np.repeat(g_s[:, :, None], ns, axis=2)
    ks = [boundstep2(sig, g_s[:, :, q], sig_max, sig_min, k_s) * 0.05 for q in range(ns)]
    models = [(eps, sig * np.exp(ks[q] * sig * g_s[:, :, q])) for q in range(ns)]
    rec_ptr = np.concatenate([[0], np.cumsum([len(r) for r in receivers])])
    return _pica(ctx, grid, src, wavelets, receivers, models, d_w, e_w, rec_ptr,
                 lambda dn, dw: dn + dw, ks)
