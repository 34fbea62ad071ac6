"""Parabolic line search of the permittivity update (src/w/w_stepparabola_e.m),
with its forward solves batched on the device.

The reference evaluates the objective at ``nparabo`` steps one shot and one
step at a time (``w_objs_e.m``, a full forward per shot and step).  Here the
perturbed models are the surveys of one ``Batch`` and every shot of every
step runs in shared launches (``gjw_batch_objective``).  The bounds search
(``w_boundstep2.m``) and the parabola fit are host arithmetic on the model,
as in the reference.
"""
from __future__ import annotations

import numpy as np

from . import survey as S


def perturb_eps(eps, g_e, ke):
    """w_objs_e.m:22-23: eps * exp(-ke * eps * g_e)."""
    return eps * np.exp(-ke * eps * g_e)


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


def stepparabola_e(ctx, grid, src, wavelets, receivers, eps, sig, g_e, d_obs, Ee0, *, k_e,
                   keprct_, keprct__, eps_max, eps_min, nparabo, band=None):
    """w_stepparabola_e.m:6-66 for one line of shots: bounds of the step on
    both sides of zero, ``nparabo`` objectives of eps * exp(-ke eps g_e)
    (batched: all shots and steps in shared launches), parabola through
    (0, Ee0) and those, its vertex.  ``band`` = (f_low, f_high) filters the
    traces as w_objs_e.m does.  Returns (step, ke, Ee)."""
    ke_ = -1.0 * boundstep2(eps, 1.0 * g_e, eps_max, eps_min, k_e) * keprct_
    ke__ = boundstep2(eps, -g_e, eps_max, eps_min, k_e) * keprct__
    ke = np.linspace(ke_, ke__, nparabo)
    surveys = []
    for k in ke:
        sv = S.Survey(grid, ctx)
        sv.set_model(perturb_eps(eps, g_e, k), sig)
        sv.set_shots(src, wavelets, receivers)
        if band is not None:
            sv.set_adj_filter(*band)
        surveys.append(sv)
    b = S.Batch(surveys)
    Ee = b.objective([d_obs] * len(surveys))
    b.close()
    for sv in surveys:
        sv.close()
    ke_all = np.concatenate([[0.0], ke])
    Ee_all = np.concatenate([[Ee0], Ee])
    return parabola_step(ke_all, Ee_all), ke_all, Ee_all
