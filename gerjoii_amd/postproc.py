"""Per-source gradient post-processing and the source reduction, on the device.

Mirrors the tail of the reference's per-source update (src/w/w_update_e.m:
108-145 for the permittivity, w_update_s.m:89-137 for the conductivity,
w_update_es.m:105-173 for both parameters) and the reduction over sources
(src/w/w_update_e_.m:28-42, w_update_s_.m:28-38) followed by the
multiplicative update (src/w/w_image_e.m:50-51, w_image_s.m:38-39).  The reference runs the
post-processing once per source inside a parfor; here every source's
gradient image goes through one launch sequence (``Context.gaussi``,
``Context.image_gaussian``, ``Context.normali``, ``Context.stack_update``).

Per-source gradients come from a survey with ``set_shot_gradients()``:
(m, n, ns) FP64 arrays, column-major, one image per shot.
"""
from __future__ import annotations

import numpy as np

from .survey import C0, Context, Grid


def taper_index(grid: Grid, src):
    """w_update_e.m:112-114: the source (iz, ix) on the padded grid (1-based,
    as gerjoii_.w.s) trimmed to the parameter matrix: ix - pjs and
    iz - pis - air + 1.  Returns (ns, 2) int32; raises where the result falls
    outside the model, where the reference's indexing would fail."""
    src = np.asarray(src, dtype=np.int64).reshape(-1, 2)
    s = np.empty(src.shape, dtype=np.int32)
    s[:, 0] = src[:, 0] - grid.pis - grid.air + 1
    s[:, 1] = src[:, 1] - grid.pjs
    bad = (s[:, 0] < 1) | (s[:, 0] > grid.m) | (s[:, 1] < 1) | (s[:, 1] > grid.n)
    if bad.any():
        k = int(np.flatnonzero(bad)[0])
        raise ValueError(f"source {k} at {tuple(src[k])} maps to model index {tuple(s[k])}, "
                         f"outside the {grid.m} x {grid.n} model")
    return s


def source_wavelength(eps_r, s, fo, c=C0):
    """w_update_e.m:116: lo = c / sqrt(eps(s)) / fo for each source (m)."""
    eps_r = np.asarray(eps_r, dtype=np.float64)
    return c / np.sqrt(eps_r[s[:, 0] - 1, s[:, 1] - 1]) / fo


def w_update_post(ctx: Context, g, grid: Grid, src, eps_r, X, Y, fo, ax, az, shape=None):
    """w_update_e.m:108-143 (permittivity) or w_update_s.m:89-137
    (conductivity; the same chain, the mask before the NaN guard, which gives
    the same result) for every source's gradient at once, in place on g
    ((m, n, ns) F-order FP64, host or device); fo is gerjoii_.w.regu.fo:

    * gaussi taper centred on the source, width 2 lo (lines 112-118);
    * image_gaussian LOW_PASS with (ax, az) = gerjoii_.w.regu.(ax, az)
      (lines 125-128);
    * normali, the NaN guard and the shape mask (lines 129-143).
    """
    s = taper_index(grid, src)
    lo = source_wavelength(eps_r, s, fo)
    ctx.gaussi(g, X, Y, s, 2 * lo)
    ctx.image_gaussian(g, ax, az, "LOW_PASS", out=g)
    ctx.normali(g, nan_guard=True, shape=shape)
    return g


def w_update_es_post(ctx: Context, g_e, g_s, grid: Grid, src, eps_r, X, Y, fo, ax, az):
    """w_update_es.m:105-173 for every source at once, in place on both
    gradients: the gaussi taper with lo from parame_.w.fo, image_gaussian
    LOW_PASS, division by max|g|.  This path has no mask and no NaN guard.
    An all-zero image stays zero here (normali's guard), where MATLAB's
    0/0 would give NaN."""
    s = taper_index(grid, src)
    lo = source_wavelength(eps_r, s, fo)
    for g in (g_e, g_s):
        ctx.gaussi(g, X, Y, s, 2 * lo)
        ctx.image_gaussian(g, ax, az, "LOW_PASS", out=g)
        ctx.normali(g)
    return g_e, g_s


def w_update_stack(ctx: Context, g, steps, model=None, neg_rule=True):
    """w_update_e.m:177 and w_update_e_.m:38-41: dparam = -sum(step_s g_s)/ns
    (zero when the steps sum below zero; neg_rule=False for the conductivity
    and joint reductions, w_update_s_.m:38 and w_update_es_.m, which have no
    such rule); with a model, w_image_e.m:50-51 (w_image_s.m:38-39) updates
    it in place: model .* exp(model .* dparam).  Returns dparam."""
    return ctx.stack_update(g, steps, model=model, neg_rule=neg_rule)
