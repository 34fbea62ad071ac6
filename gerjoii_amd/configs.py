"""Seeded synthetic surveys for the benchmark and parity workloads.

Each builder returns a ``Case``: grid, true and starting models, shot geometry
and wavelets.  Models follow the families named by the workloads: smooth
random permittivity with elliptical anomalies (C1), layered models with
lognormal lenses, lognormal conductivity and random-walk topography (C2-C5),
a water table and voids (C4).  Air above the topography is eps_r = 1,
sigma = 0, as src/shared/w-dc/topo2mat.m fills it; the shape mask is 1 below
the surface.  Sources and receivers sit on the first ground cell of their
column (w_sources.m:100-103 puts surface shots at depth index binning()-1).
Seeds: 1000 + workload number, numpy PCG64.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .survey import Grid, dt_cfl, rec_Mw, src_index, w_wavelet


@dataclass
class Case:
    name: str
    grid: Grid
    eps_true: np.ndarray
    sig_true: np.ndarray
    eps0: np.ndarray          # starting model
    sig0: np.ndarray
    src: np.ndarray           # (ns, 2) 1-based [iz, ix]
    wavelets: np.ndarray      # (ns, nt)
    receivers: list           # Mw per shot (1-based)
    shape: np.ndarray         # gradient mask (parame_.shape)
    f0: float
    extra: dict = field(default_factory=dict)


def _smooth(a, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(a, sigma, mode="nearest")


def _rescale(a, lo, hi):
    a = (a - a.min()) / max(a.max() - a.min(), 1e-30)
    return lo + (hi - lo) * a


def _topography(rng, n, base=6, amp=5):
    """1-D random walk, steps in {-1,0,1}, bounded to base +- amp cells."""
    h = np.empty(n, dtype=np.int64)
    h[0] = base
    steps = rng.integers(-1, 2, size=n)
    for j in range(1, n):
        h[j] = min(max(h[j - 1] + steps[j], base - amp), base + amp)
    return h


def _layered(rng, m, n, nlay, lo, hi, lenses=True):
    eps = np.empty((m, n))
    bounds = np.sort(rng.integers(m // 10, m - m // 10, size=nlay - 1))
    vals = rng.uniform(lo, hi, size=nlay)
    z = np.arange(m)[:, None]
    wav = (np.sin(np.arange(n) / max(n / 6.0, 1.0) + rng.uniform(0, 6)) * m / 40.0)[None, :]
    lay = np.searchsorted(bounds, z - wav, side="right")
    eps[:] = vals[lay]
    if lenses:
        for _ in range(max(3, (m * n) // 40000)):
            cz, cx = rng.uniform(0, m), rng.uniform(0, n)
            rz, rx = rng.uniform(m / 40, m / 12), rng.uniform(n / 40, n / 12)
            mask = ((np.arange(m)[:, None] - cz) / rz) ** 2 + ((np.arange(n)[None, :] - cx) / rx) ** 2 <= 1
            eps[mask] *= np.exp(rng.normal(0, 0.25))
        eps = np.clip(eps, lo, hi)
    return eps


def _lognormal_sigma(rng, m, n, lo=1e-4, hi=2e-2):
    """Lognormal conductivity around 2e-3 S/m, spatially smooth, clipped."""
    z = _smooth(rng.standard_normal((m, n)), 6)
    z = (z - z.mean()) / (z.std() + 1e-12)
    return np.clip(np.exp(np.log(2e-3) + 1.2 * z), lo, hi)


def _apply_topo(eps, sig, h):
    m, n = eps.shape
    air = np.arange(m)[:, None] < h[None, :]
    eps = eps.copy()
    sig = sig.copy()
    eps[air] = 1.0
    sig[air] = 0.0
    return eps, sig, (~air).astype(np.float64)


def _surface_geometry(g, h, src_cols, rec_cols):
    src = np.array([src_index(g, h[c], c) for c in src_cols], dtype=np.int32)
    recs = rec_Mw(g, h[rec_cols], rec_cols)
    return src, [recs.copy() for _ in src_cols]


def _case(name, g, f0, eps_t, sig_t, shape, src, recs, smooth_start, extra=None):
    T = np.arange(g.nt) * g.dt
    w = w_wavelet(T, 1e3, 1.0 / f0)  # w_natur_u.m:37, ampli 1e3 (w_geom.m:109)
    eps0 = np.where(shape > 0, _smooth(eps_t, smooth_start), eps_t)
    sig0 = np.where(shape > 0, _smooth(sig_t, smooth_start), sig_t)
    return Case(name, g, eps_t, sig_t, eps0, sig0, src, np.tile(w, (len(src), 1)), recs, shape,
                f0, extra or {})


def c1(seed=1001):
    """200 x 100 model (n x m), dx = dz = 0.02 m, 20-cell PML, nt = 1500,
    100 MHz; smooth random eps in [4, 16] + two ellipses (eps 25, sigma 1e-2);
    sigma 1e-3; 4 sources and 40 receivers one cell below the surface."""
    rng = np.random.default_rng(seed)
    m, n, P = 100, 200, 20
    dx = 0.02
    g = Grid(m=m, n=n, pml=P, air=0, nt=1500, dx=dx, dt=dt_cfl(dx))
    eps = _rescale(_smooth(rng.standard_normal((m, n)), 8), 4, 16)
    sig = np.full((m, n), 1e-3)
    for cz, cx, rz, rx in [(40, 60, 12, 25), (65, 140, 10, 20)]:
        mask = ((np.arange(m)[:, None] - cz) / rz) ** 2 + ((np.arange(n)[None, :] - cx) / rx) ** 2 <= 1
        eps[mask] = 25.0
        sig[mask] = 1e-2
    h = np.ones(n, dtype=np.int64)  # surface shots one row into the model
    shape = np.ones((m, n))
    src, recs = _surface_geometry(g, h, [25, 75, 125, 175], np.arange(2, n, 5)[:40])
    return _case("C1", g, 100e6, eps, sig, shape, src, recs, 6)


def c2(seed=1002):
    """800 x 300, dx 0.025 m, 30-cell PML, nt 4000, 200 MHz; 6 layers in
    [3, 20] + lognormal lenses; lognormal sigma in [1e-4, 2e-2]; random-walk
    topography (+-5 cells); 1 source, 160 receivers."""
    rng = np.random.default_rng(seed)
    m, n, P = 300, 800, 30
    dx = 0.025
    g = Grid(m=m, n=n, pml=P, air=0, nt=4000, dx=dx, dt=dt_cfl(dx))
    eps = _layered(rng, m, n, 6, 3.0, 20.0)
    sig = _lognormal_sigma(rng, m, n)
    h = _topography(rng, n)
    eps, sig, shape = _apply_topo(eps, sig, h)
    src, recs = _surface_geometry(g, h, [n // 2], np.arange(0, n, 5)[:160])
    return _case("C2", g, 200e6, eps, sig, shape, src, recs, 8)


def c3(seed=1003, ns=64, nt=8000):
    """2000 x 600, dx 0.025 m, 40-cell PML, nt 8000, 64 shots at 0.3 m
    (12 cells) spacing, 400 receivers per shot (every 5 cells)."""
    rng = np.random.default_rng(seed)
    m, n, P = 600, 2000, 40
    dx = 0.025
    g = Grid(m=m, n=n, pml=P, air=0, nt=nt, dx=dx, dt=dt_cfl(dx))
    eps = _layered(rng, m, n, 6, 3.0, 20.0)
    sig = _lognormal_sigma(rng, m, n)
    h = _topography(rng, n)
    eps, sig, shape = _apply_topo(eps, sig, h)
    first = (n - 12 * (ns - 1)) // 2
    src, recs = _surface_geometry(g, h, [first + 12 * k for k in range(ns)], np.arange(0, n, 5)[:400])
    return _case("C3", g, 200e6, eps, sig, shape, src, recs, 8)


def c4(seed=1004, ns=32, nt=12000, coarse=1):
    """Precision-study scale: 4000 x 1500, dx 0.02 m, 40-cell PML, nt 12000,
    250 MHz, 32 shots; layered eps in [3, 20] with a water-table layer
    (eps 81) and void inclusions (eps 1, sigma 0); lognormal sigma; 400
    receivers every 10 cells.

    ``coarse`` = c > 1 gives the same survey downsampled: the model built at
    full resolution and sampled every c-th cell in both directions, dx times
    c, dt at the same Courant number (so nt // c steps cover the same
    duration), sources and receivers at the nearest coarse column, starting
    model smoothed over the same length, PML still 40 cells."""
    rng = np.random.default_rng(seed)
    m, n, P = 1500, 4000, 40
    dx = 0.02
    eps = _layered(rng, m, n, 6, 3.0, 20.0)
    sig = _lognormal_sigma(rng, m, n)
    z = np.arange(m)[:, None]
    top = int(0.45 * m) + (np.sin(np.arange(n) / 300.0 + rng.uniform(0, 6)) * 20)[None, :]
    water = (z >= top) & (z < top + m // 20)
    eps[water] = 81.0
    sig[water] = 1e-2
    for _ in range(12):
        cz, cx = rng.uniform(0.15 * m, 0.9 * m), rng.uniform(0, n)
        rz, rx = rng.uniform(8, 25), rng.uniform(10, 40)
        void = ((z - cz) / rz) ** 2 + ((np.arange(n)[None, :] - cx) / rx) ** 2 <= 1
        eps[void] = 1.0
        sig[void] = 0.0
    h = _topography(rng, n)
    eps, sig, shape = _apply_topo(eps, sig, h)
    first = (n - 120 * (ns - 1)) // 2
    scols = np.array([first + 120 * k for k in range(ns)])
    rcols = np.arange(0, n, 10)[:400]
    c = int(coarse)
    if c > 1:
        eps, sig, shape = eps[::c, ::c].copy(), sig[::c, ::c].copy(), shape[::c, ::c].copy()
        m, n = eps.shape
        h = (shape == 0).sum(axis=0)  # air is a top-contiguous run of rows
        scols, rcols = scols // c, rcols // c
    g = Grid(m=m, n=n, pml=P, air=0, nt=nt // c, dx=dx * c, dt=dt_cfl(dx * c))
    src, recs = _surface_geometry(g, h, scols, rcols)
    return _case("C4" if c == 1 else f"C4/{c}", g, 250e6, eps, sig, shape, src, recs, 8 / c)


def c5(seed=1005):
    """Uneven batch: survey A 600 x 200 (nt 3000, 24 shots), B 1500 x 500
    (nt 7000, 48 shots), C 3000 x 800 (nt 10000, 16 shots); dx 0.025 m,
    40-cell PML, 200 MHz, layered models with topography as C3, receivers
    every 5 cells (at most 400)."""
    out = []
    for k, (n, m, nt, ns) in enumerate(((600, 200, 3000, 24), (1500, 500, 7000, 48),
                                        (3000, 800, 10000, 16))):
        rng = np.random.default_rng(seed + k)
        dx = 0.025
        g = Grid(m=m, n=n, pml=40, air=0, nt=nt, dx=dx, dt=dt_cfl(dx))
        eps = _layered(rng, m, n, 6, 3.0, 20.0)
        sig = _lognormal_sigma(rng, m, n)
        h = _topography(rng, n)
        eps, sig, shape = _apply_topo(eps, sig, h)
        step = max(1, (n - 20) // ns)
        src, recs = _surface_geometry(g, h, [10 + step * q for q in range(ns)],
                                      np.arange(0, n, 5)[:400])
        out.append(_case("C5" + "ABC"[k], g, 200e6, eps, sig, shape, src, recs, 8))
    return out


def small(seed=7, m=40, n=60, P=10, air=0, nt=300, ns=3, nr=12, dx=0.02, f0=300e6):
    """Small layered survey with topography for parity tests."""
    rng = np.random.default_rng(seed)
    g = Grid(m=m, n=n, pml=P, air=air, nt=nt, dx=dx, dt=dt_cfl(dx))
    eps = _layered(rng, m, n, 4, 3.0, 16.0)
    sig = np.clip(_lognormal_sigma(rng, m, n) * 3, 1e-4, 3e-2)
    h = _topography(rng, n, base=3, amp=2)
    eps, sig, shape = _apply_topo(eps, sig, h)
    src_cols = np.linspace(5, n - 6, ns).astype(int)
    rec_cols = np.linspace(1, n - 2, nr).astype(int)
    src, recs = _surface_geometry(g, h, src_cols, rec_cols)
    return _case("small", g, f0, eps, sig, shape, src, recs, 3)


CASES = {"C1": c1, "C2": c2, "C3": c3, "C4": c4}
