"""MATLAB v5 ``.mat`` I/O of gerjoii's radar data, without MATLAB (scipy.io).

Two kinds of file feed the radar path of a gerjoii project:

* ``line<is>.mat``: one shot gather, the struct ``radargram`` with fields
  d, t, dt, fo, dx, r, s, dr, dsr, wvlet_, Jy, std_
  (field/tinyradar-sy/readme-fwd.md:97-112), written by src/w/w_natur_u.m:41-65
  and read by src/w/w_load.m:9-51 and the project viewers
  (src/shared/graphics/graphics_py/fancy_figure.py:103-120);
* ``s_r.mat`` / ``s_r_.mat``: ns x 2 cells of source and receiver positions,
  real (metres) and indexed (grid offsets), written by experim_w.m:113-135
  (field/example-sy/base/scripts/experim_w.m) and read by w_load.m:10,27-35.

``load_survey`` turns a project's lines into the inputs of
``Survey.set_shots`` and ``Survey.fwi_gradient``; ``write_survey`` writes
them back the way w_natur_u.m does.  Only MATLAB's v5/v7 format is read
(MATLAB's default ``save``); v7.3 (HDF5) files are refused with an error.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
import scipy.io as sio

from .survey import Grid

RADARGRAM_FIELDS = ("d", "t", "dt", "fo", "dx", "r", "s", "dr", "dsr", "wvlet_", "Jy", "std_")


@dataclass
class Radargram:
    """The ``radargram`` struct of a line<is>.mat (readme-fwd.md:97-112)."""
    d: np.ndarray        # (nt, nr) single: the traces, time first
    t: np.ndarray        # (nt,) [s]
    dt: float            # [s]
    fo: float            # [Hz]
    dx: float            # [m]
    r: np.ndarray        # (nr, 2) receivers (x, z) [m]
    s: np.ndarray        # (2,) source (x, z) [m]
    dr: float            # receiver spacing [m]
    dsr: float           # source-receiver offset of the first receiver [m]
    wvlet_: np.ndarray   # (nt,) source wavelet
    Jy: np.ndarray       # (nt,) single: the field at the source
    std_: np.ndarray     # (nt*nr,) data weights
    extra: dict = field(default_factory=dict)  # v_mute, t_mute, irx_keepx_, ... when present


def _loadmat(path):
    try:
        return sio.loadmat(path)
    except NotImplementedError as e:  # scipy's message for HDF5-based files
        raise ValueError(f"{path}: MATLAB v7.3 (HDF5) files are not supported; "
                         "save with -v7 or -v6") from e


def _field(st, name):
    # the reference's access pattern: loadmat(f)[struct][field].item()
    return st[name].item()


def read_radargram(path) -> Radargram:
    """Read the ``radargram`` struct of a line<is>.mat."""
    m = _loadmat(path)
    if "radargram" not in m:
        raise ValueError(f"{path}: no variable 'radargram'")
    st = m["radargram"]
    names = st.dtype.names or ()
    missing = [f for f in ("d", "t", "dt", "r", "s") if f not in names]
    if missing:
        raise ValueError(f"{path}: radargram lacks {missing}")
    get = lambda f: _field(st, f) if f in names else None
    scalar = lambda f: float(np.asarray(get(f)).ravel()[0]) if get(f) is not None else float("nan")
    d = np.asarray(get("d"))
    nt, nr = d.shape
    vec = lambda f, dt: (np.asarray(get(f), dtype=dt).ravel() if get(f) is not None
                         else np.zeros(0, dtype=dt))
    rg = Radargram(d=d.astype(np.float32, copy=False), t=vec("t", np.float64), dt=scalar("dt"),
                   fo=scalar("fo"), dx=scalar("dx"),
                   r=np.asarray(get("r"), dtype=np.float64).reshape(nr, -1),
                   s=vec("s", np.float64), dr=scalar("dr"), dsr=scalar("dsr"),
                   wvlet_=vec("wvlet_", np.float64), Jy=vec("Jy", np.float32),
                   std_=vec("std_", np.float64))
    rg.extra = {f: get(f) for f in names if f not in RADARGRAM_FIELDS}
    return rg


def write_radargram(path, d, t, dt, fo, dx, r, s, wvlet_, Jy, std_=None, **extra):
    """w_natur_u.m:41-65: save ``radargram`` with d and Jy single, the rest
    double; dr = |r(1,1) - r(2,1)|, dsr = |s - r(1,:)| (lines 55-56); std_
    defaults to ones(nt*nr, 1) (line 35).  Vectors are stored as columns and
    s as a 1 x 2 row, as MATLAB holds them."""
    d = np.asarray(d, dtype=np.float32)
    nt, nr = d.shape
    r = np.asarray(r, dtype=np.float64).reshape(nr, 2)
    s = np.asarray(s, dtype=np.float64).reshape(1, 2)
    dr = abs(r[0, 0] - r[1, 0]) if nr > 1 else 0.0
    dsr = float(np.sqrt((s[0, 0] - r[0, 0]) ** 2 + (s[0, 1] - r[0, 1]) ** 2))
    if std_ is None:
        std_ = np.ones(nt * nr)
    col = lambda v, dt_: np.asarray(v, dtype=dt_).reshape(-1, 1)
    st = {"d": d, "t": col(t, np.float64), "dt": float(dt), "fo": float(fo), "dx": float(dx),
          "r": r, "s": s, "dr": float(dr), "dsr": dsr, "wvlet_": col(wvlet_, np.float64),
          "Jy": col(Jy, np.float32), "std_": col(std_, np.float64)}
    st.update(extra)
    sio.savemat(path, {"radargram": st}, do_compression=True)


def read_cells(path, name=None):
    """Read an ns x 2 cell (s_r or s_r_): a list of (sources, receivers)
    pairs, each a 2-D float64 array as stored."""
    m = _loadmat(path)
    if name is None:
        keys = [k for k in m if not k.startswith("__")]
        if len(keys) != 1:
            raise ValueError(f"{path}: expected one variable, found {keys}")
        name = keys[0]
    c = m[name]
    if c.dtype != object or c.ndim != 2 or c.shape[1] != 2:
        raise ValueError(f"{path}: {name} is not an ns x 2 cell")
    return [(np.atleast_2d(np.asarray(c[k, 0], dtype=np.float64)),
             np.atleast_2d(np.asarray(c[k, 1], dtype=np.float64))) for k in range(c.shape[0])]


def write_cells(path, name, pairs):
    """experim_w.m:113-135: save an ns x 2 cell of (sources, receivers)."""
    c = np.empty((len(pairs), 2), dtype=object)
    for k, (a, b) in enumerate(pairs):
        c[k, 0] = np.atleast_2d(np.asarray(a, dtype=np.float64))
        c[k, 1] = np.atleast_2d(np.asarray(b, dtype=np.float64))
    sio.savemat(path, {name: c}, do_compression=True)


def indexed_to_grid(grid: Grid, src_off, rec_off):
    """w_load.m:27-35 + w_M.m:18-20: the indexed cell entries (source
    (iz, ix), receivers (ix, iz), offsets from the last air pixel and the
    PML) -> the source on the padded Ez grid (1-based [iz, ix]) and the
    receivers' linear indices Mw = sub2ind([nx+1, ny+1], iz, ix)."""
    src_off = np.asarray(src_off, dtype=np.int64).reshape(-1, 2)
    rec_off = np.asarray(rec_off, dtype=np.int64).reshape(-1, 2)
    src = np.stack([grid.pis + grid.air + src_off[:, 0], grid.pjs + src_off[:, 1]], axis=1)
    rx = grid.pjs + rec_off[:, 0]
    rz = grid.pis + grid.air + rec_off[:, 1]
    if (rz < 1).any() or (rz > grid.nx + 1).any() or (rx < 1).any() or (rx > grid.ny + 1).any():
        raise ValueError("receiver outside the padded grid")
    return src, (rz - 1) + (rx - 1) * (grid.nx + 1) + 1


def grid_to_indexed(grid: Grid, src, Mw):
    """Inverse of indexed_to_grid: a padded-grid source [iz, ix] and Mw ->
    the cell entries (iz, ix) and (ix, iz) of s_r_.mat."""
    src = np.asarray(src, dtype=np.int64).reshape(2)
    Mw = np.asarray(Mw, dtype=np.int64)
    rz = (Mw - 1) % (grid.nx + 1) + 1
    rx = (Mw - 1) // (grid.nx + 1) + 1
    s_off = np.array([[src[0] - grid.pis - grid.air, src[1] - grid.pjs]])
    r_off = np.stack([rx - grid.pjs, rz - grid.pis - grid.air], axis=1)
    return s_off, r_off


def w_load(grid: Grid, data_path, is_, wvlets=None):
    """w_load.m:8-51 for shot ``is_`` (1-based): the observed traces d_o
    (nt, nr), std_, dr, the padded-grid source [iz, ix], Mw, the wavelet
    (column is_ of ``wvlets`` (nt, ns) as parame_.w.wvlets_, else the
    radargram's own) and the mute fields when the line holds them."""
    rg = read_radargram(os.path.join(data_path, f"line{is_}.mat"))
    cells = read_cells(os.path.join(data_path, "s_r_.mat"), "s_r_")
    s_off, r_off = cells[is_ - 1]
    src, Mw = indexed_to_grid(grid, s_off, r_off)
    out = {"d_o": rg.d, "std_": rg.std_, "dr": rg.dr, "s": src[0], "M": Mw,
           "wvlet_": (np.asarray(wvlets)[:, is_ - 1] if wvlets is not None else rg.wvlet_)}
    if "v_mute" in rg.extra:  # w_load.m:17-21
        out["v_mute"] = float(np.asarray(rg.extra["v_mute"]).ravel()[0])
        out["t_mute"] = float(np.asarray(rg.extra["t_mute"]).ravel()[0])
        out["r_real"] = rg.r
    return out


def load_survey(grid: Grid, data_path, shots, wvlets=None):
    """Lines ``shots`` (1-based numbers) of a project folder as the inputs of
    Survey.set_shots and Survey.fwi_gradient: src (ns, 2), wavelets (ns, nt)
    float32, receivers (list of Mw), d_obs (sum of nr, nt) float32 in the
    trace-major layout of the library."""
    src, wav, recs, obs = [], [], [], []
    for is_ in shots:
        L = w_load(grid, data_path, is_, wvlets)
        if L["d_o"].shape[0] != grid.nt:
            raise ValueError(f"line{is_}.mat holds {L['d_o'].shape[0]} samples, the grid {grid.nt}")
        src.append(L["s"])
        wav.append(np.asarray(L["wvlet_"], dtype=np.float32))
        recs.append(L["M"])
        obs.append(np.ascontiguousarray(L["d_o"].T))
    return (np.array(src, dtype=np.int32), np.array(wav), recs,
            np.concatenate(obs, axis=0).astype(np.float32, copy=False))


def write_survey(data_path, grid: Grid, src, receivers, traces, wavelets, fo, jz=None,
                 src_real=None, rec_real=None):
    """w_natur_u.m:41-65 and experim_w.m:113-135 for every shot of a survey:
    line<is>.mat per shot (traces in the library's trace-major layout are
    written time-first), s_r_.mat (indexed) and s_r.mat (metres; by default
    the cell positions times dx)."""
    os.makedirs(data_path, exist_ok=True)
    ns = len(src)
    t = np.arange(grid.nt) * grid.dt
    idx, real = [], []
    off = 0
    for q in range(ns):
        nr = len(receivers[q])
        s_off, r_off = grid_to_indexed(grid, src[q], receivers[q])
        idx.append((s_off, r_off))
        s_m = (np.asarray(src_real[q], dtype=np.float64).reshape(1, 2) if src_real is not None
               else np.array([[s_off[0, 1] * grid.dx, s_off[0, 0] * grid.dx]]))
        r_m = (np.asarray(rec_real[q], dtype=np.float64) if rec_real is not None
               else np.asarray(r_off, dtype=np.float64) * grid.dx)
        real.append((s_m, r_m))
        d = np.asarray(traces[off:off + nr], dtype=np.float32).T
        off += nr
        write_radargram(os.path.join(data_path, f"line{q + 1}.mat"), d, t, grid.dt, fo, grid.dx,
                        r_m, s_m, wavelets[q], jz[:, q] if jz is not None else np.zeros(grid.nt))
    write_cells(os.path.join(data_path, "s_r_.mat"), "s_r_", idx)
    write_cells(os.path.join(data_path, "s_r.mat"), "s_r", real)
