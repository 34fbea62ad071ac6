"""Host-side API over libgjw: one ``Survey`` = one padded grid, its model and
its shots, resident on one GPU.

Mirrors the reference's per-shot pipeline (src/w/w_update_e.m:50-98:
w_load -> w_fwd -> w_bozdag/w_E -> w_grad) for all shots of a survey at once.
All arrays use MATLAB's column-major convention: model arrays are (m, n)
depth-first, a shot's traces are (nt, nr) (``gerjoii_.w.d_2d``).
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _native as N

EPS0 = 8.854187817e-12  # parame_.w.eps_0
MU0 = 4e-7 * np.pi      # parame_.w.mu_0
C0 = 1.0 / np.sqrt(MU0 * EPS0)


@dataclass
class Grid:
    """Padded-grid geometry (src/w/w_geom.m:78-172; pml_w == pml_d)."""
    m: int           # model rows (depth)
    n: int           # model columns
    pml: int
    air: int
    nt: int
    dx: float
    dt: float
    eps0: float = EPS0
    mu0: float = MU0
    R0: float = 1e-8
    pml_order: int = 2

    @property
    def nx(self):  # finite_.w.nx (depth, with PML and air)
        return self.m + 2 * self.pml + self.air

    @property
    def ny(self):  # finite_.w.ny
        return self.n + 2 * self.pml

    @property
    def pis(self):
        return self.pml + 1

    @property
    def pjs(self):
        return self.pml + 1

    @property
    def pie(self):
        return self.nx - self.pml + 1

    @property
    def pje(self):
        return self.ny - self.pml + 1

    def cells(self):
        """Padded Ez cells, (nx+1)*(ny+1): the unit of one cell-update."""
        return (self.nx + 1) * (self.ny + 1)

    def c_struct(self):
        return N.GjwGrid(self.m, self.n, self.pml, self.air, self.nt, self.pml_order,
                         self.dx, self.dt, self.eps0, self.mu0, self.R0)


def dt_cfl(dx, eps_min=1.0, courant=0.9):
    """w_geom.m:120-122: dt = 0.9 / (v_max sqrt(dx^-2 + dy^-2)), dy = dx."""
    vmax = C0 / np.sqrt(eps_min)
    return courant / (vmax * np.sqrt(2.0 / dx**2))


def w_wavelet(t, ampli, tau):
    """w_wavelet.m:11-32: -(t - t0) exp(-(t - t0)^2 / tau^2), tau = 1/(pi fo),
    t0 = 5 tau, scaled to max-abs ``ampli``."""
    tau = tau / np.pi
    t0 = 5 * tau
    w = -(t - t0) * np.exp(-(((t - t0) ** 2) / (tau**2)))
    return ampli * (w / np.max(np.abs(w)))


def src_index(g: Grid, iz0, ix0):
    """w_load.m:31-32: source [iz, ix] on the padded Ez grid (1-based) from
    model offsets (iz0 rows below the top of the model, ix0 columns)."""
    return (g.pis + g.air + int(iz0), g.pjs + int(ix0))


def rec_Mw(g: Grid, rz0, rx0):
    """w_load.m:34-35 + w_M.m:20: receivers [rx, rz] -> linear Mw (1-based)."""
    rz = g.pis + g.air + np.asarray(rz0, dtype=np.int64)
    rx = g.pjs + np.asarray(rx0, dtype=np.int64)
    return (rz - 1) + (rx - 1) * (g.nx + 1) + 1


class Context:
    """One GPU (gjw_create); shared by the surveys on it."""

    def __init__(self, device=0):
        lib = N.load()
        h = C.c_void_p()
        N.check(lib.gjw_create(int(device), C.byref(h)), "gjw_create")
        self.h = h
        self.device = device
        # the native surveys and batches of this context: released before it
        self._surveys = weakref.WeakSet()
        self._batches = weakref.WeakSet()

    def close(self):
        if self.h:
            for b in list(self._batches):
                b.close()
            for sv in list(self._surveys):
                sv.close()
            N.load().gjw_destroy(self.h)
            self.h = None

    def differentiate_cube(self, v, dt):
        """differentiate_cube.m on the GPU; v is an (m, n, nt) float32 array,
        column-major (numpy F-order), modified in place."""
        assert v.dtype == np.float32
        nz, nx, nt = v.shape
        if isinstance(v, np.ndarray):
            assert v.flags.f_contiguous
        N.check(N.load().gjw_differentiate_cube(self.h, N.ptr(v), nz, nx, nt, float(dt)),
                "gjw_differentiate_cube")
        return v

    def filt_gauss(self, d, dt, f_low, f_high, steep=8):
        """filt_gauss.m on the GPU, in place: d is (ntraces, nt) float32
        (C-order numpy or torch cuda; trace-major like MATLAB's nt x nr)."""
        assert str(d.dtype) in ("float32", "torch.float32")
        ntr, nt = d.shape
        if isinstance(d, np.ndarray):
            assert d.flags.c_contiguous
        _sync_torch(d)
        N.check(N.load().gjw_filt_gauss(self.h, N.ptr(d), ntr, nt, float(dt), float(f_low),
                                        float(f_high), int(steep)), "gjw_filt_gauss")
        return d

    def lmute(self, d, d_o, t, rx, v_mute, t_mute):
        """w_lmute.m on the GPU: mutes d and d_o in place ((ntraces, nt)
        float32, C-order) and returns e = d - d_o.  t [s], rx [m], v_mute
        [m/ns], t_mute [ns]."""
        ntr, nt = d.shape
        for a in (d, d_o):
            assert isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous
        t = np.ascontiguousarray(t, dtype=np.float64)
        rx = np.ascontiguousarray(rx, dtype=np.float64)
        assert t.size == nt and rx.size == ntr
        e = np.empty_like(d)
        N.check(N.load().gjw_lmute(self.h, N.ptr(d), N.ptr(d_o), N.ptr(e), ntr, nt, N.ptr(t),
                                   N.ptr(rx), float(v_mute), float(t_mute)), "gjw_lmute")
        return e


    def w_bozdag(self, d_o, d, kind="L2"):
        """w_bozdag.m + w_E.m on the GPU for nt x nr traces (numpy, any
        layout): returns (e, d_e, d_oe, J), each array nt x nr float32
        (F-order): e the adjoint source, d_e / d_oe the synthetic and
        observed traces (L2), envelopes (env_*) or instantaneous phases
        (phase_L2)."""
        d_o = np.asfortranarray(d_o, dtype=np.float32)
        d = np.asfortranarray(d, dtype=np.float32)
        if d.ndim == 1:
            d, d_o = d[:, None], d_o[:, None]
        assert d.shape == d_o.shape
        nt, nr = d.shape
        e = np.empty_like(d, order="F")
        de = np.empty_like(d, order="F")
        doe = np.empty_like(d, order="F")
        J = C.c_double(0)
        N.check(N.load().gjw_w_bozdag(self.h, N.ptr(d_o), N.ptr(d), nr, nt, N.MISFIT_KINDS[kind],
                                      N.ptr(e), N.ptr(de), N.ptr(doe), C.byref(J)),
                "gjw_w_bozdag")
        return e, de, doe, J.value

    # -- gradient post-processing (w_update_e.m:108-145) ------------------
    @staticmethod
    def _images(a):
        """(m, n, count) view of an FP64 image batch: numpy F-order (m, n) or
        (m, n, count), or a torch tensor of shape (count, n, m) / (n, m)
        (contiguous, i.e. column-major images)."""
        if hasattr(a, "data_ptr"):
            assert str(a.dtype) == "torch.float64" and a.is_contiguous()
            sh = tuple(a.shape)
            if len(sh) == 2:
                sh = (1,) + sh
            return sh[2], sh[1], sh[0]
        assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous
        if a.ndim == 2:
            return a.shape[0], a.shape[1], 1
        return a.shape

    def gaussi(self, g, X, Y, s, lo):
        """gaussi.m taper of each image in place: s is (count, 2) 1-based
        (iz, ix) model indices, lo (count,) gaussi's lo argument."""
        m, n, cnt = self._images(g)
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        s = np.ascontiguousarray(np.reshape(s, (cnt, 2)), dtype=np.int32)
        lo = np.ascontiguousarray(np.reshape(lo, (cnt,)), dtype=np.float64)
        assert X.size == n and Y.size == m
        _sync_torch(g)
        N.check(N.load().gjw_gaussi(self.h, N.ptr(g), m, n, cnt, N.ptr(X), N.ptr(Y), N.ptr(s),
                                    N.ptr(lo)), "gjw_gaussi")
        return g

    def image_gaussian(self, img, ax, az, kind="LOW_PASS", out=None):
        """image_gaussian.m on each image (returns the filtered images)."""
        m, n, cnt = self._images(img)
        if out is None:
            out = img.clone() if hasattr(img, "data_ptr") else np.empty_like(img, order="F")
        typ = {"LOW_PASS": N.GJW_LOW_PASS, "HI_PASS": N.GJW_HI_PASS}[kind]
        _sync_torch(img, out)
        N.check(N.load().gjw_image_gaussian(self.h, N.ptr(img), N.ptr(out), m, n, cnt, float(ax),
                                            float(az), typ), "gjw_image_gaussian")
        return out

    def normali(self, a, nan_guard=False, shape=None):
        """normali.m on each image in place; nan_guard zeroes images with a
        NaN (w_update_e.m:133-137); shape (m, n) then masks every image
        (w_update_e.m:141-143)."""
        m, n, cnt = self._images(a)
        if shape is not None and not hasattr(shape, "data_ptr"):
            shape = np.asfortranarray(shape, dtype=np.float64)
        _sync_torch(a, shape)
        N.check(N.load().gjw_normali(self.h, N.ptr(a), m, n, cnt, int(bool(nan_guard)),
                                     N.ptr(shape)), "gjw_normali")
        return a

    def stack_update(self, g, steps, model=None, out=None, neg_rule=True):
        """w_update_e_.m:38-41 (and w_image_e.m:50-51 when a model is given,
        updated in place): returns dparam = sum_c(-steps[c] g_c)/count, zero
        when the steps sum below zero unless neg_rule is False (the
        conductivity reduction, w_update_s_.m:38)."""
        m, n, cnt = self._images(g)
        steps = np.ascontiguousarray(np.reshape(steps, (cnt,)), dtype=np.float64)
        if out is None:
            out = (g.new_empty((n, m)) if hasattr(g, "data_ptr")
                   else np.empty((m, n), dtype=np.float64, order="F"))
        _sync_torch(g, model, out)
        N.check(N.load().gjw_stack_update(self.h, N.ptr(g), m, n, cnt, N.ptr(steps),
                                          int(bool(neg_rule)), N.ptr(out), N.ptr(model)),
                "gjw_stack_update")
        return out


_default_ctx = {}


def context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def _sync_torch(*xs):
    for x in xs:
        if x is not None and hasattr(x, "is_cuda") and x.is_cuda:
            import torch
            torch.cuda.synchronize(x.device)
            return


class Survey:
    """A survey resident on one GPU: grid + model + shots + device state."""

    def __init__(self, grid: Grid, ctx: Context | None = None):
        self.grid = grid
        self.ctx = ctx or context(0)
        if not self.ctx.h:
            raise N.GjwError("Survey: the context is closed")
        h = C.c_void_p()
        gs = grid.c_struct()
        N.check(N.load().gjw_survey_create(self.ctx.h, C.byref(gs), C.byref(h)),
                "gjw_survey_create")
        self.h = h
        self.rec_ptr = None
        # bumped by every call that replaces the model, the shots or the kept
        # forward state (checkpoints); autograd nodes compare it
        self.generation = 0
        self.ctx._surveys.add(self)

    def close(self):
        if self.h:
            N.load().gjw_survey_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup --------------------------------------------------------------
    def set_option(self, key, value):
        N.check(N.load().gjw_set_option(self.h, int(key), int(value)), "gjw_set_option")

    def set_path(self, plain: bool):
        self.set_option(N.GJW_OPT_PATH, N.GJW_PATH_PLAIN if plain else N.GJW_PATH_BATCHED)

    def set_model(self, eps_r, sigma):
        if hasattr(eps_r, "detach"):
            eps_r = eps_r.detach().cpu().numpy()
        if hasattr(sigma, "detach"):
            sigma = sigma.detach().cpu().numpy()
        e = np.asfortranarray(eps_r, dtype=np.float64)
        s = np.asfortranarray(sigma, dtype=np.float64)
        assert e.shape == (self.grid.m, self.grid.n) and s.shape == e.shape
        self.generation += 1
        N.check(N.load().gjw_set_model(self.h, N.ptr(e), N.ptr(s)), "gjw_set_model")

    def set_shots(self, src, wavelets, receivers):
        """src: (ns, 2) 1-based [iz, ix]; wavelets: (ns, nt); receivers: list of
        1-based Mw arrays (one per shot)."""
        src = np.ascontiguousarray(np.asarray(src, dtype=np.int32).reshape(-1, 2))
        ns = src.shape[0]
        wv = np.ascontiguousarray(np.asarray(wavelets, dtype=np.float32).reshape(ns, self.grid.nt))
        recs = [np.asarray(r, dtype=np.int32).ravel() for r in receivers]
        assert len(recs) == ns
        rp = np.zeros(ns + 1, dtype=np.int64)
        rp[1:] = np.cumsum([len(r) for r in recs])
        rl = np.ascontiguousarray(np.concatenate(recs) if rp[-1] else np.zeros(1, np.int32))
        self.generation += 1
        N.check(N.load().gjw_set_shots(self.h, ns, N.ptr(src), N.ptr(wv), N.ptr(rp), N.ptr(rl)),
                "gjw_set_shots")
        self.rec_ptr = rp
        self.ns = ns

    @property
    def ntraces(self):
        return int(self.rec_ptr[-1]) if self.rec_ptr is not None else 0

    def traces_shape(self):
        return (self.ntraces, self.grid.nt)

    # -- compute ------------------------------------------------------------
    def forward(self, keep=False, out=None, jz=None):
        """Traces of all shots, (ntraces, nt) float32: row t is trace t
        (shot-major), i.e. shot s's d_2d is out[rec_ptr[s]:rec_ptr[s+1]].T.
        jz=True (or an (ns, nt) float32 array / cuda tensor) also records the
        field at each shot's source cell per step (w_solve.m:104, gerjoii_.w.Jy)
        and returns (traces, jz)."""
        if out is None and self.rec_ptr is not None:
            out = np.zeros(self.traces_shape(), dtype=np.float32)
        if jz is True:
            jz = np.zeros((len(self.rec_ptr) - 1 if self.rec_ptr is not None else 0, self.grid.nt),
                          dtype=np.float32)
        _sync_torch(out, jz)
        self.generation += 1
        N.check(N.load().gjw_forward(self.h, N.ptr(out), N.ptr(jz) if jz is not None else None,
                                     int(bool(keep))), "gjw_forward")
        return out if jz is None else (out, jz)

    def misfit(self, d_obs, e_out=None, kind="L2"):
        """Adjoint source and misfit (FP64) of the last forward against d_obs
        (w_bozdag.m / w_E.m); kind is the reference's obj_FNC: 'L2',
        'env_L2', 'phase_L2', 'env_L22' or 'env_LOG'."""
        _sync_torch(d_obs, e_out)
        if e_out is None:
            e_out = np.zeros(self.traces_shape(), dtype=np.float32)
        J = C.c_double(0)
        d = d_obs if hasattr(d_obs, "data_ptr") else np.ascontiguousarray(d_obs, dtype=np.float32)
        N.check(N.load().gjw_misfit(self.h, N.ptr(d), N.MISFIT_KINDS[kind], N.ptr(e_out),
                                    C.byref(J)), "gjw_misfit")
        return e_out, J.value

    def bozdag(self, d_obs, d, kind="L2", e_out=None):
        """w_bozdag.m + w_E.m for caller-given synthetic traces d (they replace
        the survey's traces): returns (e, J)."""
        _sync_torch(d_obs, d, e_out)
        if e_out is None:
            e_out = np.zeros(self.traces_shape(), dtype=np.float32)
        J = C.c_double(0)
        a = d_obs if hasattr(d_obs, "data_ptr") else np.ascontiguousarray(d_obs, dtype=np.float32)
        b = d if hasattr(d, "data_ptr") else np.ascontiguousarray(d, dtype=np.float32)
        N.check(N.load().gjw_bozdag(self.h, N.ptr(a), N.ptr(b), N.MISFIT_KINDS[kind], N.ptr(e_out),
                                    C.byref(J)), "gjw_bozdag")
        return e_out, J.value

    def set_misfit(self, kind):
        """Misfit of fwi_gradient (batched path): an obj_FNC name as in misfit()."""
        self.set_option(N.GJW_OPT_MISFIT, N.MISFIT_KINDS[kind])

    def set_adj_filter(self, f_low, f_high, steep=8):
        """Band filter (filt_gauss.m) of the adjoint source in fwi_gradient,
        as w_update_e.m:69-70 applies it; f_high <= f_low turns it off."""
        N.check(N.load().gjw_set_adj_filter(self.h, float(f_low), float(f_high), int(steep)),
                "gjw_set_adj_filter")

    def set_shot_gradients(self, on=True):
        """GJW_OPT_SHOT_GRADS: gradients come back per shot, (m, n, ns)
        F-order, as w_update_e.m keeps g_e per source."""
        self.set_option(N.GJW_OPT_SHOT_GRADS, int(bool(on)))
        self._shot_grads = bool(on)

    def _grad_shape(self):
        m, n = self.grid.m, self.grid.n
        return (m, n, self.ns) if getattr(self, "_shot_grads", False) else (m, n)

    def gradient(self, adj=None, mode=N.GJW_GRAD_EXACT, want_gw=False, g_eps=None, g_sig=None):
        """FP64 gradients (m, n) w.r.t. eps_r and sigma (and optionally the
        wavelet gradient (ns, nt)) for adjoint source ``adj`` (traces layout);
        (m, n, ns) per shot after set_shot_gradients()."""
        if g_eps is None:
            g_eps = np.zeros(self._grad_shape(), dtype=np.float64, order="F")
        if g_sig is None:
            g_sig = np.zeros(self._grad_shape(), dtype=np.float64, order="F")
        gw = np.zeros((self.ns, self.grid.nt), dtype=np.float64) if want_gw else None
        if adj is not None and not hasattr(adj, "data_ptr"):
            adj = np.ascontiguousarray(adj, dtype=np.float32)
        _sync_torch(adj, g_eps, g_sig)
        N.check(N.load().gjw_gradient(self.h, N.ptr(adj), int(mode), N.ptr(g_eps), N.ptr(g_sig),
                                      N.ptr(gw), None), "gjw_gradient")
        return (g_eps, g_sig, gw) if want_gw else (g_eps, g_sig)

    def fwi_gradient(self, d_obs, g_eps=None, g_sig=None):
        """Forward + misfit (L2 unless set_misfit chose another) + exact
        gradient, all on the device."""
        if g_eps is None:
            g_eps = np.zeros(self._grad_shape(), dtype=np.float64, order="F")
        if g_sig is None:
            g_sig = np.zeros(self._grad_shape(), dtype=np.float64, order="F")
        if not hasattr(d_obs, "data_ptr"):
            d_obs = np.ascontiguousarray(d_obs, dtype=np.float32)
        _sync_torch(d_obs, g_eps, g_sig)
        J = C.c_double(0)
        self.generation += 1
        N.check(N.load().gjw_fwi_gradient(self.h, N.ptr(d_obs), N.ptr(g_eps), N.ptr(g_sig),
                                          C.byref(J)), "gjw_fwi_gradient")
        return g_eps, g_sig, J.value

    def kernel_stats(self, which, reset=False):
        """(average ms, samples, launches) of a step kernel while
        GJW_OPT_KTIMING is on: 0/1 one-step forward/adjoint, 2/3 two-step,
        4/5 four-step, 6 eight-step forward (forward sweep), 7 eight-step
        adjoint, 8 six-step adjoint, 9 eight-step forward with frames
        (recompute sweep)."""
        ms, n, nl = C.c_double(0), C.c_int64(0), C.c_int64(0)
        N.check(N.load().gjw_kernel_stats(self.h, int(which), C.byref(ms), C.byref(n), C.byref(nl),
                                          int(reset)), "gjw_kernel_stats")
        return ms.value, n.value, nl.value

    def query(self, what):
        """Effective setting (gjw_query): N.GJW_Q_CKPT_INTERVAL,
        N.GJW_Q_FWD_DEPTH or N.GJW_Q_ADJ_DEPTH."""
        v = C.c_int64(0)
        N.check(N.load().gjw_query(self.h, int(what), C.byref(v)), "gjw_query")
        return v.value

    def tile(self, which=2):
        """Autotuned columns per wavefront of step kernel `which` (numbered
        as in kernel_stats: 0/1 one-step forward/adjoint, 2/3 two-step,
        4/5 four-step, 6 eight-step forward)."""
        L = C.c_int32(0)
        N.check(N.load().gjw_tile(self.h, int(which), C.byref(L)), "gjw_tile")
        return L.value

    def last_timing(self):
        ms, cells = C.c_double(0), C.c_double(0)
        N.check(N.load().gjw_last_timing(self.h, C.byref(ms), C.byref(cells)), "gjw_last_timing")
        return ms.value, cells.value


def concurrent_fwi_gradients(surveys, d_obs):
    """FWI gradients of several surveys at once: each survey on its own
    Context (its own HIP stream), each call from its own host thread (the
    ctypes calls release the GIL), so the surveys' launch sequences overlap
    on the GPU with the single-survey kernels.  For surveys too small to fill
    the GPU alone it runs about as fast as the shared launches of Batch and
    well ahead of one survey at a time (six lines of 6-8 shots on a
    2000 x 600 grid: 0.87-1.00 s, against 0.97-0.99 s and 1.23 s); results
    equal the one-at-a-time ones bit for bit, and the surveys need not share
    a checkpoint interval.  Returns [(g_eps, g_sig, J)] in survey order."""
    import threading
    ctxs = [sv.ctx.h.value for sv in surveys]
    if len(set(ctxs)) != len(ctxs):
        raise ValueError("concurrent_fwi_gradients: give each survey its own Context")
    out = [None] * len(surveys)
    err = [None] * len(surveys)

    def run(k):
        try:
            out[k] = surveys[k].fwi_gradient(d_obs[k])
        except Exception as e:  # re-raised in the caller's thread
            err[k] = e

    th = [threading.Thread(target=run, args=(k,)) for k in range(len(surveys))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for e in err:
        if e is not None:
            raise e
    return out


class Batch:
    """Several surveys (own grid, model, shots, nt, misfit each) whose FWI
    gradients run in shared launches (gjw_batch_*)."""

    def __init__(self, surveys):
        self.surveys = list(surveys)
        ctx = self.surveys[0].ctx
        arr = (C.c_void_p * len(self.surveys))(*[sv.h for sv in self.surveys])
        h = C.c_void_p()
        N.check(N.load().gjw_batch_create(ctx.h, len(self.surveys), arr, C.byref(h)),
                "gjw_batch_create")
        self.h = h
        ctx._batches.add(self)

    def close(self):
        if self.h:
            N.load().gjw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def objective(self, d_obs, objectives=None, obj_fnc="L2", fstp=None):
        """Line-search objective of each survey's current model (w_objs_e.m):
        forward only, in shared launches, traces band-filtered when the
        survey has a filter.  d_obs: one traces array per survey.
        objectives: None for w_E(d_o, d, obj_fnc) / |d_o|^2, or a list of
        (obj_FNC name, weight) pairs, gerjoii_.w.obj_FNCs with g_e_weights
        ('FREQstp' filters with fstp = (f_low, f_high) first); obj_fnc is
        gerjoii_.w.obj_FNC (w_E's transform)."""
        S = len(self.surveys)
        d_obs = [d if hasattr(d, "data_ptr") else np.ascontiguousarray(d, dtype=np.float32)
                 for d in d_obs]
        _sync_torch(*d_obs)
        J = np.zeros(S, dtype=np.float64)
        P = C.c_void_p * S
        objectives = list(objectives or [])
        kinds = np.array([N.GJW_MISFIT_FREQSTP if k == "FREQstp" else N.MISFIT_KINDS[k]
                          for k, _ in objectives] or [0], dtype=np.int32)
        weights = np.array([float(w) for _, w in objectives] or [0.0], dtype=np.float64)
        lo, hi = fstp if fstp is not None else (0.0, 0.0)
        N.check(N.load().gjw_batch_objective_multi(
            self.h, P(*[N.ptr(d).value for d in d_obs]), len(objectives), N.ptr(kinds),
            N.ptr(weights), N.MISFIT_KINDS[obj_fnc], float(lo), float(hi), N.ptr(J)),
            "gjw_batch_objective_multi")
        return J

    def forward(self, out=None):
        """Traces of each survey's current model (forward in shared
        launches, no wavefield): a list of (ntraces, nt) float32 arrays."""
        if out is None:
            out = [np.zeros(sv.traces_shape(), dtype=np.float32) for sv in self.surveys]
        _sync_torch(*out)
        P = C.c_void_p * len(self.surveys)
        N.check(N.load().gjw_batch_forward(self.h, P(*[N.ptr(o).value for o in out])),
                "gjw_batch_forward")
        return out

    def fwi_gradient(self, d_obs, g_eps=None, g_sig=None):
        """d_obs: one traces array per survey.  Returns (g_eps list, g_sig
        list, J array)."""
        S = len(self.surveys)
        if g_eps is None:
            g_eps = [np.zeros(sv._grad_shape(), dtype=np.float64, order="F")
                     for sv in self.surveys]
        if g_sig is None:
            g_sig = [np.zeros(sv._grad_shape(), dtype=np.float64, order="F")
                     for sv in self.surveys]
        d_obs = [d if hasattr(d, "data_ptr") else np.ascontiguousarray(d, dtype=np.float32)
                 for d in d_obs]
        _sync_torch(*d_obs, *g_eps, *g_sig)
        J = np.zeros(S, dtype=np.float64)
        P = C.c_void_p * S
        N.check(N.load().gjw_batch_fwi_gradient(self.h, P(*[N.ptr(d).value for d in d_obs]),
                                                P(*[N.ptr(g).value for g in g_eps]),
                                                P(*[N.ptr(g).value for g in g_sig]), N.ptr(J)),
                "gjw_batch_fwi_gradient")
        return g_eps, g_sig, J

    def last_timing(self):
        ms, cells = C.c_double(0), C.c_double(0)
        N.check(N.load().gjw_batch_last_timing(self.h, C.byref(ms), C.byref(cells)),
                "gjw_batch_last_timing")
        return ms.value, cells.value
