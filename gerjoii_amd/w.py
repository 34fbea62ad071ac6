"""The reference's MATLAB entry points, same names and struct fields, over libgjw.

Each function takes and returns dicts that stand for MATLAB's structs
(``geome_``, ``parame_``, ``finite_``, ``gerjoii_``) with the field names the
reference uses, so a Python port of a gerjoii driver calls them as the MATLAB
code does:

    finite_, gerjoii_ = w_fwd(geome_, parame_, finite_, gerjoii_)     # src/w/w_fwd.m:1
    g_w, a_max = w_grad(u_w, geome_, finite_, gerjoii_, parame_)      # src/w/w_grad.m:1
    e, d_e, d_oe = w_bozdag(d_o, d, obj_fnc)                          # src/w/w_bozdag.m:1
    Ew = w_E(d_o, d, Nw, obj_fnc)                                     # src/w/w_E.m:1
    v = differentiate_cube(v, dt)                    # src/shared/w-dc/differentiate_cube.m:1

Fields read: geome_.{dx, w.dt, w.nt}; parame_.w.{epsilon, sigma, eps_0, mu_0,
air, t}; finite_.w.{pml_w, pml_d, pml_order, R_0}; gerjoii_.w.{s, M, wvlet_,
obj_FNC, MUTE, v_mute, t_mute, r_real}; parame_.natu.w.d_2d.  Fields written:
gerjoii_.w.{d_2d, e_2d, u_2d, Jy}.  As in MATLAB, w_fwd returns finite_ and
gerjoii_ only: the muted observed data it computes stays local.  ``u_2d`` is
a handle to the wavefield held on the GPU (the plain path's stored domain
wavefield); ``w_grad`` accepts that handle, or the handle marked by
differentiate_cube, the way w_update_e.m passes differentiate_cube(u_2d).
The solve, the mute, the misfits and the gradients run in libgjw; errors
raise ``gerjoii_amd._native.GjwError`` where MATLAB would error().
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _native as N
from .survey import Grid, Survey, context


class WavefieldHandle:
    """Stands in for gerjoii_.w.u_2d: the forward wavefield stays on the GPU."""

    def __init__(self, survey, dtu=False):
        self.survey = survey
        self.dtu = dtu


def _grid(geome_, parame_, finite_):
    fw = finite_["w"]
    if int(fw["pml_w"]) != int(fw["pml_d"]):
        raise N.GjwError("pml_w must equal pml_d (w_exp2pml.m pads columns with pml_w)")
    m, n = np.asarray(parame_["w"]["epsilon"]).shape
    return Grid(m=m, n=n, pml=int(fw["pml_w"]), air=int(parame_["w"].get("air", 0)),
                nt=int(geome_["w"]["nt"]), dx=float(geome_["dx"]), dt=float(geome_["w"]["dt"]),
                eps0=float(parame_["w"]["eps_0"]), mu0=float(parame_["w"]["mu_0"]),
                R0=float(fw.get("R_0", 1e-8)), pml_order=int(fw.get("pml_order", 2)))


def w_fwd(geome_, parame_, finite_, gerjoii_, keep_wavefield=True):
    """src/w/w_fwd.m:17-35: pad, coefficients and solve (w_exp2pml, w_cpml,
    w_coeff, w_solve), the mute of w_lmute.m when gerjoii_.w.MUTE is
    'yes_MUTE', then [e_2d, d_2d, ~] = w_bozdag(d_o, d_2d, obj_FNC): d_2d
    becomes w_bozdag's second output (the envelope or the phase for the
    Hilbert-based objectives), as in the reference."""
    g = _grid(geome_, parame_, finite_)
    if "dy" in geome_ and float(geome_["dy"]) != g.dx:
        raise N.GjwError("dy must equal dx (w_geom.m:116)")
    ctx = context(0)
    sv = Survey(g, ctx)
    sv.set_model(parame_["w"]["epsilon"], parame_["w"]["sigma"])
    gw = gerjoii_["w"]
    sv.set_shots(np.asarray(gw["s"]).reshape(1, 2), np.asarray(gw["wvlet_"]).reshape(1, -1),
                 [np.asarray(gw["M"]).ravel()])
    sv.set_path(True)
    d, jz = sv.forward(keep=keep_wavefield, jz=True)   # d: (nr, nt) traces
    gerjoii_ = dict(gerjoii_)
    gerjoii_["w"] = gw = dict(gw)
    gw["Jy"] = jz[0].copy()                             # source recorder (w_solve.m:104,124)
    gw["u_2d"] = WavefieldHandle(sv) if keep_wavefield else None
    d_o = parame_.get("natu", {}).get("w", {}).get("d_2d")
    if d_o is None:
        gw["d_2d"] = d.T.copy(order="F")                # nt x nr
        return finite_, gerjoii_
    d_o = np.ascontiguousarray(np.asarray(d_o, dtype=np.float32).T)  # (nr, nt)
    if d_o.shape != d.shape:
        raise N.GjwError(f"parame_.natu.w.d_2d is {d_o.shape[::-1]}, the solve gives {d.shape[::-1]}")
    if gw.get("MUTE") == "yes_MUTE":
        # w_lmute.m:14-44 on the device: both gathers muted in place
        rx = np.asarray(gw["r_real"], dtype=np.float64)
        rx = rx[:, 0] if rx.ndim == 2 else rx
        ctx.lmute(d, d_o, np.asarray(parame_["w"]["t"], dtype=np.float64).ravel(), rx,
                  float(gw["v_mute"]), float(gw["t_mute"]))
    obj = gw.get("obj_FNC", "L2")
    gw["obj_FNC"] = obj                                 # w_fwd.m:33
    e, d_e, _, _ = ctx.w_bozdag(d_o.T, d.T, obj)
    gw["e_2d"] = e
    gw["d_2d"] = d_e
    return finite_, gerjoii_


def ref_growth_exponent(parame_, geome_):
    """Largest exponent of the per-cell growth that w_grad.m's negated
    conductivity gives its backward run: each step multiplies by
    (2 eps eps0 + dt sigma) / (2 eps eps0 - dt sigma) ~ exp(dt sigma / (eps eps0)),
    so nt steps by exp(sigma T / (eps eps0)), T = nt dt."""
    eps = np.asarray(parame_["w"]["epsilon"], dtype=np.float64)
    sig = np.asarray(parame_["w"]["sigma"], dtype=np.float64)
    T = float(geome_["w"]["nt"]) * float(geome_["w"]["dt"])
    return float(np.max(sig * T / (eps * float(parame_["w"]["eps_0"]))))


# FP32 overflows at e^88.7; a field of order 1e3 amplified by e^60 stays below
# it with margin
REF_GROWTH_LIMIT = 60.0


def w_grad(u_w, geome_, finite_, gerjoii_, parame_, mode="auto"):
    """src/w/w_grad.m: gradient for one shot.  ``u_w`` is the handle left by
    w_fwd (for the conductivity gradient, w_update_s.m:77) or
    differentiate_cube(handle) (for the permittivity gradient,
    w_update_e.m:83-98).

    mode 'REF' reproduces w_grad.m (negated conductivity, residual written
    into Ez at the receivers, correlation with u or d_t u) and also returns
    a_max; mode 'EXACT' returns the exact discrete gradient (the transpose
    of the forward solve; a_max is None).  mode 'auto' (the default) takes
    REF while its backward growth exp(sigma T / (eps eps0)) stays within
    e^60, and EXACT with a warning beyond it, where REF would overflow FP32
    (the reference's own arrays are single precision)."""
    if not isinstance(u_w, WavefieldHandle):
        raise N.GjwError("w_grad needs the wavefield handle returned in gerjoii_.w.u_2d")
    if mode == "auto":
        k = ref_growth_exponent(parame_, geome_)
        mode = "REF" if k <= REF_GROWTH_LIMIT else "EXACT"
        if mode == "EXACT":
            warnings.warn(f"w_grad: the REF backward run would grow as e^{k:.0f} (FP32 overflows "
                          f"at e^88); using the EXACT gradient", RuntimeWarning, stacklevel=2)
    if mode not in ("REF", "EXACT"):
        raise N.GjwError(f"w_grad: unknown mode {mode!r}")
    sv = u_w.survey
    e = np.ascontiguousarray(np.asarray(gerjoii_["w"]["e_2d"], dtype=np.float32).T)
    m, n = sv.grid.m, sv.grid.n
    ge = np.zeros((m, n), order="F")
    gs = np.zeros((m, n), order="F")
    if mode == "REF":
        am = np.zeros((m, n), dtype=np.float32, order="F")
        N.check(N.load().gjw_gradient(sv.h, N.ptr(e), N.GJW_GRAD_REF, N.ptr(ge), N.ptr(gs), None,
                                      N.ptr(am)), "w_grad")
        return (ge if u_w.dtu else gs), am
    N.check(N.load().gjw_gradient(sv.h, N.ptr(e), N.GJW_GRAD_EXACT, N.ptr(ge), N.ptr(gs), None,
                                  None), "w_grad")
    return (ge if u_w.dtu else gs), None


def differentiate_cube(v, dt):
    """differentiate_cube.m.  On a wavefield handle, marks it as d_t u (the
    derivative is taken on the GPU inside w_grad); on an array, differentiates
    it on the GPU."""
    if isinstance(v, WavefieldHandle):
        return WavefieldHandle(v.survey, dtu=True)
    v = np.asfortranarray(v, dtype=np.float32).copy(order="F")
    return context(0).differentiate_cube(v, dt)


def w_bozdag(d_o, d, obj_fnc="L2"):
    """src/w/w_bozdag.m:13-55 on the GPU (gjw_w_bozdag): returns (e_, d_e,
    d_oe), nt x nr float32."""
    if obj_fnc not in N.MISFIT_KINDS:
        raise N.GjwError(f"w_bozdag: unknown objective {obj_fnc!r}")
    e, de, doe, _ = context(0).w_bozdag(d_o, d, obj_fnc)
    return e, de, doe


def w_E(d_o, d, Nw=None, obj_fnc="L2"):
    """src/w/w_E.m:7-16: 0.5 e' N e in FP64 (N = identity when None); for
    env_L22 / env_LOG the inputs are squared / logged first, as w_E.m does."""
    d = np.asarray(d, dtype=np.float64).ravel(order="F")
    d_o = np.asarray(d_o, dtype=np.float64).ravel(order="F")
    if obj_fnc == "env_L22":
        d, d_o = d**2, d_o**2
    elif obj_fnc == "env_LOG":
        d, d_o = np.log(d), np.log(d_o)
    e = d - d_o
    return 0.5 * float(e @ e) if Nw is None else 0.5 * float(e @ (Nw @ e))
