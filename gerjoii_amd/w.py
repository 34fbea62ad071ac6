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
air}; finite_.w.{pml_w, pml_d, pml_order, R_0}; gerjoii_.w.{s, M, wvlet_,
obj_FNC}; parame_.natu.w.d_2d.  Fields written: gerjoii_.w.{d_2d, e_2d, u_2d}.
``u_2d`` is a handle to the wavefield held on the GPU (the plain path's
stored domain wavefield); ``w_grad`` accepts that handle, or the string
'dtu' variant, the way w_update_e.m passes differentiate_cube(u_2d).
Errors raise ``gerjoii_amd._native.GjwError`` where MATLAB would error().
"""
from __future__ import annotations

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
    """src/w/w_fwd.m:17-35: pad, coefficients, solve, misfit (L2)."""
    g = _grid(geome_, parame_, finite_)
    if "dy" in geome_ and float(geome_["dy"]) != g.dx:
        raise N.GjwError("dy must equal dx (w_geom.m:116)")
    sv = Survey(g, context(0))
    sv.set_model(parame_["w"]["epsilon"], parame_["w"]["sigma"])
    gw = gerjoii_["w"]
    sv.set_shots(np.asarray(gw["s"]).reshape(1, 2), np.asarray(gw["wvlet_"]).reshape(1, -1),
                 [np.asarray(gw["M"]).ravel()])
    sv.set_path(True)
    d, jz = sv.forward(keep=keep_wavefield, jz=True)
    gerjoii_ = dict(gerjoii_)
    gerjoii_["w"] = dict(gw)
    gerjoii_["w"]["d_2d"] = d.T.copy()                    # nt x nr
    gerjoii_["w"]["Jy"] = jz[0].copy()                     # source recorder (w_solve.m:104,124)
    gerjoii_["w"]["u_2d"] = WavefieldHandle(sv) if keep_wavefield else None
    obj = gw.get("obj_FNC", "L2")
    d_o = parame_.get("natu", {}).get("w", {}).get("d_2d")
    if d_o is not None:
        e, _, _ = w_bozdag(d_o, gerjoii_["w"]["d_2d"], obj)
        gerjoii_["w"]["e_2d"] = e
    return finite_, gerjoii_


def w_grad(u_w, geome_, finite_, gerjoii_, parame_, mode="REF"):
    """src/w/w_grad.m: gradient for one shot.  ``u_w`` is the handle left by
    w_fwd (for the conductivity gradient, w_update_s.m:77) or
    differentiate_cube(handle) (for the permittivity gradient,
    w_update_e.m:83-98).  mode='REF' reproduces w_grad.m; mode='EXACT' returns
    the exact discrete gradient instead."""
    if not isinstance(u_w, WavefieldHandle):
        raise N.GjwError("w_grad needs the wavefield handle returned in gerjoii_.w.u_2d")
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


def _analytic(x):
    """MATLAB hilbert along the time axis (columns): FFT of length nt."""
    x = np.asarray(x, dtype=np.float64)
    Nn = x.shape[0]
    X = np.fft.fft(x, axis=0)
    h = np.zeros(Nn)
    h[0] = 1
    if Nn % 2 == 0:
        h[Nn // 2] = 1
        h[1:Nn // 2] = 2
    else:
        h[1:(Nn + 1) // 2] = 2
    return np.fft.ifft(X * h.reshape((-1,) + (1,) * (x.ndim - 1)), axis=0)


def w_bozdag(d_o, d, obj_fnc="L2"):
    """src/w/w_bozdag.m:13-55 (L2 on the GPU path; the Hilbert-based misfits
    evaluated on the host in FP64)."""
    d_o = np.asarray(d_o, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    if obj_fnc in ("L2", "FREQstp"):
        return (d - d_o).astype(np.float32), d, d_o
    with np.errstate(divide="ignore", invalid="ignore"):
        A = _analytic(d)
        d_e, d_ph, d_H = np.abs(A), np.angle(A), np.imag(A)
        Ao = _analytic(d_o)
        if obj_fnc == "env_L2":
            dedoe = (d_e - np.abs(Ao)) / d_e
            e = d * dedoe - np.imag(_analytic(dedoe * d_H))
            e[np.isnan(e)] = 0
            return e.astype(np.float32), d_e, np.abs(Ao)
        if obj_fnc == "phase_L2":
            d_do_ = (d_ph - np.angle(Ao)) / d_e**2
            e = -d_do_ + np.imag(_analytic(d_do_ * d))
            return e.astype(np.float32), d_ph, np.angle(Ao)
        if obj_fnc == "env_L22":
            dedoe = d_e**2 - np.abs(Ao) ** 2
            e = d * dedoe - np.imag(_analytic(dedoe * d_H))
            return e.astype(np.float32), d_e, np.abs(Ao)
        if obj_fnc == "env_LOG":
            dedoe = np.real(np.log((np.abs(Ao) / d_e).astype(complex)))
            e = -dedoe * (d / d_e**2) + np.imag(_analytic(dedoe * (d_H / d_e**2)))
            e[np.isnan(e)] = 0
            return e.astype(np.float32), d_e, np.abs(Ao)
    raise N.GjwError(f"w_bozdag: unknown objective {obj_fnc!r}")


def w_E(d_o, d, Nw=None, obj_fnc="L2"):
    """src/w/w_E.m:7-16: 0.5 e' N e in FP64 (N = identity when None)."""
    d = np.asarray(d, dtype=np.float64).ravel(order="F")
    d_o = np.asarray(d_o, dtype=np.float64).ravel(order="F")
    if obj_fnc == "env_L22":
        d, d_o = d**2, d_o**2
    elif obj_fnc == "env_LOG":
        d, d_o = np.log(d), np.log(d_o)
    e = d - d_o
    return 0.5 * float(e @ e) if Nw is None else 0.5 * float(e @ (Nw @ e))
