"""Slow, slice-for-slice numpy reading of the reference solver -- TEST ONLY.

A second, independent restatement of w_exp2pml.m, w_cpml.m, w_coeff.m,
w_solve.m and w_grad.m that keeps MATLAB's vectorised slices (``S(a, b)`` is
MATLAB's ``a:b``).  Run on tiny grids it pins the C oracle's loop code bit for
bit in FP64: the two readings share no code, so an index slip in either shows
up as a mismatch.
"""
import numpy as np


def S(a, b):
    """MATLAB a:b (1-based, inclusive) as a Python slice."""
    return slice(a - 1, b)


def w_exp2pml(eps_model, sig_model, P, air):
    """w_exp2pml.m:17-37."""
    m, n = eps_model.shape
    eps_r_z = np.vstack([np.ones((air, n)), eps_model])
    sigma_e_z = np.vstack([np.zeros((air, n)), sig_model])
    left = np.ones((P, 1)) @ eps_r_z[:, [0]].T
    right = np.ones((P, 1)) @ eps_r_z[:, [-1]].T
    eps_r_z = np.hstack([left.T, eps_r_z, right.T])
    sigma_e_z = np.hstack([np.zeros((air + m, P)), sigma_e_z, np.zeros((air + m, P))])
    up = np.ones((P, 1)) @ eps_r_z[[0], :]
    down = np.ones((P, 1)) @ eps_r_z[[-1], :]
    eps_r_z = np.vstack([up, eps_r_z, down])
    sigma_e_z = np.vstack([np.zeros((P, n + 2 * P)), sigma_e_z, np.zeros((P, n + 2 * P))])
    return eps_r_z, sigma_e_z


def w_cpml(eps_r_z, nx, ny, P, dx, dt, eps_0, mu_0, R_0, order):
    """w_cpml.m:44-185 (pml_w = pml_d = P, dy = dx)."""
    dy = dx
    c = 1 / np.sqrt(mu_0 * eps_0)
    pis, pie, pjs, pje = P + 1, nx - P + 1, P + 1, ny - P + 1
    nxm1, nym1 = nx - 1, ny - 1
    k = {}
    sigma_max = -(order + 1) * 1 * eps_0 * (c / np.sqrt(1)) * np.log(R_0) / (2 * dx * P)
    rho_e = (np.arange(P, 0, -1) - 0.75) / P
    rho_m = (np.arange(P, 0, -1) - 0.25) / P
    sp_e = np.zeros((P, nym1))
    sp_m = np.zeros((P, nym1))
    for ind in range(P):
        sp_e[ind, :] = sigma_max * rho_e[ind] ** order
        sp_m[ind, :] = (mu_0 / eps_0) * sigma_max * rho_m[ind] ** order
    k["Chyh_xn"] = (2 * mu_0 - dt * sp_m) / (2 * mu_0 + dt * sp_m)
    k["Chyez_xn"] = (2 * dt / dx) / (2 * mu_0 + dt * sp_m)
    k["Cezxe_xn"] = (2 * eps_0 - dt * sp_e) / (2 * eps_0 + dt * sp_e)
    k["Cezxhy_xn"] = (2 * dt / dx) / (2 * eps_0 + dt * sp_e)
    k["Cezye_xn"] = 1.0
    k["Cezyhx_xn"] = -dt / (dy * eps_0)
    rho_e = (np.arange(1, P + 1) - 0.75) / P
    rho_m = (np.arange(1, P + 1) - 0.25) / P
    sp_e = np.zeros((P, nym1))
    sp_m = np.zeros((P, nym1))
    for ind in range(P):
        sp_e[ind, :] = sigma_max * rho_e[ind] ** order
        sp_m[ind, :] = (mu_0 / eps_0) * sigma_max * rho_m[ind] ** order
    sp_m = sp_m / eps_r_z[S(pie, nx), S(2, ny)]
    k["Chyh_xp"] = (2 * mu_0 - dt * sp_m) / (2 * mu_0 + dt * sp_m)
    k["Chyez_xp"] = (2 * dt / dx) / (2 * mu_0 + dt * sp_m)
    e = eps_r_z[S(pie, nx), S(2, ny)]
    k["Cezxe_xp"] = (2 * eps_0 * e - dt * sp_e) / (2 * eps_0 * e + dt * sp_e)
    k["Cezxhy_xp"] = (2 * dt / dx) / (2 * eps_0 * e + dt * sp_e)
    k["Cezye_xp"] = 1.0
    k["Cezyhx_xp"] = -dt / (dy * eps_0 * eps_r_z[S(pie, nx), S(pjs + 1, pje - 1)])
    sigma_max = -(order + 1) * 1 * eps_0 * (c / np.sqrt(1)) * np.log(R_0) / (2 * dy * P)
    rho_e = (np.arange(P, 0, -1) - 0.75) / P
    rho_m = (np.arange(P, 0, -1) - 0.25) / P
    sp_e = np.zeros((nxm1, P))
    sp_m = np.zeros((nxm1, P))
    for ind in range(P):
        sp_e[:, ind] = sigma_max * rho_e[ind] ** order
        sp_m[:, ind] = (mu_0 / eps_0) * sigma_max * rho_m[ind] ** order
    sp_m = sp_m / eps_r_z[S(2, nx), S(1, pjs - 1)]
    k["Chxh_yn"] = (2 * mu_0 - dt * sp_m) / (2 * mu_0 + dt * sp_m)
    k["Chxez_yn"] = -(2 * dt / dy) / (2 * mu_0 + dt * sp_m)
    k["Cezxe_yn"] = 1.0
    k["Cezxhy_yn"] = dt / (dx * eps_0 * eps_r_z[S(pis + 1, pie - 1), S(2, pjs)])
    e = eps_r_z[S(2, nx), S(2, pjs)]
    k["Cezye_yn"] = (2 * eps_0 * e - dt * sp_e) / (2 * eps_0 * e + dt * sp_e)
    k["Cezyhx_yn"] = -(2 * dt / dy) / (2 * eps_0 * e + dt * sp_e)
    rho_e = (np.arange(1, P + 1) - 0.75) / P
    rho_m = (np.arange(1, P + 1) - 0.25) / P
    sp_e = np.zeros((nxm1, P))
    sp_m = np.zeros((nxm1, P))
    for ind in range(P):
        sp_e[:, ind] = sigma_max * rho_e[ind] ** order
        sp_m[:, ind] = (mu_0 / eps_0) * sigma_max * rho_m[ind] ** order
    sp_m = sp_m / eps_r_z[1:, S(pje, ny)]
    k["Chxh_yp"] = (2 * mu_0 - dt * sp_m) / (2 * mu_0 + dt * sp_m)
    k["Chxez_yp"] = -(2 * dt / dy) / (2 * mu_0 + dt * sp_m)
    k["Cezxe_yp"] = 1.0
    k["Cezxhy_yp"] = dt / (dx * eps_0 * eps_r_z[S(pis + 1, pie - 1), S(pje, ny)])
    e = eps_r_z[S(2, nx), S(pje, ny)]
    k["Cezye_yp"] = (2 * eps_0 * e - dt * sp_e) / (2 * eps_0 * e + dt * sp_e)
    k["Cezyhx_yp"] = -(2 * dt / dy) / (2 * eps_0 * e + dt * sp_e)
    return k


def w_coeff(eps_r_z, sigma_e_z, s, dx, dt, eps_0, mu_0):
    """w_coeff.m:28-55 (mu_r = 1, sigma_m = 0)."""
    dy = dx
    k = {}
    den = 2 * eps_r_z * eps_0 + dt * sigma_e_z
    k["Ceze"] = (2 * eps_r_z * eps_0 - dt * sigma_e_z) / den
    k["Cezhy"] = (2 * dt / dx) / den
    k["Cezhx"] = -(2 * dt / dy) / den
    k["Chxh"] = (2 * 1 * mu_0 - dt * 0) / (2 * 1 * mu_0 + dt * 0)
    k["Chxez"] = -(2 * dt / dy) / (2 * 1 * mu_0 + dt * 0)
    k["Chyh"] = (2 * 1 * mu_0 - dt * 0) / (2 * 1 * mu_0 + dt * 0)
    k["Chyez"] = (2 * dt / dx) / (2 * 1 * mu_0 + dt * 0)
    i, j = s
    k["Cezj"] = -(2 * dt) / (2 * eps_r_z[i - 1, j - 1] * eps_0 + dt * sigma_e_z[i - 1, j - 1])
    return k


def _step(k, Ez, Hx, Hy, E, nx, ny, P):
    """new_H, new_H_pml, new_Ez of w_solve.m:136-207 (in place)."""
    pis, pie, pjs, pje = P + 1, nx - P + 1, P + 1, ny - P + 1
    Hx[:, S(pjs, pje - 1)] = k["Chxh"] * Hx[:, S(pjs, pje - 1)] + k["Chxez"] * (
        Ez[:, S(pjs + 1, pje)] - Ez[:, S(pjs, pje - 1)])
    Hy[S(pis, pie - 1), :] = k["Chyh"] * Hy[S(pis, pie - 1), :] + k["Chyez"] * (
        Ez[S(pis + 1, pie), :] - Ez[S(pis, pie - 1), :])
    Hy[S(1, pis - 1), S(2, ny)] = k["Chyh_xn"] * Hy[S(1, pis - 1), S(2, ny)] + k["Chyez_xn"] * (
        Ez[S(2, pis), S(2, ny)] - Ez[S(1, pis - 1), S(2, ny)])
    Hy[S(pie, nx), S(2, ny)] = k["Chyh_xp"] * Hy[S(pie, nx), S(2, ny)] + k["Chyez_xp"] * (
        Ez[S(pie + 1, nx + 1), S(2, ny)] - Ez[S(pie, nx), S(2, ny)])
    Hx[S(2, nx), S(1, pjs - 1)] = k["Chxh_yn"] * Hx[S(2, nx), S(1, pjs - 1)] + k["Chxez_yn"] * (
        Ez[S(2, nx), S(2, pjs)] - Ez[S(2, nx), S(1, pjs - 1)])
    Hx[S(2, nx), S(pje, ny)] = k["Chxh_yp"] * Hx[S(2, nx), S(pje, ny)] + k["Chxez_yp"] * (
        Ez[S(2, nx), S(pje + 1, ny + 1)] - Ez[S(2, nx), S(pje, ny)])
    Ez[S(pis + 1, pie - 1), S(pjs + 1, pje - 1)] = (
        k["Ceze"][S(pis + 1, pie - 1), S(pjs + 1, pje - 1)] * Ez[S(pis + 1, pie - 1), S(pjs + 1, pje - 1)]
        + k["Cezhy"][S(pis + 1, pie - 1), S(pjs + 1, pje - 1)]
        * (Hy[S(pis + 1, pie - 1), S(pjs + 1, pje - 1)] - Hy[S(pis, pie - 2), S(pjs + 1, pje - 1)])
        + k["Cezhx"][S(pis + 1, pie - 1), S(pjs + 1, pje - 1)]
        * (Hx[S(pis + 1, pie - 1), S(pjs + 1, pje - 1)] - Hx[S(pis + 1, pie - 1), S(pjs, pje - 2)]))


def _pml(k, Ez, Hx, Hy, E, nx, ny, P):
    """new_Ez_pml of w_solve.m:215-268 (in place)."""
    pis, pie, pjs, pje = P + 1, nx - P + 1, P + 1, ny - P + 1
    nxm1, nym1 = nx - 1, ny - 1
    E["Ezx_xn"] = k["Cezxe_xn"] * E["Ezx_xn"] + k["Cezxhy_xn"] * (
        Hy[S(2, pis), S(2, ny)] - Hy[S(1, pis - 1), S(2, ny)])
    E["Ezy_xn"] = k["Cezye_xn"] * E["Ezy_xn"] + k["Cezyhx_xn"] * (
        Hx[S(2, pis), S(pjs + 1, pje - 1)] - Hx[S(2, pis), S(pjs, pje - 2)])
    E["Ezx_xp"] = k["Cezxe_xp"] * E["Ezx_xp"] + k["Cezxhy_xp"] * (
        Hy[S(pie, nx), S(2, ny)] - Hy[S(pie - 1, nx - 1), S(2, ny)])
    E["Ezy_xp"] = k["Cezye_xp"] * E["Ezy_xp"] + k["Cezyhx_xp"] * (
        Hx[S(pie, nx), S(pjs + 1, pje - 1)] - Hx[S(pie, nx), S(pjs, pje - 2)])
    E["Ezx_yn"] = k["Cezxe_yn"] * E["Ezx_yn"] + k["Cezxhy_yn"] * (
        Hy[S(pis + 1, pie - 1), S(2, pjs)] - Hy[S(pis, pie - 2), S(2, pjs)])
    E["Ezy_yn"] = k["Cezye_yn"] * E["Ezy_yn"] + k["Cezyhx_yn"] * (
        Hx[S(2, nx), S(2, pjs)] - Hx[S(2, nx), S(1, pjs - 1)])
    E["Ezx_yp"] = k["Cezxe_yp"] * E["Ezx_yp"] + k["Cezxhy_yp"] * (
        Hy[S(pis + 1, pie - 1), S(pje, ny)] - Hy[S(pis, pie - 2), S(pje, ny)])
    E["Ezy_yp"] = k["Cezye_yp"] * E["Ezy_yp"] + k["Cezyhx_yp"] * (
        Hx[S(2, nx), S(pje, ny)] - Hx[S(2, nx), S(pje - 1, ny - 1)])
    Ez[S(2, pis), S(2, pjs)] = E["Ezx_xn"][:, S(1, pjs - 1)] + E["Ezy_yn"][S(1, pis - 1), :]
    Ez[S(2, pis), S(pje, ny)] = E["Ezx_xn"][:, S(pje - 1, nym1)] + E["Ezy_yp"][S(1, pis - 1), :]
    Ez[S(pie, nx), S(pje, ny)] = E["Ezx_xp"][:, S(pje - 1, nym1)] + E["Ezy_yp"][S(pie - 1, nxm1), :]
    Ez[S(pie, nx), S(2, pjs)] = E["Ezx_xp"][:, S(1, pjs - 1)] + E["Ezy_yn"][S(pie - 1, nxm1), :]
    Ez[S(pis + 1, pie - 1), S(2, pjs)] = E["Ezx_yn"] + E["Ezy_yn"][S(pis, pie - 2), :]
    Ez[S(pis + 1, pie - 1), S(pje, ny)] = E["Ezx_yp"] + E["Ezy_yp"][S(pis, pie - 2), :]
    Ez[S(2, pis), S(pjs + 1, pje - 1)] = E["Ezx_xn"][:, S(pjs, pje - 2)] + E["Ezy_xn"]
    Ez[S(pie, nx), S(pjs + 1, pje - 1)] = E["Ezx_xp"][:, S(pjs, pje - 2)] + E["Ezy_xp"]


def _zeros(nx, ny, P):
    nxm1, nym1 = nx - 1, ny - 1
    E = {"Ezx_xn": np.zeros((P, nym1)), "Ezy_xn": np.zeros((P, nym1 - 2 * P)),
         "Ezx_xp": np.zeros((P, nym1)), "Ezy_xp": np.zeros((P, nym1 - 2 * P)),
         "Ezx_yn": np.zeros((nxm1 - 2 * P, P)), "Ezy_yn": np.zeros((nxm1, P)),
         "Ezx_yp": np.zeros((nxm1 - 2 * P, P)), "Ezy_yp": np.zeros((nxm1, P))}
    return np.zeros((nx + 1, ny + 1)), np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), E


def w_solve(eps_r_z, sigma_e_z, s, wvlet, Mw, P, air, dx, dt, eps_0, mu_0,
            R_0=1e-8, order=2):
    """w_solve.m:80-118: returns (d [nt x nr], Jz [nt], u [m x n x nt])."""
    nx, ny = eps_r_z.shape
    nt = len(wvlet)
    k = w_cpml(eps_r_z, nx, ny, P, dx, dt, eps_0, mu_0, R_0, order)
    k.update(w_coeff(eps_r_z, sigma_e_z, s, dx, dt, eps_0, mu_0))
    Ez, Hx, Hy, E = _zeros(nx, ny, P)
    pis, pie, pjs, pje = P + 1, nx - P + 1, P + 1, ny - P + 1
    di = np.arange(pis + air, pie) - 1
    dj = np.arange(pjs, pje) - 1
    Mw0 = np.asarray(Mw) - 1
    d = np.zeros((nt, len(Mw0)))
    u = np.zeros((len(di), len(dj), nt))
    Jz = np.zeros(nt)
    for it in range(nt):
        _step(k, Ez, Hx, Hy, E, nx, ny, P)
        Ez[s[0] - 1, s[1] - 1] = Ez[s[0] - 1, s[1] - 1] + k["Cezj"] * wvlet[it]
        _pml(k, Ez, Hx, Hy, E, nx, ny, P)
        d[it, :] = Ez.ravel(order="F")[Mw0]
        u[:, :, it] = Ez[np.ix_(di, dj)]
        Jz[it] = Ez[s[0] - 1, s[1] - 1]
    return d, Jz, u


def w_grad(u_w, e_w, eps_r_z, sigma_e_z, s, Mw, P, air, dx, dt, eps_0, mu_0,
           R_0=1e-8, order=2):
    """w_grad.m:20-115: returns (g_w, a_max)."""
    nx, ny = eps_r_z.shape
    nt = e_w.shape[0]
    k = w_cpml(eps_r_z, nx, ny, P, dx, dt, eps_0, mu_0, R_0, order)
    k.update(w_coeff(eps_r_z, -sigma_e_z, s, dx, dt, eps_0, mu_0))
    Ez, Hx, Hy, E = _zeros(nx, ny, P)
    pis, pie, pjs, pje = P + 1, nx - P + 1, P + 1, ny - P + 1
    di = np.arange(pis + air, pie) - 1
    dj = np.arange(pjs, pje) - 1
    Mw0 = np.asarray(Mw) - 1
    g_w = np.zeros((len(di), len(dj)))
    a_w = None
    for it in range(nt):
        Ez[Mw0 % (nx + 1), Mw0 // (nx + 1)] = e_w[nt - it - 1, :]
        _step(k, Ez, Hx, Hy, E, nx, ny, P)
        _pml(k, Ez, Hx, Hy, E, nx, ny, P)
        a_w = Ez[np.ix_(di, dj)]
        g_w = g_w - (u_w[:, :, nt - it - 1] * a_w)
    return dt * g_w, np.abs(a_w)
