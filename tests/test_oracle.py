"""CPU tests that pin the oracle (no GPU).

The reference ships no tests, fixtures or golden vectors for this path, so
each oracle function is pinned by an independent check:
* the slice-literal numpy reading of w_solve.m / w_grad.m / w_cpml.m /
  w_coeff.m / w_exp2pml.m (bit for bit in FP64);
* closed forms (coefficients, PML profile, wavelet);
* mathematics: linearity and mirror symmetry of the solver, the energy decay
  of a lossy medium, the dot-product test and central finite differences for
  the exact adjoint, numpy.gradient(edge_order=2) for differentiate_cube.m,
  scipy.signal.hilbert for instant_.m;
* the committed golden fixtures (tests/golden/), produced by this oracle and
  re-checked here so that a later change to it is caught.
"""
import os

import numpy as np
import pytest

from oracle import literal_np as L

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def tiny(O, m=8, n=12, P=3, air=2, nt=40, seed=0):
    rng = np.random.default_rng(seed)
    dx = 0.05
    g = O.Grid(m, n, P, air, dx, O.dt_cfl(dx), nt)
    eps = rng.uniform(4, 16, (m, n))
    sig = rng.uniform(0, 1e-2, (m, n))
    w = O.w_wavelet(np.arange(nt) * g.dt, 1e3, 1 / 300e6)
    s = (P + air + 2, P + 4)
    Mw = O.sub2ind_Mw(g.NX, g.NY, np.full(5, P + air + 2), P + 1 + np.arange(5) * 2)
    return g, eps, sig, w, s, Mw


def test_exp2pml_literal(O):
    g, eps, sig, *_ = tiny(O)
    a = O.exp2pml(g, eps, sig)
    b = L.w_exp2pml(eps, sig, g.P, g.air)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_coefficients_literal_and_closed_form(O):
    g, eps, sig, w, s, Mw = tiny(O)
    ep, sp = O.exp2pml(g, eps, sig)
    k = O.coeffs(g, ep, sp, s)
    k2 = L.w_cpml(ep, g.NX, g.NY, g.P, g.dx, g.dt, g.eps0, g.mu0, g.R0, g.order)
    k2.update(L.w_coeff(ep, sp, s, g.dx, g.dt, g.eps0, g.mu0))
    for name in O.COEFF_NAMES:
        assert np.array_equal(k[name], k2[name]), name
    # closed forms: lossless cells give Ceze = 1; Cezhy = dt / (eps dx) there
    D = 2 * ep * g.eps0 + g.dt * sp
    np.testing.assert_allclose(k["Cezhy"], 2 * g.dt / g.dx / D, rtol=1e-15)
    np.testing.assert_allclose(k["Ceze"][sp == 0], 1.0, rtol=0, atol=0)
    assert k["Chxh"] == 1.0 and k["Chyh"] == 1.0
    np.testing.assert_allclose(k["Chyez"], g.dt / (g.mu0 * g.dx), rtol=1e-15)
    # PML profile: sigma_max from R0 and a quadratic grading, outermost cell largest
    c = 1 / np.sqrt(g.mu0 * g.eps0)
    smax = -3 * g.eps0 * c * np.log(g.R0) / (2 * g.dx * g.P)
    se = smax * ((np.arange(g.P, 0, -1) - 0.75) / g.P) ** 2
    np.testing.assert_allclose(k["Cezxe_xn"][:, 0], (2 * g.eps0 - g.dt * se) / (2 * g.eps0 + g.dt * se),
                               rtol=1e-14)
    assert np.all(np.diff(k["Cezxe_xn"][:, 0]) > 0)  # decay weakest next to the interior


def test_forward_literal_bitwise(O):
    g, eps, sig, w, s, Mw = tiny(O)
    ep, sp = O.exp2pml(g, eps, sig)
    d, jz, u = O.forward(g, ep, sp, s, w, Mw, store_u=True)
    d2, jz2, u2 = L.w_solve(ep, sp, s, w, Mw, g.P, g.air, g.dx, g.dt, g.eps0, g.mu0)
    assert np.abs(d).max() > 0
    assert np.array_equal(d, d2) and np.array_equal(u, u2) and np.array_equal(jz, jz2)
    # the per-cell form used by the adjoint and by the GPU kernels
    assert np.array_equal(O.forward(g, ep, sp, s, w, Mw, percell=True), d)


def test_grad_ref_literal_bitwise(O):
    g, eps, sig, w, s, Mw = tiny(O)
    ep, sp = O.exp2pml(g, eps, sig)
    _, _, u = O.forward(g, ep, sp, s, w, Mw, store_u=True)
    e = np.random.default_rng(1).standard_normal((g.nt, len(Mw)))
    du = O.differentiate_cube(u, g.dt)
    a = O.grad_ref(g, ep, sp, s, Mw, e, du)
    b = L.w_grad(du, e, ep, sp, s, Mw, g.P, g.air, g.dx, g.dt, g.eps0, g.mu0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_forward_linear_and_mirror(O):
    g, eps, sig, w, s, Mw = tiny(O, m=10, n=15, P=3, air=0, nt=60)
    ep, sp = O.exp2pml(g, eps, sig)
    d1 = O.forward(g, ep, sp, s, w, Mw)[0]
    d2 = O.forward(g, ep, sp, s, 2.5 * w, Mw)[0]
    np.testing.assert_allclose(d2, 2.5 * d1, rtol=1e-12, atol=1e-12 * np.abs(d1).max())
    # homogeneous medium: mirroring source and receivers across the grid's
    # vertical axis (Ez column j <-> NY+2-j) mirrors the traces; this checks
    # the left and right strips against each other
    ep, sp = O.exp2pml(g, np.full_like(eps, 7.0), np.full_like(sig, 4e-3))
    d1 = O.forward(g, ep, sp, s, w, Mw)[0]
    sm = (s[0], g.NY + 2 - s[1])
    rz = (np.asarray(Mw) - 1) % (g.NX + 1) + 1
    rx = (np.asarray(Mw) - 1) // (g.NX + 1) + 1
    dm = O.forward(g, ep, sp, sm, w, O.sub2ind_Mw(g.NX, g.NY, rz, g.NY + 2 - rx))[0]
    np.testing.assert_allclose(dm, d1, rtol=1e-9, atol=1e-9 * np.abs(d1).max())


def test_lossy_decay(O):
    """With sigma > 0 the late-time field energy is below the lossless one."""
    g, eps, sig, w, s, Mw = tiny(O, m=20, n=20, P=4, air=0, nt=200)
    ep, sp0 = O.exp2pml(g, eps, 0 * sig)
    _, _, u0 = O.forward(g, ep, sp0, s, w, Mw, store_u=True)
    ep, sp1 = O.exp2pml(g, eps, 0 * sig + 0.05)
    _, _, u1 = O.forward(g, ep, sp1, s, w, Mw, store_u=True)
    assert (u1[..., -50:] ** 2).sum() < 0.5 * (u0[..., -50:] ** 2).sum()


def test_differentiate_cube_numpy_gradient(O):
    rng = np.random.default_rng(4)
    for nt in (3, 4, 9, 10):
        u = rng.standard_normal((3, 5, nt))
        np.testing.assert_allclose(O.differentiate_cube(u, 0.3),
                                   np.gradient(u, 0.3, axis=2, edge_order=2), rtol=1e-13, atol=1e-13)


def exact_setup(O, air=2):
    rng = np.random.default_rng(11)
    m, n, P, nt = 9, 13, 4, 110
    dx = 0.05
    g = O.Grid(m, n, P, air, dx, O.dt_cfl(dx), nt)
    eps = rng.uniform(4, 16, (m, n))
    sig = rng.uniform(0, 3e-2, (m, n))
    w = O.w_wavelet(np.arange(nt) * g.dt, 1e3, 1 / 300e6)
    s = (P + air + 3, P + 5)
    rz = np.array([P + air + 2, P + air + 2, P + air + 5, 2, P + air + 1, g.NX])
    rx = np.array([P + 2, P + 9, P + 12, P + 6, P + 1, 5])
    Mw = O.sub2ind_Mw(g.NX, g.NY, rz, rx)
    return g, eps, sig, w, s, Mw


def test_exact_adjoint_dot_product(O):
    g, eps, sig, w, s, Mw = exact_setup(O)
    ep, sp = O.exp2pml(g, eps, sig)
    d = O.forward(g, ep, sp, s, w, Mw)[0]
    e = np.random.default_rng(2).standard_normal(d.shape)
    for K in (1, 7, g.nt):
        _, _, gw = O.grad_exact(g, ep, sp, s, w, Mw, e, K=K)
        lhs, rhs = (d * e).sum(), (w * gw).sum()
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (K, lhs, rhs)


@pytest.mark.parametrize("air", [0, 2])
def test_exact_gradient_finite_differences(O, air):
    g, eps, sig, w, s, Mw = exact_setup(O, air)
    ep, sp = O.exp2pml(g, eps, sig)
    d = O.forward(g, ep, sp, s, w, Mw)[0]
    e = np.random.default_rng(3).standard_normal(d.shape)
    ge, gs, _ = O.grad_exact(g, ep, sp, s, w, Mw, e, K=9)
    gem, gsm = O.fold_pad(g, ge, gs)

    def J(E, Sg):
        a, b = O.exp2pml(g, E, Sg)
        return float((O.forward(g, a, b, s, w, Mw)[0] * e).sum())

    cells = [(0, 0), (0, 5), (g.m - 1, g.n - 1), (3, 0), (5, 7), (g.m - 1, 3), (2, g.n - 1),
             (s[0] - g.P - air - 1, s[1] - g.P - 1)]
    for a, b in cells:
        h = 1e-6 * eps[a, b]
        Ep, Em = eps.copy(), eps.copy()
        Ep[a, b] += h
        Em[a, b] -= h
        fd = (J(Ep, sig) - J(Em, sig)) / (2 * h)
        assert abs(fd - gem[a, b]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gem).max()), (a, b, fd, gem[a, b])
        hs = 1e-6
        Sp, Sm = sig.copy(), sig.copy()
        Sp[a, b] += hs
        Sm[a, b] -= hs
        fds = (J(eps, Sp) - J(eps, Sm)) / (2 * hs)
        assert abs(fds - gsm[a, b]) <= 1e-5 * max(abs(fds), 1e-3 * np.abs(gsm).max()), (a, b, fds, gsm[a, b])


def test_exact_without_strip_terms_matches_interior_fd(O):
    """pml_terms=False (strip coefficients held fixed, what the GPU computes)
    equals the full gradient away from the model edges and the first column."""
    g, eps, sig, w, s, Mw = exact_setup(O, 2)
    ep, sp = O.exp2pml(g, eps, sig)
    e = np.random.default_rng(3).standard_normal((g.nt, len(Mw)))
    full = O.fold_pad(g, *O.grad_exact(g, ep, sp, s, w, Mw, e, K=9)[:2])
    part = O.fold_pad(g, *O.grad_exact(g, ep, sp, s, w, Mw, e, K=9, pml_terms=False)[:2])
    np.testing.assert_allclose(part[0][1:-1, 1:-1], full[0][1:-1, 1:-1], rtol=1e-12, atol=1e-12 * np.abs(full[0]).max())
    np.testing.assert_allclose(part[1], full[1], rtol=1e-12, atol=0)


def test_wavelet_closed_form(O):
    f0 = 100e6
    t = np.linspace(0, 60e-9, 601)
    w = O.w_wavelet(t, 1e3, 1 / f0)
    tau = 1 / (np.pi * f0)
    assert np.isclose(np.abs(w).max(), 1e3)
    # derivative-of-Gaussian: zero crossing at t0 = 5 tau, extrema at t0 +- tau/sqrt(2)
    k = np.argmax(w)
    assert abs(t[k] - (5 * tau - tau / np.sqrt(2))) <= (t[1] - t[0])
    assert abs(np.interp(5 * tau, t, w)) < 1e-4 * 1e3


def test_hilbert_matches_scipy(O):
    from scipy.signal import hilbert
    x = np.random.default_rng(0).standard_normal((33, 4))
    for N in (33, 32):
        np.testing.assert_allclose(O.hilbert_analytic(x[:N], axis=0), hilbert(x[:N], axis=0),
                                   rtol=1e-12, atol=1e-12)


def test_misfits(O):
    rng = np.random.default_rng(9)
    t = np.linspace(0, 1, 128)[:, None]
    d = np.sin(40 * t + rng.uniform(0, 1, (1, 3))) * np.exp(-((t - 0.5) / 0.2) ** 2)
    do = 0.9 * np.roll(d, 2, axis=0)
    e, _, _ = O.w_bozdag(do, d, "L2")
    np.testing.assert_array_equal(e, d - do)
    assert np.isclose(O.w_E(do, d), 0.5 * ((d - do) ** 2).sum())
    # env_L2 adjoint source = gradient of 0.5 ||A(d) - A(d_o)||^2 (Bozdag et al. 2011)
    def Jenv(dd):
        a = np.abs(O.hilbert_analytic(dd, axis=0))
        b = np.abs(O.hilbert_analytic(do, axis=0))
        return 0.5 * ((a - b) ** 2).sum()
    e, _, _ = O.w_bozdag(do, d, "env_L2")
    k = (60, 1)
    h = 1e-6
    dp, dm = d.copy(), d.copy()
    dp[k] += h
    dm[k] -= h
    fd = (Jenv(dp) - Jenv(dm)) / (2 * h)
    assert abs(fd - e[k]) <= 1e-5 * max(abs(fd), 1e-6)


def _envelope_scipy(x):
    from scipy.signal import hilbert
    return np.abs(hilbert(x, axis=0))


def _misfit_traces(seed, nt=96, nr=3):
    """Smooth wavelets on a DC floor, so that every envelope stays away from
    zero and the log and quotient misfits are differentiable at each sample."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 1, nt)[:, None]
    d = 0.6 + np.sin(30 * t + rng.uniform(0, 1, (1, nr))) * np.exp(-((t - 0.5) / 0.2) ** 2)
    do = 0.6 + 0.8 * np.sin(30 * t - 0.3 + rng.uniform(0, 1, (1, nr))) * np.exp(-((t - 0.45) / 0.25) ** 2)
    return d, do


def _central_fd(J, d, cells, h=1e-6):
    out = []
    for k in cells:
        dp, dm = d.copy(), d.copy()
        dp[k] += h
        dm[k] -= h
        out.append((J(dp) - J(dm)) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("nt", [96, 97])
def test_env_log_adjoint_is_gradient(O, nt):
    """w_bozdag.m:47-54 (env_LOG): the adjoint source equals the gradient of
    J = 1/2 ||log A(d) - log A(d_o)||^2, A = |analytic signal| (w_E.m's
    env_LOG objective on the envelopes), with A from scipy.signal.hilbert.
    A flipped sign, a missing 1/A^2, or d and H[d] swapped fail it."""
    d, do = _misfit_traces(11, nt)
    logAo = np.log(_envelope_scipy(do))

    def J(x):
        return 0.5 * ((np.log(_envelope_scipy(x)) - logAo) ** 2).sum()

    e, de, doe = O.w_bozdag(do, d, "env_LOG")
    np.testing.assert_allclose(de, _envelope_scipy(d), rtol=1e-12)
    cells = [(0, 0), (5, 1), (nt // 2, 2), (nt // 3, 0), (nt - 1, 1)]
    fd = _central_fd(J, d, cells)
    np.testing.assert_allclose(np.array([e[k] for k in cells]), fd, rtol=2e-6,
                               atol=2e-7 * np.abs(e).max())


@pytest.mark.parametrize("nt", [96, 97])
def test_env_l22_adjoint_is_half_gradient(O, nt):
    """w_bozdag.m:38-46 (env_L22): the adjoint source is one half of the
    gradient of J = 1/2 ||A(d)^2 - A(d_o)^2||^2 (d A^2 = 2 (d dd + H[d] H[dd]);
    the reference keeps no factor 2).  A wrong power of the envelope, a wrong
    sign of the Hilbert term, or the factor 2 kept fail it."""
    d, do = _misfit_traces(12, nt)
    Ao2 = _envelope_scipy(do) ** 2

    def J(x):
        return 0.5 * ((_envelope_scipy(x) ** 2 - Ao2) ** 2).sum()

    e, _, _ = O.w_bozdag(do, d, "env_L22")
    cells = [(0, 0), (7, 1), (nt // 2, 2), (nt // 4, 1), (nt - 1, 0)]
    fd = _central_fd(J, d, cells)
    np.testing.assert_allclose(2 * np.array([e[k] for k in cells]), fd, rtol=2e-6,
                               atol=2e-7 * np.abs(e).max())
    # env_LOG and env_L2 objectives through w_E.m's own transforms
    A, Ao = _envelope_scipy(d), _envelope_scipy(do)
    assert np.isclose(O.w_E(Ao, A, obj_fnc="env_L22"), 0.5 * ((A**2 - Ao**2) ** 2).sum(), rtol=1e-12)
    assert np.isclose(O.w_E(Ao, A, obj_fnc="env_LOG"), 0.5 * ((np.log(A) - np.log(Ao)) ** 2).sum(),
                      rtol=1e-12)


def _hilbert_cot(x):
    """Discrete Hilbert transform of an even-length periodic sequence as the
    closed-form circular convolution y_n = sum_m h_(n-m) x_m with
    h_k = (2/N) cot(pi k / N) for odd k and 0 for even k: the inverse DFT of
    the multiplier -i sgn(f) (0 at f = 0 and f = N/2), summed in closed form.
    No FFT and no bin ordering are involved."""
    N = x.shape[0]
    assert N % 2 == 0
    k = np.arange(N)
    h = np.zeros(N)
    odd = k % 2 == 1
    h[odd] = (2.0 / N) / np.tan(np.pi * k[odd] / N)
    idx = (k[:, None] - k[None, :]) % N
    return h[idx] @ x


def _wrap(p):
    """angle()'s branch (-pi, pi]."""
    return np.pi - np.mod(np.pi - p, 2 * np.pi)


@pytest.mark.parametrize("case", ["tone", "tone_on_dc"])
def test_phase_l2_closed_form(O, case):
    """w_bozdag.m:29-37 (phase_L2) on signals whose analytic signal is known
    exactly, so that phases and envelopes are closed forms (no Hilbert
    transform) and the one Hilbert transform left is the cot-kernel
    convolution above:
    * tone: d = a cos(w t + q), d_o = b cos(w t + q - delta), whole periods:
      A = a, phi = wrap(w t + q), phi_o = wrap(w t + q - delta), so the
      phase difference is delta or delta - 2 pi on the samples where only
      phi wrapped (angle() is not unwrapped in the reference);
    * tone on a DC floor: d = c + cos(w t), analytic signal c + e^(i w t),
      phi = atan2(sin w t, c + cos w t), A^2 = c^2 + 1 + 2 c cos w t, and the
      same for d_o with (c_o, shift delta): no wrap, a varying envelope.
    Then e = -psi + H[psi d] with psi = (phi - phi_o) / A^2.  A sign slip in
    the phase difference or in the combination, A in place of A^2, H[d] in
    place of d inside the transform, or the opposite Hilbert convention all
    fail it (checked below on the wrong variants)."""
    N, k0 = 64, 5
    n = np.arange(N)[:, None]
    wt = 2 * np.pi * k0 * n / N
    if case == "tone":
        a, b, q, delta = 1.7, 0.6, 0.123, 0.9
        d = a * np.cos(wt + q)
        do = b * np.cos(wt + q - delta)
        phi, phio, A2 = _wrap(wt + q), _wrap(wt + q - delta), np.full_like(wt, a * a)
    else:
        c, co, delta = 2.5, 1.8, 0.7
        d = c + np.cos(wt)
        do = co + np.cos(wt - delta)
        phi = np.arctan2(np.sin(wt), c + np.cos(wt))
        phio = np.arctan2(np.sin(wt - delta), co + np.cos(wt - delta))
        A2 = c * c + 1 + 2 * c * np.cos(wt)
    e, ph, pho = O.w_bozdag(do, d, "phase_L2")
    np.testing.assert_allclose(ph, phi, atol=1e-12)
    np.testing.assert_allclose(pho, phio, atol=1e-12)
    psi = (phi - phio) / A2
    want = -psi + _hilbert_cot(psi * d)
    np.testing.assert_allclose(e, want, rtol=0, atol=1e-11 * np.abs(want).max())
    if case == "tone":
        # the phase difference takes both values delta and delta - 2 pi
        dphi = phi - phio
        assert np.any(np.isclose(dphi, delta)) and np.any(np.isclose(dphi, delta - 2 * np.pi))
    psiA = (phi - phio) / np.sqrt(A2)
    wrong = [psi - _hilbert_cot(psi * d),        # phase difference flipped
             -psi - _hilbert_cot(psi * d),       # opposite Hilbert convention / combination sign
             -psiA + _hilbert_cot(psiA * d),     # A in place of A^2
             -psi + _hilbert_cot(psi * _hilbert_cot(d))]  # H[d] in place of d
    for w_ in wrong:
        assert np.abs(e - w_).max() > 1e-3 * np.abs(want).max()


def test_hilbert_cot_kernel_is_the_hilbert_of_the_oracle(O):
    """The cot-kernel convolution equals imag(hilbert(x)) of the oracle and
    maps cos to sin (k0 != 0, N/2), so it pins the convention by itself."""
    N = 64
    x = np.random.default_rng(3).standard_normal((N, 2))
    np.testing.assert_allclose(_hilbert_cot(x), np.imag(O.hilbert_analytic(x, axis=0)),
                               atol=1e-12)
    t = 2 * np.pi * 3 * np.arange(N) / N
    np.testing.assert_allclose(_hilbert_cot(np.cos(t)), np.sin(t), atol=1e-12)


def test_golden_fixtures(O):
    path = os.path.join(HERE, "golden", "tiny_forward_grad.npz")
    z = np.load(path, allow_pickle=False)
    g = O.Grid(int(z["m"]), int(z["n"]), int(z["P"]), int(z["air"]), float(z["dx"]), float(z["dt"]),
               int(z["nt"]))
    ep, sp = O.exp2pml(g, z["eps"], z["sig"])
    d = O.forward(g, ep, sp, tuple(z["src"]), z["w"], z["Mw"])[0]
    np.testing.assert_allclose(d, z["d"], rtol=1e-12, atol=1e-12 * np.abs(z["d"]).max())
    ge, gs, gw = O.grad_exact(g, ep, sp, tuple(z["src"]), z["w"], z["Mw"], z["e"], K=5)
    np.testing.assert_allclose(ge, z["g_eps_pad"], rtol=1e-10, atol=1e-12 * np.abs(z["g_eps_pad"]).max())
    np.testing.assert_allclose(gs, z["g_sig_pad"], rtol=1e-10, atol=1e-12 * np.abs(z["g_sig_pad"]).max())


# --------------------------------------------------------------------------
# 16-bit checkpoint snapshots
# --------------------------------------------------------------------------
def test_snapshot_rounding_matches_library_conversions(O):
    """The snapshot rounding equals numpy's float16 and torch's bfloat16
    round-to-nearest-even conversions of the power-of-two-scaled values, and
    the scale puts the plane's largest value just below 2^15 (FP16) or in
    [1, 2) (BF16)."""
    import torch
    rng = np.random.default_rng(11)
    for scale in (1e-7, 3.0, 2.5e4):
        v = (rng.standard_normal(4000) * scale).astype(np.float32).astype(np.float64)
        v[:8] = [0.0, -0.0, 5e-12 * scale, -3e-9 * scale, 1e-30, v[8], -v[9], v[10]]
        a = np.abs(v).max()
        r16, sh16 = O.snap_round(v, 2)
        assert 2.0 ** 14 <= a * 2.0 ** sh16 < 2.0 ** 15
        ref16 = np.ldexp(np.ldexp(v, sh16).astype(np.float32).astype(np.float16).astype(np.float64), -sh16)
        np.testing.assert_array_equal(r16, ref16)
        rb, shb = O.snap_round(v, 1)
        assert 1.0 <= a * 2.0 ** shb < 2.0
        xb = torch.from_numpy(np.ldexp(v, shb).astype(np.float32)).to(torch.bfloat16).to(torch.float64)
        np.testing.assert_array_equal(rb, np.ldexp(xb.numpy(), -shb))
    # ties go to even: 1 + 2^-8 (BF16) and 1 + 2^-11 (FP16) round down, 1 + 3*2^-8 up
    t = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.5])
    rb, _ = O.snap_round(t, 1, amax=1.5)
    np.testing.assert_array_equal(rb, [1.0, 1.0 + 4 * 2.0 ** -8, 1.5])
    t = np.array([1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11])
    r16, sh = O.snap_round(t, 2, amax=1.5)
    assert sh == 14
    np.testing.assert_array_equal(r16, [1.0, 1.0 + 4 * 2.0 ** -11])


def test_oracle_snapshot_gradient(O):
    """Checkpoints stored as FP16 / BF16 snapshots perturb the exact gradient
    by about the formats' rounding (2^-12, 2^-9 relative), and the rounding
    does reach the recomputed segments (the gradient changes)."""
    g, eps, sig, w, s, Mw = exact_setup(O)
    ep, sp = O.exp2pml(g, eps, sig)
    e = np.random.default_rng(5).standard_normal((g.nt, Mw.size))
    ref = O.grad_exact(g, ep, sp, s, w, Mw, e, K=13, pml_terms=False)
    for mode, tol in ((2, 1e-3), (1, 1e-2)):
        got = O.grad_exact(g, ep, sp, s, w, Mw, e, K=13, pml_terms=False, snap=mode)
        for a, b in zip(got[:2], ref[:2]):
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            print("snap", mode, "rel err", err)
            assert 0 < err <= tol, (mode, err)


def test_snapshot_error_follows_the_adjoint_source_band(O):
    """Why 16-bit snapshots cost far more gradient accuracy on a small case
    with a random adjoint source than on a survey with its real residual:
    snapshot rounding is grid-scale noise in the restored state, which the
    recompute carries into every step of the segment; a white-noise adjoint
    source (energy up to the Nyquist frequency) correlates with it, while a
    band-limited one (the residual of a band-limited wavelet, or white noise
    through filt_gauss) does not.  On the same model, interval and snapshot
    format, the band-limited sources give an error more than 5x smaller."""
    from gerjoii_amd import configs
    case = configs.small(m=60, n=70, P=10, nt=300, ns=1, seed=7)
    g0 = case.grid
    g = O.Grid(g0.m, g0.n, g0.pml, g0.air, g0.dx, g0.dt, g0.nt)
    ep, sp = O.exp2pml(g, case.eps0, case.sig0)
    et, st = O.exp2pml(g, case.eps_true, case.sig_true)
    src, w, Mw = case.src[0], case.wavelets[0], case.receivers[0]
    d = O.forward(g, ep, sp, src, w, Mw)[0]
    do = O.forward(g, et, st, src, w, Mw)[0]
    white = np.random.default_rng(3).standard_normal(d.shape) * np.abs(d - do).max()
    errs = {}
    for name, e in (("white", white), ("band", O.filt_gauss(white, g.dt, 5e7, 6e8, 8)),
                    ("residual", d - do)):
        ref = O.grad_exact(g, ep, sp, src, w, Mw, e, K=19, pml_terms=False)[0]
        got = O.grad_exact(g, ep, sp, src, w, Mw, e, K=19, pml_terms=False, snap=1)[0]
        errs[name] = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("BF16 snapshot eps-gradient error by adjoint source:", errs)
    assert errs["band"] * 5 < errs["white"] and errs["residual"] * 5 < errs["white"]


# --------------------------------------------------------------------------
# adjoint-source conditioning (filt_gauss.m, tukeywin, w_lmute.m)
# --------------------------------------------------------------------------

def test_tukeywin_matches_scipy(O):
    """MATLAB's tukeywin and scipy's tukey window are the same definition
    (flank length floor(r/2 (n-1)) + 1, raised cosine): equal for odd and
    even n and several taper fractions, including the Hann limit r = 1."""
    from scipy.signal import windows
    for n in (1, 2, 3, 5, 8, 11, 240, 241, 1000):
        for r in (0.1, 0.3, 0.5, 1.0):
            np.testing.assert_allclose(O.tukeywin(n, r), windows.tukey(n, r), rtol=0, atol=1e-13)


def _filt_gauss_dft(d, dt, f_low, f_high, steep):
    """The same filter by explicit sums over the signed frequencies
    f_j = (j - N/2) df, without FFT ordering or fftshift."""
    nt, nr = d.shape
    N = 1
    while N < nt:
        N *= 2
    x = np.zeros((N, nr))
    from scipy.signal import windows
    x[:nt] = d * windows.tukey(nt, 0.1)[:, None]
    df = 1.0 / dt / N
    f = (np.arange(N) - N // 2) * df
    width = f_high - f_low
    mid = f_low + width / 2
    g = np.exp(-((f - mid) ** steep) / ((df / 2) * width ** (steep - 1)))
    n = np.arange(N)
    E = np.exp(-2j * np.pi * np.outer(f * dt, n))  # (freq, time)
    X = E @ x
    y = (np.conj(E).T @ (g[:, None] * X)) / N
    return np.real(y[:nt])


def test_filt_gauss_closed_forms(O):
    """filt_gauss.m keeps only the positive-frequency band, so a tone at the
    band centre comes back at half amplitude (away from the taper), a tone
    far outside the band vanishes, the filter is linear, and it equals the
    filter written as explicit sums over the signed frequencies (odd, even
    and power-of-two lengths)."""
    dt, f_low, f_high = 1e-10, 2e8, 6e8
    nt = 2000
    t = np.arange(nt) * dt
    tone = np.cos(2 * np.pi * 4e8 * t)[:, None]
    y = O.filt_gauss(tone, dt, f_low, f_high, 8)
    mid = slice(300, 1700)
    assert np.abs(y[mid, 0] - 0.5 * tone[mid, 0]).max() < 1e-5
    far = np.cos(2 * np.pi * 2e9 * t)[:, None]
    assert np.abs(O.filt_gauss(far, dt, f_low, f_high, 8)).max() < 1e-4
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal((2, 300, 3))
    np.testing.assert_allclose(O.filt_gauss(2 * a - 3 * b, dt, f_low, f_high, 8),
                               2 * O.filt_gauss(a, dt, f_low, f_high, 8)
                               - 3 * O.filt_gauss(b, dt, f_low, f_high, 8), atol=1e-12)
    for nt in (37, 64, 100):
        d = rng.standard_normal((nt, 2))
        np.testing.assert_allclose(O.filt_gauss(d, dt, f_low, f_high, 8),
                                   _filt_gauss_dft(d, dt, f_low, f_high, 8), rtol=0, atol=1e-10)


def test_lmute_closed_forms(O):
    """w_lmute.m: the mute is 1/2 on the arrival t0 + (x - x0)/v, close to 1
    well after it, and e = d - d_o after muting both."""
    nt, nr = 50, 4
    t = np.arange(nt) * 1e-9  # [s]
    rx = np.array([2.0, 2.5, 3.0, 4.0])  # [m]
    v_mute, t_mute = 0.1, 5.0  # [m/ns], [ns]
    d = np.ones((nt, nr))
    d_o = np.full((nt, nr), 0.25)
    dm, dom, e = O.lmute(d, d_o, t, rx, v_mute, t_mute)
    to_ = t_mute * 1e-9 + (rx - rx[0]) / (v_mute * 1e9)
    for r in range(nr):
        m = 1.0 / (1.0 + np.exp(5.0 * (to_[r] - t)))
        np.testing.assert_allclose(dm[:, r], m, rtol=1e-15)
    # the arrival itself, evaluated on a time axis that holds it
    dm2, _, _ = O.lmute(np.ones((1, 1)), np.zeros((1, 1)), np.array([t_mute * 1e-9]), rx[:1],
                        v_mute, t_mute)
    assert dm2[0, 0] == 0.5
    np.testing.assert_allclose(e, dm - dom, rtol=0, atol=0)
    np.testing.assert_allclose(dom, 0.25 * dm, rtol=1e-15)


# --------------------------------------------------------------------------
# the physics of the forward stencil (analytic pins)
# --------------------------------------------------------------------------
def _line_source_field(rho, t, eps_r, amp, f0, dt, dx, mu0=1.2566370614359173e-06,
                       eps0=8.854187817e-12):
    """Ez(rho, t) of a z-directed line current I(t) in a homogeneous lossless
    medium: Ez = -(mu0 / 2 pi) d/dt int G, the 2-D Green's function
    H(t - rho/c) / (2 pi sqrt(t^2 - rho^2/c^2)); with tau' = t - (rho/c) cosh u
    the integral has no singular point:
        Ez(rho, t) = -(mu0 / 2 pi) int_0^inf I'(t - (rho/c) cosh u) du.
    The oracle's soft source adds Cezj w = -(dt/eps) w to one Ez cell per step,
    the discrete Ampere law with current density w over a dx^2 cell, and it
    acts between E^n and E^(n+1): I(t) = dx^2 w(t - dt/2), w the wavelet of
    w_wavelet.m (its max-abs scale taken from the sampled wavelet)."""
    from gerjoii_amd.survey import w_wavelet
    c = 1.0 / np.sqrt(mu0 * eps0 * eps_r)
    tau = 1.0 / (np.pi * f0)
    t0 = 5 * tau
    T = np.arange(len(t)) * dt
    raw = -(T - t0) * np.exp(-((T - t0) ** 2) / tau**2)
    scale = amp / np.max(np.abs(raw))
    assert np.allclose(scale * raw, w_wavelet(T, amp, 1.0 / f0))

    def dI(tt):  # d/dt of dx^2 w(tt - dt/2)
        s = tt - dt / 2 - t0
        return dx * dx * scale * np.exp(-(s**2) / tau**2) * (2 * s**2 / tau**2 - 1)

    out = np.zeros(len(t))
    a = rho / c
    for k, tk in enumerate(t):
        if tk <= a:
            continue
        U = np.arccosh(tk / a) + 1e-12  # beyond U the argument is negative: w is 0 there
        u = np.linspace(0.0, U, 6001)
        out[k] = -(mu0 / (2 * np.pi)) * np.trapezoid(dI(tk - a * np.cosh(u)), u)
    return out


def test_forward_matches_the_line_source_greens_function(O):
    """A receiver 40 cells from the source in a homogeneous lossless medium
    (eps_r 4, 100 MHz, at least 30 cells per wavelength where the wavelet has
    energy), before anything returns from the PML: the oracle's trace equals the
    analytic 2-D line-source field to 0.5% (measured 1.5e-3: numerical
    dispersion and the one-cell source).  The pin is tight enough to see the
    half-step timing of the source (the field one step early or late misses
    by 2.6-2.8%) and a 5% error in the wave speed's eps_r (8.5%); a slipped
    sign or dx^2 of the source fails outright."""
    from gerjoii_amd import survey as Sv
    m = n = 240
    P, dx, f0, eps_r = 20, 0.02, 100e6, 4.0
    dt = Sv.dt_cfl(dx)
    nt = 800
    g = O.Grid(m, n, P, 0, dx, dt, nt)
    sg = Sv.Grid(m=m, n=n, pml=P, air=0, nt=nt, dx=dx, dt=dt)
    src = Sv.src_index(sg, 120, 120)
    Mw = np.array([Sv.rec_Mw(sg, 120, 160)])
    ep, sp = O.exp2pml(g, np.full((m, n), eps_r), np.zeros((m, n)))
    w = Sv.w_wavelet(np.arange(nt) * dt, 1e3, 1.0 / f0)
    d = O.forward(g, ep, sp, src, w, Mw)[0][:, 0]
    t = (np.arange(nt) + 1) * dt           # sample k is Ez after step k+1
    # the first PML return: 80 cells to the nearest PML edge and back, minus the direct 40
    c = 1.0 / np.sqrt(g.mu0 * g.eps0 * eps_r)
    win = t < (40 + 2 * 80) * dx / c
    ref = _line_source_field(40 * dx, t, eps_r, 1e3, f0, dt, dx)
    err = np.linalg.norm(d[win] - ref[win]) / np.linalg.norm(ref[win])
    assert np.abs(ref[win]).max() > 0 and err <= 5e-3, err
    for e_wrong in (3.8, 4.2):
        wrong = _line_source_field(40 * dx, t, e_wrong, 1e3, f0, dt, dx)
        assert np.linalg.norm(d[win] - wrong[win]) / np.linalg.norm(wrong[win]) > 0.05
    for shift in (-1, 1):
        late = np.roll(ref, shift)
        assert np.linalg.norm(d[win] - late[win]) / np.linalg.norm(ref[win]) > 0.015
    assert np.linalg.norm(d[win] + ref[win]) / np.linalg.norm(ref[win]) > 1.5


def test_pml_absorbs(O):
    """The split-field PML (w_cpml.m, R0 = 1e-8, quadratic grading) against an
    unbounded medium: a source 10 cells from the left PML of a narrow grid and
    a receiver 30 cells to its right record the same trace as in a grid wide
    enough that no side PML echo has arrived, up to the wave the side PMLs
    return: 4.4e-5 of the direct arrival's peak with a 20-cell PML (bar 2e-4),
    falling with thickness (measured 6.6e-2, 1.5e-3 and 1.7e-4 for 2, 5 and 10
    cells)."""
    from gerjoii_amd import survey as Sv
    dx, f0, eps_r = 0.02, 100e6, 4.0
    dt = Sv.dt_cfl(dx)
    nt = 700

    def trace(m, n, P, sz, sx, rz, rx):
        g = O.Grid(m, n, P, 0, dx, dt, nt)
        sg = Sv.Grid(m=m, n=n, pml=P, air=0, nt=nt, dx=dx, dt=dt)
        ep, sp = O.exp2pml(g, np.full((m, n), eps_r), np.zeros((m, n)))
        w = Sv.w_wavelet(np.arange(nt) * dt, 1e3, 1.0 / f0)
        return O.forward(g, ep, sp, Sv.src_index(sg, sz, sx), w,
                         np.array([Sv.rec_Mw(sg, rz, rx)]))[0][:, 0]

    # source 10 cells from the left PML, receiver 30 cells to its right, both
    # 200 rows deep in grids of the same depth (the top strip is vacuum: the
    # free surface above the model reflects, alike in both grids)
    ref = trace(400, 400, 20, 200, 185, 200, 215)          # side PMLs 180+ cells away
    small20 = trace(400, 60, 20, 200, 10, 200, 40)
    small2 = trace(400, 60, 2, 200, 10, 200, 40)
    c = 1.0 / np.sqrt(O.MU0 * O.EPS0 * eps_r)
    t = (np.arange(nt) + 1) * dt
    win = t < 330 * dx / c                                 # before ref's own PML echo
    peak = np.abs(ref[win]).max()
    r20 = np.abs(small20[win] - ref[win]).max() / peak
    r2 = np.abs(small2[win] - ref[win]).max() / peak
    assert peak > 0 and r20 < 2e-4 and r2 > 100 * r20, (r20, r2)
