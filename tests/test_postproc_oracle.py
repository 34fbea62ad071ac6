"""CPU tests that pin the oracle's gradient post-processing (gaussi.m,
image_gaussian.m, normali.m, the source reduction of w_update_e_.m and the
update of w_image_e.m), and the host index logic of gerjoii_amd.postproc.

Each check is independent of the oracle's own formulation: closed forms,
eigenfunctions of the k-space filter, a separable (outer-product) form of the
taper, and invariants a transposed axis or a shifted index would break.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def test_normali_closed_forms(O):
    a = np.array([[1.0, -4.0], [2.0, 0.5]])
    np.testing.assert_array_equal(O.normali(a), a / 4.0)
    z = np.zeros((3, 2))
    np.testing.assert_array_equal(O.normali(z), z)
    # MATLAB's max skips NaN: the other entries are scaled by the finite max
    b = np.array([np.nan, 2.0, -8.0])
    out = O.normali(b)
    assert np.isnan(out[0]) and out[1] == 0.25 and out[2] == -1.0


def test_gaussi_separable_form(O):
    """1 - exp(-(a_j + b_i)) = 1 - outer(exp(-b), exp(-a)): the taper built as
    an outer product on a non-square grid with an off-centre source; it is 0
    on the source cell exactly and 1 at its maximum."""
    x = np.linspace(0.0, 3.0, 31)      # n = 31 columns
    z = np.linspace(0.0, 1.2, 13)      # m = 13 rows
    s = (4, 22)                        # (iz, ix), 1-based
    lo = 0.35
    g = O.gaussi(x, z, s, lo)
    assert g.shape == (13, 31)
    L = 2 * lo
    ex = np.exp(-((x - x[s[1] - 1]) / L) ** 2)
    ez = np.exp(-((z - z[s[0] - 1]) / L) ** 2)
    t = 1 - np.outer(ez, ex)
    np.testing.assert_allclose(g, t / t.max(), rtol=1e-13, atol=1e-15)
    assert g[s[0] - 1, s[1] - 1] == 0.0
    assert np.isclose(g.max(), 1.0)
    assert np.unravel_index(np.argmin(g), g.shape) == (s[0] - 1, s[1] - 1)


def _cake(k, N, a):
    """Filter gain at FFT bin k of an N-point axis: the cake exp(-r^2) centred
    on N/2 of the unshifted grid lands, after fftshift, on the signed
    frequency of bin k."""
    sk = k if k < N // 2 else k - N
    return np.exp(-(sk / a) ** 2)


@pytest.mark.parametrize("kx,kz", [(3, 0), (0, 5), (7, 2)])
def test_image_gaussian_eigenfunctions(O, kx, kz):
    """On power-of-two sizes (no padding) a cosine mode is an eigenfunction:
    LOW_PASS scales cos(2 pi (kx j / n + kz i / m)) by exp(-(kx/a)^2 -
    (kz/b)^2), a = n ax/2 sqrt2 along the columns and b = m az/2 sqrt2 along
    depth; a swapped axis or a shift of the cake by one bin changes the gain."""
    m, n = 32, 64
    ax, az = 0.09, 0.21
    i = np.arange(m)[:, None]
    j = np.arange(n)[None, :]
    f = np.cos(2 * np.pi * (kx * j / n + kz * i / m))
    a = n * (ax / 2) * np.sqrt(2)
    b = m * (az / 2) * np.sqrt(2)
    gain = _cake(kx, n, a) * _cake(kz, m, b)
    np.testing.assert_allclose(O.image_gaussian(f, ax, az, "LOW_PASS"), gain * f, atol=1e-12)
    # HI_PASS: 1 - exp(-r^2), divided by its maximum at the grid corner
    hmax = 1 - np.exp(-((n / 2) / a) ** 2 - ((m / 2) / b) ** 2)
    np.testing.assert_allclose(O.image_gaussian(f, ax, az, "HI_PASS"),
                               (1 - gain) / hmax * f, atol=1e-12)


def test_image_gaussian_dc_linearity_and_realness(O):
    """A constant passes LOW_PASS unchanged and HI_PASS removes it (on
    power-of-two sizes); the filter is linear; and the cake is even in the
    signed frequencies, so ifft2 of the filtered spectrum is real -- the
    property that lets the device use a real-to-real (D2Z/Z2D) transform."""
    c = np.full((16, 8), 2.5)
    np.testing.assert_allclose(O.image_gaussian(c, 0.2, 0.3, "LOW_PASS"), c, atol=1e-13)
    np.testing.assert_allclose(O.image_gaussian(c, 0.2, 0.3, "HI_PASS"), 0 * c, atol=1e-13)
    rng = np.random.default_rng(3)
    p, q = rng.standard_normal((2, 37, 53))
    np.testing.assert_allclose(O.image_gaussian(2 * p - q, 0.1, 0.15),
                               2 * O.image_gaussian(p, 0.1, 0.15) - O.image_gaussian(q, 0.1, 0.15),
                               atol=1e-12)
    for (m_, n_) in ((64, 64), (32, 128)):
        a = n_ * (0.1 / 2) * np.sqrt(2)
        b = m_ * (0.15 / 2) * np.sqrt(2)
        XY, YX = np.meshgrid(np.arange(n_), np.arange(m_))
        H = np.fft.fftshift(np.exp(-((XY - n_ / 2) / a) ** 2 - ((YX - m_ / 2) / b) ** 2))
        K = np.fft.fft2(rng.standard_normal((m_, n_)))
        z = np.fft.ifft2(H * K)
        assert np.abs(z.imag).max() <= 1e-13 * np.abs(z.real).max()


def test_w_update_post_and_stack(O):
    """The per-source chain zeroes a gradient holding a NaN (w_update_e.m:
    133-137); the reduction is -sum(step g)/ns and vanishes when the steps
    sum below zero; the update is eps .* exp(eps .* d)."""
    x = np.linspace(0, 1, 9)
    z = np.linspace(0, 0.5, 6)
    g = np.ones((6, 9))
    g[2, 3] = np.nan
    out = O.w_update_post(g, x, z, (3, 5), 0.2, 0.3, 0.3)
    np.testing.assert_array_equal(out, np.zeros((6, 9)))
    gs = np.stack([np.full((2, 3), 1.0), np.full((2, 3), 3.0)], axis=2)
    d = O.stack_update(gs, [2.0, 1.0])
    np.testing.assert_allclose(d, np.full((2, 3), -(2.0 * 1 + 1.0 * 3) / 2))
    np.testing.assert_array_equal(O.stack_update(gs, [2.0, -3.0]), np.zeros((2, 3)))
    # the conductivity reduction (w_update_s_.m:38) has no negative-sum rule
    np.testing.assert_allclose(O.stack_update(gs, [2.0, -3.0], neg_rule=False),
                               np.full((2, 3), -(2.0 * 1 - 3.0 * 3) / 2))
    eps = np.full((2, 3), 4.0)
    d, e2 = O.stack_update(gs, [0.01, 0.02], model=eps)
    np.testing.assert_allclose(e2, 4.0 * np.exp(4.0 * (-(0.01 + 0.06) / 2)), rtol=1e-15)


def test_w_update_es_post_scales_to_unit_max(O):
    """w_update_es.m's chain: both outputs have max |g| = 1, and a positive
    scalar factor on either input leaves its output unchanged (the division
    by the maximum)."""
    rng = np.random.default_rng(4)
    x = np.linspace(0, 1, 16)
    z = np.linspace(0, 0.5, 8)
    ge, gs = rng.standard_normal((2, 8, 16))
    a, b = O.w_update_es_post(ge, gs, x, z, (2, 9), 0.1, 0.3, 0.3)
    assert np.isclose(np.abs(a).max(), 1.0) and np.isclose(np.abs(b).max(), 1.0)
    a2, b2 = O.w_update_es_post(3.0 * ge, 0.5 * gs, x, z, (2, 9), 0.1, 0.3, 0.3)
    np.testing.assert_allclose(a2, a, atol=1e-14)
    np.testing.assert_allclose(b2, b, atol=1e-14)


def test_taper_index_follows_w_update_e():
    """w_update_e.m:112-114 maps the padded-grid source back to model indices
    (ix - pjs, iz - pis - air + 1); a source left of the model is refused."""
    from gerjoii_amd.postproc import taper_index
    from gerjoii_amd.survey import Grid, src_index
    g = Grid(m=20, n=30, pml=5, air=3, nt=10, dx=0.05, dt=1e-10)
    src = [src_index(g, 0, 7), src_index(g, 4, 1)]
    np.testing.assert_array_equal(taper_index(g, src), [[1, 7], [5, 1]])
    with pytest.raises(ValueError):
        taper_index(g, [src_index(g, 0, 0)])
