"""Host arithmetic of the parabolic line search (w_stepparabola_e.m,
w_boundstep2.m, w_objs_e.m): closed forms and exhaustive checks on the
search grids; no GPU."""
import numpy as np

from gerjoii_amd import linesearch as LS


def test_parabola_step_closed_forms():
    x = np.array([0.0, -0.2, 0.1, 0.4, 0.7])
    assert abs(LS.parabola_step(x, 3 * (x - 0.25) ** 2 + 1) - 0.25) < 1e-12
    assert LS.parabola_step(x, (x + 0.3) ** 2) == 0.0     # vertex at a negative step
    # least squares through noisy samples of a parabola stays near its vertex
    rng = np.random.default_rng(0)
    y = (x - 0.3) ** 2 + 1e-4 * rng.standard_normal(x.size)
    assert abs(LS.parabola_step(x, y) - 0.3) < 0.02


def _admissible(k, e, g, hi, lo):
    p = e * np.exp(k * e * g)
    return p.max() <= hi and p.min() >= lo


def test_boundstep2_is_the_last_admissible_grid_step():
    rng = np.random.default_rng(1)
    for trial in range(20):
        e = rng.uniform(3, 12, 200)
        g = rng.standard_normal(200) * rng.uniform(0.1, 3)
        hi, lo = 20.0, 1.0
        k_ = [1e-5, 10 ** rng.uniform(-2, 0)]
        k = LS.boundstep2(e, g, hi, lo, k_)
        coarse = np.logspace(np.log10(max(k_)), np.log10(min(k_)), 13)
        if k == coarse[0]:
            assert _admissible(k, e, g, hi, lo)
            continue
        # exhaustive: first admissible point of the coarse reverse grid ...
        i = next((j for j in range(13) if _admissible(coarse[j], e, g, hi, lo)), 12)
        fine = np.linspace(coarse[i], coarse[i - 1], 20)
        # ... then the last admissible point of the fine forward grid
        ok = [_admissible(f, e, g, hi, lo) for f in fine]
        j = (ok.index(False) - 1) if False in ok else 19
        assert k == fine[max(j, 0)]
        if _admissible(coarse[i], e, g, hi, lo):
            assert _admissible(k, e, g, hi, lo)


def test_perturbation_matches_w_objs_e():
    e = np.array([[4.0, 9.0], [16.0, 1.5]])
    g = np.array([[0.1, -0.2], [0.0, 1.0]])
    np.testing.assert_allclose(LS.perturb_eps(e, g, 0.05), e * np.exp(-0.05 * e * g))
    np.testing.assert_array_equal(LS.perturb_eps(e, g, 0.0), e)
