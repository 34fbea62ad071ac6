"""MATLAB .mat I/O of the radar data (gerjoii_amd.matio).

The reference ships no .mat files (``*.mat`` is gitignored), so the format
is pinned by its own writer and readers: the fields and types w_natur_u.m:
41-65 saves, the access pattern of the project's Python viewer
(fancy_figure.py:103-120: ``loadmat(f)[struct][field].item()``), and the
index arithmetic of w_load.m:27-35 / w_M.m:20, checked against the separate
implementation in gerjoii_amd.survey (src_index, rec_Mw).
"""
import os

import numpy as np
import pytest
import scipy.io as sio

from gerjoii_amd import configs
from gerjoii_amd import matio as M
from gerjoii_amd.survey import Grid, rec_Mw, src_index


def _line(rng, nt=50, nr=7):
    d = rng.standard_normal((nt, nr)).astype(np.float32)
    r = np.stack([0.5 + 0.1 * np.arange(nr), np.zeros(nr)], axis=1)
    return dict(d=d, t=np.arange(nt) * 1e-10, dt=1e-10, fo=2.5e8, dx=0.02, r=r,
                s=np.array([0.2, 0.0]), wvlet_=rng.standard_normal(nt),
                Jy=rng.standard_normal(nt).astype(np.float32))


def test_radargram_round_trip_and_reference_access(tmp_path):
    rng = np.random.default_rng(0)
    L = _line(rng)
    p = str(tmp_path / "line1.mat")
    M.write_radargram(p, **L, v_mute=0.12, t_mute=3.0)
    # the reference viewer's access: loadmat(f)['radargram'][field].item()
    st = sio.loadmat(p)["radargram"]
    d = st["d"].item()
    assert d.dtype == np.float32 and d.shape == (50, 7)
    np.testing.assert_array_equal(d, L["d"])
    assert st["t"].item().shape == (50, 1) and st["dr"].item().shape == (1, 1)
    assert st["Jy"].item().dtype == np.float32 and st["s"].item().shape == (1, 2)
    rg = M.read_radargram(p)
    np.testing.assert_array_equal(rg.d, L["d"])
    np.testing.assert_array_equal(rg.t, L["t"])
    np.testing.assert_array_equal(rg.r, L["r"])
    np.testing.assert_array_equal(rg.wvlet_, L["wvlet_"])
    np.testing.assert_array_equal(rg.Jy, L["Jy"])
    assert (rg.dt, rg.fo, rg.dx) == (1e-10, 2.5e8, 0.02)
    # w_natur_u.m:55-56 and :35
    assert np.isclose(rg.dr, 0.1) and np.isclose(rg.dsr, 0.3)
    np.testing.assert_array_equal(rg.std_, np.ones(50 * 7))
    assert float(rg.extra["v_mute"].ravel()[0]) == 0.12


def test_cells_ragged_round_trip(tmp_path):
    pairs = [(np.array([[3.0, 4.0]]), np.arange(10.0).reshape(5, 2)),
             (np.array([[1.0, 9.0]]), np.arange(6.0).reshape(3, 2))]
    p = str(tmp_path / "s_r_.mat")
    M.write_cells(p, "s_r_", pairs)
    back = M.read_cells(p)
    for (a, b), (c, d) in zip(pairs, back):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, d)
    c = sio.loadmat(p)["s_r_"]  # a MATLAB cell: object array ns x 2
    assert c.shape == (2, 2) and c.dtype == object


def test_index_maps_agree_with_survey():
    """w_load.m:27-35 + w_M.m:20 in matio agree with survey.src_index /
    rec_Mw, and grid_to_indexed inverts them."""
    g = Grid(m=30, n=40, pml=6, air=2, nt=10, dx=0.02, dt=1e-11)
    s_off = np.array([[0, 11]])
    r_off = np.array([[1, 0], [7, 0], [39, 3]])  # (ix, iz)
    src, Mw = M.indexed_to_grid(g, s_off, r_off)
    assert tuple(src[0]) == src_index(g, 0, 11)
    np.testing.assert_array_equal(Mw, rec_Mw(g, r_off[:, 1], r_off[:, 0]))
    s2, r2 = M.grid_to_indexed(g, src[0], Mw)
    np.testing.assert_array_equal(s2, s_off)
    np.testing.assert_array_equal(r2, r_off)


def test_survey_write_load_round_trip(tmp_path):
    """write_survey (w_natur_u.m / experim_w.m) then load_survey (w_load.m)
    gives back the survey's sources, receivers, wavelets and traces."""
    case = configs.small(nt=60, ns=3)
    g = case.grid
    rng = np.random.default_rng(1)
    ntr = sum(len(r) for r in case.receivers)
    traces = rng.standard_normal((ntr, g.nt)).astype(np.float32)
    M.write_survey(str(tmp_path), g, case.src, case.receivers, traces, case.wavelets, case.f0)
    assert sorted(os.listdir(tmp_path)) == ["line1.mat", "line2.mat", "line3.mat", "s_r.mat",
                                            "s_r_.mat"]
    src, wav, recs, d_obs = M.load_survey(g, str(tmp_path), [1, 2, 3])
    np.testing.assert_array_equal(src, case.src)
    for a, b in zip(recs, case.receivers):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(d_obs, traces)
    np.testing.assert_allclose(wav, case.wavelets.astype(np.float32))


def test_refuses_v73_and_missing_struct(tmp_path):
    p = str(tmp_path / "h5.mat")
    hdr = b"MATLAB 7.3 MAT-file, Platform: GLNXA64".ljust(116, b" ") + b"\0" * 8 + b"\x00\x02IM"
    with open(p, "wb") as f:
        f.write(hdr + b"\0" * 512)
    with pytest.raises(ValueError, match="v7.3"):
        M.read_radargram(p)
    q = str(tmp_path / "other.mat")
    sio.savemat(q, {"x": np.ones(3)})
    with pytest.raises(ValueError, match="radargram"):
        M.read_radargram(q)


@pytest.mark.gpu
def test_loaded_survey_gradient(gpu_ctx, tmp_path):
    """The FWI gradient from a project folder written and read back equals
    the gradient from the in-memory survey, bit for bit."""
    from gerjoii_amd import survey as S
    case = configs.small(nt=200, ns=2)
    g = case.grid
    svt = S.Survey(g, gpu_ctx)
    svt.set_model(case.eps_true, case.sig_true)
    svt.set_shots(case.src, case.wavelets, case.receivers)
    dobs = svt.forward()
    M.write_survey(str(tmp_path), g, case.src, case.receivers, dobs, case.wavelets, case.f0)
    src, wav, recs, d_obs = M.load_survey(g, str(tmp_path), [1, 2])
    a = S.Survey(g, gpu_ctx)
    a.set_model(case.eps0, case.sig0)
    a.set_shots(src, wav, recs)
    a.set_option(3, 16)
    b = S.Survey(g, gpu_ctx)
    b.set_model(case.eps0, case.sig0)
    b.set_shots(case.src, case.wavelets, case.receivers)
    b.set_option(3, 16)
    ga, sa, Ja = a.fwi_gradient(d_obs)
    gb, sb, Jb = b.fwi_gradient(dobs)
    np.testing.assert_array_equal(ga, gb)
    np.testing.assert_array_equal(sa, sb)
    assert Ja == Jb
