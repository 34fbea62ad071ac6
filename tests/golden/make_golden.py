"""Writes tests/golden/tiny_forward_grad.npz from the FP64 oracle.

The reference ships no golden vectors for this path (its *.mat files are
gitignored and it has no tests), so this fixture is produced by the oracle
after the oracle has been pinned by tests/test_oracle.py (slice-literal
numpy reading, dot-product test, finite differences).  It freezes a tiny
forward + exact-gradient case so that later edits to the oracle are caught.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def main():
    rng = np.random.default_rng(20)
    m, n, P, air, nt = 7, 11, 3, 2, 50
    dx = 0.05
    g = O.Grid(m, n, P, air, dx, O.dt_cfl(dx), nt)
    eps = rng.uniform(3, 12, (m, n))
    sig = rng.uniform(0, 2e-2, (m, n))
    w = O.w_wavelet(np.arange(nt) * g.dt, 1e3, 1 / 400e6)
    src = np.array([P + air + 2, P + 5])
    Mw = O.sub2ind_Mw(g.NX, g.NY, np.array([P + air + 2] * 4), P + 1 + np.arange(4) * 3).astype(np.int32)
    ep, sp = O.exp2pml(g, eps, sig)
    d = O.forward(g, ep, sp, tuple(src), w, Mw)[0]
    e = rng.standard_normal(d.shape)
    ge, gs, gw = O.grad_exact(g, ep, sp, tuple(src), w, Mw, e, K=5)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_forward_grad.npz")
    np.savez(out, m=m, n=n, P=P, air=air, nt=nt, dx=dx, dt=g.dt, eps=eps, sig=sig, w=w, src=src,
             Mw=Mw, d=d, e=e, g_eps_pad=ge, g_sig_pad=gs, g_w=gw)
    print("wrote", out)


if __name__ == "__main__":
    main()
