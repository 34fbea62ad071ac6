"""gerjoii_amd -- MI355X-native radar FWI hot path of diegozain/gerjoii.

The 2-D TMz Yee FDTD forward solve (src/w/w_solve.m) and its adjoint gradient
(src/w/w_grad.m) as hand-written HIP kernels for gfx950 behind a C ABI
(include/gjw.h).  This package is the Python host side over that ABI:

* ``survey``   -- Grid / Context / Survey (ctypes over libgjw.so)
* ``w``        -- the reference's MATLAB entry points (w_fwd, w_grad, ...)
                  with the same struct fields, as a drop-in for Python callers
* ``autograd`` -- torch.autograd.Function: (eps, sigma) -> traces
* ``configs``  -- seeded synthetic surveys
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose=False):
    """Compile libgjw.so for gfx950 in-tree (hipcc; runs without a GPU)."""
    cmd = ["make", "-B", "-s", "-C", os.path.join(HERE, "csrc")]
    r = subprocess.run(cmd, capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise RuntimeError("libgjw build failed:\n" + (r.stdout or "") + (r.stderr or ""))
    return os.path.join(HERE, "libgjw.so")


from .survey import Batch, Context, Grid, Survey, context, dt_cfl, w_wavelet  # noqa: E402,F401
