"""ctypes binding of libgjw.so (the HIP kernels and their C ABI, include/gjw.h).

The library is built in-tree by ``gerjoii_amd/csrc/Makefile`` (see
``gerjoii_amd.build``).  Loading fails loudly when it is missing: there is no
CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgjw.so")

GJW_OK = 0
GJW_PATH_BATCHED, GJW_PATH_PLAIN = 0, 1
GJW_OPT_PATH, GJW_OPT_CKPT_EVERY, GJW_OPT_TILE_L, GJW_OPT_SNAPSHOT, GJW_OPT_KTIMING = 1, 2, 3, 4, 5
GJW_OPT_FUSE = 6
GJW_OPT_GRAPH = 7
GJW_OPT_MISFIT = 8
GJW_OPT_ADJ_FUSE = 9
GJW_OPT_SHOT_GRADS = 10
GJW_OPT_ADJ_TILE = 11
GJW_OPT_FRAMES = 12
GJW_FRAMES_F32, GJW_FRAMES_BF16 = 0, 1
GJW_OPT_ROW_PAIRS = 13
GJW_LOW_PASS, GJW_HI_PASS = 0, 1
GJW_SNAP_F32, GJW_SNAP_BF16, GJW_SNAP_F16 = 0, 1, 2
GJW_MISFIT_L2 = 0
GJW_MISFIT_ENV_L2, GJW_MISFIT_PHASE_L2, GJW_MISFIT_ENV_L22, GJW_MISFIT_ENV_LOG = 1, 2, 3, 4
MISFIT_KINDS = {"L2": 0, "FREQstp": 0, "env_L2": 1, "phase_L2": 2, "env_L22": 3, "env_LOG": 4}
GJW_MISFIT_FREQSTP = 5  # line-search objectives: L2 after an extra band filter
GJW_GRAD_EXACT, GJW_GRAD_REF = 0, 1
GJW_Q_CKPT_INTERVAL, GJW_Q_FWD_DEPTH, GJW_Q_ADJ_DEPTH, GJW_Q_FRAMES = 1, 2, 3, 4
GJW_Q_ROW_PAIRS = 5

# every symbol include/gjw.h declares
EXPORTS = ["gjw_create", "gjw_destroy", "gjw_last_error", "gjw_version", "gjw_survey_create",
           "gjw_survey_destroy", "gjw_set_model", "gjw_set_shots", "gjw_set_option",
           "gjw_forward", "gjw_misfit", "gjw_bozdag", "gjw_gradient", "gjw_fwi_gradient",
           "gjw_differentiate_cube", "gjw_last_timing", "gjw_kernel_stats", "gjw_tile",
           "gjw_batch_create", "gjw_batch_destroy", "gjw_batch_fwi_gradient",
           "gjw_batch_last_timing", "gjw_set_adj_filter", "gjw_filt_gauss", "gjw_lmute",
           "gjw_batch_objective", "gjw_gaussi", "gjw_image_gaussian", "gjw_normali",
           "gjw_stack_update", "gjw_query", "gjw_w_bozdag", "gjw_batch_objective_multi",
           "gjw_batch_forward", "gjw_normali_noalloc"]


class GjwGrid(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("pml", C.c_int32), ("air", C.c_int32),
                ("nt", C.c_int32), ("pml_order", C.c_int32), ("dx", C.c_double),
                ("dt", C.c_double), ("eps0", C.c_double), ("mu0", C.c_double),
                ("R0", C.c_double)]


class GjwError(RuntimeError):
    pass


_lib = None


def load():
    """Load libgjw.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GjwError(f"{LIB_PATH} is missing: build it with gerjoii_amd.build() "
                       "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    # torch ships its own HIP runtime under the same soname.  Loaded first, it
    # is the one runtime of the process and libgjw binds to it; loaded after
    # libgjw, a second runtime comes in and torch finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32, i64 = C.c_int32, C.c_int64
    lib.gjw_last_error.restype = C.c_char_p
    lib.gjw_version.restype = C.c_char_p
    lib.gjw_create.argtypes = [i32, C.POINTER(vp)]
    lib.gjw_destroy.argtypes = [vp]
    lib.gjw_survey_create.argtypes = [vp, C.POINTER(GjwGrid), C.POINTER(vp)]
    lib.gjw_survey_destroy.argtypes = [vp]
    lib.gjw_set_model.argtypes = [vp, vp, vp]
    lib.gjw_set_shots.argtypes = [vp, i32, vp, vp, vp, vp]
    lib.gjw_set_option.argtypes = [vp, i32, i64]
    lib.gjw_forward.argtypes = [vp, vp, vp, i32]
    lib.gjw_misfit.argtypes = [vp, vp, i32, vp, vp]
    lib.gjw_bozdag.argtypes = [vp, vp, vp, i32, vp, vp]
    lib.gjw_batch_create.argtypes = [vp, i32, vp, vp]
    lib.gjw_batch_destroy.argtypes = [vp]
    lib.gjw_batch_fwi_gradient.argtypes = [vp, vp, vp, vp, vp]
    lib.gjw_batch_last_timing.argtypes = [vp, vp, vp]
    lib.gjw_batch_objective.argtypes = [vp, vp, vp]
    lib.gjw_gradient.argtypes = [vp, vp, i32, vp, vp, vp, vp]
    lib.gjw_fwi_gradient.argtypes = [vp, vp, vp, vp, vp]
    lib.gjw_differentiate_cube.argtypes = [vp, vp, i32, i32, i32, C.c_double]
    lib.gjw_last_timing.argtypes = [vp, vp, vp]
    lib.gjw_set_adj_filter.argtypes = [vp, C.c_double, C.c_double, i32]
    lib.gjw_filt_gauss.argtypes = [vp, vp, i32, i32, C.c_double, C.c_double, C.c_double, i32]
    lib.gjw_lmute.argtypes = [vp, vp, vp, vp, i32, i32, vp, vp, C.c_double, C.c_double]
    lib.gjw_gaussi.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp]
    lib.gjw_image_gaussian.argtypes = [vp, vp, vp, i32, i32, i32, C.c_double, C.c_double, i32]
    lib.gjw_normali.argtypes = [vp, vp, i32, i32, i32, i32, vp]
    lib.gjw_stack_update.argtypes = [vp, vp, i32, i32, i32, vp, i32, vp, vp]
    lib.gjw_kernel_stats.argtypes = [vp, i32, vp, vp, vp, i32]
    lib.gjw_tile.argtypes = [vp, C.c_int32, vp]
    lib.gjw_query.argtypes = [vp, i32, vp]
    lib.gjw_w_bozdag.argtypes = [vp, vp, vp, i32, i32, i32, vp, vp, vp, vp]
    lib.gjw_batch_objective_multi.argtypes = [vp, vp, i32, vp, vp, i32, C.c_double, C.c_double,
                                              vp]
    lib.gjw_batch_forward.argtypes = [vp, vp]
    lib.gjw_normali_noalloc.argtypes = [vp, vp, i32, i32, i32, i32, vp]
    for name in EXPORTS:
        getattr(lib, name).restype = getattr(lib, name).restype or C.c_int
    lib.gjw_last_error.restype = C.c_char_p
    lib.gjw_version.restype = C.c_char_p
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != GJW_OK:
        msg = load().gjw_last_error().decode(errors="replace")
        raise GjwError(f"{what} failed ({rc}): {msg}")
    return rc


def ptr(x):
    """Address of a numpy array or torch tensor (host or device), or None."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    if hasattr(x, "ctypes"):
        return C.c_void_p(x.ctypes.data)
    raise TypeError(type(x))
