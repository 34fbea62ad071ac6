"""The MEX gateway (gerjoii_amd/mex/gjw_mex.c), compiled against the mex.h
subset in tests/mex_stub/ and driven through the mxArray stand-in there, over
the real libgjw C ABI.

CPU: the gateway builds with -Wall -Wextra -Werror, references nothing beyond
the documented MEX / matrix names mex.h declares, the C ABI of include/gjw.h
and libc, and reports argument errors through mexErrMsgIdAndTxt.
GPU: each op against the FP64 oracle (or the drop-in wrapper that calls the
same entry point), the way the .m wrappers in INTEGRATION.md call it.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
STUB = os.path.join(HERE, "mex_stub")
GATEWAY = os.path.join(ROOT, "gerjoii_amd", "mex", "gjw_mex.c")

# class ids of tests/mex_stub/mex.h
STRUCT, CHAR, DOUBLE, SINGLE, INT32, UINT32 = 2, 4, 6, 7, 12, 13
NP = {DOUBLE: np.float64, SINGLE: np.float32, INT32: np.int32, UINT32: np.uint32}


def _build():
    subprocess.run(["make", "-s", "-C", STUB], check=True)
    return os.path.join(STUB, "libgjw_mex_stub.so")


class Mex:
    """ctypes driver of the stand-in: builds mxArrays and calls the gateway."""

    def __init__(self):
        from gerjoii_amd import _native as N
        N.load()
        C.CDLL(N.LIB_PATH, mode=C.RTLD_GLOBAL)  # the gateway's gjw_* resolve against it
        L = C.CDLL(_build())
        vp, sz = C.c_void_p, C.c_size_t
        L.stub_numeric.restype = vp
        L.stub_numeric.argtypes = [C.c_int, sz, sz, vp]
        L.stub_string.restype = vp
        L.stub_string.argtypes = [C.c_char_p]
        L.stub_struct.restype = vp
        L.stub_set_field.argtypes = [vp, C.c_char_p, vp]
        L.stub_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(sz), C.POINTER(sz), C.POINTER(vp)]
        L.stub_free.argtypes = [vp]
        L.stub_call.argtypes = [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_char_p, sz,
                                C.c_char_p, sz]
        self.L = L

    def arr(self, x, cls=None):
        """numpy array (1-D: a column) or str or dict -> mxArray."""
        if isinstance(x, str):
            return self.L.stub_string(x.encode())
        if isinstance(x, dict):
            s = self.L.stub_struct()
            for k, v in x.items():
                assert self.L.stub_set_field(s, k.encode(), self.arr(v)) == 0
            return s
        x = np.asarray(x)
        if cls is None:
            cls = {np.dtype(np.float32): SINGLE, np.dtype(np.int32): INT32,
                   np.dtype(np.uint32): UINT32}.get(x.dtype, DOUBLE)
        a = np.asfortranarray(np.atleast_1d(x), dtype=NP[cls])
        m, n = (a.shape[0], 1) if a.ndim == 1 else a.shape
        return self.L.stub_numeric(cls, m, n, a.ctypes.data)

    def get(self, p):
        cls, m, n, d = C.c_int(0), C.c_size_t(0), C.c_size_t(0), C.c_void_p()
        self.L.stub_info(p, C.byref(cls), C.byref(m), C.byref(n), C.byref(d))
        dt = NP[cls.value]
        cnt = m.value * n.value
        if cnt == 0:
            return np.zeros((m.value, n.value), dtype=dt)
        buf = (C.c_char * (cnt * np.dtype(dt).itemsize)).from_address(d.value)
        return np.frombuffer(buf, dtype=dt).reshape((m.value, n.value), order="F").copy()

    def call(self, nlhs, *args):
        """Returns the outputs as numpy arrays, or raises RuntimeError(id, msg)."""
        prhs = (C.c_void_p * len(args))(*[self.arr(a) for a in args])
        plhs = (C.c_void_p * max(nlhs, 1))()
        eid, msg = C.create_string_buffer(128), C.create_string_buffer(1024)
        rc = self.L.stub_call(nlhs, plhs, len(args), prhs, eid, 128, msg, 1024)
        for p in prhs:
            self.L.stub_free(p)
        if rc:
            raise RuntimeError(eid.value.decode(), msg.value.decode())
        out = [self.get(plhs[k]) for k in range(nlhs)]
        for k in range(nlhs):
            self.L.stub_free(plhs[k])
        return out


@pytest.fixture(scope="module")
def mex():
    return Mex()


def test_gateway_compiles_against_documented_api_only(tmp_path):
    """-Wall -Wextra -Werror; every undefined symbol of the gateway object is
    a name mex.h declares, a C-ABI entry point of include/gjw.h, or libc."""
    _build()
    obj = str(tmp_path / "gjw_mex.o")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-Wall", "-Wextra", "-Werror", "-c",
                    "-I" + STUB, "-I" + os.path.join(ROOT, "include"), GATEWAY, "-o", obj],
                   check=True)
    nm = subprocess.run(["nm", "-u", obj], check=True, capture_output=True, text=True).stdout
    und = {ln.split()[-1] for ln in nm.splitlines() if ln.strip()}
    mexh = open(os.path.join(STUB, "mex.h")).read()
    declared = set(re.findall(r"\b(mx\w+|mex\w+)\s*\(", mexh))
    gjwh = open(os.path.join(ROOT, "include", "gjw.h")).read()
    abi = set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(gjw_\w+)\s*\(", gjwh, flags=re.M))
    libc = {"memset", "memcmp", "strcmp", "snprintf", "__stack_chk_fail", "memcpy"}
    mx = {s for s in und if s.startswith(("mx", "mex"))}
    gj = {s for s in und if s.startswith("gjw_")}
    assert mx <= declared, mx - declared
    assert gj <= abi, gj - abi
    assert und - mx - gj <= libc, und - mx - gj - libc


def test_gateway_argument_errors(mex):
    """Errors come back through mexErrMsgIdAndTxt with an identifier, before
    any GPU work (these run without a GPU)."""
    for args, eid in [(("nope",), "gjw:op"), (("fwd", 1.0), "gjw:nrhs"),
                      (("grad", "EXACT", np.zeros((3, 2), np.float32), 1.0), "gjw:state"),
                      (("grad", "FOO", np.zeros((3, 2), np.float32), 1.0), "gjw:arg"),
                      ((np.ones(2),), "gjw:arg")]:
        with pytest.raises(RuntimeError) as ei:
            mex.call(1, *args)
        assert ei.value.args[0] == eid, ei.value.args


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------
def structs(case, q, d_o):
    g = case.grid
    geome_ = {"dx": g.dx, "w": {"dt": g.dt, "nt": float(g.nt)}}
    parame_ = {"w": {"epsilon": case.eps0, "sigma": case.sig0, "eps_0": g.eps0, "mu_0": g.mu0,
                     "air": float(g.air)}, "natu": {"w": {"d_2d": d_o}}}
    finite_ = {"w": {"pml_w": float(g.pml), "pml_d": float(g.pml), "pml_order": float(g.pml_order),
                     "R_0": g.R0}}
    gerjoii_ = {"w": {"s": np.asarray(case.src[q], dtype=np.float64),
                      "M": np.asarray(case.receivers[q], dtype=np.uint32),
                      "wvlet_": np.asarray(case.wavelets[q], dtype=np.float64)}}
    return geome_, parame_, finite_, gerjoii_


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.gpu
def test_gateway_fwd_grad_bozdag(oracle_mod, gpu_ctx, mex):
    """'fwd' then 'grad' (EXACT and REF, dtu 1 / 0) then 'bozdag', as the
    w_fwd.m / w_grad.m / w_bozdag.m wrappers call them."""
    from gerjoii_amd import configs
    O = oracle_mod
    case = configs.small(nt=200, ns=2)
    case.sig0[:] = np.clip(case.sig0, 0, 5e-3)
    g = case.grid
    og = O.Grid(g.m, g.n, g.pml, g.air, g.dx, g.dt, g.nt)
    ep, sp = O.exp2pml(og, case.eps0, case.sig0)
    d, jz, u = O.forward(og, ep, sp, case.src[0], case.wavelets[0], case.receivers[0], store_u=True)
    d_o = (0.9 * d).astype(np.float32)
    geome_, parame_, finite_, gerjoii_ = structs(case, 0, d_o)
    d_2d, jy = mex.call(2, "fwd", geome_, parame_, finite_, gerjoii_)
    print("mex fwd", rel_l2(d_2d, d), rel_l2(jy.ravel(), jz), flush=True)
    assert d_2d.shape == d.shape and d_2d.dtype == np.float32
    assert rel_l2(d_2d, d) <= 1e-4 and rel_l2(jy.ravel(), jz) <= 1e-4
    e, de, doe, J = mex.call(4, "bozdag", d_o, d_2d, "L2")
    np.testing.assert_array_equal(e, d_2d - d_o)
    np.testing.assert_array_equal(de, d_2d)
    assert abs(J[0, 0] - 0.5 * float(((d_2d.astype(np.float64) - d_o) ** 2).sum())) <= 1e-6 * J[0, 0]
    ge, am = mex.call(2, "grad", "EXACT", e, 1.0)
    gs, _ = mex.call(2, "grad", "EXACT", e, 0.0)
    fe, fs, _ = O.grad_exact(og, ep, sp, case.src[0], case.wavelets[0], case.receivers[0],
                             e.astype(np.float64), pml_terms=False)
    re_, rs_ = O.fold_pad(og, fe, fs)
    print("mex grad EXACT", rel_l2(ge, re_), rel_l2(gs, rs_), flush=True)
    assert am.size == 0 and rel_l2(ge, re_) <= 1e-3 and rel_l2(gs, rs_) <= 1e-3
    ga, ama = mex.call(2, "grad", "REF", e, 1.0)
    ra, amax = O.grad_ref(og, ep, sp, case.src[0], case.receivers[0], e.astype(np.float64),
                          O.differentiate_cube(u, g.dt))
    print("mex grad REF", rel_l2(ga, ra), rel_l2(ama, amax), flush=True)
    assert rel_l2(ga, ra) <= 1e-3 and rel_l2(ama, amax) <= 1e-3 and ama.dtype == np.float32
    # Hilbert objectives against the true model's traces (0.9 d would leave a
    # zero phase difference, i.e. adjoint sources of pure rounding noise)
    et, st_ = O.exp2pml(og, case.eps_true, case.sig_true)
    d_t = O.forward(og, et, st_, case.src[0], case.wavelets[0], case.receivers[0])[0]
    d_t = d_t.astype(np.float32)
    for obj in ("env_L2", "phase_L2"):
        e2, de2, doe2, _ = mex.call(4, "bozdag", d_t, d_2d, obj)
        r2, rde2, rdoe2 = O.w_bozdag(d_t.astype(np.float64), d_2d.astype(np.float64), obj)
        keep = np.abs(O.hilbert_analytic(d_2d, axis=0)) > 1e-3 * np.abs(d_2d).max()
        print("mex bozdag", obj, rel_l2(e2[keep], r2[keep]), rel_l2(de2[keep], rde2[keep]), flush=True)
        assert rel_l2(e2[keep], r2[keep]) <= 1e-3 and rel_l2(de2[keep], rde2[keep]) <= 1e-3


@pytest.mark.gpu
def test_gateway_conditioning_and_postprocessing(oracle_mod, gpu_ctx, mex):
    """'filt_gauss', 'lmute', 'gaussi', 'image_gaussian', 'normali'."""
    O = oracle_mod
    rng = np.random.default_rng(12)
    nt, nr, dt = 300, 4, 4e-11
    d = rng.standard_normal((nt, nr)).astype(np.float32)
    d_o = rng.standard_normal((nt, nr)).astype(np.float32)
    (df,) = mex.call(1, "filt_gauss", d, dt, 5e7, 4e8, 8.0)
    ref = O.filt_gauss(d.astype(np.float64), dt, 5e7, 4e8, 8)
    assert rel_l2(df, ref) <= 1e-5
    t = np.arange(nt) * dt
    rx = 0.1 * np.arange(nr)
    dm, dom, e = mex.call(3, "lmute", d, d_o, t, rx, 0.12, 1.5)
    rdm, rdom, re = O.lmute(d, d_o, t, rx, 0.12, 1.5)
    assert rel_l2(dm, rdm) <= 1e-6 and rel_l2(dom, rdom) <= 1e-6 and rel_l2(e, re) <= 1e-6
    m, n = 37, 52
    img = rng.standard_normal((m, n))
    x, z = np.arange(n) * 0.02, np.arange(m) * 0.02
    (gt,) = mex.call(1, "gaussi", np.ones((m, n)), x, z, np.array([10, 20], np.int32), 0.1)
    np.testing.assert_allclose(gt, O.gaussi(x, z, (10, 20), 0.1), rtol=0, atol=1e-13)
    for typ in ("LOW_PASS", "HI_PASS"):
        (fi,) = mex.call(1, "image_gaussian", img, 0.05, 0.07, typ)
        np.testing.assert_allclose(fi, O.image_gaussian(img, 0.05, 0.07, typ), rtol=0,
                                   atol=1e-12 * np.abs(img).max())
    (nm_,) = mex.call(1, "normali", img)
    np.testing.assert_array_equal(nm_, O.normali(img))
