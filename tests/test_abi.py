"""The C ABI (include/gjw.h): the in-tree library loads on a host without a
GPU and exports every function the header declares, and the ctypes binding
lists exactly those (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gjw.h")).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)  # drop comments
    names = re.findall(r"^\s*(?:int\s+|const char\s*\*\s*)(gjw_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = _declared()
    for must in ("gjw_create", "gjw_forward", "gjw_gradient", "gjw_fwi_gradient",
                 "gjw_batch_fwi_gradient", "gjw_batch_objective", "gjw_filt_gauss", "gjw_lmute"):
        assert must in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from gerjoii_amd import _native as N
    assert os.path.exists(N.LIB_PATH), "build the library first (gerjoii_amd/csrc/Makefile)"
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(N.EXPORTS) == _declared()


def test_one_hip_runtime_per_process():
    """Loading libgjw before torch must not bring a second HIP runtime into
    the process (torch ships its own under the same soname; two runtimes
    leave torch without a GPU).  Run in a fresh interpreter."""
    import subprocess
    import sys
    code = ("import gerjoii_amd._native as N; N.load(); import torch; "
            "m = open('/proc/self/maps').read().splitlines(); "
            "print(len({l.split()[-1] for l in m if 'libamdhip64' in l}))")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip().splitlines()[-1] == "1"


def test_context_close_releases_its_surveys_first():
    """Context.close() destroys the context's batches and surveys before the
    context itself (a survey destroyed after its context would use freed
    memory); the closed objects' later close() is a no-op."""
    from gerjoii_amd import survey as S
    calls = []

    class FakeLib:
        def __getattr__(self, name):
            def f(*a):
                calls.append(name)
                return 0
            return f

    from gerjoii_amd import _native as N
    saved = N._lib
    N._lib = FakeLib()
    try:
        import ctypes
        cx = S.Context(0)
        cx.h = ctypes.c_void_p(1)  # the stand-in library hands out no handles
        sv = S.Survey(S.Grid(m=4, n=5, pml=2, air=0, nt=3, dx=0.1, dt=1e-10), cx)
        sv.h = ctypes.c_void_p(2)
        bt = S.Batch([sv])
        bt.h = ctypes.c_void_p(3)
        cx.close()
        assert calls[-3:] == ["gjw_batch_destroy", "gjw_survey_destroy", "gjw_destroy"]
        n = len(calls)
        bt.close()
        sv.close()
        assert len(calls) == n
    finally:
        N._lib = saved
