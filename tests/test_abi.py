"""CPU checks of the boundary: libgsks.so builds/loads, exports every entry
point include/ks.h and include/omp_dgsks_list.hpp declare, keeps the
reference's ks_t layout and enum order, and fails loudly (error codes, no
CPU fallback) where there is no GPU.  No GPU compute is called here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gpu_available

LIB = os.path.join(ROOT, "ks_amd", "libgsks.so")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(LIB):
        from ks_amd import build
        build.build()
    from ks_amd import _lib
    return _lib.lib()


def _declared_c_functions(header):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       src, flags=re.M)
    return sorted({n for n in names if n not in ("if", "while", "for", "return", "sizeof")})


def test_exports_every_declared_c_symbol(L):
    names = _declared_c_functions(os.path.join(ROOT, "include", "ks.h"))
    assert {"dgsks", "dgsks_ref", "ks_malloc_aligned", "dgsks_wrapper",
            "dgsks_ref_wrapper", "dgsks_device", "dgsks_list"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_exports_cpp_list_overloads():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True,
                         text=True).stdout
    for fn in ("omp_dgsks_list_unsymmetric", "omp_dgsks_list_symmetric",
               "omp_dgsks_list_separated_u_unsymmetric", "omp_dgsks_list_separated_u_symmetric",
               "omp_dgsks_list(kernel_s*"):
        assert fn in out, fn


def test_header_compiles_as_c_and_cpp(tmp_path):
    """The reference's callers include ks.h from C and (inside extern "C") from C++."""
    c = tmp_path / "t.c"
    c.write_text('#include "ks.h"\n#include <stddef.h>\n'
                 '_Static_assert(sizeof(ks_t) == 56, "ks_t size");\n'
                 '_Static_assert(offsetof(ks_t, hi) == 32, "hi offset");\n'
                 '_Static_assert(KS_EPANECHNIKOV == 7 && KS_LAPLACE == 2, "enum order");\n'
                 'int main(void){ ks_t k; k.type = KS_GAUSSIAN; (void)k; return 0; }\n')
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-c", str(c), "-o",
                    str(tmp_path / "t.o")], check=True)
    cc = tmp_path / "t.cpp"
    cc.write_text('extern "C" {\n#include "ks.h"\n}\n#include "omp_dgsks_list.hpp"\n'
                  'static_assert(sizeof(ks_t) == 56, "size");\nint main(){ return 0; }\n')
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-I", inc, "-c", str(cc), "-o",
                    str(tmp_path / "tc.o")], check=True)


def test_kst_layout_matches_header(L):
    from ks_amd._lib import KsT
    assert ctypes.sizeof(KsT) == 56
    assert KsT.hi.offset == 32 and KsT.powe.offset == 8


def test_error_strings_and_modes(L):
    assert L.ks_error_string(0) == b"ok"
    assert b"illegal kernel type" in L.ks_error_string(1)
    old = L.ks_get_error_mode()
    L.ks_set_error_mode(1)
    assert L.ks_get_error_mode() == 1
    # a failed device call leaves its code; a later void entry point clears it
    # at its start (here an early return: m == 0), so nothing stale is reported
    assert L.sgsks_device(None, 1, 1, 1, *([None] * 10), 0, None) == 2
    for fn in (L.dgsks, L.dgsks_ref, L.sgsks):
        L.dgsks_device(None, 1, 1, 1, *([None] * 10), 0, None)
        fn(None, 0, 1, 1, *([None] * 10))
        assert L.ks_last_error() == 0
    L.ks_set_error_mode(old)


def test_malloc_aligned(L):
    p = L.ks_malloc_aligned(10, 7, 8)
    assert p and p % 64 == 0
    ctypes.CDLL(None).free(ctypes.c_void_p(p))


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU failure path")
def test_no_gpu_fails_loudly(L):
    """Without a GPU every summation entry point reports KS_ERR_DEVICE; nothing
    is computed on the CPU."""
    from ks_amd import _lib
    from ks_amd._lib import KsT
    L.ks_set_error_mode(_lib.KS_ERRMODE_RETURN)
    kt = KsT()
    kt.type = 0
    kt.scal = -0.5
    X = np.random.default_rng(0).random((8, 3))
    X2 = (X * X).sum(1)
    idx = np.arange(8, dtype=np.int32)
    u = np.zeros(8)
    w = np.ones(8)
    ptr = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = L.dgsks_device(ctypes.byref(kt), 8, 8, 3, ptr(u), ptr(idx), ptr(X), ptr(X2), ptr(idx),
                        ptr(X), ptr(X2), ptr(idx), ptr(w), ptr(idx), 0, None)
    assert rc == _lib.KS_ERR_DEVICE
    L.dgsks(ctypes.byref(kt), 8, 8, 3, ptr(u), ptr(idx), ptr(X), ptr(X2), ptr(idx), ptr(X),
            ptr(X2), ptr(idx), ptr(w), ptr(idx))
    assert L.ks_last_error() == _lib.KS_ERR_DEVICE
    assert np.all(u == 0)
    from ks_amd import gsks
    with pytest.raises(_lib.KsError):
        gsks.dgsks(3, u, idx, X, X2, idx, X, X2, idx, w, idx, type="Gaussian", scal=-0.5)
    # single precision: the same, for the device entry point and the host one
    Xf, uf, wf = X.astype(np.float32), u.astype(np.float32), w.astype(np.float32)
    rc = L.sgsks_device(ctypes.byref(kt), 8, 8, 3, ptr(uf), ptr(idx), ptr(Xf), None, ptr(idx),
                        ptr(Xf), None, ptr(idx), ptr(wf), ptr(idx), 0, None)
    assert rc == _lib.KS_ERR_DEVICE
    L.sgsks(ctypes.byref(kt), 8, 8, 3, ptr(uf), ptr(idx), ptr(Xf), None, ptr(idx), ptr(Xf), None,
            ptr(idx), ptr(wf), ptr(idx))
    assert L.ks_last_error() == _lib.KS_ERR_DEVICE
    assert np.all(uf == 0)
    with pytest.raises(_lib.KsError):
        gsks.sgsks(3, uf, idx, Xf, None, idx, Xf, None, idx, wf, idx, type="Gaussian", scal=-0.5)


def test_python_argument_checks(L):
    """python/gsks.py:19-36 asserts, raised before any library call."""
    from ks_amd import gsks
    X = np.zeros((10, 3))
    X2 = np.zeros(10)
    idx = np.arange(10, dtype=np.int32)
    u, w = np.zeros(10), np.ones(10)
    with pytest.raises(AssertionError):
        gsks.dgsks(3, u, idx, X, X2, idx.astype(np.int64), X, X2, idx, w, idx, type="Gaussian")
    with pytest.raises(AssertionError):
        gsks.dgsks(3, u, idx, X, X2[:5], idx, X, X2, idx, w, idx, type="Gaussian")
    with pytest.raises(AssertionError):
        gsks.dgsks(3, u, idx, X, X2, idx, np.zeros((10, 4)), X2, idx, w, idx, type="Gaussian")
    # sgsks: float32 arrays, norms optional, no explicit precision for host arrays
    Xf, uf, wf = X.astype(np.float32), u.astype(np.float32), w.astype(np.float32)
    with pytest.raises(AssertionError):
        gsks.sgsks(3, u, idx, Xf, None, idx, Xf, None, idx, wf, idx, type="Gaussian")
    with pytest.raises(AssertionError):
        gsks.sgsks(3, uf, idx, Xf, X2, idx, Xf, None, idx, wf, idx, type="Gaussian")
    with pytest.raises(AssertionError):
        gsks.sgsks(3, uf, idx, Xf, None, idx, Xf, None, idx, wf, idx, type="Gaussian",
                   precision="fp64")
    with pytest.raises(AssertionError):
        gsks.dgsks(3, u, idx, X, None, idx, X, X2, idx, w, idx, type="Gaussian")


def test_unsupported_type_prints_and_returns(L, capsys):
    from ks_amd import gsks
    X = np.zeros((4, 2))
    X2 = np.zeros(4)
    idx = np.arange(4, dtype=np.int32)
    u = np.zeros(4)
    gsks.dgsks(2, u, idx, X, X2, idx, X, X2, idx, np.ones(4), idx, type="Cauchy")
    assert "not suppurted" in capsys.readouterr().out


@pytest.mark.skipif(gpu_available(), reason="checks host paths that stop before the GPU")
def test_host_argument_checks_before_the_device(L):
    """Argument and list checks that return before any device work: codes in
    KS_ERRMODE_RETURN mode, and the Laplace descriptor side effect of dgsks
    (frame/dgsks.c:445-447), which happens before the summation."""
    from ks_amd import _lib
    from ks_amd._lib import KsT
    L.ks_set_error_mode(_lib.KS_ERRMODE_RETURN)
    ptr = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    X = np.random.default_rng(1).random((8, 3))
    X2 = (X * X).sum(1)
    idx = np.arange(8, dtype=np.int32)
    u, w = np.zeros(8), np.ones(8)
    # Laplace k = 3: powe = -1/2, scal = 1/(4 pi) are written into the descriptor
    kt = KsT()
    kt.type = _lib.KS_LAPLACE
    L.dgsks(ctypes.byref(kt), 8, 8, 3, ptr(u), ptr(idx), ptr(X), ptr(X2), ptr(idx), ptr(X),
            ptr(X2), ptr(idx), ptr(w), ptr(idx))
    assert kt.powe == -0.5 and abs(kt.scal - 1.0 / (4.0 * np.pi)) <= 1e-17
    assert L.ks_last_error() == _lib.KS_ERR_DEVICE
    # an illegal type and a variable bandwidth without bandwidths
    args = [ptr(u), ptr(idx), ptr(X), ptr(X2), ptr(idx), ptr(X), ptr(X2), ptr(idx), ptr(w),
            ptr(idx)]
    L.dgsks_wrapper(8, 8, 3, *args, 11, -0.5, 0.0, 0.0, None)
    assert L.ks_last_error() == _lib.KS_ERR_INVALID_TYPE
    L.dgsks_wrapper(8, 8, 3, *args, _lib.KS_GAUSSIAN_VAR_BANDWIDTH, 0.0, 0.0, 0.0, None)
    assert L.ks_last_error() == _lib.KS_ERR_NO_BANDWIDTH
    assert np.all(u == 0)
    # batched entry points: inconsistent offsets, a_off[0] != 0, bad counts
    kts = (KsT * 2)()
    off_bad = np.array([0, 5, 3], dtype=np.int64)
    off_one = np.array([1, 4, 8], dtype=np.int64)
    off = np.array([0, 4, 8], dtype=np.int64)
    lst = [ptr(idx), ptr(idx)]

    def dl(ntask, a_off, b_off):
        return L.dgsks_list(ntask, kts, 3, ptr(u), 8, ptr(a_off), *lst, ptr(b_off), *lst,
                            ptr(X), ptr(X2), ptr(X), ptr(X2), ptr(w), 0)
    assert dl(2, off_bad, off) == _lib.KS_ERR_LIST
    assert dl(2, off, off_bad) == _lib.KS_ERR_LIST
    assert dl(2, off_one, off) == _lib.KS_ERR_LIST
    assert dl(-1, off, off) == _lib.KS_ERR_INVALID_ARG
    assert dl(0, off, off) == _lib.KS_OK
    h = ctypes.c_void_p()
    assert L.ks_list_plan_create(ctypes.byref(h), 2, kts, 3, ptr(off), *lst, ptr(off), *lst,
                                 7) == _lib.KS_ERR_INVALID_ARG
    assert L.ks_list_plan_create(ctypes.byref(h), 2, kts, 3, ptr(off_bad), *lst, ptr(off), *lst,
                                 0) == _lib.KS_ERR_LIST
    assert not h.value
    assert L.ks_list_plan_execute(None, *([None] * 7)) == _lib.KS_ERR_INVALID_ARG
    L.ks_list_plan_destroy(None)
    a, b = ctypes.c_int(), ctypes.c_int()
    assert L.ks_list_autotune(9, 3, 0, ctypes.byref(a), ctypes.byref(b)) == \
        _lib.KS_ERR_INVALID_ARG
    assert np.all(u == 0)
    # mapped gradient: kernel types, outputs of the other kernel, precision, k
    one = np.zeros(1)

    def grad(kt, k=3, prec=0, dscal=None, dhi=None):
        return L.ks_grad_device(ctypes.byref(kt) if kt is not None else None, 8, 8, k, ptr(w),
                                ptr(idx), ptr(X), ptr(idx), ptr(X), ptr(idx), ptr(w), ptr(idx),
                                None, None, None, dscal, dhi, None, prec, None)
    gk = KsT()
    gk.type = _lib.KS_GAUSSIAN
    gk.scal = -1.0
    vk = KsT()
    vk.type = _lib.KS_GAUSSIAN_VAR_BANDWIDTH
    assert grad(None) == _lib.KS_ERR_INVALID_ARG
    assert grad(kt) == _lib.KS_ERR_INVALID_ARG  # Laplace: not a Gaussian kernel
    assert grad(gk, dhi=ptr(one)) == _lib.KS_ERR_INVALID_ARG
    assert grad(gk, prec=_lib.KS_PREC_MIXED_BF16X3) == _lib.KS_ERR_INVALID_ARG
    assert grad(gk, k=72) == _lib.KS_ERR_INVALID_ARG
    assert grad(gk, k=33, prec=_lib.KS_PREC_MIXED_FP32) == _lib.KS_ERR_INVALID_ARG
    assert grad(vk) == _lib.KS_ERR_NO_BANDWIDTH
    vk.hi = vk.hj = ptr(one).value
    assert grad(vk, dscal=ptr(one)) == _lib.KS_ERR_INVALID_ARG
    assert grad(vk, prec=_lib.KS_PREC_MIXED_FP32) == _lib.KS_ERR_INVALID_ARG
    assert grad(vk, k=70) == _lib.KS_ERR_INVALID_ARG
    assert L.ks_last_error() == _lib.KS_ERR_INVALID_ARG
