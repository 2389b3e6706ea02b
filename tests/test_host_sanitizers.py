"""The host layer under AddressSanitizer and UndefinedBehaviorSanitizer.

`python -m ks_amd.build --sanitize` builds the library with its host layer
(the C ABI, staging, the batched API, the autotuner and the C++ list overloads)
compiled with both sanitizers, linked with the ordinary kernel objects
(ks_amd/_build/sanitize/libgsks.so).  The ABI checks of tests/test_abi.py -- symbol
exports, the ks_t layout, error codes and modes, argument and list checks, the
no-GPU failure path -- then run in a child process that loads that build
(KS_LIBGSKS) with the clang ASan runtime preloaded.  Any report of either
sanitizer aborts the child (halt_on_error, -fno-sanitize-recover)."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT, gpu_available

pytestmark = pytest.mark.skipif(gpu_available(), reason="host-only checks; the GPU runs the "
                                                        "uninstrumented library")


def _sanitized_lib():
    from ks_amd import build as kb
    out = kb.SAN_OUT
    csrc = kb.CSRC
    newest = max(os.path.getmtime(os.path.join(csrc, f)) for f in os.listdir(csrc))
    newest = max(newest, os.path.getmtime(os.path.join(ROOT, "include", "ks.h")),
                 os.path.getmtime(kb.__file__))
    kobj = [os.path.join(kb.OBJ, f + ".o") for f in kb.SOURCES if f.endswith(".hip")]
    if any(os.path.exists(o) and os.path.getmtime(o) > newest for o in kobj):
        newest = max(os.path.getmtime(o) for o in kobj if os.path.exists(o))
    if not os.path.exists(out) or os.path.getmtime(out) < newest:
        kb.build_sanitized()
    return out, kb.sanitizer_runtime()


def _env(lib, runtime):
    pre = os.environ.get("LD_PRELOAD", "")
    return dict(os.environ, KS_LIBGSKS=lib, LD_PRELOAD=runtime + (":" + pre if pre else ""),
                ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:halt_on_error=1",
                UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def test_sanitized_build_is_the_one_loaded():
    lib, runtime = _sanitized_lib()
    code = ("import sys; sys.path.insert(0, %r); from ks_amd import _lib; _lib.lib(); "
            "maps = open('/proc/self/maps').read(); "
            "print('sanitized' if %r in maps else 'plain'); "
            "print('asan' if 'libclang_rt.asan' in maps else 'no-asan')") % (ROOT, lib)
    r = subprocess.run([sys.executable, "-c", code], env=_env(lib, runtime), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.split() == ["sanitized", "asan"]


def test_abi_checks_under_asan_and_ubsan():
    lib, runtime = _sanitized_lib()
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_abi.py"),
                        "-q", "-s", "-p", "no:cacheprovider", "-m", "not gpu"],
                       env=_env(lib, runtime), capture_output=True, text=True, timeout=600,
                       cwd=ROOT)
    print(r.stdout[-3000:])
    # -s: a sanitizer report reaches the child's stderr before it halts
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
