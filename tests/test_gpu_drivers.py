"""The reference-facing ABI of libgsks.so executed from C and C++ programs
(tests/drivers/): dgsks, dgsks_ref, dgsks_wrapper, dgsks_ref_wrapper and
ks_malloc_aligned for all eight kernel types (abi_driver.c), and the five
omp_dgsks_list overloads of include/omp_dgsks_list.hpp (list_driver.cpp).
Each driver checks its results against the CPU oracle it links and exits
non-zero on a mismatch; the bars are stated in the drivers (relative L2 1e-12,
1e-13 for the reference algorithm, and the reference list driver's 1e-10 per
element).  The drivers are built by build() (tests/drivers/Makefile)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "drivers", "bin")


def _run(args, timeout=300):
    exe = os.path.join(BIN, args[0])
    assert os.path.exists(exe), f"{exe} is not built: run build() (tests/drivers/Makefile)"
    print("running", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([exe] + [str(a) for a in args[1:]], capture_output=True, text=True,
                       timeout=timeout)
    print(r.stdout[-6000:])
    if r.stderr:
        print(r.stderr[-2000:])
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,k", [(1000, 1000, 16), (513, 257, 3), (300, 200, 300),
                                   (2100, 1900, 64), (777, 1200, 1)])
def test_abi_driver_single_calls(m, n, k):
    """Shapes of the reference's own runs (1000^2 k16, 513x257 k3, 300x200
    k300), a matrix-core shape and k = 1."""
    r = _run(["abi_driver", m, n, k, 11 + k])
    assert r.returncode == 0
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("k", [3, 8, 16, 32])
def test_list_driver_all_overloads(k):
    r = _run(["list_driver", k, 24, 5 + k])
    assert r.returncode == 0
    assert r.stdout.count("PASS") == (15 if k >= 3 else 10)


def test_drivers_link():
    """Without a GPU: the drivers build from include/ and resolve every symbol
    they use in libgsks.so (the dynamic loader binds them at start-up)."""
    root = os.path.dirname(HERE)
    if not os.path.exists(os.path.join(root, "ks_amd", "libgsks.so")):
        pytest.skip("libgsks.so is not built")
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "drivers", "Makefile")], check=True)
    for exe in ("abi_driver", "list_driver"):
        env = dict(os.environ, LD_BIND_NOW="1")
        r = subprocess.run([os.path.join(BIN, exe)], capture_output=True, text=True, env=env,
                           timeout=60)
        # usage message, exit 2: every symbol resolved, nothing ran
        assert r.returncode == 2, (exe, r.returncode, r.stderr)
        assert "usage" in r.stderr
