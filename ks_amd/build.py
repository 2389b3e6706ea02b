"""Builds ks_amd/libgsks.so (HIP kernels + C-ABI) for gfx950 with hipcc, in tree.

    python -m ks_amd.build [--force]
    python -m ks_amd.build --sanitize     # ks_amd/_build/sanitize/libgsks.so

The --sanitize build is the same library with its host layer (the .cpp
sources: the C ABI, staging, the batched API, the autotuner, the C++ list
overloads) compiled with AddressSanitizer and UndefinedBehaviorSanitizer, linked
with the kernel objects of the ordinary build (whose host code is launch code
that only runs where a GPU is).  It needs the clang ASan runtime preloaded into
the process that loads it (sanitizer_runtime()) and serves the CPU checks of the
host layer (argument checks, error paths, the no-GPU failure path).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgsks.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SAN_DIR = os.path.join(OBJ, "sanitize")
SAN_OUT = os.path.join(SAN_DIR, "libgsks.so")
# host pass only, and every sanitizer flag bound to it
SAN_FLAGS = ["--offload-host-only", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
             "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host",
             "-fno-omit-frame-pointer", "-g", "-shared-libsan"]
SANITIZER_RUNTIME = "libclang_rt.asan-x86_64.so"

SOURCES = ["fused_f64.hip", "fused_f64_sym.hip", "fused_f64_mfma.hip", "mixed.hip", "multirhs.hip", "multirhs_mixed.hip", "grad.hip", "ref_f64.hip", "norms.hip", "list.hip", "sgsks.hip", "noalloc.hip",
           "api.cpp", "staging.cpp", "list_api.cpp", "list_tune.cpp", "omp_dgsks_list.cpp"]

# Packed FP32 instructions are disabled for all device code: some of their
# forms have given wrong results on gfx950 without any error.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-result", "-Wno-unused-function"]


def _compile(src: str, sanitize: bool = False) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(SAN_DIR if sanitize else OBJ, src + ".o")
    cmd = [HIPCC] + COMMON + DEVICE_FLAGS + (SAN_FLAGS if sanitize else []) + \
          ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(force: bool = True, jobs: int | None = None) -> str:
    """Compile every source and link libgsks.so.  Always recompiles when force."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not force and os.path.exists(OUT):
        newest = max(os.path.getmtime(os.path.join(CSRC, f)) for f in os.listdir(CSRC))
        newest = max(newest, os.path.getmtime(os.path.join(HERE, "..", "include", "ks.h")))
        if os.path.getmtime(OUT) >= newest:
            return OUT
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + \
          ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def sanitizer_runtime() -> str:
    """Path of the clang ASan runtime the sanitized build links against."""
    r = subprocess.run([HIPCC, "-print-file-name=" + SANITIZER_RUNTIME], capture_output=True,
                       text=True)
    path = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        raise RuntimeError(f"{SANITIZER_RUNTIME} not found next to {HIPCC}")
    return path


def build_sanitized(jobs: int | None = None) -> str:
    """Compile the host layer with ASan + UBSan and link it with the ordinary
    kernel objects into ks_amd/_build/sanitize/libgsks.so (always recompiles
    the host layer; builds the kernel objects first if they are missing)."""
    os.makedirs(SAN_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    host = [s for s in srcs if s.endswith(".cpp")]
    kern = [s for s in srcs if s.endswith(".hip")]
    if not all(os.path.exists(os.path.join(OBJ, s + ".o")) for s in kern):
        build(force=True, jobs=jobs)
    jobs = jobs or min(8, len(host))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda f: _compile(f, True), host))
    objs += [os.path.join(OBJ, s + ".o") for s in kern]
    tmp = SAN_OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + SAN_FLAGS + objs + \
          ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitized link failed:\n{r.stderr}")
    os.replace(tmp, SAN_OUT)
    return SAN_OUT


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized())
    else:
        print(build(force="--force" in sys.argv or True))
