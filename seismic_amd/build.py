"""Builds the HIP engine and the drop-in libraries in-tree (seismic_amd/lib).

    python -m seismic_amd.build            # or __graft_entry__.build()

Outputs
  lib/libseismic_xcorr.so      engine + extended C ABI (include/seismic_xcorr.h)
  lib/correlation_cuda.so      drop-in for Correlation_in_CUDA (include/correlation_cuda.h)
  lib/correlation_multicuda.so drop-in for Correlation_in_MultiCUDA (include/correlation_multicuda.h)
The kernels are compiled for gfx950 only, one object per FFT size so the
seven instantiations (P = 1024 ... 8192, including 3072, 5120 and 6144)
build in parallel.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
# SEISMIC_AMD_BUILD_DIR builds into <dir>/lib and <dir>/obj instead of the tree
_ALT = os.environ.get("SEISMIC_AMD_BUILD_DIR")
LIB = os.path.join(_ALT, "lib") if _ALT else os.path.join(HERE, "lib")
OBJ = os.path.join(_ALT, "obj") if _ALT else os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CXX = os.path.join(ROCM, "llvm", "bin", "clang++")

ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
              f"-I{ROCM}/include", "-Wall", "-Wno-unused-result"]
KERNEL_RADICES = (8, 16, 24, 32, 40, 48, 64)
HEADERS = ("xc_fft.hpp", "xc_kernels.h", "xc_engine.h", "xc_lag.h", "xc_plan.h", "xc_hostio.h")


def _run(cmd):
    t0 = time.time()
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return time.time() - t0


def _stale(out, deps):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = False, verbose: bool = True, jobs: int | None = None) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS] + [
        os.path.join(HERE, "..", "include", h) for h in
        ("seismic_xcorr.h", "correlation_cuda.h", "correlation_multicuda.h", "correlation_c.h")]
    steps = []
    kobjs = []
    for r in KERNEL_RADICES:
        src = os.path.join(CSRC, "xc_kernels.hip")
        out = os.path.join(OBJ, f"xc_kernels_r{r}.o")
        kobjs.append(out)
        if force or _stale(out, [src] + hdrs):
            steps.append([HIPCC] + HIP_FLAGS + [f"-DSXC_R={r}", "-c", src, "-o", out])
    for name in ("xc_lag.hip", "xc_noalloc.hip"):
        src = os.path.join(CSRC, name)
        out = os.path.join(OBJ, name.replace(".hip", ".o"))
        kobjs.append(out)
        if force or _stale(out, [src] + hdrs):
            steps.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", out])
    hobjs = []
    for name in ("xc_engine.cpp", "xc_dispatch.cpp", "xc_serve.cpp", "xc_tiled.cpp", "xc_hostio.cpp"):
        src = os.path.join(CSRC, name)
        out = os.path.join(OBJ, name.replace(".cpp", ".o"))
        hobjs.append(out)
        if force or _stale(out, [src] + hdrs):
            steps.append([CXX] + HOST_FLAGS + ["-c", src, "-o", out])
    jobs = jobs or min(8, max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, steps))
    core = os.path.join(LIB, "libseismic_xcorr.so")
    if force or steps or not os.path.exists(core):
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + kobjs + hobjs +
             ["-o", core + ".tmp"])
        os.replace(core + ".tmp", core)
    for shim, src in (("correlation_cuda.so", "abi_cuda.cpp"),
                      ("correlation_multicuda.so", "abi_multicuda.cpp"),
                      ("correlation_c.so", "abi_omp.cpp")):
        out = os.path.join(LIB, shim)
        s = os.path.join(CSRC, src)
        if force or _stale(out, [s, core] + hdrs):
            _run([CXX] + HOST_FLAGS + ["-shared", s, "-o", out + ".tmp", f"-L{LIB}",
                  "-lseismic_xcorr", f"-L{ROCM}/lib", "-lamdhip64",
                  "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
            os.replace(out + ".tmp", out)
    if verbose:
        print(f"seismic_amd: built {len(steps)} objects -> {LIB}")


if __name__ == "__main__":
    build(force="--force" in sys.argv)
