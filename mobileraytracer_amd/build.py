"""Build recipe of the native libraries (hipcc for gfx950; no cmake, no JIT).

Outputs, in-tree next to this file:
  libmrt_gpu.so   C ABI of include/mobilert_gpu.h: HIP kernels + host runtime
  libmobilert.so  C++ plugin API of include/mobilert/*.hpp on top of it

Every HIP translation unit is compiled with -ffp-contract=off (no fused
multiply-add contraction, so GPU arithmetic rounds like the reference's
statement-by-statement code) and without the packed-FP32 feature.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
ARCH = os.environ.get("MRT_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          f"-I{ROOT / 'include'}", "-Wall", "-Wno-unused-function"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

GPU_DEVICE_SOURCES = ["mrt_kernels.hip", "mrt_wavefront.hip", "mrt_denoise.hip", "mrt_denoise_train.hip",
                      "mrt_bvh_gpu.hip", "mrt_grid_gpu.hip", "mrt_walk_gpu.hip"]
GPU_HOST_SOURCES = ["mrt_api.cpp", "mrt_build.cpp", "mrt_scenes.cpp", "mrt_server.cpp", "mrt_obj.cpp",
                    "mrt_denoise_api.cpp", "mrt_denoise_train_api.cpp"]
HOST_API_SOURCES = ["host/mobilert_api.cpp", "host/c_wrapper.cpp"]

LIB_GPU = HERE / "libmrt_gpu.so"
LIB_HOST = HERE / "libmobilert.so"
API_CHECK_SRC = ROOT / "tests" / "cpp" / "api_check.cpp"
API_CHECK_BIN = ROOT / "tests" / "cpp" / "api_check"


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build")
    return exe


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def build(verbose: bool = False) -> list[Path]:
    """Compile both libraries (always recompiles). Returns the library paths."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    jobs = []
    objs_gpu = []
    for src in GPU_DEVICE_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        jobs.append([hipcc, *COMMON, *DEVICE, "-c", str(CSRC / src), "-o", str(obj)])
        objs_gpu.append(obj)
    for src in GPU_HOST_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        jobs.append([hipcc, *COMMON, "-c", str(CSRC / src), "-o", str(obj)])
        objs_gpu.append(obj)
    objs_host = []
    for src in HOST_API_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
                     f"-I{ROOT / 'include'}", "-c", str(CSRC / src), "-o", str(obj)])
        objs_host.append(obj)
    workers = min(len(jobs), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs_gpu), "-o", str(LIB_GPU),
          "-Wl,--no-undefined"], verbose)
    if not objs_host:
        return [LIB_GPU]
    _run(["g++", "-shared", "-fPIC", *map(str, objs_host), "-o", str(LIB_HOST), f"-L{HERE}",
          "-lmrt_gpu", "-Wl,-rpath,$ORIGIN", "-pthread"], verbose)
    # The C++ API's test program (tests/cpp), built here so that no test
    # compiles anything.
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(API_CHECK_SRC), "-o",
          str(API_CHECK_BIN), f"-L{HERE}", "-lmobilert", "-lmrt_gpu", "-Wl,-rpath,$ORIGIN/../../mobileraytracer_amd", "-pthread"], verbose)
    return [LIB_GPU, LIB_HOST, API_CHECK_BIN]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
