"""Build the gfx950 HIP library ``neon_amd/libneon_hip.so`` in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 only, with packed-FP32 instructions
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``), and linked into one shared
library exporting the C ABI of ``include/neon_hip.h``.
"""
from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libneon_hip.so"

SOURCES = ["engine.hip", "newton.hip", "eigen.hip", "capi.hip", "meshgen.cpp", "loads.cpp"]
# the measurement library (include/neon_hip_measure.h): timing probes linked against LIB
MEASURE_SOURCES = ["measure.hip"]
MEASURE_LIB = HERE / "libneon_hip_measure.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fopenmp",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-result",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _compile(src: str) -> pathlib.Path:
    path = CSRC / src
    obj = BUILD / (src + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += ["-x", "hip", *DEVICE]
    cmd += ["-c", str(path), "-o", str(obj)]
    subprocess.run(cmd, check=True)
    return obj


def build(jobs: int = 4) -> pathlib.Path:
    """Compile every source and link ``libneon_hip.so`` and the measurement library
    ``libneon_hip_measure.so`` (always rebuilds)."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, SOURCES + MEASURE_SOURCES))
    product, measure = objs[: len(SOURCES)], objs[len(SOURCES):]
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run(
        [HIPCC, "-shared", "-fopenmp", "--offload-arch=gfx950", *map(str, product), "-o", str(tmp)],
        check=True,
    )
    os.replace(tmp, LIB)
    tmp = MEASURE_LIB.with_suffix(".so.tmp")
    subprocess.run(
        [HIPCC, "-shared", "-fopenmp", "--offload-arch=gfx950", *map(str, measure), f"-L{HERE}", "-lneon_hip",
         "-Wl,-rpath,$ORIGIN", "-o", str(tmp)],
        check=True,
    )
    os.replace(tmp, MEASURE_LIB)
    return LIB


HOST_TEST_SRC = ROOT / "tests" / "cpp" / "host_api_test.cpp"
HOST_TEST = ROOT / "tests" / "cpp" / "host_api_test"


def build_host_test() -> pathlib.Path:
    """Compile the C++ host-class test program against ``libneon_hip.so`` with g++ (host code
    only: it uses the C ABI through ``include/neon_hip.hpp``)."""
    cxx = os.environ.get("CXX", "g++")
    tmp = HOST_TEST.with_suffix(".tmp")
    subprocess.run(
        [cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(HOST_TEST_SRC), f"-L{HERE}", "-lneon_hip",
         "-Wl,-rpath,$ORIGIN/../../neon_amd", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(tmp)],
        check=True,
    )
    os.replace(tmp, HOST_TEST)
    return HOST_TEST


if __name__ == "__main__":
    print(build())
    print(build_host_test())
