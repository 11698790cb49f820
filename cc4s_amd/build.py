"""Builds libcc4s_ccsd.so (HIP for gfx950) in-tree.

Usage: python -m cc4s_amd.build [--jobs N]
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libcc4s_ccsd.so"
BENCH = HERE / "gett_bench"
DRIVER = HERE / "cc4s_ccsd"
DRIVER_SOURCES = ["main.cpp", "yaml.cpp", "tensor_io.cpp"]
SOURCES = ["contract.hip", "mixed.hip", "elementwise.hip", "adjoint.hip", "triples.hip", "ccsd.hip",
           "coupled_cluster.hip", "finite_size.hip", "abi.hip", "noalloc.hip", "ladder.hip", "ladder_pairs.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CC4S_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unknown-pragmas",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _compile(src: str) -> pathlib.Path:
    out = OBJ / (pathlib.Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs: int = 8, verbose: bool = True) -> pathlib.Path:
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}",
           *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    # contraction-kernel benchmark, linked against the library
    cmd = [HIPCC, *COMMON, "-x", "hip", str(CSRC / "gett_bench.cpp"),
           "-o", str(BENCH), f"-L{HERE}", "-lcc4s_ccsd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"benchmark build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {BENCH}")
    # host drop-in driver for cc4s operation files (host C++ over the C ABI)
    drv = CSRC / "driver"
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall",
           f"-I{ROOT / 'include'}", f"-I{drv}",
           *(str(drv / f) for f in DRIVER_SOURCES),
           "-o", str(DRIVER), f"-L{HERE}", "-lcc4s_ccsd", "-Wl,-rpath,$ORIGIN",
           "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {DRIVER}")
    return LIB


if __name__ == "__main__":
    jobs = 8
    if "--jobs" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs)
