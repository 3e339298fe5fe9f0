"""Build recipe for the native libraries (in-tree, no JIT cache).

- ``xatu_amd/lib/libxatu_bse.so``: the C-ABI (include/xatu_bse.h) with the HIP
  kernels for gfx950, built with hipcc.
- ``xatu_amd/_xatu*.so``: the C++ ExcitonTB/SystemTB/ResultTB facade with its
  pybind11 bindings, built with g++; it reaches the GPU only through the C-ABI.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shlex
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = ROOT / "_build" / "obj"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
    # never emit packed FP32 VALU instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}",
]

DEVICE_SOURCES = [
    "device/fourier.hip",
    "device/assemble.hip",
    "device/matvec.hip",
    "device/vec_kernels.hip",
    "device/hf.hip",
    "device/band_adjoint.hip",
    "device/convolution.hip",
    "device/split_product.hip",
    "device/projected_eigen.hip",
    "device/capi.hip",
    "device/davidson.cpp",
]

HOST_SOURCES = [
    "host/configuration.cpp",
    "host/system.cpp",
    "host/exciton.cpp",
    "host/optics.cpp",
    "host/results.cpp",
    "host/crystal.cpp",
    "host/pybind.cpp",
]


def _run(cmd: list[str], quiet: bool = False) -> None:
    if not quiet:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {cmd[0]} ... {cmd[-1]}")


def _compile_device(src: str) -> Path:
    srcp = CSRC / src
    obj = OBJDIR / (src.replace("/", "_") + ".o")
    lang = ["-x", "hip"] if srcp.suffix == ".cpp" else []
    _run([str(ROCM / "bin" / "hipcc"), *HIPCC_FLAGS, *lang, "-c", str(srcp), "-o", str(obj)], quiet=True)
    return obj


def build_device(jobs: int = 8) -> Path:
    OBJDIR.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_device, DEVICE_SOURCES))
    out = LIBDIR / "libxatu_bse.so"
    tmp = LIBDIR / "libxatu_bse.so.tmp"
    _run([str(ROCM / "bin" / "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs),
          "-o", str(tmp), f"-L{ROCM / 'lib'}", "-lrocsolver", "-lrocblas", "-lrocfft",
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    return out


def _host_compile(src: str, pybind_inc: list[str]) -> Path:
    obj = OBJDIR / (src.replace("/", "_") + ".o")
    _run(["g++", "-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wno-sign-compare",
          f"-I{ROOT / 'include'}", *pybind_inc, "-c", str(CSRC / src), "-o", str(obj)], quiet=True)
    return obj


def build_host(jobs: int = 8) -> Path:
    import pybind11

    OBJDIR.mkdir(parents=True, exist_ok=True)
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _host_compile(s, inc), HOST_SOURCES))
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = PKG / f"_xatu{suffix}"
    tmp = PKG / f"_xatu{suffix}.tmp"
    _run(["g++", "-shared", "-fPIC", "-fopenmp", *map(str, objs), "-o", str(tmp),
          f"-L{LIBDIR}", "-lxatu_bse", "-Wl,-rpath,$ORIGIN/lib"])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_device()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
