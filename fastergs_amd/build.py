"""Builds libfastergs.so (the C-ABI library of hand-written gfx950 kernels)
and _C.so, the PyTorch-ROCm extension of its torch-typed operations.

    python fastergs_amd/build.py

Every HIP translation unit is compiled for gfx950 only, with
  -ffp-contract=off                         no implicit FMA contraction (bit-identity
                                            with the CPU oracle)
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt
  -munsafe-fp-atomics                       hardware float atomics (device memory only)
  -target-feature -packed-fp32-ops          no packed FP32 VALU instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfastergs.so")
SOURCES = ["scan_sort.hip", "preprocess.hip", "render.hip", "backward.hip", "sched.hip", "optim.hip",
           "ssim.hip", "api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


# The torch extension is host code only: g++ against torch's headers (the
# compiler torch was built with), linked to libfastergs.so and torch's HIP
# libraries; __HIP_PLATFORM_AMD__ selects the AMD platform in the HIP
# runtime headers that c10/hip includes.
EXT_SRC = os.path.join(CSRC, "torch_ext.cpp")
EXT = os.path.join(HERE, "_C.so")


def _ext_flags():
    import sysconfig

    import torch
    import torch.utils.cpp_extension as ce
    inc = [*ce.include_paths(), sysconfig.get_paths()["include"], "/opt/rocm/include"]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cflags = ["-O2", "-std=c++17", "-fPIC", *[f"-I{d}" for d in inc], "-D__HIP_PLATFORM_AMD__",
              "-DUSE_ROCM", "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H",
              f"-D_GLIBCXX_USE_CXX11_ABI={abi}"]
    tlib = ce.library_paths()[0]
    ldflags = [f"-L{tlib}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip",
               "-ltorch_python", f"-L{HERE}", "-lfastergs", "-L/opt/rocm/lib", "-lamdhip64",
               "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tlib}"]
    return cflags, ldflags


def _compile_ext(cflags) -> str:
    obj = os.path.join(OBJ, "torch_ext.o")
    r = subprocess.run(["g++", *cflags, "-c", EXT_SRC, "-o", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for torch_ext.cpp:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    cflags, ldflags = _ext_flags()
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES) + 1, os.cpu_count() or 4)) as ex:
        ext_obj = ex.submit(_compile_ext, cflags)
        objs = list(ex.map(_compile, SOURCES))
        ext_obj = ext_obj.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libfastergs.so",
           "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    tmp = EXT + ".tmp"
    r = subprocess.run(["g++", "-shared", "-o", tmp, ext_obj, *ldflags], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"extension link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, EXT)
    if verbose:
        print("built", LIB, "and", EXT)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
