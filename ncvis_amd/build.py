"""Builds the native parts of ncvis_amd in-tree.

* ``libncvis_amd.so``: the HIP kernels and the C ABI (``include/ncvis_amd.h``), compiled by hipcc
  for gfx950 only.
* ``_ncvis*.so``: the Cython extension (``_ncvis.pyx``) holding the Python API and the
  ``ncvis::NCVis`` C++ class (``csrc/ncvis.cpp``), linked against ``libncvis_amd.so``.
* ``ncvis_cli``: the command-line harness (``csrc/ncvis_main.cpp``) on the same class.

Usage: ``python -m ncvis_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libncvis_amd.so")

# the kNN tile instantiations are split over several sources so they compile in parallel
KNN_TILE_SOURCES = ["knn_q1_f32dma.hip", "knn_q1_f32.hip", "knn_q1_bf16.hip", "knn_q1_fp16.hip", "knn_t16.hip",
                    "knn_wide.hip", "knn_wide_f32.hip", "knn_sym.hip"]
HIP_SOURCES = KNN_TILE_SOURCES + ["knn.hip", "symmetrize.hip", "embed.hip", "abi.hip", "autotune.hip", "calibrate.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: some of their forms have given wrong results on this part
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]
# per-source extras: the kNN tiles keep MFMA accumulators in VGPRs (no AGPR copies before the
# VALU scoring of every block)
HIP_EXTRA = {src: ["-mllvm", "-amdgpu-mfma-vgpr-form"] for src in KNN_TILE_SOURCES + ["knn.hip"]}


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, *HIP_EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(len(HIP_SOURCES), max(2, os.cpu_count() or 2))) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cython(verbose: bool = False) -> str:
    import numpy as np
    pyx = os.path.join(PKG, "_ncvis.pyx")
    cpp = os.path.join(BUILD, "_ncvis.cpp")
    os.makedirs(BUILD, exist_ok=True)
    _run([sys.executable, "-m", "cython", "--cplus", "-3", "-I", PKG, "-o", cpp, pyx])
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_ncvis" + suffix)
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++14", "-fPIC", "-shared",
          "-I", sysconfig.get_paths()["include"], "-I", np.get_include(), "-I", INCLUDE, "-I", CSRC,
          "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
          cpp, os.path.join(CSRC, "ncvis.cpp"), "-o", tmp,
          "-L", PKG, "-lncvis_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_cli(verbose: bool = False) -> str:
    """``ncvis_cli``: the reference's command-line harness (src/main.cpp) on the C++ class."""
    out = os.path.join(PKG, "ncvis_cli")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++14", "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "ncvis_main.cpp"), os.path.join(CSRC, "ncvis.cpp"), "-o", tmp,
          "-L", PKG, "-lncvis_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_cython(verbose)
    build_cli(verbose)


if __name__ == "__main__":
    build(verbose=True)
