"""One HIP runtime per process.

``libncvis_amd.so`` needs ``libamdhip64.so.7``. PyTorch-ROCm wheels bundle their own copy (same
soname, older release, with its own ``libhsa-runtime64``). If both copies end up in one process,
two HSA runtimes compete for the GPU and the one initialised second sees no device
(``hipErrorNoDevice``). Loading torch's copy first makes the dynamic linker resolve the library's
dependency to it by soname, so the library and torch share one runtime (and one device context,
which also lets the library run on torch's streams and memory). Without torch installed the
library loads the ROCm runtime it was linked against.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

_done = False


def share_torch_hip_runtime() -> None:
    global _done
    if _done:
        return
    _done = True
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
