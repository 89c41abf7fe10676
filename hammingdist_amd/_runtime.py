"""One HIP runtime per process.

PyTorch-ROCm wheels ship their own libamdhip64.so, which torch's libraries
link by the unversioned name; the engine links libamdhip64.so.7.  If the
engine's runtime (from /opt/rocm) were loaded first, a later ``import torch``
would map a second copy of the HIP runtime and torch could no longer see the
GPU.  Loading torch's copy first (without importing torch) makes both the
engine and torch resolve to the same runtime whatever the import order.
Set HD_HIP_RUNTIME=system to skip this and use the system runtime.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

_done = False


def ensure_single_runtime() -> None:
    global _done
    if _done:
        return
    _done = True
    if os.environ.get("HD_HIP_RUNTIME", "") == "system":
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    lib = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(lib):
        ctypes.CDLL(lib, mode=ctypes.RTLD_GLOBAL)
