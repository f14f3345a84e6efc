"""MI355X-native NCVis fit path (gfx950).

``NCVis`` / ``find_ab_params`` mirror the reference's Python API (``import ncvis_amd as ncvis``).
The native pieces live in ``libncvis_amd.so`` (HIP kernels + C ABI, ``include/ncvis_amd.h``) and
the Cython extension ``_ncvis`` (C++ ``ncvis::NCVis`` class); build them with
``python -m ncvis_amd.build``. There is no CPU fallback: without the built extension, or without a
GPU, the API raises.
"""
from ._runtime import share_torch_hip_runtime as _share_rt

_share_rt()
try:
    from ._ncvis import NCVis, NCVisWrapper, find_ab_params  # noqa: F401
except ImportError as _e:  # pragma: no cover - exercised only when the extension is missing
    _import_error = _e

    def _missing(*_a, **_k):
        raise ImportError("ncvis_amd native extension is not built (python -m ncvis_amd.build): "
                          f"{_import_error}")

    NCVis = NCVisWrapper = find_ab_params = _missing  # type: ignore

__all__ = ["NCVis", "NCVisWrapper", "find_ab_params"]
