"""ctypes binding of the C ABI in ``libncvis_amd.so`` (declarations: ``include/ncvis_amd.h``).

Device-pointer stage entry points take integer device addresses (e.g. ``tensor.data_ptr()``) and
a stream handle (``torch.cuda.current_stream().cuda_stream``). Every call raises
``RuntimeError`` with the library's message on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libncvis_amd.so")

# every symbol include/ncvis_amd.h declares
EXPORTS = (
    "ncvis_create", "ncvis_destroy", "ncvis_last_error", "ncvis_set_option", "ncvis_fit_transform",
    "ncvis_fit_transform_device", "ncvis_get_stats", "ncvis_get_sgd_mode", "ncvis_get_last_error", "ncvis_knn_device",
    "ncvis_symmetrize_device", "ncvis_init_embedding_device", "ncvis_optimize_device",
    "ncvis_nce_loss_device", "ncvis_tune", "ncvis_fit_transform_many", "ncvis_optimize_device_noalloc",
    "ncvis_autotune_knn", "ncvis_load_tune_cache", "ncvis_tune_cache_size", "ncvis_autotune_last_error",
    "ncvis_knn_plan",
    "ncvis_calibrate_rmw", "ncvis_calibrate_rmw_copies", "ncvis_calibrate_rmw_eighths", "ncvis_calibrate_last_error",
)

PREC = {"fp32": 0, "bf16": 1, "fp16": 2}
DIST = {"euclidean": 0, "inner_product": 1, "cosine": 2, "correlation": 3}

_lib = None
vp = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (python -m ncvis_amd.build)")
        from ._runtime import share_torch_hip_runtime
        share_torch_hip_runtime()
        L = C.CDLL(LIB_PATH)
        L.ncvis_create.argtypes = [C.c_size_t] * 6 + [C.c_int, C.c_int] + [C.c_float] * 4 + [
            C.POINTER(C.c_size_t), C.c_int, C.POINTER(vp)]
        L.ncvis_destroy.argtypes = [vp]
        L.ncvis_destroy.restype = None
        L.ncvis_last_error.argtypes = [vp]
        L.ncvis_last_error.restype = C.c_char_p
        L.ncvis_get_last_error.argtypes = []
        L.ncvis_get_last_error.restype = C.c_char_p
        L.ncvis_set_option.argtypes = [vp, C.c_char_p, C.c_double]
        L.ncvis_fit_transform.argtypes = [vp, vp, C.c_size_t, C.c_size_t, vp]
        L.ncvis_fit_transform_device.argtypes = [vp, vp, C.c_size_t, C.c_size_t, vp, vp]
        L.ncvis_fit_transform_many.argtypes = [vp, C.c_size_t, C.POINTER(vp), C.POINTER(C.c_size_t),
                                               C.POINTER(C.c_size_t), C.POINTER(vp), C.c_int, vp]
        L.ncvis_get_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_float), C.POINTER(C.c_int64)]
        L.ncvis_get_sgd_mode.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ncvis_knn_device.argtypes = [vp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int, C.c_int, vp,
                                       vp, vp]
        L.ncvis_symmetrize_device.argtypes = [vp, vp, C.c_size_t, C.c_size_t, vp, vp, vp,
                                              C.POINTER(C.c_int64), vp]
        L.ncvis_init_embedding_device.argtypes = [C.c_size_t, C.c_size_t, vp, vp, C.c_float, C.c_int,
                                                  C.c_uint64, C.c_int64, vp, vp]
        L.ncvis_optimize_device.argtypes = [C.c_size_t, C.c_size_t, vp, vp, C.c_int64, vp,
                                            C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_size_t),
                                            C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint64,
                                            C.c_int64, C.c_int, vp, vp]
        L.ncvis_optimize_device_noalloc.argtypes = L.ncvis_optimize_device.argtypes
        L.ncvis_tune.argtypes = [C.c_char_p, C.c_double]
        L.ncvis_autotune_knn.argtypes = [vp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int, C.c_int, C.c_char_p,
                                         C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ncvis_knn_plan.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ncvis_load_tune_cache.argtypes = [C.c_char_p]
        L.ncvis_calibrate_rmw.argtypes = [C.c_size_t, C.c_int64, C.c_int, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
        L.ncvis_calibrate_rmw_copies.argtypes = [C.c_size_t, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.ncvis_calibrate_rmw_eighths.argtypes = [C.c_size_t, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.ncvis_calibrate_last_error.argtypes = []
        L.ncvis_calibrate_last_error.restype = C.c_char_p
        L.ncvis_tune_cache_size.argtypes = []
        L.ncvis_autotune_last_error.argtypes = []
        L.ncvis_autotune_last_error.restype = C.c_char_p
        L.ncvis_nce_loss_device.argtypes = [C.c_size_t, C.c_size_t, vp, vp, C.c_int64, vp, C.c_int,
                                            vp, C.c_float, C.c_float, C.c_float, vp, vp, vp, vp]
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("ncvis_destroy", "ncvis_last_error", "ncvis_get_last_error", "ncvis_autotune_last_error",
                            "ncvis_calibrate_last_error"):
                f.restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib().ncvis_get_last_error().decode())


def _stream(stream) -> int:
    return 0 if stream is None else int(stream)


def autotune_knn(dX, N: int, D: int, k: int, distance: str = "euclidean", precision: str = "fp16",
                 cache_path=None):
    """Times the kNN tile variants for this search shape (dX: device pointer, or None for synthetic
    rows) and keeps the fastest; cache_path: None (memory only), "" ($NCVIS_AMD_TUNE_CACHE) or a
    JSON file. Returns {"split", "stage_mult", "dma", "sym", "ms"}."""
    sp, sm, dm, sy, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
    path = None if cache_path is None else str(cache_path).encode()
    rc = lib().ncvis_autotune_knn(dX or None, N, D, k, DIST[distance], PREC[precision], path, C.byref(sp),
                                  C.byref(sm), C.byref(dm), C.byref(sy), C.byref(ms))
    if rc != 0:
        raise RuntimeError(lib().ncvis_autotune_last_error().decode())
    return {"split": sp.value, "stage_mult": sm.value, "dma": dm.value, "sym": sy.value, "ms": ms.value}


def knn_plan(N: int, D: int, k: int, distance: str = "euclidean", precision: str = "fp16"):
    """The kNN launch variant a search of this shape uses now: {"split", "stage_mult", "dma", "sym"}
    (sym: the symmetric search's bounding-row stride, 0 for the list tiles)."""
    sp, sm, dm, sy = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    if lib().ncvis_knn_plan(N, D, k, DIST[distance], PREC[precision], C.byref(sp), C.byref(sm), C.byref(dm),
                            C.byref(sy)) != 0:
        raise RuntimeError(lib().ncvis_autotune_last_error().decode())
    return {"split": sp.value, "stage_mult": sm.value, "dma": dm.value, "sym": sy.value}


def calibrate_rmw(rows: int, threads: int, iters: int = 512):
    """Random 8-byte-row load and load + 64-bit CAS rates (rows/s) on a table of `rows` entries."""
    lo, ca = C.c_double(), C.c_double()
    if lib().ncvis_calibrate_rmw(rows, threads, iters, C.byref(lo), C.byref(ca)) != 0:
        raise RuntimeError(lib().ncvis_calibrate_last_error().decode())
    return {"load_rows_per_s": lo.value, "load_cas_rows_per_s": ca.value}


def calibrate_rmw_copies(rows: int, threads: int, iters: int = 512) -> float:
    """Rate (rows/s) of the per-XCD-copy update: random row read, read again, plain store, on the
    XCD's own copy of a table of `rows` 8-byte entries."""
    r = C.c_double()
    if lib().ncvis_calibrate_rmw_copies(rows, threads, iters, C.byref(r)) != 0:
        raise RuntimeError(lib().ncvis_calibrate_last_error().decode())
    return r.value


def calibrate_rmw_eighths(rows: int, threads: int, iters: int = 512) -> float:
    """The epochs' update on permuted copies: each XCD's rows drawn from its own eighth of the
    copy, each read once and stored, rows/s."""
    r = C.c_double()
    if lib().ncvis_calibrate_rmw_eighths(rows, threads, iters, C.byref(r)) != 0:
        raise RuntimeError(lib().ncvis_calibrate_last_error().decode())
    return r.value


def load_tune_cache(path=None) -> int:
    """Replaces the tuning table from a JSON file (None: empties it); returns its entry count."""
    if lib().ncvis_load_tune_cache(None if path is None else str(path).encode()) != 0:
        raise RuntimeError(lib().ncvis_autotune_last_error().decode())
    return lib().ncvis_tune_cache_size()


def tune(key: str, value: float) -> None:
    """Set a process-wide tuning knob (ncvis_tune)."""
    _check(lib().ncvis_tune(key.encode(), float(value)))


def knn(dX: int, N: int, D: int, k: int, distance: str, precision: str, d_inds: int, d_dists: int,
        stream=None) -> None:
    _check(lib().ncvis_knn_device(dX, N, D, k, DIST[distance], PREC[precision], d_inds, d_dists,
                                  _stream(stream)))


def symmetrize(d_inds: int, d_dists: int, N: int, k: int, d_row_ptr: int, d_col: int, d_dist: int,
               stream=None) -> int:
    E = C.c_int64(0)
    _check(lib().ncvis_symmetrize_device(d_inds, d_dists, N, k, d_row_ptr, d_col, d_dist,
                                         C.byref(E), _stream(stream)))
    return E.value


def init_embedding(N: int, d: int, d_row_ptr: int, d_col: int, alpha: float, n_init_epochs: int,
                   seed: int, n_threads: int, dY: int, stream=None) -> None:
    _check(lib().ncvis_init_embedding_device(N, d, d_row_ptr, d_col, alpha, n_init_epochs, seed,
                                             n_threads, dY, _stream(stream)))


def optimize(N: int, d: int, d_row_ptr: int, d_col: int, E: int, dY: int, Q: float, n_noise,
             a: float, b: float, alpha: float, alpha_Q: float, seed: int, n_threads: int,
             atomic_updates: int, stream=None, noalloc: bool = False):
    """Runs len(n_noise) epochs; returns (final Q, per-epoch Q list). noalloc: the entry point
    that allocates no device memory (ncvis_optimize_device_noalloc)."""
    n_epochs = len(n_noise)
    arr = (C.c_size_t * max(n_epochs, 1))(*[int(v) for v in n_noise])
    trace = (C.c_float * max(n_epochs, 1))()
    q = C.c_float(Q)
    fn = lib().ncvis_optimize_device_noalloc if noalloc else lib().ncvis_optimize_device
    _check(fn(N, d, d_row_ptr, d_col, E, dY, C.byref(q), n_epochs, arr, a, b, alpha, alpha_Q, seed,
              n_threads, atomic_updates, C.cast(trace, vp), _stream(stream)))
    return q.value, [trace[i] for i in range(n_epochs)]


def nce_loss(N: int, d: int, d_row_ptr: int, d_col: int, E: int, d_neg: int, n: int, dY: int,
             Q: float, a: float, b: float, d_loss: int, d_gradY, d_gradQ, stream=None) -> None:
    _check(lib().ncvis_nce_loss_device(N, d, d_row_ptr, d_col, E, d_neg, n, dY, Q, a, b, d_loss,
                                       d_gradY or None, d_gradQ or None, _stream(stream)))


class Handle:
    """The handle API (ncvis_create / ncvis_fit_transform*)."""

    def __init__(self, d=2, n_threads=0, n_neighbors=15, M=16, ef_construction=200, random_seed=42,
                 n_epochs=50, n_init_epochs=20, a=1.0, b=1.0, alpha=1.0, alpha_Q=1.0, n_noise=None,
                 distance="euclidean"):
        L = lib()
        self._h = vp()
        plan = None
        if n_noise is not None:
            n_epochs = len(n_noise)
            plan = (C.c_size_t * max(n_epochs, 1))(*[int(v) for v in n_noise])
        _check(L.ncvis_create(d, n_threads, n_neighbors, M, ef_construction, random_seed, n_epochs,
                              n_init_epochs, a, b, alpha, alpha_Q, plan, DIST[distance],
                              C.byref(self._h)))
        self.d = d

    def _hcheck(self, rc):
        if rc != 0:
            raise RuntimeError(lib().ncvis_last_error(self._h).decode())

    def set_option(self, key: str, value: float) -> None:
        self._hcheck(lib().ncvis_set_option(self._h, key.encode(), float(value)))

    def fit_transform_device(self, dX: int, N: int, D: int, dY: int, stream=None) -> None:
        self._hcheck(lib().ncvis_fit_transform_device(self._h, dX, N, D, dY, _stream(stream)))

    def fit_transform_many(self, jobs, n_streams: int = 4, stream=None) -> None:
        """jobs: sequence of (dX, N, D, dY) device pointers and sizes, fitted together in shared
        launches (replayed from a hipGraph for repeated calls); n_streams is ignored."""
        n = len(jobs)
        px = (vp * n)(*[j[0] for j in jobs])
        pn = (C.c_size_t * n)(*[j[1] for j in jobs])
        pd = (C.c_size_t * n)(*[j[2] for j in jobs])
        py = (vp * n)(*[j[3] for j in jobs])
        self._hcheck(lib().ncvis_fit_transform_many(self._h, n, px, pn, pd, py, int(n_streams),
                                                    _stream(stream)))

    def fit_transform_host(self, X, Y) -> None:
        self._hcheck(lib().ncvis_fit_transform(self._h, X.ctypes.data, X.shape[0], X.shape[1],
                                               Y.ctypes.data))

    def stats(self):
        ms = (C.c_double * 8)()
        e, c, q = C.c_int64(), C.c_int64(), C.c_float()
        lib().ncvis_get_stats(self._h, ms, C.byref(e), C.byref(q), C.byref(c))
        mode, launches = C.c_int(), C.c_int()
        lib().ncvis_get_sgd_mode(self._h, C.byref(mode), C.byref(launches))
        names = ("knn", "symmetrize", "init", "optimize", "total", "knn_tiles", "knn_rerank",
                 "sgd_kernel")
        return {"stage_ms": {n: ms[i] for i, n in enumerate(names)}, "n_edges": e.value,
                "Q": q.value, "chains": c.value, "sgd_mode": mode.value, "sgd_launches": launches.value}

    def __del__(self):
        # at interpreter shutdown the module globals may already be gone: nothing left to free then
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ncvis_destroy(self._h)
            self._h = None
