"""CPU oracle for the NCVis fit path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this
module, and only as the checker (or as the timed CPU baseline). The product path in
``ncvis_amd`` never imports it.

ctypes bindings over ``oracle/liboracle.so`` (the C restatement in ``ncvis_oracle.c``; see its
header for the reference file:line each function follows and for how each one is pinned), plus
numpy float64 helpers (exact kNN for recall, closed-form preprocessing), plus the binding of the
reference's own ``src/knntable.cpp`` compiled into ``oracle/_ref`` when that is present.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_knntable.so")

DIST_CODES = {"euclidean": 0, "inner_product": 1, "cosine": 2, "correlation": 3}

_lib = None
_ref = None

_p = np.ctypeslib.ndpointer
_f32 = _p(dtype=np.float32, flags="C_CONTIGUOUS")
_f64 = _p(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = _p(dtype=np.int64, flags="C_CONTIGUOUS")
_u64 = _p(dtype=np.uint64, flags="C_CONTIGUOUS")


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference harness when /root/reference exists)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_pcg64_stream.argtypes = [C.c_uint64, C.c_uint64, _u64]
        L.orc_uniform_index_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, _u64]
        L.orc_uniform_float_stream.argtypes = [C.c_uint64, C.c_uint64, _f32]
        L.orc_static_chunk.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]
        L.orc_preprocess.argtypes = [_f32, C.c_int64, C.c_int64, C.c_int, _f32]
        L.orc_knn_bruteforce.argtypes = [_f32, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                         _i64, _f32]
        L.orc_intro_sort_idx.argtypes = [_i64, C.c_int64, _f32]
        L.orc_symmetrize.argtypes = [_i64, _f32, C.c_int64, C.c_int64, _i64, _i64, _f32]
        L.orc_symmetrize.restype = C.c_int64
        L.orc_init_embedding.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64, _i64, C.c_float,
                                         C.c_int, C.c_uint64, C.c_int, _f32]
        L.orc_sgd_sample.argtypes = [_f32, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_uint64,
                                     C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_sgd_sample.restype = C.c_float
        L.orc_optimize.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64, _i64, _f32,
                                   C.POINTER(C.c_float), C.c_int, _u64, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_uint64, C.c_int, C.c_void_p]
        L.orc_optimize_chains.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64, _i64, _f32,
                                          C.POINTER(C.c_float), C.c_int, _u64, C.c_float, C.c_float,
                                          C.c_float, C.c_float, C.c_uint64, C.c_int64, C.c_void_p]
        L.orc_nce_loss.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64, _i64, _i64, C.c_int64,
                                   _f32, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                   C.c_void_p]
        L.orc_nce_loss.restype = C.c_double
        L.orc_fit_transform.argtypes = [_f32, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int64,
                                        C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float,
                                        C.c_float, C.c_float, _u64, C.c_int, _f32, _f64,
                                        C.POINTER(C.c_float)]
        L.orc_fit_transform.restype = C.c_int
        L.orc_hnsw_knn.argtypes = [_f32, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_uint64, C.c_int, _i64, _f32, _f64]
        L.orc_hnsw_knn.restype = C.c_long
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own KNNTable::symmetrize (src/knntable.cpp), via oracle/_ref."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_symmetrize.argtypes = [_i64, _f32, C.c_long, C.c_long, C.c_int, _i64, _i64, _f32]
        R.ref_symmetrize.restype = C.c_long
        _ref = R
    return _ref


# ----------------------------------------------------------------------------- RNG
def pcg64_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().orc_pcg64_stream(seed, n, out)
    return out


def uniform_index_stream(seed: int, rng: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().orc_uniform_index_stream(seed, rng, n, out)
    return out


def uniform_float_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().orc_uniform_float_stream(seed, n, out)
    return out


def static_chunk(n: int, T: int, t: int):
    s0, e0 = C.c_int64(), C.c_int64()
    lib().orc_static_chunk(n, T, t, C.byref(s0), C.byref(e0))
    return s0.value, e0.value


# ------------------------------------------------------------------------ kNN
def preprocess(X: np.ndarray, distance: str = "euclidean") -> np.ndarray:
    X = np.ascontiguousarray(X, np.float32)
    out = np.empty_like(X)
    lib().orc_preprocess(X, X.shape[0], X.shape[1], DIST_CODES[distance], out)
    return out


def knn_bruteforce(Xp: np.ndarray, k: int, distance: str = "euclidean", n_threads: int = 0):
    """Exact float32 kNN in the reference's row format: [N, k] int64 ids / float32 distances,
    farthest first, with the single smallest (dist, label) pair of the top k+1 dropped."""
    Xp = np.ascontiguousarray(Xp, np.float32)
    N, D = Xp.shape
    inds = np.empty((N, k), np.int64)
    dists = np.empty((N, k), np.float32)
    lib().orc_knn_bruteforce(Xp, N, D, k, DIST_CODES[distance], n_threads, inds, dists)
    return inds, dists


def hnsw_knn(Xp: np.ndarray, k: int, distance: str = "euclidean", M: int = 8, ef_construction: int = 100,
             seed: int = 42, n_threads: int = 0):
    """HNSW kNN in the reference's row format (hnswlib's role in buildKNN/findKNN), plus
    (build_seconds, search_seconds)."""
    Xp = np.ascontiguousarray(Xp, np.float32)
    N, D = Xp.shape
    inds = np.empty((N, k), np.int64)
    dists = np.empty((N, k), np.float32)
    times = np.zeros(2, np.float64)
    short = lib().orc_hnsw_knn(Xp, N, D, DIST_CODES[distance], M, ef_construction, k, seed, n_threads,
                               inds, dists, times)
    return inds, dists, (float(times[0]), float(times[1])), int(short)


def knn_fp64(Xp: np.ndarray, k: int, distance: str = "euclidean", queries=None, block: int = 2048):
    """Exact float64 top-(k+1) by (dist, label), smallest dropped, farthest first.

    Distances are computed as sum((x-y)^2) (euclidean) or 1 - <x, y> in float64 from the
    float32 preprocessed rows. ``queries`` selects a subset of query rows (recall checks)."""
    X64 = np.asarray(Xp, np.float64)
    N = X64.shape[0]
    q_idx = np.arange(N) if queries is None else np.asarray(queries)
    inds = np.empty((len(q_idx), k), np.int64)
    dists = np.empty((len(q_idx), k), np.float64)
    sq = (X64 * X64).sum(1)
    for s in range(0, len(q_idx), block):
        qi = q_idx[s:s + block]
        Qm = X64[qi]
        if distance == "euclidean":
            # exact enough in float64; ties are broken by label below
            Dm = sq[qi][:, None] + sq[None, :] - 2.0 * (Qm @ X64.T)
            Dm = np.maximum(Dm, 0.0)
        else:
            Dm = 1.0 - Qm @ X64.T
        thr = np.partition(Dm, k, axis=1)[:, k]
        for r in range(len(qi)):
            # every entry not above the (k+1)-th smallest value, ordered by (dist, label)
            cand = np.nonzero(Dm[r] <= thr[r])[0]
            sel = cand[np.lexsort((cand, Dm[r][cand]))][:k + 1]
            inds[s + r] = sel[1:][::-1]
            dists[s + r] = Dm[r][sel][1:][::-1]
    return inds, dists


# ---------------------------------------------------------------- symmetrize
def symmetrize(inds: np.ndarray, dists: np.ndarray):
    """Oracle restatement of KNNTable::symmetrize at one thread -> (row_ptr, col, dist)."""
    inds = np.ascontiguousarray(inds, np.int64)
    dists = np.ascontiguousarray(dists, np.float32)
    N, k = inds.shape
    row_ptr = np.empty(N + 1, np.int64)
    col = np.empty(2 * N * k + 1, np.int64)
    dout = np.empty(2 * N * k + 1, np.float32)
    E = lib().orc_symmetrize(inds, dists, N, k, row_ptr, col, dout)
    return row_ptr, col[:E].copy(), dout[:E].copy()


def ref_symmetrize(inds: np.ndarray, dists: np.ndarray, n_threads: int = 1):
    """The reference's own KNNTable::symmetrize (oracle/_ref) -> (row_ptr, col, dist)."""
    inds = np.ascontiguousarray(inds, np.int64)
    dists = np.ascontiguousarray(dists, np.float32)
    N, k = inds.shape
    row_ptr = np.empty(N + 1, np.int64)
    col = np.empty(2 * N * k + 1, np.int64)
    dout = np.empty(2 * N * k + 1, np.float32)
    E = ref().ref_symmetrize(inds, dists, N, k, n_threads, row_ptr, col, dout)
    return row_ptr, col[:E].copy(), dout[:E].copy()


def intro_sort_idx(key: np.ndarray) -> np.ndarray:
    key = np.ascontiguousarray(key, np.float32)
    idx = np.arange(len(key), dtype=np.int64)
    lib().orc_intro_sort_idx(idx, len(key), key)
    return idx


def edges_from_csr(row_ptr: np.ndarray, col: np.ndarray):
    N = len(row_ptr) - 1
    src = np.repeat(np.arange(N, dtype=np.int64), np.diff(row_ptr))
    return src, np.ascontiguousarray(col, np.int64)


# ---------------------------------------------------------------- embedding
def init_embedding(N: int, d: int, row_ptr, col, alpha: float, n_init_epochs: int, seed: int,
                   n_threads: int = 1) -> np.ndarray:
    src, dst = edges_from_csr(row_ptr, col)
    Y = np.empty((N, d), np.float32)
    lib().orc_init_embedding(N, d, len(src), src, dst, alpha, n_init_epochs, seed, n_threads, Y)
    return Y


def optimize(Y: np.ndarray, Q: float, row_ptr, col, n_noise, a: float, b: float, alpha: float,
             alpha_Q: float, seed: int, n_threads: int = 1, trace: bool = False):
    """ncvis.cpp:275-352 on n_threads OpenMP threads. Returns (Y, Q) or, with trace, (Y, Q, Q per
    epoch)."""
    src, dst = edges_from_csr(row_ptr, col)
    Y = np.ascontiguousarray(Y, np.float32).copy()
    N, d = Y.shape
    nn = np.ascontiguousarray(n_noise, np.uint64)
    q = C.c_float(Q)
    qt = np.zeros(max(len(nn), 1), np.float32)
    lib().orc_optimize(N, d, len(src), src, dst, Y, C.byref(q), len(nn), nn, a, b, alpha, alpha_Q,
                       seed, n_threads, qt.ctypes.data)
    return (Y, q.value, qt[:len(nn)]) if trace else (Y, q.value)


def optimize_chains(Y: np.ndarray, Q: float, row_ptr, col, n_noise, a: float, b: float,
                    alpha: float, alpha_Q: float, seed: int, chains: int):
    """The epochs with `chains` chains run one after another inside each epoch (chain t = the
    reference's thread t). Returns (Y, Q, Q per epoch)."""
    src, dst = edges_from_csr(row_ptr, col)
    Y = np.ascontiguousarray(Y, np.float32).copy()
    N, d = Y.shape
    nn = np.ascontiguousarray(n_noise, np.uint64)
    q = C.c_float(Q)
    qt = np.zeros(max(len(nn), 1), np.float32)
    lib().orc_optimize_chains(N, d, len(src), src, dst, Y, C.byref(q), len(nn), nn, a, b, alpha,
                              alpha_Q, seed, chains, qt.ctypes.data)
    return Y, q.value, qt[:len(nn)]


def sgd_sample(Y: np.ndarray, i: int, o: int, positive: bool, cur_noise: int, Q: float,
               step: float, a: float, b: float, alpha_Q: float):
    Y = np.ascontiguousarray(Y, np.float32).copy()
    q = lib().orc_sgd_sample(Y, Y.shape[1], i, o, int(positive), cur_noise, Q, step, a, b,
                             alpha_Q)
    return Y, q


def nce_loss(Y: np.ndarray, Q: float, src, dst, neg, a: float, b: float, grad: bool = True):
    Y = np.ascontiguousarray(Y, np.float32)
    N, d = Y.shape
    src = np.ascontiguousarray(src, np.int64)
    dst = np.ascontiguousarray(dst, np.int64)
    neg = np.ascontiguousarray(neg, np.int64)
    n = neg.shape[1]
    dY = np.zeros((N, d), np.float64) if grad else None
    dQ = C.c_double(0.0)
    L = lib().orc_nce_loss(N, d, len(src), src, dst, neg, n, Y, Q, a, b,
                           dY.ctypes.data if grad else None,
                           C.cast(C.pointer(dQ), C.c_void_p) if grad else None)
    return (L, dY, dQ.value) if grad else L


def fit_transform(X: np.ndarray, d: int = 2, n_threads: int = 1, n_neighbors: int = 15,
                  random_seed: int = 42, n_init_epochs: int = 20, a: float = 1.0, b: float = 1.0,
                  alpha: float = 1.0, alpha_Q: float = 1.0, n_noise=None, n_epochs: int = 50,
                  distance: str = "euclidean"):
    """Full CPU pipeline (exact brute-force kNN in place of HNSW). Returns (Y, stage_times, Q)."""
    X = np.ascontiguousarray(X, np.float32)
    N, D = X.shape
    if n_noise is None:
        n_noise = np.full(n_epochs, 3, np.uint64)
    nn = np.ascontiguousarray(n_noise, np.uint64)
    Y = np.empty((N, d), np.float32)
    times = np.zeros(6, np.float64)
    q = C.c_float(0.0)
    rc = lib().orc_fit_transform(X, N, D, d, n_threads, n_neighbors, random_seed, len(nn),
                                 n_init_epochs, a, b, alpha, alpha_Q, nn, DIST_CODES[distance], Y,
                                 times, C.byref(q))
    if rc != 0:
        raise RuntimeError("oracle fit_transform rejected its input")
    return Y, times, q.value
