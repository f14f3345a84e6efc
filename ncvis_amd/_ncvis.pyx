# cython: language_level=3
"""Python API of the MI355X NCVis fit path.

Same surface as the reference's ``wrapper/ncvis.pyx``: ``find_ab_params`` (:7-24), the Cython
class ``NCVisWrapper`` over the C++ ``ncvis::NCVis`` (:26-46) and ``NCVis`` with identical keyword
arguments, defaults, noise plan and ValueErrors (:48-163). The C++ class runs the fit on the GPU
through ``libncvis_amd.so``.

Differences a user can observe:
  * ``n_threads`` is accepted, but the GPU runs a GPU-sized number of SGD chains (each chain plays
    one of the reference's OpenMP threads) whatever it says. Extra keyword ``exact_threads=True``
    runs exactly ``n_threads`` chains; ``n_threads=1, exact_threads=True`` reproduces the
    reference at one thread.
  * ``M`` / ``ef_construction`` are accepted and ignored: the neighbour search is exact.
  * extra keyword ``knn_precision`` ('fp16' default, 'bf16', 'fp32'): precision of the candidate
    GEMM; final neighbour distances are always exact FP32.
  * C++ errors inside ``fit_transform`` surface as ``RuntimeError``.
"""
from libcpp.string cimport string
cimport numpy as cnp
import numpy as np

from . cimport cncvis

cnp.import_array()

_PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2}


def find_ab_params(spread=1., min_dist=0.1):
    """Fit (a, b) of 1 / (1 + a x^(2b)) to the offset exponential decay defined by ``spread``
    and ``min_dist`` (least squares on 300 points of [0, 3 spread]), as the reference does
    (wrapper/ncvis.pyx:7-24, after UMAP)."""
    from scipy.optimize import curve_fit

    def curve(x, a, b):
        return 1.0 / (1.0 + a * x ** (2 * b))

    xv = np.linspace(0, spread * 3, 300)
    yv = np.where(xv < min_dist, 1.0, np.exp(-(xv - min_dist) / spread))
    params, _ = curve_fit(curve, xv, yv)
    return params[0], params[1]


def negative_plan_for(n_noise, n_epochs):
    """Per-epoch noise-sample counts and the epoch count (wrapper/ncvis.pyx:103-123): None ramps
    linspace(0,1,n_epochs)**3 up to 5 noise samples per edge on average (rounded, at least 1); an
    int >= 1 repeats; a 1-D ndarray is used as given and sets n_epochs."""
    if n_noise is None:
        n_negative = 5
        plan = np.linspace(0, 1, n_epochs) ** 3
        plan /= plan.sum()
        plan *= n_epochs * n_negative
        plan = plan.round().astype(np.uintp)
        plan[plan < 1] = 1
    elif type(n_noise) is np.ndarray:
        if len(n_noise.shape) != 1:
            raise ValueError("n_noise should have exactly one dimension, but shape {} was passed".format(n_noise.shape))
        plan = n_noise.astype(np.uintp)
        n_epochs = plan.size
    elif type(n_noise) is int:
        if n_noise < 1:
            raise ValueError("n_noise should be at least 1, but {} was passed".format(n_noise))
        plan = np.full(n_epochs, n_noise).astype(np.uintp)
    else:
        raise ValueError("n_noise has unsupported type")
    return plan, n_epochs


cdef class NCVisWrapper:
    cdef cncvis.NCVis* c_ncvis
    cdef size_t d

    def __cinit__(self, size_t d, size_t n_threads, size_t n_neighbors, size_t M, size_t ef_construction,
                  size_t random_seed, int n_epochs, int n_init_epochs, float a, float b, float alpha,
                  float alpha_Q, object n_noise, cncvis.Distance distance):
        cdef cnp.uintp_t[:] n_noise_arr
        cdef size_t* plan = NULL
        if isinstance(n_noise, int):
            n_noise_arr = np.full(n_epochs, n_noise, dtype=np.uintp)
        elif isinstance(n_noise, np.ndarray):
            if len(n_noise.shape) > 1:
                raise ValueError("Expected 1D n_noise array.")
            n_epochs = n_noise.shape[0]
            n_noise_arr = np.ascontiguousarray(n_noise.astype(np.uintp))
        else:
            raise ValueError("n_noise has unsupported type")
        if n_epochs > 0:
            plan = <size_t*>&n_noise_arr[0]
        self.c_ncvis = new cncvis.NCVis(d, n_threads, n_neighbors, M, ef_construction, random_seed, n_epochs,
                                        n_init_epochs, a, b, alpha, alpha_Q, plan, distance)
        self.d = d

    def __dealloc__(self):
        del self.c_ncvis

    def set_option(self, str key, double value):
        self.c_ncvis.set_option(key.encode(), value)

    def stats(self):
        cdef double ms[8]
        cdef long long e = 0, c = 0
        cdef float q = 0
        self.c_ncvis.stats(ms, &e, &q, &c)
        names = ("knn", "symmetrize", "init", "optimize", "total", "knn_tiles", "knn_rerank",
                 "sgd_kernel")
        return {"stage_ms": {n: ms[i] for i, n in enumerate(names)}, "n_edges": e, "Q": q, "chains": c}

    def fit_transform(self, float[:, ::1] X, float[:, ::1] Y):
        if X.shape[0] == 0 or X.shape[1] == 0:
            raise RuntimeError("[ncvis::NCVis::fit_transform] Dataset should have at least one element.")
        self.c_ncvis.fit_transform(&X[0, 0], X.shape[0], X.shape[1], &Y[0, 0])


class NCVis:
    def __init__(self, d=2, n_threads=-1, n_neighbors=15, M=8, ef_construction=100, random_seed=42, n_epochs=50,
                 n_init_epochs=20, spread=1., min_dist=0.4, a=None, b=None, alpha=1., alpha_Q=1., n_noise=None,
                 distance="euclidean", knn_precision="fp16", exact_threads=False):
        """
        Creates new NCVis instance (parameters as the reference's wrapper/ncvis.pyx:49-101).

        d : embedding dimensionality. n_threads : the reference's thread count; the GPU picks its
        own chain count unless exact_threads is set (then n_threads chains). n_neighbors :
        neighbours per point. M, ef_construction : accepted for compatibility (exact search).
        random_seed : chain t uses pcg64(random_seed + t). n_epochs / n_init_epochs : optimisation
        and initialisation epochs. spread, min_dist, a, b : kernel 1/(1+a|x-y|^(2b)). alpha,
        alpha_Q : learning rates of the embedding and of the normaliser Q. n_noise : int, 1-D
        ndarray (per-epoch plan) or None (plan ramping up as linspace(0,1,n_epochs)**3). distance :
        'euclidean', 'cosine', 'correlation' or 'inner_product'. knn_precision : 'fp16', 'bf16' or
        'fp32' candidate GEMM. exact_threads : run exactly n_threads SGD chains (chain t is the
        reference's thread t; n_threads=1 is the reference at one thread).
        """
        self.d = d
        negative_plan, n_epochs = negative_plan_for(n_noise, n_epochs)

        if n_threads < 1:
            n_threads = 0

        distances = {
            'euclidean': cncvis.squared_L2,
            'cosine': cncvis.cosine_similarity,
            'correlation': cncvis.correlation,
            'inner_product': cncvis.inner_product,
        }
        if distance not in distances:
            raise ValueError(f"Unsupported distance, expected one of: {'euclidean', 'cosine', 'correlation', 'inner_product'}, but got {distance}")
        if knn_precision not in _PRECISIONS:
            raise ValueError(f"knn_precision must be one of {tuple(_PRECISIONS)}, but got {knn_precision}")

        if (a is None) or (b is None):
            if (a is None) and (b is None):
                a, b = find_ab_params(spread, min_dist)
            else:
                raise ValueError(f'Expected (a, b) to be (float, float) or (None, None), but got (a, b) = ({a}, {b})')

        self.a, self.b = a, b
        self.negative_plan = negative_plan
        self.model = NCVisWrapper(d, n_threads, n_neighbors, M, ef_construction, random_seed, n_epochs, n_init_epochs,
                                  a, b, alpha, alpha_Q, negative_plan, distances[distance])
        self.model.set_option("knn_precision", _PRECISIONS[knn_precision])
        if exact_threads:
            self.model.set_option("exact_threads", 1)

    def fit_transform(self, X):
        """
        Builds an embedding for given points.

        X : ndarray [n_samples, n_high_dimensions], converted to C-contiguous float32.
        Returns Y : float32 ndarray [n_samples, d].
        """
        Y = np.empty((X.shape[0], self.d), dtype=np.float32)
        self.model.fit_transform(np.ascontiguousarray(X, dtype=np.float32), Y)
        return Y

    def stats(self):
        """Stage times (ms), edge count, final Q and chain count of the last fit."""
        return self.model.stats()
