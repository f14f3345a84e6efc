"""GPU parity of the symmetric threshold kNN search (16-bit GEMMs, k + 1 <= 16): every pair of
rows is scored once, candidates within either row's score bound go to per-row buckets, and rows
whose bucket overflows take the list tiles. On exactly scored (integer) data the table must be
the oracle's brute-force table bit for bit; on float data it keeps the exact search's recall."""
import sys

import numpy as np
import pytest

from oracle import oracle
from ncvis_amd import synthetic

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ncvis_amd import abi
    abi.lib()
    return torch


def _knn(torch, X, k, distance, precision, sym, chunk=128):
    """GPU table with the symmetric search forced on with row bounds from every sym-th row (0: the
    list tiles); chunk: candidate stages per chunk of the emission pass's work order (0: each
    workgroup sweeps a pair of query blocks)."""
    from ncvis_amd import abi
    N, D = X.shape
    dX = torch.from_numpy(np.ascontiguousarray(X.astype(np.float32))).cuda()
    inds = torch.empty((N, k), dtype=torch.int32, device="cuda")
    dists = torch.empty((N, k), dtype=torch.float32, device="cuda")
    try:
        abi.tune("knn_sym", sym)
        abi.tune("knn_sym_chunk", chunk)
        abi.knn(dX.data_ptr(), N, D, k, distance, precision, inds.data_ptr(), dists.data_ptr(),
                torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
    finally:
        abi.tune("knn_sym", -1)
        abi.tune("knn_sym_chunk", 128)
    return inds.cpu().numpy().astype(np.int64), dists.cpu().numpy()


# D picks each padded width the search takes (32, 64, 128); N = 1100 leaves an odd count of
# 256-row query blocks (the middle block pairs with itself) and padding rows in the last block
@pytest.mark.parametrize("D", [20, 64, 128])
@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("N,stride", [(1100, 2), (5000, 8)])
def test_sym_exact_on_integer_data(torch_cuda, D, precision, N, stride):
    _say(f"sym knn N={N} D={D} {precision} stride={stride}")
    X, _ = synthetic.digits_like(N=N, D=D, seed=31 + D)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    gi, gd = _knn(torch_cuda, X, 15, "euclidean", precision, stride)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


# the emission pass's work order: pairs of query blocks per workgroup (0), chunk-major items with
# chunks of 2 query blocks (4 stages: many items, a partial last chunk) or 3 (6), and the default
@pytest.mark.parametrize("chunk", [0, 4, 6, 128])
@pytest.mark.parametrize("D", [20, 128])
def test_sym_work_orders_give_the_same_table(torch_cuda, chunk, D):
    _say(f"sym knn chunk={chunk} D={D}")
    X, _ = synthetic.digits_like(N=5000, D=D, seed=41 + D)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    gi, gd = _knn(torch_cuda, X, 15, "euclidean", "fp16", 4, chunk)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("k", [1, 5, 10])
def test_sym_exact_short_rows(torch_cuda, k):
    # k + 1 < 16: the buckets still hold each row's 32 smallest pairs; the rerank keeps k + 1
    _say(f"sym knn k={k}")
    X, _ = synthetic.digits_like(N=2300, D=64, seed=7)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), k)
    gi, gd = _knn(torch_cuda, X, k, "euclidean", "fp16", 4)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


def test_sym_inner_product_exact(torch_cuda):
    # beta = -1 and zero score offsets; integer rows keep the GEMM exact
    _say("sym knn inner_product")
    X, _ = synthetic.digits_like(N=1500, D=64, seed=9)
    X = X / 16.0  # exact in FP16 and small enough that 1 - <x, y> stays exact in FP32
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X, "inner_product"), 15, "inner_product")
    gi, gd = _knn(torch_cuda, X, 15, "inner_product", "fp16", 4)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("distance", ["cosine", "correlation"])
def test_sym_cosine_family(torch_cuda, distance):
    _say(f"sym knn {distance}")
    X, _ = synthetic.gaussian_mixture(1800, 24, 6, seed=5)
    gi, gd = _knn(torch_cuda, X, 15, distance, "fp16", 4)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X, distance), 15, distance)
    agree = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi, oi)])
    assert agree > 0.995
    np.testing.assert_allclose(np.sort(gd, 1), np.sort(od, 1), atol=2e-5, rtol=1e-5)


def test_sym_overflow_rows_take_the_list_tiles(torch_cuda):
    # 600 copies of each of 20 integer rows: a copy's bound is its distance to its duplicates, so
    # its bucket (512 entries) overflows, while the 500 other rows' buckets do not; both kinds
    # must give the oracle's table
    _say("sym knn overflow")
    rng = np.random.default_rng(5)
    base = rng.integers(0, 3, size=(20, 16)).astype(np.float32)
    X = np.repeat(base, 600, axis=0)
    X = np.concatenate([X, rng.integers(0, 8, size=(500, 16)).astype(np.float32)])
    X = X[rng.permutation(X.shape[0])]
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    gi, gd = _knn(torch_cuda, X, 15, "euclidean", "fp16", 2)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


def test_sym_all_rows_identical(torch_cuda):
    # every score ties: every bucket overflows and every row goes to the list tiles
    _say("sym knn identical rows")
    X = np.ones((1300, 32), dtype=np.float32)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    gi, gd = _knn(torch_cuda, X, 15, "euclidean", "fp16", 4)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_sym_recall_float_data(torch_cuda, precision):
    _say(f"sym knn recall {precision}")
    X, _ = synthetic.anisotropic_mixture(N=6000, D=128, n_clusters=10, seed=1)
    gi, gd = _knn(torch_cuda, X, 15, "euclidean", precision, 8)
    q = np.arange(0, 6000, 7)
    oi, _ = oracle.knn_fp64(oracle.preprocess(X), 15, "euclidean", queries=q)
    recall = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi[q], oi)])
    assert recall >= 0.99, recall
    assert np.all(np.isfinite(gd)) and np.all(np.diff(gd, axis=1) <= 0)


@pytest.mark.parametrize("transform", ["offset1e4", "scale1e5"])
def test_sym_recall_any_offset_and_scale(torch_cuda, transform):
    _say(f"sym knn {transform}")
    X, _ = synthetic.anisotropic_mixture(N=3000, D=128, n_clusters=10, seed=2)
    X = X.astype(np.float64)
    X = {"offset1e4": X + 1e4, "scale1e5": X * 1e5}[transform].astype(np.float32)
    gi, _ = _knn(torch_cuda, X, 15, "euclidean", "fp16", 4)
    q = np.arange(0, 3000, 5)
    oi, _ = oracle.knn_fp64(X, 15, "euclidean", queries=q)
    recall = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi[q], oi)])
    assert recall >= 0.99, recall


def test_sym_matches_list_tiles_at_scale(torch_cuda):
    # size-independent check at a size the default rule sends to the symmetric search: on integer
    # data both GPU searches are exact, so their tables agree bit for bit
    _say("sym knn vs list tiles N=300000")
    X, _ = synthetic.digits_like(N=300_000, D=64, seed=3)
    gi, gd = _knn(torch_cuda, X, 15, "euclidean", "fp16", 8)
    li, ld = _knn(torch_cuda, X, 15, "euclidean", "fp16", 0)
    assert np.array_equal(gi, li) and np.array_equal(gd, ld)
