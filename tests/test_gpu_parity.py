"""GPU parity of every stage of the fit path, through the C ABI, against the CPU oracle and the
committed golden fixtures (made from the reference's own symmetrize)."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle
from ncvis_amd import synthetic

pytestmark = pytest.mark.gpu

A_DEF, B_DEF = 0.7668736385558027, 1.2225638162909107  # find_ab_params(1, 0.4)


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


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _stream(torch):
    return torch.cuda.current_stream().cuda_stream


def gpu_knn(torch, X, k, distance="euclidean", precision="fp32"):
    from ncvis_amd import abi
    N, D = X.shape
    dX = _dev(torch, X.astype(np.float32))
    inds = torch.empty((N, k), dtype=torch.int32, device="cuda")
    dists = torch.empty((N, k), dtype=torch.float32, device="cuda")
    abi.knn(dX.data_ptr(), N, D, k, distance, precision, inds.data_ptr(), dists.data_ptr(),
            _stream(torch))
    torch.cuda.synchronize()
    return inds.cpu().numpy().astype(np.int64), dists.cpu().numpy()


def gpu_symmetrize(torch, inds, dists):
    from ncvis_amd import abi
    N, k = inds.shape
    di = _dev(torch, inds.astype(np.int32))
    dd = _dev(torch, dists.astype(np.float32))
    rp = torch.empty(N + 1, dtype=torch.int64, device="cuda")
    col = torch.empty(2 * N * k, dtype=torch.int32, device="cuda")
    dist = torch.empty(2 * N * k, dtype=torch.float32, device="cuda")
    E = abi.symmetrize(di.data_ptr(), dd.data_ptr(), N, k, rp.data_ptr(), col.data_ptr(),
                       dist.data_ptr(), _stream(torch))
    torch.cuda.synchronize()
    return rp.cpu().numpy(), col[:E].cpu().numpy().astype(np.int64), dist[:E].cpu().numpy()


def gpu_init(torch, N, d, row_ptr, col, alpha, n_init, seed, T):
    from ncvis_amd import abi
    rp = _dev(torch, row_ptr.astype(np.int64))
    cl = _dev(torch, col.astype(np.int32)) if len(col) else torch.zeros(1, dtype=torch.int32,
                                                                          device="cuda")
    Y = torch.empty((N, d), dtype=torch.float32, device="cuda")
    abi.init_embedding(N, d, rp.data_ptr(), cl.data_ptr(), alpha, n_init, seed, T, Y.data_ptr(),
                       _stream(torch))
    torch.cuda.synchronize()
    return Y.cpu().numpy()


def gpu_optimize(torch, Y0, Q, row_ptr, col, plan, a, b, alpha, alpha_Q, seed, T, atomic=0):
    from ncvis_amd import abi
    N, d = Y0.shape
    rp = _dev(torch, row_ptr.astype(np.int64))
    cl = _dev(torch, col.astype(np.int32))
    Y = _dev(torch, Y0.astype(np.float32))
    q, trace = abi.optimize(N, d, rp.data_ptr(), cl.data_ptr(), len(col), Y.data_ptr(), Q, plan, a,
                            b, alpha, alpha_Q, seed, T, atomic, _stream(torch))
    torch.cuda.synchronize()
    return Y.cpu().numpy(), q, trace


# ------------------------------------------------------------------------------ RNG / init
@pytest.mark.parametrize("T", [1, 3, 64])
def test_init_random_layout_bit_exact(torch_cuda, T):
    # n_init_epochs = 0 exposes the raw pcg64 / uniform_real<float> draws under T threads' static
    # partition (ncvis.cpp:195-201)
    _say(f"init random T={T}")
    N, d = 1001, 2
    rp = np.zeros(N + 1, np.int64)
    col = np.zeros(0, np.int64)
    got = gpu_init(torch_cuda, N, d, rp, col, 1.0, 0, 42, T)
    want = oracle.init_embedding(N, d, rp, col, 1.0, 0, 42, T)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _digits_graph(N=600, D=32, k=10, seed=3):
    X, _ = synthetic.digits_like(N=N, D=D, seed=seed)
    inds, dists = oracle.knn_bruteforce(oracle.preprocess(X), k)
    row_ptr, col, _ = oracle.symmetrize(inds, dists)
    return X, inds, dists, row_ptr, col


@pytest.mark.parametrize("n_init", [1, 4, 7])
def test_init_embedding_matches_oracle(torch_cuda, n_init):
    _say(f"init n_init={n_init}")
    X, inds, dists, row_ptr, col = _digits_graph()
    N, k = inds.shape
    got = gpu_init(torch_cuda, N, 2, row_ptr, col, np.float32(1.0 / k), n_init, 42, 1)
    want = oracle.init_embedding(N, 2, row_ptr, col, np.float32(1.0 / k), n_init, 42, 1)
    np.testing.assert_allclose(got, want, atol=1e-5)


# ------------------------------------------------------------------------------ kNN
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("shape", [(1500, 64, 15), (700, 32, 5), (333, 100, 31), (40, 16, 15),
                                   (600, 300, 15), (500, 257, 10), (300, 512, 31), (450, 384, 15),
                                   (700, 64, 50), (400, 128, 40), (200, 32, 63), (300, 300, 45)])
def test_knn_exact_on_integer_data(torch_cuda, precision, shape):
    # integer data: every precision computes exact candidate scores, so the table equals the
    # exact (distance, label) order bit for bit, ties included
    N, D, k = shape
    _say(f"knn int N={N} D={D} k={k} {precision}")
    X, _ = synthetic.digits_like(N=N, D=D, seed=N + D)
    gi, gd = gpu_knn(torch_cuda, X, k, "euclidean", precision)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), k)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_knn_exact_with_duplicate_rows(torch_cuda, precision):
    # 40 distinct integer rows repeated ~75 times each: every query has dozens of neighbours at
    # distance 0, so the table is decided by the label tie-break alone (smaller label first,
    # as the (distance, label) max-heap of the reference keeps them)
    _say(f"knn duplicates {precision}")
    rng = np.random.default_rng(11)
    protos = rng.integers(0, 8, (40, 32)).astype(np.float32)
    X = np.ascontiguousarray(protos[rng.integers(0, 40, 3000)])
    gi, gd = gpu_knn(torch_cuda, X, 15, "euclidean", precision)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od)


@pytest.mark.parametrize("shape", [(300, 600, 15), (300, 300, 45), (200, 20, 70), (5, 3, 4)])
def test_knn_generic_path(torch_cuda, shape):
    # rows over 512 elements, FP32 rows over 256 with k + 1 > 32, k + 1 > 64: the exact scan
    N, D, k = shape
    _say(f"knn generic N={N} D={D} k={k}")
    X, _ = synthetic.digits_like(N=N, D=D, seed=7)
    gi, gd = gpu_knn(torch_cuda, X, k, "euclidean", "fp32")
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), k)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("distance", ["cosine", "correlation", "inner_product"])
@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_knn_inner_product_family(torch_cuda, distance, precision):
    _say(f"knn {distance} {precision}")
    X, _ = synthetic.gaussian_mixture(800, 24, 6, seed=5)
    gi, gd = gpu_knn(torch_cuda, X, 15, distance, precision)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X, distance), 15, distance)
    # exact FP32 rerank: distances agree to rounding; identical neighbour sets almost everywhere
    agree = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi, oi)])
    assert agree > 0.995
    np.testing.assert_allclose(np.sort(gd, 1), np.sort(od, 1), atol=2e-5, rtol=1e-5)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_knn_recall_wide_rows_cosine(torch_cuda, precision):
    # rows of 300 elements (the 100k x 300 cosine configuration's shape) take the one-set-per-wave
    # tiles of the 16-bit GEMM, or the 16x16x4 tiles of the FP32 GEMM
    _say(f"knn wide cosine {precision}")
    X, _ = synthetic.normalized_mixture(N=3000, D=300, n_clusters=12, seed=3)
    gi, gd = gpu_knn(torch_cuda, X, 15, "cosine", precision)
    Xp = oracle.preprocess(X, "cosine")
    oi, od = oracle.knn_bruteforce(Xp, 15, "cosine")
    recall = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi, oi)])
    assert recall >= 0.99, recall
    assert np.all(np.diff(gd, axis=1) <= 0)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_knn_recall_float_data(torch_cuda, precision):
    _say(f"knn recall {precision}")
    X, _ = synthetic.anisotropic_mixture(N=4000, D=128, n_clusters=10, seed=1)
    gi, _ = gpu_knn(torch_cuda, X, 15, "euclidean", precision)
    q = np.arange(0, 4000, 7)
    oi, _ = oracle.knn_fp64(oracle.preprocess(X), 15, "euclidean", queries=q)
    recall = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi[q], oi)])
    assert recall >= 0.99, recall


@pytest.mark.parametrize("transform", ["offset1e3", "offset1e4", "scale1e5", "scale1e-6"])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_knn_recall_any_offset_and_scale(torch_cuda, precision, transform):
    # the candidate GEMM works on centred, power-of-two scaled copies of the rows, so data far
    # from the origin (cancellation in ||c||^2 - 2<q,c>, lost FP16/BF16 resolution), beyond FP16's
    # range (|x| > 65504) or below its normal range keeps the exact search's recall
    _say(f"knn {transform} {precision}")
    X, _ = synthetic.anisotropic_mixture(N=3000, D=128, n_clusters=10, seed=2)
    X = X.astype(np.float64)
    X = {"offset1e3": X + 1e3, "offset1e4": X + 1e4, "scale1e5": X * 1e5, "scale1e-6": X * 1e-6}[transform]
    X = X.astype(np.float32)
    gi, gd = gpu_knn(torch_cuda, X, 15, "euclidean", precision)
    q = np.arange(0, 3000, 5)
    oi, _ = oracle.knn_fp64(X, 15, "euclidean", queries=q)
    recall = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi[q], oi)])
    assert recall >= 0.99, recall
    assert np.all(np.isfinite(gd)) and np.all(np.diff(gd, axis=1) <= 0)


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_knn_inner_product_beyond_fp16_range(torch_cuda, precision):
    # inner_product ranks by <x, y>, which translation would change: the rows are only scaled
    _say(f"knn inner_product x1e5 {precision}")
    X, _ = synthetic.anisotropic_mixture(N=2000, D=64, n_clusters=8, seed=4)
    X = (X.astype(np.float64) * 1e5).astype(np.float32)
    gi, _ = gpu_knn(torch_cuda, X, 15, "inner_product", precision)
    q = np.arange(0, 2000, 5)
    oi, _ = oracle.knn_fp64(X, 15, "inner_product", queries=q)
    recall = np.mean([len(set(a) & set(b)) / 15 for a, b in zip(gi[q], oi)])
    assert recall >= 0.99, recall


# ------------------------------------------------------------------------------ symmetrize
@pytest.mark.parametrize("name", ["sym_digits1500.npz", "sym_ties_hubs.npz", "sym_unsorted.npz",
                                  "sym_tiny.npz"])
def test_symmetrize_bit_exact_vs_reference_fixture(torch_cuda, golden_dir, name):
    _say(f"symmetrize fixture {name}")
    z = np.load(os.path.join(golden_dir, name))
    rp, col, dist = gpu_symmetrize(torch_cuda, z["inds"].astype(np.int64), z["dists"])
    assert np.array_equal(rp, z["row_ptr"])
    assert np.array_equal(col, z["col"].astype(np.int64))
    assert np.array_equal(dist.view(np.uint32), z["dist"].view(np.uint32))


@pytest.mark.parametrize("seed", range(4))
def test_symmetrize_bit_exact_random_tables(torch_cuda, seed):
    _say(f"symmetrize random seed={seed}")
    rng = np.random.default_rng(seed)
    N, k = int(rng.integers(50, 3000)), int(rng.integers(1, 24))
    levels = int(rng.choice([2, 3, 50, 100000]))
    inds = np.stack([rng.choice(np.delete(np.arange(N), i), k, replace=False) for i in range(N)])
    hubs = rng.choice(N, 5, replace=False)
    mask = rng.random((N, k)) < 0.2
    for i, j in zip(*np.nonzero(mask)):
        h = hubs[j % 5]
        if h != i and h not in inds[i]:
            inds[i, j] = h
    dists = (rng.integers(1, levels + 1, (N, k)) / levels).astype(np.float32)
    got = gpu_symmetrize(torch_cuda, inds, dists)
    want = oracle.symmetrize(inds, dists)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_knn_then_symmetrize_digits(torch_cuda):
    _say("knn+symmetrize digits 5620x64")
    X, _ = synthetic.digits_like(seed=0)
    gi, gd = gpu_knn(torch_cuda, X, 15, "euclidean", "fp16")
    got = gpu_symmetrize(torch_cuda, gi, gd)
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    want = oracle.symmetrize(oi, od)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


# ------------------------------------------------------------------------------ SGD epochs
@pytest.mark.parametrize("plan", [[1], [3], [0], [5, 2]])
def test_optimize_single_chain_matches_reference(torch_cuda, plan):
    # n_threads = 1, fixed seed, sequential edge order: one chain == the reference at one thread.
    # One epoch must match within 1e-4 abs; a second epoch compounds last-ulp differences of the
    # transcendental functions (device vs host libm) and is held to 1e-3.
    _say(f"optimize T=1 plan={plan}")
    tol = 1e-4 if len(plan) == 1 else 1e-3
    X, inds, dists, row_ptr, col = _digits_graph()
    N, k = inds.shape
    Y0 = oracle.init_embedding(N, 2, row_ptr, col, np.float32(1.0 / k), 5, 42, 1)
    gy, gq, _ = gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 1.0, 1.0, 42, 1)
    oy, oq = oracle.optimize(Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 1.0, 1.0, 42, 1)
    np.testing.assert_allclose(gy, oy, atol=tol, rtol=0)
    assert abs(gq - oq) < tol * max(1.0, abs(oq))


def test_optimize_single_chain_d1_and_d3(torch_cuda):
    X, inds, dists, row_ptr, col = _digits_graph(N=300)
    N, k = inds.shape
    for d in (1, 3, 5):
        _say(f"optimize T=1 d={d}")
        Y0 = oracle.init_embedding(N, d, row_ptr, col, np.float32(1.0 / k), 3, 42, 1)
        gy, gq, _ = gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, [2], A_DEF, B_DEF, 1.0, 1.0, 7, 1)
        oy, oq = oracle.optimize(Y0, 0.0, row_ptr, col, [2], A_DEF, B_DEF, 1.0, 1.0, 7, 1)
        np.testing.assert_allclose(gy, oy, atol=1e-4, rtol=0)


@pytest.mark.parametrize("T", [3, 64, 4096])
@pytest.mark.parametrize("atomic", [0, 1, 2])
def test_optimize_chains_alpha0_q_trace(torch_cuda, T, atomic):
    # alpha = 0: every step is 0 (ncvis.cpp:287), Y never changes and each chain's Q copy depends
    # only on its static edge chunk, its pcg64(seed + t) stream (kept across epochs) and the
    # epoch's Q, so the chain partition, the seeding and the per-epoch Q average (:342-349) are
    # deterministic. atomic = 1 at d = 2 is the many-chain epoch on device atomics; atomic = 2 the
    # epoch on per-XCD copies of Y, split into sub-epochs with the copies merged after each.
    _say(f"optimize alpha=0 T={T} atomic={atomic}")
    X, inds, dists, row_ptr, col = _digits_graph(N=2000, D=32, k=15)
    N, k = inds.shape
    Y0 = oracle.init_embedding(N, 2, row_ptr, col, np.float32(1.0 / k), 20, 42, 1)
    plan = [3, 1, 5, 2]
    gy, gq, gtrace = gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 0.0, 1.0,
                                  42, T, atomic)
    _, oq, otrace = oracle.optimize_chains(Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 0.0, 1.0, 42, T)
    assert np.array_equal(gy, Y0)
    print(f"T={T} atomic={atomic} gpu trace {gtrace} oracle {otrace}", flush=True)
    np.testing.assert_allclose(gtrace, otrace, rtol=1e-5, atol=1e-5)


def test_optimize_chains_hogwild_q_within_reference_scatter(torch_cuda):
    # alpha = 1, many chains: Hogwild is nondeterministic in the reference too; the GPU's final Q
    # must lie within the scatter of the oracle's T-thread runs over seeds.
    _say("optimize hogwild Q vs oracle seeds")
    X, inds, dists, row_ptr, col = _digits_graph(N=2000, D=32, k=15)
    N, k = inds.shape
    Y0 = oracle.init_embedding(N, 2, row_ptr, col, np.float32(1.0 / k), 20, 42, 1)
    plan = [3] * 10
    oq = np.array([oracle.optimize(Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 1.0, 1.0, s, 8)[1]
                   for s in (42, 43, 44, 45)])
    for atomic, T in ((0, 64), (1, 64), (1, 0), (2, 64), (2, 0)):
        gy, gq, _ = gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 1.0, 1.0,
                                 42, T, atomic)
        print(f"atomic={atomic} T={T}: gpu Q {gq:.4f}; oracle Q over seeds {oq}", flush=True)
        assert np.all(np.isfinite(gy))
        tol = 3 * max(float(oq.std()), 0.01 * abs(float(oq.mean())))
        assert abs(gq - oq.mean()) <= tol, (gq, oq)


@pytest.mark.parametrize("d", [1, 3, 4])
def test_xcd_copies_every_embedding_width(torch_cuda, d):
    # the per-XCD-copy epochs read and write rows as coordinate pairs (even d) or single floats
    # (odd d). alpha = 0: Y stays, the merges keep it bit for bit, and the Q trace is the oracle's;
    # alpha = 1: a finite embedding that moves
    _say(f"xcd copies d={d}")
    X, inds, dists, row_ptr, col = _digits_graph(N=2000, D=32, k=15)
    N, k = inds.shape
    Y0 = oracle.init_embedding(N, d, row_ptr, col, np.float32(1.0 / k), 5, 42, 1)
    plan = [3, 1, 2]
    gy, gq, gtrace = gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 0.0, 1.0, 42, 64, 2)
    _, oq, otrace = oracle.optimize_chains(Y0, 0.0, row_ptr, col, plan, A_DEF, B_DEF, 0.0, 1.0, 42, 64)
    assert np.array_equal(gy, Y0)
    np.testing.assert_allclose(gtrace, otrace, rtol=1e-5, atol=1e-5)
    gy, gq, _ = gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, [3] * 10, A_DEF, B_DEF, 1.0, 1.0, 42, 0, 2)
    assert np.all(np.isfinite(gy)) and np.isfinite(gq)
    assert np.abs(gy - Y0).max() > 0.1 and gy.std() > 0.5 * Y0.std()


def test_sub_epochs_keep_the_single_chain_fit(torch_cuda):
    # the epochs run as sgd_sync consecutive launches that carry every chain's pcg64 state, Q copy
    # and current row over; one exact chain must give the same bits however the epoch is split
    from ncvis_amd import abi
    torch = torch_cuda
    X, _ = synthetic.digits_like(N=700, D=32, seed=4)
    dX = _dev(torch, X)
    out = []
    try:
        for S in (1, 3, 7):
            _say(f"single chain, {S} sub-epochs")
            abi.tune("sgd_sync", S)
            h = abi.Handle(d=2, n_threads=1, n_neighbors=15, n_noise=[2, 3, 1], a=A_DEF, b=B_DEF)
            h.set_option("exact_threads", 1)
            dY = torch.empty((700, 2), dtype=torch.float32, device="cuda")
            h.fit_transform_device(dX.data_ptr(), 700, 32, dY.data_ptr(), _stream(torch))
            torch.cuda.synchronize()
            out.append((dY.cpu().numpy(), h.stats()["Q"]))
    finally:
        abi.tune("sgd_sync", 16)
    for Y, q in out[1:]:
        assert np.array_equal(Y, out[0][0]) and q == out[0][1]


def test_replica_option_validation(torch_cuda):
    from ncvis_amd import abi
    h = abi.Handle(d=2, n_threads=0, n_neighbors=10, n_noise=[1])
    for bad in (-2, 3):
        with pytest.raises(RuntimeError, match="atomic_updates"):
            h.set_option("atomic_updates", bad)
    X, inds, dists, row_ptr, col = _digits_graph(N=300)
    Y0 = np.zeros((300, 5), np.float32)
    with pytest.raises(RuntimeError, match="d <= 4"):
        gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, [1], A_DEF, B_DEF, 1.0, 1.0, 42, 8, 2)
    with pytest.raises(RuntimeError, match="allocates"):
        from ncvis_amd import abi as _a
        rp = _dev(torch_cuda, row_ptr.astype(np.int64))
        cl = _dev(torch_cuda, col.astype(np.int32))
        Y = _dev(torch_cuda, np.zeros((300, 2), np.float32))
        _a.optimize(300, 2, rp.data_ptr(), cl.data_ptr(), len(col), Y.data_ptr(), 0.0, [1], A_DEF, B_DEF, 1.0,
                    1.0, 42, 8, 2, _stream(torch_cuda), noalloc=True)


# ------------------------------------------------------------------------------ NCE loss
@pytest.mark.parametrize("d", [2, 16])
def test_nce_loss_and_adjoint_vs_oracle(torch_cuda, d):
    from ncvis_amd import abi
    torch = torch_cuda
    _say(f"nce loss d={d}")
    X, inds, dists, row_ptr, col = _digits_graph(N=400)
    N = inds.shape[0]
    rng = np.random.default_rng(d)
    Y = rng.normal(size=(N, d)).astype(np.float32) * 3
    E, n = len(col), 5
    neg = rng.integers(0, N, (E, n))
    Q = -2.5
    src, dst = oracle.edges_from_csr(row_ptr, col)
    L, dY, dQ = oracle.nce_loss(Y, Q, src, dst, neg, A_DEF, B_DEF)
    rp = _dev(torch, row_ptr)
    cl = _dev(torch, col.astype(np.int32))
    ng = _dev(torch, neg.astype(np.int32))
    Yd = _dev(torch, Y)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    gQ = torch.zeros(1, dtype=torch.float64, device="cuda")
    gY = torch.zeros((N, d), dtype=torch.float32, device="cuda")
    abi.nce_loss(N, d, rp.data_ptr(), cl.data_ptr(), E, ng.data_ptr(), n, Yd.data_ptr(), Q, A_DEF,
                 B_DEF, loss.data_ptr(), gY.data_ptr(), gQ.data_ptr(), _stream(torch))
    torch.cuda.synchronize()
    assert abs(loss.item() - L) < 1e-5 * abs(L)
    assert abs(gQ.item() - dQ) < 1e-4 * max(1, abs(dQ))
    np.testing.assert_allclose(gY.cpu().numpy(), dY, rtol=1e-3, atol=1e-4 * np.abs(dY).max())


# ------------------------------------------------------------------------------ full pipeline
def test_fit_transform_single_chain_matches_reference(torch_cuda):
    import ncvis_amd
    _say("fit_transform T=1 vs oracle")
    X, _ = synthetic.digits_like(N=800, D=32, seed=9)
    plan = np.array([1, 2], np.uint64)
    vis = ncvis_amd.NCVis(n_threads=1, n_neighbors=15, n_noise=plan, n_init_epochs=20,
                          knn_precision="fp32", random_seed=42, exact_threads=True)
    Y = vis.fit_transform(X)
    Yo, _, _ = oracle.fit_transform(X, d=2, n_threads=1, n_neighbors=15, random_seed=42,
                                 n_init_epochs=20, a=vis.a, b=vis.b, n_noise=plan)
    np.testing.assert_allclose(Y, Yo, atol=1e-4, rtol=0)


def test_fit_transform_handle_device_api(torch_cuda):
    from ncvis_amd import abi
    torch = torch_cuda
    _say("handle device api")
    X, _ = synthetic.digits_like(N=500, D=16, seed=2)
    h = abi.Handle(d=2, n_threads=0, n_neighbors=10, n_noise=[1, 2, 3], a=A_DEF, b=B_DEF)
    dX = _dev(torch, X)
    dY = torch.empty((500, 2), dtype=torch.float32, device="cuda")
    h.fit_transform_device(dX.data_ptr(), 500, 16, dY.data_ptr(), _stream(torch))
    torch.cuda.synchronize()
    s = h.stats()
    assert s["n_edges"] > 500 * 10 and s["chains"] >= 1
    assert np.all(np.isfinite(dY.cpu().numpy()))
    with pytest.raises(RuntimeError, match="at least one element"):
        h.fit_transform_device(dX.data_ptr(), 0, 16, dY.data_ptr())


# ------------------------------------------------------------------------------ tuning knobs
_KNN_DEFAULTS = {"knn_stage_mult": 2, "knn_dma": -1, "knn_split": 0}


@pytest.mark.parametrize("knobs", [{"knn_stage_mult": 1}, {"knn_dma": 0}, {"knn_dma": 0, "knn_stage_mult": 1},
                                   {"knn_split": 1}, {"knn_split": 2}, {"knn_split": 4},
                                   {"knn_split": 4, "knn_dma": 0}])
@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_knn_tuning_knobs_keep_the_table(torch_cuda, knobs, precision):
    # every kernel variant a knob selects gives the exact table on integer data
    from ncvis_amd import abi
    _say(f"knn knobs {knobs} {precision}")
    X, _ = synthetic.digits_like(N=1100, D=64, seed=21)
    k = 15 if precision == "fp16" else 25  # FP32: also the two-list-per-query tiles of k + 1 > 16
    oi, od = oracle.knn_bruteforce(oracle.preprocess(X), k)
    try:
        for key, v in knobs.items():
            abi.tune(key, v)
        gi, gd = gpu_knn(torch_cuda, X, k, "euclidean", precision)
    finally:
        for key, v in _KNN_DEFAULTS.items():
            abi.tune(key, v)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("knobs", [{"sgd_prefetch": 4}, {"sgd_fast_math": 0}, {"sgd_fast_math": 1}, {"sgd_reread": 0},
                                   {"sgd_reread": 1}, {"max_chains": 64},
                                   {"sgd_edge_flush": 1}, {"sgd_sync": 1}, {"sgd_replicas": 0}])
def test_sgd_tuning_knobs_run(torch_cuda, knobs):
    # the many-chain epoch variants the SGD knobs select all descend the same objective
    from ncvis_amd import abi
    _say(f"sgd knobs {knobs}")
    X, inds, dists, row_ptr, col = _digits_graph(N=1500, D=32, k=15, seed=5)
    N = X.shape[0]
    Y0 = oracle.init_embedding(N, 2, row_ptr, col, np.float32(1.0 / 15), 5, 42, 1)
    defaults = {"sgd_prefetch": 1, "sgd_fast_math": 2, "sgd_reread": -1, "max_chains": 1 << 17, "sgd_edge_flush": 0,
                "sgd_sync": 16,
                "sgd_replicas": 1}
    runs = []
    try:
        for key, v in knobs.items():
            abi.tune(key, v)
        for atomic in (1, 2):  # device atomics, per-XCD copies
            runs.append(gpu_optimize(torch_cuda, Y0, 0.0, row_ptr, col, [3] * 20, A_DEF, B_DEF, 1.0, 1.0, 42, 0,
                                     atomic))
    finally:
        for key, v in defaults.items():
            abi.tune(key, v)
    for Y, q, _ in runs:
        assert np.all(np.isfinite(Y)) and np.isfinite(q)
        assert Y.std() > 0.5 * Y0.std()


def test_knn_autotune_cache_round_trip(torch_cuda, tmp_path):
    # the autotuner times the tile variants of a shape, keeps the fastest and writes the JSON
    # cache; the table it selects still gives the exact table on integer data, and the cache
    # loads back into a fresh table
    from ncvis_amd import abi
    import json
    torch = torch_cuda
    _say("knn autotune")
    X, _ = synthetic.digits_like(N=6000, D=64, seed=31)
    dX = _dev(torch, X)
    path = tmp_path / "tiles.json"
    try:
        best = abi.autotune_knn(dX.data_ptr(), 6000, 64, 15, "euclidean", "fp16", cache_path=path)
        assert best["split"] in (1, 2, 4) and best["stage_mult"] in (1, 2) and best["ms"] > 0
        doc = json.loads(path.read_text())
        assert len(doc["entries"]) == 1 and doc["entries"][0]["split"] == best["split"]
        gi, gd = gpu_knn(torch, X, 15, "euclidean", "fp16")
        oi, od = oracle.knn_bruteforce(oracle.preprocess(X), 15)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od)
        synth = abi.autotune_knn(None, 20000, 128, 15, "euclidean", "fp32", cache_path=path)
        assert synth["dma"] in (0, 1)
        assert abi.load_tune_cache(path) == 2
    finally:
        abi.load_tune_cache(None)


def test_knn_autotune_times_the_symmetric_search(torch_cuda, tmp_path):
    # from 2^16 rows the autotuner also times the symmetric threshold search; whichever variant
    # it keeps, the plan reports it and the table equals the list tiles' on integer data
    from ncvis_amd import abi
    import json
    torch = torch_cuda
    _say("knn autotune symmetric")
    N = 70_000
    X, _ = synthetic.digits_like(N=N, D=64, seed=41)
    dX = _dev(torch, X)
    path = tmp_path / "tiles.json"
    try:
        best = abi.autotune_knn(dX.data_ptr(), N, 64, 15, "euclidean", "fp16", cache_path=path)
        assert best["sym"] in (0, 4, 8, 16) and best["ms"] > 0
        assert json.loads(path.read_text())["entries"][0]["sym"] == best["sym"]
        plan = abi.knn_plan(N, 64, 15, "euclidean", "fp16")
        assert plan["sym"] == best["sym"] and plan["split"] == best["split"]
        gi, gd = gpu_knn(torch, X, 15, "euclidean", "fp16")
    finally:
        abi.load_tune_cache(None)
    try:
        abi.tune("knn_sym", 0)
        li, ld = gpu_knn(torch, X, 15, "euclidean", "fp16")
    finally:
        abi.tune("knn_sym", -1)
    assert np.array_equal(gi, li) and np.array_equal(gd, ld)
