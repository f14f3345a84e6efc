"""The cosine configuration end to end: a 50-cluster Gaussian mixture (N(0,1) centres,
within-cluster sigma 0.3) with L2-normalised rows of 300 elements, cosine distance, k=15,
n_noise=3. Small size against the oracle stage by stage; full size (100,000 points) through
size-independent properties."""
import sys

import numpy as np
import pytest

from oracle import oracle
from ncvis_amd import synthetic

pytestmark = pytest.mark.gpu

A_DEF, B_DEF = 0.7668736385558027, 1.2225638162909107


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


def _knn(torch, X, k, precision):
    from ncvis_amd import abi
    N, D = X.shape
    dX = torch.from_numpy(X).cuda()
    inds = torch.empty((N, k), dtype=torch.int32, device="cuda")
    dists = torch.empty((N, k), dtype=torch.float32, device="cuda")
    abi.knn(dX.data_ptr(), N, D, k, "cosine", precision, inds.data_ptr(), dists.data_ptr(),
            torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return inds.cpu().numpy().astype(np.int64), dists.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_config2_small_stage_by_stage(torch_cuda, precision):
    from ncvis_amd import abi
    torch = torch_cuda
    _say(f"config2 4000x300 cosine {precision}")
    X, _ = synthetic.normalized_mixture(N=4000, D=300, n_clusters=50, seed=12)
    N, k = X.shape[0], 15
    gi, gd = _knn(torch, X, k, precision)
    Xp = oracle.preprocess(X, "cosine")
    oi, od = oracle.knn_bruteforce(Xp, k, "cosine")
    recall = np.mean([len(set(a) & set(b)) / k for a, b in zip(gi, oi)])
    assert recall >= 0.995, recall
    # exact FP32 rerank: a neighbour found by both has the oracle's distance to rounding
    same = gi == oi
    np.testing.assert_allclose(gd[same], od[same], atol=2e-6, rtol=0)
    # symmetrize of the GPU's table: bit-exact with the oracle's symmetrize of the same table
    di = torch.from_numpy(gi.astype(np.int32)).cuda()
    dd = torch.from_numpy(gd).cuda()
    rp = torch.empty(N + 1, dtype=torch.int64, device="cuda")
    col = torch.empty(2 * N * k, dtype=torch.int32, device="cuda")
    cdist = torch.empty(2 * N * k, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    E = abi.symmetrize(di.data_ptr(), dd.data_ptr(), N, k, rp.data_ptr(), col.data_ptr(), cdist.data_ptr(), st)
    torch.cuda.synchronize()
    orp, ocol, odist = oracle.symmetrize(gi, gd)
    assert np.array_equal(rp.cpu().numpy(), orp)
    assert np.array_equal(col[:E].cpu().numpy().astype(np.int64), ocol)
    # init (T = 1) and one SGD epoch with n_noise = 3 at one chain against the oracle
    Y = torch.empty((N, 2), dtype=torch.float32, device="cuda")
    abi.init_embedding(N, 2, rp.data_ptr(), col.data_ptr(), float(np.float32(1.0 / k)), 20, 42, 1, Y.data_ptr(), st)
    torch.cuda.synchronize()
    Y0 = oracle.init_embedding(N, 2, orp, ocol, np.float32(1.0 / k), 20, 42, 1)
    np.testing.assert_allclose(Y.cpu().numpy(), Y0, atol=1e-5, rtol=0)
    Y.copy_(torch.from_numpy(Y0))
    q, _ = abi.optimize(N, 2, rp.data_ptr(), col.data_ptr(), E, Y.data_ptr(), 0.0, [3], A_DEF, B_DEF, 1.0, 1.0, 42,
                        1, 0, st)
    torch.cuda.synchronize()
    oy, oq = oracle.optimize(Y0, 0.0, orp, ocol, [3], A_DEF, B_DEF, 1.0, 1.0, 42, 1)
    np.testing.assert_allclose(Y.cpu().numpy(), oy, atol=1e-4, rtol=0)
    assert abs(q - oq) < 1e-4 * max(1.0, abs(oq))


@pytest.mark.timeout(300)
def test_config2_full_size(torch_cuda):
    # N = 100,000: recall@15 of both GEMM precisions against float64 exact cosine search on
    # 2,000 queries, and a whole fit (default many-chain mode) that stays finite and separates
    # the clusters: its 15 nearest neighbours in the embedding mostly share the query's cluster
    import ncvis
    torch = torch_cuda
    _say("config2 100000x300 cosine full size")
    X, lab = synthetic.normalized_mixture(N=100_000, D=300, n_clusters=50, seed=0)
    N, k = X.shape[0], 15
    q = np.sort(np.random.default_rng(3).choice(N, 2000, replace=False))
    X64 = torch.from_numpy(X).cuda().double()
    X64 = X64 / X64.norm(dim=1, keepdim=True)
    sim = X64[torch.from_numpy(q).cuda()] @ X64.T
    exact = torch.topk(sim, k + 1, dim=1).indices[:, 1:].cpu().numpy()
    del sim
    for precision in ("fp32", "fp16"):
        gi, gd = _knn(torch, X, k, precision)
        recall = np.mean([len(set(a) & set(b)) / k for a, b in zip(gi[q], exact)])
        print(f"config2 {precision} recall@15 {recall:.5f}", flush=True)
        assert recall >= 0.99, (precision, recall)
        assert np.all(np.diff(gd, axis=1) <= 0)
    vis = ncvis.NCVis(distance="cosine", n_noise=3, knn_precision="fp32")
    Y = vis.fit_transform(X)
    assert np.all(np.isfinite(Y))
    Yt = torch.from_numpy(Y).cuda().double()
    qt = torch.from_numpy(q).cuda()
    d2 = ((Yt[qt][:, None, :] - Yt[None, :, :]) ** 2).sum(-1)
    nn = torch.topk(d2, k + 1, dim=1, largest=False).indices[:, 1:].cpu().numpy()
    agree = float(np.mean(lab[nn] == lab[q][:, None]))
    print(f"config2 fit: {vis.stats()['stage_ms']}, label agreement@15 {agree:.4f}", flush=True)
    assert agree > 0.9, agree
