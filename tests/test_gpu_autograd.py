"""NCE loss forward + hand-written adjoint behind torch.autograd.Function: checked against float64
central finite differences, the oracle's analytic float64 gradient, and the reference's in-loop
per-sample update."""
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
    return torch


def _graph(N=1000, D=64, k=15, seed=4):
    X, _ = synthetic.gaussian_mixture(N, D, 20, seed=seed)
    inds, dists = oracle.knn_bruteforce(oracle.preprocess(X), k)
    return oracle.symmetrize(inds, dists)


@pytest.mark.parametrize("d", [2, 16])
def test_autograd_matches_fd_and_oracle(torch_cuda, d):
    torch = torch_cuda
    from ncvis_amd.autograd import nce_loss
    _say(f"autograd d={d}")
    row_ptr, col, _ = _graph()
    N, E, n = len(row_ptr) - 1, len(col), 5
    rng = np.random.default_rng(d)
    Y = (rng.normal(size=(N, d)) * 2).astype(np.float32)
    neg = rng.integers(0, N, (E, n))
    Q = -3.0
    Yt = torch.tensor(Y, device="cuda", requires_grad=True)
    Qt = torch.tensor(Q, dtype=torch.float32, device="cuda", requires_grad=True)
    rp = torch.tensor(row_ptr, device="cuda")
    cl = torch.tensor(col.astype(np.int32), device="cuda")
    ng = torch.tensor(neg.astype(np.int32), device="cuda")
    loss = nce_loss(Yt, Qt, rp, cl, ng, A_DEF, B_DEF)
    loss.backward()
    src, dst = oracle.edges_from_csr(row_ptr, col)
    L, dY, dQ = oracle.nce_loss(Y, Q, src, dst, neg, A_DEF, B_DEF)
    assert abs(loss.item() - L) <= 1e-5 * abs(L)
    gY = Yt.grad.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(gY, dY, rtol=2e-3, atol=2e-4 * np.abs(dY).max())
    assert abs(Qt.grad.item() - dQ) <= 1e-4 * max(1.0, abs(dQ))
    # float64 central differences of the loss on a few coordinates
    eps = 1e-3
    for r, c in [(0, 0), (17, d - 1), (500, 1 % d)]:
        Yp = Y.astype(np.float64).copy(); Yp[r, c] += eps
        Ym = Y.astype(np.float64).copy(); Ym[r, c] -= eps
        fd = (oracle.nce_loss(Yp.astype(np.float32), Q, src, dst, neg, A_DEF, B_DEF, grad=False)
              - oracle.nce_loss(Ym.astype(np.float32), Q, src, dst, neg, A_DEF, B_DEF, grad=False)) / (2 * eps)
        assert abs(gY[r, c] - fd) <= 5e-3 * max(1.0, abs(fd)), (r, c, gY[r, c], fd)


def test_adjoint_equals_in_loop_update(torch_cuda):
    """One edge (i, j) with one negative m: the reference's per-sample updates, each taken from
    the same Y, equal -step * (the adjoint of that sample); their sum is the adjoint of the
    two-sample loss. Q's update is -alpha_Q * dL/dQ."""
    torch = torch_cuda
    from ncvis_amd.autograd import nce_loss
    _say("adjoint vs in-loop update")
    rng = np.random.default_rng(3)
    N, d = 4, 2
    Y = rng.normal(size=(N, d)).astype(np.float32)
    i, j, m = 0, 2, 3
    row_ptr = np.array([0, 1, 1, 1, 1], np.int64)
    col = np.array([j], np.int64)
    neg = np.array([[m]], np.int64)
    Q, step = -0.4, 1e-3
    Yp, Qp = oracle.sgd_sample(Y, i, j, True, 1, Q, step, A_DEF, B_DEF, 1.0)
    Yn, Qn = oracle.sgd_sample(Y, i, m, False, 1, Q, step, A_DEF, B_DEF, 1.0)
    upd = ((Yp.astype(np.float64) - Y) + (Yn.astype(np.float64) - Y)) / (-step)
    dq = (Q - Qp) + (Q - Qn)
    Yt = torch.tensor(Y, device="cuda", requires_grad=True)
    Qt = torch.tensor(Q, dtype=torch.float32, device="cuda", requires_grad=True)
    loss = nce_loss(Yt, Qt, torch.tensor(row_ptr, device="cuda"),
                    torch.tensor(col.astype(np.int32), device="cuda"),
                    torch.tensor(neg.astype(np.int32), device="cuda"), A_DEF, B_DEF)
    loss.backward()
    np.testing.assert_allclose(Yt.grad.cpu().numpy(), upd, rtol=2e-3, atol=1e-5)
    assert abs(Qt.grad.item() - dq) < 1e-4


def test_adjoint_large_graph_runs(torch_cuda):
    """200,000-point, 20-cluster D=64 graph, d in {2, 16}: the adjoint completes and is finite
    (the size the differentiation workload names; numerics are checked on the subgraphs above)."""
    torch = torch_cuda
    from ncvis_amd import abi
    from ncvis_amd.autograd import nce_loss
    _say("adjoint 200k")
    X, _ = synthetic.gaussian_mixture(200_000, 64, 20, seed=8)
    N, D, k = X.shape[0], X.shape[1], 15
    dX = torch.tensor(X, device="cuda")
    inds = torch.empty((N, k), dtype=torch.int32, device="cuda")
    dists = torch.empty((N, k), dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    abi.knn(dX.data_ptr(), N, D, k, "euclidean", "fp16", inds.data_ptr(), dists.data_ptr(), st)
    rp = torch.empty(N + 1, dtype=torch.int64, device="cuda")
    col = torch.empty(2 * N * k, dtype=torch.int32, device="cuda")
    cd = torch.empty(2 * N * k, dtype=torch.float32, device="cuda")
    E = abi.symmetrize(inds.data_ptr(), dists.data_ptr(), N, k, rp.data_ptr(), col.data_ptr(),
                       cd.data_ptr(), st)
    col = col[:E].contiguous()
    ng = torch.randint(0, N, (E, 5), dtype=torch.int32, device="cuda")
    for d in (2, 16):
        Yt = torch.randn((N, d), device="cuda", requires_grad=True)
        loss = nce_loss(Yt, -5.0, rp, col, ng, A_DEF, B_DEF)
        loss.backward()
        assert torch.isfinite(loss) and torch.isfinite(Yt.grad).all()
