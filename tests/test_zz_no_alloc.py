"""ncvis_optimize_device_noalloc (the epoch loop with its scratch in a static per-device array,
for callers that cannot allocate) against ncvis_optimize_device, through the C ABI: identical
results on two calls in a row, the second on a smaller graph, and the size limits."""
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


def _graph(N, seed):
    X, _ = synthetic.digits_like(N=N, D=16, seed=seed)
    inds, dists = oracle.knn_bruteforce(X, 10)
    row_ptr, col, _ = oracle.symmetrize(inds, dists)
    Y0 = oracle.init_embedding(N, 2, row_ptr, col, np.float32(0.1), 5, 42, 1)
    return row_ptr, col, Y0


def _run(torch, row_ptr, col, Y0, plan, alpha, T, atomic, noalloc):
    from ncvis_amd import abi
    rp = torch.from_numpy(row_ptr.astype(np.int64)).cuda()
    cl = torch.from_numpy(col.astype(np.int32)).cuda()
    Y = torch.from_numpy(Y0.copy()).cuda()
    q, trace = abi.optimize(len(Y0), 2, rp.data_ptr(), cl.data_ptr(), len(col), Y.data_ptr(), 0.0, plan,
                            A_DEF, B_DEF, alpha, 1.0, 42, T, atomic, torch.cuda.current_stream().cuda_stream,
                            noalloc=noalloc)
    torch.cuda.synchronize()
    return Y.cpu().numpy(), q, np.array(trace, np.float32)


@pytest.mark.parametrize("T,atomic,alpha", [(1, 0, 1.0), (64, 1, 0.0), (257, 0, 0.0)])
def test_noalloc_equals_allocating_entry_point(torch_cuda, T, atomic, alpha):
    # single chain (deterministic at any alpha) and many chains at alpha = 0 (deterministic):
    # bit-identical Y, Q and Q trace; the second call runs a smaller graph on the same scratch
    for N, seed in ((1500, 3), (700, 8)):
        _say(f"noalloc N={N} T={T} atomic={atomic} alpha={alpha}")
        row_ptr, col, Y0 = _graph(N, seed)
        plan = [3, 1, 4]
        ya, qa, ta = _run(torch_cuda, row_ptr, col, Y0, plan, alpha, T, atomic, noalloc=False)
        yb, qb, tb = _run(torch_cuda, row_ptr, col, Y0, plan, alpha, T, atomic, noalloc=True)
        assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))
        assert qa == qb and np.array_equal(ta, tb)


def test_noalloc_rejects_sizes_beyond_its_scratch(torch_cuda):
    _say("noalloc limits")
    row_ptr, col, Y0 = _graph(300, 1)
    with pytest.raises(RuntimeError, match="scratch limit"):
        _run(torch_cuda, row_ptr, col, Y0, [1], 1.0, (1 << 17) + 1, 1, noalloc=True)
    with pytest.raises(RuntimeError, match="scratch limit"):
        _run(torch_cuda, row_ptr, col, Y0, [1] * ((1 << 16) + 1), 1.0, 4, 1, noalloc=True)
