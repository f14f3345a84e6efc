"""ncvis_fit_transform_many: many independent data sets in one call, each fitted exactly as a
single fit_transform_device call fits it."""
import numpy as np
import pytest

from ncvis_amd import synthetic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ncvis_amd import abi
    abi.lib()
    return torch


def _sets(shapes, seed=0):
    out = []
    for i, (n, d, c) in enumerate(shapes):
        X, _ = synthetic.gaussian_mixture(n, d, c, seed=seed + i)
        out.append(X.astype(np.float32))
    return out


@pytest.mark.parametrize("n_streams", [1, 3, 8])
def test_many_equals_one_by_one_single_chain(torch_cuda, n_streams):
    torch = torch_cuda
    from ncvis_amd import abi
    print(f"[case] batch single-chain n_streams={n_streams}", flush=True)
    sets = _sets([(500, 32, 4), (1200, 64, 6), (2500, 16, 8), (800, 100, 3), (1500, 48, 5)])
    h = abi.Handle(d=2, n_threads=1, n_neighbors=15, n_epochs=12, n_init_epochs=5, a=0.77, b=1.22)
    h.set_option("exact_threads", 1)
    dX = [torch.from_numpy(X).cuda() for X in sets]
    ref = []
    for x in dX:
        y = torch.empty((x.shape[0], 2), dtype=torch.float32, device="cuda")
        h.fit_transform_device(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr())
        torch.cuda.synchronize()
        ref.append(y.cpu().numpy())
    ys = [torch.full((x.shape[0], 2), float("nan"), dtype=torch.float32, device="cuda") for x in dX]
    h.fit_transform_many([(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr()) for x, y in zip(dX, ys)],
                         n_streams=n_streams)
    torch.cuda.synchronize()
    for r, y in zip(ref, ys):
        assert np.array_equal(r.view(np.uint32), y.cpu().numpy().view(np.uint32))


def test_many_multi_chain_finite(torch_cuda):
    torch = torch_cuda
    from ncvis_amd import abi
    print("[case] batch multi-chain", flush=True)
    sets = _sets([(3000, 64, 10), (6000, 32, 12), (2000, 128, 5), (4000, 64, 8)], seed=10)
    h = abi.Handle(d=2, n_threads=0, n_neighbors=15, n_epochs=30, n_init_epochs=10, a=0.77, b=1.22)
    dX = [torch.from_numpy(X).cuda() for X in sets]
    ys = [torch.full((x.shape[0], 2), float("nan"), dtype=torch.float32, device="cuda") for x in dX]
    h.fit_transform_many([(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr()) for x, y in zip(dX, ys)],
                         n_streams=4)
    torch.cuda.synchronize()
    for y in ys:
        a = y.cpu().numpy()
        assert np.all(np.isfinite(a)) and a.std() > 0.1


def test_many_reports_the_failing_data_set(torch_cuda):
    torch = torch_cuda
    from ncvis_amd import abi
    print("[case] batch error", flush=True)
    h = abi.Handle(d=2, n_threads=1, n_neighbors=5, n_epochs=3, n_init_epochs=1)
    ok = torch.from_numpy(_sets([(300, 8, 3)])[0]).cuda()
    one = torch.zeros((1, 8), dtype=torch.float32, device="cuda")
    y0 = torch.empty((300, 2), dtype=torch.float32, device="cuda")
    y1 = torch.empty((1, 2), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="data set 1"):
        h.fit_transform_many([(ok.data_ptr(), 300, 8, y0.data_ptr()), (one.data_ptr(), 1, 8, y1.data_ptr())],
                             n_streams=2)
    torch.cuda.synchronize()
    assert np.all(np.isfinite(y0.cpu().numpy()))
