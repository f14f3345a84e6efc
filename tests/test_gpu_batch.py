"""ncvis_fit_transform_many: many independent data sets in one call (one segmented launch per
stage, captured in a hipGraph and replayed), each fitted exactly as a single
fit_transform_device call fits it."""
import numpy as np
import pytest

from oracle import oracle
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


@pytest.mark.parametrize("n_streams", [1, 8])
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


def config5_shapes(count=64, seed=5):
    """The batching workload: N log-uniform in [2,000, 50,000], D in {32, 64, 128}, 5-30 clusters."""
    rng = np.random.default_rng(seed)
    return [(int(np.exp(rng.uniform(np.log(2000), np.log(50000)))), int(rng.choice([32, 64, 128])),
             int(rng.integers(5, 31))) for _ in range(count)]


def _fit_all(torch, h, dX, graph):
    ys = [torch.full((x.shape[0], 2), float("nan"), dtype=torch.float32, device="cuda") for x in dX]
    h.set_option("graph", 1 if graph else 0)
    h.fit_transform_many([(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr()) for x, y in zip(dX, ys)])
    torch.cuda.synchronize()
    return ys


@pytest.mark.timeout(300)
def test_config5_batch_equals_single_fits(torch_cuda):
    # 64 data sets, default (GPU-sized) chain counts, alpha = 0: the SGD epochs run but leave Y
    # unchanged, so Y is deterministic (kNN, symmetrize, init with each set's chain count) and
    # every set's batched Y equals its single fit bit for bit
    torch = torch_cuda
    from ncvis_amd import abi
    shapes = config5_shapes()
    print(f"[case] config5 alpha=0 batch of {len(shapes)} ({sum(n for n, _, _ in shapes)} points)", flush=True)
    sets = _sets(shapes, seed=100)
    h = abi.Handle(d=2, n_threads=0, n_neighbors=15, n_epochs=10, n_init_epochs=20, a=0.77, b=1.22, alpha=0.0)
    dX = [torch.from_numpy(X).cuda() for X in sets]
    ys = _fit_all(torch, h, dX, graph=True)
    hs = abi.Handle(d=2, n_threads=0, n_neighbors=15, n_epochs=10, n_init_epochs=20, a=0.77, b=1.22, alpha=0.0)
    for i, (x, y) in enumerate(zip(dX, ys)):
        ref = torch.empty((x.shape[0], 2), dtype=torch.float32, device="cuda")
        hs.fit_transform_device(x.data_ptr(), x.shape[0], x.shape[1], ref.data_ptr())
        torch.cuda.synchronize()
        assert hs.stats()["chains"] > 1
        assert np.array_equal(ref.cpu().numpy().view(np.uint32), y.cpu().numpy().view(np.uint32)), i


def test_config5_subset_single_chain_vs_single_fits_and_oracle(torch_cuda):
    # the 8 smallest config-5 sets at one chain and one epoch: batched == single fits bit for bit,
    # and each matches the oracle's symmetrize + init + one-thread epoch on the GPU's kNN table
    torch = torch_cuda
    from ncvis_amd import abi
    shapes = sorted(config5_shapes())[:8]
    print(f"[case] config5 subset single chain {shapes}", flush=True)
    sets = _sets(shapes, seed=300)
    a, b = 0.7668736385558027, 1.2225638162909107
    plan = [3]
    kw = dict(d=2, n_threads=1, n_neighbors=15, n_init_epochs=5, a=a, b=b, n_noise=plan)
    h = abi.Handle(**kw)
    h.set_option("exact_threads", 1)
    dX = [torch.from_numpy(X).cuda() for X in sets]
    ys = _fit_all(torch, h, dX, graph=True)
    hs = abi.Handle(**kw)
    hs.set_option("exact_threads", 1)
    st = torch.cuda.current_stream().cuda_stream
    for X, x, y in zip(sets, dX, ys):
        N, k = X.shape[0], 15
        ref = torch.empty((N, 2), dtype=torch.float32, device="cuda")
        hs.fit_transform_device(x.data_ptr(), N, X.shape[1], ref.data_ptr())
        torch.cuda.synchronize()
        got = y.cpu().numpy()
        assert np.array_equal(ref.cpu().numpy().view(np.uint32), got.view(np.uint32))
        inds = torch.empty((N, k), dtype=torch.int32, device="cuda")
        dists = torch.empty((N, k), dtype=torch.float32, device="cuda")
        abi.knn(x.data_ptr(), N, X.shape[1], k, "euclidean", "fp16", inds.data_ptr(), dists.data_ptr(), st)
        torch.cuda.synchronize()
        rp, col, _ = oracle.symmetrize(inds.cpu().numpy().astype(np.int64), dists.cpu().numpy())
        Y0 = oracle.init_embedding(N, 2, rp, col, np.float32(1.0 / k), 5, 42, 1)
        Yo, _ = oracle.optimize(Y0, 0.0, rp, col, plan, a, b, 1.0, 1.0, 42, 1)
        # float data: the oracle itself built with and without -ffast-math (the reference's flags
        # let the compiler reassociate the sample formula) differs by up to 5.9e-5 after this
        # epoch, so the bound is 2e-4 here (1e-4 on the integer-valued data of the other tests)
        np.testing.assert_allclose(got, Yo, atol=2e-4, rtol=0)


def test_graph_replays_match_eager(torch_cuda):
    # the captured batch replayed three times, with eager fits of other shapes in between, gives
    # the eager result every time (one chain per set: deterministic)
    torch = torch_cuda
    from ncvis_amd import abi
    print("[case] graph replays", flush=True)
    sets = _sets([(900, 32, 4), (2500, 64, 7), (1600, 128, 5)], seed=40)
    h = abi.Handle(d=2, n_threads=1, n_neighbors=10, n_epochs=4, n_init_epochs=3, a=0.77, b=1.22)
    h.set_option("exact_threads", 1)
    dX = [torch.from_numpy(X).cuda() for X in sets]
    jobs_y = [torch.empty((x.shape[0], 2), dtype=torch.float32, device="cuda") for x in dX]
    jobs = [(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr()) for x, y in zip(dX, jobs_y)]
    h.set_option("graph", 0)
    h.fit_transform_many(jobs)
    torch.cuda.synchronize()
    eager = [y.cpu().numpy().copy() for y in jobs_y]
    h.set_option("graph", 1)
    other = torch.from_numpy(_sets([(700, 16, 3)], seed=9)[0]).cuda()
    yo = torch.empty((700, 2), dtype=torch.float32, device="cuda")
    for rep in range(3):
        for y in jobs_y:
            y.fill_(float("nan"))
        h.fit_transform_many(jobs)
        torch.cuda.synchronize()
        for e, y in zip(eager, jobs_y):
            assert np.array_equal(e.view(np.uint32), y.cpu().numpy().view(np.uint32)), rep
        h.fit_transform_device(other.data_ptr(), 700, 16, yo.data_ptr())
        torch.cuda.synchronize()


def test_single_fit_graph_replays_match_eager(torch_cuda):
    # a device-pointer fit repeated on the same buffers runs eagerly once, is captured at its
    # second call and replayed from then on: every replay gives the eager result bit for bit (one
    # chain: deterministic), with eager fits of another shape in between, and records its own
    # stage times (the epoch launches' sum is NaN for a replay)
    torch = torch_cuda
    from ncvis_amd import abi
    print("[case] single-fit graph replays", flush=True)
    X = _sets([(1800, 48, 6)], seed=21)[0]
    dX = torch.from_numpy(X).cuda()
    dY = torch.empty((1800, 2), dtype=torch.float32, device="cuda")
    h = abi.Handle(d=2, n_threads=1, n_neighbors=15, n_epochs=3, n_init_epochs=4, a=0.77, b=1.22)
    h.set_option("exact_threads", 1)
    h.fit_transform_device(dX.data_ptr(), 1800, 48, dY.data_ptr())
    torch.cuda.synchronize()
    eager = dY.cpu().numpy().copy()
    s0 = h.stats()["stage_ms"]
    assert all(np.isfinite(v) and v >= 0 for v in s0.values())
    other = torch.from_numpy(_sets([(700, 16, 3)], seed=9)[0]).cuda()
    yo = torch.empty((700, 2), dtype=torch.float32, device="cuda")
    h2 = abi.Handle(d=2, n_threads=1, n_neighbors=15, n_epochs=3, n_init_epochs=4, a=0.77, b=1.22)
    totals = []
    for rep in range(4):
        dY.fill_(float("nan"))
        h.fit_transform_device(dX.data_ptr(), 1800, 48, dY.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(eager.view(np.uint32), dY.cpu().numpy().view(np.uint32)), rep
        s = h.stats()["stage_ms"]
        for name, v in s.items():
            if name == "sgd_kernel":
                assert np.isnan(v), (rep, v)
            else:
                assert np.isfinite(v) and v >= 0, (rep, name, v)
        totals.append(s["total"])
        h2.fit_transform_device(other.data_ptr(), 700, 16, yo.data_ptr())
        torch.cuda.synchronize()
    assert len(set(totals)) > 1, totals  # recorded by each replay, not left from the eager fit
    # option "graph" 0: eager again, every launch timed
    h.set_option("graph", 0)
    h.fit_transform_device(dX.data_ptr(), 1800, 48, dY.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(eager.view(np.uint32), dY.cpu().numpy().view(np.uint32))
    assert np.isfinite(h.stats()["stage_ms"]["sgd_kernel"])
