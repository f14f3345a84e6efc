"""Benchmark of the NCVis fit path on one MI355X.

Headline (default): ``fit_transform`` of N=1,000,000 x D=128 float32 points (100-cluster
anisotropic Gaussian mixture, N(0,1) centres, per-cluster diagonal scales log-uniform in
[0.1, 1]), euclidean, d=2, k=15, 50 epochs with the default noise plan, 20 init epochs; kNN
candidate GEMM in FP16 (MFMA) with exact FP32 rerank. X is resident in HBM when the timed region
starts; a "step" is one whole fit_transform (kNN -> symmetrize -> init -> SGD epochs).

Also reported: the same fit with the FP32 candidate GEMM, recall@15 of both against float64
exact search on 10,000 queries, per-stage times, the roofline of the dominant kernel, the
host-buffer API time (X and Y over PCIe), and the CPU baseline: the oracle's C/OpenMP
restatement on this host's cores with an HNSW kNN stage at the reference's Python defaults (its
recall@15 on the same queries is reported too), init and SGD timed on a bounded sample and scaled
to the whole fit.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1m|digits|100k|batch64]

`--workload batch64` times 64 small data sets through ncvis_fit_transform_many against the same
fits one call at a time (a separate line; the headline is the 1m workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp16": 2500.0, "bf16": 2500.0, "fp32": 157.3}

WORKLOADS = {
    "1m": dict(desc="anisotropic Gaussian mixture N=1,000,000 D=128, 100 clusters", N=1_000_000,
               D=128, clusters=100, distance="euclidean"),
    "100k": dict(desc="anisotropic Gaussian mixture N=100,000 D=128, 100 clusters", N=100_000,
                 D=128, clusters=100, distance="euclidean"),
    "digits": dict(desc="digits-shaped integer data N=5,620 D=64, 10 clusters", N=5620, D=64,
                   clusters=10, distance="euclidean"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_data(w, seed=0):
    from ncvis_amd import synthetic
    if w["N"] == 5620:
        X, _ = synthetic.digits_like(seed=seed)
    else:
        X, _ = synthetic.anisotropic_mixture(N=w["N"], D=w["D"], n_clusters=w["clusters"], seed=seed)
    return X


def measured_traffic(workload_desc, precision):
    """HBM bytes per launch of the two dominant kernels, from the calibrated FETCH_SIZE/WRITE_SIZE
    counter runs committed for the 1M x 128 workload; {} for other workloads or precisions."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_fit_1m_d128.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return {}
    if t.get("workload") != workload_desc or precision != "fp16":
        return {}
    return {name: float(v["traffic_bytes"]) for name, v in t["kernels"].items()}


def recall_at_k(torch, dX, inds, k, queries):
    """recall@k of the GPU table against float64 exact search (same self-drop convention)."""
    X64 = dX.double()
    xn = (X64 * X64).sum(1)
    hits = 0
    for s in range(0, len(queries), 256):
        q = torch.as_tensor(queries[s:s + 256], device=dX.device)
        Dm = xn[q][:, None] + xn[None, :] - 2.0 * (X64[q] @ X64.T)
        top = torch.topk(Dm, k + 1, dim=1, largest=False).indices[:, 1:]
        got = inds[q].long()
        hits += (got[:, :, None] == top[:, None, :]).any(2).sum().item()
        del Dm
    return hits / (len(queries) * k)


def cpu_baseline(X, plan, a, b, k, n_init_epochs, seed, distance="euclidean"):
    """The oracle (C/OpenMP restatement of the reference) timed on the host cores, stage by stage
    as fit_transform runs them: HNSW kNN (M=8, ef_construction=100, search ef=max(10, k+1), the
    Python defaults) over the whole data set, symmetrize of that table, 2 of the init epochs and
    one SGD epoch at the plan's mean noise, the last two scaled to the whole fit."""
    from oracle import oracle
    threads = len(os.sched_getaffinity(0))
    N = X.shape[0]
    parts = {}
    t0 = time.perf_counter()
    h_inds, h_dists, (t_build, t_search), _ = oracle.hnsw_knn(X, k, distance, 8, 100, seed, threads)
    parts["knn"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    row_ptr, col, _ = oracle.symmetrize(h_inds, h_dists)
    parts["symmetrize"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    Y = oracle.init_embedding(N, 2, row_ptr, col, np.float32(1.0 / k), 2, seed, threads)
    parts["init"] = (time.perf_counter() - t0) * max(n_init_epochs, 1) / 2
    mean_noise = int(round(float(np.mean(plan))))
    t0 = time.perf_counter()
    oracle.optimize(Y, 0.0, row_ptr, col, [mean_noise], a, b, 1.0, 1.0, seed, threads)
    t_ep = time.perf_counter() - t0
    parts["optimize"] = t_ep * float(np.sum(1 + plan)) / (1 + mean_noise)
    total = sum(parts.values())
    sample = (f"HNSW kNN (M=8, ef_construction=100) over all {N} rows ({t_build:.1f} s build, "
              f"{t_search:.1f} s search); symmetrize of that table; 2 of {n_init_epochs} init epochs "
              f"scaled x{max(n_init_epochs, 1) / 2:.0f}; 1 SGD epoch at n_noise={mean_noise} scaled to "
              f"the plan's {int(np.sum(1 + plan))} sample passes")
    out = {"value": round(total, 2), "unit": "s", "cores": threads, "kind": "port", "sample": sample,
           "stage_s": {k_: round(v, 2) for k_, v in parts.items()}, "n_edges": int(row_ptr[-1])}
    return out, h_inds


def run_batch(args):
    """64 independent data sets (N log-uniform in [2,000, 50,000], D in {32, 64, 128}, 5-30
    clusters each): one ncvis_fit_transform_many call vs the same fits one by one."""
    import torch
    from ncvis_amd import abi, synthetic
    from ncvis_amd._ncvis import find_ab_params, negative_plan_for

    rng = np.random.default_rng(5)
    shapes = [(int(np.exp(rng.uniform(np.log(2000), np.log(50000)))), int(rng.choice([32, 64, 128])),
               int(rng.integers(5, 31))) for _ in range(64)]
    sets = [synthetic.gaussian_mixture(n, dd, c, seed=100 + i)[0] for i, (n, dd, c) in enumerate(shapes)]
    a, b = find_ab_params(1.0, 0.4)
    plan, n_epochs = negative_plan_for(None, 50)
    h = abi.Handle(d=2, n_threads=0, n_neighbors=15, random_seed=42, n_init_epochs=20, a=a, b=b,
                   n_noise=list(plan), distance="euclidean")
    h.set_option("knn_precision", abi.PREC[args.precision])
    dX = [torch.from_numpy(X).cuda() for X in sets]
    dY = [torch.empty((x.shape[0], 2), dtype=torch.float32, device="cuda") for x in dX]
    jobs = [(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr()) for x, y in zip(dX, dY)]

    def one_by_one():
        for x, y in zip(dX, dY):
            h.fit_transform_device(x.data_ptr(), x.shape[0], x.shape[1], y.data_ptr())

    def batched():
        h.fit_transform_many(jobs, n_streams=args.streams)

    res = {}
    for name, fn in (("one_by_one", one_by_one), ("batched", batched)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / args.steps
        log(f"{name}: {res[name]:.3f} s")
    assert all(bool(torch.isfinite(y).all()) for y in dY), "non-finite embedding"
    out = {
        "metric": "fit_transform_many sec @64 data sets (N 2k-50k, D 32-128)",
        "value": round(res["batched"], 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(res["batched"] * 1e3, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": f"{args.precision}-knn/f32", "data": "synthetic",
        "config": {"workload": "64 Gaussian mixtures, N log-uniform [2000, 50000], D in {32,64,128}",
                   "total_points": int(sum(n for n, _, _ in shapes)), "streams": args.streams},
        "one_by_one_s": round(res["one_by_one"], 4),
        "speedup_batched": round(res["one_by_one"] / res["batched"], 3),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="1m", choices=sorted(WORKLOADS) + ["batch64"])
    ap.add_argument("--streams", type=int, default=8, help="streams of the batch64 workload")
    ap.add_argument("--precision", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the fit path runs on a single MI355X in this build "
              "(sharding over several GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch
    from ncvis_amd import abi
    from ncvis_amd._ncvis import find_ab_params, negative_plan_for

    if args.workload == "batch64":
        run_batch(args)
        return
    w = WORKLOADS[args.workload]
    N, D, k, d = w["N"], w["D"], 15, 2
    seed = 42
    log(f"workload {w['desc']}")
    X = make_data(w)
    a, b = find_ab_params(1.0, 0.4)
    plan, n_epochs = negative_plan_for(None, 50)
    n_init = 20
    dev = torch.device("cuda:0")
    dX = torch.from_numpy(X).to(dev)
    dY = torch.empty((N, d), dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream(dev)

    h = abi.Handle(d=d, n_threads=0, n_neighbors=k, random_seed=seed, n_init_epochs=n_init, a=a, b=b,
                   n_noise=list(plan), distance=w["distance"])
    h.set_option("knn_precision", abi.PREC[args.precision])

    def fit():
        h.fit_transform_device(dX.data_ptr(), N, D, dY.data_ptr(), st.cuda_stream)

    for i in range(args.warmup):
        log(f"warmup {i}")
        fit()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        fit()
        stats.append(h.stats())
        log(f"step {i}: {stats[-1]['stage_ms']['total']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sec = elapsed / args.steps
    Y = dY.cpu().numpy()
    assert np.all(np.isfinite(Y)), "non-finite embedding"
    s_last = stats[-1]
    E = s_last["n_edges"]
    stage = {n: float(np.mean([s["stage_ms"][n] for s in stats])) for n in s_last["stage_ms"]}

    # kNN tables of both precisions for recall (outside the timed region)
    rng = np.random.default_rng(1)
    queries = np.sort(rng.choice(N, size=min(N, 10_000), replace=False))
    inds = torch.empty((N, k), dtype=torch.int32, device=dev)
    dists = torch.empty((N, k), dtype=torch.float32, device=dev)
    abi.knn(dX.data_ptr(), N, D, k, w["distance"], args.precision, inds.data_ptr(), dists.data_ptr(),
            st.cuda_stream)
    torch.cuda.synchronize()
    rec_mixed = recall_at_k(torch, dX, inds, k, queries)
    fp32 = {}
    if not args.no_fp32 and args.precision != "fp32":
        log("fp32 kNN fit")
        h32 = abi.Handle(d=d, n_threads=0, n_neighbors=k, random_seed=seed, n_init_epochs=n_init, a=a,
                         b=b, n_noise=list(plan), distance=w["distance"])
        h32.set_option("knn_precision", abi.PREC["fp32"])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        h32.fit_transform_device(dX.data_ptr(), N, D, dY.data_ptr(), st.cuda_stream)
        torch.cuda.synchronize()
        fp32["fit_s"] = time.perf_counter() - t1
        s32 = h32.stats()
        fp32["knn_ms"] = s32["stage_ms"]["knn"]
        fp32["knn_tiles_ms"] = s32["stage_ms"]["knn_tiles"]
        abi.knn(dX.data_ptr(), N, D, k, w["distance"], "fp32", inds.data_ptr(), dists.data_ptr(),
                st.cuda_stream)
        torch.cuda.synchronize()
        fp32["recall@15"] = recall_at_k(torch, dX, inds, k, queries)

    # roofline of the dominant kernel: algorithmic work per launch / event-timed launch duration
    n_samples = E * float(np.sum(1 + plan))
    sgd_bytes = E * 8.0 * n_epochs + n_samples * 4 * d * 4
    knn_flops = 2.0 * N * N * D
    knn_tiles_s = stage["knn_tiles"] / 1e3
    sgd_s = stage["sgd_kernel"] / 1e3
    traffic = measured_traffic(w["desc"], args.precision)
    roof_knn = {"bound": "mfma", "kernel": "k_knn_tiles", "achieved": knn_flops / knn_tiles_s / 1e12,
                "peak": MFMA_PEAK_TFLOPS[args.precision], "unit": "TFLOP/s",
                "traffic": traffic.get("k_knn_tiles"), "launch_ms": stage["knn_tiles"]}
    roof_knn["frac"] = roof_knn["achieved"] / roof_knn["peak"]
    roof_sgd = {"bound": "hbm", "kernel": "k_sgd_epoch", "achieved": sgd_bytes / sgd_s / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": traffic.get("k_sgd_epoch"),
                "launch_ms": stage["sgd_kernel"] / n_epochs}
    roof_sgd["frac"] = roof_sgd["achieved"] / roof_sgd["peak"]
    dominant = roof_knn if knn_tiles_s >= sgd_s else roof_sgd

    # the host-buffer entry point (X and Y cross PCIe): reported beside, never as `value`
    Xh = np.ascontiguousarray(X)
    Yh = np.empty((N, d), np.float32)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    h.fit_transform_host(Xh, Yh)
    host_api_s = time.perf_counter() - t1

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        from oracle import oracle
        Xc = X if w["distance"] == "euclidean" else oracle.preprocess(X, w["distance"])
        cpu, h_inds = cpu_baseline(Xc, plan, a, b, k, n_init, seed, w["distance"])
        inds.copy_(torch.from_numpy(h_inds.astype(np.int32)))
        cpu["knn_recall@15"] = round(recall_at_k(torch, dX, inds, k, queries), 5)

    out = {
        "metric": "fit_transform sec @1M×128 (mixed vs FP32 GPU vs CPU); recall@15 vs FP64",
        "value": round(sec, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.precision}-knn/f32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "N": N, "D": D, "k": k, "d": d, "n_epochs": n_epochs,
                   "n_init_epochs": n_init, "noise_plan": "default (sum 1+n = %d)" % int(np.sum(1 + plan)),
                   "knn_precision": args.precision, "chains": s_last["chains"]},
        "recall@15": round(rec_mixed, 5),
        "fp32": {k_: (round(v, 5) if isinstance(v, float) else v) for k_, v in fp32.items()},
        "stage_ms": {k_: round(v, 2) for k_, v in stage.items()},
        "n_edges": E,
        "host_api_s": round(host_api_s, 4),
        "edge_updates_per_s": n_samples / (stage["optimize"] / 1e3),
        "roofline": {k_: (round(v, 5) if isinstance(v, float) else v) for k_, v in dominant.items()},
        "roofline_all": [{k_: (round(v, 5) if isinstance(v, float) else v) for k_, v in r.items()}
                         for r in (roof_knn, roof_sgd)],
        "cpu_baseline": cpu,
    }
    if fp32.get("fit_s"):
        out["speedup_mixed_vs_fp32_knn"] = round(fp32["knn_ms"] / stage["knn"], 3)
    if cpu:
        out["speedup_vs_cpu"] = round(cpu["value"] / sec, 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
