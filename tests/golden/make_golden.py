"""Regenerates the committed golden fixtures (run in the build container only).

* ``rng_golden.json``: pcg64 raw outputs, ``uniform_int_distribution<size_t>`` and
  ``uniform_real_distribution<float>`` draws, from ``make_golden_rng.cpp`` compiled against the
  pcg-cpp header vendored by pyarrow and libstdc++ -- the same engine and distributions the
  reference uses at src/ncvis.cpp:195-196 and :281-283.
* ``sym_*.npz``: kNN tables (inputs) and their symmetrized CSR (expected outputs) produced by
  the reference's own ``KNNTable::symmetrize`` (src/knntable.cpp:24-99) at one thread, compiled
  into ``oracle/_ref`` by ``oracle/Makefile``.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from ncvis_amd import synthetic  # noqa: E402


def make_rng():
    import pyarrow
    inc = pyarrow.get_include()
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "gen")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", inc, "-o", exe,
                               os.path.join(HERE, "make_golden_rng.cpp")])
        out = subprocess.check_output([exe]).decode()
    data = json.loads(out)
    with open(os.path.join(HERE, "rng_golden.json"), "w") as f:
        json.dump(data, f, indent=1)


def tie_heavy_table(rng, N, k, n_levels, hub_frac=0.0, n_hubs=0, sort_desc=True):
    inds = np.empty((N, k), np.int64)
    dists = np.empty((N, k), np.float32)
    hubs = rng.choice(N, size=max(n_hubs, 1), replace=False)
    for i in range(N):
        pool = []
        while len(pool) < k:
            if n_hubs and rng.random() < hub_frac:
                c = int(rng.choice(hubs))
            else:
                c = int(rng.integers(0, N))
            if c != i and c not in pool:
                pool.append(c)
        dv = (rng.integers(1, n_levels + 1, k) / n_levels).astype(np.float32)
        if sort_desc:
            o = np.lexsort((np.array(pool), dv))[::-1]
            pool = list(np.array(pool)[o])
            dv = dv[o]
        inds[i] = pool
        dists[i] = dv
    return inds, dists


def save_sym(name, inds, dists):
    row_ptr, col, dist = oracle.ref_symmetrize(inds, dists, n_threads=1)
    np.savez_compressed(os.path.join(HERE, name), inds=inds.astype(np.int32), dists=dists,
                        row_ptr=row_ptr.astype(np.int64), col=col.astype(np.int32), dist=dist)
    lens = np.diff(row_ptr)
    print(f"{name}: N={inds.shape[0]} k={inds.shape[1]} E={len(col)} max_row={lens.max()}")


def make_sym():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    # digits-shaped integer data: exact distances with many ties
    X, _ = synthetic.digits_like(N=1500, D=64, seed=1)
    inds, dists = oracle.knn_bruteforce(oracle.preprocess(X), 15)
    save_sym("sym_digits1500.npz", inds, dists)
    rng = np.random.default_rng(7)
    # few distance levels + hub targets: long rows (introsort with partitions), heavy ties
    inds, dists = tie_heavy_table(rng, 400, 15, 3, hub_frac=0.3, n_hubs=4)
    save_sym("sym_ties_hubs.npz", inds, dists)
    # unsorted own lists (any table is legal input to symmetrize)
    inds, dists = tie_heavy_table(rng, 300, 6, 50, sort_desc=False)
    save_sym("sym_unsorted.npz", inds, dists)
    # tiny tables, k = N-1
    inds, dists = tie_heavy_table(rng, 5, 4, 2)
    save_sym("sym_tiny.npz", inds, dists)


if __name__ == "__main__":
    make_rng()
    make_sym()
