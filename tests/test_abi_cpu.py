"""CPU-only checks of the boundary: the C-ABI library loads and exports every symbol the header
declares; the Python layer validates arguments like the reference; no CPU fallback exists."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "ncvis_amd.h")).read()
    return sorted(set(re.findall(r"\b(ncvis_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from ncvis_amd import abi
    lib = ctypes.CDLL(abi.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 13
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(abi.EXPORTS)


def test_cython_extension_links_the_library():
    import ncvis_amd
    from ncvis_amd import _ncvis  # noqa: F401
    assert callable(ncvis_amd.NCVis)


def test_find_ab_params_closed_forms():
    import ncvis_amd
    a, b = ncvis_amd.find_ab_params(1.0, 0.4)
    assert abs(a - 0.7668736385558027) < 1e-6 and abs(b - 1.2225638162909107) < 1e-6
    a, b = ncvis_amd.find_ab_params(1.0, 0.1)
    assert abs(a - 1.5769434602697652) < 1e-6 and abs(b - 0.8950608778515733) < 1e-6


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except Exception:
        return True


@pytest.mark.parametrize("kwargs,exc", [
    (dict(n_noise=np.ones((2, 2), np.uint64)), ValueError),
    (dict(n_noise=0), ValueError),
    (dict(n_noise=2.5), ValueError),
    (dict(distance="manhattan"), ValueError),
    (dict(a=1.0), ValueError),
    (dict(knn_precision="int8"), ValueError),
])
def test_python_layer_validation(kwargs, exc):
    import ncvis_amd
    with pytest.raises(exc):
        ncvis_amd.NCVis(**kwargs)


@pytest.mark.skipif(not _no_gpu(), reason="checks the no-GPU failure mode")
def test_no_gpu_fails_loudly():
    import ncvis_amd
    with pytest.raises(RuntimeError):
        ncvis_amd.NCVis()


def test_default_negative_plan():
    # linspace(0,1,50)**3 normalised to 250, rounded, clamped >= 1 (wrapper/ncvis.pyx:103-112)
    want = [1] * 21 + [2] * 4 + [3] * 3 + [4] * 3 + [5] * 2 + [6] + [7] * 2 + [8] * 2 + [9, 10, 11, 11, 12, 13, 14, 15, 16, 17, 18, 20]
    from ncvis_amd._ncvis import negative_plan_for
    plan, ne = negative_plan_for(None, 50)
    assert list(plan) == want and int(plan.sum()) == 262 and ne == 50
    plan, ne = negative_plan_for(np.array([0, 2, 3]), 50)
    assert list(plan) == [0, 2, 3] and ne == 3
    plan, ne = negative_plan_for(4, 7)
    assert list(plan) == [4] * 7 and ne == 7


def test_one_hip_runtime_whatever_the_import_order():
    # the library must bind to the HIP runtime torch bundles, whether it is loaded before or
    # after torch: two runtimes in one process compete for the GPU
    import subprocess
    import sys
    pytest.importorskip("torch")
    code = ("import re, sys; sys.path.insert(0, %r)\n"
            "from ncvis_amd import abi, synthetic\n"
            "abi.lib()\n"
            "import torch\n"
            "m = open('/proc/self/maps').read()\n"
            "print(len(set(re.findall(r'\\S*libamdhip64\\S*', m))))\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"


def test_create_rejects_n_noise_that_overflows_the_sample_count():
    # the epoch kernels count n_noise + 1 samples per edge in 32 bits; 2^32 - 1 must be refused
    # before any GPU work (the check precedes the stream creation, so it runs without a GPU)
    from ncvis_amd import abi
    L = abi.lib()
    h = ctypes.c_void_p()
    for bad in (0xFFFFFFFF, 1 << 40):
        plan = (ctypes.c_size_t * 2)(3, bad)
        rc = L.ncvis_create(2, 1, 15, 16, 200, 42, 2, 20, 1.0, 1.0, 1.0, 1.0, plan, 0, ctypes.byref(h))
        assert rc == -1, rc  # NCVIS_EINVAL
        assert b"n_noise" in L.ncvis_get_last_error()
        assert not h.value


def test_ncvis_module_exposes_the_reference_names():
    # the reference's extension is the module `ncvis` (setup.py:133-134; recipe/test.py:2)
    import ncvis
    import ncvis_amd
    assert ncvis.NCVis is ncvis_amd.NCVis
    assert ncvis.find_ab_params is ncvis_amd.find_ab_params
    assert ncvis.NCVisWrapper is ncvis_amd.NCVisWrapper


def test_exact_threads_is_a_keyword():
    # exact_threads selects n_threads chains; it is validated like the reference's kwargs
    import inspect
    import ncvis_amd
    sig = inspect.signature(ncvis_amd.NCVis.__init__)
    assert sig.parameters["exact_threads"].default is False
    assert sig.parameters["n_threads"].default == -1


def test_tune_cache_reader(tmp_path):
    # the JSON cache of the kNN autotuner loads without a GPU; malformed files are refused
    from ncvis_amd import abi
    good = tmp_path / "tiles.json"
    good.write_text('''{"format": "ncvis_amd kNN tile variants", "version": 1, "entries": [
      {"arch": "gfx950", "prec": 2, "dk": 4, "kl": 16, "wide": 0, "nlog2": 14, "split": 4, "stage_mult": 2, "dma": 1, "ms": 0.5},
      {"arch": "gfx950", "prec": 0, "dk": 2, "kl": 32, "wide": 0, "nlog2": 20, "split": 1, "stage_mult": 2, "dma": 0, "ms": 9.0}
    ]}''')
    try:
        assert abi.load_tune_cache(good) == 2
        bad = tmp_path / "bad.json"
        bad.write_text('{"version": 1}')
        with pytest.raises(RuntimeError, match="entries"):
            abi.load_tune_cache(bad)
        with pytest.raises(RuntimeError, match="cannot read"):
            abi.load_tune_cache(tmp_path / "missing.json")
    finally:
        assert abi.load_tune_cache(None) == 0


def test_sgd_copy_knobs_are_validated():
    # the per-XCD copy knobs are range-checked without touching a GPU
    from ncvis_amd import abi
    for key, bad in (("sgd_replicas", 2), ("sgd_replicas", -1), ("sgd_sync", 0), ("sgd_sync", 65)):
        with pytest.raises(RuntimeError):
            abi.tune(key, bad)
    for key, good in (("sgd_replicas", 1), ("sgd_sync", 16), ("sgd_edge_flush", 0)):
        abi.tune(key, good)
