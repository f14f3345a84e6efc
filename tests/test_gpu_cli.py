"""The command-line harness (the reference's src/main.cpp arguments and data) on the GPU."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ncvis_amd", "ncvis_cli")


def _run(*args):
    if not os.path.exists(CLI):
        pytest.fail("ncvis_cli is not built (python -m ncvis_amd.build)")
    print(f"[case] ncvis_cli {' '.join(args)}", flush=True)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=240)


def test_cli_usage_without_arguments():
    p = _run()
    assert p.returncode == 1 and p.stdout.startswith("Usage: ncvis_cli")


def test_cli_prints_small_embedding():
    p = _run("5", "3", "1", "10", "5")
    assert p.returncode == 0, p.stderr
    rows = re.findall(r"^\[(.*)\]$", p.stdout, re.M)
    assert len(rows) == 5
    for r in rows:
        vals = [float(v) for v in r.split()]
        assert len(vals) == 2 and all(abs(v) < 1e6 for v in vals)


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_cli_stage_times(precision):
    p = _run("3000", "15", "0", "20", "5", precision)
    assert p.returncode == 0, p.stderr
    for stage in ("knn", "symmetrize", "initialize", "optimize", "fit_transform"):
        m = re.search(rf"^{stage}: ([0-9.]+) ms", p.stdout, re.M)
        assert m and float(m.group(1)) >= 0.0, p.stdout
