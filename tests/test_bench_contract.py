"""The bench's output contract on the CPU: the record's fields, the roofline block's arithmetic,
and the refusal of a run on several GPUs (no GPU is touched)."""
import importlib.util
import os
import subprocess
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_several_gpus_are_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_record_fields(bench):
    args = types.SimpleNamespace(warmup=5)
    rec = bench.base_record(args, 500, 0.2, 2500.0, {"workload": "w"})
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config"):
        assert key in rec
    assert rec["n_gpus"] == 1 and rec["steps"] == 500 and rec["warmup"] == 5
    assert rec["ms_per_step"] == pytest.approx(1000.0 * 0.2 / 500)
    assert rec["higher_is_better"] is True and rec["vs_baseline"] is None
    assert "model" not in rec["config"]


def test_roofline_block(bench):
    s0 = {"interactions_total": 1_000}
    s1 = {"interactions_total": 1_000 + 50 * 2_000_000, "cells": 400_000}
    roof = bench.walk_roofline(0.1, s0, s1, 50, 600_000, False)
    # 2e6 interactions per launch, 12 flops each, in 0.1 ms
    assert roof["interactions_per_iteration"] == pytest.approx(2e6)
    assert roof["achieved"] == pytest.approx(12 * 2e6 / 1e-4 / 1e12)
    assert roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"])
    assert roof["unit"] == "TFLOP/s" and roof["bound"] == "valu"
    assert roof["avg_launch_ms"] == 0.1
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in roof
    # every record once (16 B each) plus each body's 32 B of forces, positions and ids
    assert roof["algorithmic_bytes_per_launch"] == bench.walk_algorithmic_bytes(600_000, 400_000, False)
