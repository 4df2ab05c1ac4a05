"""CPU tier: the host-only legs of bench.py (the CPU baseline's bond sample
and the sweep's cost units), at a bond dimension that runs in seconds."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_cpu_baseline_times_every_part_at_full_size():
    bond, parts = bench.cpu_baseline(48)
    for k in ("heff_apply", "density_matrix", "eigh_sample", "eigh_scaled", "environment"):
        assert parts[k] > 0.0
    # the eigendecomposition is timed at n = d chi itself, not scaled
    assert parts["eigh_scaled"] == parts["eigh_sample"]
    want = 3 * parts["heff_apply"] + parts["density_matrix"] + parts["eigh_scaled"] + \
        parts["environment"]
    assert bond == pytest.approx(want)


def test_cpu_baseline_scaled_sample():
    _, parts = bench.cpu_baseline(48, eig_n=32)
    assert parts["eigh_scaled"] == pytest.approx(parts["eigh_sample"] * (96 / 32) ** 3)


def test_sweep_cost_units():
    # every bond of a chain far longer than log2(chi) sits at the full shape
    # except the 2 log2(chi) or so near each end, so the units stay below the
    # 2 (N - 1) bond updates of a sweep and grow with N
    u100 = bench.sweep_cost_units(100, 2, 2048, 5)
    u200 = bench.sweep_cost_units(200, 2, 2048, 5)
    assert 0 < u100 < 2 * 99
    assert u200 > u100
