"""bench.py reads each roofline's HBM traffic from the committed rocprofv3
summaries (FETCH_SIZE + WRITE_SIZE per launch, calibrated on a copy
kernel). Every kernel shape the library runs by default must have its
traffic there, at a plausible multiple of its algorithmic bytes (the
summaries' own traffic_over_algorithmic lines: 1.03 to 1.28)."""
from __future__ import annotations

import importlib.util
from pathlib import Path

import pytest

from psyclonebench_amd import _native

ROOT = Path(__file__).resolve().parents[1]
_spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)

N4096 = 4096 * 4096
ENS_POINTS = 30903964  # the 256-domain ensemble of seed 0


@pytest.mark.parametrize("label, alg", [
    ("mixed rows 4 prefetch 2", bench.B_ALG_MIXED * N4096),
    ("mixed rows 4 cols 56 prefetch 2", bench.B_ALG_MIXED * N4096),
    ("ensemble f64 4x0p2", bench.B_ALG_F64 * ENS_POINTS),
    ("ensemble f32 4x0p2", bench.B_ALG_MIXED * ENS_POINTS),
])
def test_default_shapes_have_profiled_traffic(label, alg):
    t = bench.traffic_from_profile(label)
    assert t is not None, label
    assert 1.0 <= t / alg <= 1.3, (label, t / alg)


def test_default_fused_variant_has_profiled_traffic():
    v = _native.hip_lib().nemo_hip_fused_variant()
    t = bench.traffic_from_profile(f"variant {v}")
    assert t is not None, v
    assert 1.0 <= t / (bench.B_ALG_F64 * N4096) <= 1.3
