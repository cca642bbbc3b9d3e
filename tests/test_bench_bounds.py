"""bench.py's work-weighted Amdahl bound of batching over the plain loop (host logic only)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def test_bound_from_the_full_size_jobs_rate():
    # 1024^2 jobs: 100 units of work in 0.5 s, so the GPU sustains 200 units/s; the mix's 150
    # units then need at least 0.75 s against the plain loop's 1.5 s
    b = bench.amdahl_bound(1.5, {128: 0.4, 1024: 0.5}, {128: 50, 1024: 100})
    assert abs(b - 2.0) < 1e-12


def test_no_bound_without_full_size_jobs():
    # jobs that do not fill the GPU alone bound nothing (64 Cornell rooms at 256^2 batch 4x)
    assert bench.amdahl_bound(0.2, {256: 0.2}, {256: 100}) is None
    assert bench.amdahl_bound(1.0, {1024: 0.0}, {1024: 10}) is None
