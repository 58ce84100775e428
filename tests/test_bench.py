"""The benchmark's host side: its workloads are the ones it claims (CPU only)."""

from __future__ import annotations

import json
import subprocess
import sys

import numpy as np

from helpers import ROOT

sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_the_2d_workload_has_the_stated_geometry():
    sim, true, start, sources, sensors = bench.workload(shots=4, steps=10, roi=64, receivers=16)
    assert sim.Nx == (64 + 2 * bench.LAYER + 2,) * 2
    assert len(sources) == 4 and sensors.shape == (2, 16)
    for s in sources:  # the shots sit on the first row of the region
        assert int(s.position[0, 0]) == bench.LAYER + 1
    assert true.shape == tuple(sim.Nx_padded) and start.shape == true.shape
    sim.require_interior(sensors, "sensor")


def test_the_3d_workload_has_two_fields_and_an_air_sphere():
    sim, indicator, source, sensors = bench.workload_3d(nt=5, roi=32, layer=4)
    assert indicator.shape == (2, *sim.Nx_padded)
    kappa, rho = indicator
    assert rho.min() == np.float32(1.204) and kappa.min() == np.float32(1.419e5)
    assert rho.max() <= 1300.0 + 1e-3
    assert sensors.shape == (3, 1024)
    sim.require_interior(source.position, "source")
    sim.require_interior(sensors, "sensor")


def test_more_than_one_gpu_is_refused():
    done = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120)
    assert done.returncode == 2
    assert "one GPU" in json.loads(done.stdout.strip().splitlines()[-1])["error"]
