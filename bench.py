"""Throughput of the shot-batched 2D scalar forward + adjoint on one MI355X.

Workload: a 1024 x 1024 region behind a 32-cell PML (Nx = 1090 x 1090, padded to
1090 x 1120), dx = 1 mm, order 8, float32, 4000 steps, 32 Ricker shots along the
top row and 256 receivers every 4 cells; the true model is 1500 m/s with
smoothed 10 % noise and five elliptical inclusions at 2000-3000 m/s, the start
model its sigma-40 smoothing. One step of this benchmark is one full FWI
gradient: every shot's forward sweep (traces and boundary strip recorded),
the L2 trace misfit, and the adjoint sweep that rebuilds the forward field
backwards and sums all shots' velocity gradients, inputs resident in HBM.

Work unit: one interior node (PML included) of one shot advanced one step by
one sweep; a step is W = 2 B n_int N units (forward and adjoint sweeps).
Gcell-updates/s = W / wall time of the step.

The JSON line adds `roofline` for the dominant work, the adjoint level (its two
launches back to back: the core kernel over the layer-free rectangle and the
general kernels over the frame around it, timed together with HIP events on the
launch stream), and `cpu_baseline`, the C/OpenMP oracle on the host cores over
a bounded sample.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # B/s, MI355X published
ROI, LAYER, DX = 1024, 32, 1e-3
C1, C2 = 1500.0, 3000.0
FREQ, SHOTS, RECEIVERS, STEPS = 1e5, 32, 256, 4000


def models(roi=ROI, seed_noise=2, seed_ellipses=3):
    """True and start velocity over the region: smoothed noise plus five ellipses."""
    from scipy.ndimage import gaussian_filter

    noise = gaussian_filter(np.random.default_rng(seed_noise).standard_normal((roi, roi)), 8)
    noise /= noise.std()
    true = 1500.0 * (1.0 + 0.1 * noise)
    rng = np.random.default_rng(seed_ellipses)
    a0, a1 = np.meshgrid(np.arange(roi), np.arange(roi), indexing="ij")
    for _ in range(5):
        c = rng.uniform(2000, 3000)
        ax = rng.uniform(20, 80, 2)
        ang = rng.uniform(0, np.pi)
        cen = rng.uniform(0, roi, 2)
        x, y = a0 - cen[0], a1 - cen[1]
        u = x * np.cos(ang) + y * np.sin(ang)
        v = -x * np.sin(ang) + y * np.cos(ang)
        true[(u / ax[0]) ** 2 + (v / ax[1]) ** 2 <= 1.0] = c
    start = gaussian_filter(true, 40)
    return true, start


def workload(shots=SHOTS, steps=STEPS, roi=ROI, receivers=RECEIVERS):
    """The simulation, the true and start indicators, the shots and the receiver nodes."""
    from cuwave_amd import boundary, scalar, wave
    from cuwave_amd.signals import ricker

    true, start = models(roi)
    width = LAYER + 1  # the wall node is part of the layer
    n = roi + 2 * LAYER + 2
    dt = 0.4 * DX / float(true.max())
    sim = scalar.AcousticWave((n, n), (DX, DX), steps, dt, precision="float32", space_order=8,
                              rho1=1.0, rho2=1.0, kappa1=C1**2, kappa2=C2**2,
                              pml=boundary.Pml(width, float(true.max()), 1e-6))

    def indicator(v):
        full = np.pad(v, ((width, n - width - roi), (width, sim.Nx_padded[1] - width - roi)), mode="edge")
        return sim.indicator_of_velocity(full)

    t = np.arange(steps) * dt
    wavelet = (ricker(t, 1.0, FREQ) / (DX * DX))[:, None]
    cols = np.round(np.linspace(16, roi - 16, shots)).astype(int) + width
    sources = [wave.Source(np.array([[width], [c]]), wavelet) for c in cols]
    rec_cols = np.arange(0, roi, max(roi // receivers, 1))[:receivers] + width
    sensors = np.stack([np.full(len(rec_cols), width), rec_cols])
    return sim, indicator(true), indicator(start), sources, sensors


def gpu_run(args):
    import torch

    from cuwave_amd import utils, wave
    from cuwave_amd.engine import Propagator

    sim, true, start, sources, sensors = workload(args.shots, args.nt, args.roi)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    B = len(sources)

    class NodeSensors(utils.Sensors):
        """Receivers on grid nodes: the record is the traces."""

        def __init__(self, nodes):
            self.nodes = nodes
            self.count = nodes.shape[1]
            self.weights = np.ones((self.count, 1), dtype=np.float32)

    rec = NodeSensors(sensors)
    _, observed = wave.simulate_shots(sim, sources, true, sensors)
    torch.cuda.synchronize()

    def step():
        return utils.misfit_gradient(sim, sources, start, rec, list(observed),
                                     adjoint=utils.reconstruction_sensitivity)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    Propagator.timings = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cost, gradient = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timings, Propagator.timings = Propagator.timings, None
    work = 2.0 * B * n_int * sim.N * args.steps
    # the dominant work: an adjoint level (core + frame launches) for the batch
    per = {}
    for kind, launches, s, e in timings:
        ms = s.elapsed_time(e)
        tot = per.setdefault(kind, [0.0, 0])
        tot[0] += ms
        tot[1] += launches
    bwd_ms = per["backward"][0] / per["backward"][1]
    fwd_ms = per["forward"][0] / per["forward"][1]
    bytes_bwd = (24.0 * B + 16.0) * n_int  # algorithmic bytes of one adjoint level
    bytes_fwd = (12.0 * B + 8.0) * n_int
    achieved = bytes_bwd / (bwd_ms * 1e-3)
    return {
        "value": work / elapsed / 1e9,
        "ms_per_step": elapsed / args.steps * 1e3,
        "cost": cost,
        "grad_norm": float(torch.linalg.vector_norm(gradient)),
        "n_int": n_int,
        "B": B,
        "sim": sim,
        "kernels": {
            "backward_ms": bwd_ms,
            "forward_ms": fwd_ms,
            "backward_GBps": achieved / 1e9,
            "forward_GBps": bytes_fwd / (fwd_ms * 1e-3) / 1e9,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic_per_launch(),
        },
    }


def traffic_per_launch():
    """HBM bytes per adjoint level from the committed rocprofv3 PMC summary, if any."""
    path = ROOT / "profiles" / "traffic_adjoint_level_scalar2d_b32.json"
    if not path.exists():
        return None
    return json.loads(path.read_text()).get("bytes_per_launch")


def cpu_baseline(sample_steps: int, shots: int = 1):
    """The oracle (C/OpenMP restatement of the reference's CPU path) on the host cores.

    A bounded sample of the same workload: `shots` shot(s), the full grid, the
    first `sample_steps` steps, through the oracle's reconstruction adjoint.
    """
    from oracle import driver as D

    sim, true, start, sources, sensors = workload(shots, sample_steps)
    osim = D.Sim("AcousticWave", sim.Nx, sim.dx, sim.N, sim.dt, "float32", 8,
                 materials=(1.0, 1.0, C1**2, C2**2),
                 pml=D.Pml(sim.pml.width, sim.pml.wavespeed, sim.pml.reflection))
    _, observed = D.simulate(osim, sources[0].position, sources[0].signal, true, sensors)
    t0 = time.perf_counter()
    for s in sources:
        D.reconstruction_sensitivity(osim, s.position, s.signal, start, sensors,
                                     D.l2_misfit(observed))
    elapsed = time.perf_counter() - t0
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    return 2.0 * shots * n_int * sample_steps / elapsed / 1e9, elapsed


def probe_copy(mib: int = 256, reps: int = 5) -> None:
    """A plain device copy of `mib` MiB, `reps` times: the known-size kernel that
    calibrates the FETCH_SIZE / WRITE_SIZE counters the traffic figures use."""
    import torch

    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "copy", "bytes_read": src.numel() * 4,
                      "bytes_written": dst.numel() * 4, "reps": reps}))


def workload_3d(nt: int, roi: int = 320, layer: int = 16):
    """The 3D acoustic case: independent bulk modulus and density.

    ROI roi^3 behind a `layer`-cell PML, dx = 1 mm, order 8, float32. Both
    fields are smoothed noise (sigma 10), min-max scaled: rho in [900, 1300]
    (seed 4), vp in [1400, 1800] (seed 5), kappa = rho vp^2. An air sphere of
    radius roi/16 sits at the centre (rho 1.204, kappa 1.419e5). One Ricker
    source at the centre of the top face; a 32 x 32 receiver grid on the top
    plane. Returns the simulation, the (kappa, rho) indicator, the source and
    the receiver nodes.
    """
    from scipy.ndimage import gaussian_filter

    from cuwave_amd import boundary, scalar, wave
    from cuwave_amd.signals import ricker

    def field(seed, lo, hi):
        f = gaussian_filter(np.random.default_rng(seed).standard_normal((roi,) * 3), 10)
        return lo + (hi - lo) * (f - f.min()) / (f.max() - f.min())

    rho, vp = field(4, 900.0, 1300.0), field(5, 1400.0, 1800.0)
    kappa = rho * vp**2
    c = (np.arange(roi) - roi / 2) ** 2
    air = c[:, None, None] + c[None, :, None] + c[None, None, :] <= (roi / 16) ** 2
    rho[air], kappa[air] = 1.204, 1.419e5
    width = layer + 1
    n = roi + 2 * width
    dt = 0.4 * DX / 1800.0
    sim = scalar.BulkDensityWave((n, n, n), (DX, DX, DX), nt, dt, precision="float32",
                                 space_order=8, pml=boundary.Pml(width, 1800.0, 1e-6))
    pad = [(width, n - width - roi), (width, n - width - roi),
           (width, sim.Nx_padded[2] - width - roi)]
    indicator = np.stack([np.pad(kappa, pad, mode="edge"), np.pad(rho, pad, mode="edge")])
    wavelet = (ricker(np.arange(nt) * dt, 1.0, FREQ) / DX**3)[:, None]
    source = wave.Source(np.array([[width], [width + roi // 2], [width + roi // 2]]), wavelet)
    g = np.arange(32) * (roi // 32) + roi // 64 + width
    sensors = np.stack([np.full(32 * 32, width), np.repeat(g, 32), np.tile(g, 32)])
    return sim, indicator.astype(np.float32), source, sensors


def gpu_run_3d(args):
    """Forward + reconstruction adjoint of the 3D case: W = 2 n_int N units a step."""
    import torch

    from cuwave_amd import sensitivity, wave
    from cuwave_amd.engine import Propagator

    sim, indicator, source, sensors = workload_3d(args.nt)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    observed = 0.9 * wave.simulate(sim, source, indicator, sensors)[1].cpu().numpy()
    objective = sensitivity.l2_misfit(observed)

    def step():
        return sensitivity.reconstruction_sensitivity(sim, source, indicator, sensors,
                                                      objective)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    Propagator.timings = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timings, Propagator.timings = Propagator.timings, None
    per = {}
    for kind, launches, s, e in timings:
        tot = per.setdefault(kind, [0.0, 0])
        tot[0] += s.elapsed_time(e)
        tot[1] += launches
    bwd_ms = per["backward"][0] / per["backward"][1]
    fwd_ms = per["forward"][0] / per["forward"][1]
    # one shot: the reverse march 20 B, the adjoint 12 B and two accumulators
    # read and written 16 B a node a level; the forward 20 B a node a step
    bytes_bwd, bytes_fwd = 48.0 * n_int, 20.0 * n_int
    achieved = bytes_bwd / (bwd_ms * 1e-3)
    return {
        "metric": "Gcell-updates/s (3D acoustic forward+adjoint)",
        "value": 2.0 * n_int * sim.N * args.steps / elapsed / 1e9,
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3D acoustic gradient: forward + reconstruction adjoint, "
                               "independent bulk modulus and density",
                   "grid": f"{sim.Nx[0]}^3 (ROI 320^3 + 17-cell PML)", "space_order": 8,
                   "shots": 1, "receivers": 1024, "time_steps": args.nt,
                   "interior_nodes": n_int, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
        "kernels": {"backward_ms": bwd_ms, "forward_ms": fwd_ms,
                    "backward_GBps": achieved / 1e9,
                    "forward_GBps": bytes_fwd / (fwd_ms * 1e-3) / 1e9},
    }


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--shots", type=int, default=SHOTS)
    p.add_argument("--nt", type=int, default=STEPS)
    p.add_argument("--roi", type=int, default=ROI)
    p.add_argument("--cpu-sample", type=int, default=1500, help="oracle steps for the CPU baseline")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--probe-copy", action="store_true",
                   help="run only a 256 MiB device copy (counter calibration)")
    p.add_argument("--workload", choices=("2d", "3d"), default="2d",
                   help="3d: the 3D acoustic gradient on a 354^3 grid instead (its own line)")
    args = p.parse_args()
    if args.probe_copy:
        probe_copy()
        return
    if args.workload == "3d":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        print(json.dumps(gpu_run_3d(args)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the shot batch shards over "
                                   "GPUs by design only; run with --gpus 1"}))
        sys.exit(2)
    res = gpu_run(args)
    cpu = None
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        value, seconds = cpu_baseline(args.cpu_sample)
        cpu = {"value": value, "unit": "Gcell-updates/s", "cores": min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))),
               "kind": "port",
               "sample": f"1 shot, full 1090x1090 grid, first {args.cpu_sample} of {STEPS} steps, "
                         f"forward + reconstruction adjoint through the C oracle ({seconds:.1f} s)"}
    line = {
        "metric": "Gcell-updates/s (shot-batched 2D scalar forward+adjoint)",
        "value": res["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "2D scalar FWI gradient: forward + reconstruction adjoint, shot-batched",
            "grid": f"{res['sim'].Nx[0]}x{res['sim'].Nx[1]} (ROI {args.roi}^2 + {LAYER}-cell PML)",
            "space_order": 8,
            "shots": res["B"],
            "receivers": RECEIVERS,
            "time_steps": args.nt,
            "interior_nodes": res["n_int"],
            "parallelism": "single GPU, shots batched per launch",
        },
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "kernels": res["kernels"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
