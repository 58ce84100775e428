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


def workload(shots=SHOTS, steps=STEPS, roi=ROI, receivers=RECEIVERS, precision="float32"):
    """The simulation, the true and start indicators, the shots and the receiver nodes."""
    from cuwave_amd import boundary, scalar, wave
    from cuwave_amd.signals import ricker

    true, start = models(roi)
    width = LAYER + 1  # the wall node is part of the layer
    n = roi + 2 * LAYER + 2
    dt = 0.4 * DX / float(true.max())
    sim = scalar.AcousticWave((n, n), (DX, DX), steps, dt, precision=precision, space_order=8,
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


class _GridSensors:
    """Receivers on grid nodes, as `utils.Sensors` (the record is the traces)."""

    def __new__(cls, nodes):
        from cuwave_amd import utils

        class NodeSensors(utils.Sensors):
            def __init__(self, nodes):
                self.nodes = nodes
                self.count = nodes.shape[1]
                self.weights = np.ones((self.count, 1), dtype=np.float32)

        return NodeSensors(nodes)


def velocity_of(sim, indicator):
    """The velocity field an AcousticWave(rho1 = rho2 = 1) indicator encodes."""
    m = 1.0 / C1**2 + np.asarray(indicator, dtype=np.float64) * (1.0 / C2**2 - 1.0 / C1**2)
    return 1.0 / np.sqrt(m)


def gpu_run(args):
    """The 2D gradient (32 shots, 1090 x 1090, order 8): velocity -> indicator -> traces
    (the autograd Function over every shot in one batch) -> L2 misfit -> backward to the
    velocity."""
    import torch

    from cuwave_amd import autograd, wave
    from cuwave_amd.engine import Propagator

    sim, true, start, sources, sensors = workload(args.shots, args.nt, args.roi)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    B = len(sources)
    rec = _GridSensors(sensors)
    _, observed = wave.simulate_shots(sim, sources, true, sensors)
    observed = observed.contiguous()
    c_start = torch.as_tensor(velocity_of(sim, start), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def step():
        c = c_start.clone().requires_grad_(True)
        gamma = (1.0 / c**2 - 1.0 / C1**2) / (1.0 / C2**2 - 1.0 / C1**2)
        traces = autograd.traces(gamma, sim, sources, rec)
        loss = 0.5 * torch.sum((traces - observed) ** 2)
        loss.backward()
        return float(loss.detach()), c.grad

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
    measured_at = B == SHOTS and args.roi == ROI
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
            # measured with the profiler at the default 32 shots only
            "traffic": traffic_per_launch() if measured_at else None,
        },
    }


def one_shot_checks(args, shots: int = 4):
    """Per the workload: the one-shot-per-launch rate over `shots` shots run one after
    another (W / wall time, as the batched line), and for the first shot the
    reconstruction drift and the float32 gradient's error against the float64 run of
    the same scheme."""
    import torch

    from cuwave_amd import sensitivity, utils, wave

    sim, true, start, sources, sensors = workload(args.shots, args.nt, args.roi)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    rec = _GridSensors(sensors)
    part = sources[:shots]
    _, observed = wave.simulate_shots(sim, part, true, sensors)
    observed = [o for o in observed]
    utils.misfit_gradient(sim, part[:1], start, rec, observed[:1], batch=1)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    utils.misfit_gradient(sim, part, start, rec, observed, batch=1)
    torch.cuda.synchronize()
    b1 = 2.0 * len(part) * n_int * sim.N / (time.perf_counter() - t0) / 1e9
    grads = {}
    drift = None
    for precision in ("float32", "float64"):
        s64 = workload(1, args.nt, args.roi, precision=precision)[0]
        obs = wave.simulate(s64, part[0], true.astype(s64.np_dtype), sensors)[1]
        _, g, _, info = sensitivity.reconstruction_sensitivity(
            s64, part[0], start.astype(s64.np_dtype), sensors, sensitivity.l2_misfit(obs.cpu()))
        jm, js = s64.parametrization_jacobian()
        grads[precision] = (jm * g["mass"].double() + js * g["stiff"].double()).cpu()
        if precision == "float32":
            drift = info["drift"]
    g32, g64 = grads["float32"], grads["float64"]
    err = float(torch.linalg.vector_norm(g32 - g64) / torch.linalg.vector_norm(g64))
    return {"b1_value": b1, "b1_shots": len(part), "fp64_rel_l2": err, "drift": drift}


def traffic_per_launch(name: str = "traffic_adjoint_level_scalar2d_b32.json"):
    """HBM bytes per adjoint level from a committed rocprofv3 PMC summary, if any."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    return json.loads(path.read_text()).get("bytes_per_launch")


def cpu_baseline(sample_steps: int, runs: int = 3):
    """The oracle (C/OpenMP restatement of the reference's CPU path) on the host cores.

    A bounded sample of the same workload: one shot, the full grid, the first
    `sample_steps` steps, forward + reconstruction adjoint through the oracle;
    the median of `runs` timed runs after one warm-up run.
    """
    from oracle import driver as D

    sim, true, start, sources, sensors = workload(1, sample_steps)
    osim = D.Sim("AcousticWave", sim.Nx, sim.dx, sim.N, sim.dt, "float32", 8,
                 materials=(1.0, 1.0, C1**2, C2**2),
                 pml=D.Pml(sim.pml.width, sim.pml.wavespeed, sim.pml.reflection))
    s = sources[0]
    _, observed = D.simulate(osim, s.position, s.signal, true, sensors)
    times = []
    for _ in range(runs + 1):
        t0 = time.perf_counter()
        D.reconstruction_sensitivity(osim, s.position, s.signal, start, sensors,
                                     D.l2_misfit(observed))
        times.append(time.perf_counter() - t0)
    median = float(np.median(times[1:]))
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    return 2.0 * n_int * sample_steps / median / 1e9, median


def cpu_baseline_3d(sample_steps: int, runs: int = 3):
    """The oracle on the host cores over a bounded sample of the 3D workload: the
    full 354^3 grid, its first `sample_steps` steps, forward + reconstruction
    adjoint; the median of `runs` timed runs after one warm-up run."""
    from oracle import driver as D

    sim, indicator, source, sensors = workload_3d(sample_steps)
    osim = D.Sim("BulkDensityWave", sim.Nx, sim.dx, sim.N, sim.dt, "float32", 8,
                 pml=D.Pml(sim.pml.width, sim.pml.wavespeed, sim.pml.reflection))
    _, observed = D.simulate(osim, source.position, source.signal, indicator, sensors)
    observed = 0.9 * observed
    times = []
    for _ in range(runs + 1):
        t0 = time.perf_counter()
        D.reconstruction_sensitivity(osim, source.position, source.signal, indicator, sensors,
                                     D.l2_misfit(observed))
        times.append(time.perf_counter() - t0)
    median = float(np.median(times[1:]))
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    return 2.0 * n_int * sample_steps / median / 1e9, median


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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic_per_launch("traffic_adjoint_level_acoustic3d_b1.json")},
        "cpu_baseline": None,
        "kernels": {"backward_ms": bwd_ms, "forward_ms": fwd_ms,
                    "backward_GBps": achieved / 1e9,
                    "forward_GBps": bytes_fwd / (fwd_ms * 1e-3) / 1e9},
    }


def workload_forward(nt: int = 2000):
    """The plain 2D forward: ROI 256^2 behind a 20-cell PML (Nx = 298 x 298), dx = 1 mm,
    order 8, float32; 1500 m/s with a 2000 m/s disk of radius 30 cells (centre
    seed 1); one Ricker source (100 kHz) at ROI (0, 128), 16 receivers on ROI row
    0; dt = 0.4 dx / 2000."""
    from cuwave_amd import boundary, scalar, wave
    from cuwave_amd.signals import ricker

    roi, layer = 256, 20
    width = layer + 1
    n = roi + 2 * layer + 2
    dt = 0.4 * DX / 2000.0
    sim = scalar.AcousticWave((n, n), (DX, DX), nt, dt, precision="float32", space_order=8,
                              rho1=1.0, rho2=1.0, kappa1=1500.0**2, kappa2=2000.0**2,
                              pml=boundary.Pml(width, 2000.0, 1e-6))
    centre = np.random.default_rng(1).integers(64, 192, 2)
    a0, a1 = np.meshgrid(np.arange(sim.Nx_padded[0]), np.arange(sim.Nx_padded[1]), indexing="ij")
    disk = (a0 - width - centre[0]) ** 2 + (a1 - width - centre[1]) ** 2 <= 30**2
    indicator = disk.astype(np.float32)
    wavelet = ricker(np.arange(nt) * dt, 1.0, 1e5)[:, None]
    source = wave.Source(np.array([[width], [width + 128]]), wavelet)
    cols = np.round(np.linspace(0, roi - 1, 16)).astype(int) + width
    sensors = np.stack([np.full(16, width), cols])
    return sim, indicator, source, sensors


def forward_line(args) -> dict:
    """Gcell-updates/s of the plain forward (W = n_int N per run), timed over whole runs."""
    import torch

    from cuwave_amd import wave
    from cuwave_amd.engine import Propagator

    from cuwave_amd import tuning

    nt = args.nt if args.nt != STEPS else 2000
    sim, indicator, source, sensors = workload_forward(nt)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))

    def runs(count):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(count):
            wave.simulate(sim, source, indicator, sensors)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    window = wave.graph_window(sim, 1)
    runs(args.warmup)
    Propagator.timings = []
    elapsed = runs(args.steps)
    timings, Propagator.timings = Propagator.timings, None
    # the same runs replayed from captured windows of 24 steps, for comparison
    below, wave.GRAPH_BELOW = wave.GRAPH_BELOW, 1 << 40
    try:
        graphed_s = runs(args.steps)
    finally:
        wave.GRAPH_BELOW = below
    tuned = tuning.tune(sim, 1, indicator, save=False)
    launch_ms = sum(s.elapsed_time(e) for _, _, s, e in timings) / sum(n for _, n, _, _ in timings)
    achieved = 20.0 * n_int / (launch_ms * 1e-3)
    cpu = None
    if not args.no_cpu:
        from oracle import driver as D

        osim = D.Sim("AcousticWave", sim.Nx, sim.dx, sim.N, sim.dt, "float32", 8,
                     materials=(1.0, 1.0, 1500.0**2, 2000.0**2),
                     pml=D.Pml(sim.pml.width, sim.pml.wavespeed, sim.pml.reflection))
        times = []
        for _ in range(4):
            c0 = time.perf_counter()
            D.simulate(osim, source.position, source.signal, indicator, sensors)
            times.append(time.perf_counter() - c0)
        med = float(np.median(times[1:]))
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": n_int * nt / med / 1e9, "unit": "Gcell-updates/s",
               "cores": min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))),
               "kind": "port",
               "sample": f"the whole {nt}-step forward through the C oracle; median of 3 runs "
                         f"after a warm-up ({med:.2f} s a run)"}
    return {
        "metric": "Gcell-updates/s (2D scalar forward, one shot)",
        "value": n_int * nt * args.steps / elapsed / 1e9, "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D scalar forward: traces and final field, one shot",
                   "grid": "298x298 (ROI 256^2 + 20-cell PML)", "space_order": 8,
                   "shots": 1, "receivers": 16, "time_steps": nt, "interior_nodes": n_int,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
        "kernels": {"step_us": launch_ms * 1e3,
                    "note": "a step's launches on the stream, averaged over every step run"},
        "graph_window": window,
        "graphed": {"value": n_int * nt * args.steps / graphed_s / 1e9,
                    "ms_per_step": graphed_s / args.steps * 1e3,
                    "note": "the same runs replayed from one captured 24-step window"},
        "tuner": {"stream_threshold": tuned["stream_threshold"],
                  "graph_window": tuned["graph_window"],
                  "us_per_step": {k: round(v * 1e3, 2) for k, v in tuned["timings"].items()},
                  "note": "tuning.tune on this sweep: threshold/window candidates, best of 3"},
    }


def learned_line(args) -> dict:
    """The learned-FWI workload: the U-Net preconditioner trained on solver data, then plain against
    learned FWI on held-out models, timed to the plain run's final misfit.

    ROI `roi`^2 behind a 32-cell PML, dx = 1 mm, order 8, float32, `shots`
    Ricker shots (100 kHz) along the top row, `receivers` receivers every
    roi / receivers cells; velocity 1500-2500 m/s layered models with 1-6
    ellipses (seeds 0.. for training, 1000.. held out), the start model each
    one's Gaussian smoothing (sigma roi / 32). Training: one sample a model
    (the features at the start model against the step to the true one),
    AdamW lr 1e-3, batch 16. The shots run `shot_batch` a sweep.
    """
    import torch
    from scipy.ndimage import gaussian_filter

    from cuwave_amd import boundary, learned, scalar, utils, wave
    from cuwave_amd.signals import ricker

    roi, shots, nt = args.roi, args.shots, args.nt
    width = LAYER + 1
    n = roi + 2 * LAYER + 2
    dt = 0.4 * DX / 2500.0
    sim = scalar.AcousticWave((n, n), (DX, DX), nt, dt, precision="float32", space_order=8,
                              rho1=1.0, rho2=1.0, kappa1=1500.0**2, kappa2=2500.0**2,
                              pml=boundary.Pml(width, 2500.0, 1e-6))

    def indicator(v):
        full = np.pad(v, ((width, n - width - roi), (width, sim.Nx_padded[1] - width - roi)),
                      mode="edge")
        return sim.indicator_of_velocity(full)

    wavelet = (ricker(np.arange(nt) * dt, 1.0, FREQ) / (DX * DX))[:, None]
    cols = np.round(np.linspace(roi / (2 * shots), roi - roi / (2 * shots), shots)).astype(int)
    sources = [wave.Source(np.array([[width], [c + width]]), wavelet) for c in cols]
    every = max(roi // args.receivers, 1)
    coords = np.stack([np.full(roi // every, width * DX),
                       (np.arange(0, roi, every) + width) * DX], axis=1)
    sensors = utils.Sensors(sim, coords)
    sigma = roi / 32

    def pair(seed):
        true = learned.layered_model(seed, (roi, roi))
        return true, gaussian_filter(true, sigma)

    def note(text):
        print(f"[learned {time.perf_counter() - start:7.1f}s] {text}", file=sys.stderr,
              flush=True)

    start = t0 = time.perf_counter()
    samples = []
    for seed in range(args.train_models):
        t, s = pair(seed)
        samples.append(learned.training_sample(sim, sources, sensors, indicator(t),
                                               indicator(s), batch=args.shot_batch))
        if (seed + 1) % 16 == 0:
            note(f"{seed + 1} of {args.train_models} training samples")
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    torch.manual_seed(0)
    net = learned.UNet().cuda()
    t0 = time.perf_counter()
    every = max(args.epochs // 10, 1)
    losses = learned.train(net, samples, epochs=args.epochs, lr=1e-3, batch=16,
                           progress=lambda e, loss: e % every or note(f"epoch {e} loss {loss:.4g}"))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del samples
    held = {1000 + k: pair(1000 + k) for k in range(args.eval_models)}
    results = learned.evaluate(
        sim, sources, sensors, indicator, held, net, args.iterations, step=args.fwi_step,
        batch=args.shot_batch,
        progress=lambda r: note(f"model {r['seed']}: ratio {r['ratio']}, plain "
                                f"{r['plain']['history'][-1][1]:.4g}, learned "
                                f"{r['learned']['history'][-1][1]:.4g}"))
    ratios = [r["ratio"] for r in results]
    reached = [x for x in ratios if x is not None]
    return {
        "metric": "FWI time-to-misfit, learned-preconditioned vs plain (ratio of seconds)",
        "value": float(np.median([x if x is not None else np.inf for x in ratios])),
        "unit": f"learned seconds / plain seconds to the plain {args.iterations}-iteration "
                "misfit",
        "n_gpus": 1, "steps": args.iterations, "warmup": 0, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 solver, bf16 network",
        "data": "synthetic",
        "config": {"workload": "learned-preconditioned FWI vs plain normalised-gradient FWI",
                   "grid": f"{n}x{n} (ROI {roi}^2 + {LAYER}-cell PML)", "space_order": 8,
                   "shots": shots, "receivers": len(coords), "time_steps": nt,
                   "train_models": args.train_models, "eval_models": args.eval_models,
                   "iterations": args.iterations, "epochs": args.epochs,
                   "shots_a_sweep": min(args.shot_batch, shots)},
        "reached": f"{len(reached)} of {len(ratios)} held-out models",
        "per_model": [{"seed": r["seed"], "ratio": r["ratio"], "plain_s": r["plain_s"],
                       "learned_s": r["learned_s"],
                       "plain_misfit": r["plain"]["history"][-1][1],
                       "learned_misfit": r["learned"]["history"][-1][1],
                       "start_misfit": r["plain"]["history"][0][1],
                       "plain_model_error": r["plain"]["model_error"],
                       "learned_model_error": r["learned"]["model_error"],
                       "plain_history": [[round(t, 3), c] for t, c in r["plain"]["history"]],
                       "learned_history": [[round(t, 3), c] for t, c in r["learned"]["history"]]}
                      for r in results],
        "training": {"data_s": t_data, "train_s": t_train, "first_loss": losses[0],
                     "last_loss": losses[-1]},
    }


EL_ROI, EL_LAYER, EL_DX, EL_FREQ = (1024, 2048), 32, 5e-4, 5e5
EL_SHOTS, EL_RECEIVERS, EL_STEPS = 16, 512, 3000
EL_RHO0, EL_VP0 = 2200.0, 3000.0


def workload_elastic(shots=EL_SHOTS, nt=EL_STEPS, roi=EL_ROI):
    """The 2D elastic case in the reference's parametrization.

    Depth roi[0] on axis 0 under a traction-free top face, width roi[1] on
    axis 1, a PML of EL_LAYER cells on the other three faces (clamped walls,
    CFS shift 1e5 Hz), dx = 0.5 mm, order 4, float32. The reference's
    ElasticWave holds vp and vs fixed (vp0 = 3000 m/s, vs = vp / 1.8) and lets
    gamma scale density and both Lame parameters alike, so the model's 8
    layers (boundaries at sorted seed-6 depths) are density layers, gamma =
    rho / rho0 with rho ~ U[1800, 2600] and rho0 = 2200, and the 10 % anomaly
    a seed-7 rectangle of gamma x 1.1; the start model has no anomaly. 16
    vertical forces (Ricker, 500 kHz) on the free surface at ROI columns
    linspace(64, 1984), 512 receivers of the vertical displacement every 4
    columns on the same row.
    """
    from cuwave_amd import boundary, elastic, wave
    from cuwave_amd.signals import ricker

    width = EL_LAYER + 1  # the wall node is part of the layer
    n0, n1 = roi[0] + EL_LAYER + 2, roi[1] + 2 * EL_LAYER + 2
    dt = 0.4 * EL_DX / EL_VP0
    sim = elastic.ElasticWave(
        (n0, n1), (EL_DX, EL_DX), nt, dt, precision="float32", space_order=4,
        density=EL_RHO0, wavespeed_p=EL_VP0, wavespeed_s=EL_VP0 / 1.8,
        boundary=((boundary.Traction, boundary.Clamped), (boundary.Clamped, boundary.Clamped)),
        pml=boundary.Pml(width, EL_VP0, 1e-6, 1e5, (1, 2, 3)))
    rng = np.random.default_rng(6)
    tops = np.sort(rng.integers(1, roi[0], 7))
    rho = rng.uniform(1800.0, 2600.0, 8)
    depth = np.arange(sim.Nx_padded[0]) - 1
    layer = np.searchsorted(tops, depth, side="right")
    start = np.repeat((rho[layer] / EL_RHO0)[:, None], sim.Nx_padded[1], axis=1)
    box = np.random.default_rng(7)
    r0 = int(box.integers(roi[0] // 4, roi[0] // 2))
    c0 = int(box.integers(roi[1] // 4, roi[1] // 2))
    true = start.copy()
    true[1 + r0:1 + r0 + roi[0] // 8, width + c0:width + c0 + roi[1] // 8] *= 1.1
    t = np.arange(nt) * dt
    wavelet = (ricker(t, 1.0, EL_FREQ) / (EL_DX * EL_DX))[:, None]
    cols = np.round(np.linspace(64, roi[1] - 64, shots)).astype(int) + width
    sources = [wave.Source(np.array([[0], [1], [c]]), wavelet) for c in cols]
    rec = np.arange(0, roi[1], max(roi[1] // EL_RECEIVERS, 1))[:EL_RECEIVERS] + width
    sensors = np.stack([np.zeros(len(rec), int), np.ones(len(rec), int), rec])
    return sim, true.astype(np.float32), start.astype(np.float32), sources, sensors


class _NodeReceivers:
    """Receivers on unknowns: the record is the traces."""

    def __init__(self, nodes):
        self.nodes = nodes

    def objective(self, objective):
        return objective

    def traces(self, record):
        return record


def gpu_run_elastic(args):
    """The elastic FWI gradient: every shot's forward, the L2 misfit, the reverse sweep."""
    import torch

    from cuwave_amd import elastic, utils, wave

    sim, true, start, sources, sensors = workload_elastic(args.shots, args.nt)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    B = len(sources)
    _, observed = wave.simulate_shots(sim, sources, true, sensors)
    observed = list(observed)
    rec = _NodeReceivers(sensors)

    def step():
        return utils.misfit_gradient(sim, sources, start, rec, observed,
                                     adjoint=utils.reconstruction_sensitivity)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    elastic._Reconstruction.timings = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cost, gradient = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timings, elastic._Reconstruction.timings = elastic._Reconstruction.timings, None
    per = {}
    for kind, steps, s, e in timings:
        tot = per.setdefault(kind, [0.0, 0])
        tot[0] += s.elapsed_time(e)
        tot[1] += steps
    bwd_ms = per["backward"][0] / per["backward"][1]
    fwd_ms = per["forward"][0] / per["forward"][1]
    # a forward step: u1 (2 components) read, u0 read and u2 written over it,
    # gamma and the prepared material once a node; an adjoint level: the
    # reverse and adjoint steps, the triplet and lambda read by the gradient,
    # three accumulators read and written
    bytes_fwd = (24.0 * B + 12.0) * n_int
    bytes_bwd = (48.0 * B + 36.0) * n_int
    achieved = bytes_bwd / (bwd_ms * 1e-3)
    return {
        "value": 2.0 * B * n_int * sim.N * args.steps / elapsed / 1e9,
        "ms_per_step": elapsed / args.steps * 1e3,
        "cost": cost, "B": B, "n_int": n_int, "sim": sim,
        "grad_norm": float(torch.linalg.vector_norm(gradient)),
        "kernels": {"backward_ms": bwd_ms, "forward_ms": fwd_ms,
                    "backward_GBps": achieved / 1e9,
                    "forward_GBps": bytes_fwd / (fwd_ms * 1e-3) / 1e9},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic_per_launch("traffic_adjoint_level_elastic2d_b16.json")},
    }


def cpu_baseline_elastic(sample_steps: int, runs: int = 3):
    """The elastic oracle on the host cores: one shot, the full grid, the first
    `sample_steps` steps of forward + reconstruction adjoint; the median of
    `runs` timed runs after one warm-up run."""
    from oracle import driver as D
    from oracle import elastic as E

    sim, true, start, sources, sensors = workload_elastic(1, sample_steps)
    osim = E.ElSim(sim.Nx, sim.dx, sim.N, sim.dt, sim.precision, sim.space_order, sim.stencil,
                   ((0, 1), (1, 1)), sim.density, sim.wavespeed_p, sim.wavespeed_s, "strain",
                   D.Pml(sim.pml.width, sim.pml.wavespeed, sim.pml.reflection,
                         sim.pml.frequency, sim.pml.faces))
    src = sources[0]
    _, observed = E.simulate(osim, src.position, src.signal, true, sensors)
    times = []
    for _ in range(runs + 1):
        t0 = time.perf_counter()
        E.reconstruction_sensitivity(osim, src.position, src.signal, start, sensors,
                                     D.l2_misfit(observed))
        times.append(time.perf_counter() - t0)
    median = float(np.median(times[1:]))
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    return 2.0 * n_int * sample_steps / median / 1e9, median


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--shots", type=int, default=SHOTS)
    p.add_argument("--nt", type=int, default=STEPS)
    p.add_argument("--roi", type=int, default=ROI)
    p.add_argument("--cpu-sample", type=int, default=300,
                   help="oracle steps of the CPU baseline's sample (median of 3 after a warm-up)")
    p.add_argument("--no-checks", action="store_true",
                   help="skip the one-shot-per-launch rate, the drift and the float64 error")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--probe-copy", action="store_true",
                   help="run only a 256 MiB device copy (counter calibration)")
    p.add_argument("--train-models", type=int, default=256)
    p.add_argument("--eval-models", type=int, default=16)
    p.add_argument("--iterations", type=int, default=30)
    p.add_argument("--receivers", type=int, default=128)
    p.add_argument("--fwi-step", type=float, default=0.2,
                   help="the first FWI step, the largest indicator change")
    p.add_argument("--shot-batch", type=int, default=16,
                   help="learned: shots a sweep (24 shots as 16 + 8 leave no lane padding)")
    p.add_argument("--epochs", type=int, default=100)
    p.add_argument("--workload",
                   choices=("2d", "3d", "elastic", "forward", "learned", "maxwell", "aniso"),
                   default="2d",
                   help="3d: the 3D acoustic gradient on a 354^3 grid instead (its own line); "
                        "elastic: the 2D elastic gradient on 1058 x 2114, 16 shots; "
                        "forward: the plain 298 x 298 forward, one shot; "
                        "maxwell: the 3D vector Maxwell forward on 162^3, 8 shots; "
                        "aniso: the 2D cell-assembled elastic forward on 1026^2, 16 shots")
    args = p.parse_args()
    if args.probe_copy:
        probe_copy()
        return
    if args.workload == "learned":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        # config 5's sizes where the 2D line's defaults were left
        args.roi = 512 if args.roi == ROI else args.roi
        args.shots = 24 if args.shots == SHOTS else args.shots
        args.nt = 2000 if args.nt == STEPS else args.nt
        print(json.dumps(learned_line(args)))
        return
    if args.workload == "forward":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        print(json.dumps(forward_line(args)))
        return
    if args.workload == "aniso":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        print(json.dumps(aniso_line(args)))
        return
    if args.workload == "maxwell":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        print(json.dumps(maxwell_line(args)))
        return
    if args.workload == "elastic":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        print(json.dumps(elastic_line(args)))
        return
    if args.workload == "3d":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU"}))
            sys.exit(2)
        args.nt = 1500 if args.nt == STEPS else args.nt  # the 3D case's own step count
        res = gpu_run_3d(args)
        if not args.no_cpu:
            cores = len(os.sched_getaffinity(0))
            sample = 8
            value, seconds = cpu_baseline_3d(sample)
            res["cpu_baseline"] = {
                "value": value, "unit": "Gcell-updates/s",
                "cores": min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))),
                "kind": "port",
                "sample": f"the full 354^3 grid, first {sample} of {args.nt} steps, forward + "
                          f"reconstruction adjoint through the C oracle, its strip set-up "
                          f"included; median of 3 runs after a warm-up ({seconds:.1f} s a run)"}
        print(json.dumps(res))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the shot batch shards over "
                                   "GPUs by design only; run with --gpus 1"}))
        sys.exit(2)
    res = gpu_run(args)
    checks = None if args.no_checks else one_shot_checks(args)
    cpu = None
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        value, seconds = cpu_baseline(args.cpu_sample)
        cpu = {"value": value, "unit": "Gcell-updates/s", "cores": min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))),
               "kind": "port",
               "sample": f"1 shot, full 1090x1090 grid, first {args.cpu_sample} of {STEPS} steps, "
                         f"forward + reconstruction adjoint through the C oracle; median of 3 "
                         f"runs after a warm-up ({seconds:.1f} s a run)"}
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
            "workload": "2D scalar FWI gradient wrt velocity: velocity -> indicator -> traces "
                        "(torch.autograd.Function) -> L2 -> backward; forward + reconstruction "
                        "adjoint, shot-batched",
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
    if checks is not None:
        line["one_shot_per_launch"] = {
            "value": checks["b1_value"], "unit": "Gcell-updates/s",
            "sample": f"{checks['b1_shots']} shots one after another, each its own launches",
            "speedup": res["value"] / checks["b1_value"]}
        line["accuracy"] = {"fp32_vs_fp64_gradient_rel_l2": checks["fp64_rel_l2"],
                            "reconstruction_drift": checks["drift"],
                            "sample": "the first shot, the same scheme in float32 and float64"}
    print(json.dumps(line))


def elastic_line(args) -> dict:
    shots = args.shots if args.shots != SHOTS else EL_SHOTS
    nt = args.nt if args.nt != STEPS else EL_STEPS
    args.shots, args.nt = shots, nt
    res = gpu_run_elastic(args)
    cpu = None
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        sample = 20
        value, seconds = cpu_baseline_elastic(sample)
        cpu = {"value": value, "unit": "Gcell-updates/s",
               "cores": min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))),
               "kind": "port",
               "sample": f"1 shot, full {res['sim'].Nx[0]}x{res['sim'].Nx[1]} grid, the first "
                         f"{sample} steps, forward + reconstruction adjoint through the C oracle; "
                         f"median of 3 runs after a warm-up ({seconds:.1f} s a run)"}
    sim = res["sim"]
    return {
        "metric": "Gcell-updates/s (shot-batched 2D elastic forward+adjoint)",
        "value": res["value"], "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D elastic FWI gradient: forward + reconstruction adjoint, "
                               "shot-batched, gamma-layered (the reference's parametrization)",
                   "grid": f"{sim.Nx[0]}x{sim.Nx[1]} (ROI 1024x2048, free top, 32-cell PML)",
                   "space_order": 4, "shots": res["B"], "receivers": EL_RECEIVERS,
                   "time_steps": nt, "interior_nodes": res["n_int"],
                   "parallelism": "single GPU, shots batched per launch"},
        "roofline": res["roofline"], "cpu_baseline": cpu, "kernels": res["kernels"],
    }


MX_N, MX_SHOTS, MX_STEPS = 162, 8, 200


class _Traces:
    """Receivers on unknowns for `misfit_gradient`: the record is the traces."""

    def __init__(self, nodes):
        self.nodes = nodes

    def objective(self, objective):
        return objective


def gradient_rate(sim, indicator, sources, sensors) -> dict:
    """The misfit gradient of every shot in one batch against zero data (one forward
    and one reverse sweep and the host chain rule), timed whole: W = 2 B n_int N."""
    import torch

    from cuwave_amd import utils

    observed = [np.zeros((sim.N, sensors.shape[1]), dtype=sim.np_dtype)] * len(sources)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    utils.misfit_gradient(sim, sources, indicator, _Traces(sensors), observed)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    utils.misfit_gradient(sim, sources, indicator, _Traces(sensors), observed)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    return {"value": 2 * n_int * len(sources) * sim.N / seconds / 1e9,
            "unit": "Gcell-updates/s", "seconds": seconds,
            "note": "misfit_gradient of every shot in one batch, the field rebuilt by the "
                    "reverse march; forward and adjoint updates counted, host chain rule "
                    "included"}


def workload_maxwell(nt: int = MX_STEPS, shots: int = MX_SHOTS, n: int = MX_N):
    """The 3D vector Maxwell forward: a conducting box of n^3 nodes (n - 2 interior a
    side), dx = 1 / (n - 3), order 4, float32; eps 1 with an eps 4 sphere of radius
    n / 5 at the centre; `shots` z dipoles (a Ricker current of 4 periods across the
    run) on a ring in the mid plane, 16 probes of every component on a line;
    dt = 0.4 dx / sqrt(3)."""
    from cuwave_amd import maxwell, wave
    from cuwave_amd.signals import ricker

    dx = 1.0 / (n - 3)
    dt = 0.4 * dx / np.sqrt(3.0)
    sim = maxwell.DielectricWave((n,) * 3, (dx,) * 3, nt, dt, precision="float32",
                                 space_order=4, permittivity1=1.0, permittivity2=4.0)
    a = np.meshgrid(*(np.arange(m) for m in sim.Nx_padded), indexing="ij")
    sphere = sum((x - n // 2) ** 2 for x in a) <= (n // 5) ** 2
    indicator = sphere.astype(np.float32)
    wavelet = ricker(np.arange(nt) * dt, 1.0, 4.0 / (nt * dt))
    sources = []
    for b in range(shots):
        angle = 2 * np.pi * b / shots
        at = [n // 2 + int(0.35 * n * np.cos(angle)), n // 2 + int(0.35 * n * np.sin(angle))]
        sources.append(wave.Source(np.array([[2], [n // 2], [at[0]], [at[1]]]), wavelet[:, None]))
    line = np.round(np.linspace(4, n - 5, 16)).astype(int)
    sensors = np.concatenate([np.stack([np.full(16, c), np.full(16, n // 3), line,
                                        np.full(16, n // 2)]) for c in range(3)], axis=1)
    return sim, indicator, sources, sensors


def maxwell_line(args) -> dict:
    """Gcell-updates/s of the 3D vector Maxwell forward (interior nodes x shots x steps),
    the steps' launches timed with HIP events on the library's stream."""
    import torch

    from cuwave_amd import maxwell

    nt = args.nt if args.nt != STEPS else MX_STEPS
    shots = args.shots if args.shots != SHOTS else MX_SHOTS
    sim, indicator, sources, sensors = workload_maxwell(nt, shots)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    mat = sim.build_materials(indicator)
    eng = maxwell.MaxwellEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    U = eng.levels(2)
    stream = torch.cuda.current_stream()  # the stream the library launches on

    def run():
        U.zero_()
        for t in range(sim.N):
            u = eng.step(U[t % 2], U[1 - t % 2], U[t % 2])
            src.inject(u, signal, t)
            rec.record(u, record, t)

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        run()
    stop.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_ms = start.elapsed_time(stop) / (args.steps * sim.N)
    # a step's algorithmic bytes: the curl reads u1's 3 components and writes 3 pair
    # curls, the update reads the 3 curls, u0 and u1 and writes u2, per lane; the
    # point inverse inertia (3 a node) once for every lane
    per_node = (18 * eng.lanes + 3) * 4
    achieved = per_node * n_int / (step_ms * 1e-3)
    cpu = None
    if not args.no_cpu:
        from oracle import maxwell as M
        from cuwave_amd.stencils import staggered_weights

        sample = 4
        inner = (slice(None), *(slice(0, n) for n in sim.Nx))
        minv = mat["minv"][inner]
        pos = np.asarray(sources[0].position)
        weight = sim.excitation_weights(mat, sim.flatten_indices(pos))
        times = []
        for _ in range(3):
            c0 = time.perf_counter()
            M.simulate(minv, staggered_weights(2, sim.stencil).astype(np.float32),
                       [np.float32(1.0 / d) for d in sim.dx], sim.step_factors(),
                       np.float32(1.0), sim.abi_grid(1).conductor, sample, tuple(pos), weight,
                       np.asarray(sources[0].signal, dtype=np.float32), tuple(sensors))
            times.append(time.perf_counter() - c0)
        med = float(np.median(times))
        cpu = {"value": n_int * sample / med / 1e9, "unit": "Gcell-updates/s", "cores": 1,
               "kind": "port",
               "sample": f"1 shot, the full {sim.Nx[0]}^3 grid, the first {sample} steps through "
                         f"the numpy oracle; median of 3 runs ({med:.1f} s a run)"}
    return {
        "metric": "Gcell-updates/s (shot-batched 3D vector Maxwell forward)",
        "value": n_int * shots * nt * args.steps / elapsed / 1e9, "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3D vector Maxwell forward: traces of every shot, a conducting "
                               "box with a dielectric sphere",
                   "grid": f"{sim.Nx[0]}^3", "space_order": 4, "shots": shots,
                   "receivers": int(sensors.shape[1]), "time_steps": nt, "interior_nodes": n_int,
                   "parallelism": "single GPU, shots batched per launch"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
        "kernels": {"step_ms": step_ms, "bytes_per_node": per_node,
                    "note": "a step's launches (curl, update, sources, probes) on the stream"},
        "gradient": gradient_rate(sim, indicator, sources, sensors),
    }


AN_N, AN_SHOTS, AN_STEPS = 1026, 16, 200


def workload_aniso(nt: int = AN_STEPS, shots: int = AN_SHOTS, n: int = AN_N):
    """The 2D cell-assembled elastic forward: n^2 nodes, dx = 1 m, float32, traction
    faces but a clamped bottom; rho0 2000, vp 3000, vs 1500 with an orthotropic C
    (C11 2 rho0 vp^2, C22 rho0 vp^2, C12 rho0 (vp^2 - 2 vs^2), C66 rho0 vs^2); gamma
    1 with a gamma 1.5 disk of radius n / 8; `shots` vertical forces (Ricker, 4
    periods across the run) along row n / 8, 32 receivers of both components on row
    2; dt = 0.25 dx / (sqrt(2) c), c = sqrt(2) vp the fastest speed C carries."""
    from cuwave_amd import anisotropic, wave
    from cuwave_amd.boundary import Clamped, Traction
    from cuwave_amd.signals import ricker

    rho, vp, vs = 2000.0, 3000.0, 1500.0
    C = rho * np.array([[2 * vp**2, vp**2 - 2 * vs**2, 0.0],
                        [vp**2 - 2 * vs**2, vp**2, 0.0], [0.0, 0.0, vs**2]])
    dt = 0.25 / (np.sqrt(2.0) * np.sqrt(2.0) * vp)
    sim = anisotropic.AnisotropicElasticWave(
        (n, n), (1.0, 1.0), nt, dt, precision="float32", density=rho, wavespeed_p=vp,
        wavespeed_s=vs, C=C, boundary=((Traction, Clamped), (Traction, Traction)))
    a0, a1 = np.meshgrid(*(np.arange(m) for m in sim.Nx_padded), indexing="ij")
    disk = (a0 - n // 2) ** 2 + (a1 - n // 2) ** 2 <= (n // 8) ** 2
    indicator = np.where(disk, 1.5, 1.0).astype(np.float32)
    wavelet = ricker(np.arange(nt) * dt, 1.0, 4.0 / (nt * dt))
    cols = np.round(np.linspace(8, n - 9, shots)).astype(int)
    sources = [wave.Source(np.array([[0], [n // 8], [c]]), wavelet[:, None]) for c in cols]
    line = np.round(np.linspace(4, n - 5, 32)).astype(int)
    sensors = np.concatenate([np.stack([np.full(32, c), np.full(32, 2), line]) for c in range(2)],
                             axis=1)
    return sim, indicator, sources, sensors


def aniso_line(args) -> dict:
    """Gcell-updates/s of the 2D cell-assembled elastic forward (interior nodes x shots x
    steps), the steps' launches timed with HIP events on the library's stream."""
    import torch

    from cuwave_amd import anisotropic

    nt = args.nt if args.nt != STEPS else AN_STEPS
    shots = args.shots if args.shots != SHOTS else AN_SHOTS
    sim, indicator, sources, sensors = workload_aniso(nt, shots)
    n_int = int(np.prod([n - 2 for n in sim.Nx]))
    mat = sim.build_materials(indicator)
    eng = anisotropic.AnisoEngine(sim, mat, len(sources))
    src, signal = eng.sources(mat, sources)
    rec = eng.sensors(sensors)
    record = torch.zeros((sim.N, rec.count, eng.lanes), dtype=sim.dtype, device=eng.device)
    U = eng.levels(2)
    stream = torch.cuda.current_stream()  # the stream the library launches on

    def run():
        U.zero_()
        for t in range(sim.N):
            u = eng.step(U[t % 2], U[1 - t % 2], U[t % 2])
            src.inject(u, signal, t)
            rec.record(u, record, t)

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        run()
    stop.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_ms = start.elapsed_time(stop) / (args.steps * sim.N)
    # a step's algorithmic bytes: u0 and u1 read and u2 written, 2 components a lane;
    # the inverse inertia and the cell material once for every lane
    per_node = (6 * eng.lanes + 2) * 4
    achieved = per_node * n_int / (step_ms * 1e-3)
    cpu = None
    if not args.no_cpu:
        from oracle import anisotropic as O

        sample = 16
        inner = tuple(slice(0, n) for n in sim.Nx)
        pos = np.asarray(sources[0].position)
        weight = (np.float32(sim.dt**2) * mat["minv"][inner][tuple(pos[1:])]).astype(np.float32)
        times = []
        for _ in range(3):
            c0 = time.perf_counter()
            O.simulate(mat["minv"][inner], mat["cell"][inner], mat["table"],
                       np.float32(sim.dt**2), sample, tuple(pos), weight,
                       np.asarray(sources[0].signal, dtype=np.float32), tuple(sensors))
            times.append(time.perf_counter() - c0)
        med = float(np.median(times))
        cpu = {"value": n_int * sample / med / 1e9, "unit": "Gcell-updates/s", "cores": 1,
               "kind": "port",
               "sample": f"1 shot, the full {sim.Nx[0]}^2 grid, the first {sample} steps through "
                         f"the numpy oracle; median of 3 runs ({med:.1f} s a run)"}
    return {
        "metric": "Gcell-updates/s (shot-batched 2D cell-assembled elastic forward)",
        "value": n_int * shots * nt * args.steps / elapsed / 1e9, "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D anisotropic elastic forward: traces of every shot, an "
                               "orthotropic medium with a denser disk",
                   "grid": f"{sim.Nx[0]}^2", "space_order": 2, "shots": shots,
                   "receivers": int(sensors.shape[1]), "time_steps": nt, "interior_nodes": n_int,
                   "parallelism": "single GPU, shots batched per launch"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
        "kernels": {"step_ms": step_ms, "bytes_per_node": per_node,
                    "note": "a step's launches (the step, sources, probes) on the stream"},
        "gradient": gradient_rate(sim, indicator, sources, sensors),
    }


if __name__ == "__main__":
    main()
