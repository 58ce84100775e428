"""Measured choices of how a sweep is launched, cached on disk.

The kernels' tile shapes are compile-time constants. What a sweep can still
choose is how its steps are launched:

- `stream_threshold`: the fewest blocks a 2D stream box must hold for the
  stream kernels to march it (`cw_set_stream_threshold`); below it every node
  runs in the general kernels;
- `graph_window`: the steps of one captured graph window that a forward march
  replays (`Propagator.march`), 0 for plain launches.

`tune(sim, lanes)` times a short forward march for each candidate pair on the
current GPU and keeps the fastest. Its key is the equation family, dimension,
order, precision, padded shape, lanes and device name. The choices go to a
JSON file (`CUWAVE_AMD_TUNING`, default `~/.cache/cuwave_amd/tuning.json`),
which later processes read back. Every candidate gives the same bits; only the
time differs. The reference picks its launch tiers from device attributes and
hand tables (cuwave/scalar.py:422-737); here they are measured.
"""

from __future__ import annotations

import contextlib
import json
import os
import threading
from pathlib import Path

import numpy as np
import torch

from . import _abi

THRESHOLDS = (0, 512, 1 << 30)  # march every box, the default, never march
WINDOWS = (0, 12, 24, 48)
DEFAULT_THRESHOLD = 512

_lock = threading.Lock()
_cache: dict | None = None


def cache_path() -> Path:
    env = os.environ.get("CUWAVE_AMD_TUNING")
    return Path(env) if env else Path.home() / ".cache" / "cuwave_amd" / "tuning.json"


def key(sim, lanes: int) -> str:
    """The sweep shape the choices are kept under."""
    dev = torch.cuda.get_device_name() if torch.cuda.is_available() else "cpu"
    shape = "x".join(map(str, sim.Nx_padded))
    return (f"{type(sim).__name__}/{sim.ndim}d/o{sim.space_order}/{sim.precision}/"
            f"{shape}/b{lanes}/{dev}")


def _load() -> dict:
    global _cache
    if _cache is None:
        path = cache_path()
        try:
            _cache = json.loads(path.read_text()) if path.exists() else {}
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _save() -> None:
    path = cache_path()
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(_cache, indent=1, sort_keys=True))
    tmp.replace(path)


def forget() -> None:
    """Drop the choices read into this process (the file stays)."""
    global _cache
    with _lock:
        _cache = None


def choice(sim, lanes: int, knob: str):
    """The tuned value of `knob` for this sweep, or None when it was never tuned."""
    with _lock:
        entry = _load().get(key(sim, lanes))
    return None if entry is None else entry.get(knob)


@contextlib.contextmanager
def applied(sim, lanes: int):
    """Run the enclosed launches with the sweep's tuned stream threshold (the
    process-wide setting is put back after)."""
    value = choice(sim, lanes, "stream_threshold")
    if value is None:
        yield
        return
    previous = _abi.call_value("cw_set_stream_threshold", int(value))
    try:
        yield
    finally:
        _abi.call_value("cw_set_stream_threshold", int(previous))


def tune(sim, lanes: int = 1, indicator=None, steps: int | None = None, repeats: int = 3,
         thresholds=THRESHOLDS, windows=WINDOWS, save: bool = True) -> dict:
    """Time a forward march of each (stream threshold, graph window) pair and keep the
    fastest for this sweep; returns {"stream_threshold", "graph_window", "ms_per_step",
    "timings"}. The march runs `steps` steps (at most sim.N; by default enough for
    three windows of the largest candidate), source-free, best of `repeats`."""
    from .engine import Propagator

    if sim.ncomp > 1:
        raise NotImplementedError("the tuner covers the pressure family's sweeps")
    if indicator is None:
        indicator = np.ones(sim.indicator_shape(), dtype=sim.np_dtype)
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, lanes)
    if prop.lanes != lanes:
        raise ValueError(f"lanes must be a power of two up to 64, not {lanes}")
    steps = min(sim.N, steps or 3 * max(windows) + 8)
    U = prop.levels_memory(2)
    aux = prop.aux_memory()
    timings = {}
    previous = _abi.call_value("cw_set_stream_threshold", DEFAULT_THRESHOLD)
    try:
        for threshold in thresholds:
            _abi.call_value("cw_set_stream_threshold", int(threshold))
            for window in windows:
                if window and steps < 3 * window:
                    continue
                best = float("inf")
                for _ in range(repeats + 1):  # the first run warms
                    start = torch.cuda.Event(enable_timing=True)
                    end = torch.cuda.Event(enable_timing=True)
                    torch.cuda.synchronize()
                    start.record()
                    prop.march(U, 0, steps, window, aux=aux)
                    end.record()
                    end.synchronize()
                    best = min(best, start.elapsed_time(end) / steps)
                timings[f"{threshold}/{window}"] = best
    finally:
        _abi.call_value("cw_set_stream_threshold", int(previous))
    threshold, window = min(timings, key=timings.get).split("/")
    result = {"stream_threshold": int(threshold), "graph_window": int(window),
              "ms_per_step": timings[f"{threshold}/{window}"], "timings": timings}
    with _lock:
        _load()[key(sim, lanes)] = {k: result[k] for k in ("stream_threshold", "graph_window",
                                                          "ms_per_step")}
        if save:
            _save()
    return result
