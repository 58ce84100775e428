"""Forward marches replayed from one captured graph window, and the tuner.

`cw_scalar_forward_clocked` takes its step index from a device counter, so a
window of steps captured once replays for every window of a march
(`Propagator.march`). The march must give the bits of one plain call: in 1D,
2D (CPML, three shots, sources, sensors, the illumination sum, core blocks
marched) and 3D, for marches of one, several and many replays with a
remainder. Each march captures its own graph. Plain calls run between
the graphed ones, so every replay runs after other work and is checked
against the plain result. A clock at or past its end launches steps that do
nothing. The tuner's measured choices go to its file, come back in a fresh
read, and leave the simulated bits unchanged. Adjoint levels whose frame
launches fork onto the side stream and join back capture into one graph, and
every replay from the same state gives the eager call's bits.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from helpers import golden, product_sim

pytestmark = pytest.mark.gpu

from cuwave_amd import _abi, tuning, wave  # noqa: E402
from cuwave_amd.engine import Propagator, sensor_points, source_points  # noqa: E402


def _case(name, N, Nx=None, shots=1, seed=0):
    if name == "1d":
        from cuwave_amd import scalar

        dx = (1e-3,)
        sim = scalar.ScalarWave(Nx or (400,), dx, N, 0.5 * wave.stable_dt(dx, 1500.0, 2),
                                precision="float64", wavespeed=1500.0, density=1.0)
        sig0 = np.sin(np.arange(N) * 0.3)[:, None] * np.exp(-np.arange(N) / 20.0)[:, None]
    else:
        base = golden(name)
        z = {k: base[k] for k in base.files}
        z["N"] = np.array(N)
        if Nx is not None:
            z["Nx"] = np.array(Nx)
        sim = product_sim(z)
        sig0 = base["src_sig"][:, :1]
    rng = np.random.default_rng(seed)
    indicator = (0.3 + 0.4 * rng.random(sim.indicator_shape())).astype(sim.np_dtype)
    sig = np.resize(sig0, (N, 1)).astype(sim.np_dtype)
    sources = []
    for k in range(shots):
        pos = np.array([[int(rng.integers(3, n - 3))] for n in sim.Nx])
        sources.append(wave.Source(pos, sig * (1 + k)))
    sensors = np.array([[int(rng.integers(3, n - 3)) for _ in range(5)] for n in sim.Nx])
    return sim, indicator, sources, sensors


def _sweep(sim, indicator, sources, sensors, window, illum=True):
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, len(sources))
    src, signal = source_points(prop, mat, sources)
    rec = sensor_points(prop, sensors)
    record = prop.record_memory(rec)
    U = prop.levels_memory(2)
    acc = prop.accumulator() if illum else None
    kw = dict(aux=prop.aux_memory(), sources=src, signal=signal, sensors=rec, record=record,
              illum=acc)
    if window:
        prop.march(U, 0, sim.N, window, **kw)
    else:
        prop.forward(U, 0, sim.N, **kw)
    torch.cuda.synchronize()
    return [t.clone() for t in (U, record, acc) if t is not None]


@pytest.fixture
def _march_every_box():
    previous = _abi.call_value("cw_set_stream_threshold", 0)
    yield
    _abi.call_value("cw_set_stream_threshold", previous)


@pytest.mark.parametrize("name,Nx,shots", [
    ("fwd_pml_acoustic_o8_f32", (134, 200), 3),
    ("fwd_pml_acoustic_o8_f32", None, 1),
    ("fwd_pml_3d_o4_f32", None, 2),
    ("fwd_dirichlet_o4_f64", None, 2),
    ("1d", None, 1),
])
def test_a_graphed_march_is_bitwise_one_plain_call(name, Nx, shots, _march_every_box):
    for N, window in ((72, 24), (77, 24), (131, 12), (100, 48)):
        if not name.startswith("fwd_pml_acoustic") and N != 77:
            continue
        sim, indicator, sources, sensors = _case(name, N, Nx, shots)
        print(f"case graphed march {name} {Nx} N={N} window={window}", flush=True)
        plain = _sweep(sim, indicator, sources, sensors, 0)
        graphed = _sweep(sim, indicator, sources, sensors, window)
        again = _sweep(sim, indicator, sources, sensors, 0)
        for a, b, c in zip(plain, graphed, again):
            assert torch.equal(a, b) and torch.equal(a, c), (name, N, window)


def test_a_clock_past_its_end_steps_nothing():
    sim, indicator, sources, sensors = _case("fwd_pml_acoustic_o8_f32", 60)
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, 1)
    src, signal = source_points(prop, mat, sources)
    U = prop.levels_memory(2)
    a = prop.forward_args(U, aux=prop.aux_memory(), sources=src, signal=signal)
    import ctypes

    prop.forward(U, 0, 40, aux=prop.aux_memory(), sources=src, signal=signal)
    for start in (sim.N, sim.N + 7):
        clock = torch.full((1,), start, dtype=torch.int32, device=prop.device)
        before = U.clone()
        _abi.call("cw_scalar_forward_clocked", ctypes.byref(a), _abi.ptr(clock), 5, sim.N,
                  _abi.stream())
        torch.cuda.synchronize()
        assert torch.equal(before, U)
        assert int(clock.item()) == start + 5
    with pytest.raises(RuntimeError, match="clock"):
        _abi.call("cw_scalar_forward_clocked", ctypes.byref(a), None, 5, sim.N, _abi.stream())


def test_tuned_choices_are_kept_and_change_no_bits(tmp_path, monkeypatch):
    path = tmp_path / "tuning.json"
    monkeypatch.setenv("CUWAVE_AMD_TUNING", str(path))
    tuning.forget()
    sim, indicator, sources, sensors = _case("fwd_pml_acoustic_o8_f32", 160, (134, 200), 2)
    print("case tuner 134x200 two lanes", flush=True)
    before = wave.simulate_shots(sim, sources, indicator, sensors)
    result = tuning.tune(sim, lanes=2, indicator=indicator, repeats=2)
    assert result["stream_threshold"] in tuning.THRESHOLDS
    assert result["graph_window"] in tuning.WINDOWS
    assert len(result["timings"]) >= 9 and result["ms_per_step"] > 0
    stored = json.loads(path.read_text())
    assert stored[tuning.key(sim, 2)]["graph_window"] == result["graph_window"]
    tuning.forget()  # a later process reads the file back
    assert tuning.choice(sim, 2, "stream_threshold") == result["stream_threshold"]
    assert wave.graph_window(sim, 2) == result["graph_window"]
    after = wave.simulate_shots(sim, sources, indicator, sensors)
    for a, b in zip(before, after):
        assert torch.equal(a, b)
    tuning.forget()


@pytest.mark.parametrize("combine", [True, False])
def test_captured_adjoint_levels_replay_the_eager_ones(combine, _march_every_box):
    """Adjoint levels whose frame launches fork onto the side stream and join back
    (two streams, events between them) captured into one graph: every replay, from
    the same starting state and with eager work between replays, gives the bits of
    the eager call."""
    sim, indicator, sources, sensors = _case("fwd_pml_acoustic_o8_f32", 40, (134, 200), 3)
    print(f"case captured adjoint levels combine={combine}", flush=True)
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, len(sources))
    gen = torch.Generator(device=prop.device).manual_seed(0)

    def noise(t):
        return t.copy_(torch.randn(t.shape, generator=gen, device=t.device, dtype=t.dtype))

    rec = sensor_points(prop, sensors,
                        weight=sim.excitation_weights(mat, sim.flatten_indices(sensors)))
    adjoint = noise(prop.record_memory(rec))
    state = {"P": noise(prop.levels_memory(2)), "U": noise(prop.levels_memory(2)),
             "gm": noise(prop.accumulator()), "gs": noise(prop.accumulator())}
    for k, e in enumerate(prop.eta_memory()):
        state[f"eta{k}"] = noise(e)
    start = {k: v.clone() for k, v in state.items()}
    eta = [state[f"eta{k}"] for k in range(len(prop.slabs))]
    kw = dict(eta=eta, g_mass=state["gm"], sensors=rec, adjoint=adjoint)
    kw.update(combine=(0.7, -0.3)) if combine else kw.update(g_stiff=state["gs"])

    def levels():
        prop.backward(state["P"], state["U"], sim.N - 9, sim.N - 1, **kw)

    def restore():
        for k, v in start.items():
            state[k].copy_(v)

    levels()
    torch.cuda.synchronize()
    eager = {k: v.clone() for k, v in state.items()}
    restore()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        levels()
    torch.cuda.synchronize()
    for replay in range(3):
        restore()
        noise(prop.levels_memory(2))  # other work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for k, v in eager.items():
            assert torch.equal(state[k], v), (replay, k)
