"""The allocation-free forward (cw_scalar_forward_static) against cw_scalar_forward.

The static entry point takes its CPML memories from the library's own scratch
array instead of the caller's `pml.aux` buffers. Called through the C ABI on
the same inputs it must give the same bits as the allocating entry point:
twice in a row, the second call on a smaller grid with fewer shots (so a
scratch that was not cleared, or laid out for the first call's sizes, shows),
then one sweep split over two calls (the second continues the scratch). A
sweep whose memories pass the scratch is refused before any launch.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from helpers import golden, product_sim

pytestmark = pytest.mark.gpu

from cuwave_amd import _abi, wave  # noqa: E402
from cuwave_amd.engine import Propagator, sensor_points, source_points  # noqa: E402


def _case(Nx, shots, N, seed):
    """An eighth-order acoustic CPML case (a golden case's physics) on a grid of Nx."""
    base = golden("adj_pml_acoustic_o8_f32")
    z = {k: base[k] for k in ("kind", "dx", "dt", "space_order", "stencil", "boundary",
                              "materials", "pml", "pml_faces")}
    z.update(Nx=np.array(Nx), N=np.array(N), precision=np.array("float32"))
    sim = product_sim(z)
    rng = np.random.default_rng(seed)
    indicator = rng.uniform(0.2, 0.8, tuple(sim.Nx_padded)).astype(sim.np_dtype)
    mat = sim.build_materials(indicator)
    prop = Propagator(sim, mat, shots)
    sig = rng.standard_normal((N, 1)).astype(sim.np_dtype)
    sources = [wave.Source(np.array([[rng.integers(3, Nx[0] - 3)], [rng.integers(3, Nx[1] - 3)]]),
                           sig * (1 + k)) for k in range(shots)]
    src, signal = source_points(prop, mat, sources)
    sensors = np.stack([rng.integers(3, n - 3, 9) for n in Nx]).astype(np.int32)
    rec = sensor_points(prop, sensors)
    return sim, prop, src, signal, rec


def _run(entry, sim, prop, src, signal, rec, spans):
    U = prop.levels_memory(2)
    record = prop.record_memory(rec)
    aux = prop.aux_memory() if entry == "cw_scalar_forward" else None
    a = prop.forward_args(U, aux=aux, sources=src, signal=signal, sensors=rec, record=record)
    for t0, t1 in spans:
        _abi.call(entry, ctypes.byref(a), t0, t1, _abi.stream())
    torch.cuda.synchronize()
    return U.cpu().numpy(), record.cpu().numpy()


@pytest.mark.parametrize("Nx, shots, N", [((134, 200), 5, 40), ((70, 90), 2, 30)],
                         ids=["134x200-5shots", "70x90-2shots"])
def test_the_static_forward_gives_the_bits_of_the_allocating_one(Nx, shots, N):
    print(f"case static forward {Nx} {shots} shots {N} steps", flush=True)
    case = _case(Nx, shots, N, seed=Nx[0])
    want_u, want_rec = _run("cw_scalar_forward", *case, [(0, N)])
    got_u, got_rec = _run("cw_scalar_forward_static", *case, [(0, N)])
    np.testing.assert_array_equal(got_u, want_u)
    np.testing.assert_array_equal(got_rec, want_rec)
    assert np.abs(want_rec).max() > 0.0


def test_a_sweep_split_over_two_static_calls_continues_the_layer():
    print("case static forward split 134x200 3 shots", flush=True)
    case = _case((134, 200), 3, 40, seed=5)
    want_u, want_rec = _run("cw_scalar_forward", *case, [(0, 40)])
    got_u, got_rec = _run("cw_scalar_forward_static", *case, [(0, 17), (17, 40)])
    np.testing.assert_array_equal(got_u, want_u)
    np.testing.assert_array_equal(got_rec, want_rec)


def test_memories_past_the_scratch_are_refused_before_any_launch():
    print("case static forward refused", flush=True)
    _abi.require_device()
    base = golden("adj_pml_acoustic_o8_f32")
    z = {k: base[k] for k in ("kind", "dx", "dt", "space_order", "stencil", "boundary",
                              "materials", "pml_faces")}
    z.update(Nx=np.array((1090, 1090)), N=np.array(10), precision=np.array("float32"),
             pml=np.array((32, 1500.0, 1e-6, 0.0)))
    sim = product_sim(z)
    a = _abi.ForwardArgs()
    a.grid = sim.abi_grid(64)
    for d in range(2):
        a.pml.lo[d] = a.pml.hi[d] = 32
    a.steps = 10
    lib = _abi.load()
    assert lib.cw_scalar_forward_static(ctypes.byref(a), 0, 10, None) != 0
    assert b"static scratch" in lib.cw_last_error()
