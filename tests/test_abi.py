"""The C ABI on the host: the library loads, exports every declared symbol, buckets
points and refuses bad arguments before it would launch anything (no GPU needed)."""

from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest

from helpers import ROOT

from cuwave_amd import _abi, scalar


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "cuwave_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(cw_\w+)\s*\(", text, flags=re.M)


def test_every_declared_symbol_is_exported_and_bound():
    lib = _abi.load()
    names = declared_functions()
    assert "cw_scalar_forward" in names and "cw_scalar_backward" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_abi.SIGNATURES)
    assert lib.cw_version() == _abi.ABI_VERSION


def test_struct_layouts_match_the_header():
    # the field order of every struct, as the header declares it
    text = (ROOT / "include" / "cuwave_amd.h").read_text()
    for struct, cls in (("cw_grid", _abi.Grid), ("cw_points", _abi.Points),
                        ("cw_strip", _abi.Strip), ("cw_levels", _abi.Levels),
                        ("cw_backward_args", _abi.BackwardArgs),
                        ("cw_maxwell_grid", _abi.MaxwellGrid),
                        ("cw_maxwell_step_args", _abi.MaxwellStepArgs),
                        ("cw_aniso_grid", _abi.AnisoGrid),
                        ("cw_aniso_step_args", _abi.AnisoStepArgs),
                        ("cw_aniso_gradient_args", _abi.AnisoGradientArgs)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = re.findall(r"(\w+)\s*(?:\[\d+\])*\s*[,;]", body)
        assert [f[0] for f in cls._fields_] == fields, struct


def _sim(**kw):
    return scalar.ScalarWave((30, 40), (0.1, 0.1), 10, 0.01, wavespeed=1.0, density=1.0,
                             space_order=4, **kw)


def test_points_are_bucketed_by_tile_stably():
    sim = _sim()
    grid = sim.abi_grid(2)
    tiles = _abi.I32()
    _abi.call("cw_tile_count", ctypes.byref(grid), ctypes.byref(tiles))
    rng = np.random.default_rng(0)
    pos = np.stack([rng.integers(1, 29, 200), rng.integers(1, 39, 200)])
    lin = sim.flatten_indices(pos)
    starts = np.zeros(tiles.value + 1, np.int32)
    order = np.zeros(200, np.int32)
    _abi.call("cw_point_lists", ctypes.byref(grid), lin.ctypes.data, 200, starts.ctypes.data,
              order.ctypes.data)
    assert starts[0] == 0 and starts[-1] == 200 and np.all(np.diff(starts) >= 0)
    assert sorted(order) == list(range(200))
    for k in range(tiles.value):
        bucket = order[starts[k] : starts[k + 1]]
        assert np.all(np.diff(bucket) > 0)  # stable: original order kept inside a tile
        # a tile holds nodes of a contiguous run of rows and columns only
        if len(bucket):
            rows, cols = pos[0][bucket], pos[1][bucket]
            assert rows.max() - rows.min() < 4 and cols.max() - cols.min() < 64


def test_bad_arguments_are_refused_before_any_launch():
    sim = _sim()
    grid = sim.abi_grid(1)
    args = _abi.ForwardArgs()
    args.grid = grid
    with pytest.raises(RuntimeError, match="prepared materials"):
        _abi.call("cw_scalar_forward", ctypes.byref(args), 0, 1, None)
    args.grid.order = 10
    with pytest.raises(RuntimeError, match="space order"):
        _abi.call("cw_scalar_forward", ctypes.byref(args), 0, 1, None)
    args.grid.order = 4
    args.grid.np[1] = 41
    with pytest.raises(RuntimeError, match="padded"):
        _abi.call("cw_scalar_forward", ctypes.byref(args), 0, 1, None)
    b = _abi.BackwardArgs()
    b.grid = sim.abi_grid(1)
    with pytest.raises(RuntimeError):
        _abi.call("cw_scalar_backward", ctypes.byref(b), 0, 1, None)


def test_the_vector_families_refuse_bad_arguments_before_any_launch():
    from cuwave_amd import anisotropic, maxwell

    mx = maxwell.DielectricWave((10, 11, 12), (0.1,) * 3, 5, 0.01, permittivity1=1.0,
                                permittivity2=2.0)
    a = _abi.MaxwellStepArgs()
    a.grid = mx.abi_grid(3)
    with pytest.raises(RuntimeError, match="curl scratch"):
        _abi.call("cw_maxwell_step", ctypes.byref(a), None)
    a.grid.ndim = 1
    with pytest.raises(RuntimeError, match="no curl"):
        _abi.call("cw_maxwell_step", ctypes.byref(a), None)
    a.grid = mx.abi_grid(3)
    a.grid.lanes = 3
    with pytest.raises(RuntimeError, match="power of two"):
        _abi.call("cw_maxwell_step", ctypes.byref(a), None)
    a.grid = mx.abi_grid(3)
    a.grid.conductor = 1 << 6
    with pytest.raises(RuntimeError, match="one bit a face"):
        _abi.call("cw_maxwell_step", ctypes.byref(a), None)
    an = anisotropic.AnisotropicElasticWave((10, 12), (0.1, 0.1), 5, 1e-3, density=1.0,
                                            wavespeed_p=2.0, wavespeed_s=1.0)
    s = _abi.AnisoStepArgs()
    s.grid = an.abi_grid(2)
    with pytest.raises(RuntimeError, match="table"):
        _abi.call("cw_aniso_step", ctypes.byref(s), None)
    s.grid.np[1] = 13
    with pytest.raises(RuntimeError, match="padded"):
        _abi.call("cw_aniso_step", ctypes.byref(s), None)
    g = _abi.AnisoGradientArgs()
    g.grid = an.abi_grid(2)
    with pytest.raises(RuntimeError, match="accumulators"):
        _abi.call("cw_aniso_gradient", ctypes.byref(g), None)
    p = _abi.ElasticPointArgs()
    p.runs, p.kind = 1, 3
    with pytest.raises(RuntimeError, match="points need"):
        _abi.call("cw_aniso_points", ctypes.byref(an.abi_grid(2)), ctypes.byref(p), None)


def test_lanes_hold_the_batch_as_a_power_of_two():
    sim = _sim()
    assert [sim.abi_grid(b).lanes for b in (1, 2, 3, 5, 32, 33, 64)] == [1, 2, 4, 8, 32, 64, 64]
    with pytest.raises(ValueError, match="1 .. 64"):
        sim.abi_grid(65)
    args = _abi.ForwardArgs()
    args.grid = sim.abi_grid(3)
    for lanes in (3, 2, 128):
        args.grid.lanes = lanes
        with pytest.raises(RuntimeError, match="lanes must be a power of two"):
            _abi.call("cw_scalar_forward", ctypes.byref(args), 0, 1, None)


def test_the_product_refuses_to_run_without_its_device(monkeypatch):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="AMD GPU"):
        _abi.require_device()
