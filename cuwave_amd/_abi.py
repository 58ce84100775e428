"""ctypes bindings of `libcuwave_amd.so`, the C ABI declared in `include/cuwave_amd.h`.

This replaces the reference's kernel launch path (`cp.RawModule(...).get_function(name)`
launched as `kernel(grid, block, args)`, cuwave/wave.py:637-658; its CPU form
`launch_<name>(gx, gy, gz, bx, by, bz, void **args)`, cuwave/kernels/cpu.hpp:93-98):
the host holds torch tensors on the GPU and hands their raw pointers and sizes to
the library, which launches the HIP kernels on torch's current stream.

There is no fallback: without the built library, or without a GPU, every entry
point raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

LIBRARY = Path(__file__).resolve().parent / "libcuwave_amd.so"
ABI_VERSION = 15
DEVICE = "cuda"  # where every tensor handed to the library must live

I32, I64, F64, VP = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class Grid(ctypes.Structure):
    _fields_ = [
        ("ndim", I32),
        ("order", I32),
        ("n", I32 * 3),
        ("np", I32 * 3),
        ("nbatch", I32),
        ("lanes", I32),
        ("dtype", I32),
        ("dirichlet", I32),
        ("neumann", I32),
        ("op", F64 * 4),
        ("f", F64 * 3),
    ]


class Materials(ctypes.Structure):
    _fields_ = [("stiff", VP), ("minv", VP), ("prep", VP), ("damping", VP),
                ("damping_reverse", VP), ("dt", F64), ("cells", VP),
                ("rows", (F64 * 4) * 4)]


class Pml(ctypes.Structure):
    _fields_ = [
        ("lo", I32 * 3),
        ("hi", I32 * 3),
        ("profile", VP * 3),
        ("aux", VP * 3),
        ("eta", (VP * 2) * 3),
    ]


class Points(ctypes.Structure):
    _fields_ = [
        ("starts", VP),
        ("lin", VP),
        ("shot", VP),
        ("col", VP),
        ("weight", VP),
        ("row_stride", I64),
        ("col_stride", I64),
        ("shot_stride", I64),
    ]


class Strip(ctypes.Structure):
    _fields_ = [
        ("map", VP),
        ("record", VP),
        ("count", I32),
        ("first", I32),
        ("rows", I32),
        ("bf16", I32),
    ]


class Levels(ctypes.Structure):
    _fields_ = [
        ("base", VP),
        ("slot_stride", I64),
        ("slots", I32),
        ("offset", I32),
    ]


class ForwardArgs(ctypes.Structure):
    _fields_ = [
        ("grid", Grid),
        ("mat", Materials),
        ("pml", Pml),
        ("u", Levels),
        ("sources", Points),
        ("signal", VP),
        ("sensors", Points),
        ("record", VP),
        ("strip", Strip),
        ("steps", I32),
        ("illum", VP),
    ]


class BackwardArgs(ctypes.Structure):
    _fields_ = [
        ("grid", Grid),
        ("mat", Materials),
        ("pml", Pml),
        ("lam", Levels),
        ("u", Levels),
        ("reverse", I32),
        ("sensors", Points),
        ("adjoint_signal", VP),
        ("sources", Points),
        ("signal", VP),
        ("strip", Strip),
        ("g_mass", VP),
        ("g_stiff", VP),
        ("combine", I32),
        ("jac_mass", F64),
        ("jac_stiff", F64),
        ("skip_layer", I32),
        ("mass_factor", F64),
        ("steps", I32),
    ]


class SuperposedArgs(ctypes.Structure):
    _fields_ = [
        ("grid", Grid),
        ("mat", Materials),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("acc_mass", VP),
        ("acc_stiff", VP),
        ("ft", F64),
        ("fs", F64),
        ("step", I32),
    ]


class GradientArgs(ctypes.Structure):
    _fields_ = [
        ("grid", Grid),
        ("mat", Materials),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("l1", VP),
        ("g_mass", VP),
        ("g_stiff", VP),
        ("inv_dt2", F64),
        ("factor", F64 * 3),
    ]


class ElasticGrid(ctypes.Structure):
    _fields_ = [
        ("ndim", I32),
        ("order", I32),
        ("n", I32 * 3),
        ("np", I32 * 3),
        ("nbatch", I32),
        ("lanes", I32),
        ("dtype", I32),
        ("clamped", I32),
        ("closure", I32 * 4),
        ("stag", F64 * 4),
        ("lame", F64),
        ("shear", F64),
        ("density", F64),
        ("inv_dx", F64 * 3),
        ("step", F64 * 3),
    ]


class ElasticMaterial(ctypes.Structure):
    _fields_ = [("gamma", VP), ("prep", VP), ("rows", VP)]


class ElasticLayer(ctypes.Structure):
    _fields_ = [
        ("lo", I32 * 3),
        ("hi", I32 * 3),
        ("profile", VP * 3),
        ("memory", VP * 3),
        ("phase", I32),
    ]


class ElasticStepArgs(ctypes.Structure):
    _fields_ = [
        ("grid", ElasticGrid),
        ("mat", ElasticMaterial),
        ("pml", ElasticLayer),
        ("sigma", VP),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("g_mass", VP),
        ("g_normal", VP),
        ("g_shear", VP),
        ("v1", VP),
        ("mass_factor", F64),
    ]


class ElasticGradientArgs(ctypes.Structure):
    _fields_ = [
        ("grid", ElasticGrid),
        ("rows", VP),
        ("g_mass", VP),
        ("g_normal", VP),
        ("g_shear", VP),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("l1", VP),
        ("mass_factor", F64),
    ]


class MaxwellGrid(ctypes.Structure):
    _fields_ = [
        ("ndim", I32),
        ("order", I32),
        ("n", I32 * 3),
        ("np", I32 * 3),
        ("nbatch", I32),
        ("lanes", I32),
        ("dtype", I32),
        ("conductor", I32),
        ("stag", F64 * 4),
        ("nu", F64),
        ("inv_dx", F64 * 3),
        ("step", F64 * 3),
    ]


class MaxwellStepArgs(ctypes.Structure):
    _fields_ = [
        ("grid", MaxwellGrid),
        ("minv", VP),
        ("nu_pair", VP),
        ("g_nu", VP),
        ("curl", VP),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("g_mass", VP),
        ("v1", VP),
        ("mass_factor", F64),
        ("damping", VP),
        ("dt", F64),
    ]


class AnisoGrid(ctypes.Structure):
    _fields_ = [
        ("ndim", I32),
        ("n", I32 * 3),
        ("np", I32 * 3),
        ("nbatch", I32),
        ("lanes", I32),
        ("dtype", I32),
        ("step", F64),
    ]


class AnisoStepArgs(ctypes.Structure):
    _fields_ = [
        ("grid", AnisoGrid),
        ("table", VP),
        ("minv", VP),
        ("cell", VP),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("damping", VP),
        ("dt", F64),
    ]


class AnisoGradientArgs(ctypes.Structure):
    _fields_ = [
        ("grid", AnisoGrid),
        ("table", VP),
        ("g_mass", VP),
        ("g_cell", VP),
        ("u0", VP),
        ("u1", VP),
        ("u2", VP),
        ("l1", VP),
        ("mass_factor", F64),
    ]


class ElasticPointArgs(ctypes.Structure):
    _fields_ = [
        ("starts", VP),
        ("lin", VP),
        ("shot", VP),
        ("col", VP),
        ("runs", I32),
        ("kind", I32),
        ("row_stride", I64),
        ("col_stride", I64),
        ("shot_stride", I64),
        ("row", I64),
        ("field", VP),
        ("table", VP),
        ("weight", VP),
        ("g_mass", VP),
        ("u1", VP),
        ("mass_factor", F64),
    ]


class Conv3x3Args(ctypes.Structure):
    _fields_ = [
        ("n", I32),
        ("h", I32),
        ("w", I32),
        ("cin", I32),
        ("cout", I32),
        ("x", VP),
        ("weight", VP),
        ("weight_flipped", VP),
        ("bias", VP),
        ("relu", I32),
        ("y", VP),
        ("dy", VP),
        ("dx", VP),
        ("dweight", VP),
        ("workspace", VP),
        ("splits", I32),
    ]


class AdamWArgs(ctypes.Structure):
    _fields_ = [
        ("count", I64),
        ("param", VP),
        ("grad", VP),
        ("m", VP),
        ("v", VP),
        ("lr", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("weight_decay", ctypes.c_float),
        ("step", I32),
    ]


# the bytes of the library's own scratch on each GPU (CW_STATIC_SCRATCH_BYTES)
STATIC_SCRATCH_BYTES = 64 << 20

# every symbol include/cuwave_amd.h declares, with its argument types
SIGNATURES = {
    "cw_version": ([], I32),
    "cw_last_error": ([], ctypes.c_char_p),
    "cw_set_side_stream": ([I32], I32),
    "cw_set_stream_threshold": ([I32], I32),
    "cw_tile_count": ([ctypes.POINTER(Grid), ctypes.POINTER(I32)], I32),
    "cw_point_lists": ([ctypes.POINTER(Grid), VP, I32, VP, VP], I32),
    "cw_scalar_prepare": ([ctypes.POINTER(Grid), ctypes.POINTER(Materials), VP], I32),
    "cw_scalar_forward": ([ctypes.POINTER(ForwardArgs), I32, I32, VP], I32),
    "cw_scalar_forward_clocked": ([ctypes.POINTER(ForwardArgs), VP, I32, I32, VP], I32),
    "cw_scalar_backward": ([ctypes.POINTER(BackwardArgs), I32, I32, VP], I32),
    "cw_scalar_forward_static": ([ctypes.POINTER(ForwardArgs), I32, I32, VP], I32),
    "cw_scalar_superposed": ([ctypes.POINTER(SuperposedArgs), VP], I32),
    "cw_scalar_gradient": ([ctypes.POINTER(GradientArgs), VP], I32),
    "cw_elastic_prepare": ([ctypes.POINTER(ElasticGrid), ctypes.POINTER(ElasticMaterial), VP], I32),
    "cw_elastic_step": ([ctypes.POINTER(ElasticStepArgs), VP], I32),
    "cw_elastic_gradient": ([ctypes.POINTER(ElasticGradientArgs), VP], I32),
    "cw_elastic_points": ([ctypes.POINTER(ElasticGrid), ctypes.POINTER(ElasticPointArgs), VP], I32),
    "cw_maxwell_step": ([ctypes.POINTER(MaxwellStepArgs), VP], I32),
    "cw_maxwell_points": ([ctypes.POINTER(MaxwellGrid), ctypes.POINTER(ElasticPointArgs), VP], I32),
    "cw_aniso_step": ([ctypes.POINTER(AnisoStepArgs), VP], I32),
    "cw_aniso_gradient": ([ctypes.POINTER(AnisoGradientArgs), VP], I32),
    "cw_aniso_points": ([ctypes.POINTER(AnisoGrid), ctypes.POINTER(ElasticPointArgs), VP], I32),
    "cw_conv3x3_bf16_fwd": ([ctypes.POINTER(Conv3x3Args), VP], I32),
    "cw_conv3x3_bf16_bwd_data": ([ctypes.POINTER(Conv3x3Args), VP], I32),
    "cw_conv3x3_bf16_bwd_weight": ([ctypes.POINTER(Conv3x3Args), VP], I32),
    "cw_conv3x3_weights": ([I32, I32, VP, VP, VP, VP], I32),
    "cw_adamw": ([ctypes.POINTER(AdamWArgs), VP], I32),
}

_LIB: ctypes.CDLL | None = None


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """The library, loaded once; raises when it is not built or does not match."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = Path(path) if path is not None else LIBRARY
    if not path.exists():
        raise RuntimeError(
            f"{path} is not built: run `python -c 'import __graft_entry__ as g; "
            f"g.build()'` (hipcc --offload-arch=gfx950)"
        )
    lib = ctypes.CDLL(str(path))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.cw_version() != ABI_VERSION:
        raise RuntimeError(f"{path} has ABI {lib.cw_version()}, expected {ABI_VERSION}")
    _LIB = lib
    return lib


def call(name: str, *args) -> None:
    """Call entry point `name`, raising with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc:
        raise RuntimeError(f"{name}: {lib.cw_last_error().decode()}")


def call_value(name: str, *args) -> int:
    """Call a setter that returns a value (the previous setting), not a status."""
    return int(getattr(load(), name)(*args))


def ptr(t: torch.Tensor | None) -> int | None:
    """The raw pointer of a contiguous tensor on the library's device, None for None."""
    if t is None:
        return None
    if t.device.type != DEVICE:
        raise ValueError(f"tensors handed to the HIP kernels live on {DEVICE}, not {t.device}")
    if not t.is_contiguous():
        raise ValueError("tensors handed to the HIP kernels must be contiguous")
    return t.data_ptr()


def stream() -> int | None:
    """torch's current HIP stream, the one every launch goes on."""
    if DEVICE != "cuda":
        return None
    return torch.cuda.current_stream().cuda_stream


def require_device() -> torch.device:
    """The device the kernels run on; raises where there is none."""
    if DEVICE == "cuda" and not torch.cuda.is_available():
        raise RuntimeError("cuwave_amd runs its kernels on an AMD GPU and finds none")
    load()
    return torch.device(DEVICE)
