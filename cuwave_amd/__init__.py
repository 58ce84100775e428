"""cuwave_amd: the finite-difference time-stepping and adjoint path of CuWave, for MI355X.

Hand-written HIP kernels for gfx950 behind a C ABI (`include/cuwave_amd.h`),
driven from torch tensors through ctypes, with the reference's module layout and
Python API: `cuwave_amd.scalar` (`ScalarWave`, `AcousticWave`, ...),
`cuwave_amd.wave.simulate`, `cuwave_amd.sensitivity` (`sensitivity`,
`reconstruction_sensitivity`, `superposition_sensitivity`) and
`cuwave_amd.utils.misfit_gradient`.
"""

__version__ = "0.1.0"
