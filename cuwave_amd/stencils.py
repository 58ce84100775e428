"""Finite-difference weights of the pressure operator, in float64 on the host.

The step reads one row of cell-flux weights at the full stencil radius
(cuwave/stencils.py:186-213): "taylor" is the formal-order row, "optimized" the
least-squares row over the widest band whose phase-velocity error stays under
`DISPERSION_TOL` (cuwave/stencils.py:40-58), second order under refinement.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

STENCILS = ("taylor", "optimized")
DISPERSION_TOL = 1e-3  # phase velocity error an optimized row keeps over its band


def weights(R: int) -> np.ndarray:
    """Central second-derivative weights of order 2R, `w[-R..R]`, from the moment conditions."""
    x = np.arange(-R, R + 1)
    moments = x[None, :] ** np.arange(2 * R + 1)[:, None]
    rhs = np.zeros(2 * R + 1)
    rhs[2] = 2.0
    return np.linalg.solve(moments, rhs)


def cell_coefficients(R: int, stencil: str = "taylor") -> np.ndarray:
    """The cell-flux weights at radius R: tail sums of `weights(R)`, or the fitted row."""
    if stencil not in STENCILS:
        raise ValueError(f"unknown stencil {stencil!r}, expected one of {STENCILS}")
    if stencil == "optimized":
        return _band_fit(R, staggered=False)
    w = weights(R)
    return np.array([w[R + 1 + k :].sum() for k in range(R)])


def staggered_weights(R: int, stencil: str = "taylor") -> np.ndarray:
    """First-derivative weights on the half offsets (2k - 1) / 2, `w[1..R]`."""
    if stencil not in STENCILS:
        raise ValueError(f"unknown stencil {stencil!r}, expected one of {STENCILS}")
    if stencil == "optimized":
        return _band_fit(R, staggered=True)
    x = np.arange(1, R + 1) - 0.5
    odd = x[None, :] ** (2 * np.arange(R)[:, None] + 1)
    rhs = np.zeros(R)
    rhs[0] = 0.5
    return np.linalg.solve(odd, rhs)


@lru_cache
def _band_fit(R: int, staggered: bool) -> np.ndarray:
    """Least squares of the symbol ratio against one over the widest admissible band.

    The band shrinks from pi in 400 steps; at each, the ratio of the discrete to
    the exact symbol is fitted to one over 400 wavenumbers under the consistency
    constraint sum (2k - 1) c_k = 1 (a Lagrange row), and the first band whose
    phase-velocity error stays within `DISPERSION_TOL` is kept.
    """
    k = np.arange(1, R + 1)
    for band in np.linspace(np.pi, 0.0, 400, endpoint=False):
        theta = np.linspace(band / 400, band, 400)
        G = 2.0 * np.sin(np.outer(theta, k - 0.5)) / theta[:, None]
        if not staggered:
            G *= 2.0 * np.sin(theta / 2.0)[:, None] / theta[:, None]
        consistency = (2 * k - 1.0)[None]
        system = np.block([[G.T @ G, consistency.T], [consistency, np.zeros((1, 1))]])
        c = np.linalg.solve(system, np.append(G.sum(axis=0), 1.0))[:R]
        ratio = G @ c
        error = ratio - 1.0 if staggered else np.sqrt(ratio) - 1.0
        if np.abs(error).max() <= DISPERSION_TOL:
            return c
    raise ValueError(f"no band keeps radius {R} within {DISPERSION_TOL}")


def flux_row(space_order: int, stencil: str = "optimized") -> np.ndarray:
    """The four cell weights the kernels take: the full-radius row, zero padded.

    At radius 1 the weight is exactly one, whatever the fit returns, as the
    reference's kernels hard-code it (cuwave/kernels/common.cuh:26-29).
    """
    R = space_order // 2
    if R > 4:
        raise ValueError(f"space_order {space_order} is past the kernels' 8")
    row = np.zeros(4)
    row[:R] = cell_coefficients(R, stencil)
    if R == 1:
        row[0] = 1.0
    return row


def cell_rows(space_order: int, stencil: str = "optimized") -> np.ndarray:
    """(4, 4) cell weights per radius, row r - 1 the radius-r row zero padded.

    A cell graded at a domain wall reads the row of its own radius: the formal-order
    (taylor) row below the full radius, the full radius' row as `flux_row` gives it
    (cuwave/stencils.py:186-200).
    """
    R = space_order // 2
    rows = np.zeros((4, 4))
    for r in range(1, R):
        rows[r - 1, :r] = cell_coefficients(r, "taylor")
    rows[R - 1] = flux_row(space_order, stencil)
    return rows


def stable_dt(
    dx: tuple[float, ...],
    wavespeed: float,
    space_order: int = 2,
    stencil: str = "optimized",
) -> float:
    """CFL-stable timestep: the flux symbol peaks at Nyquist, at the alternating sum."""
    c = cell_coefficients(space_order // 2, stencil)
    lam = 4.0 * abs(float(c @ (-1.0) ** np.arange(len(c))))
    return 2.0 / (wavespeed * float(np.sqrt(lam * sum(1.0 / d**2 for d in dx))))
