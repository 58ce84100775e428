"""CPU oracle of the cell-assembled elastic family: TEST INFRASTRUCTURE ONLY.

Only tests/ import this; the product (`cuwave_amd.anisotropic`) never does. It
restates, in numpy on the logical grid, one shot, whole arrays at a time:

- `step`: fd_kernel without damping (cuwave/kernels/anisotropic.cu:72-116,
  334-397): the force gathered from the 2^ndim cells around each interior
  node, cells in block order (the cell below first, axis 0 fastest), then
  the cell's corners, then the components, each term subtracted as
  (g K) u; then u2 = -u0 + 2 u1 + (f minv) force;
- `gradient`: gradient_kernel (anisotropic_sensitivity.cu:111-156);
- `inverse_inertia` and `cell_average`: the lumped inertia and the harmonic
  cell means (cuwave/anisotropic.py:137-154, 260-280);
- `simulate`: the forward march with point sources and probes.

The one-cell stencil K is an input (the product's table, pinned on its own by
its symmetry, its rigid-body null space and the continuum limit). The oracle
is pinned by a plane wave's acceleration against the Lame speeds, a rigid
motion carrying no force, the checkerboard not being a null mode, and the
exact symmetry of the mass-weighted operator with traction and clamped faces
(tests/test_anisotropic.py).
"""

from __future__ import annotations

import numpy as np


def axis_index(shape, d):
    s = [1] * len(shape)
    s[d] = shape[d]
    return np.arange(shape[d]).reshape(s)


def shift(f: np.ndarray, offsets) -> np.ndarray:
    """g[a] = f[a + offsets] where that lies on the grid, else 0."""
    out = np.zeros_like(f)
    src, dst = [], []
    for o, n in zip(offsets, f.shape):
        src.append(slice(max(o, 0), n + min(o, 0)))
        dst.append(slice(max(-o, 0), n - max(o, 0)))
    out[tuple(dst)] = f[tuple(src)]
    return out


def interior(shape) -> np.ndarray:
    m = np.ones(shape, dtype=bool)
    for d in range(len(shape)):
        a = axis_index(shape, d)
        m = m & (a >= 1) & (a <= shape[d] - 2)
    return m


def force(u: np.ndarray, cell: np.ndarray, K: np.ndarray) -> np.ndarray:
    """The cell-gathered force on every component (anisotropic.cu:334-365)."""
    ndim = u.shape[0]
    shape = u.shape[1:]
    ncorner = 2**ndim
    out = np.zeros_like(u)
    for c in range(ncorner):
        e = [((c >> d) & 1) - 1 for d in range(ndim)]
        inside = interior(shape)
        for d in range(ndim):
            low = axis_index(shape, d) + e[d]
            inside = inside & (low >= 1) & (low <= shape[d] - 3)
        me = ncorner - 1 - c
        gc = shift(cell, e)
        for l in range(ncorner):
            off = [e[d] + ((l >> d) & 1) for d in range(ndim)]
            for j in range(ndim):
                uj = shift(u[j], off)
                for i in range(ndim):
                    t = K[me * ndim + i, l * ndim + j]
                    out[i] = np.where(inside, out[i] - (gc * t) * uj, out[i])
    return out


def step(u0, u1, minv, cell, K, f, damping=None, dt_=0.0):
    """u2 = -u0 + 2 u1 + (f minv) force on the interior nodes, zero elsewhere; with a
    nodal `damping` the damped update (2 u1 - u0 (1 - beta) + f minv force) / (1 + beta),
    beta = minv d dt / 2 (anisotropic.cu:380-389)."""
    F = force(u1, cell, K)
    dt = u1.dtype.type
    inside = interior(u1.shape[1:])
    if damping is None:
        return np.where(inside, -u0 + dt(2) * u1 + (dt(f) * minv) * F, 0)
    beta = dt(0.5) * minv * damping * dt(dt_)
    return np.where(inside, (dt(2) * u1 - u0 * (dt(1) - beta) + (dt(f) * minv) * F) /
                    (dt(1) + beta), 0)


def gradient(g_mass, g_cell, u0, u1, u2, l1, K, mf):
    """Both densities of the triplet against l1, taken off the accumulators in place
    (anisotropic_sensitivity.cu:111-156)."""
    ndim = u1.shape[0]
    shape = u1.shape[1:]
    dt = u1.dtype.type
    ncorner = 2**ndim
    inside = interior(shape)
    acc = np.zeros(shape, dtype=u1.dtype)
    for i in range(ndim):
        acc = acc + l1[i] * (u2[i] - dt(2) * u1[i] + u0[i])
    g_mass[...] = np.where(inside, g_mass - dt(mf) * acc, g_mass)
    own = inside.copy()
    for d in range(ndim):
        own = own & (axis_index(shape, d) <= shape[d] - 3)
    bil = np.zeros(shape, dtype=u1.dtype)
    for l in range(ncorner):
        lo = [(l >> d) & 1 for d in range(ndim)]
        for m in range(ncorner):
            mo = [(m >> d) & 1 for d in range(ndim)]
            for i in range(ndim):
                li = shift(l1[i], lo)
                for j in range(ndim):
                    bil = bil + (li * K[l * ndim + i, m * ndim + j]) * shift(u1[j], mo)
    g_cell[...] = np.where(own, g_cell - bil, g_cell)


def cell_weights(shape) -> np.ndarray:
    w = np.ones(shape)
    for d in range(len(shape)):
        a = axis_index(shape, d)
        w = np.where((a == 1) | (a == shape[d] - 2), w * 0.5, w)
    return w


def inverse_inertia(gamma: np.ndarray, density: float, volume: float, clamped: int) -> np.ndarray:
    """1 / (gamma rho0 V W), zero on a clamped wall node (cuwave/anisotropic.py:260-280)."""
    dt = gamma.dtype.type
    mass = gamma * dt(density * volume) * cell_weights(gamma.shape).astype(gamma.dtype)
    minv = (1.0 / np.maximum(mass, np.finfo(gamma.dtype).tiny)).astype(gamma.dtype)
    for face in range(2 * gamma.ndim):
        if (clamped >> face) & 1:
            d = face // 2
            wall = [slice(None)] * gamma.ndim
            wall[d] = 1 if face % 2 == 0 else gamma.shape[d] - 2
            minv[tuple(wall)] = 0.0
    return minv


def cell_average(gamma: np.ndarray) -> np.ndarray:
    """The harmonic mean over a cell's corners at its low corner (cuwave/anisotropic.py:
    137-154), corners in the order bit d = plus side of axis d."""
    dt = gamma.dtype
    out = np.zeros_like(gamma)
    safe = np.maximum(gamma, np.finfo(dt).tiny)
    inner = tuple(slice(0, n - 1) for n in gamma.shape)
    for c in range(2**gamma.ndim):
        bits = [(c >> d) & 1 for d in range(gamma.ndim)]
        shifted = tuple(slice(b, n - 1 + b) for b, n in zip(bits, gamma.shape))
        out[inner] += 1.0 / safe[shifted]
    out[inner] = dt.type(2.0**gamma.ndim) / out[inner]
    return out


def simulate(minv, cell, K, f, nsteps, src, weight, signal, probes, damping=None, dt_=0.0):
    """The forward march of one shot; `src`, `probes` are (component, *node) index
    tuples. Returns (the last level, (N, count))."""
    shape = (minv.ndim, *minv.shape)
    dt = minv.dtype
    u0 = np.zeros(shape, dtype=dt)
    u1 = np.zeros(shape, dtype=dt)
    traces = np.zeros((nsteps, len(probes[0])), dtype=dt)
    for t in range(nsteps):
        u2 = step(u0, u1, minv, cell, K, f, damping, dt_)
        for e in range(len(src[0])):
            at = tuple(int(s[e]) for s in src)
            u2[at] = u2[at] + weight[e] * signal[t, e]
        traces[t] = u2[probes]
        u0, u1 = u1, u2
    return u1, traces
