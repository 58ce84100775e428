"""CPU oracle of the vector Maxwell family: TEST INFRASTRUCTURE ONLY.

Only tests/ import this; the product (`cuwave_amd.maxwell`) never does. It
restates the reference's curl-curl pair on the Yee lattice in numpy, whole
arrays at a time, one shot, on the logical grid (no padding):

- `curl` is curl_kernel (cuwave/kernels/maxwell.cu:49-73, 209-239);
- `force` and `step` are fd_kernel (maxwell.cu:171-204, 242-274), with a
  nodal damping field its USE_DAMPING build, `step(..., v1=, g_mass=)` its
  GRADIENT build;
- `point_inverse_inertia` is MaxwellWave.build_materials with the
  permittivity's harmonic point means and the wall weights
  (cuwave/maxwell.py:247-272, cuwave/wave.py:386-425), and
  `pair_inverse_permeability` its pair-point nu W where the permeability is
  designed (the USE_MAGNETIC build);
- `simulate` is the forward march of cuwave/wave.py:1071-1163 with the
  point sources and probes of cuwave/kernels/common.cuh:456-490.

Every expression keeps the reference's operation order, so a float64 run is
the kernels' to the bit. The oracle is pinned by the fdtdx traces the
reference's own tests compare with (tests/reference/fdtdx_vector3D_*.npz,
copied into tests/golden/), by the exact symmetry of the mass-weighted
operator, by the null space a discrete gradient field spans, and by a
conductor holding the tangential field at zero.
"""

from __future__ import annotations

import numpy as np


def pair_row(ndim: int, k: int, l: int) -> int:
    """The curl row of the pair (k, l), k < l: 3 - k - l in 3D, 0 in 2D."""
    return 3 - k - l if ndim == 3 else 0


def node_image(b: np.ndarray, N: int, odd: int):
    """Node b past the wall node: its mirror about it, and the sign across a face of `odd`
    (cuwave/kernels/common.cuh:68-74)."""
    neg = ((b < 1) & bool(odd & 1)) | ((b > N - 2) & bool(odd & 2))
    return np.where(b < 1, 2 - b, np.where(b > N - 2, 2 * (N - 2) - b, b)), neg


def half_image(b: np.ndarray, N: int, odd: int):
    """Half point b past the wall's half point: its mirror about that, and the sign
    (cuwave/kernels/common.cuh:76-81)."""
    neg = ((b < 1) & bool(odd & 1)) | ((b > N - 3) & bool(odd & 2))
    return np.where(b < 1, 1 - b, np.where(b > N - 3, 2 * N - 5 - b, b)), neg


def shifted(f: np.ndarray, axis: int, j: int, odd: int, half: bool) -> np.ndarray:
    """g[..., a, ...] = f at a + j along `axis`, through the image (a in 1 .. N - 2;
    the ghost entries are junk the caller masks)."""
    N = f.shape[axis]
    a = np.clip(np.arange(N), 1, N - 2)
    b, neg = (half_image if half else node_image)(a + j, N, odd)
    b = np.clip(b, 0, N - 1)
    out = np.take(f, b, axis=axis)
    shape = [1] * f.ndim
    shape[axis] = N
    return np.where(neg.reshape(shape), -out, out)


def axis_index(shape, d):
    """The node index along axis d, broadcast to `shape`."""
    s = [1] * len(shape)
    s[d] = shape[d]
    return np.arange(shape[d]).reshape(s)


def interior(shape) -> np.ndarray:
    m = np.ones(shape, dtype=bool)
    for d in range(len(shape)):
        a = axis_index(shape, d)
        m = m & (a >= 1) & (a <= shape[d] - 2)
    return m


def pair_curl(u: np.ndarray, k: int, l: int, w, inv_dx, conductor: int) -> np.ndarray:
    """The (k, l) curl b at its pair points (maxwell.cu:49-73)."""
    dt = u.dtype.type
    b = np.zeros(u.shape[1:], dtype=u.dtype)
    ol, ok = conductor >> (2 * l), conductor >> (2 * k)
    for j in range(1, len(w) + 1):
        dk = shifted(u[k], l, j, ol, False) - shifted(u[k], l, 1 - j, ol, False)
        dl = shifted(u[l], k, j, ok, False) - shifted(u[l], k, 1 - j, ok, False)
        b = b - (dt(inv_dx[l]) * dt(w[j - 1])) * dk
        b = b + (dt(inv_dx[k]) * dt(w[j - 1])) * dl
    return b


def curl(u: np.ndarray, w: np.ndarray, inv_dx, nu, conductor: int, nu_pair=None, v1=None,
         g_nu=None) -> np.ndarray:
    """nu W b of every pair point (maxwell.cu:209-239): u (ndim, *Nx). With `nu_pair`
    (npairs, *Nx) the product nu W is read per pair point, and with `g_nu` the adjoint
    curl b against the forward one of v1 is taken off it in place."""
    ndim = u.shape[0]
    shape = u.shape[1:]
    dt = u.dtype.type
    h = np.zeros((1 if ndim == 2 else 3, *shape), dtype=u.dtype)
    for k in range(ndim - 1):
        for l in range(k + 1, ndim):
            b = pair_curl(u, k, l, w, inv_dx, conductor)
            weight = np.ones(shape, dtype=u.dtype)
            for d in range(ndim):
                if d not in (k, l):
                    a = axis_index(shape, d)
                    weight = np.where((a == 1) | (a == shape[d] - 2), weight * dt(0.5), weight)
            valid = interior(shape) & (axis_index(shape, k) <= shape[k] - 3) & (
                axis_index(shape, l) <= shape[l] - 3)
            p = pair_row(ndim, k, l)
            if g_nu is not None:
                bv = pair_curl(v1, k, l, w, inv_dx, conductor)
                g_nu[p] = np.where(valid, g_nu[p] - b * bv, g_nu[p])
            scaled = nu_pair[p] * b if nu_pair is not None else (dt(nu) * weight) * b
            h[p] = np.where(valid, scaled, 0)
    return h


def force(h: np.ndarray, c: int, w: np.ndarray, f, conductor: int, ndim: int) -> np.ndarray:
    """The force on component c from the curls (maxwell.cu:171-204)."""
    shape = h.shape[1:]
    dt = h.dtype.type
    R = len(w)
    out = np.zeros(shape, dtype=h.dtype)
    for m in range(ndim):
        if m == c:
            continue
        hp = h[pair_row(ndim, min(c, m), max(c, m))]
        odd = ~conductor >> (2 * m)
        acc = np.zeros(shape, dtype=h.dtype)
        for j in range(1, R + 1):
            acc = acc + dt(w[j - 1]) * shifted(hp, m, j - 1, odd, True)
            acc = acc - dt(w[j - 1]) * shifted(hp, m, -j, odd, True)
        a = axis_index(shape, m)
        acc = np.where((a == 1) | (a == shape[m] - 2), acc * dt(0.5), acc)
        out = out + (dt(1.0 if m < c else -1.0) * dt(f[m])) * acc
    return out


def step(u0, u1, minv, w, inv_dx, step_f, nu, conductor, v1=None, g_mass=None, mf=0.0,
         nu_pair=None, g_nu=None, damping=None, dt_=0.0):
    """u2 of the pair; with v1 and g_mass the inertia density taken off g_mass in place,
    with g_nu too the permeability density; with a nodal `damping` the damped update
    (2 u1 - u0 (1 - beta) + minv force) / (1 + beta), beta = minv d dt / 2
    (maxwell.cu:209-274)."""
    ndim = u1.shape[0]
    shape = u1.shape[1:]
    dt = u1.dtype.type
    h = curl(u1, w, inv_dx, nu, conductor, nu_pair, v1, g_nu)
    u2 = np.zeros_like(u1)
    for c in range(ndim):
        valid = interior(shape) & (axis_index(shape, c) <= shape[c] - 3)
        F = force(h, c, w, step_f, conductor, ndim)
        if damping is None:
            nxt = -u0[c] + dt(2) * u1[c] + minv[c] * F
        else:
            beta = dt(0.5) * minv[c] * damping * dt(dt_)
            nxt = (dt(2) * u1[c] - u0[c] * (dt(1) - beta) + minv[c] * F) / (dt(1) + beta)
        if g_mass is not None:
            dens = (dt(mf) * v1[c]) * (nxt - dt(2) * u1[c] + u0[c])
            g_mass[c] = np.where(valid, g_mass[c] - dens, g_mass[c])
        u2[c] = np.where(valid, nxt, 0)
    return u2


def wall_weights(shape, axes) -> np.ndarray:
    """Halved on the walls of the axes a point family is not staggered along
    (cuwave/wave.py:386-394)."""
    w = np.ones(shape)
    for d in range(len(shape)):
        if d in axes:
            continue
        a = axis_index(shape, d)
        w = np.where((a == 1) | (a == shape[d] - 2), w * 0.5, w)
    return w


def point_harmonic(field: np.ndarray, c: int) -> np.ndarray:
    """The harmonic mean of the two nodes point c sits between (cuwave/wave.py:396-425)."""
    dt = field.dtype
    tiny = np.finfo(dt).tiny
    inv = 1.0 / np.maximum(field, tiny)
    mean = np.zeros_like(field)
    N = field.shape[c]
    lo = [slice(None)] * field.ndim
    hi = [slice(None)] * field.ndim
    lo[c], hi[c] = slice(0, N - 1), slice(1, N)
    mean[tuple(lo)] = 0.5 * (inv[tuple(lo)] + inv[tuple(hi)])
    return np.where(mean > 0.0, 1.0 / np.maximum(mean, tiny), 0.0).astype(dt)


def point_inverse_inertia(eps: np.ndarray, conductor: int) -> np.ndarray:
    """1 / (W eps_c) per component, zeroed on a conductor's wall node for the tangential
    components (cuwave/maxwell.py:247-263)."""
    ndim = eps.ndim
    shape = eps.shape
    dt = eps.dtype
    tiny = np.finfo(dt).tiny
    minv = np.zeros((ndim, *shape), dtype=dt)
    for c in range(ndim):
        mass = wall_weights(shape, (c,)).astype(dt) * point_harmonic(eps, c)
        minv[c] = 1.0 / np.maximum(mass, tiny)
    for face in range(2 * ndim):
        if not (conductor >> face) & 1:
            continue
        d = face // 2
        wall = [slice(None)] * ndim
        wall[d] = 1 if face % 2 == 0 else shape[d] - 2
        for c in range(ndim):
            if c != d:
                minv[(c, *wall)] = 0.0
    return minv


def pair_average(field: np.ndarray, axes) -> np.ndarray:
    """The harmonic mean of the four nodes a pair point straddles (cuwave/wave.py:436-457)."""
    dt = field.dtype
    out = np.zeros_like(field)
    safe = np.maximum(field, np.finfo(dt).tiny)
    inner = tuple(slice(0, n - 1) if d in axes else slice(0, n) for d, n in enumerate(field.shape))
    for b0 in (0, 1):
        for b1 in (0, 1):
            bits = dict(zip(axes, (b0, b1)))
            shifted_ = tuple(slice(bits[d], n - 1 + bits[d]) if d in axes else slice(0, n)
                             for d, n in enumerate(field.shape))
            out[inner] += 1.0 / safe[shifted_]
    out[inner] = dt.type(4.0) / out[inner]
    return out


def pair_inverse_permeability(nu: np.ndarray) -> np.ndarray:
    """nu W of every pair point: the wall weights of the axes the pair does not span
    times the pair mean of nu (cuwave/maxwell.py:264-270)."""
    axes = [(0, 1)] if nu.ndim == 2 else [(1, 2), (0, 2), (0, 1)]
    return np.stack([wall_weights(nu.shape, a).astype(nu.dtype) * pair_average(nu, a)
                     for a in axes])


def simulate(minv, w, inv_dx, step_f, nu, conductor, nsteps, src, weight, signal, probes,
             nu_pair=None, damping=None, dt_=0.0):
    """The forward march of one shot: u^t = step(u^(t-2), u^(t-1)), then the sources
    (field += weight signal, entry by entry), then the probes. `src`, `probes` are
    (component, *node) index tuples of arrays; returns (the last level, (N, count))."""
    shape = minv.shape
    dt = minv.dtype
    u0 = np.zeros(shape, dtype=dt)
    u1 = np.zeros(shape, dtype=dt)
    traces = np.zeros((nsteps, len(probes[0])), dtype=dt)
    for t in range(nsteps):
        u2 = step(u0, u1, minv, w, inv_dx, step_f, nu, conductor, nu_pair=nu_pair,
                  damping=damping, dt_=dt_)
        for e in range(len(src[0])):
            at = tuple(int(s[e]) for s in src)
            u2[at] = u2[at] + weight[e] * signal[t, e]
        traces[t] = u2[probes]
        u0, u1 = u1, u2
    return u1, traces
