// Device code of the elastic (staggered displacement) path for gfx950.
//
// Isotropic elasticity on a staggered grid (cuwave/elastic.py:1-10):
// displacement component c sits half a node up axis c, the normal strains on
// the nodes, the shear (k, l) half a node up both axes. A step is a pair of
// launches over the interior nodes: the stress launch forms the strains of
// u1 and writes the stresses (and, in the adjoint pair, the stiffness
// densities), the update launch takes the divergence of the stresses into
// u2 (and the inertia density). A convolutional PML folds into both
// launches through a memory per lined axis. The walls close with the
// staggered summation-by-parts rows of Prochnow et al. 2017, read from a
// table the caller passes.
//
// Layout as in the pressure path: per-shot arrays carry the shots
// innermost, (component, *np, lanes); a thread holds V neighbouring lanes of
// one node and runs every expression lane by lane in the reference's
// operation order (cuwave/kernels/elastic.cu:121-600, 754-853;
// cuwave/kernels/common.cuh:152-298), so a lane's bits are those of the
// one-shot arithmetic. The densities a node's lanes produce are summed over
// the lanes (cw::shot_sum) into one per-node accumulator.
#pragma once

#include "common.hpp"

namespace cwe {

using cw::fmat;
using cw::lanes_at;
using cw::lanes_put;
using cw::LanesOf;

constexpr int kThreads = 256;

template <typename T> struct Tiny;
template <> struct Tiny<float> {
  static constexpr float value = 1.17549435e-38f;
};
template <> struct Tiny<double> {
  static constexpr double value = 2.2250738585072014e-308;
};

// the grid, the operator and the material constants, by value
template <typename T> struct ElGeom {
  int n[3], np[3], s[3];
  int64_t nodes;          // nodes of one component
  int lanes, nbatch;
  int clamped;            // faces held at zero, bit 2 axis + side
  int nc, hc, nspan, hspan; // closure rows and the points they read
  int interior;           // interior nodes, every axis 1 .. n - 2
  T sg[4];                // full-radius staggered weights, sg[k - 1]
  T lam, mu;              // Lame parameters
  T f[3];                 // per-axis factors: 1 / dx (stress), dt^2 / dx (update)
  const T *rows;          // node closure rows (nc x nspan), then half rows (hc x hspan)
};

// the PML fold: per lined axis (lo + hi > 0) its memory and profile
template <typename T> struct ElLayer {
  T *aux[3];              // (rows, slab, lanes)
  const T *profile[3];    // (4, n[d]): a, b at nodes, then at half points
  int lo[3], hi[3];
  int phase;              // 0 forward; 1 / 2 the adjoint reads block 0 / 2 ndim
};

template <int NDIM>
__device__ __forceinline__ constexpr int stress_row(int c, int e) {
  if (c == e)
    return c;
  const int k = c < e ? c : e, l = c < e ? e : c;
  return NDIM == 3 ? NDIM + 3 - k - l : 2;
}

template <int NDIM> constexpr int nvoigt() { return NDIM + NDIM * (NDIM - 1) / 2; }

template <typename T>
__device__ __forceinline__ bool clamped_face(const ElGeom<T> &g, int d, int side) {
  return (g.clamped >> (2 * d + side)) & 1;
}

// the slab row of node a of an axis (rows lo off the low wall, hi off the
// high one), -1 off the slab; its index in the axis' memory
__device__ __forceinline__ int64_t slab_at(int a, int N, int lo, int hi, bool slow,
                                           int64_t plane, int64_t off) {
  const int row = cw::slab_row(a, N, lo, hi);
  if (row < 0)
    return -1;
  return slow ? (int64_t)row * plane + off : off * (int64_t)(lo + hi) + row;
}

// A line of points a derivative reads: a field component (or a stress row)
// along an axis, and optionally through the PML memory (the value coef *
// (base + memory) of the reference's Memory accessor,
// cuwave/kernels/common.cuh:421-437).
template <typename T> struct Line {
  const T *p;        // field or stresses, offset to the thread's first lane
  int64_t idx;       // the node
  int64_t nodes;     // nodes between components
  int step;          // node stride of the axis
  int lanes;
  int e;             // the axis
  bool stress;       // rows by stress_row(k, e)
  bool memory;
  const T *aux;      // offset to the thread's first lane
  const T *profile;
  int64_t size, plane, off;
  int row0, a, lo, hi, N;
  bool own_half, slow;
};

template <int NDIM, typename Q, typename T>
__device__ __forceinline__ Q line_at(const Line<T> &L, int k, int j) {
  const int row = L.stress ? stress_row<NDIM>(k, L.e) : k;
  const Q base = lanes_at<Q>(L.p, ((int64_t)row * L.nodes + L.idx + (int64_t)j * L.step) * L.lanes);
  if (!L.memory)
    return base;
  const int at = L.a + j;
  const int64_t i = slab_at(at, L.N, L.lo, L.hi, L.slow, L.plane, L.off);
  if (i < 0)
    return Q((T)0);
  const bool half = (k == L.e) == L.own_half;
  const T coef = L.profile[(half ? 2 * L.N : 0) + at];
  return coef * (base + lanes_at<Q>(L.aux, ((int64_t)(L.row0 + k) * L.size + i) * L.lanes));
}

template <typename T>
__device__ __forceinline__ Line<T> field_line(const T *p, int64_t idx, const ElGeom<T> &g, int d) {
  Line<T> L{};
  L.p = p;
  L.idx = idx;
  L.nodes = g.nodes;
  L.step = g.s[d];
  L.lanes = g.lanes;
  L.e = d;
  return L;
}

// ------------------------------------------- staggered SBP derivatives
// closure rows off a wall (cuwave/kernels/common.cuh:214-242)
template <int NDIM, typename Q, typename T>
__device__ Q node_rows(const ElGeom<T> &g, const Line<T> &L, int k, int lo, int hi) {
  Q acc((T)0);
  if (lo < g.nc) {
    for (int m = 0; m < g.nspan; ++m)
      acc += g.rows[lo * g.nspan + m] * line_at<NDIM, Q>(L, k, m - lo);
  } else {
    for (int m = 0; m < g.nspan; ++m)
      acc -= g.rows[hi * g.nspan + m] * line_at<NDIM, Q>(L, k, hi - 1 - m);
  }
  return acc;
}

template <int NDIM, typename Q, typename T>
__device__ Q half_rows(const ElGeom<T> &g, const Line<T> &L, int k, int lo, int hi) {
  const T *rows = g.rows + g.nc * g.nspan;
  Q acc((T)0);
  if (lo < g.hc) {
    for (int i = 0; i < g.hspan; ++i)
      acc += rows[lo * g.hspan + i] * line_at<NDIM, Q>(L, k, i - lo);
  } else {
    for (int i = 0; i < g.hspan; ++i)
      acc -= rows[hi * g.hspan + i] * line_at<NDIM, Q>(L, k, hi + 1 - i);
  }
  return acc;
}

// d/dx at node a of a half-point field (common.cuh:247-258)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q node_derivative(const ElGeom<T> &g, const Line<T> &L, int k, int a,
                                             int N) {
  const int lo = a - 1, hi = N - 2 - a;
  if (lo < g.nc || hi < g.nc)
    return node_rows<NDIM, Q>(g, L, k, lo, hi);
  Q acc((T)0);
#pragma unroll
  for (int q = 1; q <= R; ++q)
    acc += g.sg[q - 1] * (line_at<NDIM, Q>(L, k, q - 1) - line_at<NDIM, Q>(L, k, -q));
  return acc;
}

// d/dx at half point a of a nodal field (common.cuh:261-271)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q half_derivative(const ElGeom<T> &g, const Line<T> &L, int k, int a,
                                             int N) {
  const int lo = a - 1, hi = N - 3 - a;
  if (lo < g.hc || hi < g.hc)
    return half_rows<NDIM, Q>(g, L, k, lo, hi);
  Q acc((T)0);
#pragma unroll
  for (int q = 1; q <= R; ++q)
    acc += g.sg[q - 1] * (line_at<NDIM, Q>(L, k, q) - line_at<NDIM, Q>(L, k, 1 - q));
  return acc;
}

// the divergence of a nodal stress at half point a (common.cuh:276-288)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q node_divergence(const ElGeom<T> &g, const Line<T> &L, int k, int a,
                                             int N) {
  const int lo = a - 1, hi = N - 3 - a;
  if (lo < g.hc || hi < g.hc)
    return half_rows<NDIM, Q>(g, L, k, lo, hi);
  Q acc((T)0);
#pragma unroll
  for (int q = 1; q <= R; ++q) {
    acc += g.sg[q - 1] * line_at<NDIM, Q>(L, k, q);
    acc -= g.sg[q - 1] * line_at<NDIM, Q>(L, k, 1 - q);
  }
  return acc;
}

// the divergence of a half-point stress at node a (common.cuh:292-304)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q half_divergence(const ElGeom<T> &g, const Line<T> &L, int k, int a,
                                             int N) {
  const int lo = a - 1, hi = N - 2 - a;
  if (lo < g.nc || hi < g.nc)
    return node_rows<NDIM, Q>(g, L, k, lo, hi);
  Q acc((T)0);
#pragma unroll
  for (int q = 1; q <= R; ++q) {
    acc += g.sg[q - 1] * line_at<NDIM, Q>(L, k, q - 1);
    acc -= g.sg[q - 1] * line_at<NDIM, Q>(L, k, -q);
  }
  return acc;
}

// -------------------------------------------------- strains and stresses
template <int NDIM, typename T>
__device__ __forceinline__ bool traction_zero(const ElGeom<T> &g, const int *A, int d) {
  return (A[d] == 1 && !clamped_face(g, d, 0)) || (A[d] == g.n[d] - 2 && !clamped_face(g, d, 1));
}

template <int NDIM, typename T>
__device__ __forceinline__ int zeroed_axes(const ElGeom<T> &g, const int *A) {
  int zeroed = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    if (traction_zero<NDIM>(g, A, d))
      zeroed |= 1 << d;
  return zeroed;
}

// the normal strains of u at node idx, zero on a traction wall (elastic.cu:121-141)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ int normal_strains(const ElGeom<T> &g, const T *u, int64_t idx,
                                              const int *A, Q *eps) {
  int zeroed = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    if (traction_zero<NDIM>(g, A, d)) {
      eps[d] = Q((T)0);
      zeroed |= 1 << d;
    } else {
      const Line<T> L = field_line(u, idx, g, d);
      eps[d] = g.f[d] * node_derivative<NDIM, R, Q>(g, L, d, A[d], g.n[d]);
    }
  }
  return zeroed;
}

// the Lame parameter condensed over the traction-free axes (elastic.cu:145-154)
template <int NDIM, typename T>
__device__ __forceinline__ T condensed_lame(const ElGeom<T> &g, int zeroed) {
  T lam = g.lam;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    if ((zeroed >> d) & 1)
      lam = (T)2.f * lam * g.mu / (lam + (T)2.f * g.mu);
  return lam;
}

// the engineering shear strain (k, l) at its point (elastic.cu:159-170)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q shear_strain(const ElGeom<T> &g, const T *u, int64_t idx, int k,
                                          int l, const int *A) {
  const Line<T> Lk = field_line(u, idx, g, k);
  const Line<T> Ll = field_line(u, idx, g, l);
  const Q dk = half_derivative<NDIM, R, Q>(g, Lk, l, A[k], g.n[k]);
  const Q dl = half_derivative<NDIM, R, Q>(g, Ll, k, A[l], g.n[l]);
  return fmat(g.f[k], dk, g.f[l] * dl);
}

template <typename T>
__device__ __forceinline__ bool shear_point(const ElGeom<T> &g, int k, int l, const int *A) {
  return A[k] <= g.n[k] - 3 && A[l] <= g.n[l] - 3;
}

// every strain of u at node idx (elastic.cu:216-233)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ int strain_point(const ElGeom<T> &g, const T *u, int64_t idx,
                                            const int *A, Q *eps) {
  const int zeroed = normal_strains<NDIM, R>(g, u, idx, A, eps);
#pragma unroll
  for (int k = 0; k < NDIM - 1; ++k)
#pragma unroll
    for (int l = k + 1; l < NDIM; ++l)
      eps[stress_row<NDIM>(k, l)] =
          shear_point(g, k, l, A) ? shear_strain<NDIM, R, Q>(g, u, idx, k, l, A) : Q((T)0);
  return zeroed;
}

// the stresses of the strains (elastic.cu:236-263); the shear gamma of the
// pair points comes prepared (`shear`: the pair rows of prep)
template <int NDIM, typename Q, typename T>
__device__ __forceinline__ void stress_of(const ElGeom<T> &g, const Q *eps, int zeroed, T gn,
                                          const T *shear, int64_t idx, const int *A, Q *sig) {
  const T lam = condensed_lame<NDIM>(g, zeroed);
  Q tr((T)0);
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    tr += eps[d];
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    sig[d] = ((zeroed >> d) & 1) ? Q((T)0) : gn * ((T)2.f * g.mu * eps[d] + lam * tr);
#pragma unroll
  for (int k = 0; k < NDIM - 1; ++k)
#pragma unroll
    for (int l = k + 1; l < NDIM; ++l) {
      const int v = stress_row<NDIM>(k, l);
      sig[v] = shear_point(g, k, l, A) ? shear[(v - NDIM) * g.nodes + idx] * g.mu * eps[v]
                                       : Q((T)0);
    }
}

// the force on component c at its point: the divergence of the stresses
// (elastic.cu:281-303)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q nodal_force(const ElGeom<T> &g, const T *sigma, int64_t idx, int c,
                                         const int *A) {
  const Line<T> Lc = field_line(sigma, idx, g, c);
  Q force = g.f[c] * node_divergence<NDIM, R, Q>(g, Lc, c, A[c], g.n[c]);
#pragma unroll
  for (int l = 0; l < NDIM; ++l) {
    if (l == c)
      continue;
    const Line<T> Lv = field_line(sigma, idx, g, l);
    force = fmat(g.f[l], half_divergence<NDIM, R, Q>(g, Lv, stress_row<NDIM>(c, l), A[l], g.n[l]),
                 force);
  }
  return force;
}

// the normal-strain energy form of two strain sets (elastic.cu:339-351)
template <int NDIM, typename Q, typename T>
__device__ __forceinline__ Q normal_form(const ElGeom<T> &g, const Q *ea, const Q *eb, int zeroed) {
  const T lam = condensed_lame<NDIM>(g, zeroed);
  Q tra((T)0), trb((T)0), dot((T)0);
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    tra += ea[d];
    trb += eb[d];
    dot += ea[d] * eb[d];
  }
  return (T)2.f * g.mu * dot + lam * tra * trb;
}

// ------------------------------------------------------------ the pml fold
template <int NDIM, typename T>
__device__ __forceinline__ bool in_slab(const ElGeom<T> &g, int a, int e, int lo, int hi, int ext) {
  return (lo && a - 1 < lo + ext) || (hi && g.n[e] - 2 - a < hi + ext);
}

template <int NDIM, typename T>
__device__ __forceinline__ void slab_geometry(const ElGeom<T> &g, int e, int lo, int hi,
                                              const int *A, int64_t *plane, int64_t *size,
                                              int64_t *off) {
  int64_t p = 1, o = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    if (d != e) {
      p *= g.np[d];
      o = o * g.np[d] + A[d];
    }
  *plane = p;
  *size = p * (int64_t)(lo + hi);
  *off = o;
}

// d u_k / d x_e at the strain point (k, e) (elastic.cu:417-431)
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q strain_term(const ElGeom<T> &g, const Line<T> &L, int k, int e,
                                         const int *A) {
  if (k == e) {
    if (traction_zero<NDIM>(g, A, e))
      return Q((T)0);
    return g.f[e] * node_derivative<NDIM, R, Q>(g, L, e, A[e], g.n[e]);
  }
  if (A[k] > g.n[k] - 3 || A[e] > g.n[e] - 3)
    return Q((T)0);
  return g.f[e] * half_derivative<NDIM, R, Q>(g, L, k, A[e], g.n[e]);
}

// elastic.cu:434-442
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q divergence_term(const ElGeom<T> &g, const Line<T> &L, int c, int e,
                                             const int *A, T fe) {
  if (c == e)
    return fe * node_divergence<NDIM, R, Q>(g, L, c, A[c], g.n[c]);
  return fe * half_divergence<NDIM, R, Q>(g, L, c, A[e], g.n[e]);
}

// the stresses of the memory strains t of axis e (elastic.cu:445-467)
template <int NDIM, typename Q, typename T>
__device__ __forceinline__ void add_stress(const ElGeom<T> &g, Q *sig, T gn, const T *shear,
                                           int64_t idx, int e, const Q *t, const int *A) {
  const int zeroed = zeroed_axes<NDIM>(g, A);
  if (!((zeroed >> e) & 1)) {
    const T lam = condensed_lame<NDIM>(g, zeroed);
    const Q gg = gn * t[e];
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
      if (!((zeroed >> d) & 1))
        sig[d] += gg * ((d == e ? (T)2.f * g.mu : (T)0) + lam);
  }
#pragma unroll
  for (int k = 0; k < NDIM; ++k)
    if (k != e && A[k] <= g.n[k] - 3 && A[e] <= g.n[e] - 3) {
      const int v = stress_row<NDIM>(k, e);
      sig[v] += shear[(v - NDIM) * g.nodes + idx] * g.mu * t[k];
    }
}

// the strain memory of axis e (elastic.cu:486-528)
template <int NDIM, int R, int V, typename T>
__device__ __forceinline__ void stress_memory(const ElGeom<T> &g, const ElLayer<T> &P, int e,
                                              const T *u, T gn, const T *shear, int64_t idx,
                                              int lane0, const int *A, LanesOf<T, V> *sig) {
  using Q = LanesOf<T, V>;
  const int a = A[e], lo = P.lo[e], hi = P.hi[e];
  if (!in_slab<NDIM>(g, a, e, lo, hi, P.phase ? R : 0))
    return;
  int64_t plane, size, off;
  slab_geometry<NDIM>(g, e, lo, hi, A, &plane, &size, &off);
  const int N = g.n[e];
  const T *prof = P.profile[e];
  Q q[3];
  if (P.phase) {
    Line<T> h = field_line(u, idx, g, e);
    h.memory = true;
    h.aux = P.aux[e] + lane0;
    h.profile = prof;
    h.size = size;
    h.plane = plane;
    h.off = off;
    h.row0 = NDIM;
    h.a = a;
    h.lo = lo;
    h.hi = hi;
    h.N = N;
    h.own_half = true;
    h.slow = e < NDIM - 1;
#pragma unroll
    for (int k = 0; k < NDIM; ++k)
      q[k] = strain_term<NDIM, R, Q>(g, h, k, e, A);
  } else {
    const Line<T> f = field_line(u, idx, g, e);
    const int64_t t = slab_at(a, N, lo, hi, e < NDIM - 1, plane, off);
    T *aux = P.aux[e] + lane0;
#pragma unroll
    for (int k = 0; k < NDIM; ++k) {
      const int at = (k == e ? 0 : 2 * N) + a;
      const int64_t m = ((int64_t)k * size + t) * g.lanes;
      q[k] = prof[N + at] * lanes_at<Q>(aux, m) + prof[at] * strain_term<NDIM, R, Q>(g, f, k, e, A);
      lanes_put(aux, m, q[k]);
    }
  }
  add_stress<NDIM>(g, sig, gn, shear, idx, e, q, A);
}

// the divergence memory of component c along axis e (elastic.cu:532-573)
template <int NDIM, int R, int V, typename T>
__device__ __forceinline__ LanesOf<T, V> force_memory(const ElGeom<T> &g, const ElLayer<T> &P,
                                                      const T *sigma, const T *u1, int64_t idx,
                                                      int lane0, int c, int e, const int *A) {
  using Q = LanesOf<T, V>;
  const int a = A[e], lo = P.lo[e], hi = P.hi[e];
  if (!in_slab<NDIM>(g, a, e, lo, hi, P.phase ? R : 0))
    return Q((T)0);
  int64_t plane, size, off;
  slab_geometry<NDIM>(g, e, lo, hi, A, &plane, &size, &off);
  const int N = g.n[e];
  const bool slow = e < NDIM - 1;
  const int64_t t = slab_at(a, N, lo, hi, slow, plane, off);
  const T *prof = P.profile[e];
  const int at = (c == e ? 2 * N : 0) + a;
  T *aux = P.aux[e] + lane0;
  const int64_t z = ((int64_t)(NDIM + c) * size + t) * g.lanes;
  Line<T> k = field_line(sigma, idx, g, e);
  k.stress = true;
  if (P.phase) {
    k.memory = true;
    k.aux = aux;
    k.profile = prof;
    k.size = size;
    k.plane = plane;
    k.off = off;
    k.row0 = P.phase == 1 ? 0 : 2 * NDIM;
    k.a = a;
    k.lo = lo;
    k.hi = hi;
    k.N = N;
    k.own_half = false;
    k.slow = slow;
    const Q term = divergence_term<NDIM, R, Q>(g, k, c, e, A, g.f[e]);
    if (t >= 0) {
      const Q u = lanes_at<Q>(u1, ((int64_t)c * g.nodes + idx) * g.lanes);
      lanes_put(aux, z, prof[N + at] * (lanes_at<Q>(aux, z) + u));
    }
    return term;
  }
  const Q next = prof[N + at] * lanes_at<Q>(aux, z) +
                 prof[at] * divergence_term<NDIM, R, Q>(g, k, c, e, A, g.f[e]);
  lanes_put(aux, z, next);
  return next;
}

// the adjoint's strain memory settled into its other block (elastic.cu:577-595)
template <int NDIM, int V, typename T>
__device__ __forceinline__ void settle_memory(const ElGeom<T> &g, const ElLayer<T> &P,
                                              const T *sigma, int64_t idx, int lane0, int e,
                                              const int *A) {
  using Q = LanesOf<T, V>;
  const int a = A[e], lo = P.lo[e], hi = P.hi[e];
  if (!in_slab<NDIM>(g, a, e, lo, hi, 0))
    return;
  int64_t plane, size, off;
  slab_geometry<NDIM>(g, e, lo, hi, A, &plane, &size, &off);
  const int N = g.n[e];
  const int64_t t = slab_at(a, N, lo, hi, e < NDIM - 1, plane, off);
  const int from = P.phase == 1 ? 0 : 2 * NDIM, to = 2 * NDIM - from;
  const T *prof = P.profile[e];
  T *aux = P.aux[e] + lane0;
#pragma unroll
  for (int k = 0; k < NDIM; ++k) {
    const int at = (k == e ? 0 : 2 * N) + a;
    const Q s = lanes_at<Q>(sigma, ((int64_t)stress_row<NDIM>(k, e) * g.nodes + idx) * g.lanes);
    lanes_put(aux, ((int64_t)(to + k) * size + t) * g.lanes,
              prof[N + at] * (lanes_at<Q>(aux, ((int64_t)(from + k) * size + t) * g.lanes) + s));
  }
}

// ---------------------------------------------------------------- kernels
// thread -> (interior node, first lane): a node's lanes / V threads are
// neighbours in the wave; false past the last node (whole node groups).
// Blocks go to the XCDs in contiguous runs of nodes (cw::xcd_tile), so the
// rows a block's taps read along the slow axes were fetched into the same
// XCD's L2 by the blocks before it.
template <int NDIM, int V, typename T>
__device__ __forceinline__ bool node_of(const ElGeom<T> &g, int *A, int64_t *idx, int *lane0) {
  const int per = g.lanes / V;
  const int64_t t = (int64_t)cw::xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int64_t i = t / per;
  if (i >= g.interior)
    return false;
  *lane0 = (int)(t % per) * V;
  int64_t rest = i;
#pragma unroll
  for (int d = NDIM - 1; d >= 0; --d) {
    const int w = g.n[d] - 2;
    A[d] = 1 + (int)(rest % w);
    rest /= w;
  }
  int64_t at = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    at += (int64_t)A[d] * g.s[d];
  *idx = at;
  return true;
}

// the per-node sum over the shots of a density, taken off accumulator `acc`
template <typename T, int V>
__device__ __forceinline__ void take_off(T *acc, int64_t at, const LanesOf<T, V> &x, int lanes,
                                         int lane0) {
  const T sum = cw::shot_sum(x, lanes);
  if (lane0 == 0)
    acc[at] -= sum;
}

// the prepared material of each interior node: the inverse inertia of each
// component point (elastic.cu:200-211), then the shear gamma of each pair
// point (elastic.cu:189-197). The reference's CPU fallback (the oracle's
// IEEE path) forms the pair's reciprocals in double precision, `fmax` being
// the C library's double function there, and rounds the sum to T after each
// term; the float build does the same so its bits match.
template <typename T, int NDIM>
__global__ void __launch_bounds__(kThreads)
el_prepare(ElGeom<T> g, const T *__restrict__ gamma, T rho, T *__restrict__ prep) {
  int A[3];
  int64_t idx;
  int lane0;
  ElGeom<T> one = g;
  one.lanes = 1;
  if (!node_of<NDIM, 1>(one, A, &idx, &lane0))
    return;
  const T tiny = Tiny<T>::value;
#pragma unroll
  for (int c = 0; c < NDIM; ++c) {
    bool held = false;
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
      if (d != c && ((A[d] == 1 && clamped_face(g, d, 0)) ||
                     (A[d] == g.n[d] - 2 && clamped_face(g, d, 1))))
        held = true;
    T minv = (T)0;
    if (A[c] <= g.n[c] - 2) {
      const T mean = (T)0.5 * (gamma[idx] + gamma[idx + g.s[c]]);
      minv = held ? (T)0 : (T)1 / fmax(rho * mean, tiny);
    }
    prep[(int64_t)c * g.nodes + idx] = minv;
  }
#pragma unroll
  for (int k = 0; k < NDIM - 1; ++k)
#pragma unroll
    for (int l = k + 1; l < NDIM; ++l) {
      const int v = stress_row<NDIM>(k, l);
      T out = (T)0;
      if (shear_point(g, k, l, A)) {
        const int64_t at[4] = {idx, idx + g.s[l], idx + g.s[k], idx + g.s[k] + g.s[l]};
        T sum = (T)(1.0 / fmax((double)gamma[at[0]], (double)tiny));
#pragma unroll
        for (int i = 1; i < 4; ++i)
          sum = (T)((double)sum + 1.0 / fmax((double)gamma[at[i]], (double)tiny));
        out = (T)4 / sum;
      }
      prep[(int64_t)v * g.nodes + idx] = out;
    }
}

// stress_kernel (elastic.cu:756-795): the stresses of u1 into sigma, the PML
// fold, and in the adjoint pair (GRAD) the stiffness densities of u1 against
// the forward field v1 (stiffness_density, elastic.cu:356-383)
template <typename T, int NDIM, int R, int V, bool GRAD>
__global__ void __launch_bounds__(kThreads)
el_stress(ElGeom<T> g, ElLayer<T> P, const T *__restrict__ gamma, const T *__restrict__ prep,
          const T *__restrict__ u1, T *__restrict__ sigma, T *__restrict__ g_normal,
          T *__restrict__ g_shear, const T *__restrict__ v1) {
  using Q = LanesOf<T, V>;
  constexpr int NV = nvoigt<NDIM>();
  int A[3];
  int64_t idx;
  int lane0;
  if (!node_of<NDIM, V>(g, A, &idx, &lane0))
    return;
  const T *u = u1 + lane0;
  const T *shear = prep + NDIM * g.nodes;
  const T gn = gamma[idx];
  Q eps[NV], sig[NV];
  const int zeroed = strain_point<NDIM, R, Q>(g, u, idx, A, eps);
  if constexpr (GRAD) {
    const T *v = v1 + lane0;
    Q eu[NDIM];
    const int z2 = normal_strains<NDIM, R, Q>(g, v, idx, A, eu);
    take_off<T, V>(g_normal, idx, normal_form<NDIM, Q>(g, eps, eu, z2), g.lanes, lane0);
#pragma unroll
    for (int k = 0; k < NDIM - 1; ++k)
#pragma unroll
      for (int l = k + 1; l < NDIM; ++l) {
        if (!shear_point(g, k, l, A))
          continue;
        const int w = stress_row<NDIM>(k, l);
        const Q d = g.mu * eps[w] * shear_strain<NDIM, R, Q>(g, v, idx, k, l, A);
        take_off<T, V>(g_shear, (int64_t)(w - NDIM) * g.nodes + idx, d, g.lanes, lane0);
      }
  }
  stress_of<NDIM, Q>(g, eps, zeroed, gn, shear, idx, A, sig);
#pragma unroll
  for (int e = 0; e < NDIM; ++e)
    if (P.lo[e] + P.hi[e])
      stress_memory<NDIM, R, V>(g, P, e, u, gn, shear, idx, lane0, A, sig);
#pragma unroll
  for (int w = 0; w < NV; ++w)
    lanes_put(sigma + lane0, ((int64_t)w * g.nodes + idx) * g.lanes, sig[w]);
}

// fd_kernel (elastic.cu:798-853): u2 from u0, u1 and the stresses, the PML
// fold, and in the adjoint pair the inertia density against v1; u2 may be u0
template <typename T, int NDIM, int R, int V, bool GRAD>
__global__ void __launch_bounds__(kThreads)
el_update(ElGeom<T> g, ElLayer<T> P, const T *__restrict__ prep, const T *sigma, const T *u0,
          const T *__restrict__ u1, T *u2, T *__restrict__ g_mass, const T *__restrict__ v1, T mf) {
  using Q = LanesOf<T, V>;
  int A[3];
  int64_t idx;
  int lane0;
  if (!node_of<NDIM, V>(g, A, &idx, &lane0))
    return;
#pragma unroll
  for (int c = 0; c < NDIM; ++c) {
    if (A[c] > g.n[c] - 3)
      continue;
    const Q force = nodal_force<NDIM, R, Q>(g, sigma + lane0, idx, c, A);
    const int64_t n = (int64_t)c * g.nodes + idx;
    const int64_t at = n * g.lanes + lane0;
    const T mi = prep[n];
    const Q w0 = lanes_at<Q>(u0, at), w1 = lanes_at<Q>(u1, at);
    Q next = -w0 + (T)2.f * w1 + mi * force;
#pragma unroll
    for (int e = 0; e < NDIM; ++e)
      if (P.lo[e] + P.hi[e])
        next += mi * force_memory<NDIM, R, V>(g, P, sigma + lane0, u1 + lane0, idx, lane0, c, e, A);
    if constexpr (GRAD) {
      const Q v = lanes_at<Q>(v1, at);
      take_off<T, V>(g_mass, n, mf * v * (next - (T)2.f * w1 + w0), g.lanes, lane0);
    }
    lanes_put(u2, at, next);
  }
  if (P.phase) {
#pragma unroll
    for (int e = 0; e < NDIM; ++e)
      if (P.lo[e] + P.hi[e])
        settle_memory<NDIM, V>(g, P, sigma + lane0, idx, lane0, e, A);
  }
}

// gradient_kernel (elastic_sensitivity.cu:115-158): both densities of the
// forward triplet u0, u1, u2 against the adjoint l1; g.f holds 1 / dx
template <typename T, int NDIM, int R, int V>
__global__ void __launch_bounds__(kThreads)
el_gradient(ElGeom<T> g, T *__restrict__ g_mass, T *__restrict__ g_normal,
            T *__restrict__ g_shear, const T *__restrict__ u0, const T *__restrict__ u1,
            const T *__restrict__ u2, const T *__restrict__ l1, T mf) {
  using Q = LanesOf<T, V>;
  int A[3];
  int64_t idx;
  int lane0;
  if (!node_of<NDIM, V>(g, A, &idx, &lane0))
    return;
#pragma unroll
  for (int c = 0; c < NDIM; ++c) {
    if (A[c] > g.n[c] - 3)
      continue;
    const int64_t n = (int64_t)c * g.nodes + idx;
    const int64_t at = n * g.lanes + lane0;
    const Q l = lanes_at<Q>(l1, at);
    const Q w0 = lanes_at<Q>(u0, at), w1 = lanes_at<Q>(u1, at), w2 = lanes_at<Q>(u2, at);
    take_off<T, V>(g_mass, n, mf * l * (w2 - (T)2.f * w1 + w0), g.lanes, lane0);
  }
  Q eu[NDIM], el[NDIM];
  const int zeroed = normal_strains<NDIM, R, Q>(g, u1 + lane0, idx, A, eu);
  normal_strains<NDIM, R, Q>(g, l1 + lane0, idx, A, el);
  take_off<T, V>(g_normal, idx, normal_form<NDIM, Q>(g, el, eu, zeroed), g.lanes, lane0);
#pragma unroll
  for (int k = 0; k < NDIM - 1; ++k)
#pragma unroll
    for (int l = k + 1; l < NDIM; ++l) {
      if (!shear_point(g, k, l, A))
        continue;
      const int w = stress_row<NDIM>(k, l);
      const Q d = g.mu * shear_strain<NDIM, R, Q>(g, l1 + lane0, idx, k, l, A) *
                  shear_strain<NDIM, R, Q>(g, u1 + lane0, idx, k, l, A);
      take_off<T, V>(g_shear, (int64_t)(w - NDIM) * g.nodes + idx, d, g.lanes, lane0);
    }
}

// --------------------------------------------------------------- points
// entry e of a point list reads or writes element lin[e] * lanes + b of a
// field for shot lane b (every lane when shot[e] < 0, else lane shot[e]) and
// element row * rs + col[e] * cs + b * ss of a table
struct ElPoints {
  const int32_t *starts; // (runs + 1) CSR: the entries of a run share one node
  const int32_t *lin, *shot, *col;
  int32_t runs;
  int64_t rs, cs, ss;
};

// kind 0: u += weight * table (injection, a run's entries in order);
// 1: table = u (record); 2: u = table (strip replay); 3: the adjoint
// injection, u += weight * table and g_mass += -mf u1 weight * table, lane
// after lane (cuwave/kernels/common.cuh:456-503)
template <typename T>
__global__ void __launch_bounds__(kThreads)
el_points(ElPoints p, int kind, int lanes, int nbatch, int64_t row, T *__restrict__ u,
          T *__restrict__ table, const T *__restrict__ weight, T *__restrict__ g_mass,
          const T *__restrict__ u1, T mf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kind == 3) {
    if (t >= p.runs)
      return;
    for (int e = p.starts[t]; e < p.starts[t + 1]; ++e) {
      const int64_t n = p.lin[e];
      for (int b = 0; b < nbatch; ++b) {
        if (p.shot[e] >= 0 && p.shot[e] != b)
          continue;
        const T load = weight[e] * table[row * p.rs + p.col[e] * p.cs + b * p.ss];
        u[n * lanes + b] += load;
        g_mass[n] += -mf * u1[n * lanes + b] * load;
      }
    }
    return;
  }
  const int64_t run = t / lanes;
  const int b = (int)(t % lanes);
  if (run >= p.runs || b >= nbatch)
    return;
  for (int e = p.starts[run]; e < p.starts[run + 1]; ++e) {
    if (p.shot[e] >= 0 && p.shot[e] != b)
      continue;
    const int64_t n = (int64_t)p.lin[e] * lanes + b;
    const int64_t m = row * p.rs + p.col[e] * p.cs + b * p.ss;
    if (kind == 0)
      u[n] += weight[e] * table[m];
    else if (kind == 1)
      table[m] = u[n];
    else
      u[n] = table[m];
  }
}

} // namespace cwe
