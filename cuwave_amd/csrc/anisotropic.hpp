// Device code of the cell-assembled (anisotropic) elastic family for gfx950.
//
// Collocated displacement components on the nodes, the operator assembled
// cell by cell as -B^T C B with C on the cell (cuwave/anisotropic.py:1-12):
// each interior node gathers, from each of the 2^ndim cells it is a corner
// of, the cell's material g times the row of the one-cell stencil K that
// belongs to its corner, applied to the displacements of the cell's corners
// (cuwave/kernels/anisotropic.cu:72-116, 334-365). A cell takes part while
// both of its corners along every axis are interior nodes, so a traction-free
// face is the natural condition of the interior-cell sum and a clamped one a
// zeroed inertia. Order 2 only: one cell reaches one node each way.
//
// Layout as in the other families: per-shot arrays keep the shots
// innermost, (component, *np, lanes); a thread holds V neighbouring lanes of
// one interior node and runs every expression lane by lane in the
// reference's operation order. K sits in device memory, read with uniform
// (scalar) loads: every index into it is a compile-time constant.
#pragma once

#include "common.hpp"

namespace cwa {

using cw::lanes_at;
using cw::lanes_put;
using cw::LanesOf;

constexpr int kThreads = 256;

template <typename T> struct AnGeom {
  int n[3], s[3];
  int64_t nodes;  // nodes of one component
  int lanes;
  int interior;   // interior nodes, every axis 1 .. n - 2
  T f;            // the update factor dt^2 (the spacing sits in K)
};

template <int NDIM> constexpr int corners() { return 1 << NDIM; }
template <int NDIM> constexpr int nloc() { return corners<NDIM>() * NDIM; }

// the interior node of this thread and the first of its V lanes
template <int NDIM, int V, typename T>
__device__ __forceinline__ bool an_node(const AnGeom<T> &g, int *A, int64_t *idx, int *lane0) {
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

// fd_kernel without damping (anisotropic.cu:367-397): the cell-gathered force
// on every component, then u2 = -u0 + 2 u1 + f minv force. The cells run in
// block order (axis 0 fastest, the cell below the node first), then the
// corners of a cell in the same order, then the components.
template <typename T, int NDIM, int V>
__global__ void __launch_bounds__(kThreads)
an_step(AnGeom<T> g, const T *__restrict__ K, const T *__restrict__ minv,
        const T *__restrict__ cell, const T *u0, const T *__restrict__ u1, T *u2,
        const T *__restrict__ damping, T dt) {
  using Q = LanesOf<T, V>;
  constexpr int NL = nloc<NDIM>();
  int A[3];
  int64_t idx;
  int lane0;
  if (!an_node<NDIM, V>(g, A, &idx, &lane0))
    return;
  const int64_t comp = g.nodes * g.lanes;
  const T *v = u1 + idx * g.lanes + lane0;
  Q force[NDIM];
#pragma unroll
  for (int i = 0; i < NDIM; ++i)
    force[i] = (T)0;
#pragma unroll
  for (int c = 0; c < corners<NDIM>(); ++c) {
    bool inside = true;
    int base = 0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const int e = ((c >> d) & 1) - 1; // the cell's low corner, from the node
      inside &= A[d] + e >= 1 && A[d] + e <= g.n[d] - 3;
      base += e * g.s[d];
    }
    if (!inside)
      continue;
    const int self = corners<NDIM>() - 1 - c; // the node's corner in the cell
    const T gc = cell[idx + base];
#pragma unroll
    for (int l = 0; l < corners<NDIM>(); ++l) {
      int off = base;
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
        off += ((l >> d) & 1) * g.s[d];
#pragma unroll
      for (int j = 0; j < NDIM; ++j) {
        const Q uj = lanes_at<Q>(v, j * comp + (int64_t)off * g.lanes);
#pragma unroll
        for (int i = 0; i < NDIM; ++i)
          force[i] = force[i] - (gc * K[(self * NDIM + i) * NL + l * NDIM + j]) * uj;
      }
    }
  }
  const T mi = minv[idx];
  // with a damping field the damped update (anisotropic.cu:380-389), nodal beta
  const T beta = damping ? (T)0.5 * mi * damping[idx] * dt : (T)0;
#pragma unroll
  for (int i = 0; i < NDIM; ++i) {
    const int64_t at = (i * g.nodes + idx) * g.lanes + lane0;
    const Q w0 = lanes_at<Q>(u0, at), w1 = lanes_at<Q>(u1, at);
    const Q next = damping ? ((T)2 * w1 - w0 * ((T)1 - beta) + g.f * mi * force[i]) / ((T)1 + beta)
                           : -w0 + (T)2 * w1 + (g.f * mi) * force[i];
    lanes_put(u2, at, next);
  }
}

// gradient_kernel (anisotropic_sensitivity.cu:111-156): the inertia density of
// the forward triplet against the adjoint l1 at the node, and the stiffness
// density of the cell whose low corner the node is, the one-cell stencil
// paired between l1 and u1; each summed over the lanes into its accumulator
template <typename T, int NDIM, int V>
__global__ void __launch_bounds__(kThreads)
an_gradient(AnGeom<T> g, const T *__restrict__ K, T *__restrict__ g_mass,
            T *__restrict__ g_cell, const T *__restrict__ u0, const T *__restrict__ u1,
            const T *__restrict__ u2, const T *__restrict__ l1, T mass_factor) {
  using Q = LanesOf<T, V>;
  constexpr int NL = nloc<NDIM>();
  int A[3];
  int64_t idx;
  int lane0;
  if (!an_node<NDIM, V>(g, A, &idx, &lane0))
    return;
  const int64_t comp = g.nodes * g.lanes;
  const int64_t e = idx * g.lanes + lane0;
  Q acc = (T)0;
#pragma unroll
  for (int i = 0; i < NDIM; ++i) {
    const int64_t at = i * comp + e;
    acc = acc + lanes_at<Q>(l1, at) *
                    (lanes_at<Q>(u2, at) - (T)2 * lanes_at<Q>(u1, at) + lanes_at<Q>(u0, at));
  }
  const T mass = cw::shot_sum(mass_factor * acc, g.lanes);
  bool own = true;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    own &= A[d] <= g.n[d] - 3;
  Q bilinear = (T)0;
  if (own) {
#pragma unroll
    for (int l = 0; l < corners<NDIM>(); ++l) {
      int lo = 0;
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
        lo += ((l >> d) & 1) * g.s[d];
#pragma unroll
      for (int m = 0; m < corners<NDIM>(); ++m) {
        int mo = 0;
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
          mo += ((m >> d) & 1) * g.s[d];
#pragma unroll
        for (int i = 0; i < NDIM; ++i)
#pragma unroll
          for (int j = 0; j < NDIM; ++j)
            bilinear = bilinear + (lanes_at<Q>(l1, i * comp + e + (int64_t)lo * g.lanes) *
                                   K[(l * NDIM + i) * NL + m * NDIM + j]) *
                                      lanes_at<Q>(u1, j * comp + e + (int64_t)mo * g.lanes);
      }
    }
  }
  // every thread of the node takes part in both sums; one writer a node
  const T stiff = cw::shot_sum(bilinear, g.lanes);
  if (lane0 == 0) {
    g_mass[idx] -= mass;
    if (own)
      g_cell[idx] -= stiff;
  }
}

} // namespace cwa
