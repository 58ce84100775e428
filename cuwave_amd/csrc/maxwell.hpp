// Device code of the vector Maxwell family for gfx950.
//
// Maxwell's equations in second-order curl-curl form on the Yee lattice
// (cuwave/maxwell.py:162-260): electric component c sits half a node up axis
// c, the curl of the pair (k, l) half a node up both axes. A step is a pair
// of launches over the interior nodes: the curl launch writes nu W b of each
// pair point of the node into a scratch (b the pair's curl of u1, W its cell
// weight), the update launch gathers the curls into each component's force
// and writes u2 = -u0 + 2 u1 + minv force, with the adjoint's accumulator
// also the inertia density of u1 against the forward v1
// (cuwave/kernels/maxwell.cu:44-275).
//
// The walls read images, never a ghost: a field node past the wall node reads
// its mirror about it, negated across a conductor; a pair point past the
// wall's half point reads its mirror about that, negated across a magnetic
// wall, the opposite parity, so the gather is the curl's exact transpose.
//
// Layout as in the other families: per-shot arrays keep the shots
// innermost, (component, *np, lanes); a thread holds V neighbouring lanes of
// one interior node and runs every expression lane by lane in the
// reference's operation order, so a lane's bits are those of the one-shot
// arithmetic. The densities a node's lanes produce are summed over the lanes
// into one per-point accumulator.
#pragma once

#include "common.hpp"

namespace cwm {

using cw::lanes_at;
using cw::lanes_put;
using cw::LanesOf;

constexpr int kThreads = 256;

// the grid and the operator constants, by value
template <typename T> struct MxGeom {
  int n[3], np[3], s[3];
  int64_t nodes;  // nodes of one component
  int lanes;
  int conductor;  // Conductor faces, bit 2 axis + side
  int interior;   // interior nodes, every axis 1 .. n - 2
  T sg[4];        // full-radius staggered weights, sg[k - 1]
  T nu;           // 1 / permeability
  T f[3];         // per-axis factors: 1 / dx (curl), dt^2 / dx (update)
};

// the curl scratch row of the pair (k, l), k < l: 3 - k - l in 3D
template <int NDIM> __device__ __forceinline__ constexpr int pair_row(int k, int l) {
  return NDIM == 3 ? 3 - k - l : 0;
}

// node b of an axis of N nodes, past the wall node its mirror about it;
// `neg` across a face of `odd` (bit 0 low, bit 1 high)
__device__ __forceinline__ int node_image(int b, int N, int odd, bool &neg) {
  neg = (b < 1 && (odd & 1)) || (b > N - 2 && (odd & 2));
  return b < 1 ? 2 - b : b > N - 2 ? 2 * (N - 2) - b : b;
}

// half point b (half a node up from node b), past the wall's half point its
// mirror about it; `neg` across a face of `odd`
__device__ __forceinline__ int half_image(int b, int N, int odd, bool &neg) {
  neg = (b < 1 && (odd & 1)) || (b > N - 3 && (odd & 2));
  return b < 1 ? 1 - b : b > N - 3 ? 2 * N - 5 - b : b;
}

// the V lanes of point a + j of the line through element e (node a, element
// stride se), through `node_image` or `half_image`
template <bool HALF, typename Q, typename T>
__device__ __forceinline__ Q imaged(const T *__restrict__ u, int64_t e, int64_t se, int a,
                                    int N, int odd, int j) {
  bool neg;
  const int b = HALF ? half_image(a + j, N, odd, neg) : node_image(a + j, N, odd, neg);
  const Q v = lanes_at<Q>(u, e + (int64_t)(b - a) * se);
  return neg ? -v : v;
}

// the interior node of this thread and the first of its V lanes
template <int NDIM, int V, typename T>
__device__ __forceinline__ bool mx_node(const MxGeom<T> &g, int *A, int64_t *idx,
                                        int *lane0) {
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

// the (k, l) curl at the node's pair point (maxwell.cu:49-73): u_k along l
// and u_l along k, nodes 1 - R .. R away, the first term negated
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q curl_of(const MxGeom<T> &g, const T *__restrict__ u, int64_t e,
                                     int k, int l, const int *A) {
  const int64_t comp = g.nodes * g.lanes;
  const T *uk = u + k * comp, *ul = u + l * comp;
  const int64_t sl = (int64_t)g.s[l] * g.lanes, sk = (int64_t)g.s[k] * g.lanes;
  const int ol = g.conductor >> (2 * l), ok = g.conductor >> (2 * k);
  Q b = (T)0;
#pragma unroll
  for (int j = 1; j <= R; ++j) {
    const Q dk = imaged<false, Q>(uk, e, sl, A[l], g.n[l], ol, j) -
                 imaged<false, Q>(uk, e, sl, A[l], g.n[l], ol, 1 - j);
    const Q dl = imaged<false, Q>(ul, e, sk, A[k], g.n[k], ok, j) -
                 imaged<false, Q>(ul, e, sk, A[k], g.n[k], ok, 1 - j);
    b = b - (g.f[l] * g.sg[j - 1]) * dk;
    b = b + (g.f[k] * g.sg[j - 1]) * dl;
  }
  return b;
}

// curl_kernel (maxwell.cu:209-239): every pair point of the node off the
// staggered ghosts of its two axes, nu times its cell weight (halved on the
// walls of the axis it does not span) times the curl; MAG reads that product
// per pair point from nu_pair (a designed permeability), and GRAD then takes
// the adjoint curl against the forward one off g_nu, summed over the lanes
template <typename T, int NDIM, int R, int V, bool MAG, bool GRAD>
__global__ void __launch_bounds__(kThreads)
mx_curl(MxGeom<T> g, const T *__restrict__ u1, T *__restrict__ h,
        const T *__restrict__ nu_pair, T *__restrict__ g_nu, const T *__restrict__ v1) {
  using Q = LanesOf<T, V>;
  int A[3];
  int64_t idx;
  int lane0;
  if (!mx_node<NDIM, V>(g, A, &idx, &lane0))
    return;
  const int64_t e = idx * g.lanes + lane0;
#pragma unroll
  for (int k = 0; k < NDIM - 1; ++k)
#pragma unroll
    for (int l = k + 1; l < NDIM; ++l) {
      if (A[k] > g.n[k] - 3 || A[l] > g.n[l] - 3)
        continue;
      T w = (T)1;
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
        if (d != k && d != l && (A[d] == 1 || A[d] == g.n[d] - 2))
          w *= (T)0.5;
      const Q b = curl_of<NDIM, R, Q>(g, u1, e, k, l, A);
      const int64_t p = pair_row<NDIM>(k, l) * g.nodes + idx;
      if constexpr (GRAD) {
        const T s = cw::shot_sum(b * curl_of<NDIM, R, Q>(g, v1, e, k, l, A), g.lanes);
        if (lane0 == 0)
          g_nu[p] -= s;
      }
      if constexpr (MAG)
        lanes_put(h, p * g.lanes + lane0, nu_pair[p] * b);
      else
        lanes_put(h, p * g.lanes + lane0, (g.nu * w) * b);
    }
}

// the force on component c (maxwell.cu:171-204): the curls of its pairs
// differenced along the other axis, pair points -R .. R - 1 away, halved on
// a wall node of that axis
template <int NDIM, int R, typename Q, typename T>
__device__ __forceinline__ Q force_on(const MxGeom<T> &g, const T *__restrict__ h, int64_t e,
                                      int c, const int *A) {
  Q force = (T)0;
#pragma unroll
  for (int m = 0; m < NDIM; ++m) {
    if (m == c)
      continue;
    const T *hp = h + pair_row<NDIM>(c < m ? c : m, c < m ? m : c) * g.nodes * g.lanes;
    const int64_t sm = (int64_t)g.s[m] * g.lanes;
    const int odd = ~g.conductor >> (2 * m);
    Q acc = (T)0;
#pragma unroll
    for (int j = 1; j <= R; ++j) {
      acc = acc + g.sg[j - 1] * imaged<true, Q>(hp, e, sm, A[m], g.n[m], odd, j - 1);
      acc = acc - g.sg[j - 1] * imaged<true, Q>(hp, e, sm, A[m], g.n[m], odd, -j);
    }
    if (A[m] == 1 || A[m] == g.n[m] - 2)
      acc = acc * (T)0.5;
    // u_c enters its pair as +d_m u_c for m < c, as -d_m u_c for m > c
    force = force + ((m < c ? (T)1 : (T)-1) * g.f[m]) * acc;
  }
  return force;
}

// fd_kernel (maxwell.cu:242-274): every component of the node off the
// staggered ghost of its own axis; GRAD adds the inertia density of u1
// against the forward v1, summed over the lanes
template <typename T, int NDIM, int R, int V, bool GRAD>
__global__ void __launch_bounds__(kThreads)
mx_update(MxGeom<T> g, const T *__restrict__ minv, const T *__restrict__ h, const T *u0,
          const T *__restrict__ u1, T *u2, T *__restrict__ g_mass, const T *__restrict__ v1,
          T mf, const T *__restrict__ damping, T dt) {
  using Q = LanesOf<T, V>;
  int A[3];
  int64_t idx;
  int lane0;
  if (!mx_node<NDIM, V>(g, A, &idx, &lane0))
    return;
  const int64_t e = idx * g.lanes + lane0;
#pragma unroll
  for (int c = 0; c < NDIM; ++c) {
    if (A[c] > g.n[c] - 3)
      continue;
    const Q force = force_on<NDIM, R, Q>(g, h, e, c, A);
    const int64_t n = (int64_t)c * g.nodes + idx;
    const int64_t at = n * g.lanes + lane0;
    const T mi = minv[n];
    const Q w0 = lanes_at<Q>(u0, at), w1 = lanes_at<Q>(u1, at);
    Q next;
    if (damping) {
      // the damped update (maxwell.cu:257-264), damping nodal
      const T beta = (T)0.5 * mi * damping[idx] * dt;
      next = ((T)2 * w1 - w0 * ((T)1 - beta) + mi * force) / ((T)1 + beta);
    } else {
      next = -w0 + (T)2 * w1 + mi * force;
    }
    if constexpr (GRAD) {
      const Q v = lanes_at<Q>(v1, at);
      const T sum = cw::shot_sum(mf * v * (next - (T)2 * w1 + w0), g.lanes);
      if (lane0 == 0)
        g_mass[n] -= sum;
    }
    lanes_put(u2, at, next);
  }
}

} // namespace cwm
