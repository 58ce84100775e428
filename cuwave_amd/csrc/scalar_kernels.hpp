// The pressure-family kernels for gfx950, general form: any dimension, walls,
// PML, points. On a 2D grid they cover only the band of nodes within reach of
// a wall (the streaming kernels of stream.hpp take everything else); in 3D and
// on grids too small for a stream they cover every node.
//
// Every kernel rounds as the reference's kernels do on its IEEE CPU path: no
// contraction (built with -ffp-contract=off), an fma exactly where the
// reference pins one (cuwave/kernels/scalar.cu:14-33, 416-436), correctly
// rounded division. The harmonic cell means sc sp / (sc + sp)
// (cuwave/kernels/scalar.cu:63) are computed once by `prepare_kernel`, in
// the same operand order, so reading them is bitwise the in-place division;
// the inverse inertia 1 / sc likewise.
//
// A launch's second grid axis is the shot group: a thread steps its node for
// the `group` shots of its group, which share the material loads and, in the
// adjoint, one gradient accumulator of the caller's (cw_scalar_grad_groups).
#pragma once

#include "common.hpp"

namespace cw {

template <typename T> struct Points {
  const int32_t *starts; // (ntiles + 1) CSR offsets, or null
  const int32_t *lin;
  const int32_t *shot; // -1 for every shot
  const int32_t *col;
  const T *weight;
  int columns;
  int64_t shot_stride;
};

template <typename T> struct Levels {
  T *base;
  int64_t slot, batch;
  int slots, offset;
  __device__ __forceinline__ T *at(int level) const {
    int k = (level + offset) % slots;
    k += k < 0 ? slots : 0;
    return base + (int64_t)k * slot;
  }
};

// the adjoint memory of a slab node: the recursion's three variables, then
// the two tap fields its neighbours' gathers read (scalar_sensitivity.cu:
// 232-251 `pml_tap`, computed once by the owner instead of by every reader)
constexpr int kEta = 5;

template <typename T> struct Layer {
  int lo[3], hi[3];
  const T *profile[3];
  T *aux[3];
  T *eta[3][2];
};

template <typename T> struct Strip {
  const int32_t *map;
  void *record;
  int count, first, rows;
  int bf16;
};

// where the general kernels run: every tile, or (2D) the band of nodes
// outside the stream box rows [y0, y1) x columns [x0, x1)
struct Region {
  int band;
  int y0, y1, x0, x1;
  int count; // band nodes
};

template <typename T> struct FwdParams {
  Geom<T> g;
  Region region;
  int group; // shots a thread steps
  const T *inv, *cell;
  Layer<T> pml;
  Levels<T> u;
  Points<T> src;
  const T *signal;
  Points<T> rec;
  T *record;
  Strip<T> strip;
  int steps;
};

template <typename T> struct BwdParams {
  Geom<T> g;
  Region region;
  int group;
  const T *inv, *cell, *stiff;
  Layer<T> pml;
  Levels<T> lam, u;
  int reverse;
  Points<T> rec;
  const T *adjoint;
  Points<T> src;
  const T *signal;
  Strip<T> strip;
  T *gmass, *gstiff;
  int64_t gbatch;
  int combine, stiff_grad, skip_layer;
  T jm, js, mf;
  int steps;
};

// ------------------------------------------------------------------ helpers
// the cell flux divergence of one axis over taps t[0 .. 2R] centred on R,
// rounded as cuwave/kernels/scalar.cu:14-33 pins it
template <typename T, int R>
__device__ __forceinline__ T flux(const T *t, T gp, T gm, const T *op) {
  const T uc = t[R];
  T Dp = t[R + 1] - uc, Dm = uc - t[R - 1];
  if (R > 1) {
    Dp = op[0] * Dp;
    Dm = op[0] * Dm;
  }
#pragma unroll
  for (int k = 2; k <= R; ++k) {
    Dp = fmat(op[k - 1], t[R + k] - t[R + 1 - k], Dp);
    Dm = fmat(op[k - 1], t[R + k - 1] - t[R - k], Dm);
  }
  const T mm = Dm * gm;
  return fmat(Dp, gp, -mm);
}

// the axes combined as cuwave/kernels/scalar.cu:416-436 rounds them
template <typename T, int NDIM, int R>
__device__ __forceinline__ T combine_axes(const T *d, const T *f) {
  if (NDIM == 1)
    return f[0] * d[0];
  if (NDIM == 2) {
    if (R == 1) {
      const T p = f[1] * d[1];
      return fmat(f[0], d[0], p);
    }
    const T p = f[0] * d[0];
    return fmat(f[1], d[1], p);
  }
  const T p = f[1] * d[1];
  return fmat(f[2], d[2], fmat(f[0], d[0], p));
}

// the two cell differences with plain sums (cuwave/kernels/scalar.cu:68-82,
// scalar_sensitivity.cu:116-135), over taps centred on R
template <typename T, int R>
__device__ __forceinline__ void cell_diffs(const T *t, const T *op, T &Dp,
                                           T &Dm) {
  const T uc = t[R];
  Dp = op[0] * (t[R + 1] - uc);
  Dm = op[0] * (uc - t[R - 1]);
#pragma unroll
  for (int k = 2; k <= R; ++k) {
    Dp += op[k - 1] * (t[R + k] - t[R + 1 - k]);
    Dm += op[k - 1] * (t[R + k - 1] - t[R - k]);
  }
}

template <typename T> __device__ __forceinline__ T harmonic(T a, T b) {
  return a * b / (a + b);
}

// the forward CPML of axis d at a slab node (cuwave/kernels/scalar.cu:
// 499-532): the three memories updated in place, their share of the field
// returned. `an .. bh1` are the profile at the node and at the half point
// below it.
template <typename T>
__device__ __forceinline__ T pml_forward(T *aux, int64_t size, T Dp, T Dm,
                                         T gp, T gm, T f, T mi, T an, T bn,
                                         T ah, T bh, T ah1, T bh1) {
  const T pp = bh * aux[0] + ah * Dp;
  const T pm = bh1 * aux[size] + ah1 * Dm;
  const T z = bn * aux[2 * size] + an * f * (gp * (Dp + pp) - gm * (Dm + pm));
  aux[0] = pp;
  aux[size] = pm;
  aux[2 * size] = z;
  return mi * (f * (gp * pp - gm * pm) + z);
}

// the adjoint memory of a slab node stepped by lambda (scalar_sensitivity.cu:
// 517-542 with q = W lambda), then the two tap fields of the new memory
// (:232-251): cur and nxt address the node's kEta rows, `size` apart
template <typename T>
__device__ __forceinline__ void pml_adjoint_update(const T *cur, T *nxt,
                                                   int64_t size, T q, T gp,
                                                   T gm, T f, T an, T bn,
                                                   T ah, T bh, T ah1, T bh1) {
  const T ez = cur[2 * size];
  const T ep = bh * (cur[0] + an * f * gp * ez) + f * gp * q;
  const T em = bh1 * (cur[size] - an * f * gm * ez) - f * gm * q;
  const T zz = bn * ez + q;
  nxt[0] = ep;
  nxt[size] = em;
  nxt[2 * size] = zz;
  const T e2 = an * f * zz;
  nxt[3 * size] = ah * ep + gp * (1 + ah) * e2;
  nxt[4 * size] = ah1 * em - gm * (1 + ah1) * e2;
}

// d(Dp)/d(u at row q) and d(Dm)/d(u at row q) (scalar_sensitivity.cu:219-230)
template <typename T, int R>
__device__ __forceinline__ T plus_tap(int q, const T *op) {
  if (q >= 1)
    return q <= R ? op[q - 1] : (T)0;
  return 1 - q <= R ? -op[-q] : (T)0;
}
template <typename T, int R>
__device__ __forceinline__ T minus_tap(int q, const T *op) {
  if (q >= 0)
    return q + 1 <= R ? op[q] : (T)0;
  return -q <= R ? -op[-q - 1] : (T)0;
}

// scalar_sensitivity.cu:137-154 transposed_divergence_taps over taps
// l[0 .. 2R] and cell means: hi[j - 1] the cell (a + j - 1, a + j),
// lo[j - 1] the cell (a - j, a - j + 1)
template <typename T, int R>
__device__ __forceinline__ T transposed(const T *l, const T *chi,
                                        const T *clo, const T *op) {
  T div = 0, lh = l[R], ll = l[R];
#pragma unroll
  for (int j = 1; j <= R; ++j) {
    const T lh1 = l[R + j], ll1 = l[R - j];
    div += op[j - 1] * (chi[j - 1] * (lh1 - lh) - clo[j - 1] * (ll - ll1));
    lh = lh1;
    ll = ll1;
  }
  return div;
}

// scalar_sensitivity.cu:116-135 stiffness_gradient_taps
template <typename T, int R>
__device__ __forceinline__ T stiffness_density(const T *u, const T *l, T lc,
                                               T sc, T sp, T sm, T factor,
                                               const T *op) {
  const T dgp = sp * sp / ((sc + sp) * (sc + sp));
  const T dgm = sm * sm / ((sc + sm) * (sc + sm));
  T Dp, Dm;
  cell_diffs<T, R>(u, op, Dp, Dm);
  return factor * (dgp * Dp * (l[R + 1] - lc) + dgm * Dm * (lc - l[R - 1]));
}

// ------------------------------------------------------------ node mapping
// the 2D band node t, or false past the band
template <typename T>
__device__ __forceinline__ bool band_node(const Region &r, const Geom<T> &g,
                                          int t, int *a) {
  const int W = g.n[1] - 2;
  const int top = (r.y0 - 1) * W;
  if (t < top) {
    a[0] = 1 + t / W;
    a[1] = 1 + t % W;
    return true;
  }
  t -= top;
  const int bot = (g.n[0] - 1 - r.y1) * W;
  if (t < bot) {
    a[0] = r.y1 + t / W;
    a[1] = 1 + t % W;
    return true;
  }
  t -= bot;
  const int left = r.x0 - 1, right = g.n[1] - 1 - r.x1, per = left + right;
  if (per <= 0 || t >= per * (r.y1 - r.y0))
    return false;
  const int k = t % per;
  a[0] = r.y0 + t / per;
  a[1] = k < left ? 1 + k : r.x1 + (k - left);
  return true;
}

// this thread's node: its axes a, its tile, and whether every tap of its
// tile lies on stored nodes at distance `reach` (a tile-wide flag)
template <int NDIM, typename T>
__device__ __forceinline__ bool general_node(const Region &r, const Geom<T> &g,
                                             int reach, int *a, int &tile,
                                             bool &deep) {
  if (NDIM == 2 && r.band) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= r.count || !band_node(r, g, t, a))
      return false;
    tile = (a[0] / TileShape<2>::TY) * g.tf + a[1] / TileShape<2>::TX;
    deep = false;
    return true;
  }
  tile = xcd_tile(blockIdx.x, g.ntiles);
  int lo[3], hi[3];
  tile_node<NDIM>(g, tile, a, lo, hi);
  bool inside = true;
  deep = true;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    inside &= a[d] > 0 && a[d] < g.n[d] - 1;
    deep &= lo[d] >= reach && hi[d] <= g.n[d] - 1 - reach;
  }
  return inside;
}

// ----------------------------------------------------------- preparation
// inv = minv or 1 / stiff; cell[d] = the harmonic mean of the cell (a, a + 1)
// along axis d, for every node that has a next node along d
template <typename T, int NDIM>
__global__ __launch_bounds__(kThreads) void prepare_kernel(
    Geom<T> g, const T *__restrict__ stiff, const T *__restrict__ minv,
    T *__restrict__ inv, T *__restrict__ cell) {
  const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= g.field)
    return;
  int a[3];
  int64_t rest = idx;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    a[d] = (int)(rest / g.s[d]);
    rest -= (int64_t)a[d] * g.s[d];
  }
  const T sc = stiff[idx];
  inv[idx] = minv ? minv[idx] : (T)1 / sc;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    cell[d * g.field + idx] =
        a[d] + 1 < g.np[d] ? harmonic(sc, stiff[idx + g.s[d]]) : (T)0;
}

// ------------------------------------------------------------ forward step
template <typename T, int NDIM, int R>
__global__ __launch_bounds__(kThreads) void forward_kernel(const FwdParams<T> p,
                                                           const int t) {
  const Geom<T> &g = p.g;
  int a[3], tile;
  bool deep;
  if (!general_node<NDIM>(p.region, g, R, a, tile, deep))
    return;
  int64_t idx = 0;
  bool held = false;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    idx += (int64_t)a[d] * g.s[d];
    held |= held_wall(a[d], g.n[d], g.dirichlet >> (2 * d));
  }
  const T mi = p.inv[idx];
  T gp[NDIM], gm[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    gp[d] = p.cell[d * g.field + idx];
    gm[d] = p.cell[d * g.field + idx - g.s[d]];
  }
  int prow[NDIM];
  bool layer = false;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    prow[d] = p.pml.lo[d] + p.pml.hi[d]
                  ? slab_row(a[d], g.n[d], p.pml.lo[d], p.pml.hi[d])
                  : -1;
    layer |= prow[d] >= 0;
  }
  int s0 = 0, s1 = 0, r0 = 0, r1 = 0;
  if (p.src.starts) {
    s0 = p.src.starts[tile];
    s1 = p.src.starts[tile + 1];
  }
  if (p.rec.starts) {
    r0 = p.rec.starts[tile];
    r1 = p.rec.starts[tile + 1];
  }
  int strip_col = -1;
  if (layer && p.strip.map)
    strip_col = p.strip.map[idx];
  T *const u2s = p.u.at(t);
  const T *const u1s = p.u.at(t - 1);
  const T *const u0s = p.u.at(t - 2);
  const int b0 = blockIdx.y * p.group;
  const int b1 = min(b0 + p.group, g.nbatch);

  for (int b = b0; b < b1; ++b) {
    const int64_t ob = (int64_t)b * p.u.batch;
    const T *__restrict__ u1 = u1s + ob;
    T taps[NDIM][2 * R + 1];
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
#pragma unroll
      for (int j = -R; j <= R; ++j)
        taps[d][j + R] = deep ? u1[idx + (int64_t)j * g.s[d]]
                              : mirrored(u1, idx, g.s[d], a[d], g.n[d],
                                         g.dirichlet >> (2 * d), j);
    T div[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
      div[d] = flux<T, R>(taps[d], gp[d], gm[d], g.op);
    const T lap = combine_axes<T, NDIM, R>(div, g.f);
    const T uc = taps[0][R];
    T v = -u0s[ob + idx] + (T)2 * uc + mi * lap;
    if (held)
      v = 0;
    if (layer) {
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        if (prow[d] < 0)
          continue;
        const int N = g.n[d], L = p.pml.lo[d], H = p.pml.hi[d], ad = a[d];
        const T *an = p.pml.profile[d], *bn = an + N, *ah = an + 2 * N,
                *bh = an + 3 * N;
        const int64_t size = slab_size<NDIM>(g, d, L, H);
        T *aux = p.pml.aux[d] + (int64_t)b * 3 * size +
                 slab_index<NDIM>(g, a, d, prow[d], L, H);
        T Dp, Dm;
        cell_diffs<T, R>(taps[d], g.op, Dp, Dm);
        const T share = pml_forward(aux, size, Dp, Dm, gp[d], gm[d], g.f[d], mi,
                                    an[ad], bn[ad], ah[ad], bh[ad], ah[ad - 1],
                                    bh[ad - 1]);
        if (!held)
          v += share;
      }
    }
    for (int e = s0; e < s1; ++e)
      if (p.src.lin[e] == idx && (p.src.shot[e] < 0 || p.src.shot[e] == b))
        v += p.src.weight[e] *
             p.signal[b * p.src.shot_stride + (int64_t)t * p.src.columns +
                      p.src.col[e]];
    T *__restrict__ u2 = u2s + ob;
    u2[idx] = v;
    // the ghost refill folded in: a mirror node writes its face ghost
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const int faces = (g.neumann | g.dirichlet) >> (2 * d);
      const int odd = g.dirichlet >> (2 * d);
      if ((faces & 1) && a[d] == 2)
        u2[idx - 2 * (int64_t)g.s[d]] = (odd & 1) ? -v : v;
      if ((faces & 2) && a[d] == g.n[d] - 3)
        u2[idx + 2 * (int64_t)g.s[d]] = (odd & 2) ? -v : v;
    }
    for (int e = r0; e < r1; ++e)
      if (p.rec.lin[e] == idx)
        p.record[(int64_t)b * p.steps * p.rec.columns +
                 (int64_t)t * p.rec.columns + p.rec.col[e]] = v;
    if (strip_col >= 0) {
      const int64_t at =
          ((int64_t)b * p.strip.rows + (t - p.strip.first)) * p.strip.count +
          strip_col;
      if (p.strip.bf16)
        ((uint16_t *)p.strip.record)[at] = bf16_bits((float)v);
      else
        ((T *)p.strip.record)[at] = v;
    }
  }
}

// --------------------------------------------------------- adjoint step
// level n of the adjoint (see cw_backward_args): lambda^n over the slot of
// lambda^(n + 2), its gradients against u^n, the transposed layer, and the
// reverse march u^(n - 1) over the slot of u^(n + 1)
template <typename T, int NDIM, int R>
__global__ __launch_bounds__(kThreads) void backward_kernel(
    const BwdParams<T> p, const int n) {
  constexpr int REACH = R + 2; // scalar_sensitivity.cu:44 WALL_REACH
  const Geom<T> &g = p.g;
  int a[3], tile;
  bool deep;
  if (!general_node<NDIM>(p.region, g, REACH, a, tile, deep))
    return;
  int64_t idx = 0;
  bool held = false;
  T w = 1; // the cell weight W of the node
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    idx += (int64_t)a[d] * g.s[d];
    held |= held_wall(a[d], g.n[d], g.dirichlet >> (2 * d));
    w *= (a[d] == 1 || a[d] == g.n[d] - 2) ? (T)0.5 : (T)1;
  }
  bool near[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    near[d] = !deep && (a[d] < REACH || a[d] > g.n[d] - 1 - REACH);
  const T mi = p.inv[idx];
  const T sc = p.stiff[idx];
  T sp[NDIM], sm[NDIM];
  if (p.stiff_grad) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      sp[d] = near[d] ? reflected(p.stiff, idx, g.s[d], a[d], g.n[d], 0, 1)
                      : p.stiff[idx + g.s[d]];
      sm[d] = near[d] ? reflected(p.stiff, idx, g.s[d], a[d], g.n[d], 0, -1)
                      : p.stiff[idx - g.s[d]];
    }
  }
  // the layer: slab rows (the memory this node owns) and the reach of the
  // transposed taps, R rows past the front
  int prow[NDIM];
  bool ext[NDIM], layer = false, reach = false;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const int L = p.pml.lo[d], H = p.pml.hi[d], N = g.n[d];
    prow[d] = L + H ? slab_row(a[d], N, L, H) : -1;
    ext[d] = (L && a[d] - 1 < L + R) || (H && N - 2 - a[d] < H + R);
    layer |= prow[d] >= 0;
    reach |= ext[d];
  }
  const bool grad = !(p.skip_layer && layer);
  const int m = p.steps - 1 - n;
  int r0 = 0, r1 = 0, s0 = 0, s1 = 0;
  if (p.rec.starts) {
    r0 = p.rec.starts[tile];
    r1 = p.rec.starts[tile + 1];
  }
  const bool reverse = p.reverse && n <= p.steps - 2;
  if (reverse && p.src.starts) {
    s0 = p.src.starts[tile];
    s1 = p.src.starts[tile + 1];
  }
  int strip_col = -1;
  if (reverse && layer && p.strip.map)
    strip_col = p.strip.map[idx];
  T *const l0s = p.lam.at(n);
  const T *const l1s = p.lam.at(n + 1);
  const T *const u1s = p.u.at(n);
  T *const uos = p.u.at(n + 1);
  const int gi = blockIdx.y;
  const int b0 = gi * p.group;
  const int b1 = min(b0 + p.group, g.nbatch);
  const bool shared = p.gbatch == 0;
  T accm = 0, accs = 0;
  if (grad && shared) {
    accm = p.gmass[(int64_t)gi * g.field + idx];
    if (!p.combine)
      accs = p.gstiff[(int64_t)gi * g.field + idx];
  }

  for (int b = b0; b < b1; ++b) {
    const T *__restrict__ l1 = l1s + (int64_t)b * p.lam.batch;
    T *__restrict__ l0 = l0s + (int64_t)b * p.lam.batch;
    const T *__restrict__ u1 = u1s + (int64_t)b * p.u.batch;
    const T lc = l1[idx], lo_ = l0[idx], uc = u1[idx];
    const T lw = held ? (T)0 : lc;
    T div_l = 0, gsum = 0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const int odd = g.dirichlet >> (2 * d);
      T lt[2 * R + 1];
      T ch[R], cl[R];
      if (near[d]) {
#pragma unroll
        for (int j = -R; j <= R; ++j)
          lt[j + R] = reflected(l1, idx, g.s[d], a[d], g.n[d], odd, j);
        lt[R] = lc;
        T ks[2 * R + 1];
#pragma unroll
        for (int j = -R; j <= R; ++j)
          ks[j + R] = reflected(p.stiff, idx, g.s[d], a[d], g.n[d], 0, j);
        ks[R] = sc;
#pragma unroll
        for (int j = 1; j <= R; ++j) {
          ch[j - 1] = harmonic(ks[R + j - 1], ks[R + j]);
          cl[j - 1] = harmonic(ks[R - j + 1], ks[R - j]);
        }
      } else {
#pragma unroll
        for (int j = -R; j <= R; ++j)
          lt[j + R] = l1[idx + (int64_t)j * g.s[d]];
#pragma unroll
        for (int j = 1; j <= R; ++j) {
          ch[j - 1] = p.cell[d * g.field + idx + (int64_t)(j - 1) * g.s[d]];
          cl[j - 1] = p.cell[d * g.field + idx - (int64_t)j * g.s[d]];
        }
      }
      div_l += g.f[d] * transposed<T, R>(lt, ch, cl, g.op);
      if (p.stiff_grad && grad) {
        T ut[2 * R + 1];
#pragma unroll
        for (int j = -R; j <= R; ++j)
          ut[j + R] = near[d] ? reflected(u1, idx, g.s[d], a[d], g.n[d], odd, j)
                              : u1[idx + (int64_t)j * g.s[d]];
        ut[R] = uc;
        if (near[d])
          lt[R] = lw;
        gsum += stiffness_density<T, R>(ut, lt, near[d] ? lw : lc, sc, sp[d],
                                        sm[d], g.f[d], g.op);
      }
    }
    T ln = -lo_ + (T)2 * lc + mi * div_l;
    if (held)
      ln = 0;
    if (grad && !shared) {
      accm = p.gmass[(int64_t)b * p.gbatch + idx];
      if (!p.combine)
        accs = p.gstiff[(int64_t)b * p.gbatch + idx];
    }
    // the inertia density by parts in time (scalar_sensitivity.cu:512-514)
    T cm = p.combine ? (T)0 : accm;
    cm = cm - p.mf * uc * (ln - (T)2 * lc + lo_);
    T cs = p.combine ? (T)0 : accs;
    cs = cs - p.mf * gsum;
    if (reach) {
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        if (!ext[d])
          continue;
        const int N = g.n[d], L = p.pml.lo[d], H = p.pml.hi[d];
        const int64_t size = slab_size<NDIM>(g, d, L, H);
        const T *eta = p.pml.eta[d][n & 1] + (int64_t)b * kEta * size;
        T sum = 0;
        int at[3];
#pragma unroll
        for (int e = 0; e < NDIM; ++e)
          at[e] = a[e];
#pragma unroll
        for (int k = -R; k <= R; ++k) {
          const int row = slab_row(a[d] + k, N, L, H);
          if (row >= 0) {
            at[d] = a[d] + k;
            const int64_t i = slab_index<NDIM>(g, at, d, row, L, H);
            sum += eta[3 * size + i] * plus_tap<T, R>(-k, g.op) +
                   eta[4 * size + i] * minus_tap<T, R>(-k, g.op);
          }
        }
#pragma unroll
        for (int high = 0; high < 2; ++high) {
          // the reads past the wall that land on a, folded back
          const int ghost = high ? 2 * (N - 2) - a[d] : 2 - a[d];
          if (high ? ghost < N - 1 : ghost > 0)
            continue;
          T fold = 0;
#pragma unroll
          for (int k = 1; k <= R; ++k) {
            const int ai = high ? N - 1 - k : k;
            const int row = slab_row(ai, N, L, H);
            if (row >= 0 && ghost - ai >= -R && ghost - ai <= R) {
              at[d] = ai;
              const int64_t i = slab_index<NDIM>(g, at, d, row, L, H);
              fold += eta[3 * size + i] * plus_tap<T, R>(ghost - ai, g.op) +
                      eta[4 * size + i] * minus_tap<T, R>(ghost - ai, g.op);
            }
          }
          sum += ((g.dirichlet >> (2 * d + high)) & 1) ? -fold : fold;
        }
        const T delta = mi * sum / w;
        ln += delta;
        cm -= p.mf * uc * delta;
      }
    }
    for (int e = r0; e < r1; ++e)
      if (p.rec.lin[e] == idx) {
        const T load =
            p.rec.weight[e] *
            p.adjoint[(int64_t)b * p.rec.shot_stride +
                      (int64_t)m * p.rec.columns + p.rec.col[e]];
        ln += load;
        cm += -p.mf * uc * load;
      }
    l0[idx] = ln;
    // the memory the next level gathers, from this level's final lambda
    if (layer) {
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        if (prow[d] < 0)
          continue;
        const int N = g.n[d], L = p.pml.lo[d], H = p.pml.hi[d], ad = a[d];
        const int64_t size = slab_size<NDIM>(g, d, L, H);
        const int64_t i = slab_index<NDIM>(g, a, d, prow[d], L, H);
        const T *an = p.pml.profile[d], *bn = an + N, *ah = an + 2 * N,
                *bh = an + 3 * N;
        pml_adjoint_update(p.pml.eta[d][n & 1] + (int64_t)b * kEta * size + i,
                           p.pml.eta[d][(n + 1) & 1] +
                               (int64_t)b * kEta * size + i,
                           size, held ? (T)0 : w * ln,
                           p.cell[d * g.field + idx],
                           p.cell[d * g.field + idx - g.s[d]], g.f[d], an[ad],
                           bn[ad], ah[ad], bh[ad], ah[ad - 1], bh[ad - 1]);
      }
    }
    if (grad) {
      if (p.combine) {
        accm = accm + (p.jm * cm + p.js * cs);
      } else {
        accm = cm;
        accs = cs;
      }
      if (!shared) {
        p.gmass[(int64_t)b * p.gbatch + idx] = accm;
        if (!p.combine)
          p.gstiff[(int64_t)b * p.gbatch + idx] = accs;
      }
    }
    if (reverse) {
      // the lossless step run backwards: u^(n - 1) from u^n and u^(n + 1)
      T taps[NDIM][2 * R + 1];
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
#pragma unroll
        for (int j = -R; j <= R; ++j)
          taps[d][j + R] = deep ? u1[idx + (int64_t)j * g.s[d]]
                                : mirrored(u1, idx, g.s[d], a[d], g.n[d],
                                           g.dirichlet >> (2 * d), j);
      T div[NDIM];
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
        div[d] = flux<T, R>(taps[d], p.cell[d * g.field + idx],
                            p.cell[d * g.field + idx - g.s[d]], g.op);
      const T lap = combine_axes<T, NDIM, R>(div, g.f);
      T *__restrict__ uo = uos + (int64_t)b * p.u.batch;
      T v = -uo[idx] + (T)2 * uc + mi * lap;
      if (held)
        v = 0;
      for (int e = s0; e < s1; ++e)
        if (p.src.lin[e] == idx && (p.src.shot[e] < 0 || p.src.shot[e] == b))
          v += p.src.weight[e] *
               p.signal[b * p.src.shot_stride +
                        (int64_t)(n + 1) * p.src.columns + p.src.col[e]];
      if (strip_col >= 0) {
        const int64_t at = ((int64_t)b * p.strip.rows + (n - 1 - p.strip.first)) *
                               p.strip.count +
                           strip_col;
        v = p.strip.bf16
                ? (T)bf16_value(((const uint16_t *)p.strip.record)[at])
                : ((const T *)p.strip.record)[at];
      }
      uo[idx] = v;
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        const int faces = (g.neumann | g.dirichlet) >> (2 * d);
        const int odd = g.dirichlet >> (2 * d);
        if ((faces & 1) && a[d] == 2)
          uo[idx - 2 * (int64_t)g.s[d]] = (odd & 1) ? -v : v;
        if ((faces & 2) && a[d] == g.n[d] - 3)
          uo[idx + 2 * (int64_t)g.s[d]] = (odd & 2) ? -v : v;
      }
    }
  }
  if (grad && shared) {
    p.gmass[(int64_t)gi * g.field + idx] = accm;
    if (!p.combine)
      p.gstiff[(int64_t)gi * g.field + idx] = accs;
  }
}

// --------------------------------------------------- superposition step
// scalar_sensitivity.cu:599-668 (step = 1) and :352-385 (step = 0)
template <typename T, int NDIM, int R>
__global__ __launch_bounds__(kThreads) void superposed_kernel(
    Geom<T> g, const T *__restrict__ inv, const T *__restrict__ cell,
    const T *__restrict__ u0, const T *__restrict__ u1, T *__restrict__ u2,
    T *__restrict__ acc_mass, T *__restrict__ acc_stiff, T ft, T fs,
    int step) {
  const int tile = xcd_tile(blockIdx.x, g.ntiles);
  int a[3], lo[3], hi[3];
  tile_node<NDIM>(g, tile, a, lo, hi);
  bool inside = true;
#pragma unroll
  for (int d = 0; d < NDIM; ++d)
    inside &= a[d] > 0 && a[d] < g.n[d] - 1;
  if (!inside)
    return;
  int64_t idx = 0;
  bool held = false;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    idx += (int64_t)a[d] * g.s[d];
    held |= held_wall(a[d], g.n[d], g.dirichlet >> (2 * d));
  }
  T sum = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const T gd = u1[idx + g.s[d]] - u1[idx - g.s[d]];
    if (d == 0)
      sum = g.f[0] * gd * gd;
    else
      sum += g.f[d] * gd * gd;
  }
  if (!step) {
    const T dudt = u2[idx] - u0[idx];
    acc_mass[idx] += ft * dudt * dudt;
    acc_stiff[idx] += sum;
    return;
  }
  T div[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    T t[2 * R + 1];
#pragma unroll
    for (int j = -R; j <= R; ++j)
      t[j + R] = mirrored(u1, idx, g.s[d], a[d], g.n[d],
                          g.dirichlet >> (2 * d), j);
    div[d] = flux<T, R>(t, cell[d * g.field + idx],
                        cell[d * g.field + idx - g.s[d]], g.op);
  }
  const T lap = combine_axes<T, NDIM, R>(div, g.f);
  const T uc = u1[idx];
  const T dudt = uc - u2[idx];
  acc_mass[idx] += ft * dudt * dudt;
  acc_stiff[idx] += fs * sum;
  u2[idx] = held ? (T)0 : -u0[idx] + (T)2 * uc + inv[idx] * lap;
}

} // namespace cw
