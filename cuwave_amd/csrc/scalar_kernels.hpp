// The pressure-family kernels for gfx950, general form: any dimension, walls,
// PML, points. On a 2D grid they cover only the band of nodes within reach of
// a wall (the streaming kernels of stream.hpp take everything else); in 1D
// and 3D they cover every node.
//
// A thread steps one shot of one node: a block's 256 threads are
// 256 / lanes nodes times their shot lanes, so the lanes of a node are
// neighbouring threads of one wave (see common.hpp for the layout).
//
// Every kernel rounds as the reference's kernels do on its IEEE CPU path: no
// contraction (built with -ffp-contract=off), an fma exactly where the
// reference pins one (cuwave/kernels/scalar.cu:14-33, 416-436), correctly
// rounded division. The harmonic cell means sc sp / (sc + sp)
// (cuwave/kernels/scalar.cu:63) are computed once by `prepare_kernel`, in
// the same operand order, so reading them is bitwise the in-place division;
// the inverse inertia 1 / sc likewise.
#pragma once

#include "common.hpp"

namespace cw {

template <typename T> struct Points {
  const int32_t *starts; // (ntiles + 1) CSR offsets, or null
  const int32_t *lin;
  const int32_t *shot; // -1 for every shot
  const int32_t *col;
  const T *weight;
  int64_t rs, cs, ss; // row, column and shot strides of the signal / record
  __device__ __forceinline__ int64_t at(int row, int e, int b) const {
    return (int64_t)row * rs + (int64_t)col[e] * cs + (int64_t)b * ss;
  }
  __device__ __forceinline__ bool takes(int e, int64_t idx, int b) const {
    return lin[e] == idx && (shot[e] < 0 || shot[e] == b);
  }
};

template <typename T> struct Levels {
  T *base;
  int64_t slot;
  int slots, offset;
  __device__ __forceinline__ uint32_t slot_of(int level) const {
    int k = (level + offset) % slots;
    return (uint32_t)(k + (k < 0 ? slots : 0));
  }
  __device__ __forceinline__ T *at(int level) const {
    return base + (int64_t)slot_of(level) * slot;
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

// where the general kernels run: every interior node, or the band of nodes
// outside a core box: (2D) rows [y0, y1) x columns [x0, x1), (3D) planes
// [z0, z1) x rows [y0, y1) x columns [x0, x1)
struct Region {
  int band;
  int y0, y1, x0, x1;
  int count; // band nodes
  int q;     // (2D) the rows and columns by each wall that its folds reach, R + 1
  int z0, z1;
};

template <typename T> struct FwdParams {
  Geom<T> g;
  Region region;
  const T *inv, *cell;
  Layer<T> pml;
  Levels<T> u;
  Points<T> src;
  const T *signal;
  Points<T> rec;
  T *record;
  Strip<T> strip;
  T *illum; // per node: sum over steps and lanes of u^2, or null
  // a step index offset in device memory (a launch's step t is t + *clock),
  // or null; the steps from t_end on are skipped
  const int *clock;
  int t_end;
  const T *damp; // nodal damping d, or null (lossless)
  T dt;
  // a domain mask: per node the packed cell radii (read by forward_masked
  // only), the stiffness a held cell reads, the cell weights per radius
  const int32_t *cells;
  const T *stiff;
  T rows[4][4];
};

template <typename T> struct BwdParams {
  Geom<T> g;
  Region region;
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
  int combine, stiff_grad, skip_layer;
  T jm, js, mf;
  int steps;
  const T *damp, *rdamp; // the damping, and the reverse march's (or null)
  T dt;
};

// the damped update (cuwave/kernels/scalar.cu:455-459): the step
// (2 u1 - u0 (1 - beta) + minv L u1) / (1 + beta), beta = minv d dt / 2,
// rounded as the reference's expression is
template <typename T, typename Q>
__device__ __forceinline__ Q damped(Q uc, Q u0, Q mlap, T mi, T d, T dt) {
  const T beta = (T)0.5 * mi * d * dt;
  return ((T)2 * uc - ((T)1 - beta) * u0 + mlap) / ((T)1 + beta);
}

// the cell code of a node whose cells all keep the full radius, and the
// radius code of a cell open onto a held node (cuwave/wave.py:661-662)
constexpr int kDeepCells = 1 << 30;
constexpr int kHeldCell = 7;

// ------------------------------------------------------------------ helpers
// the cell flux divergence of one axis over taps t[0 .. 2R] centred on R,
// rounded as cuwave/kernels/scalar.cu:14-33 pins it
template <typename T, int R, typename Q = T>
__device__ __forceinline__ Q flux(const Q *t, T gp, T gm, const T *op) {
  const Q uc = t[R];
  Q Dp = t[R + 1] - uc, Dm = uc - t[R - 1];
  if (R > 1) {
    Dp = op[0] * Dp;
    Dm = op[0] * Dm;
  }
#pragma unroll
  for (int k = 2; k <= R; ++k) {
    Dp = fmat(op[k - 1], t[R + k] - t[R + 1 - k], Dp);
    Dm = fmat(op[k - 1], t[R + k - 1] - t[R - k], Dm);
  }
  const Q mm = Dm * gm;
  return fmat(Dp, gp, -mm);
}

// the axes combined as cuwave/kernels/scalar.cu:416-436 rounds them
template <typename T, int NDIM, int R, typename Q = T>
__device__ __forceinline__ Q combine_axes(const Q *d, const T *f) {
  if (NDIM == 1)
    return f[0] * d[0];
  if (NDIM == 2) {
    if (R == 1) {
      const Q p = f[1] * d[1];
      return fmat(f[0], d[0], p);
    }
    const Q p = f[0] * d[0];
    return fmat(f[1], d[1], p);
  }
  const Q p = f[1] * d[1];
  return fmat(f[2], d[2], fmat(f[0], d[0], p));
}

// the two cell differences with plain sums (cuwave/kernels/scalar.cu:68-82,
// scalar_sensitivity.cu:116-135), over taps centred on R
template <typename T, int R, typename Q = T>
__device__ __forceinline__ void cell_diffs(const Q *t, const T *op, Q &Dp,
                                           Q &Dm) {
  const Q uc = t[R];
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
// 499-532): the three memories, `size` elements apart, updated in place,
// their share of the field returned. `an .. bh1` are the profile at the node
// and at the half point below it.
// `m` holds the three memories as read (so a caller may load them early)
template <typename T, typename Q = T>
__device__ __forceinline__ Q pml_forward_from(const Q *m, T *aux,
                                              int64_t size, Q Dp, Q Dm, T gp,
                                              T gm, T f, T mi, T an, T bn,
                                              T ah, T bh, T ah1, T bh1) {
  const Q pp = bh * m[0] + ah * Dp;
  const Q pm = bh1 * m[1] + ah1 * Dm;
  const Q z = bn * m[2] + an * f * (gp * (Dp + pp) - gm * (Dm + pm));
  lanes_put(aux, 0, pp);
  lanes_put(aux, size, pm);
  lanes_put(aux, 2 * size, z);
  return mi * (f * (gp * pp - gm * pm) + z);
}
template <typename T, typename Q = T>
__device__ __forceinline__ Q pml_forward(T *aux, int64_t size, Q Dp, Q Dm,
                                         T gp, T gm, T f, T mi, T an, T bn,
                                         T ah, T bh, T ah1, T bh1) {
  const Q m[3] = {lanes_at<Q>(aux, 0), lanes_at<Q>(aux, size),
                  lanes_at<Q>(aux, 2 * size)};
  return pml_forward_from(m, aux, size, Dp, Dm, gp, gm, f, mi, an, bn, ah, bh,
                          ah1, bh1);
}

// the adjoint memory of a slab node stepped by lambda (scalar_sensitivity.cu:
// 517-542 with q = W lambda), then the two tap fields of the new memory
// (:232-251): cur and nxt address the node's kEta rows, `size` apart
template <typename T, typename Q = T>
__device__ __forceinline__ void pml_adjoint_from(const Q *c, T *nxt,
                                                 int64_t size, Q q, T gp, T gm,
                                                 T f, T an, T bn, T ah, T bh,
                                                 T ah1, T bh1) {
  const Q ez = c[2];
  const Q ep = bh * (c[0] + an * f * gp * ez) + f * gp * q;
  const Q em = bh1 * (c[1] - an * f * gm * ez) - f * gm * q;
  const Q zz = bn * ez + q;
  lanes_put(nxt, 0, ep);
  lanes_put(nxt, size, em);
  lanes_put(nxt, 2 * size, zz);
  const Q e2 = an * f * zz;
  lanes_put(nxt, 3 * size, ah * ep + gp * (1 + ah) * e2);
  lanes_put(nxt, 4 * size, ah1 * em - gm * (1 + ah1) * e2);
}
// the same, `cur` the memory to read (`size` apart)
template <typename T, typename Q = T>
__device__ __forceinline__ void pml_adjoint_update(const T *cur, T *nxt,
                                                   int64_t size, Q q, T gp,
                                                   T gm, T f, T an, T bn,
                                                   T ah, T bh, T ah1, T bh1) {
  const Q c[3] = {lanes_at<Q>(cur, 0), lanes_at<Q>(cur, size),
                  lanes_at<Q>(cur, 2 * size)};
  pml_adjoint_from(c, nxt, size, q, gp, gm, f, an, bn, ah, bh, ah1, bh1);
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

// the transposed layer of one axis at a node: the tap fields of the slab
// nodes within R along the axis, `at(k)` the element of the node k away or
// -1 off the slab, `size` the elements between memory rows
// Every tap's two loads are issued whether or not its node is on the slab
// (off it from element 0, always in range) and the term is then kept or
// dropped: no branch splits the loads, so they are all in flight together.
template <typename T, int R, typename Q = T, typename At>
__device__ __forceinline__ Q gather(const T *__restrict__ eta, int64_t size,
                                    const T *op, At at) {
  bool on[2 * R + 1];
  Q e3[2 * R + 1], e4[2 * R + 1];
#pragma unroll
  for (int k = -R; k <= R; ++k) {
    const int64_t i = at(k);
    on[k + R] = i >= 0;
    const int64_t j = on[k + R] ? i : 0;
    e3[k + R] = lanes_at<Q>(eta, 3 * size + j);
    e4[k + R] = lanes_at<Q>(eta, 4 * size + j);
  }
#pragma unroll
  for (int k = 0; k <= 2 * R; ++k) {
    keep(e3[k]);
    keep(e4[k]);
  }
  Q sum = (T)0;
#pragma unroll
  for (int k = -R; k <= R; ++k) {
    const Q term = e3[k + R] * plus_tap<T, R>(-k, op) +
                   e4[k + R] * minus_tap<T, R>(-k, op);
    sum = on[k + R] ? sum + term : sum;
  }
  return sum;
}

// scalar_sensitivity.cu:137-154 transposed_divergence_taps over taps
// l[0 .. 2R] and cell means: hi[j - 1] the cell (a + j - 1, a + j),
// lo[j - 1] the cell (a - j, a - j + 1)
template <typename T, int R, typename Q = T>
__device__ __forceinline__ Q transposed(const Q *l, const T *chi,
                                        const T *clo, const T *op) {
  Q div = (T)0, lh = l[R], ll = l[R];
#pragma unroll
  for (int j = 1; j <= R; ++j) {
    const Q lh1 = l[R + j], ll1 = l[R - j];
    div += op[j - 1] * (chi[j - 1] * (lh1 - lh) - clo[j - 1] * (ll - ll1));
    lh = lh1;
    ll = ll1;
  }
  return div;
}

// scalar_sensitivity.cu:116-135 stiffness_gradient_taps
template <typename T, int R, typename Q = T>
__device__ __forceinline__ Q stiffness_density(const Q *u, const Q *l, Q lc,
                                               T sc, T sp, T sm, T factor,
                                               const T *op) {
  const T dgp = sp * sp / ((sc + sp) * (sc + sp));
  const T dgm = sm * sm / ((sc + sm) * (sc + sm));
  Q Dp, Dm;
  cell_diffs<T, R>(u, op, Dp, Dm);
  return factor * (dgp * Dp * (l[R + 1] - lc) + dgm * Dm * (lc - l[R - 1]));
}

// the gradient of one level added into the per-node accumulators: with one
// lane the reference's chain (the densities subtracted from the accumulator
// itself, which `cm` and `cs` then hold), with several the shots' densities
// summed across the lanes first; every thread of the node takes part, the
// one holding lane 0 (`first`) stores
template <typename T, typename Q>
__device__ __forceinline__ void add_gradient(const BwdParams<T> &p,
                                             int64_t idx, bool first, T accm,
                                             T accs, Q cm, Q cs) {
  const int lanes = p.g.lanes;
  if (p.combine) {
    const T x = shot_sum(p.jm * cm + p.js * cs, lanes);
    if (first)
      p.gmass[idx] = accm + x;
  } else if (lanes == 1) {
    p.gmass[idx] = lane(cm, 0);
    p.gstiff[idx] = lane(cs, 0);
  } else {
    const T xm = shot_sum(cm, lanes), xs = shot_sum(cs, lanes);
    if (first) {
      p.gmass[idx] = accm + xm;
      p.gstiff[idx] = accs + xs;
    }
  }
}

// the flux divergence of a node at a domain wall (cuwave/kernels/scalar.cu:
// 85-158): along each axis the cell to the next node at radius rp and the
// cell to the previous one at radius rm, read from `code`; a closed cell
// (radius 0) carries no flux, a cell open onto a held node (kHeldCell) the
// difference to the zero halfway across at the node's own stiffness sc.
// A cell of radius r reads the radius-r weights and no node past r, so no
// node outside the stepped set. Plain sums, the axes added in order, as the
// reference rounds its graded cells; gp, gm are the harmonic cell means.
template <typename T, int NDIM, int R, typename Q>
__device__ __forceinline__ Q graded(const T *__restrict__ u1, int64_t e,
                                    int sh, const int *s, int code, Q uc,
                                    T sc, const T *gp, const T *gm,
                                    const T *f, const T (*rows)[4]) {
  Q sum = (T)0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const int64_t se = (int64_t)s[d] << sh;
    const int rp = (code >> (6 * d)) & 7, rm = (code >> (6 * d + 3)) & 7;
    Q Dp = (T)0, Dm = (T)0;
    if (rp == kHeldCell) {
      Dp = -uc * sc;
    } else if (rp) {
      const T *w = rows[rp - 1];
      Dp = w[0] * (lanes_at<Q>(u1, e + se) - uc);
#pragma unroll
      for (int k = 2; k <= R; ++k)
        if (k <= rp)
          Dp = Dp + w[k - 1] * (lanes_at<Q>(u1, e + k * se) -
                                lanes_at<Q>(u1, e - (k - 1) * se));
      Dp = Dp * gp[d];
    }
    if (rm == kHeldCell) {
      Dm = uc * sc;
    } else if (rm) {
      const T *w = rows[rm - 1];
      Dm = w[0] * (uc - lanes_at<Q>(u1, e - se));
#pragma unroll
      for (int k = 2; k <= R; ++k)
        if (k <= rm)
          Dm = Dm + w[k - 1] * (lanes_at<Q>(u1, e + (k - 1) * se) -
                                lanes_at<Q>(u1, e - k * se));
      Dm = Dm * gm[d];
    }
    const Q term = f[d] * (Dp - Dm);
    sum = d == 0 ? term : sum + term;
  }
  return sum;
}

// ------------------------------------------------------------ node mapping
// the 2D band node t, or false past the band. The nodes within q of a wall
// come first (the rows by the two walls, then the columns by the side walls
// in the rows between): their mirrored taps, folds and, in the layer
// corners, two memories make them the band's slowest nodes, and dispatched
// first they run under the rest instead of trailing it (measured 1.9 % off a
// 32-shot 1090 x 1090 gradient against the band in row order). Then the rest
// of the rows above and below the box, then the side columns beside it.
template <typename T>
__device__ __forceinline__ bool band_node(const Region &r, const Geom<T> &g,
                                          int t, int *a) {
  const int N0 = g.n[0], N1 = g.n[1], W = N1 - 2, q = r.q;
  const int rows = 2 * q * W;
  if (t < rows) {
    const int k = t / W;
    a[0] = k < q ? 1 + k : N0 - 1 - 2 * q + k;
    a[1] = 1 + t % W;
    return true;
  }
  t -= rows;
  if (t < (N0 - 2 - 2 * q) * 2 * q) {
    const int k = t % (2 * q);
    a[0] = q + 1 + t / (2 * q);
    a[1] = k < q ? 1 + k : N1 - 1 - 2 * q + k;
    return true;
  }
  t -= (N0 - 2 - 2 * q) * 2 * q;
  const int Wi = W - 2 * q; // the columns off the side walls
  const int top = (r.y0 - 1 - q) * Wi;
  if (t < top) {
    a[0] = q + 1 + t / Wi;
    a[1] = q + 1 + t % Wi;
    return true;
  }
  t -= top;
  const int bot = (N0 - 1 - q - r.y1) * Wi;
  if (t < bot) {
    a[0] = r.y1 + t / Wi;
    a[1] = q + 1 + t % Wi;
    return true;
  }
  t -= bot;
  const int left = r.x0 - 1 - q, right = N1 - 1 - q - r.x1, per = left + right;
  if (per <= 0 || t >= per * (r.y1 - r.y0))
    return false;
  const int k = t % per;
  a[0] = r.y0 + t / per;
  a[1] = k < left ? q + 1 + k : r.x1 + (k - left);
  return true;
}

// the 3D shell node t, or false past the shell: the whole planes before and
// after the box's, then in the box's planes the whole rows above and below
// its rows, then in its rows the columns either side of it
template <typename T>
__device__ __forceinline__ bool shell_node(const Region &r, const Geom<T> &g,
                                           int t, int *a) {
  const int H = g.n[1] - 2, W = g.n[2] - 2;
  const int before = r.z0 - 1, planes = before + (g.n[0] - 1 - r.z1);
  if (t < planes * H * W) {
    const int k = t / (H * W), rest = t % (H * W);
    a[0] = k < before ? 1 + k : r.z1 + (k - before);
    a[1] = 1 + rest / W;
    a[2] = 1 + rest % W;
    return true;
  }
  t -= planes * H * W;
  const int depth = r.z1 - r.z0;
  const int above = r.y0 - 1, rows = above + (g.n[1] - 1 - r.y1);
  if (t < depth * rows * W) {
    const int rest = t % (rows * W), k = rest / W;
    a[0] = r.z0 + t / (rows * W);
    a[1] = k < above ? 1 + k : r.y1 + (k - above);
    a[2] = 1 + rest % W;
    return true;
  }
  t -= depth * rows * W;
  const int left = r.x0 - 1, cols = left + (g.n[2] - 1 - r.x1);
  const int span = r.y1 - r.y0;
  if (cols <= 0 || t >= depth * span * cols)
    return false;
  const int rest = t % (span * cols), k = rest % cols;
  a[0] = r.z0 + t / (span * cols);
  a[1] = r.y0 + rest / cols;
  a[2] = k < left ? 1 + k : r.x1 + (k - left);
  return true;
}

// this thread's node, every thread of a node (2^ts of them) the same: the
// band's, or the interior's in row-major order; block `bid` of `nb`.
// The 2D band's parts cost very different amounts a node (the wall rows and
// the slab corners several times the slab rows), so its blocks are cut into
// chunks of kBandChunk dealt to the XCDs in turn: every XCD gets a share of
// each part (a contiguous run each, as xcd_tile deals the rest). Measured
// 2 % off a 1090 x 1090 adjoint level against contiguous runs.
template <int NDIM, typename T>
__device__ __forceinline__ bool general_node(const Region &r, const Geom<T> &g,
                                             int *a, int bid, int nb, int ts) {
  const int per = kThreads >> ts;
  if (NDIM == 3 && r.band) {
    const int t = xcd_tile(bid, nb) * per + (int)(threadIdx.x >> ts);
    return t < r.count && shell_node(r, g, t, a);
  }
  if (NDIM == 2 && r.band) {
    // the hardware deals block bid to XCD bid % kXcds as its (bid / kXcds)-th
    const int x = bid % kXcds, k = bid / kXcds;
    const int chunk = (k / kBandChunk) * kXcds + x;
    const int tb = (chunk * kBandChunk + k % kBandChunk) * per +
                   (int)(threadIdx.x >> ts);
    return tb < r.count && band_node(r, g, tb, a);
  }
  const int t = xcd_tile(bid, nb) * per + (int)(threadIdx.x >> ts);
  if (t >= g.interior)
    return false;
  if (NDIM == 1) {
    a[0] = 1 + t;
  } else if (NDIM == 2) {
    const int W = g.n[1] - 2;
    a[0] = 1 + t / W;
    a[1] = 1 + t % W;
  } else {
    // the rows cut into kXcds stripes, a stripe's nodes in (plane, row,
    // column) order: the run of items an XCD takes is one stripe marched
    // plane by plane, so the planes the slow-axis taps read (R either side)
    // are the stripe's share of them, which its L2 holds (whole planes it
    // would not)
    const int W = g.n[2] - 2, H = g.n[1] - 2;
    const int hs = (H + kXcds - 1) / kXcds;
    const int64_t full = (int64_t)(g.n[0] - 2) * hs * W;
    const int k = (int)(t / full);
    const int rem = (int)(t - k * full);
    const int rows = min(hs, H - k * hs);
    a[0] = 1 + rem / (rows * W);
    a[1] = 1 + k * hs + (rem % (rows * W)) / W;
    a[2] = 1 + rem % W;
  }
  return true;
}

// ----------------------------------------------------------- preparation
// inv = minv or 1 / stiff; cell[d] = the harmonic mean of the cell (a, a + 1)
// along axis d, for every node that has a next node along d
template <typename T, int NDIM>
__global__ __launch_bounds__(kThreads) void prepare_kernel(
    Geom<T> g, const T *__restrict__ stiff, const T *__restrict__ minv,
    T *__restrict__ inv, T *__restrict__ cell) {
  const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= g.nodes)
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
    cell[d * g.nodes + idx] =
        a[d] + 1 < g.np[d] ? harmonic(sc, stiff[idx + g.s[d]]) : (T)0;
}

// ------------------------------------------------------------ forward step
// step t at the nodes of general block `bid` of `nb`, V shot lanes a thread;
// MASK reads the domain mask's cells: a node outside the stepped set is left
// as it is, a node at a domain wall takes the graded flux divergence
template <typename T, int NDIM, int R, int V, bool MASK = false>
__device__ __forceinline__ void forward_general(const FwdParams<T> &p,
                                                const int t, int bid, int nb) {
  using Q = LanesOf<T, V>;
  const Geom<T> &g = p.g;
  const int ts = g.lshift - log2_of(V);
  int a[3];
  if (!general_node<NDIM>(p.region, g, a, bid, nb, ts))
    return;
  const int sh = g.lshift;
  const int b = (int)(threadIdx.x & ((1u << ts) - 1)) * V; // first lane
  // node and element indices in 32 bits for the taps (the grid check
  // refuses a level of 2^31 elements or more)
  int node = 0;
  bool held = false, deep = true;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    node += a[d] * g.s[d];
    held |= held_wall(a[d], g.n[d], g.dirichlet >> (2 * d));
    deep &= a[d] >= R && a[d] <= g.n[d] - 1 - R;
  }
  const int64_t idx = node;
  int code = kDeepCells;
  if constexpr (MASK) {
    code = p.cells[idx];
    if (!code)
      return; // outside the domain, or held: never stepped
  }
  const int elem = (node << sh) + b;
  const int64_t e = elem;
  const T mi = p.inv[idx];
  T gp[NDIM], gm[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    gp[d] = p.cell[d * g.nodes + idx];
    gm[d] = p.cell[d * g.nodes + idx - g.s[d]];
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
  const int tile = point_tile<NDIM>(g, a);
  // the point-list bounds and the strip column first, in flight with the taps
  int r0 = 0, r1 = 0, s0 = 0, s1 = 0, col = -1;
  if (p.src.starts) {
    s0 = p.src.starts[tile];
    s1 = p.src.starts[tile + 1];
  }
  if (p.rec.starts) {
    r0 = p.rec.starts[tile];
    r1 = p.rec.starts[tile + 1];
  }
  if ((layer || p.damp) && p.strip.map) // a damped node may be a strip node
    col = p.strip.map[idx];
  const T *__restrict__ u1 = p.u.at(t - 1);
  Q taps[NDIM][2 * R + 1];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const int se = g.s[d] << sh;
    // mirrored only along an axis near a wall (the same bits as along every
    // axis), there by moving the one load of a tap and setting its sign
    const bool wall = !(deep || (a[d] >= R && a[d] <= g.n[d] - 1 - R));
#pragma unroll
    for (int j = -R; j <= R; ++j) {
      bool neg = false;
      const int k =
          wall ? mirror_step(a[d], g.n[d], g.dirichlet >> (2 * d), j, neg) : j;
      const Q v = lanes_at<Q>(u1, elem + k * se);
      taps[d][j + R] = neg ? -v : v;
    }
  }
  const Q uc = taps[0][R];
  Q lap;
  if (MASK && !(code & kDeepCells)) {
    lap = graded<T, NDIM, R, Q>(u1, e, sh, g.s, code, uc, p.stiff[idx], gp,
                                gm, g.f, p.rows);
  } else {
    Q div[NDIM];
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
      div[d] = flux<T, R>(taps[d], gp[d], gm[d], g.op);
    lap = combine_axes<T, NDIM, R>(div, g.f);
  }
  Q v = p.damp ? damped<T, Q>(uc, lanes_at<Q>(p.u.at(t - 2), e), mi * lap, mi,
                              p.damp[idx], p.dt)
               : -lanes_at<Q>(p.u.at(t - 2), e) + (T)2 * uc + mi * lap;
  if (held)
    v = (T)0;
  if (layer) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      if (prow[d] < 0)
        continue;
      const int N = g.n[d], L = p.pml.lo[d], H = p.pml.hi[d], ad = a[d];
      const T *an = p.pml.profile[d], *bn = an + N, *ah = an + 2 * N,
              *bh = an + 3 * N;
      const int64_t size = slab_size<NDIM>(g, d, L, H) << sh;
      T *aux = p.pml.aux[d] +
               (slab_index<NDIM>(g, a, d, prow[d], L, H) << sh) + b;
      Q Dp, Dm;
      cell_diffs<T, R>(taps[d], g.op, Dp, Dm);
      const Q share =
          pml_forward(aux, size, Dp, Dm, gp[d], gm[d], g.f[d], mi, an[ad],
                      bn[ad], ah[ad], bh[ad], ah[ad - 1], bh[ad - 1]);
      if (!held)
        v += share;
    }
  }
  for (int k = s0; k < s1; ++k)
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (p.src.takes(k, idx, b + i))
          lane(v, i) += p.src.weight[k] * p.signal[p.src.at(t, k, b + i)];
  T *__restrict__ u2 = p.u.at(t);
  lanes_put(u2, e, v);
  if (p.illum) {
    // every thread of the node takes part in the sum; one writer a node
    const T s2 = shot_sum(v * v, g.lanes);
    if (b == 0)
      atomicAdd(p.illum + idx, s2);
  }
  // the ghost refill folded in: a mirror node writes its face ghost
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const int faces = (g.neumann | g.dirichlet) >> (2 * d);
    const int odd = g.dirichlet >> (2 * d);
    const int64_t se = (int64_t)g.s[d] << sh;
    if ((faces & 1) && a[d] == 2)
      lanes_put(u2, e - 2 * se, (odd & 1) ? -v : v);
    if ((faces & 2) && a[d] == g.n[d] - 3)
      lanes_put(u2, e + 2 * se, (odd & 2) ? -v : v);
  }
  for (int k = r0; k < r1; ++k)
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (p.rec.takes(k, idx, b + i))
          p.record[p.rec.at(t, k, b + i)] = lane(v, i);
  {
    if (col >= 0) {
      const int64_t at =
          (((int64_t)(t - p.strip.first) * p.strip.count + col) << sh) + b;
      if (p.strip.bf16) {
#pragma unroll
        for (int i = 0; i < V; ++i)
          ((uint16_t *)p.strip.record)[at + i] = bf16_bits((float)lane(v, i));
      } else {
        lanes_put((T *)p.strip.record, at, v);
      }
    }
  }
}

// --------------------------------------------------------- adjoint step
// level n of the adjoint (see cw_backward_args) at the nodes of general
// block `bid` of `nb`: lambda^n over the slot of lambda^(n + 2), its
// gradients against u^n, the transposed layer, and the reverse march
// u^(n - 1) over the slot of u^(n + 1)
template <typename T, int NDIM, int R, int V>
__device__ __forceinline__ void backward_general(const BwdParams<T> &p,
                                                 const int n, int bid,
                                                 int nb) {
  using Q = LanesOf<T, V>;
  constexpr int REACH = R + 2; // scalar_sensitivity.cu:44 WALL_REACH
  const Geom<T> &g = p.g;
  const int ts = g.lshift - log2_of(V);
  int a[3];
  if (!general_node<NDIM>(p.region, g, a, bid, nb, ts))
    return;
  const int sh = g.lshift;
  const int b = (int)(threadIdx.x & ((1u << ts) - 1)) * V; // first lane
  int64_t idx = 0;
  bool held = false, any_near = false;
  T w = 1; // the cell weight W of the node
  bool near[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    idx += (int64_t)a[d] * g.s[d];
    held |= held_wall(a[d], g.n[d], g.dirichlet >> (2 * d));
    w *= (a[d] == 1 || a[d] == g.n[d] - 2) ? (T)0.5 : (T)1;
    near[d] = a[d] < REACH || a[d] > g.n[d] - 1 - REACH;
    any_near |= near[d];
  }
  const int64_t e = (idx << sh) + b;
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
  const int tile = point_tile<NDIM>(g, a);
  T *__restrict__ l0 = p.lam.at(n);
  const T *__restrict__ l1 = p.lam.at(n + 1);
  const T *__restrict__ u1 = p.u.at(n);
  T *__restrict__ uo = p.u.at(n + 1);
  const bool chain = g.lanes == 1 && !p.combine;

  // The loads that need only the node come first, so that they are all in
  // flight together rather than one branch after another.
  int strip_col = -1;
  if (p.reverse && (layer || p.damp) && p.strip.map)
    strip_col = p.strip.map[idx];
  int r0 = 0, r1 = 0, s0 = 0, s1 = 0;
  if (p.rec.starts) {
    r0 = p.rec.starts[tile];
    r1 = p.rec.starts[tile + 1];
  }
  if (p.reverse && p.src.starts) {
    s0 = p.src.starts[tile];
    s1 = p.src.starts[tile + 1];
  }
  T accm = 0, accs = 0;
  if (grad) {
    accm = p.gmass[idx];
    if (!p.combine)
      accs = p.gstiff[idx];
  }
  const T mi = p.inv[idx];
  const T sc = p.stiff[idx];
  // the layer memory this node owns, as the previous level left it
  Q cur[NDIM][3];
  int64_t own[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    own[d] = 0;
    if (prow[d] >= 0) {
      const int L = p.pml.lo[d], H = p.pml.hi[d];
      const int64_t size = slab_size<NDIM>(g, d, L, H) << sh;
      own[d] = (slab_index<NDIM>(g, a, d, prow[d], L, H) << sh) + b;
      const T *c = p.pml.eta[d][n & 1] + own[d];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        cur[d][k] = lanes_at<Q>(c, k * size);
    }
  }
  // the reverse march rebuilds u^(n - 1) where something reads it: with the
  // layer's gradient skipped, a layer node off the strip is out of reach of
  // every reversible node and of every strip node (the strip holds all the
  // layer nodes within reach of a reversible one), so its value is never read
  const bool reverse = p.reverse && n <= p.steps - 2 &&
                       !(p.skip_layer && layer && p.strip.map && strip_col < 0);
  Q replay = (T)0;
  if (reverse && strip_col >= 0) {
    // a layer node is replayed from the forward sweep's record
    const int64_t at =
        (((int64_t)(n - 1 - p.strip.first) * p.strip.count + strip_col) << sh) +
        b;
    if (p.strip.bf16) {
#pragma unroll
      for (int i = 0; i < V; ++i)
        lane(replay, i) =
            (T)bf16_value(((const uint16_t *)p.strip.record)[at + i]);
    } else {
      replay = lanes_at<Q>((const T *)p.strip.record, at);
    }
  }
  T sp[NDIM], sm[NDIM];
  if (p.stiff_grad) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      sp[d] = near[d] ? reflected<T>(p.stiff, idx, g.s[d], a[d], g.n[d], 0, 1)
                      : p.stiff[idx + g.s[d]];
      sm[d] = near[d]
                  ? reflected<T>(p.stiff, idx, g.s[d], a[d], g.n[d], 0, -1)
                  : p.stiff[idx - g.s[d]];
    }
  }

  const Q lc = lanes_at<Q>(l1, e), lo_ = lanes_at<Q>(l0, e);
  const Q uc = lanes_at<Q>(u1, e);
  const Q lw = held ? Q((T)0) : lc;
  Q div_l = (T)0, gsum = (T)0;
  if (!any_near) {
    // off the walls: every axis' taps and cell means in one run of loads
    Q lt[NDIM][2 * R + 1];
    T ch[NDIM][R], cl[NDIM][R];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const int64_t se = (int64_t)g.s[d] << sh;
#pragma unroll
      for (int j = -R; j <= R; ++j)
        if (j)
          lt[d][j + R] = lanes_at<Q>(l1, e + j * se);
      lt[d][R] = lc;
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        ch[d][j - 1] = p.cell[d * g.nodes + idx + (int64_t)(j - 1) * g.s[d]];
        cl[d][j - 1] = p.cell[d * g.nodes + idx - (int64_t)j * g.s[d]];
      }
    }
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
      div_l += g.f[d] * transposed<T, R>(lt[d], ch[d], cl[d], g.op);
    if (p.stiff_grad && grad) {
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        const int64_t se = (int64_t)g.s[d] << sh;
        Q ut[2 * R + 1];
#pragma unroll
        for (int j = -R; j <= R; ++j)
          ut[j + R] = j ? lanes_at<Q>(u1, e + j * se) : uc;
        gsum += stiffness_density<T, R>(ut, lt[d], lc, sc, sp[d], sm[d],
                                        g.f[d], g.op);
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const int odd = g.dirichlet >> (2 * d);
      const int64_t se = (int64_t)g.s[d] << sh;
      Q lt[2 * R + 1];
      T ch[R], cl[R];
      if (near[d]) {
        reflected_line<R>(l1, e, se, a[d], g.n[d], odd, lt);
        lt[R] = lc;
        T ks[2 * R + 1];
        reflected_line<R>(p.stiff, idx, (int64_t)g.s[d], a[d], g.n[d], 0, ks);
        ks[R] = sc;
#pragma unroll
        for (int j = 1; j <= R; ++j) {
          ch[j - 1] = harmonic(ks[R + j - 1], ks[R + j]);
          cl[j - 1] = harmonic(ks[R - j + 1], ks[R - j]);
        }
      } else {
#pragma unroll
        for (int j = -R; j <= R; ++j)
          lt[j + R] = lanes_at<Q>(l1, e + j * se);
#pragma unroll
        for (int j = 1; j <= R; ++j) {
          ch[j - 1] = p.cell[d * g.nodes + idx + (int64_t)(j - 1) * g.s[d]];
          cl[j - 1] = p.cell[d * g.nodes + idx - (int64_t)j * g.s[d]];
        }
      }
      div_l += g.f[d] * transposed<T, R>(lt, ch, cl, g.op);
      if (p.stiff_grad && grad) {
        Q ut[2 * R + 1];
        if (near[d]) {
          reflected_line<R>(u1, e, se, a[d], g.n[d], odd, ut);
        } else {
#pragma unroll
          for (int j = -R; j <= R; ++j)
            ut[j + R] = lanes_at<Q>(u1, e + j * se);
        }
        ut[R] = uc;
        if (near[d])
          lt[R] = lw;
        gsum += stiffness_density<T, R>(ut, lt, near[d] ? lw : lc, sc, sp[d],
                                        sm[d], g.f[d], g.op);
      }
    }
  }
  Q ln = p.damp ? damped<T, Q>(lc, lo_, mi * div_l, mi, p.damp[idx], p.dt)
                : -lo_ + (T)2 * lc + mi * div_l;
  if (held)
    ln = (T)0;
  // the inertia density by parts in time (scalar_sensitivity.cu:512-514)
  Q cm = chain ? accm : (T)0;
  cm = cm - p.mf * uc * (ln - (T)2 * lc + lo_);
  Q cs = chain ? accs : (T)0;
  cs = cs - p.mf * gsum;
  if (reach) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      if (!ext[d])
        continue;
      const int N = g.n[d], L = p.pml.lo[d], H = p.pml.hi[d];
      const int64_t size = slab_size<NDIM>(g, d, L, H) << sh;
      const T *eta = p.pml.eta[d][n & 1] + b;
      // the slab index along axis d is base + row * stride (the other axes'
      // offset is the node's own), so each tap costs one 32-bit product (a
      // slab holds fewer nodes than a field, which stays under 2^31)
      const int base = (int)slab_index<NDIM>(g, a, d, 0, L, H);
      const int stride = (int)slab_index<NDIM>(g, a, d, 1, L, H) - base;
      Q sum = gather<T, R, Q>(eta, size, g.op, [&](int k) {
        const int row = slab_row(a[d] + k, N, L, H);
        return row < 0 ? (int64_t)-1 : (int64_t)(base + row * stride) << sh;
      });
#pragma unroll
      for (int high = 0; high < 2; ++high) {
        // the reads past the wall that land on a, folded back
        const int ghost = high ? 2 * (N - 2) - a[d] : 2 - a[d];
        if (high ? ghost < N - 1 : ghost > 0)
          continue;
        bool on[R];
        Q e3[R], e4[R];
#pragma unroll
        for (int k = 1; k <= R; ++k) {
          const int ai = high ? N - 1 - k : k;
          const int row = slab_row(ai, N, L, H);
          on[k - 1] = row >= 0 && ghost - ai >= -R && ghost - ai <= R;
          const int64_t i =
              on[k - 1] ? (int64_t)(base + row * stride) << sh : 0;
          e3[k - 1] = lanes_at<Q>(eta, 3 * size + i);
          e4[k - 1] = lanes_at<Q>(eta, 4 * size + i);
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
          keep(e3[k]);
          keep(e4[k]);
        }
        Q fold = (T)0;
#pragma unroll
        for (int k = 1; k <= R; ++k) {
          const int ai = high ? N - 1 - k : k;
          const Q term = e3[k - 1] * plus_tap<T, R>(ghost - ai, g.op) +
                         e4[k - 1] * minus_tap<T, R>(ghost - ai, g.op);
          fold = on[k - 1] ? fold + term : fold;
        }
        sum += ((g.dirichlet >> (2 * d + high)) & 1) ? -fold : fold;
      }
      const Q delta = mi * sum / w;
      ln += delta;
      cm -= p.mf * uc * delta;
    }
  }
  for (int k = r0; k < r1; ++k)
#pragma unroll
    for (int i = 0; i < V; ++i)
      if (p.rec.takes(k, idx, b + i)) {
        const T load = p.rec.weight[k] * p.adjoint[p.rec.at(m, k, b + i)];
        lane(ln, i) += load;
        lane(cm, i) += -p.mf * lane(uc, i) * load;
      }
  lanes_put(l0, e, ln);
  // the memory the next level gathers, from this level's final lambda
  if (layer) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      if (prow[d] < 0)
        continue;
      const int N = g.n[d], L = p.pml.lo[d], H = p.pml.hi[d], ad = a[d];
      const int64_t size = slab_size<NDIM>(g, d, L, H) << sh;
      const T *an = p.pml.profile[d], *bn = an + N, *ah = an + 2 * N,
              *bh = an + 3 * N;
      pml_adjoint_from(cur[d], p.pml.eta[d][(n + 1) & 1] + own[d], size,
                       held ? Q((T)0) : w * ln, p.cell[d * g.nodes + idx],
                       p.cell[d * g.nodes + idx - g.s[d]], g.f[d], an[ad],
                       bn[ad], ah[ad], bh[ad], ah[ad - 1], bh[ad - 1]);
    }
  }
  if (grad)
    add_gradient(p, idx, b == 0, accm, accs, cm, cs);
  if (reverse) {
    // the lossless step run backwards: u^(n - 1) from u^n and u^(n + 1)
    Q v = replay;
    if (strip_col < 0) {
      bool deep = true;
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
        deep &= a[d] >= R && a[d] <= g.n[d] - 1 - R;
      Q taps[NDIM][2 * R + 1];
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        const int64_t se = (int64_t)g.s[d] << sh;
        const bool wall = !(deep || (a[d] >= R && a[d] <= g.n[d] - 1 - R));
#pragma unroll
        for (int j = -R; j <= R; ++j) {
          if constexpr (NDIM == 2) {
            // one load a tap, moved and signed at a wall, in 32 bits (in
            // the plane the faster form; in 3D the one below)
            bool neg = false;
            const int k = wall ? mirror_step(a[d], g.n[d],
                                             g.dirichlet >> (2 * d), j, neg)
                               : j;
            const Q v = lanes_at<Q>(u1, (int)e + k * (int)se);
            taps[d][j + R] = neg ? -v : v;
          } else {
            taps[d][j + R] = wall ? mirrored<Q>(u1, e, se, a[d], g.n[d],
                                                g.dirichlet >> (2 * d), j)
                                  : lanes_at<Q>(u1, e + j * se);
          }
        }
      }
      Q div[NDIM];
#pragma unroll
      for (int d = 0; d < NDIM; ++d)
        div[d] = flux<T, R>(taps[d], p.cell[d * g.nodes + idx],
                            p.cell[d * g.nodes + idx - g.s[d]], g.op);
      const Q lap = combine_axes<T, NDIM, R>(div, g.f);
      v = p.rdamp ? damped<T, Q>(uc, lanes_at<Q>(uo, e), mi * lap, mi,
                                 p.rdamp[idx], p.dt)
                  : -lanes_at<Q>(uo, e) + (T)2 * uc + mi * lap;
      if (held)
        v = (T)0;
      for (int k = s0; k < s1; ++k)
#pragma unroll
        for (int i = 0; i < V; ++i)
          if (p.src.takes(k, idx, b + i))
            lane(v, i) +=
                p.src.weight[k] * p.signal[p.src.at(n + 1, k, b + i)];
    }
    lanes_put(uo, e, v);
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
      const int faces = (g.neumann | g.dirichlet) >> (2 * d);
      const int odd = g.dirichlet >> (2 * d);
      const int64_t se = (int64_t)g.s[d] << sh;
      if ((faces & 1) && a[d] == 2)
        lanes_put(uo, e - 2 * se, (odd & 1) ? -v : v);
      if ((faces & 2) && a[d] == g.n[d] - 3)
        lanes_put(uo, e + 2 * se, (odd & 2) ? -v : v);
    }
  }
}

// ------------------------------------------------------ field gradients
// The two densities of a forward triplet (u0, u1, u2) against an adjoint
// field l1 (cuwave/kernels/scalar_sensitivity.cu:303-350), each subtracted
// from its accumulator: mass inv_dt2 l1 (u2 - 2 u1 + u0); stiffness the
// harmonic cell means differentiated, per axis as the adjoint kernels form
// it (backward_general), mirrored past the walls and with l1 zero on a held
// wall node. One lane, every interior node; g.f holds the step factors over
// dt^2 (the operator L, not dt^2 L).
template <typename T, int NDIM, int R>
__global__ __launch_bounds__(kThreads) void gradient_kernel(
    const Geom<T> g, const T *__restrict__ stiff, const T *__restrict__ u0,
    const T *__restrict__ u1, const T *__restrict__ u2,
    const T *__restrict__ l1, T *__restrict__ gmass, T *__restrict__ gstiff,
    const T inv_dt2) {
  constexpr int REACH = R + 2; // scalar_sensitivity.cu:44 WALL_REACH
  int a[3];
  if (!general_node<NDIM>(Region{}, g, a, blockIdx.x, gridDim.x, 0))
    return;
  int64_t idx = 0;
  bool held = false, near[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    idx += (int64_t)a[d] * g.s[d];
    held |= held_wall(a[d], g.n[d], g.dirichlet >> (2 * d));
    near[d] = a[d] < REACH || a[d] > g.n[d] - 1 - REACH;
  }
  const T uc = u1[idx], lc = l1[idx];
  gmass[idx] -= inv_dt2 * lc * (u2[idx] - (T)2 * uc + u0[idx]);
  const T sc = stiff[idx];
  const T lw = held ? (T)0 : lc;
  T gsum = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const int odd = g.dirichlet >> (2 * d);
    const int64_t s = g.s[d];
    T ut[2 * R + 1], lt[2 * R + 1], sp, sm;
    if (near[d]) {
      reflected_line<R>(u1, idx, s, a[d], g.n[d], odd, ut);
      reflected_line<R>(l1, idx, s, a[d], g.n[d], odd, lt);
      sp = reflected<T>(stiff, idx, s, a[d], g.n[d], 0, 1);
      sm = reflected<T>(stiff, idx, s, a[d], g.n[d], 0, -1);
      lt[R] = lw;
    } else {
#pragma unroll
      for (int j = -R; j <= R; ++j) {
        ut[j + R] = u1[idx + j * s];
        lt[j + R] = l1[idx + j * s];
      }
      sp = stiff[idx + s];
      sm = stiff[idx - s];
    }
    ut[R] = uc;
    gsum += stiffness_density<T, R>(ut, lt, near[d] ? lw : lc, sc, sp, sm,
                                    g.f[d], g.op);
  }
  gstiff[idx] -= gsum;
}

// --------------------------------------------------- superposition step
// scalar_sensitivity.cu:599-668 (step = 1) and :352-385 (step = 0), one shot
template <typename T, int NDIM, int R>
__global__ __launch_bounds__(kThreads) void superposed_kernel(
    Geom<T> g, const T *__restrict__ inv, const T *__restrict__ cell,
    const T *__restrict__ u0, const T *__restrict__ u1, T *__restrict__ u2,
    T *__restrict__ acc_mass, T *__restrict__ acc_stiff, T ft, T fs,
    int step) {
  const Region full{};
  int a[3];
  if (!general_node<NDIM>(full, g, a, blockIdx.x, gridDim.x, g.lshift))
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
      t[j + R] = mirrored<T>(u1, idx, g.s[d], a[d], g.n[d],
                          g.dirichlet >> (2 * d), j);
    div[d] = flux<T, R>(t, cell[d * g.nodes + idx],
                        cell[d * g.nodes + idx - g.s[d]], g.op);
  }
  const T lap = combine_axes<T, NDIM, R>(div, g.f);
  const T uc = u1[idx];
  const T dudt = uc - u2[idx];
  acc_mass[idx] += ft * dudt * dudt;
  acc_stiff[idx] += fs * sum;
  u2[idx] = held ? (T)0 : -u0[idx] + (T)2 * uc + inv[idx] * lap;
}

} // namespace cw
