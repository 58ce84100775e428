// Streaming kernels for the bulk of a 2D grid: every node at least
// order / 2 + 2 nodes off the walls, the PML included. The general kernels
// (scalar_kernels.hpp) take the thin band that is left, where the wall
// mirror, the ghost refill and the held nodes live.
//
// A thread is one shot lane of one column and marches a chunk of rows down
// the slow axis, keeping the 2R + 1 values of its column that the slow-axis
// taps read in a register queue: each row costs one new load per queued
// field. The fast-axis taps are the same lane of the neighbouring columns,
// each a contiguous run of the wave (common.hpp), served by L1 and L2. A
// block is 256 / lanes neighbouring columns, and blocks are dealt to the
// XCDs in contiguous runs, rows of blocks across the box, so the columns
// whose taps overlap share one L2.
//
// Every load of a row is issued a row ahead into one of two row buffers
// that alternate (the loop is unrolled by two, so no buffer is copied while
// its loads are in flight), and all of them are issued on every row, the
// last row's repeated at the end of a chunk: the compiler then waits for
// exactly the older buffer and each wave keeps a whole row of loads in
// flight while it computes. Fields and materials go through buffer
// resources: one vector offset per thread advanced a row at a time, the tap
// offsets uniform.
//
// A row of the slow axis is either in that axis' slab or not, so its layer
// terms are uniform over the block; the fast axis' slab columns are fixed
// per thread for the whole chunk.
//
// The arithmetic is the general kernels' term by term, in the same order,
// so both paths give the same bits; the cell weight of every node here is 1.
#pragma once

#include "scalar_kernels.hpp"

namespace cw {

// the box of a sweep: rows [y0, y1), columns [x0, x1)
struct Box {
  int y0, y1, x0, x1;
  int chunk; // rows a thread marches
};

// stream block `bid` of `nb` (a multiple of the XCD count, at least the
// box's blocks): its rows [ya, yb) and first column xb, or false for a
// padding block. Blocks start at multiples of their width, so no wave
// straddles a point tile.
__device__ __forceinline__ bool box_item(const Box &box, int cols, int bid,
                                         int nb, int &ya, int &yb, int &xb) {
  const int xa = box.x0 / cols * cols;
  const int ncx = (box.x1 - xa + cols - 1) / cols;
  const int nchunk = (box.y1 - box.y0 + box.chunk - 1) / box.chunk;
  const int w = xcd_tile(bid, nb);
  if (w >= nchunk * ncx)
    return false;
  ya = box.y0 + (w / ncx) * box.chunk;
  yb = min(ya + box.chunk, box.y1);
  xb = xa + (w % ncx) * cols;
  return true;
}

// the profile of axis d at node a and at the half point below it
template <typename T> struct Prof {
  T an, bn, ah, bh, ah1, bh1;
};
template <typename T>
__device__ __forceinline__ Prof<T> profile_at(const T *pr, int N, int a) {
  return {pr[a], pr[N + a], pr[2 * N + a], pr[3 * N + a], pr[2 * N + a - 1],
          pr[3 * N + a - 1]};
}

// whether a point list holds a point in the tiles of rows [ya, yb) at tile
// column tx: looked up once a chunk, so rows without points (nearly all)
// skip the lists without waiting on a load
template <typename T>
__device__ __forceinline__ bool chunk_has(const Points<T> &pts,
                                          const Geom<T> &g, int ya, int yb,
                                          int tx) {
  if (!pts.starts)
    return false;
  int any = 0;
  for (int ty = ya / TileShape<2>::TY; ty <= (yb - 1) / TileShape<2>::TY; ++ty)
    any |= pts.starts[ty * g.tf + tx + 1] - pts.starts[ty * g.tf + tx];
  return __builtin_amdgcn_readfirstlane(any) != 0;
}

// taps[0 .. 2R] from the fast-axis taps x[0 .. 2R) (k = -R .. -1, 1 .. R)
// and the centre
template <typename T, int R>
__device__ __forceinline__ void row_line(T *t, const T *x, T c) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    t[k] = x[k];
    t[R + 1 + k] = x[R + k];
  }
  t[R] = c;
}

// Offsets of a thread's march. `f` and `n` are the byte offsets of the
// field element and of the node at (y - R, x - R) for the current row y;
// F(j, k) and N(j, k) the uniform byte offsets of (y + j, x + k) from them.
template <typename T, int R> struct March {
  uint32_t f, n;         // at row ya
  uint32_t frow, nrow;   // bytes a row
  uint32_t es0, el, s0;  // element strides of a row and a column, node row
  __device__ __forceinline__ uint32_t F(int j, int k) const {
    return ((uint32_t)(j + R) * es0 + (uint32_t)(k + R) * el) * sizeof(T);
  }
  __device__ __forceinline__ uint32_t N(int j, int k) const {
    return ((uint32_t)(j + R) * s0 + (uint32_t)(k + R)) * sizeof(T);
  }
};

template <typename T, int R>
__device__ __forceinline__ March<T, R> march_at(const Geom<T> &g, int ya,
                                                int xl, int b) {
  March<T, R> m;
  const int sh = g.lshift;
  m.s0 = (uint32_t)g.s[0];
  m.es0 = m.s0 << sh;
  m.el = 1u << sh;
  const uint32_t node = (uint32_t)(ya - R) * m.s0 + (uint32_t)(xl - R);
  m.f = ((node << sh) + (uint32_t)b) * sizeof(T);
  m.n = node * sizeof(T);
  m.frow = m.es0 * sizeof(T);
  m.nrow = m.s0 * sizeof(T);
  return m;
}

// --------------------------------------------------------------- forward
template <typename T, int R> struct FwdRow {
  T feed;       // u^(t-1) at (y + R, x): the queue's new entry
  T x[2 * R];   // u^(t-1) at (y, x -+ k)
  T u0, mi, gp1, gm1, gp0;
  int col;      // the node's strip column (or -1)
};

template <typename T, int R>
__device__ __forceinline__ void forward_stream(const FwdParams<T> &p,
                                               const Box &box, const int t,
                                               int bid, int nb) {
  const Geom<T> &g = p.g;
  const int sh = g.lshift;
  int ya, yb, xb;
  if (!box_item(box, kThreads >> sh, bid, nb, ya, yb, xb))
    return;
  const int b = threadIdx.x & (g.lanes - 1);
  const int x = xb + (int)(threadIdx.x >> sh);
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1); // every thread loads
  const int tx = __builtin_amdgcn_readfirstlane(xl) / TileShape<2>::TX;
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = ((int64_t)(L0 + H0) * g.np[1]) << sh;
  const int64_t size1 = ((int64_t)(L1 + H1) * g.np[0]) << sh;
  const int prow1 = L1 + H1 ? slab_row(xl, N1, L1, H1) : -1;
  Prof<T> pf1{};
  if (prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, xl);
  const uint32_t fbytes = (uint32_t)((g.nodes << sh) * sizeof(T));
  const uint32_t nbytes = (uint32_t)(g.nodes * sizeof(T));
  const Buf U1 = make_buf(p.u.at(t - 1), fbytes);
  const Buf U0 = make_buf(p.u.at(t - 2), fbytes);
  const Buf U2 = make_buf(p.u.at(t), fbytes);
  const Buf INV = make_buf(p.inv, nbytes);
  const Buf CY = make_buf(p.cell, nbytes);           // cells along the slow axis
  const Buf CX = make_buf(p.cell + g.nodes, nbytes); // along the fast axis
  const Buf MAP = make_buf(p.strip.map, p.strip.map ? nbytes / sizeof(T) * 4 : 0u);
  const March<T, R> M = march_at<T, R>(g, ya, xl, b);
  const uint32_t mrow = M.nrow / sizeof(T) * 4, m0 = M.n / sizeof(T) * 4;
  const bool has_src = chunk_has(p.src, g, ya, yb, tx);
  const bool has_rec = chunk_has(p.rec, g, ya, yb, tx);

  auto load = [&](FwdRow<T, R> &r, int y) {
    const uint32_t f = M.f + (uint32_t)(y - ya) * M.frow;
    const uint32_t n = M.n + (uint32_t)(y - ya) * M.nrow;
    r.feed = bload<T>(U1, f, M.F(R, 0));
#pragma unroll
    for (int k = 1; k <= R; ++k) {
      r.x[R - k] = bload<T>(U1, f, M.F(0, -k));
      r.x[R - 1 + k] = bload<T>(U1, f, M.F(0, k));
    }
    r.u0 = bload<T>(U0, f, M.F(0, 0));
    r.mi = bload<T>(INV, n, M.N(0, 0));
    r.gp1 = bload<T>(CX, n, M.N(0, 0));
    r.gm1 = bload<T>(CX, n, M.N(0, -1));
    r.gp0 = bload<T>(CY, n, M.N(0, 0));
    r.col = bload_int(MAP, m0 + (uint32_t)(y - ya) * mrow, M.N(0, 0) / sizeof(T) * 4);
  };

  T q[2 * R + 1];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j)
    q[j] = bload<T>(U1, M.f, M.F(j - R, 0));
  T gm0 = bload<T>(CY, M.n, M.N(-1, 0));

  auto step = [&](const FwdRow<T, R> &r, int y) {
    q[2 * R] = r.feed;
    if (active) {
      T tf[2 * R + 1];
      row_line<T, R>(tf, r.x, q[R]);
      const T mi = r.mi;
      T div[2];
      div[0] = flux<T, R>(q, r.gp0, gm0, g.op);
      div[1] = flux<T, R>(tf, r.gp1, r.gm1, g.op);
      const T lap = combine_axes<T, 2, R>(div, g.f);
      T v = -r.u0 + (T)2 * q[R] + mi * lap;
      const int prow0 = L0 + H0 ? slab_row(y, N0, L0, H0) : -1;
      if (prow0 >= 0) {
        const Prof<T> pf = profile_at(p.pml.profile[0], N0, y);
        T Dp, Dm;
        cell_diffs<T, R>(q, g.op, Dp, Dm);
        v += pml_forward(p.pml.aux[0] +
                             (((int64_t)prow0 * g.np[1] + x) << sh) + b,
                         size0, Dp, Dm, r.gp0, gm0, g.f[0], mi, pf.an, pf.bn,
                         pf.ah, pf.bh, pf.ah1, pf.bh1);
      }
      if (prow1 >= 0) {
        T Dp, Dm;
        cell_diffs<T, R>(tf, g.op, Dp, Dm);
        v += pml_forward(p.pml.aux[1] +
                             (((int64_t)y * (L1 + H1) + prow1) << sh) + b,
                         size1, Dp, Dm, r.gp1, r.gm1, g.f[1], mi, pf1.an,
                         pf1.bn, pf1.ah, pf1.bh, pf1.ah1, pf1.bh1);
      }
      const int nc = y * g.s[0] + x;
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      if (has_src)
        for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
          if (p.src.takes(k, nc, b))
            v += p.src.weight[k] * p.signal[p.src.at(t, k, b)];
      bstore(U2, v, M.f + (uint32_t)(y - ya) * M.frow, M.F(0, 0));
      if (has_rec)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b))
            p.record[p.rec.at(t, k, b)] = v;
      if ((prow0 >= 0 || prow1 >= 0) && p.strip.map) {
        const int col = r.col;
        if (col >= 0) {
          const int64_t at =
              (((int64_t)(t - p.strip.first) * p.strip.count + col) << sh) + b;
          if (p.strip.bf16)
            ((uint16_t *)p.strip.record)[at] = bf16_bits((float)v);
          else
            ((T *)p.strip.record)[at] = v;
        }
      }
    }
    gm0 = r.gp0;
#pragma unroll
    for (int j = 0; j < 2 * R; ++j)
      q[j] = q[j + 1];
  };

  FwdRow<T, R> A, B;
  load(A, ya);
  for (int y = ya; y < yb; y += 2) {
    load(B, min(y + 1, yb - 1));
    step(A, y);
    load(A, min(y + 2, yb - 1));
    if (y + 1 < yb)
      step(B, y + 1);
  }
}

// --------------------------------------------------------------- adjoint
template <typename T, int R> struct BwdRow {
  T lfeed, ufeed;   // lambda^(n+1), u^n at (y + R, x)
  T lx[2 * R], ux[2 * R]; // their fast-axis taps at row y
  T cx[2 * R];      // fast-axis cell means at x - R .. x + R - 1
  T cfeed;          // slow-axis cell mean of row y + R - 1
  T lo, uo, mi;     // lambda^(n+2), u^(n+1), 1 / m at (y, x)
  T accm, accs;     // the gradient accumulators at the node
  T sk[5];          // stiffness at the node and its four neighbours
  int col;          // the node's strip column (or -1)
};

template <typename T, int R, bool STIFF>
__device__ __forceinline__ void backward_stream(const BwdParams<T> &p,
                                                const Box &box, const int n,
                                                int bid, int nb) {
  const Geom<T> &g = p.g;
  const int sh = g.lshift;
  int ya, yb, xb;
  if (!box_item(box, kThreads >> sh, bid, nb, ya, yb, xb))
    return;
  const int b = threadIdx.x & (g.lanes - 1);
  const int x = xb + (int)(threadIdx.x >> sh);
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1);
  const int tx = __builtin_amdgcn_readfirstlane(xl) / TileShape<2>::TX;
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = ((int64_t)(L0 + H0) * g.np[1]) << sh;
  const int64_t size1 = ((int64_t)(L1 + H1) * g.np[0]) << sh;
  const int m = p.steps - 1 - n;
  const bool reverse = p.reverse && n <= p.steps - 2;
  const bool chain = g.lanes == 1 && !p.combine;
  // the column: its slab row along the fast axis, whether the transposed
  // taps of that axis reach the slab, its profile
  const int prow1 = L1 + H1 ? slab_row(xl, N1, L1, H1) : -1;
  const bool ext1 = (L1 && xl - 1 < L1 + R) || (H1 && N1 - 2 - xl < H1 + R);
  Prof<T> pf1{};
  if (prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, xl);
  const uint32_t fbytes = (uint32_t)((g.nodes << sh) * sizeof(T));
  const uint32_t nbytes = (uint32_t)(g.nodes * sizeof(T));
  const Buf L1b = make_buf(p.lam.at(n + 1), fbytes);
  const Buf U1b = make_buf(p.u.at(n), fbytes);
  const Buf L0b = make_buf(p.lam.at(n), fbytes);
  // without the reverse march u^(n + 1) is neither read nor written: an
  // empty resource turns its loads into zeros that touch no memory
  const Buf UOb = make_buf(p.u.at(n + 1), reverse ? fbytes : 0u);
  const Buf INV = make_buf(p.inv, nbytes);
  const Buf CY = make_buf(p.cell, nbytes);
  const Buf CX = make_buf(p.cell + g.nodes, nbytes);
  const Buf GM = make_buf(p.gmass, nbytes);
  const Buf GS = make_buf(p.gstiff, p.combine ? 0u : nbytes);
  const Buf SK = make_buf(p.stiff, STIFF ? nbytes : 0u);
  const Buf MAP = make_buf(p.strip.map, reverse && p.strip.map ? nbytes / sizeof(T) * 4 : 0u);
  const March<T, R> M = march_at<T, R>(g, ya, xl, b);
  const uint32_t mrow = M.nrow / sizeof(T) * 4, m0 = M.n / sizeof(T) * 4;
  const bool has_rec = chunk_has(p.rec, g, ya, yb, tx);
  const bool has_src = reverse && chunk_has(p.src, g, ya, yb, tx);

  auto load = [&](BwdRow<T, R> &r, int y) {
    const uint32_t f = M.f + (uint32_t)(y - ya) * M.frow;
    const uint32_t nn = M.n + (uint32_t)(y - ya) * M.nrow;
    r.lfeed = bload<T>(L1b, f, M.F(R, 0));
    r.ufeed = bload<T>(U1b, f, M.F(R, 0));
#pragma unroll
    for (int k = 1; k <= R; ++k) {
      r.lx[R - k] = bload<T>(L1b, f, M.F(0, -k));
      r.lx[R - 1 + k] = bload<T>(L1b, f, M.F(0, k));
      r.ux[R - k] = bload<T>(U1b, f, M.F(0, -k));
      r.ux[R - 1 + k] = bload<T>(U1b, f, M.F(0, k));
    }
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
      r.cx[k] = bload<T>(CX, nn, M.N(0, k - R));
    r.cfeed = bload<T>(CY, nn, M.N(R - 1, 0));
    r.lo = bload<T>(L0b, f, M.F(0, 0));
    r.uo = bload<T>(UOb, f, M.F(0, 0));
    r.mi = bload<T>(INV, nn, M.N(0, 0));
    r.accm = bload<T>(GM, nn, M.N(0, 0));
    r.accs = bload<T>(GS, nn, M.N(0, 0));
    if (STIFF) {
      r.sk[0] = bload<T>(SK, nn, M.N(0, 0));
      r.sk[1] = bload<T>(SK, nn, M.N(1, 0));
      r.sk[2] = bload<T>(SK, nn, M.N(-1, 0));
      r.sk[3] = bload<T>(SK, nn, M.N(0, 1));
      r.sk[4] = bload<T>(SK, nn, M.N(0, -1));
    }
    r.col = bload_int(MAP, m0 + (uint32_t)(y - ya) * mrow, M.N(0, 0) / sizeof(T) * 4);
  };

  // queues along the slow axis: lambda^(n + 1) and u^n rows y - R .. y + R,
  // the cell means of rows y - R .. y + R - 1
  T ql[2 * R + 1], qu[2 * R + 1], qc[2 * R];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j) {
    ql[j] = bload<T>(L1b, M.f, M.F(j - R, 0));
    qu[j] = bload<T>(U1b, M.f, M.F(j - R, 0));
    if (j < 2 * R - 1)
      qc[j] = bload<T>(CY, M.n, M.N(j - R, 0));
  }

  auto step = [&](const BwdRow<T, R> &r, int y) {
    ql[2 * R] = r.lfeed;
    qu[2 * R] = r.ufeed;
    qc[2 * R - 1] = r.cfeed;
    if (active) {
      const uint32_t f = M.f + (uint32_t)(y - ya) * M.frow;
      const int nc = y * g.s[0] + x;
      T lt1[2 * R + 1], ut1[2 * R + 1], chi1[R], clo1[R], chi0[R], clo0[R];
      row_line<T, R>(lt1, r.lx, ql[R]);
      row_line<T, R>(ut1, r.ux, qu[R]);
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        chi1[j - 1] = r.cx[R + j - 1];
        clo1[j - 1] = r.cx[R - j];
        chi0[j - 1] = qc[R + j - 1];
        clo0[j - 1] = qc[R - j];
      }
      const T mi = r.mi, lo_ = r.lo;
      // this row along the slow axis: slab row, reach of the transposed taps
      const int prow0 = L0 + H0 ? slab_row(y, N0, L0, H0) : -1;
      const bool ext0 = (L0 && y - 1 < L0 + R) || (H0 && N0 - 2 - y < H0 + R);
      const bool layer = prow0 >= 0 || prow1 >= 0;
      const bool grad = !(p.skip_layer && layer);
      const T accm = r.accm, accs = r.accs;
      const T lc = ql[R], uc = qu[R];
      T div_l = 0;
      div_l += g.f[0] * transposed<T, R>(ql, chi0, clo0, g.op);
      div_l += g.f[1] * transposed<T, R>(lt1, chi1, clo1, g.op);
      T ln = -lo_ + (T)2 * lc + mi * div_l;
      // the inertia density by parts in time (scalar_sensitivity.cu:512-514)
      T cm = chain ? accm : (T)0;
      cm = cm - p.mf * uc * (ln - (T)2 * lc + lo_);
      T cs = chain ? accs : (T)0;
      if (STIFF && grad) {
        T gsum = 0;
        gsum += stiffness_density<T, R>(qu, ql, lc, r.sk[0], r.sk[1], r.sk[2],
                                        g.f[0], g.op);
        gsum += stiffness_density<T, R>(ut1, lt1, lc, r.sk[0], r.sk[3],
                                        r.sk[4], g.f[1], g.op);
        cs = cs - p.mf * gsum;
      }
      // the transposed layer: the slab nodes' tap fields within reach
      if (ext0) {
        const T sum =
            gather<T, R>(p.pml.eta[0][n & 1] + b, size0, g.op, [&](int k) {
              const int row = slab_row(y + k, N0, L0, H0);
              return row >= 0 ? ((int64_t)row * g.np[1] + x) << sh
                              : (int64_t)-1;
            });
        const T delta = mi * sum;
        ln += delta;
        cm -= p.mf * uc * delta;
      }
      if (ext1) {
        const T sum =
            gather<T, R>(p.pml.eta[1][n & 1] + b, size1, g.op, [&](int k) {
              const int row = slab_row(x + k, N1, L1, H1);
              return row >= 0 ? ((int64_t)y * (L1 + H1) + row) << sh
                              : (int64_t)-1;
            });
        const T delta = mi * sum;
        ln += delta;
        cm -= p.mf * uc * delta;
      }
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      if (has_rec)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b)) {
            const T load = p.rec.weight[k] * p.adjoint[p.rec.at(m, k, b)];
            ln += load;
            cm += -p.mf * uc * load;
          }
      bstore(L0b, ln, f, M.F(0, 0));
      // the memory the next level gathers, from this level's final lambda
      if (prow0 >= 0) {
        const Prof<T> pf0 = profile_at(p.pml.profile[0], N0, y);
        const int64_t i = (((int64_t)prow0 * g.np[1] + x) << sh) + b;
        pml_adjoint_update(p.pml.eta[0][n & 1] + i,
                           p.pml.eta[0][(n + 1) & 1] + i, size0, ln, chi0[0],
                           clo0[0], g.f[0], pf0.an, pf0.bn, pf0.ah, pf0.bh,
                           pf0.ah1, pf0.bh1);
      }
      if (prow1 >= 0) {
        const int64_t i = (((int64_t)y * (L1 + H1) + prow1) << sh) + b;
        pml_adjoint_update(p.pml.eta[1][n & 1] + i,
                           p.pml.eta[1][(n + 1) & 1] + i, size1, ln, chi1[0],
                           clo1[0], g.f[1], pf1.an, pf1.bn, pf1.ah, pf1.bh,
                           pf1.ah1, pf1.bh1);
      }
      if (grad)
        add_gradient(p, nc, b, accm, accs, cm, cs);
      if (reverse) {
        const int strip_col = layer && p.strip.map ? r.col : -1;
        T v;
        if (strip_col >= 0) {
          // a layer node is replayed from the forward sweep's record
          const int64_t at =
              (((int64_t)(n - 1 - p.strip.first) * p.strip.count + strip_col)
               << sh) +
              b;
          v = p.strip.bf16
                  ? (T)bf16_value(((const uint16_t *)p.strip.record)[at])
                  : ((const T *)p.strip.record)[at];
        } else {
          T div[2];
          div[0] = flux<T, R>(qu, chi0[0], clo0[0], g.op);
          div[1] = flux<T, R>(ut1, chi1[0], clo1[0], g.op);
          const T lap = combine_axes<T, 2, R>(div, g.f);
          v = -r.uo + (T)2 * uc + mi * lap;
          if (has_src)
            for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
              if (p.src.takes(k, nc, b))
                v += p.src.weight[k] * p.signal[p.src.at(n + 1, k, b)];
        }
        bstore(UOb, v, f, M.F(0, 0));
      }
    }
#pragma unroll
    for (int j = 0; j < 2 * R; ++j) {
      ql[j] = ql[j + 1];
      qu[j] = qu[j + 1];
    }
#pragma unroll
    for (int j = 0; j < 2 * R - 1; ++j)
      qc[j] = qc[j + 1];
  };

  BwdRow<T, R> A, B;
  load(A, ya);
  for (int y = ya; y < yb; y += 2) {
    load(B, min(y + 1, yb - 1));
    step(A, y);
    load(A, min(y + 2, yb - 1));
    if (y + 1 < yb)
      step(B, y + 1);
  }
}

// ------------------------------------------------------------ the steps
// One launch a step: the first `streams` blocks (a multiple of the XCD
// count) march the stream box, the rest take the general nodes (the 2D wall
// band, or every node in 1D and 3D).
template <typename T, int NDIM, int R>
__global__ __launch_bounds__(kThreads) void forward_step(const FwdParams<T> p,
                                                         const Box box,
                                                         const int t,
                                                         const int streams) {
  if constexpr (NDIM == 2) {
    if ((int)blockIdx.x < streams) {
      forward_stream<T, R>(p, box, t, blockIdx.x, streams);
      return;
    }
  }
  forward_general<T, NDIM, R>(p, t, blockIdx.x - streams,
                              gridDim.x - streams);
}

template <typename T, int NDIM, int R, bool STIFF>
__global__ __launch_bounds__(kThreads) void backward_step(const BwdParams<T> p,
                                                          const Box box,
                                                          const int n,
                                                          const int streams) {
  if constexpr (NDIM == 2) {
    if ((int)blockIdx.x < streams) {
      backward_stream<T, R, STIFF>(p, box, n, blockIdx.x, streams);
      return;
    }
  }
  backward_general<T, NDIM, R>(p, n, blockIdx.x - streams,
                               gridDim.x - streams);
}

} // namespace cw
