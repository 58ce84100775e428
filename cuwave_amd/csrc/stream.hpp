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
// block is 256 / W neighbouring columns (W the shot lanes of a node, a
// compile-time constant here), and blocks are dealt to the XCDs in
// contiguous runs, rows of blocks across the box, so the columns whose taps
// overlap share one L2.
//
// Every load of a row is issued a row ahead into one of two row buffers
// that alternate (the loop is unrolled by two, so no buffer is copied while
// its loads are in flight), and all of them are issued on every row, the
// last row's repeated at the end of a chunk: the compiler then waits for
// exactly the older buffer and each wave keeps a whole row of loads in
// flight while it computes. Fields and materials go through buffer
// resources: one vector offset per thread for the fields and one for the
// nodes, advanced a row at a time; a fast-axis tap is a constant in the
// instruction, a row or a time slot one scalar offset.
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

// the profile of axis d at node a and at the half point below it
template <typename T> struct Prof {
  T an, bn, ah, bh, ah1, bh1;
};
template <typename T>
__device__ __forceinline__ Prof<T> profile_at(const T *pr, int N, int a) {
  return {pr[a], pr[N + a], pr[2 * N + a], pr[3 * N + a], pr[2 * N + a - 1],
          pr[3 * N + a - 1]};
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

constexpr int log2_of(int w) { return w <= 1 ? 0 : 1 + log2_of(w / 2); }

// The offsets of a thread's march over its W-lane columns. `f` is the byte
// offset of the field element (y, x - R) of its lane, `n` that of the node
// (y - 1, x - R) and `m` the same node's in an int32 map, for the block's
// first row; each row adds frow, nrow, mrow. Fast-axis taps are the
// constants fx(k), nx(k), mx(k).
template <typename T, int R, int W> struct March {
  static constexpr uint32_t Z = sizeof(T);
  static constexpr int SH = log2_of(W);
  uint32_t f, n, m;
  uint32_t frow, nrow, mrow;
  static constexpr uint32_t fx(int k) { return (uint32_t)((k + R) * W) * Z; }
  static constexpr uint32_t nx(int k) { return (uint32_t)(k + R) * Z; }
  static constexpr uint32_t mx(int k) { return (uint32_t)(k + R) * 4u; }
};

template <typename T, int R, int W>
__device__ __forceinline__ March<T, R, W> march_at(const Geom<T> &g, int ya,
                                                   int xl, int b) {
  using M = March<T, R, W>;
  M m;
  const uint32_t s0 = (uint32_t)g.s[0];
  const uint32_t node = (uint32_t)ya * s0 + (uint32_t)(xl - R);
  m.f = ((node << M::SH) + (uint32_t)b) * M::Z;
  m.n = (node - s0) * M::Z;
  m.m = (node - s0) * 4u;
  m.frow = (s0 << M::SH) * M::Z;
  m.nrow = s0 * M::Z;
  m.mrow = s0 * 4u;
  return m;
}

// --------------------------------------------------------------- forward
template <typename T, int R> struct FwdRow {
  T feed;       // u^(t-1) at (y + R, x): the queue's new entry
  T x[2 * R];   // u^(t-1) at (y, x -+ k)
  T u0, mi, gp1, gm1, gp0;
  int col;      // the node's strip column (or -1)
};

template <typename T, int R, int W>
__device__ __forceinline__ void forward_stream(const FwdParams<T> &p,
                                               const Box &box, const int t,
                                               int bid, int nb) {
  using Mch = March<T, R, W>;
  constexpr int SH = Mch::SH;
  const Geom<T> &g = p.g;
  int ya, yb, xb;
  if (!box_item(box, kThreads / W, bid, nb, ya, yb, xb))
    return;
  const int b = threadIdx.x & (W - 1);
  const int x = xb + (int)(threadIdx.x >> SH);
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1); // every thread loads
  const int tx = __builtin_amdgcn_readfirstlane(xl) / TileShape<2>::TX;
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = ((int64_t)(L0 + H0) * g.np[1]) << SH;
  const int64_t size1 = ((int64_t)(L1 + H1) * g.np[0]) << SH;
  const int prow1 = L1 + H1 ? slab_row(xl, N1, L1, H1) : -1;
  Prof<T> pf1{};
  if (prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, xl);
  const uint32_t LB = (uint32_t)((g.nodes << SH) * sizeof(T));
  const uint32_t NB = (uint32_t)(g.nodes * sizeof(T));
  const Buf U1 = make_buf(p.u.at(t - 1), LB);
  const Buf U0 = make_buf(p.u.at(t - 2), LB);
  const Buf U2 = make_buf(p.u.at(t), LB);
  // the prepared materials: cell means of the slow axis, of the fast axis,
  // the inverse inertia
  const Buf MAT = make_buf(p.cell, 3 * NB);
  const Buf MAP =
      make_buf(p.strip.map, p.strip.map ? (uint32_t)(g.nodes * 4) : 0u);
  const Mch M = march_at<T, R, W>(g, ya, xl, b);
  const uint32_t feed = R * M.frow;     // the field row y + R
  const uint32_t cy = M.nrow;           // node row y, slow-axis cells
  const uint32_t cx = NB + M.nrow;      // fast-axis cells
  const uint32_t inv = 2 * NB + M.nrow; // inverse inertia
  const bool has_src = chunk_has(p.src, g, ya, yb, tx);
  const bool has_rec = chunk_has(p.rec, g, ya, yb, tx);

  auto load = [&](FwdRow<T, R> &r, int y) {
    const uint32_t f = M.f + (uint32_t)(y - ya) * M.frow;
    const uint32_t n = M.n + (uint32_t)(y - ya) * M.nrow;
    r.feed = bload<T>(U1, f + Mch::fx(0), feed);
#pragma unroll
    for (int k = 1; k <= R; ++k) {
      r.x[R - k] = bload<T>(U1, f + Mch::fx(-k), 0);
      r.x[R - 1 + k] = bload<T>(U1, f + Mch::fx(k), 0);
    }
    r.u0 = bload<T>(U0, f + Mch::fx(0), 0);
    r.mi = bload<T>(MAT, n + Mch::nx(0), inv);
    r.gp1 = bload<T>(MAT, n + Mch::nx(0), cx);
    r.gm1 = bload<T>(MAT, n + Mch::nx(-1), cx);
    r.gp0 = bload<T>(MAT, n + Mch::nx(0), cy);
    r.col = bload_int(MAP, M.m + (uint32_t)(y - ya) * M.mrow + Mch::mx(0),
                      M.mrow);
  };

  // the queue's first rows ya - R .. ya + R - 1 (every offset non-negative:
  // the box starts R + 2 rows down)
  T q[2 * R + 1];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j)
    q[j] = bload<T>(U1, M.f + Mch::fx(0) + (uint32_t)j * M.frow - R * M.frow,
                    0);
  T gm0 = bload<T>(MAT, M.n + Mch::nx(0), 0);

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
                             (((int64_t)prow0 * g.np[1] + x) << SH) + b,
                         size0, Dp, Dm, r.gp0, gm0, g.f[0], mi, pf.an, pf.bn,
                         pf.ah, pf.bh, pf.ah1, pf.bh1);
      }
      if (prow1 >= 0) {
        T Dp, Dm;
        cell_diffs<T, R>(tf, g.op, Dp, Dm);
        v += pml_forward(p.pml.aux[1] +
                             (((int64_t)y * (L1 + H1) + prow1) << SH) + b,
                         size1, Dp, Dm, r.gp1, r.gm1, g.f[1], mi, pf1.an,
                         pf1.bn, pf1.ah, pf1.bh, pf1.ah1, pf1.bh1);
      }
      const int nc = y * g.s[0] + x;
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      if (has_src)
        for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
          if (p.src.takes(k, nc, b))
            v += p.src.weight[k] * p.signal[p.src.at(t, k, b)];
      bstore(U2, v, M.f + (uint32_t)(y - ya) * M.frow + Mch::fx(0), 0);
      if (has_rec)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b))
            p.record[p.rec.at(t, k, b)] = v;
      if ((prow0 >= 0 || prow1 >= 0) && p.strip.map) {
        const int col = r.col;
        if (col >= 0) {
          const int64_t at =
              (((int64_t)(t - p.strip.first) * p.strip.count + col) << SH) + b;
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

template <typename T, int R, int W, bool STIFF>
__device__ __forceinline__ void backward_stream(const BwdParams<T> &p,
                                                const Box &box, const int n,
                                                int bid, int nb) {
  using Mch = March<T, R, W>;
  constexpr int SH = Mch::SH;
  const Geom<T> &g = p.g;
  int ya, yb, xb;
  if (!box_item(box, kThreads / W, bid, nb, ya, yb, xb))
    return;
  const int b = threadIdx.x & (W - 1);
  const int x = xb + (int)(threadIdx.x >> SH);
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1);
  const int tx = __builtin_amdgcn_readfirstlane(xl) / TileShape<2>::TX;
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = ((int64_t)(L0 + H0) * g.np[1]) << SH;
  const int64_t size1 = ((int64_t)(L1 + H1) * g.np[0]) << SH;
  const int m = p.steps - 1 - n;
  const bool reverse = p.reverse && n <= p.steps - 2;
  const bool chain = W == 1 && !p.combine;
  // the column: its slab row along the fast axis, whether the transposed
  // taps of that axis reach the slab, its profile
  const int prow1 = L1 + H1 ? slab_row(xl, N1, L1, H1) : -1;
  const bool ext1 = (L1 && xl - 1 < L1 + R) || (H1 && N1 - 2 - xl < H1 + R);
  Prof<T> pf1{};
  if (prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, xl);
  const uint32_t LB = (uint32_t)((g.nodes << SH) * sizeof(T));
  const uint32_t NB = (uint32_t)(g.nodes * sizeof(T));
  // lambda: both slots under one resource, each level a slot offset; u the
  // same when the reverse march writes its second slot, else one level
  const uint32_t l1s = p.lam.slot_of(n + 1) * LB, l0s = p.lam.slot_of(n) * LB;
  const Buf LAM = make_buf(p.lam.base, 2 * LB);
  const Buf U = reverse ? make_buf(p.u.base, 2 * LB) : make_buf(p.u.at(n), LB);
  const uint32_t u1s = reverse ? p.u.slot_of(n) * LB : 0u;
  const uint32_t uos = reverse ? p.u.slot_of(n + 1) * LB : 0u;
  const Buf MAT = make_buf(p.cell, 3 * NB);
  const Buf GM = make_buf(p.gmass, NB);
  const Buf GS = make_buf(p.gstiff, p.combine ? 0u : NB);
  const Buf SK = make_buf(p.stiff, STIFF ? NB : 0u);
  const Buf MAP = make_buf(p.strip.map, reverse && p.strip.map
                                            ? (uint32_t)(g.nodes * 4)
                                            : 0u);
  const Mch M = march_at<T, R, W>(g, ya, xl, b);
  const uint32_t feed = R * M.frow;
  const uint32_t cy = M.nrow, cx = NB + M.nrow, inv = 2 * NB + M.nrow;
  const uint32_t cfeed = R * M.nrow; // slow-axis cell of row y + R - 1
  const bool has_rec = chunk_has(p.rec, g, ya, yb, tx);
  const bool has_src = reverse && chunk_has(p.src, g, ya, yb, tx);

  auto load = [&](BwdRow<T, R> &r, int y) {
    const uint32_t f = M.f + (uint32_t)(y - ya) * M.frow;
    const uint32_t nn = M.n + (uint32_t)(y - ya) * M.nrow;
    r.lfeed = bload<T>(LAM, f + Mch::fx(0), l1s + feed);
    r.ufeed = bload<T>(U, f + Mch::fx(0), u1s + feed);
#pragma unroll
    for (int k = 1; k <= R; ++k) {
      r.lx[R - k] = bload<T>(LAM, f + Mch::fx(-k), l1s);
      r.lx[R - 1 + k] = bload<T>(LAM, f + Mch::fx(k), l1s);
      r.ux[R - k] = bload<T>(U, f + Mch::fx(-k), u1s);
      r.ux[R - 1 + k] = bload<T>(U, f + Mch::fx(k), u1s);
    }
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
      r.cx[k] = bload<T>(MAT, nn + Mch::nx(k - R), cx);
    r.cfeed = bload<T>(MAT, nn + Mch::nx(0), cfeed);
    r.lo = bload<T>(LAM, f + Mch::fx(0), l0s);
    r.uo = bload<T>(U, f + Mch::fx(0), uos);
    r.mi = bload<T>(MAT, nn + Mch::nx(0), inv);
    r.accm = bload<T>(GM, nn + Mch::nx(0), M.nrow);
    r.accs = bload<T>(GS, nn + Mch::nx(0), M.nrow);
    if (STIFF) {
      r.sk[0] = bload<T>(SK, nn + Mch::nx(0), M.nrow);
      r.sk[1] = bload<T>(SK, nn + Mch::nx(0), 2 * M.nrow);
      r.sk[2] = bload<T>(SK, nn + Mch::nx(0), 0);
      r.sk[3] = bload<T>(SK, nn + Mch::nx(1), M.nrow);
      r.sk[4] = bload<T>(SK, nn + Mch::nx(-1), M.nrow);
    }
    r.col = bload_int(MAP, M.m + (uint32_t)(y - ya) * M.mrow + Mch::mx(0),
                      M.mrow);
  };

  // queues along the slow axis: lambda^(n + 1) and u^n rows y - R .. y + R,
  // the cell means of rows y - R .. y + R - 1
  T ql[2 * R + 1], qu[2 * R + 1], qc[2 * R];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j) {
    // rows ya - R + j (the box starts R + 2 rows down: offsets stay positive)
    const uint32_t f = M.f + Mch::fx(0) + (uint32_t)j * M.frow - R * M.frow;
    ql[j] = bload<T>(LAM, f, l1s);
    qu[j] = bload<T>(U, f, u1s);
    if (j < 2 * R - 1)
      qc[j] = bload<T>(MAT, M.n + Mch::nx(0) + (uint32_t)(j + 1) * M.nrow -
                                R * M.nrow,
                       0);
  }

  auto step = [&](const BwdRow<T, R> &r, int y) {
    ql[2 * R] = r.lfeed;
    qu[2 * R] = r.ufeed;
    qc[2 * R - 1] = r.cfeed;
    if (active) {
      const uint32_t f = M.f + (uint32_t)(y - ya) * M.frow + Mch::fx(0);
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
              return row >= 0 ? ((int64_t)row * g.np[1] + x) << SH
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
              return row >= 0 ? ((int64_t)y * (L1 + H1) + row) << SH
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
      bstore(LAM, ln, f, l0s);
      // the memory the next level gathers, from this level's final lambda
      if (prow0 >= 0) {
        const Prof<T> pf0 = profile_at(p.pml.profile[0], N0, y);
        const int64_t i = (((int64_t)prow0 * g.np[1] + x) << SH) + b;
        pml_adjoint_update(p.pml.eta[0][n & 1] + i,
                           p.pml.eta[0][(n + 1) & 1] + i, size0, ln, chi0[0],
                           clo0[0], g.f[0], pf0.an, pf0.bn, pf0.ah, pf0.bh,
                           pf0.ah1, pf0.bh1);
      }
      if (prow1 >= 0) {
        const int64_t i = (((int64_t)y * (L1 + H1) + prow1) << SH) + b;
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
               << SH) +
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
        bstore(U, v, f, uos);
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
// band, or every node in 1D and 3D). W is the shot lanes of a node for the
// stream blocks (0 where there are none).
template <typename T, int NDIM, int R, int W>
__global__ __launch_bounds__(kThreads) void forward_step(const FwdParams<T> p,
                                                         const Box box,
                                                         const int t,
                                                         const int streams) {
  if constexpr (NDIM == 2 && W > 0) {
    if ((int)blockIdx.x < streams) {
      forward_stream<T, R, W>(p, box, t, blockIdx.x, streams);
      return;
    }
  }
  forward_general<T, NDIM, R>(p, t, blockIdx.x - streams,
                              gridDim.x - streams);
}

template <typename T, int NDIM, int R, int W, bool STIFF>
__global__ __launch_bounds__(kThreads) void backward_step(const BwdParams<T> p,
                                                          const Box box,
                                                          const int n,
                                                          const int streams) {
  if constexpr (NDIM == 2 && W > 0) {
    if ((int)blockIdx.x < streams) {
      backward_stream<T, R, W, STIFF>(p, box, n, blockIdx.x, streams);
      return;
    }
  }
  backward_general<T, NDIM, R>(p, n, blockIdx.x - streams,
                               gridDim.x - streams);
}

} // namespace cw
