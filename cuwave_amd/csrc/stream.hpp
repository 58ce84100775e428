// Streaming kernels for the bulk of a 2D grid: the rectangle of nodes at
// least order / 2 + 2 off the walls that neither lie in a PML slab nor reach
// one with their transposed taps (the stream box). The general kernels
// (scalar_kernels.hpp) take the frame around it: the slabs, the ring their
// transposed taps reach, and the wall band with the mirror, the ghost
// refill and the held nodes.
//
// A thread is one shot lane of one column and marches a chunk of rows down
// the slow axis, keeping the 2R + 1 values of its column that the slow-axis
// taps read in a register queue. A block is C = 256 / W neighbouring columns
// (W the shot lanes of a node, a compile-time constant here). Each row of
// a tapped field reaches the block once, as one contiguous run staged in
// LDS by the DMA loads of gfx950 (global_load_lds, 16 bytes a lane, no
// registers on the way): C + 2R columns with the halo, in a ring of R + 2
// rows. The per-node materials of a row are staged the same way.
//
// A row first reads its fast-axis taps (LDS row y), its queue feed (LDS row
// y + R) and its materials into registers; then it issues the DMA of row
// y + R + 1 and the thread's own values of the next row; then it computes
// from registers while those loads fly. Reading LDS first matters: the
// compiler cannot tell which LDS bytes an issued DMA writes, so any LDS
// read after one waits for every load in flight. One wait and one barrier
// end the row.
//
// Blocks are dealt to the XCDs in contiguous runs, rows of blocks across
// the box, so vertically neighbouring chunks share one L2.
//
// The box holds no layer node, so the core kernel carries no layer, point
// or strip code. A box block holding a source or sensor runs in the second
// launch of the step instead, with the point code (the layer code there is
// never reached).
//
// The arithmetic is the general kernels' term by term, in the same order,
// so both paths give the same bits; the cell weight of every node here is 1.
#pragma once

#include "scalar_kernels.hpp"

namespace cw {

// The box of a sweep, rows [y0, y1) and columns [x0, x1), cut into blocks
// of `chunk` rows by C columns (the first block column starting at the
// multiple of C below x0). Blocks [c0, c1) x [k0, k1) are the core, which
// the core kernel runs; with the box layer-free that is every block, and
// the blocks holding points are handed to the second launch.
struct Box {
  int y0, y1, x0, x1;
  int chunk;
  int c0, c1, k0, k1;
};

// whether a point list holds a point in the tiles of rows [ya, yb) at tile
// columns [tx0, tx1]: looked up once a block, so rows without points
// (nearly all) skip the lists without waiting on a load
template <typename T>
__device__ __forceinline__ bool chunk_has(const Points<T> &pts,
                                          const Geom<T> &g, int ya, int yb,
                                          int tx0, int tx1) {
  if (!pts.starts)
    return false;
  int any = 0;
  for (int ty = ya / TileShape<2>::TY; ty <= (yb - 1) / TileShape<2>::TY; ++ty)
    any |= pts.starts[ty * g.tf + tx1 + 1] - pts.starts[ty * g.tf + tx0];
  return __builtin_amdgcn_readfirstlane(any) != 0;
}

// the rows [ya, yb) (bit y - ya; a chunk is at most 32 rows) at which the
// wave's point tile column `tx` holds a point of `pts`: the wave's lanes scan
// the lists together, once a block, so that the rows without a point
// (nearly all) skip the lists without a load
template <typename T>
__device__ __forceinline__ uint32_t point_rows(const Points<T> &pts,
                                               const Geom<T> &g, int ya, int yb,
                                               int tx) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t mask = 0;
  for (int ty = ya / TileShape<2>::TY; ty <= (yb - 1) / TileShape<2>::TY;
       ++ty) {
    const int tile = ty * g.tf + tx;
    const int e = pts.starts[tile + 1];
    for (int k = pts.starts[tile] + lane; k < e; k += kWave) {
      const int row = pts.lin[k] / g.s[0];
      if (row >= ya && row < yb)
        mask |= 1u << (row - ya);
    }
  }
  for (int m = 1; m < kWave; m <<= 1)
    mask |= (uint32_t)__shfl_xor((int)mask, m, kWave);
  return __builtin_amdgcn_readfirstlane(mask);
}

// stream block `bid` of `nb` (a multiple of the XCD count) of the core or
// of the second launch: its rows [ya, yb) and first column xb, or false for a
// block the other kernel runs (or a padding block). A block starts at a
// multiple of its width, so no wave straddles a point tile.
template <bool EDGE, typename T>
__device__ __forceinline__ bool box_item(const Box &box, const Geom<T> &g,
                                         const Points<T> &src, bool use_src,
                                         const Points<T> &rec, int cols,
                                         int bid, int nb, int &ya, int &yb,
                                         int &xb) {
  const int xa = box.x0 / cols * cols;
  const int ncx = (box.x1 - xa + cols - 1) / cols;
  const int nchunk = (box.y1 - box.y0 + box.chunk - 1) / box.chunk;
  const int w = xcd_tile(bid, nb);
  int c, k;
  if (EDGE) {
    if (w >= nchunk * ncx)
      return false;
    c = w / ncx;
    k = w % ncx;
  } else {
    const int kw = box.k1 - box.k0;
    if (w >= (box.c1 - box.c0) * kw)
      return false;
    c = box.c0 + w / kw;
    k = box.k0 + w % kw;
  }
  ya = box.y0 + c * box.chunk;
  yb = min(ya + box.chunk, box.y1);
  xb = xa + k * cols;
  const bool core = c >= box.c0 && c < box.c1 && k >= box.k0 && k < box.k1;
  if (EDGE && !core)
    return true;
  const int tx0 = xb / TileShape<2>::TX;
  const int tx1 = min(xb + cols - 1, g.np[1] - 1) / TileShape<2>::TX;
  const bool points = (use_src && chunk_has(src, g, ya, yb, tx0, tx1)) ||
                      chunk_has(rec, g, ya, yb, tx0, tx1);
  return EDGE == points;
}

// Copy `bytes` contiguous bytes from global `src` to LDS `dst` by DMA, the
// block's four waves taking 64-lane chunks in turn from wave `first`; lanes
// past the end stay idle. SIZE is the bytes a lane moves (4 or 16).
template <int SIZE>
__device__ __forceinline__ void dma_block(const char *src, char *dst,
                                          int bytes, int first) {
  const int wave = ((int)(threadIdx.x >> 6) - first) & 3;
  const int lane = threadIdx.x & 63;
  constexpr int CH = 64 * SIZE;
  for (int c = wave * CH; c < bytes; c += 4 * CH)
    if (c + lane * SIZE < bytes)
      lds_dma<SIZE>(src + c + lane * SIZE, dst + c);
}

// The staging geometry of a block of a W-lane field: C columns, each row of
// the field staged with an R-column halo on both sides (SEGC columns, SEGB
// bytes) in a ring of R + 2 rows: while a step reads rows y .. y + R (the
// fast-axis taps of row y and the queue feed of row y + R), row y + R + 1
// arrives. DMA moves 16 bytes a lane where a node's lanes allow.
template <typename T, int R, int W> struct Stage {
  static constexpr int SH = log2_of(W);
  static constexpr int C = kThreads / W;
  static constexpr int SEGC = C + 2 * R;
  static constexpr int SEGB = SEGC * W * (int)sizeof(T);
  static constexpr int RING = R + 2;
  static constexpr int DMA = (W * sizeof(T)) % 16 == 0 ? 16 : 4;
};

// stage row y of a W-lane field level `f` (np1 the padded row length):
// nodes xb - R .. xb + C + R - 1 of the row, clipped at the row's end
template <typename T, int R, int W>
__device__ __forceinline__ void stage_row(const T *f, int y, int xb, int s0,
                                          int np1, char *ring, int first) {
  using S = Stage<T, R, W>;
  const int end = min(xb + S::C + R, np1);
  dma_block<S::DMA>((const char *)(f + ((int64_t)y * s0 + xb - R) * W),
                    ring + (y % S::RING) * S::SEGB,
                    (end - (xb - R)) * W * (int)sizeof(T), first);
}

// stage `count` node values from node `node` of a per-node array (clipped
// at the row end `lim`, exclusive, counted from `node`)
template <typename T>
__device__ __forceinline__ void stage_nodes(const T *a, int64_t node, int count,
                                            int lim, T *dst, int first) {
  dma_block<4>((const char *)(a + node), (char *)dst,
               min(count, lim) * (int)sizeof(T), first);
}

// --------------------------------------------------------------- forward
template <typename T, int R, int W, bool EDGE, bool ILLUM>
__device__ __forceinline__ void forward_stream(const FwdParams<T> &p,
                                               const Box &box, const int t,
                                               int bid, int nb) {
  using S = Stage<T, R, W>;
  constexpr int SH = S::SH, C = S::C;
  const Geom<T> &g = p.g;
  int ya, yb, xb;
  if (!box_item<EDGE>(box, g, p.src, true, p.rec, C, bid, nb, ya, yb, xb))
    return;
  // the staged rows of u^(t-1), and per material row slot (two alternating)
  // the slow-axis cells [C], the fast-axis cells of x - 1 .. [C + 1], 1 / m
  __shared__ __align__(16) char ring[S::RING * S::SEGB];
  __shared__ __align__(16) T lmat[2][3][C + 1];
  const int b = threadIdx.x & (W - 1);
  const int xc = (int)(threadIdx.x >> SH);
  const int x = xb + xc;
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1);
  const int tx = __builtin_amdgcn_readfirstlane(xl) / TileShape<2>::TX;
  const int s0 = g.s[0], np1 = g.np[1];
  const T *u1 = p.u.at(t - 1);
  const uint32_t LB = (uint32_t)((g.nodes << SH) * sizeof(T));
  const Buf U0 = make_buf(p.u.at(t - 2), LB);
  const Buf U2 = make_buf(p.u.at(t), LB);
  const T *cy = p.cell, *cx = p.cell + g.nodes, *inv = p.cell + 2 * g.nodes;
  // this thread's element at row ya, and its row stride
  const uint32_t frow = ((uint32_t)s0 << SH) * sizeof(T);
  const uint32_t e0 = ((((uint32_t)ya * s0 + xl) << SH) + b) * sizeof(T);
  const bool has_src = EDGE && chunk_has(p.src, g, ya, yb, tx, tx);
  const bool has_rec = EDGE && chunk_has(p.rec, g, ya, yb, tx, tx);
  const uint32_t src_rows = has_src ? point_rows(p.src, g, ya, yb, tx) : 0u;
  const uint32_t rec_rows = has_rec ? point_rows(p.rec, g, ya, yb, tx) : 0u;
  const int lane0 = ((xc + R) * W + b); // this thread in a staged row

  auto stage_mat = [&](int y) {
    const int64_t r = (int64_t)y * s0;
    stage_nodes(cy, r + xb, C, np1 - xb, lmat[y & 1][0], 1);
    stage_nodes(cx, r + xb - 1, C + 1, np1 - xb + 1, lmat[y & 1][1], 2);
    stage_nodes(inv, r + xb, C, np1 - xb, lmat[y & 1][2], 3);
  };

  // the queue's first rows ya - R .. ya + R - 1, straight from memory
  T q[2 * R + 1];
  {
    const Buf U1 = make_buf(u1, LB);
#pragma unroll
    for (int j = 0; j < 2 * R; ++j)
      q[j] = bload<T>(U1, e0 + (uint32_t)j * frow - R * frow, 0);
  }
  T gm0 = cy[(int64_t)(ya - 1) * s0 + xl];
  for (int y = ya; y <= ya + R; ++y)
    stage_row<T, R, W>(u1, y, xb, s0, np1, ring, 0);
  stage_mat(ya);
  T u0n = bload<T>(U0, e0, 0);
  wait_vmcnt<0>();
  __syncthreads();

  for (int y = ya; y < yb; ++y) {
    // this row's staged values into registers first: once the next row's
    // copies are in flight, every read of LDS would wait for them
    const T *row = (const T *)(ring + (y % S::RING) * S::SEGB) + lane0;
    T tf[2 * R + 1];
#pragma unroll
    for (int k = -R; k <= R; ++k)
      if (k)
        tf[k + R] = row[k * W];
    q[2 * R] = ((const T *)(ring + ((y + R) % S::RING) * S::SEGB))[lane0];
    const T gp0 = lmat[y & 1][0][xc], mi = lmat[y & 1][2][xc];
    const T gp1 = lmat[y & 1][1][xc + 1], gm1 = lmat[y & 1][1][xc];
    __builtin_amdgcn_sched_barrier(0);
    // the next row's loads: its staged rows, then this thread's own values
    stage_row<T, R, W>(u1, y + R + 1, xb, s0, np1, ring, 0);
    stage_mat(y + 1);
    const T u0c = u0n;
    u0n = bload<T>(U0, e0 + (uint32_t)(y + 1 - ya) * frow, 0);
    __builtin_amdgcn_sched_barrier(0);

    tf[R] = q[R];
    T v = 0;
    if (active) {
      T div[2];
      div[0] = flux<T, R>(q, gp0, gm0, g.op);
      div[1] = flux<T, R>(tf, gp1, gm1, g.op);
      const T lap = combine_axes<T, 2, R>(div, g.f);
      v = -u0c + (T)2 * q[R] + mi * lap;
      const int nc = y * s0 + x;
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      if ((src_rows >> (y - ya)) & 1u)
        for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
          if (p.src.takes(k, nc, b))
            v += p.src.weight[k] * p.signal[p.src.at(t, k, b)];
      if ((rec_rows >> (y - ya)) & 1u)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b))
            p.record[p.rec.at(t, k, b)] = v;
    }
    // every thread stores (off the box, out of range: dropped), so the
    // count of operations after the staging copies is fixed
    bstore(U2, v, active ? e0 + (uint32_t)(y - ya) * frow : kDrop, 0);
    if (ILLUM && p.illum) {
      // a column's lanes sum u^2 (off the box v is zero); one writer a node
      const T s2 = shot_sum(v * v, W);
      if (active && b == 0)
        atomicAdd(p.illum + (int64_t)y * s0 + x, s2);
    }
    gm0 = gp0;
#pragma unroll
    for (int j = 0; j < 2 * R; ++j)
      q[j] = q[j + 1];
    // the next row's loads done (the store may still run), then every
    // wave's copies
    wait_vmcnt<1>();
    __syncthreads();
  }
}

// --------------------------------------------------------------- adjoint
template <typename T, int R, int W, bool STIFF, bool EDGE>
__device__ __forceinline__ void backward_stream(const BwdParams<T> &p,
                                                const Box &box, const int n,
                                                int bid, int nb) {
  using S = Stage<T, R, W>;
  constexpr int SH = S::SH, C = S::C;
  const Geom<T> &g = p.g;
  const bool reverse = p.reverse && n <= p.steps - 2;
  int ya, yb, xb;
  if (!box_item<EDGE>(box, g, p.src, reverse, p.rec, C, bid, nb, ya, yb, xb))
    return;
  // the staged rows of lambda^(n+1) and u^n; per material row slot the
  // fast-axis cells of x - R .. [C + 2R], 1 / m [C] and the slow-axis cell
  // of row y + R - 1 [C]
  __shared__ __align__(16) char lring[S::RING * S::SEGB];
  __shared__ __align__(16) char uring[S::RING * S::SEGB];
  __shared__ __align__(16) T lmat[2][3][C + 2 * R];
  const int b = threadIdx.x & (W - 1);
  const int xc = (int)(threadIdx.x >> SH);
  const int x = xb + xc;
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1);
  const int tx = __builtin_amdgcn_readfirstlane(xl) / TileShape<2>::TX;
  const int s0 = g.s[0], np1 = g.np[1];
  const int m = p.steps - 1 - n;
  const bool chain = W == 1 && !p.combine;
  const T *l1 = p.lam.at(n + 1), *u1 = p.u.at(n);
  const uint32_t LB = (uint32_t)((g.nodes << SH) * sizeof(T));
  const uint32_t NB = (uint32_t)(g.nodes * sizeof(T));
  const Buf L0b = make_buf(p.lam.at(n), LB);
  // without the reverse march u^(n + 1) is neither read nor written: an
  // empty resource turns its loads into zeros that touch no memory
  const Buf UOb = make_buf(p.u.at(n + 1), reverse ? LB : 0u);
  const Buf GM = make_buf(p.gmass, NB);
  const Buf GS = make_buf(p.gstiff, p.combine ? 0u : NB);
  const Buf SK = make_buf(p.stiff, STIFF ? NB : 0u);
  const T *cy = p.cell, *cx = p.cell + g.nodes, *inv = p.cell + 2 * g.nodes;
  const uint32_t frow = ((uint32_t)s0 << SH) * sizeof(T);
  const uint32_t e0 = ((((uint32_t)ya * s0 + xl) << SH) + b) * sizeof(T);
  const uint32_t n0 = ((uint32_t)ya * s0 + xl) * sizeof(T);
  const uint32_t nrow = (uint32_t)s0 * sizeof(T);
  const bool has_rec = EDGE && chunk_has(p.rec, g, ya, yb, tx, tx);
  const bool has_src = EDGE && reverse && chunk_has(p.src, g, ya, yb, tx, tx);
  const uint32_t rec_rows = has_rec ? point_rows(p.rec, g, ya, yb, tx) : 0u;
  const uint32_t src_rows = has_src ? point_rows(p.src, g, ya, yb, tx) : 0u;
  const int lane0 = ((xc + R) * W + b);

  auto stage_mat = [&](int y) {
    const int64_t r = (int64_t)y * s0;
    T(*s)[C + 2 * R] = lmat[y & 1];
    stage_nodes(cx, r + xb - R, C + 2 * R, np1 - xb + R, s[0], 2);
    stage_nodes(inv, r + xb, C, np1 - xb, s[1], 3);
    stage_nodes(cy, r + (int64_t)(R - 1) * s0 + xb, C, np1 - xb, s[2], 0);
  };
  // the thread's own values of a row, one row ahead
  struct Own {
    T lo, uo, accm, accs, sk[5];
  };
  auto own = [&](Own &o, int y) {
    const uint32_t e = e0 + (uint32_t)(y - ya) * frow;
    const uint32_t nn = n0 + (uint32_t)(y - ya) * nrow;
    o.lo = bload<T>(L0b, e, 0);
    o.uo = bload<T>(UOb, e, 0);
    o.accm = bload<T>(GM, nn, 0);
    o.accs = bload<T>(GS, nn, 0);
    if (STIFF) {
      o.sk[0] = bload<T>(SK, nn, 0);
      o.sk[1] = bload<T>(SK, nn + nrow, 0);
      o.sk[2] = bload<T>(SK, nn - nrow, 0);
      o.sk[3] = bload<T>(SK, nn + (uint32_t)sizeof(T), 0);
      o.sk[4] = bload<T>(SK, nn - (uint32_t)sizeof(T), 0);
    }
  };

  // queues along the slow axis: lambda^(n + 1) and u^n rows y - R .. y + R,
  // the cell means of rows y - R .. y + R - 1; the first rows from memory
  T ql[2 * R + 1], qu[2 * R + 1], qc[2 * R];
  {
    const Buf L1b = make_buf(l1, LB), U1b = make_buf(u1, LB);
#pragma unroll
    for (int j = 0; j < 2 * R; ++j) {
      ql[j] = bload<T>(L1b, e0 + (uint32_t)j * frow - R * frow, 0);
      qu[j] = bload<T>(U1b, e0 + (uint32_t)j * frow - R * frow, 0);
    }
#pragma unroll
    for (int j = 0; j < 2 * R - 1; ++j)
      qc[j] = cy[(int64_t)(ya - R + j) * s0 + xl];
  }
  for (int y = ya; y <= ya + R; ++y) {
    stage_row<T, R, W>(l1, y, xb, s0, np1, lring, 0);
    stage_row<T, R, W>(u1, y, xb, s0, np1, uring, 2);
  }
  stage_mat(ya);
  Own on;
  own(on, ya);
  wait_vmcnt<0>();
  __syncthreads();

  for (int y = ya; y < yb; ++y) {
    // this row's staged values into registers first: once the next row's
    // copies are in flight, every read of LDS would wait for them
    const T(*mt)[C + 2 * R] = lmat[y & 1];
    const int ls = (y % S::RING) * S::SEGB, fs = ((y + R) % S::RING) * S::SEGB;
    ql[2 * R] = ((const T *)(lring + fs))[lane0];
    qu[2 * R] = ((const T *)(uring + fs))[lane0];
    qc[2 * R - 1] = mt[2][xc];
    T lt1[2 * R + 1], ut1[2 * R + 1], chi1[R], clo1[R];
    {
      const T *lrow = (const T *)(lring + ls) + lane0;
      const T *urow = (const T *)(uring + ls) + lane0;
#pragma unroll
      for (int k = -R; k <= R; ++k)
        if (k) {
          lt1[k + R] = lrow[k * W];
          ut1[k + R] = urow[k * W];
        }
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        chi1[j - 1] = mt[0][xc + R + j - 1];
        clo1[j - 1] = mt[0][xc + R - j];
      }
    }
    const T mi = mt[1][xc];
    __builtin_amdgcn_sched_barrier(0);
    // the next row's loads: its staged rows, then this thread's own values
    stage_row<T, R, W>(l1, y + R + 1, xb, s0, np1, lring, 0);
    stage_row<T, R, W>(u1, y + R + 1, xb, s0, np1, uring, 2);
    stage_mat(y + 1);
    const Own o = on;
    own(on, y + 1);
    __builtin_amdgcn_sched_barrier(0);

    lt1[R] = ql[R];
    ut1[R] = qu[R];
    const uint32_t e = e0 + (uint32_t)(y - ya) * frow;
    const uint32_t nn = n0 + (uint32_t)(y - ya) * nrow;
    T ln = 0, v = 0, gm_new = 0, gs_new = 0;
    bool gstore = false;
    if (active) {
      const int nc = y * s0 + x;
      T chi0[R], clo0[R];
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        chi0[j - 1] = qc[R + j - 1];
        clo0[j - 1] = qc[R - j];
      }
      const T lo_ = o.lo;
      const T accm = o.accm, accs = o.accs;
      const T lc = ql[R], uc = qu[R];
      T div_l = 0;
      div_l += g.f[0] * transposed<T, R>(ql, chi0, clo0, g.op);
      div_l += g.f[1] * transposed<T, R>(lt1, chi1, clo1, g.op);
      ln = -lo_ + (T)2 * lc + mi * div_l;
      // the inertia density by parts in time (scalar_sensitivity.cu:512-514)
      T cm = chain ? accm : (T)0;
      cm = cm - p.mf * uc * (ln - (T)2 * lc + lo_);
      T cs = chain ? accs : (T)0;
      if (STIFF) {
        T gsum = 0;
        gsum += stiffness_density<T, R>(qu, ql, lc, o.sk[0], o.sk[1], o.sk[2],
                                        g.f[0], g.op);
        gsum += stiffness_density<T, R>(ut1, lt1, lc, o.sk[0], o.sk[3],
                                        o.sk[4], g.f[1], g.op);
        cs = cs - p.mf * gsum;
      }
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      if ((rec_rows >> (y - ya)) & 1u)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b)) {
            const T load = p.rec.weight[k] * p.adjoint[p.rec.at(m, k, b)];
            ln += load;
            cm += -p.mf * uc * load;
          }
      // the gradient as add_gradient forms it, stored below
      if (p.combine) {
        gm_new = accm + shot_sum(p.jm * cm + p.js * cs, W);
        gstore = b == 0;
      } else if (W == 1) {
        gm_new = cm;
        gs_new = cs;
        gstore = true;
      } else {
        gm_new = accm + shot_sum(cm, W);
        gs_new = accs + shot_sum(cs, W);
        gstore = b == 0;
      }
      if (reverse) {
        T div[2];
        div[0] = flux<T, R>(qu, chi0[0], clo0[0], g.op);
        div[1] = flux<T, R>(ut1, chi1[0], clo1[0], g.op);
        const T lap = combine_axes<T, 2, R>(div, g.f);
        v = -o.uo + (T)2 * uc + mi * lap;
        if ((src_rows >> (y - ya)) & 1u)
          for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
            if (p.src.takes(k, nc, b))
              v += p.src.weight[k] * p.signal[p.src.at(n + 1, k, b)];
      }
    }
    // four stores every thread issues (off the box or without a value, out
    // of range: dropped; an empty resource drops them all), so the count of
    // operations after the staging copies is fixed
    bstore(L0b, ln, active ? e : kDrop, 0);
    bstore(GM, gm_new, gstore ? nn : kDrop, 0);
    bstore(GS, gs_new, gstore ? nn : kDrop, 0);
    bstore(UOb, v, active ? e : kDrop, 0);
#pragma unroll
    for (int j = 0; j < 2 * R; ++j) {
      ql[j] = ql[j + 1];
      qu[j] = qu[j + 1];
    }
#pragma unroll
    for (int j = 0; j < 2 * R - 1; ++j)
      qc[j] = qc[j + 1];
    // the next row's loads done (the four stores may still run), then
    // every wave's copies
    wait_vmcnt<4>();
    __syncthreads();
  }
}

// ------------------------------------------------------------ the steps
// The edge launch of a step: the first `streams` blocks (a multiple of the
// XCD count) march the stream box's blocks that hold sources or sensors,
// the rest take the general nodes (the 2D frame around the box, or every
// node in 1D and 3D), V shot lanes a thread. W is the shot lanes of a node
// for the stream blocks (0 where there are none).
// the step a launch makes: `t`, offset by the device clock where one is given
// (-1 outside 0 .. t_end - 1: the launch does nothing)
template <typename T>
__device__ __forceinline__ int clocked_step(const FwdParams<T> &p, int t) {
  if (!p.clock)
    return t;
  const int tc = t + *p.clock;
  return tc >= 0 && tc < p.t_end ? tc : -1;
}

template <typename T, int NDIM, int R, int W, int V>
__global__ __launch_bounds__(kThreads) void forward_step(const FwdParams<T> p,
                                                         const Box box,
                                                         const int step,
                                                         const int streams) {
  const int t = clocked_step(p, step);
  if (t < 0)
    return;
  if constexpr (NDIM == 2 && W > 0) {
    if ((int)blockIdx.x < streams) {
      forward_stream<T, R, W, true, true>(p, box, t, blockIdx.x, streams);
      return;
    }
  }
  forward_general<T, NDIM, R, V>(p, t, blockIdx.x - streams,
                                 gridDim.x - streams);
}

// a step of a grid under a domain mask: every interior node by the general
// kernels, each reading its cell code (see forward_general)
template <typename T, int NDIM, int R, int V>
__global__ __launch_bounds__(kThreads) void forward_masked(const FwdParams<T> p,
                                                           const int step) {
  const int t = clocked_step(p, step);
  if (t < 0)
    return;
  forward_general<T, NDIM, R, V, true>(p, t, blockIdx.x, gridDim.x);
}

// The one-lane float 3D sweep is asked for four blocks a CU: its registers
// then just fit four waves a SIMD (the default packing left it one register
// over, at three), which measured 18 % faster; every other build keeps the
// default, as a tighter budget would make it spill.
#ifndef CW_FRAME_BLOCKS_2D
#define CW_FRAME_BLOCKS_2D 1
#endif
template <typename T, int NDIM, int R, int W, int V, bool STIFF>
__global__ __launch_bounds__(kThreads,
                             NDIM == 3 && V == 1 && sizeof(T) == 4 ? 4
                             : NDIM == 2 && sizeof(T) == 4 ? CW_FRAME_BLOCKS_2D : 1) void backward_step(
    const BwdParams<T> p,
                                                          const Box box,
                                                          const int n,
                                                          const int streams) {
  if constexpr (NDIM == 2 && W > 0) {
    if ((int)blockIdx.x < streams) {
      backward_stream<T, R, W, STIFF, true>(p, box, n, blockIdx.x, streams);
      return;
    }
  }
  backward_general<T, NDIM, R, V>(p, n, blockIdx.x - streams,
                                  gridDim.x - streams);
}

// the core blocks of a 2D step: no layer, no wall, no point, no strip;
// ILLUM builds in the illumination sum (a separate build, so a step that
// does not gather it compiles as if it did not exist)
template <typename T, int R, int W, bool ILLUM>
__global__ __launch_bounds__(kThreads) void forward_core(const FwdParams<T> p,
                                                         const Box box,
                                                         const int step) {
  const int t = clocked_step(p, step);
  if (t < 0)
    return;
  forward_stream<T, R, W, false, ILLUM>(p, box, t, blockIdx.x, gridDim.x);
}

template <typename T, int R, int W, bool STIFF>
__global__ __launch_bounds__(kThreads) void backward_core(const BwdParams<T> p,
                                                          const Box box,
                                                          const int n) {
  backward_stream<T, R, W, STIFF, false>(p, box, n, blockIdx.x, gridDim.x);
}

} // namespace cw
