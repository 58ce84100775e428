// Streaming kernels for the bulk of a 2D grid: every node at least
// order / 2 + 2 nodes off the walls, the PML included. The general kernels
// (scalar_kernels.hpp) take the thin band that is left, where the wall
// mirror, the ghost refill and the held nodes live.
//
// A thread is one shot lane of one column and marches a chunk of rows down
// the slow axis, keeping the 2R + 1 values of its column that the slow-axis
// taps read in a register queue: each row costs one new load per queued
// field. The fast-axis taps are the same lane of the neighbouring columns,
// each a contiguous run of the wave (common.hpp), served by L1 and L2; the
// next row's loads are all issued before the current row is computed. A
// block is 256 / lanes neighbouring columns, and blocks are dealt to the
// XCDs in contiguous runs, rows of blocks across the box, so the columns
// whose taps overlap share one L2.
//
// A row of the slow axis is either in that axis' slab or not, so its layer
// terms are uniform over the block and its profile values scalar loads; the
// fast axis' slab columns are fixed per thread for the whole chunk.
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

// this block's rows [ya, yb) and first column xb; blocks start at multiples
// of their width, so no wave straddles a point tile
__device__ __forceinline__ void box_item(const Box &box, int cols, int &ya,
                                         int &yb, int &xb) {
  const int xa = box.x0 / cols * cols;
  const int ncx = (box.x1 - xa + cols - 1) / cols;
  const int nchunk = (box.y1 - box.y0 + box.chunk - 1) / box.chunk;
  const int w = xcd_tile(blockIdx.x, nchunk * ncx);
  ya = box.y0 + (w / ncx) * box.chunk;
  yb = min(ya + box.chunk, box.y1);
  xb = xa + (w % ncx) * cols;
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

// the fast-axis taps k = -R .. -1, 1 .. R of element e into x[0 .. 2R)
template <typename T, int R>
__device__ __forceinline__ void load_row_taps(T *x, const T *__restrict__ u,
                                              uint32_t e, uint32_t el) {
#pragma unroll
  for (int k = 1; k <= R; ++k) {
    x[R - k] = u[e - k * el];
    x[R - 1 + k] = u[e + k * el];
  }
}

// taps[0 .. 2R] from the fast-axis taps and the centre
template <typename T, int R>
__device__ __forceinline__ void row_line(T *t, const T *x, T c) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    t[k] = x[k];
    t[R + 1 + k] = x[R + k];
  }
  t[R] = c;
}

// --------------------------------------------------------------- forward
template <typename T, int R>
__global__ __launch_bounds__(kThreads) void forward_stream(
    const FwdParams<T> p, const Box box, const int t) {
  const Geom<T> &g = p.g;
  const int sh = g.lshift;
  int ya, yb, xb;
  box_item(box, kThreads >> sh, ya, yb, xb);
  const int b = threadIdx.x & (g.lanes - 1);
  const int x = xb + (int)(threadIdx.x >> sh);
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1); // every thread loads
  const int s0 = g.s[0];
  const uint32_t es0 = (uint32_t)s0 << sh, el = 1u << sh;
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = ((int64_t)(L0 + H0) * g.np[1]) << sh;
  const int64_t size1 = ((int64_t)(L1 + H1) * g.np[0]) << sh;
  const int prow1 = L1 + H1 ? slab_row(xl, N1, L1, H1) : -1;
  Prof<T> pf1{};
  if (prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, xl);
  const int tx = xl / TileShape<2>::TX;
  const T *__restrict__ u1 = p.u.at(t - 1);
  const T *__restrict__ u0 = p.u.at(t - 2);
  T *__restrict__ u2 = p.u.at(t);
  const T *__restrict__ cx = p.cell + g.nodes; // cells along the fast axis
  const T *__restrict__ cyc = p.cell;          // cells along the slow axis
  uint32_t node = (uint32_t)(ya * s0 + xl);
  uint32_t e = (node << sh) + b;

  T q[2 * R + 1];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j)
    q[j] = u1[e + (j - R) * es0];
  T gm0 = cyc[node - s0];
  // the next row's loads, issued a row ahead
  T nq = u1[e + R * es0], nx[2 * R];
  load_row_taps<T, R>(nx, u1, e, el);
  T nu0 = u0[e], nmi = p.inv[node];
  T ngp1 = cx[node], ngm1 = cx[node - 1], ngp0 = cyc[node];

  for (int y = ya; y < yb; ++y) {
    q[2 * R] = nq;
    T tf[2 * R + 1];
    row_line<T, R>(tf, nx, q[R]);
    const T u0c = nu0, mi = nmi, gp1 = ngp1, gm1 = ngm1, gp0 = ngp0;
    const uint32_t ec = e, nc = node;
    e += es0;
    node += s0;
    if (y + 1 < yb) {
      nq = u1[e + R * es0];
      load_row_taps<T, R>(nx, u1, e, el);
      nu0 = u0[e];
      nmi = p.inv[node];
      ngp1 = cx[node];
      ngm1 = cx[node - 1];
      ngp0 = cyc[node];
    }
    if (active) {
      T div[2];
      div[0] = flux<T, R>(q, gp0, gm0, g.op);
      div[1] = flux<T, R>(tf, gp1, gm1, g.op);
      const T lap = combine_axes<T, 2, R>(div, g.f);
      T v = -u0c + (T)2 * q[R] + mi * lap;
      const int prow0 = L0 + H0 ? slab_row(y, N0, L0, H0) : -1;
      if (prow0 >= 0) {
        const Prof<T> pf = profile_at(p.pml.profile[0], N0, y);
        T Dp, Dm;
        cell_diffs<T, R>(q, g.op, Dp, Dm);
        v += pml_forward(p.pml.aux[0] +
                             (((int64_t)prow0 * g.np[1] + x) << sh) + b,
                         size0, Dp, Dm, gp0, gm0, g.f[0], mi, pf.an, pf.bn,
                         pf.ah, pf.bh, pf.ah1, pf.bh1);
      }
      if (prow1 >= 0) {
        T Dp, Dm;
        cell_diffs<T, R>(tf, g.op, Dp, Dm);
        v += pml_forward(p.pml.aux[1] +
                             (((int64_t)y * (L1 + H1) + prow1) << sh) + b,
                         size1, Dp, Dm, gp1, gm1, g.f[1], mi, pf1.an, pf1.bn,
                         pf1.ah, pf1.bh, pf1.ah1, pf1.bh1);
      }
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      if (p.src.starts)
        for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
          if (p.src.takes(k, nc, b))
            v += p.src.weight[k] * p.signal[p.src.at(t, k, b)];
      u2[ec] = v;
      if (p.rec.starts)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b))
            p.record[p.rec.at(t, k, b)] = v;
      if ((prow0 >= 0 || prow1 >= 0) && p.strip.map) {
        const int col = p.strip.map[nc];
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
    gm0 = gp0;
#pragma unroll
    for (int j = 0; j < 2 * R; ++j)
      q[j] = q[j + 1];
  }
}

// --------------------------------------------------------------- adjoint
template <typename T, int R, bool STIFF>
__global__ __launch_bounds__(kThreads) void backward_stream(
    const BwdParams<T> p, const Box box, const int n) {
  const Geom<T> &g = p.g;
  const int sh = g.lshift;
  int ya, yb, xb;
  box_item(box, kThreads >> sh, ya, yb, xb);
  const int b = threadIdx.x & (g.lanes - 1);
  const int x = xb + (int)(threadIdx.x >> sh);
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(max(x, box.x0), box.x1 - 1);
  const int s0 = g.s[0];
  const uint32_t es0 = (uint32_t)s0 << sh, el = 1u << sh;
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
  const int tx = xl / TileShape<2>::TX;
  const T *__restrict__ l1 = p.lam.at(n + 1);
  const T *__restrict__ u1 = p.u.at(n);
  T *__restrict__ l0 = p.lam.at(n);
  T *__restrict__ uo = p.u.at(n + 1);
  const T *__restrict__ cx = p.cell + g.nodes;
  const T *__restrict__ cyc = p.cell;
  uint32_t node = (uint32_t)(ya * s0 + xl);
  uint32_t e = (node << sh) + b;

  // queues along the slow axis: lambda^(n + 1) and u^n rows y - R .. y + R,
  // the cell means of rows y - R .. y + R - 1
  T ql[2 * R + 1], qu[2 * R + 1], qc[2 * R];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j) {
    ql[j] = l1[e + (j - R) * es0];
    qu[j] = u1[e + (j - R) * es0];
    if (j < 2 * R - 1)
      qc[j] = cyc[node + (j - R) * s0];
  }
  // the next row's loads: the queue feeds, the fast-axis taps of lambda, u
  // and the fast-axis cell means (x - R .. x + R - 1), the slots overwritten
  T nl, nu, nlx[2 * R], nux[2 * R], ncx[2 * R], nlo, nuo, nmi, ncy;
  auto fetch = [&]() {
    nl = l1[e + R * es0];
    nu = u1[e + R * es0];
    load_row_taps<T, R>(nlx, l1, e, el);
    load_row_taps<T, R>(nux, u1, e, el);
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
      ncx[k] = cx[node + k - R];
    nlo = l0[e];
    nuo = reverse ? uo[e] : (T)0;
    nmi = p.inv[node];
    ncy = cyc[node + (R - 1) * s0];
  };
  fetch();

  for (int y = ya; y < yb; ++y) {
    ql[2 * R] = nl;
    qu[2 * R] = nu;
    qc[2 * R - 1] = ncy;
    T lt1[2 * R + 1], ut1[2 * R + 1], chi1[R], clo1[R];
    row_line<T, R>(lt1, nlx, ql[R]);
    row_line<T, R>(ut1, nux, qu[R]);
#pragma unroll
    for (int j = 1; j <= R; ++j) {
      chi1[j - 1] = ncx[R + j - 1];
      clo1[j - 1] = ncx[R - j];
    }
    const T lo_ = nlo, uo_ = nuo, mi = nmi;
    const uint32_t ec = e, nc = node;
    e += es0;
    node += s0;
    if (y + 1 < yb)
      fetch();
    if (active) {
      // this row along the slow axis: slab row, reach of the transposed taps
      const int prow0 = L0 + H0 ? slab_row(y, N0, L0, H0) : -1;
      const bool ext0 = (L0 && y - 1 < L0 + R) || (H0 && N0 - 2 - y < H0 + R);
      const bool layer = prow0 >= 0 || prow1 >= 0;
      const bool grad = !(p.skip_layer && layer);
      T chi0[R], clo0[R];
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        chi0[j - 1] = qc[R + j - 1];
        clo0[j - 1] = qc[R - j];
      }
      T accm = 0, accs = 0;
      if (grad) {
        accm = p.gmass[nc];
        if (!p.combine)
          accs = p.gstiff[nc];
      }
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
        const T sc = p.stiff[nc];
        T gsum = 0;
        gsum += stiffness_density<T, R>(qu, ql, lc, sc, p.stiff[nc + s0],
                                        p.stiff[nc - s0], g.f[0], g.op);
        gsum += stiffness_density<T, R>(ut1, lt1, lc, sc, p.stiff[nc + 1],
                                        p.stiff[nc - 1], g.f[1], g.op);
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
      if (p.rec.starts)
        for (int k = p.rec.starts[tile]; k < p.rec.starts[tile + 1]; ++k)
          if (p.rec.takes(k, nc, b)) {
            const T load = p.rec.weight[k] * p.adjoint[p.rec.at(m, k, b)];
            ln += load;
            cm += -p.mf * uc * load;
          }
      l0[ec] = ln;
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
        int strip_col = -1;
        if (layer && p.strip.map)
          strip_col = p.strip.map[nc];
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
          v = -uo_ + (T)2 * uc + mi * lap;
          if (p.src.starts)
            for (int k = p.src.starts[tile]; k < p.src.starts[tile + 1]; ++k)
              if (p.src.takes(k, nc, b))
                v += p.src.weight[k] * p.signal[p.src.at(n + 1, k, b)];
        }
        uo[ec] = v;
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
  }
}

} // namespace cw
