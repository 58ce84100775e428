// Streaming kernels for the bulk of a 2D grid: every node at least
// order / 2 + 2 nodes off the walls, the PML included. The general kernels
// (scalar_kernels.hpp) take the thin band that is left, where the wall
// mirror, the ghost refill and the held nodes live.
//
// One wave (64 lanes) owns a strip of 64 columns aligned with the step
// tiles and marches a chunk of rows down the slow axis. Each lane keeps the
// 2R + 1 values of its own column that the slow-axis taps read in a register
// queue, and the row being updated goes through LDS with an R-column halo on
// each side for the fast-axis taps, so every field value is loaded from
// memory once per chunk. The next row's loads are issued before the current
// row is computed. Work items (chunk, strip, shot group) are dealt to the
// XCDs in contiguous runs with the shots innermost, so the shots of one strip
// read the same material lines out of one L2 and vertically adjacent chunks,
// which share their halo rows, run side by side.
//
// A row of the slow axis is either in that axis' slab or not, so its layer
// terms are uniform over the wave and its profile values scalar loads; the
// fast axis' slab columns are fixed per lane for the whole chunk.
//
// The arithmetic is the general kernels' term by term, in the same order,
// so both paths give the same bits; the cell weight of every node here is 1.
#pragma once

#include "scalar_kernels.hpp"

namespace cw {

constexpr int kLanes = 64;

// the box of a sweep: rows [y0, y1), columns [x0, x1), and the tile columns
// [tx0, tx1) its strips follow
struct Box {
  int y0, y1, x0, x1;
  int tx0, tx1;
  int chunk; // rows a wave marches
};

__device__ __forceinline__ void box_item(const Box &box, int nbatch, int group,
                                         int &cy, int &tx, int &gi) {
  const int nstrip = box.tx1 - box.tx0;
  const int nchunk = (box.y1 - box.y0 + box.chunk - 1) / box.chunk;
  const int ngroup = (nbatch + group - 1) / group;
  const int w = xcd_tile(blockIdx.x, nchunk * nstrip * ngroup);
  gi = w % ngroup;
  const int rest = w / ngroup;
  tx = box.tx0 + rest % nstrip;
  cy = rest / nstrip;
}

// the entries of tile `tile` a node takes: sources add
template <typename T>
__device__ __forceinline__ T add_sources(const Points<T> &src,
                                         const T *__restrict__ signal,
                                         int tile, int64_t idx, int b, int row,
                                         T v) {
  if (!src.starts)
    return v;
  const int e0 = src.starts[tile], e1 = src.starts[tile + 1];
  for (int e = e0; e < e1; ++e)
    if (src.lin[e] == idx && (src.shot[e] < 0 || src.shot[e] == b))
      v += src.weight[e] * signal[b * src.shot_stride +
                                  (int64_t)row * src.columns + src.col[e]];
  return v;
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

// the transposed layer of one axis at a node: the tap fields of the slab
// nodes within R along the axis, `at(k)` the memory index of the node k
// away or -1 off the slab (the general kernels' sum, same order)
template <typename T, int R, typename At>
__device__ __forceinline__ T gather(const T *__restrict__ eta, int64_t size,
                                    const T *op, At at) {
  T sum = 0;
#pragma unroll
  for (int k = -R; k <= R; ++k) {
    const int64_t i = at(k);
    if (i >= 0)
      sum += eta[3 * size + i] * plus_tap<T, R>(-k, op) +
             eta[4 * size + i] * minus_tap<T, R>(-k, op);
  }
  return sum;
}

// --------------------------------------------------------------- forward
template <typename T, int R>
__global__ __launch_bounds__(kLanes) void forward_stream(
    const FwdParams<T> p, const Box box, const int t) {
  const Geom<T> &g = p.g;
  int cy, tx, b;
  box_item(box, g.nbatch, 1, cy, tx, b);
  const int lane = threadIdx.x;
  const int x = tx * kLanes + lane;
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(x, g.np[1] - 1); // every lane loads its column
  const int ya = box.y0 + cy * box.chunk;
  const int yb = min(ya + box.chunk, box.y1);
  const int s0 = g.s[0];
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = (int64_t)(L0 + H0) * g.np[1];
  const int64_t size1 = (int64_t)(L1 + H1) * g.np[0];
  const int prow1 = L1 + H1 ? slab_row(x, N1, L1, H1) : -1;
  Prof<T> pf1{};
  if (active && prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, x);
  T *const aux0 = L0 + H0 ? p.pml.aux[0] + (int64_t)b * 3 * size0 : nullptr;
  T *const aux1 = L1 + H1 ? p.pml.aux[1] + (int64_t)b * 3 * size1 : nullptr;
  const int64_t ob = (int64_t)b * p.u.batch;
  const T *__restrict__ u1 = p.u.at(t - 1) + ob;
  const T *__restrict__ u0 = p.u.at(t - 2) + ob;
  T *__restrict__ u2 = p.u.at(t) + ob;
  const T *__restrict__ cx = p.cell + g.field; // cells along the fast axis
  const T *__restrict__ cyc = p.cell;          // cells along the slow axis
  // halo lanes: the first R load the columns left of the strip, the last R
  // those right of it
  const int hx = lane < R ? max(tx * kLanes - R + lane, 0)
                 : lane >= kLanes - R
                     ? min(tx * kLanes + R + lane, g.np[1] - 1)
                     : -1;
  const int hslot = lane < R ? lane : lane >= kLanes - R ? lane + 2 * R : -1;
  __shared__ T rows[2][kLanes + 2 * R];

  T q[2 * R + 1];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j)
    q[j] = u1[(int64_t)(ya - R + j) * s0 + xl];
  T gm0 = cyc[(int64_t)(ya - 1) * s0 + xl];
  // the next row's loads, issued a row ahead
  T nq = u1[(int64_t)(ya + R) * s0 + xl];
  T nh = hx >= 0 ? u1[(int64_t)ya * s0 + hx] : (T)0;
  T nu0 = u0[(int64_t)ya * s0 + xl];
  T nmi = p.inv[(int64_t)ya * s0 + xl];
  T ngp1 = cx[(int64_t)ya * s0 + xl], ngm1 = cx[(int64_t)ya * s0 + xl - 1];
  T ngp0 = cyc[(int64_t)ya * s0 + xl];

  for (int y = ya; y < yb; ++y) {
    q[2 * R] = nq;
    const T h = nh, u0c = nu0, mi = nmi, gp1 = ngp1, gm1 = ngm1, gp0 = ngp0;
    if (y + 1 < yb) {
      const int64_t r1 = (int64_t)(y + 1) * s0;
      nq = u1[r1 + (int64_t)R * s0 + xl];
      nh = hx >= 0 ? u1[r1 + hx] : (T)0;
      nu0 = u0[r1 + xl];
      nmi = p.inv[r1 + xl];
      ngp1 = cx[r1 + xl];
      ngm1 = cx[r1 + xl - 1];
      ngp0 = cyc[r1 + xl];
    }
    T *L = rows[y & 1];
    L[R + lane] = q[R];
    if (hslot >= 0)
      L[hslot] = h;
    __syncthreads();
    if (active) {
      T tf[2 * R + 1];
#pragma unroll
      for (int j = 0; j <= 2 * R; ++j)
        tf[j] = L[lane + j];
      T div[2];
      div[0] = flux<T, R>(q, gp0, gm0, g.op);
      div[1] = flux<T, R>(tf, gp1, gm1, g.op);
      const T lap = combine_axes<T, 2, R>(div, g.f);
      const int64_t idx = (int64_t)y * s0 + x;
      T v = -u0c + (T)2 * q[R] + mi * lap;
      const int prow0 = L0 + H0 ? slab_row(y, N0, L0, H0) : -1;
      if (prow0 >= 0) {
        const Prof<T> pf = profile_at(p.pml.profile[0], N0, y);
        T Dp, Dm;
        cell_diffs<T, R>(q, g.op, Dp, Dm);
        v += pml_forward(aux0 + (int64_t)prow0 * g.np[1] + x, size0, Dp, Dm,
                         gp0, gm0, g.f[0], mi, pf.an, pf.bn, pf.ah, pf.bh,
                         pf.ah1, pf.bh1);
      }
      if (prow1 >= 0) {
        T Dp, Dm;
        cell_diffs<T, R>(tf, g.op, Dp, Dm);
        v += pml_forward(aux1 + (int64_t)y * (L1 + H1) + prow1, size1, Dp, Dm,
                         gp1, gm1, g.f[1], mi, pf1.an, pf1.bn, pf1.ah, pf1.bh,
                         pf1.ah1, pf1.bh1);
      }
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      v = add_sources(p.src, p.signal, tile, idx, b, t, v);
      u2[idx] = v;
      if (p.rec.starts) {
        const int e0 = p.rec.starts[tile], e1 = p.rec.starts[tile + 1];
        for (int e = e0; e < e1; ++e)
          if (p.rec.lin[e] == idx)
            p.record[(int64_t)b * p.steps * p.rec.columns +
                     (int64_t)t * p.rec.columns + p.rec.col[e]] = v;
      }
      if ((prow0 >= 0 || prow1 >= 0) && p.strip.map) {
        const int col = p.strip.map[idx];
        if (col >= 0) {
          const int64_t at =
              ((int64_t)b * p.strip.rows + (t - p.strip.first)) *
                  p.strip.count +
              col;
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
// S shots a wave share the material loads and one gradient accumulator
// (group gi of the caller's ceil(B / S), or the shot's own in per-shot mode)
template <typename T, int R, int S, bool STIFF>
__global__ __launch_bounds__(kLanes) void backward_stream(
    const BwdParams<T> p, const Box box, const int n) {
  const Geom<T> &g = p.g;
  int cy, tx, gi;
  box_item(box, g.nbatch, S, cy, tx, gi);
  const int b0 = gi * S;
  const int lane = threadIdx.x;
  const int x = tx * kLanes + lane;
  const bool active = x >= box.x0 && x < box.x1;
  const int xl = min(x, g.np[1] - 1);
  const int ya = box.y0 + cy * box.chunk;
  const int yb = min(ya + box.chunk, box.y1);
  const int s0 = g.s[0];
  const int N0 = g.n[0], N1 = g.n[1];
  const int L0 = p.pml.lo[0], H0 = p.pml.hi[0];
  const int L1 = p.pml.lo[1], H1 = p.pml.hi[1];
  const int64_t size0 = (int64_t)(L0 + H0) * g.np[1];
  const int64_t size1 = (int64_t)(L1 + H1) * g.np[0];
  const int m = p.steps - 1 - n;
  const bool reverse = p.reverse && n <= p.steps - 2;
  // the lane's column: its slab row along the fast axis, whether the
  // transposed taps of that axis reach the slab, its profile
  const int prow1 = L1 + H1 ? slab_row(x, N1, L1, H1) : -1;
  const bool ext1 = (L1 && x - 1 < L1 + R) || (H1 && N1 - 2 - x < H1 + R);
  Prof<T> pf1{};
  if (active && prow1 >= 0)
    pf1 = profile_at(p.pml.profile[1], N1, x);
  const T *__restrict__ cx = p.cell + g.field;
  const T *__restrict__ cyc = p.cell;
  const int hx = lane < R ? max(tx * kLanes - R + lane, 0)
                 : lane >= kLanes - R
                     ? min(tx * kLanes + R + lane, g.np[1] - 1)
                     : -1;
  const int hslot = lane < R ? lane : lane >= kLanes - R ? lane + 2 * R : -1;
  __shared__ T rl[S][2][kLanes + 2 * R];
  __shared__ T ru[S][2][kLanes + 2 * R];
  __shared__ T rg[2][kLanes + 2 * R];

  const T *l1s[S], *u1s[S];
  T *l0s[S], *uos[S];
  bool on[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int b = min(b0 + s, g.nbatch - 1);
    on[s] = b0 + s < g.nbatch;
    l1s[s] = p.lam.at(n + 1) + (int64_t)b * p.lam.batch;
    l0s[s] = p.lam.at(n) + (int64_t)b * p.lam.batch;
    u1s[s] = p.u.at(n) + (int64_t)b * p.u.batch;
    uos[s] = p.u.at(n + 1) + (int64_t)b * p.u.batch;
  }
  // queues along the slow axis: lambda^(n + 1) and u^n rows y - R .. y + R,
  // the cell means of rows y - R .. y + R - 1
  T ql[S][2 * R + 1], qu[S][2 * R + 1], qc[2 * R];
#pragma unroll
  for (int j = 0; j < 2 * R; ++j) {
    const int64_t r = (int64_t)(ya - R + j) * s0 + xl;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      ql[s][j] = l1s[s][r];
      qu[s][j] = u1s[s][r];
    }
    if (j < 2 * R - 1)
      qc[j] = cyc[r];
  }
  T nl[S], nu[S], nhl[S], nhu[S], nlo[S], nuo[S];
  T nmi, ncx, nhc, ncy;
  auto fetch = [&](int y) {
    const int64_t r = (int64_t)y * s0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      nl[s] = l1s[s][r + (int64_t)R * s0 + xl];
      nu[s] = u1s[s][r + (int64_t)R * s0 + xl];
      nhl[s] = hx >= 0 ? l1s[s][r + hx] : (T)0;
      nhu[s] = hx >= 0 ? u1s[s][r + hx] : (T)0;
      nlo[s] = l0s[s][r + xl];
      nuo[s] = reverse ? uos[s][r + xl] : (T)0;
    }
    nmi = p.inv[r + xl];
    ncx = cx[r + xl];
    nhc = hx >= 0 ? cx[r + hx] : (T)0;
    ncy = cyc[r + (int64_t)(R - 1) * s0 + xl];
  };
  fetch(ya);

  for (int y = ya; y < yb; ++y) {
    T lo_[S], uo_[S], hl[S], hu[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      ql[s][2 * R] = nl[s];
      qu[s][2 * R] = nu[s];
      lo_[s] = nlo[s];
      uo_[s] = nuo[s];
      hl[s] = nhl[s];
      hu[s] = nhu[s];
    }
    qc[2 * R - 1] = ncy;
    const T mi = nmi, ccx = ncx, hc = nhc;
    if (y + 1 < yb)
      fetch(y + 1);
    const int p2 = y & 1;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      rl[s][p2][R + lane] = ql[s][R];
      ru[s][p2][R + lane] = qu[s][R];
      if (hslot >= 0) {
        rl[s][p2][hslot] = hl[s];
        ru[s][p2][hslot] = hu[s];
      }
    }
    rg[p2][R + lane] = ccx;
    if (hslot >= 0)
      rg[p2][hslot] = hc;
    __syncthreads();
    if (active) {
      const int64_t idx = (int64_t)y * s0 + x;
      const int tile = (y / TileShape<2>::TY) * g.tf + tx;
      // this row along the slow axis: slab row, reach of the transposed taps
      const int prow0 = L0 + H0 ? slab_row(y, N0, L0, H0) : -1;
      const bool ext0 = (L0 && y - 1 < L0 + R) || (H0 && N0 - 2 - y < H0 + R);
      const bool layer = prow0 >= 0 || prow1 >= 0;
      const bool grad = !(p.skip_layer && layer);
      // the cell means of both axes
      T chi0[R], clo0[R], chi1[R], clo1[R];
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        chi0[j - 1] = qc[R + j - 1];
        clo0[j - 1] = qc[R - j];
        chi1[j - 1] = rg[p2][R + lane + j - 1];
        clo1[j - 1] = rg[p2][R + lane - j];
      }
      T sc = 0, sp0 = 0, sm0 = 0, sp1 = 0, sm1 = 0;
      if (STIFF && grad) {
        sc = p.stiff[idx];
        sp0 = p.stiff[idx + s0];
        sm0 = p.stiff[idx - s0];
        sp1 = p.stiff[idx + 1];
        sm1 = p.stiff[idx - 1];
      }
      const bool shared = p.gbatch == 0;
      T accm = 0, accs = 0;
      if (grad && shared) {
        accm = p.gmass[(int64_t)gi * g.field + idx];
        if (!p.combine)
          accs = p.gstiff[(int64_t)gi * g.field + idx];
      }
      int strip_col = -1;
      if (reverse && layer && p.strip.map)
        strip_col = p.strip.map[idx];
      Prof<T> pf0{};
      if (prow0 >= 0)
        pf0 = profile_at(p.pml.profile[0], N0, y);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (!on[s])
          continue;
        const int b = b0 + s;
        T lt1[2 * R + 1], ut1[2 * R + 1];
#pragma unroll
        for (int j = 0; j <= 2 * R; ++j) {
          lt1[j] = rl[s][p2][lane + j];
          ut1[j] = ru[s][p2][lane + j];
        }
        const T lc = ql[s][R], uc = qu[s][R];
        T div_l = 0;
        div_l += g.f[0] * transposed<T, R>(ql[s], chi0, clo0, g.op);
        div_l += g.f[1] * transposed<T, R>(lt1, chi1, clo1, g.op);
        T ln = -lo_[s] + (T)2 * lc + mi * div_l;
        if (grad && !shared) {
          accm = p.gmass[(int64_t)b * p.gbatch + idx];
          if (!p.combine)
            accs = p.gstiff[(int64_t)b * p.gbatch + idx];
        }
        // the inertia density by parts in time (scalar_sensitivity.cu:512-514)
        T cm = p.combine ? (T)0 : accm;
        cm = cm - p.mf * uc * (ln - (T)2 * lc + lo_[s]);
        T cs = p.combine ? (T)0 : accs;
        if (STIFF && grad) {
          T gsum = 0;
          gsum += stiffness_density<T, R>(qu[s], ql[s], lc, sc, sp0, sm0,
                                          g.f[0], g.op);
          gsum += stiffness_density<T, R>(ut1, lt1, lc, sc, sp1, sm1, g.f[1],
                                          g.op);
          cs = cs - p.mf * gsum;
        }
        // the transposed layer: the slab nodes' tap fields within reach
        if (ext0) {
          const T *eta = p.pml.eta[0][n & 1] + (int64_t)b * kEta * size0;
          const T sum = gather<T, R>(eta, size0, g.op, [&](int k) {
            const int row = slab_row(y + k, N0, L0, H0);
            return row >= 0 ? (int64_t)row * g.np[1] + x : (int64_t)-1;
          });
          const T delta = mi * sum;
          ln += delta;
          cm -= p.mf * uc * delta;
        }
        if (ext1) {
          const T *eta = p.pml.eta[1][n & 1] + (int64_t)b * kEta * size1;
          const T sum = gather<T, R>(eta, size1, g.op, [&](int k) {
            const int row = slab_row(x + k, N1, L1, H1);
            return row >= 0 ? (int64_t)y * (L1 + H1) + row : (int64_t)-1;
          });
          const T delta = mi * sum;
          ln += delta;
          cm -= p.mf * uc * delta;
        }
        if (p.rec.starts) {
          const int e0 = p.rec.starts[tile], e1 = p.rec.starts[tile + 1];
          for (int e = e0; e < e1; ++e)
            if (p.rec.lin[e] == idx) {
              const T load =
                  p.rec.weight[e] *
                  p.adjoint[(int64_t)b * p.rec.shot_stride +
                            (int64_t)m * p.rec.columns + p.rec.col[e]];
              ln += load;
              cm += -p.mf * uc * load;
            }
        }
        l0s[s][idx] = ln;
        // the memory the next level gathers, from this level's final lambda
        if (prow0 >= 0) {
          const int64_t i = (int64_t)prow0 * g.np[1] + x;
          const int64_t o = (int64_t)b * kEta * size0 + i;
          pml_adjoint_update(p.pml.eta[0][n & 1] + o,
                             p.pml.eta[0][(n + 1) & 1] + o, size0, ln,
                             chi0[0], clo0[0], g.f[0], pf0.an, pf0.bn, pf0.ah,
                             pf0.bh, pf0.ah1, pf0.bh1);
        }
        if (prow1 >= 0) {
          const int64_t i = (int64_t)y * (L1 + H1) + prow1;
          const int64_t o = (int64_t)b * kEta * size1 + i;
          pml_adjoint_update(p.pml.eta[1][n & 1] + o,
                             p.pml.eta[1][(n + 1) & 1] + o, size1, ln,
                             chi1[0], clo1[0], g.f[1], pf1.an, pf1.bn, pf1.ah,
                             pf1.bh, pf1.ah1, pf1.bh1);
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
          T v;
          if (strip_col >= 0) {
            // a layer node is replayed from the forward sweep's record
            const int64_t at =
                ((int64_t)b * p.strip.rows + (n - 1 - p.strip.first)) *
                    p.strip.count +
                strip_col;
            v = p.strip.bf16
                    ? (T)bf16_value(((const uint16_t *)p.strip.record)[at])
                    : ((const T *)p.strip.record)[at];
          } else {
            T div[2];
            div[0] = flux<T, R>(qu[s], chi0[0], clo0[0], g.op);
            div[1] = flux<T, R>(ut1, chi1[0], clo1[0], g.op);
            const T lap = combine_axes<T, 2, R>(div, g.f);
            v = -uo_[s] + (T)2 * uc + mi * lap;
            v = add_sources(p.src, p.signal, tile, idx, b, n + 1, v);
          }
          uos[s][idx] = v;
        }
      }
      if (grad && shared) {
        p.gmass[(int64_t)gi * g.field + idx] = accm;
        if (!p.combine)
          p.gstiff[(int64_t)gi * g.field + idx] = accs;
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) {
        ql[s][j] = ql[s][j + 1];
        qu[s][j] = qu[s][j + 1];
      }
#pragma unroll
    for (int j = 0; j < 2 * R - 1; ++j)
      qc[j] = qc[j + 1];
  }
}

} // namespace cw
