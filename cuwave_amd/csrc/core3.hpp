// The core of a 3D forward sweep of one shot lane: the box of nodes at least
// order / 2 + 2 off the walls that neither lie in a PML slab nor within
// R + 1 of one (the rule stream_box applies in 2D). The general kernels take
// the shell around it.
//
// A block is kCoreY rows by kCoreX columns of the box, and it marches a
// chunk of kCoreZ planes down the slow axis; a thread is one column. The
// 2R + 1 values of its column that the slow-axis taps read sit in a register
// queue, fed one plane ahead. Each plane of the tapped field reaches the block
// once, staged in LDS with an R-node halo along the two fast axes (the
// corners, which no tap reads, left out): a thread writes its own column's
// value from the queue and loads at most two halo values, and the middle- and
// fast-axis taps are then LDS reads. Two LDS planes alternate, so one barrier
// a plane orders both the staging and the reads. A plane's global loads are
// issued while the plane before computes (its taps read from LDS first).
//
// Blocks start at multiples of kCoreY rows and kCoreX columns, so a block
// plane lies in one point-tile column and two point-tile rows (TileShape<3>
// is 4 rows by 64 columns): the point lists are looked up per plane from
// those two tiles only. The threads of a block past the box (its first and
// last block row and column) stage their values but store nothing; their
// loads are clamped into the field.
//
// The arithmetic is the general kernels' deep-node path term by term, in the
// same order, so both give the same bits.
#pragma once

#include "stream.hpp"

namespace cw {

constexpr int kCoreX = 32, kCoreY = 8, kCoreZ = 32;
static_assert(kCoreX * kCoreY == kThreads, "a 3D core block is one column a thread");
static_assert(TileShape<3>::TX % kCoreX == 0 && kCoreY % TileShape<3>::TY == 0 &&
                  kCoreY / TileShape<3>::TY == 2,
              "a 3D core block plane spans one point-tile column and two rows");

// the box, and its blocks along each axis (nx, ny from the multiples of
// kCoreX, kCoreY at or below x0, y0)
struct Box3 {
  int z0, z1, y0, y1, x0, x1;
  int nz, ny, nx;
};

__device__ __forceinline__ int clamp_to(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

template <typename T, int R>
__global__ __launch_bounds__(kThreads) void forward_core3(const FwdParams<T> p,
                                                          const Box3 box,
                                                          const int step) {
  constexpr int LX = kCoreX + 2 * R, LY = kCoreY + 2 * R;
  __shared__ T plane[2][LY][LX];
  const int t = clocked_step(p, step);
  if (t < 0)
    return;
  const Geom<T> &g = p.g;
  // block item w: x fastest, then y, then the plane chunk, so the blocks an
  // XCD runs together are neighbours in the plane
  const int w = xcd_tile(blockIdx.x, gridDim.x);
  if (w >= box.nx * box.ny * box.nz)
    return; // a padding block, whole
  const int cx = w % box.nx, cy = (w / box.nx) % box.ny,
            cz = w / (box.nx * box.ny);
  const int tx = (int)threadIdx.x % kCoreX, ty = (int)threadIdx.x / kCoreX;
  const int xb = (box.x0 / kCoreX + cx) * kCoreX;
  const int yb = (box.y0 / kCoreY + cy) * kCoreY;
  const int x = xb + tx, y = yb + ty;
  const bool mine = x >= box.x0 && x < box.x1 && y >= box.y0 && y < box.y1;
  const int za = box.z0 + cz * kCoreZ, ze = min(za + kCoreZ, box.z1);
  const int s0 = g.s[0], s1 = g.s[1];
  const int xmax = g.np[2] - 1, ymax = g.n[1] - 1;
  const int64_t col = (int64_t)clamp_to(y, 0, ymax) * s1 + clamp_to(x, 0, xmax);
  const T *__restrict__ u1 = p.u.at(t - 1);
  const T *__restrict__ u0 = p.u.at(t - 2);
  T *__restrict__ u2 = p.u.at(t);
  // the halo a thread stages: entries h = threadIdx.x and h + kThreads of
  // the 4 R (kCoreX + kCoreY) halo values (rows above, rows below, columns
  // left, columns right), each as its (LDS row, LDS column, field offset)
  constexpr int HX = R * kCoreX, HY = R * kCoreY, NH = 2 * HX + 2 * HY;
  int hr[2], hc[2];
  int64_t ho[2];
  bool hon[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int h = (int)threadIdx.x + k * kThreads;
    hon[k] = h < NH;
    int r = 0, c = 0;
    if (h < HX) {
      r = h / kCoreX;
      c = R + h % kCoreX;
    } else if (h < 2 * HX) {
      r = R + kCoreY + (h - HX) / kCoreX;
      c = R + (h - HX) % kCoreX;
    } else if (h < 2 * HX + HY) {
      r = R + (h - 2 * HX) / R;
      c = (h - 2 * HX) % R;
    } else {
      r = R + (h - 2 * HX - HY) / R;
      c = R + kCoreX + (h - 2 * HX - HY) % R;
    }
    hr[k] = r;
    hc[k] = c;
    ho[k] = (int64_t)clamp_to(yb - R + r, 0, ymax) * s1 +
            clamp_to(xb - R + c, 0, xmax);
  }
  // the point tiles of this block's two halves, per plane
  constexpr int PX = TileShape<3>::TX, PY = TileShape<3>::TY;
  const int tile_base = (yb / PY) * g.tf + xb / PX;
  const int tile_mine = tile_base + (ty >= PY ? g.tf : 0);
  const bool second = yb + PY < g.n[1];
  // a plane's node loads, issued a plane ahead so they fly while the plane
  // before computes: the queue feed, the halo, u^(t-2), the inverse inertia
  // and the cell means (the slow axis' lower one is the plane before's)
  struct Ahead {
    T feed, halo[2], u0, inv, cz, cy, cym, cx, cxm;
  };
  auto fetch = [&](int z, Ahead &f) {
    const int64_t idx = (int64_t)z * s0 + col;
    f.feed = u1[idx + (int64_t)R * s0];
#pragma unroll
    for (int k = 0; k < 2; ++k)
      f.halo[k] = hon[k] ? u1[(int64_t)z * s0 + ho[k]] : (T)0;
    f.u0 = u0[idx];
    f.inv = p.inv[idx];
    f.cz = p.cell[idx];
    f.cy = p.cell[g.nodes + idx];
    f.cym = p.cell[g.nodes + idx - s1];
    f.cx = p.cell[2 * g.nodes + idx];
    f.cxm = p.cell[2 * g.nodes + idx - 1];
  };
  T q[2 * R + 1];
#pragma unroll
  for (int k = 0; k < 2 * R; ++k)
    q[k] = u1[(int64_t)(za - R + k) * s0 + col];
  T czm = p.cell[(int64_t)(za - 1) * s0 + col]; // the slow axis' lower cell
  Ahead cur;
  fetch(za, cur);
  q[2 * R] = cur.feed;
  {
    T(*pl)[LX] = plane[za & 1];
    pl[R + ty][R + tx] = q[R];
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (hon[k])
        pl[hr[k]][hc[k]] = cur.halo[k];
  }
  for (int z = za; z < ze; ++z) {
    __syncthreads();
    T(*pl)[LX] = plane[z & 1];
    T taps[3][2 * R + 1];
#pragma unroll
    for (int j = -R; j <= R; ++j) {
      taps[0][j + R] = q[j + R];
      taps[1][j + R] = pl[R + ty + j][R + tx];
      taps[2][j + R] = pl[R + ty][R + tx + j];
    }
    Ahead next = cur;
    if (z + 1 < ze)
      fetch(z + 1, next);
    // the point lists of the block plane, looked up once for the block
    int a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    if (p.src.starts) {
      const int tile = z * g.tm * g.tf + tile_base;
      const int lo = p.src.starts[tile];
      const int hi = p.src.starts[second ? tile + g.tf + 1 : tile + 1];
      if (lo != hi) {
        a0 = p.src.starts[z * g.tm * g.tf + tile_mine];
        a1 = p.src.starts[z * g.tm * g.tf + tile_mine + 1];
      }
    }
    if (p.rec.starts) {
      const int tile = z * g.tm * g.tf + tile_base;
      const int lo = p.rec.starts[tile];
      const int hi = p.rec.starts[second ? tile + g.tf + 1 : tile + 1];
      if (lo != hi) {
        b0 = p.rec.starts[z * g.tm * g.tf + tile_mine];
        b1 = p.rec.starts[z * g.tm * g.tf + tile_mine + 1];
      }
    }
    if (mine) {
      const int64_t idx = (int64_t)z * s0 + col;
      T div[3];
      div[0] = flux<T, R>(taps[0], cur.cz, czm, g.op);
      div[1] = flux<T, R>(taps[1], cur.cy, cur.cym, g.op);
      div[2] = flux<T, R>(taps[2], cur.cx, cur.cxm, g.op);
      const T lap = combine_axes<T, 3, R>(div, g.f);
      const T uc = taps[0][R];
      T v = -cur.u0 + (T)2 * uc + cur.inv * lap;
      for (int k = a0; k < a1; ++k)
        if (p.src.takes(k, idx, 0))
          v += p.src.weight[k] * p.signal[p.src.at(t, k, 0)];
      u2[idx] = v;
      if (p.illum)
        atomicAdd(p.illum + idx, v * v);
      for (int k = b0; k < b1; ++k)
        if (p.rec.takes(k, idx, 0))
          p.record[p.rec.at(t, k, 0)] = v;
    }
    if (z + 1 < ze) {
#pragma unroll
      for (int k = 0; k < 2 * R; ++k)
        q[k] = q[k + 1];
      q[2 * R] = next.feed;
      czm = cur.cz;
      cur = next;
      T(*nl)[LX] = plane[(z + 1) & 1];
      nl[R + ty][R + tx] = q[R];
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (hon[k])
          nl[hr[k]][hc[k]] = cur.halo[k];
    }
  }
}

// The adjoint level of a 3D sweep of one shot lane over its core box (R + 1
// past the PML slabs, so no transposed-layer tap reaches in): lambda^n over
// the slot of lambda^(n + 2), the gradients against u^n and, with `reverse`,
// u^(n - 1) over the slot of u^(n + 1), as backward_general's deep path
// forms them term by term. Both tapped fields (lambda^(n + 1), u^n) march
// like the forward's field: a register queue down the slow axis, the plane
// with its halo in LDS. The cell means the transposed taps read come from
// LDS along the two fast axes (R either side) and from a queue of 2R planes
// down the slow one; the stiffness of the six neighbours (STIFF) from a
// one-node LDS halo and a queue of three. A plane's global loads are issued
// while the plane before computes.
template <typename T, int R, bool STIFF>
__global__ __launch_bounds__(kThreads) void backward_core3(
    const BwdParams<T> p, const Box3 box, const int n) {
  constexpr int LX = kCoreX + 2 * R, LY = kCoreY + 2 * R;
  constexpr int SX = kCoreX + 2, SY = kCoreY + 2;
  struct Plane {
    T l[LY][LX], u[LY][LX];
    T cy[LY][kCoreX]; // rows yb - R .. yb + kCoreY + R - 1 (the last unread)
    T cx[kCoreY][LX]; // columns xb - R .. xb + kCoreX + R - 1 (the last unread)
    T st[STIFF ? SY : 1][STIFF ? SX : 1];
  };
  __shared__ Plane plane[2];
  const Geom<T> &g = p.g;
  const int w = xcd_tile(blockIdx.x, gridDim.x);
  if (w >= box.nx * box.ny * box.nz)
    return; // a padding block, whole
  const int cx = w % box.nx, cy = (w / box.nx) % box.ny,
            cz = w / (box.nx * box.ny);
  const int tx = (int)threadIdx.x % kCoreX, ty = (int)threadIdx.x / kCoreX;
  const int xb = (box.x0 / kCoreX + cx) * kCoreX;
  const int yb = (box.y0 / kCoreY + cy) * kCoreY;
  const int x = xb + tx, y = yb + ty;
  const bool mine = x >= box.x0 && x < box.x1 && y >= box.y0 && y < box.y1;
  const int za = box.z0 + cz * kCoreZ, ze = min(za + kCoreZ, box.z1);
  const int s0 = g.s[0], s1 = g.s[1];
  const int64_t N = g.nodes;
  const int xmax = g.np[2] - 1, ymax = g.n[1] - 1;
  const int64_t col = (int64_t)clamp_to(y, 0, ymax) * s1 + clamp_to(x, 0, xmax);
  T *__restrict__ l0 = p.lam.at(n);
  const T *__restrict__ l1 = p.lam.at(n + 1);
  const T *__restrict__ u1 = p.u.at(n);
  T *__restrict__ uo = p.u.at(n + 1);
  const bool reverse = p.reverse && n <= p.steps - 2;
  const bool chain = !p.combine;
  const int m = p.steps - 1 - n;
  // the halo entries a thread stages, as (LDS row, LDS column, field offset):
  // the two tapped fields (at most two each), the cell means along y (rows
  // above and below) and x (columns either side), the stiffness' ring
  constexpr int HX = R * kCoreX, HY = R * kCoreY, NH = 2 * HX + 2 * HY;
  int hr[2], hc[2];
  int64_t ho[2];
  bool hon[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int h = (int)threadIdx.x + k * kThreads;
    hon[k] = h < NH;
    int r = 0, c = 0;
    if (h < HX) {
      r = h / kCoreX;
      c = R + h % kCoreX;
    } else if (h < 2 * HX) {
      r = R + kCoreY + (h - HX) / kCoreX;
      c = R + (h - HX) % kCoreX;
    } else if (h < 2 * HX + HY) {
      r = R + (h - 2 * HX) / R;
      c = (h - 2 * HX) % R;
    } else {
      r = R + (h - 2 * HX - HY) / R;
      c = R + kCoreX + (h - 2 * HX - HY) % R;
    }
    hr[k] = r;
    hc[k] = c;
    ho[k] = (int64_t)clamp_to(yb - R + r, 0, ymax) * s1 +
            clamp_to(xb - R + c, 0, xmax);
  }
  // cell means along y: the 2 R - 1 halo rows (R above, R - 1 below) of
  // kCoreX, one a thread (2R - 1 <= 7 rows, 224 <= kThreads)
  const int yr = (int)threadIdx.x / kCoreX; // 0 .. 7
  const bool yon = yr < 2 * R - 1;
  const int yrow = yr < R ? yr : R + kCoreY + (yr - R);
  const int64_t yoff = (int64_t)clamp_to(yb - R + yrow, 0, ymax) * s1 +
                       clamp_to(xb + tx, 0, xmax);
  // cell means along x: the 2 R - 1 halo columns of kCoreY rows
  const bool xon = (int)threadIdx.x < kCoreY * (2 * R - 1);
  const int xr = (int)threadIdx.x / (2 * R - 1), xk = (int)threadIdx.x % (2 * R - 1);
  const int xcol = xk < R ? xk : R + kCoreX + (xk - R);
  const int64_t xoff = (int64_t)clamp_to(yb + (xon ? xr : 0), 0, ymax) * s1 +
                       clamp_to(xb - R + xcol, 0, xmax);
  // the stiffness ring: 2 kCoreX + 2 kCoreY values
  const int sh_ = (int)threadIdx.x;
  const bool son = STIFF && sh_ < 2 * kCoreX + 2 * kCoreY;
  int sr = 0, sc_ = 0;
  if (sh_ < kCoreX) {
    sr = 0;
    sc_ = 1 + sh_;
  } else if (sh_ < 2 * kCoreX) {
    sr = kCoreY + 1;
    sc_ = 1 + sh_ - kCoreX;
  } else if (sh_ < 2 * kCoreX + kCoreY) {
    sr = 1 + sh_ - 2 * kCoreX;
    sc_ = 0;
  } else {
    sr = 1 + (sh_ - 2 * kCoreX - kCoreY) % kCoreY;
    sc_ = kCoreX + 1;
  }
  const int64_t soff = (int64_t)clamp_to(yb - 1 + sr, 0, ymax) * s1 +
                       clamp_to(xb - 1 + sc_, 0, xmax);
  // the point tiles of this block's two halves, per plane
  constexpr int PX = TileShape<3>::TX, PY = TileShape<3>::TY;
  const int tile_base = (yb / PY) * g.tf + xb / PX;
  const int tile_mine = tile_base + (ty >= PY ? g.tf : 0);
  const bool second = yb + PY < g.n[1];
  struct Ahead {
    T lf, uf, czf, sf;             // the queue feeds (planes z + R, z + R,
                                   // z + R - 1, z + 1)
    T lh[2], uh[2], cyh, cxh, sth; // the halos
    T cyc, cxc;                    // this node's own cell means along y, x
    T l0v, uov, inv, gm, gs;       // this node's other values
  };
  auto fetch = [&](int z, Ahead &f) {
    const int64_t idx = (int64_t)z * s0 + col, pz = (int64_t)z * s0;
    f.lf = l1[idx + (int64_t)R * s0];
    f.uf = u1[idx + (int64_t)R * s0];
    f.czf = p.cell[idx + (int64_t)(R - 1) * s0];
    if (STIFF)
      f.sf = p.stiff[idx + s0];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      f.lh[k] = hon[k] ? l1[pz + ho[k]] : (T)0;
      f.uh[k] = hon[k] ? u1[pz + ho[k]] : (T)0;
    }
    f.cyh = yon ? p.cell[N + pz + yoff] : (T)0;
    f.cxh = xon ? p.cell[2 * N + pz + xoff] : (T)0;
    if (STIFF)
      f.sth = son ? p.stiff[pz + soff] : (T)0;
    f.cyc = p.cell[N + idx];
    f.cxc = p.cell[2 * N + idx];
    f.l0v = l0[idx];
    f.uov = reverse ? uo[idx] : (T)0;
    f.inv = p.inv[idx];
    f.gm = p.gmass[idx];
    f.gs = chain ? p.gstiff[idx] : (T)0;
  };
  auto stage = [&](int z, const Ahead &f, const T *ql, const T *qu, T sc) {
    Plane &pl = plane[z & 1];
    pl.l[R + ty][R + tx] = ql[R];
    pl.u[R + ty][R + tx] = qu[R];
    pl.cy[R + ty][tx] = f.cyc;
    pl.cx[ty][R + tx] = f.cxc;
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (hon[k]) {
        pl.l[hr[k]][hc[k]] = f.lh[k];
        pl.u[hr[k]][hc[k]] = f.uh[k];
      }
    if (yon)
      pl.cy[yrow][tx] = f.cyh;
    if (xon)
      pl.cx[xr][xcol] = f.cxh;
    if constexpr (STIFF) {
      pl.st[1 + ty][1 + tx] = sc;
      if (son)
        pl.st[sr][sc_] = f.sth;
    }
  };
  T ql[2 * R + 1], qu[2 * R + 1], qc[2 * R], qs[3];
#pragma unroll
  for (int k = 0; k < 2 * R; ++k) {
    ql[k] = l1[(int64_t)(za - R + k) * s0 + col];
    qu[k] = u1[(int64_t)(za - R + k) * s0 + col];
  }
  // cell means along z: planes za - R .. za + R - 2 now, za + R - 1 the feed
#pragma unroll
  for (int k = 0; k < 2 * R - 1; ++k)
    qc[k] = p.cell[(int64_t)(za - R + k) * s0 + col];
  if (STIFF) {
    qs[0] = p.stiff[(int64_t)(za - 1) * s0 + col];
    qs[1] = p.stiff[(int64_t)za * s0 + col];
  }
  Ahead cur;
  fetch(za, cur);
  ql[2 * R] = cur.lf;
  qu[2 * R] = cur.uf;
  qc[2 * R - 1] = cur.czf;
  if (STIFF)
    qs[2] = cur.sf;
  stage(za, cur, ql, qu, STIFF ? qs[1] : (T)0);
  for (int z = za; z < ze; ++z) {
    __syncthreads();
    const Plane &pl = plane[z & 1];
    T lt[3][2 * R + 1], ut[3][2 * R + 1], ch[3][R], cl[3][R];
    T sp[3], sm[3];
#pragma unroll
    for (int j = -R; j <= R; ++j) {
      lt[0][j + R] = ql[j + R];
      lt[1][j + R] = pl.l[R + ty + j][R + tx];
      lt[2][j + R] = pl.l[R + ty][R + tx + j];
      ut[0][j + R] = qu[j + R];
      ut[1][j + R] = pl.u[R + ty + j][R + tx];
      ut[2][j + R] = pl.u[R + ty][R + tx + j];
    }
#pragma unroll
    for (int j = 1; j <= R; ++j) {
      ch[0][j - 1] = qc[R + j - 1];
      cl[0][j - 1] = qc[R - j];
      ch[1][j - 1] = pl.cy[R + ty + j - 1][tx];
      cl[1][j - 1] = pl.cy[R + ty - j][tx];
      ch[2][j - 1] = pl.cx[ty][R + tx + j - 1];
      cl[2][j - 1] = pl.cx[ty][R + tx - j];
    }
    T sc = (T)0;
    if constexpr (STIFF) {
      sc = qs[1];
      sp[0] = qs[2];
      sm[0] = qs[0];
      sp[1] = pl.st[2 + ty][1 + tx];
      sm[1] = pl.st[ty][1 + tx];
      sp[2] = pl.st[1 + ty][2 + tx];
      sm[2] = pl.st[1 + ty][tx];
    }
    Ahead next = cur;
    if (z + 1 < ze)
      fetch(z + 1, next);
    int a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    if (p.rec.starts) {
      const int tile = z * g.tm * g.tf + tile_base;
      if (p.rec.starts[tile] != p.rec.starts[second ? tile + g.tf + 1 : tile + 1]) {
        b0 = p.rec.starts[z * g.tm * g.tf + tile_mine];
        b1 = p.rec.starts[z * g.tm * g.tf + tile_mine + 1];
      }
    }
    if (reverse && p.src.starts) {
      const int tile = z * g.tm * g.tf + tile_base;
      if (p.src.starts[tile] != p.src.starts[second ? tile + g.tf + 1 : tile + 1]) {
        a0 = p.src.starts[z * g.tm * g.tf + tile_mine];
        a1 = p.src.starts[z * g.tm * g.tf + tile_mine + 1];
      }
    }
    if (mine) {
      const int64_t idx = (int64_t)z * s0 + col;
      const T mi = cur.inv;
      const T lc = lt[0][R], lo_ = cur.l0v, uc = ut[0][R];
      T div_l = (T)0, gsum = (T)0;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        div_l += g.f[d] * transposed<T, R>(lt[d], ch[d], cl[d], g.op);
      if constexpr (STIFF) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          gsum += stiffness_density<T, R>(ut[d], lt[d], lc, sc, sp[d], sm[d],
                                          g.f[d], g.op);
      }
      T ln = -lo_ + (T)2 * lc + mi * div_l;
      T cm = chain ? cur.gm : (T)0;
      cm = cm - p.mf * uc * (ln - (T)2 * lc + lo_);
      T cs = chain ? cur.gs : (T)0;
      cs = cs - p.mf * gsum;
      for (int k = b0; k < b1; ++k)
        if (p.rec.takes(k, idx, 0)) {
          const T load = p.rec.weight[k] * p.adjoint[p.rec.at(m, k, 0)];
          ln += load;
          cm += -p.mf * uc * load;
        }
      l0[idx] = ln;
      if (p.combine) {
        p.gmass[idx] = cur.gm + (p.jm * cm + p.js * cs);
      } else {
        p.gmass[idx] = cm;
        p.gstiff[idx] = cs;
      }
      if (reverse) {
        T div[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
          div[d] = flux<T, R>(ut[d], ch[d][0], cl[d][0], g.op);
        const T lap = combine_axes<T, 3, R>(div, g.f);
        T v = -cur.uov + (T)2 * uc + mi * lap;
        for (int k = a0; k < a1; ++k)
          if (p.src.takes(k, idx, 0))
            v += p.src.weight[k] * p.signal[p.src.at(n + 1, k, 0)];
        uo[idx] = v;
      }
    }
    if (z + 1 < ze) {
#pragma unroll
      for (int k = 0; k < 2 * R; ++k) {
        ql[k] = ql[k + 1];
        qu[k] = qu[k + 1];
      }
      ql[2 * R] = next.lf;
      qu[2 * R] = next.uf;
#pragma unroll
      for (int k = 0; k < 2 * R - 1; ++k)
        qc[k] = qc[k + 1];
      qc[2 * R - 1] = next.czf;
      if constexpr (STIFF) {
        qs[0] = qs[1];
        qs[1] = qs[2];
        qs[2] = next.sf;
      }
      cur = next;
      stage(z + 1, cur, ql, qu, STIFF ? qs[1] : (T)0);
    }
  }
}

} // namespace cw
