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

} // namespace cw
