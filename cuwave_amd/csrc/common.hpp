// Device-side geometry shared by the pressure kernels for gfx950.
//
// A launch walks tiles of one shot's grid: 64 consecutive nodes of the fast
// (contiguous) axis by 4 rows of the next axis (256 nodes on a 1D grid), one
// thread a node, each thread stepping its node for every shot of the batch so
// the material loads and the precomputed cell means serve all of them. Tiles
// are dealt to the 8 XCDs in contiguous runs, so vertically adjacent tiles,
// which share their halo rows, sit behind the same L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cw {

constexpr int kThreads = 256;
constexpr int kXcds = 8;

template <int NDIM> struct TileShape {
  static constexpr int TX = NDIM == 1 ? 256 : 64;
  static constexpr int TY = kThreads / TX;
};

// the grid and operator a kernel reads, all by value
template <typename T> struct Geom {
  int n[3];          // logical extents (ghosts included)
  int np[3];         // padded extents
  int s[3];          // element strides, s[ndim - 1] == 1
  int64_t field;     // elements of one shot's field
  int nbatch;
  int dirichlet;     // held (odd) faces, bit 2 axis + side
  int neumann;       // refilled even faces
  int tf, tm, ts;    // tiles along the fast, middle and slow axes
  int ntiles;
  T op[4];           // op[k - 1], the full-radius cell weights
  T f[3];            // step factors per axis
};

// block -> tile: contiguous runs of tiles per XCD, the hardware dealing
// consecutive blocks round-robin over the XCDs
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int q = ntiles / kXcds, r = ntiles % kXcds;
  const int xcd = bid % kXcds, k = bid / kXcds;
  return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// the node of this thread in tile `tile`: a[d] per axis, the tile's first
// node per axis in lo[d] and its last in hi[d]
template <int NDIM, typename T>
__device__ __forceinline__ void tile_node(const Geom<T> &g, int tile, int *a,
                                          int *lo, int *hi) {
  constexpr int TX = TileShape<NDIM>::TX, TY = TileShape<NDIM>::TY;
  const int tx = tile % g.tf, rest = tile / g.tf;
  const int ty = rest % g.tm, tz = rest / g.tm;
  const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
  if (NDIM == 1) {
    a[0] = tx * TX + lx;
    lo[0] = tx * TX;
    hi[0] = lo[0] + TX - 1;
  } else if (NDIM == 2) {
    a[1] = tx * TX + lx;
    lo[1] = tx * TX;
    hi[1] = lo[1] + TX - 1;
    a[0] = ty * TY + ly;
    lo[0] = ty * TY;
    hi[0] = lo[0] + TY - 1;
  } else {
    a[2] = tx * TX + lx;
    lo[2] = tx * TX;
    hi[2] = lo[2] + TX - 1;
    a[1] = ty * TY + ly;
    lo[1] = ty * TY;
    hi[1] = lo[1] + TY - 1;
    a[0] = tz;
    lo[0] = hi[0] = tz;
  }
}

template <typename T> __device__ __forceinline__ T fmat(T a, T b, T c);
template <> __device__ __forceinline__ float fmat(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <> __device__ __forceinline__ double fmat(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// whether node a is the wall node of a face of `odd` (bit 0 low, bit 1 high)
__device__ __forceinline__ bool held_wall(int a, int N, int odd) {
  return ((odd & 1) && a == 1) || ((odd & 2) && a == N - 2);
}

// tap j of the line through idx at node a, past the stored ghost the value
// of its mirror about the wall node, negated past a face of `odd`
template <typename T>
__device__ __forceinline__ T mirrored(const T *__restrict__ u, int64_t idx,
                                      int s, int a, int N, int odd, int j) {
  int b = a + j;
  const bool neg = (b < 0 && (odd & 1)) || (b > N - 1 && (odd & 2));
  b = b < 0 ? 2 - b : b > N - 1 ? 2 * (N - 2) - b : b;
  const T v = u[idx + (int64_t)(b - a) * s];
  return neg ? -v : v;
}

// tap j through the image about the wall node, for a field no refill keeps
// past it, zero on a held wall node
template <typename T>
__device__ __forceinline__ T reflected(const T *__restrict__ u, int64_t idx,
                                       int s, int a, int N, int odd, int j) {
  int b = a + j;
  const bool neg = (b < 1 && (odd & 1)) || (b > N - 2 && (odd & 2));
  b = b < 1 ? 2 - b : b > N - 2 ? 2 * (N - 2) - b : b;
  if (held_wall(b, N, odd))
    return (T)0;
  const T v = u[idx + (int64_t)(b - a) * s];
  return neg ? -v : v;
}

// the row of a slab node of axis d (rows lo off the low wall, hi off the
// high one), or -1 off the slab
__device__ __forceinline__ int slab_row(int a, int N, int lo, int hi) {
  return (a >= 1 && a - 1 < lo) ? a - 1
         : (a <= N - 2 && N - 2 - a < hi) ? lo + N - 2 - a
                                          : -1;
}

// a slab node's index in its axis' memory: the slow axes keep whole slab
// rows together, the fast axis its slab columns per row
template <int NDIM, typename T>
__device__ __forceinline__ int64_t slab_index(const Geom<T> &g, const int *a,
                                              int d, int row, int lo, int hi) {
  int64_t off = 0, plane = 1;
  for (int e = 0; e < NDIM; ++e)
    if (e != d) {
      off = off * g.np[e] + a[e];
      plane *= g.np[e];
    }
  return d < NDIM - 1 ? (int64_t)row * plane + off
                      : off * (int64_t)(lo + hi) + row;
}

template <int NDIM, typename T>
__device__ __forceinline__ int64_t slab_size(const Geom<T> &g, int d, int lo,
                                             int hi) {
  int64_t plane = 1;
  for (int e = 0; e < NDIM; ++e)
    if (e != d)
      plane *= g.np[e];
  return plane * (int64_t)(lo + hi);
}

// round-to-nearest-even bfloat16 bits of a float, as a narrowed record keeps
__device__ __forceinline__ uint16_t bf16_bits(float v) {
  const uint32_t bits = __float_as_uint(v);
  return (uint16_t)((bits + 0x7FFFu + ((bits >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_value(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}

} // namespace cw
