// Device-side geometry shared by the pressure kernels for gfx950.
//
// Field layout: a level of a batched field is (*np, lanes): the shots of one
// node sit next to each other (lanes, a power of two >= the shot count, the
// extra lanes padding that stays zero). A wave's 64 lanes are then 64 / lanes
// neighbouring nodes times their shots: every field access of a wave is one
// contiguous run (128 B for 32 float shots), a stencil tap along any axis is
// another such run, and a material value is one address the node's lanes
// share. Materials, cell means and gradients are per node.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cw {

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr int kXcds = 8;

// the point-list tiles: 64 nodes of the fast axis by 4 rows of the next one
// (256 nodes on a 1D grid); a node's tile names the list it scans
template <int NDIM> struct TileShape {
  static constexpr int TX = NDIM == 1 ? 256 : 64;
  static constexpr int TY = kThreads / TX;
};

// the grid and operator a kernel reads, all by value
template <typename T> struct Geom {
  int n[3];          // logical extents (ghosts included)
  int np[3];         // padded extents
  int s[3];          // node strides, s[ndim - 1] == 1
  int64_t nodes;     // nodes of one field
  int lanes, lshift; // shot lanes a node holds, log2 of it
  int nbatch;        // shots that run (lanes beyond are padding)
  int dirichlet;     // held (odd) faces, bit 2 axis + side
  int neumann;       // refilled even faces
  int tf, tm, ts;    // point tiles along the fast, middle and slow axes
  int ntiles;
  int interior;      // interior nodes (every axis 1 .. n - 2)
  T op[4];           // op[k - 1], the full-radius cell weights
  T f[3];            // step factors per axis
};

// blocks of the 2D band a chunk dealt to one XCD (general_node)
constexpr int kBandChunk = 16;

// block -> work item: contiguous runs of items per XCD, the hardware dealing
// consecutive blocks round-robin over the XCDs
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int q = ntiles / kXcds, r = ntiles % kXcds;
  const int xcd = bid % kXcds, k = bid / kXcds;
  return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// the point tile of node a
template <int NDIM, typename T>
__device__ __forceinline__ int point_tile(const Geom<T> &g, const int *a) {
  constexpr int TX = TileShape<NDIM>::TX, TY = TileShape<NDIM>::TY;
  if (NDIM == 1)
    return a[0] / TX;
  if (NDIM == 2)
    return (a[0] / TY) * g.tf + a[1] / TX;
  return (a[0] * g.tm + a[1] / TY) * g.tf + a[2] / TX;
}

template <typename T> __device__ __forceinline__ T fmat(T a, T b, T c);
template <> __device__ __forceinline__ float fmat(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <> __device__ __forceinline__ double fmat(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// ------------------------------------------------------------ lane groups
// V neighbouring shot lanes of one node that one thread of a general kernel
// holds: 16 bytes (a single wide load or store) where the node has that
// many lanes. Every operation acts lane by lane in the order the scalar
// expression states, so a lane's bits are those of the one-lane arithmetic.
template <typename T, int V> struct alignas(sizeof(T) * V) Lanes {
  T v[V];
  Lanes() = default;
  __device__ __forceinline__ Lanes(T x) {
#pragma unroll
    for (int i = 0; i < V; ++i)
      v[i] = x;
  }
  __device__ __forceinline__ T &operator[](int i) { return v[i]; }
  __device__ __forceinline__ const T &operator[](int i) const { return v[i]; }
};

// the lanes a thread of a general kernel holds at `lanes` shot lanes
template <typename T>
__host__ __device__ constexpr int lanes_per_thread(int lanes) {
  return lanes < (int)(16 / sizeof(T)) ? lanes : (int)(16 / sizeof(T));
}

// the lanes a thread of a scalar general kernel holds: CW_GENERAL_BYTES'
// worth (one 16-byte access by default) or the node's lane count if smaller
#ifndef CW_GENERAL_BYTES
#define CW_GENERAL_BYTES 16
#endif
template <typename T>
__host__ __device__ constexpr int general_lanes(int lanes) {
  return lanes < (int)(CW_GENERAL_BYTES / sizeof(T)) ? lanes
                                                      : (int)(CW_GENERAL_BYTES / sizeof(T));
}

// the value type of V lanes: T itself for one
template <typename T, int V> struct LaneType { using type = Lanes<T, V>; };
template <typename T> struct LaneType<T, 1> { using type = T; };
template <typename T, int V> using LanesOf = typename LaneType<T, V>::type;

#define CW_LANES_OP(OP)                                                        \
  template <typename T, int V>                                                 \
  __device__ __forceinline__ Lanes<T, V> operator OP(const Lanes<T, V> &a,     \
                                                     const Lanes<T, V> &b) {   \
    Lanes<T, V> r;                                                             \
    _Pragma("unroll") for (int i = 0; i < V; ++i) r.v[i] = a.v[i] OP b.v[i];   \
    return r;                                                                  \
  }                                                                            \
  template <typename T, int V>                                                 \
  __device__ __forceinline__ Lanes<T, V> operator OP(T a,                      \
                                                     const Lanes<T, V> &b) {   \
    Lanes<T, V> r;                                                             \
    _Pragma("unroll") for (int i = 0; i < V; ++i) r.v[i] = a OP b.v[i];        \
    return r;                                                                  \
  }                                                                            \
  template <typename T, int V>                                                 \
  __device__ __forceinline__ Lanes<T, V> operator OP(const Lanes<T, V> &a,     \
                                                     T b) {                    \
    Lanes<T, V> r;                                                             \
    _Pragma("unroll") for (int i = 0; i < V; ++i) r.v[i] = a.v[i] OP b;        \
    return r;                                                                  \
  }                                                                            \
  template <typename T, int V>                                                 \
  __device__ __forceinline__ Lanes<T, V> &operator OP##=(Lanes<T, V> &a,       \
                                                         const Lanes<T, V> &b) \
  {                                                                            \
    a = a OP b;                                                                \
    return a;                                                                  \
  }
CW_LANES_OP(+)
CW_LANES_OP(-)
CW_LANES_OP(*)
CW_LANES_OP(/)
#undef CW_LANES_OP

template <typename T, int V>
__device__ __forceinline__ Lanes<T, V> operator-(const Lanes<T, V> &a) {
  Lanes<T, V> r;
#pragma unroll
  for (int i = 0; i < V; ++i)
    r.v[i] = -a.v[i];
  return r;
}

template <typename T, int V>
__device__ __forceinline__ Lanes<T, V> fmat(const Lanes<T, V> &a,
                                           const Lanes<T, V> &b,
                                           const Lanes<T, V> &c) {
  Lanes<T, V> r;
#pragma unroll
  for (int i = 0; i < V; ++i)
    r.v[i] = fmat(a.v[i], b.v[i], c.v[i]);
  return r;
}
template <typename T, int V>
__device__ __forceinline__ Lanes<T, V> fmat(T a, const Lanes<T, V> &b,
                                           const Lanes<T, V> &c) {
  return fmat(Lanes<T, V>(a), b, c);
}
template <typename T, int V>
__device__ __forceinline__ Lanes<T, V> fmat(const Lanes<T, V> &a, T b,
                                           const Lanes<T, V> &c) {
  return fmat(a, Lanes<T, V>(b), c);
}

// lane i of a value (a one-lane value is its own lane 0)
template <typename T>
__device__ __forceinline__ T &lane(T &x, int) { return x; }
template <typename T>
__device__ __forceinline__ const T &lane(const T &x, int) { return x; }
template <typename T, int V>
__device__ __forceinline__ T &lane(Lanes<T, V> &x, int i) { return x.v[i]; }
template <typename T, int V>
__device__ __forceinline__ const T &lane(const Lanes<T, V> &x, int i) {
  return x.v[i];
}

constexpr int log2_of(int v) { return v <= 1 ? 0 : 1 + log2_of(v >> 1); }

// Pin a loaded value in place: an empty asm that reads and writes it, so the
// compiler cannot sink the load into the arm of a select that uses it (it
// turns such selects into branches, each waiting on its own load).
#ifndef CW_KEEP
#define CW_KEEP(x) asm volatile("" : "+v"(x))
#endif
template <typename T> __device__ __forceinline__ void keep(T &x) { CW_KEEP(x); }
template <typename T, int V>
__device__ __forceinline__ void keep(Lanes<T, V> &x) {
#pragma unroll
  for (int i = 0; i < V; ++i)
    CW_KEEP(x.v[i]);
}

// the V lanes at element e of a per-shot array (e a multiple of V)
template <typename Q, typename T>
__device__ __forceinline__ Q lanes_at(const T *__restrict__ p, int64_t e) {
  return *reinterpret_cast<const Q *>(p + e);
}
template <typename Q, typename T>
__device__ __forceinline__ void lanes_put(T *__restrict__ p, int64_t e,
                                          const Q &v) {
  *reinterpret_cast<Q *>(p + e) = v;
}

// whether node a is the wall node of a face of `odd` (bit 0 low, bit 1 high)
__device__ __forceinline__ bool held_wall(int a, int N, int odd) {
  return ((odd & 1) && a == 1) || ((odd & 2) && a == N - 2);
}

// tap j of the line through element e at node a (s the element stride of
// the axis), past the stored ghost the value of its mirror about the wall
// node, negated past a face of `odd`
template <typename Q, typename T>
__device__ __forceinline__ Q mirrored(const T *__restrict__ u, int64_t e,
                                      int64_t s, int a, int N, int odd, int j) {
  int b = a + j;
  const bool neg = (b < 0 && (odd & 1)) || (b > N - 1 && (odd & 2));
  b = b < 0 ? 2 - b : b > N - 1 ? 2 * (N - 2) - b : b;
  const Q v = lanes_at<Q>(u, e + (int64_t)(b - a) * s);
  return neg ? -v : v;
}

// the node step from a to the node `mirrored` reads for tap j, and whether
// that tap is negated
__device__ __forceinline__ int mirror_step(int a, int N, int odd, int j,
                                           bool &neg) {
  const int b = a + j;
  neg = (b < 0 && (odd & 1)) || (b > N - 1 && (odd & 2));
  return (b < 0 ? 2 - b : b > N - 1 ? 2 * (N - 2) - b : b) - a;
}

// tap j through the image about the wall node, for a field no refill keeps
// past it, zero on a held wall node
template <typename Q, typename T>
__device__ __forceinline__ Q reflected(const T *__restrict__ u, int64_t e,
                                       int64_t s, int a, int N, int odd,
                                       int j) {
  int b = a + j;
  const bool neg = (b < 1 && (odd & 1)) || (b > N - 2 && (odd & 2));
  b = b < 1 ? 2 - b : b > N - 2 ? 2 * (N - 2) - b : b;
  // loaded either way (no branch splits a line's loads), zeroed if held
  Q v = lanes_at<Q>(u, e + (int64_t)(b - a) * s);
  keep(v);
  return held_wall(b, N, odd) ? Q((T)0) : neg ? -v : v;
}

// taps -R .. R of a line as `reflected` gives them, t[R] included: every
// tap's load issued first, then all of them kept, then the signs and zeros
template <int R, typename Q, typename T>
__device__ __forceinline__ void reflected_line(const T *__restrict__ u,
                                               int64_t e, int64_t s, int a,
                                               int N, int odd, Q *t) {
  bool neg[2 * R + 1], zero[2 * R + 1];
#pragma unroll
  for (int j = -R; j <= R; ++j) {
    int b = a + j;
    neg[j + R] = (b < 1 && (odd & 1)) || (b > N - 2 && (odd & 2));
    b = b < 1 ? 2 - b : b > N - 2 ? 2 * (N - 2) - b : b;
    zero[j + R] = held_wall(b, N, odd);
    t[j + R] = lanes_at<Q>(u, e + (int64_t)(b - a) * s);
  }
#pragma unroll
  for (int j = 0; j <= 2 * R; ++j)
    keep(t[j]);
#pragma unroll
  for (int j = 0; j <= 2 * R; ++j)
    t[j] = zero[j] ? Q((T)0) : neg[j] ? -t[j] : t[j];
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

// A buffer resource over `bytes` bytes from a base: a load or store takes a
// byte offset from a vector register plus one from a scalar register (or
// the instruction), with no 64-bit address arithmetic per access; the
// hardware drops an access whose vector offset lies past the end (a read
// gives 0). The word-3 value is the gfx9-family raw-buffer format.
struct Buf {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ Buf make_buf(const void *base, uint32_t bytes) {
  return {__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                            (int)bytes, 0x00020000)};
}
// a vector offset past every resource: the access is dropped (a read gives
// 0), the way a lane takes part in an instruction without touching memory
constexpr uint32_t kDrop = 0xFFFFFFF0u;

template <typename T>
__device__ __forceinline__ T bload(const Buf &b, uint32_t vo, uint32_t so);
template <>
__device__ __forceinline__ float bload<float>(const Buf &b, uint32_t vo,
                                              uint32_t so) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_raw_buffer_load_b32(b.r, vo, so, 0));
}
template <>
__device__ __forceinline__ double bload<double>(const Buf &b, uint32_t vo,
                                                uint32_t so) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(b.r, vo, so, 0));
}
__device__ __forceinline__ void bstore(const Buf &b, float v, uint32_t vo,
                                       uint32_t so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), b.r,
                                        vo, so, 0);
}
__device__ __forceinline__ void bstore(const Buf &b, double v, uint32_t vo,
                                       uint32_t so) {
  using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(b.r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), b.r, vo, so,
                                        0);
}

// The LDS DMA of gfx950: each active lane of the wave copies SIZE (4 or 16)
// bytes from its global address `src` to LDS at the wave-uniform `dst` plus
// lane * SIZE, without passing through registers; completion is counted by
// vmcnt like any vector load.
#ifndef CW_LDS_CAST
#define CW_LDS_CAST(p) ((__attribute__((address_space(3))) void *)(p))
#endif
template <int SIZE>
__device__ __forceinline__ void lds_dma(const void *src, void *dst) {
  static_assert(SIZE == 4 || SIZE == 16, "a lane moves 4 or 16 bytes");
  if constexpr (SIZE == 16)
    __builtin_amdgcn_global_load_lds(src, CW_LDS_CAST(dst), 16, 0, 0);
  else
    __builtin_amdgcn_global_load_lds(src, CW_LDS_CAST(dst), 4, 0, 0);
}

// s_waitcnt on vmcnt only: wait until at most n vector memory operations of
// this wave are outstanding (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt
// and lgkmcnt left at their no-wait maxima)
template <int N> __device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt is six bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// the sum over a node's shot lanes (a butterfly: every lane of the node ends
// with the same bits); every lane of the node must take part
template <typename T>
__device__ __forceinline__ T shot_sum(T v, int lanes) {
  for (int m = 1; m < lanes; m <<= 1)
    v += __shfl_xor(v, m, kWave);
  return v;
}

// the same sum over `lanes` lanes held V to a thread: the pairs, quads, ...
// inside the thread, then the butterfly over the node's threads, so the sum
// has the bits of the butterfly over one lane a thread
template <typename T, int V>
__device__ __forceinline__ T shot_sum(Lanes<T, V> x, int lanes) {
#pragma unroll
  for (int m = 1; m < V; m <<= 1)
#pragma unroll
    for (int i = 0; i < V; i += 2 * m)
      x.v[i] += x.v[i + m];
  return shot_sum(x.v[0], lanes / V);
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
