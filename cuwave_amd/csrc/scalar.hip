// The C ABI of include/cuwave_amd.h for the pressure family: argument
// checks, template dispatch over (precision, dimension, radius), the
// material preparation and the superposition step. The step launchers live
// in steps_*.hip, one translation unit per (precision, radius).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"

namespace cwh {

thread_local std::string g_error;

int fail(const std::string &what) {
  g_error = what;
  return 1;
}

namespace {
constexpr int kDevices = 64;
std::mutex g_side_mutex;
std::vector<Side *> g_side_free[kDevices];
std::atomic<int> g_side_enabled{1};
} // namespace

Side *lease_side(bool *off) {
  *off = g_side_enabled.load() == 0;
  if (*off)
    return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) {
    fail("side stream: no current device");
    return nullptr;
  }
  {
    std::lock_guard<std::mutex> hold(g_side_mutex);
    if (!g_side_free[dev].empty()) {
      Side *s = g_side_free[dev].back();
      g_side_free[dev].pop_back();
      return s;
    }
  }
  Side *s = new Side;
  s->device = dev;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->core, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->edge, hipEventDisableTiming) != hipSuccess) {
    fail("side stream: creation failed");
    if (s->stream)
      (void)hipStreamDestroy(s->stream);
    if (s->core)
      (void)hipEventDestroy(s->core);
    if (s->edge)
      (void)hipEventDestroy(s->edge);
    delete s;
    return nullptr;
  }
  return s;
}

void return_side(Side *s) {
  std::lock_guard<std::mutex> hold(g_side_mutex);
  g_side_free[s->device].push_back(s);
}

int side_enabled() { return g_side_enabled.load(); }
void set_side_enabled(int on) { g_side_enabled.store(on ? 1 : 0); }

// two blocks a CU: a box that cannot fill the GPU once over marches too few
// chunks to pay for its row-by-row latency
std::atomic<int> g_stream_threshold{512};
int stream_threshold() { return g_stream_threshold.load(); }

} // namespace cwh

namespace cwh {
int side_enabled();
void set_side_enabled(int on);
} // namespace cwh

namespace {

using cwh::fail;
using cwh::hip_status;
using cwh::make_geom;

int check_grid(const cw_grid *g) {
  if (!g)
    return fail("grid is NULL");
  if (g->ndim < 1 || g->ndim > 3)
    return fail("ndim must be 1, 2 or 3");
  if (g->order != 2 && g->order != 4 && g->order != 6 && g->order != 8)
    return fail("space order must be 2, 4, 6 or 8");
  if (g->dtype != 0 && g->dtype != 1)
    return fail("dtype must be 0 (float32) or 1 (float64)");
  if (g->nbatch < 1)
    return fail("nbatch must be at least 1");
  if (g->lanes < g->nbatch || g->lanes > cw::kWave ||
      (g->lanes & (g->lanes - 1)))
    return fail("lanes must be a power of two from nbatch to 64");
  const int R = g->order / 2;
  int64_t nodes = 1;
  for (int d = 0; d < g->ndim; ++d) {
    if (g->n[d] < R + 3)
      return fail("every axis needs order / 2 + 3 nodes to mirror its taps");
    const int want = d == g->ndim - 1 ? (g->n[d] + 31) / 32 * 32 : g->n[d];
    if (g->np[d] != want)
      return fail("np must be n with the last axis padded to a multiple of 32");
    nodes *= g->np[d];
  }
  if (nodes * g->lanes >= (int64_t)1 << 31)
    return fail("one level of the batch overflows the int32 element index");
  return 0;
}

int check_levels(const cw_levels &l, const char *what) {
  if (!l.base || l.slots < 1)
    return fail(std::string(what) + ": levels need a base and slots");
  return 0;
}

int check_layer(const cw_grid &g, const cw_pml &p, bool forward) {
  for (int d = 0; d < g.ndim; ++d) {
    if (p.lo[d] < 0 || p.hi[d] < 0)
      return fail("pml widths must be non-negative");
    if (!(p.lo[d] + p.hi[d]))
      continue;
    const int R = g.order / 2;
    if ((p.lo[d] ? p.lo[d] + R : 0) + (p.hi[d] ? p.hi[d] + R : 0) >= g.n[d] - 2)
      return fail("the pml layers and their reach leave no interior");
    if (!p.profile[d])
      return fail("a lined axis needs its profile");
    if (forward ? !p.aux[d] : (!p.eta[d][0] || !p.eta[d][1]))
      return fail("a lined axis needs its memory");
  }
  return 0;
}

int check_points(const cw_points &p, bool weighted, const char *what) {
  if (!p.starts)
    return 0;
  if (!p.lin || !p.shot || !p.col || (weighted && !p.weight))
    return fail(std::string(what) + ": a point list needs lin, shot, col" +
                (weighted ? " and weight" : ""));
  if (p.row_stride < 0 || p.col_stride < 0 || p.shot_stride < 0)
    return fail(std::string(what) + ": strides must be non-negative");
  return 0;
}

// ------------------------------------------------------------ dispatch
template <template <typename, int, int> class F, typename... A>
int dispatch(const cw_grid &g, A &&...args) {
  const int R = g.order / 2;
#define CW_CASE(T, NDIM, RR)                                                   \
  if ((g.dtype == 1) == (sizeof(T) == 8) && g.ndim == NDIM && R == RR)         \
    return F<T, NDIM, RR>::run(args...);
#define CW_RADII(T, NDIM)                                                      \
  CW_CASE(T, NDIM, 1) CW_CASE(T, NDIM, 2) CW_CASE(T, NDIM, 3)                  \
  CW_CASE(T, NDIM, 4)
  CW_RADII(float, 1)
  CW_RADII(float, 2)
  CW_RADII(float, 3)
  CW_RADII(double, 1)
  CW_RADII(double, 2)
  CW_RADII(double, 3)
#undef CW_RADII
#undef CW_CASE
  return fail("no kernel for this precision, dimension and order");
}

template <typename T, int NDIM, int R> struct Prepare {
  static int run(const cw_grid &gr, const cw_materials &m,
                 hipStream_t stream) {
    const cw::Geom<T> g = make_geom<T>(gr);
    const int blocks = (int)((g.nodes + cw::kThreads - 1) / cw::kThreads);
    if (R == 1) // one preparation serves every radius
      hipLaunchKernelGGL((cw::prepare_kernel<T, NDIM>), dim3(blocks),
                         dim3(cw::kThreads), 0, stream, g, (const T *)m.stiff,
                         (const T *)m.minv, (T *)m.prep + NDIM * g.nodes,
                         (T *)m.prep);
    else
      return Prepare<T, NDIM, 1>::run(gr, m, stream);
    return hip_status("prepare_kernel");
  }
};

template <typename T, int NDIM, int R> struct Forward {
  static int run(const cw_forward_args &a, int t0, int t1,
                 hipStream_t stream) {
    return cwh::forward_steps<T, NDIM, R>(a, t0, t1, stream);
  }
};

// the clock's advance after a window of `steps` steps (one thread)
__global__ void clock_tick(int *clock, int steps) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    clock[threadIdx.x] += steps;
}

// a window of steps 0 .. steps - 1 past the device clock, then its advance
template <typename T, int NDIM, int R> struct Clocked {
  static int run(const cw_forward_args &a, int32_t *clock, int steps,
                 int t_end, hipStream_t stream) {
    if (int e = cwh::forward_steps<T, NDIM, R>(a, 0, steps, stream, clock,
                                               t_end))
      return e;
    hipLaunchKernelGGL(clock_tick, dim3(1), dim3(64), 0, stream, clock, steps);
    return hip_status("clock_tick");
  }
};

template <typename T, int NDIM, int R> struct Backward {
  static int run(const cw_backward_args &a, int n0, int n1,
                 hipStream_t stream) {
    return cwh::backward_steps<T, NDIM, R>(a, n0, n1, stream);
  }
};

template <typename T, int NDIM, int R> struct Superposed {
  static int run(const cw_superposed_args &a, hipStream_t stream) {
    const cw::Geom<T> g = make_geom<T>(a.grid);
    const int blocks = cwh::general_blocks(a.grid, cw::Region{}, 1);
    const T *cell = (const T *)a.mat.prep;
    hipLaunchKernelGGL((cw::superposed_kernel<T, NDIM, R>), dim3(blocks),
                       dim3(cw::kThreads), 0, stream, g, cell + NDIM * g.nodes,
                       cell, (const T *)a.u0, (const T *)a.u1,
                       (T *)a.u2, (T *)a.acc_mass, (T *)a.acc_stiff, (T)a.ft,
                       (T)a.fs, a.step);
    return hip_status("superposed_kernel");
  }
};

template <typename T, int NDIM, int R> struct Gradient {
  static int run(const cw_gradient_args &a, hipStream_t stream) {
    cw::Geom<T> g = make_geom<T>(a.grid);
    for (int d = 0; d < NDIM; ++d)
      g.f[d] = (T)a.factor[d];
    const int blocks = cwh::general_blocks(a.grid, cw::Region{}, 1);
    hipLaunchKernelGGL((cw::gradient_kernel<T, NDIM, R>), dim3(blocks),
                       dim3(cw::kThreads), 0, stream, g, (const T *)a.mat.stiff,
                       (const T *)a.u0, (const T *)a.u1, (const T *)a.u2,
                       (const T *)a.l1, (T *)a.g_mass, (T *)a.g_stiff,
                       (T)a.inv_dt2);
    return hip_status("gradient_kernel");
  }
};

int check_forward(const cw_forward_args &a, int32_t t0, int32_t t1);

// ------------------------------------------------------- static scratch
// The workspace of cw_scalar_forward_static: one array on every GPU the
// library runs on, held until the process ends. Host code never names it
// past hipGetSymbolAddress: `scratch` on the host is a stand-in, not a device
// address, so every launch and fill takes the address that call returned for
// the current device.
static __device__ __align__(256) unsigned char scratch[CW_STATIC_SCRATCH_BYTES];

std::mutex g_scratch_mutex;
unsigned char *g_scratch_of[64] = {};

// the scratch of the current device, looked up once per device
int static_scratch(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return fail("static scratch: no current device");
  std::lock_guard<std::mutex> hold(g_scratch_mutex);
  if (!g_scratch_of[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("static scratch: hipGetSymbolAddress: ") +
                  hipGetErrorString(e));
    g_scratch_of[dev] = (unsigned char *)p;
  }
  *out = g_scratch_of[dev];
  return 0;
}

// zero `words` 16-byte words (a fill kernel, so a caller may capture it)
__global__ void clear_words(uint4 *p, int64_t words) {
  const uint4 z = {0u, 0u, 0u, 0u};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z;
}

} // namespace

extern "C" {

int cw_version(void) { return CW_ABI_VERSION; }

int cw_set_side_stream(int32_t enabled) {
  const int was = cwh::side_enabled();
  cwh::set_side_enabled(enabled);
  return was;
}

int cw_set_stream_threshold(int32_t blocks) {
  return cwh::g_stream_threshold.exchange(blocks < 0 ? 0 : blocks);
}

const char *cw_last_error(void) { return cwh::g_error.c_str(); }

int cw_tile_count(const cw_grid *grid, int32_t *tiles) {
  if (int e = check_grid(grid))
    return e;
  if (!tiles)
    return fail("tiles is NULL");
  *tiles = make_geom<float>(*grid).ntiles;
  return 0;
}

int cw_point_lists(const cw_grid *grid, const int32_t *lin, int32_t count,
                   int32_t *starts, int32_t *order) {
  if (int e = check_grid(grid))
    return e;
  if (count < 0 || (count && (!lin || !order)) || !starts)
    return fail("point lists need lin, order and starts");
  const cw::Geom<float> g = make_geom<float>(*grid);
  const int nd = grid->ndim;
  const int TX = nd == 1 ? cw::TileShape<1>::TX : cw::TileShape<2>::TX;
  const int TY = cw::kThreads / TX;
  std::vector<int32_t> tile(count);
  for (int j = 0; j < count; ++j) {
    if (lin[j] < 0 || lin[j] >= g.nodes)
      return fail("a point lies outside the field");
    int a[3] = {0, 0, 0};
    int64_t rest = lin[j];
    for (int d = 0; d < nd; ++d) {
      a[d] = (int)(rest / g.s[d]);
      rest -= (int64_t)a[d] * g.s[d];
    }
    const int tx = a[nd - 1] / TX;
    const int ty = nd >= 2 ? a[nd - 2] / TY : 0;
    const int tz = nd == 3 ? a[0] : 0;
    tile[j] = (tz * g.tm + ty) * g.tf + tx;
  }
  for (int k = 0; k <= g.ntiles; ++k)
    starts[k] = 0;
  for (int j = 0; j < count; ++j)
    ++starts[tile[j] + 1];
  for (int k = 0; k < g.ntiles; ++k)
    starts[k + 1] += starts[k];
  std::vector<int32_t> fill(starts, starts + g.ntiles);
  for (int j = 0; j < count; ++j)
    order[fill[tile[j]]++] = j;
  return 0;
}

int cw_scalar_prepare(const cw_grid *grid, const cw_materials *mat,
                      void *stream) {
  if (int e = check_grid(grid))
    return e;
  if (!mat || !mat->stiff || !mat->prep)
    return fail("prepare needs stiff and prep");
  return dispatch<Prepare>(*grid, *grid, *mat, (hipStream_t)stream);
}

int cw_scalar_forward_clocked(const cw_forward_args *args, int32_t *clock,
                              int32_t steps, int32_t t_end, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (!clock)
    return fail("clock is NULL");
  if (steps < 0 || t_end < 0)
    return fail("steps and t_end must be non-negative");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->mat.prep)
    return fail("forward needs the prepared materials");
  // every step the clock can reach lies in 0 .. t_end - 1; the kernels skip
  // the rest
  if (int e = check_forward(*args, 0, t_end))
    return e;
  if (steps == 0)
    return 0;
  return dispatch<Clocked>(args->grid, *args, clock, (int)steps, (int)t_end,
                           (hipStream_t)stream);
}

int cw_scalar_gradient(const cw_gradient_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (args->grid.lanes != 1)
    return fail("the field gradient takes one shot");
  if (!args->mat.stiff || !args->u0 || !args->u1 || !args->u2 || !args->l1 ||
      !args->g_mass || !args->g_stiff)
    return fail("the field gradient needs stiff, u0, u1, u2, l1 and both accumulators");
  if (args->mat.cells)
    return fail("a domain mask is forward only: the adjoint kernels grade no domain wall");
  return dispatch<Gradient>(args->grid, *args, (hipStream_t)stream);
}

int cw_scalar_forward(const cw_forward_args *args, int32_t t0, int32_t t1,
                      void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->mat.prep)
    return fail("forward needs the prepared materials");
  if (int e = check_forward(*args, t0, t1))
    return e;
  if (t0 == t1)
    return 0;
  return dispatch<Forward>(args->grid, *args, t0, t1, (hipStream_t)stream);
}

} // extern "C"

namespace {

int check_forward(const cw_forward_args &a0, int32_t t0, int32_t t1) {
  const cw_forward_args *args = &a0;
  if (int e = check_levels(args->u, "u"))
    return e;
  if (int e = check_layer(args->grid, args->pml, true))
    return e;
  if (int e = check_points(args->sources, true, "sources"))
    return e;
  if (int e = check_points(args->sensors, false, "sensors"))
    return e;
  if (args->sources.starts && !args->signal)
    return fail("sources need a signal");
  if (args->sensors.starts && !args->record)
    return fail("sensors need a record");
  if (args->strip.map && (!args->strip.record || args->strip.count < 1))
    return fail("a strip needs its record and count");
  if (args->strip.bf16 && args->grid.dtype != 0)
    return fail("a bfloat16 strip needs float32 fields");
  if (t0 < 0 || t1 > args->steps || t0 > t1)
    return fail("steps t0 .. t1 must lie within 0 .. steps");
  if (args->mat.damping && !(args->mat.dt > 0.0))
    return fail("a damping field needs dt > 0");
  if (args->mat.cells) {
    if (!args->mat.stiff)
      return fail("a domain mask needs the stiffness its held cells read");
    for (int d = 0; d < args->grid.ndim; ++d)
      if (args->pml.lo[d] || args->pml.hi[d])
        return fail("a pml takes no domain mask");
    if (args->strip.map)
      return fail("a domain mask takes no strip (its reverse march is not built)");
  }
  if (args->strip.map && t0 < t1 &&
      (t0 < args->strip.first || t1 > args->strip.first + args->strip.rows))
    return fail("the strip record does not hold steps t0 .. t1");
  return 0;
}

} // namespace

extern "C" {

int cw_scalar_backward(const cw_backward_args *args, int32_t n0, int32_t n1,
                       void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->mat.prep || !args->mat.stiff)
    return fail("backward needs stiff and the prepared materials");
  if (args->mat.cells)
    return fail("a domain mask is forward only: the adjoint kernels grade no domain wall");
  if (int e = check_levels(args->lam, "lambda"))
    return e;
  if (int e = check_levels(args->u, "u"))
    return e;
  if (args->lam.slots != 2)
    return fail("lambda takes two slots");
  if ((args->mat.damping || args->mat.damping_reverse) && !(args->mat.dt > 0.0))
    return fail("a damping field needs dt > 0");
  if (int e = check_layer(args->grid, args->pml, false))
    return e;
  if (int e = check_points(args->sensors, true, "sensors"))
    return e;
  if (int e = check_points(args->sources, true, "sources"))
    return e;
  if (args->sensors.starts && !args->adjoint_signal)
    return fail("sensors need the adjoint signal");
  if (args->reverse && args->sources.starts && !args->signal)
    return fail("the reverse march needs the source signal");
  if (args->reverse && args->u.slots != 2)
    return fail("the reverse march takes two slots of u");
  if (args->strip.map && (!args->strip.record || args->strip.count < 1))
    return fail("a strip needs its record and count");
  if (args->strip.bf16 && args->grid.dtype != 0)
    return fail("a bfloat16 strip needs float32 fields");
  if (args->reverse && args->strip.map && n0 < n1) {
    // level n replays the strip of step n - 1, for n up to steps - 2
    const int top = n1 - 1 < args->steps - 2 ? n1 - 1 : args->steps - 2;
    const int lo = n0 - 1, hi = top - 1;
    if (lo <= hi && (lo < args->strip.first ||
                     hi >= args->strip.first + args->strip.rows))
      return fail("the strip record does not hold the replayed steps");
  }
  if (!args->g_mass || (!args->combine && !args->g_stiff))
    return fail("backward needs its gradient accumulators");
  if (n0 < 0 || n1 > args->steps || n0 > n1)
    return fail("levels n0 .. n1 must lie within 0 .. steps");
  if (n0 == n1)
    return 0;
  return dispatch<Backward>(args->grid, *args, n0, n1, (hipStream_t)stream);
}

int cw_scalar_forward_static(const cw_forward_args *args, int32_t t0,
                             int32_t t1, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  cw_forward_args a = *args;
  const int64_t elem = a.grid.dtype ? 8 : 4;
  int64_t nodes = 1;
  for (int d = 0; d < a.grid.ndim; ++d)
    nodes *= a.grid.np[d];
  // the layout of the memories in the scratch: each lined axis' (3, slab,
  // lanes) block, rounded up to 256 bytes
  int64_t offset[3] = {0, 0, 0}, used = 0;
  for (int d = 0; d < a.grid.ndim; ++d) {
    if (a.pml.lo[d] < 0 || a.pml.hi[d] < 0)
      return fail("pml widths must be non-negative");
    if (!(a.pml.lo[d] + a.pml.hi[d]))
      continue;
    const int64_t slab = (nodes / a.grid.np[d]) * (a.pml.lo[d] + a.pml.hi[d]);
    offset[d] = used;
    used += (3 * slab * a.grid.lanes * elem + 255) / 256 * 256;
  }
  if (used > (int64_t)CW_STATIC_SCRATCH_BYTES)
    return fail("the layer memories need " + std::to_string(used) +
                " bytes, past the static scratch of " +
                std::to_string((int64_t)CW_STATIC_SCRATCH_BYTES) + " bytes");
  unsigned char *base = nullptr;
  if (used) {
    if (int e = static_scratch(&base))
      return e;
  }
  for (int d = 0; d < a.grid.ndim; ++d)
    a.pml.aux[d] = (a.pml.lo[d] + a.pml.hi[d]) ? base + offset[d] : nullptr;
  if (!a.mat.prep)
    return fail("forward needs the prepared materials");
  if (int e = check_forward(a, t0, t1))
    return e;
  if (t0 == t1)
    return 0;
  if (t0 == 0 && used) {
    const int64_t words = used / 16;
    const int blocks = (int)std::min<int64_t>((words + cw::kThreads - 1) / cw::kThreads, 4096);
    hipLaunchKernelGGL(clear_words, dim3(blocks), dim3(cw::kThreads), 0,
                       (hipStream_t)stream, (uint4 *)base, words);
    if (int e = hip_status("clear_words"))
      return e;
  }
  return dispatch<Forward>(a.grid, a, t0, t1, (hipStream_t)stream);
}

int cw_scalar_superposed(const cw_superposed_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (args->grid.nbatch != 1 || args->grid.lanes != 1)
    return fail("the superposition step takes one shot a call");
  if (!args->mat.prep || !args->u0 || !args->u1 ||
      !args->u2 || !args->acc_mass || !args->acc_stiff)
    return fail("superposed needs prep, u0, u1, u2 and both accumulators");
  if (args->mat.cells)
    return fail("a domain mask is forward only: the adjoint kernels grade no domain wall");
  return dispatch<Superposed>(args->grid, *args, (hipStream_t)stream);
}

} // extern "C"
