// The C ABI of include/cuwave_amd.h for the pressure family: argument
// checks, template dispatch over (precision, dimension, radius), and the
// time loops that launch the fused kernels of a step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cuwave_amd.h"
#include "scalar_kernels.hpp"
#include "stream.hpp"

namespace {

thread_local std::string g_error;

int fail(const std::string &what) {
  g_error = what;
  return 1;
}

int hip_status(const char *where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_error = std::string(where) + ": " + hipGetErrorString(e);
    return 2;
  }
  return 0;
}

int log2_of(int v) {
  int k = 0;
  while ((1 << k) < v)
    ++k;
  return k;
}

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

template <typename T> cw::Geom<T> make_geom(const cw_grid &gr) {
  cw::Geom<T> g{};
  const int nd = gr.ndim;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = d < nd ? gr.n[d] : 1;
    g.np[d] = d < nd ? gr.np[d] : 1;
  }
  int64_t stride = 1;
  for (int d = nd - 1; d >= 0; --d) {
    g.s[d] = (int)stride;
    stride *= g.np[d];
  }
  for (int d = nd; d < 3; ++d)
    g.s[d] = 0;
  g.nodes = stride;
  g.lanes = gr.lanes;
  g.lshift = log2_of(gr.lanes);
  g.nbatch = gr.nbatch;
  g.dirichlet = gr.dirichlet;
  g.neumann = gr.neumann;
  const int TX = nd == 1 ? cw::TileShape<1>::TX : cw::TileShape<2>::TX;
  const int TY = cw::kThreads / TX;
  g.tf = (g.np[nd - 1] + TX - 1) / TX;
  g.tm = nd >= 2 ? (g.n[nd - 2] + TY - 1) / TY : 1;
  g.ts = nd == 3 ? g.n[0] : 1;
  g.ntiles = g.tf * g.tm * g.ts;
  g.interior = 1;
  for (int d = 0; d < nd; ++d)
    g.interior *= g.n[d] - 2;
  for (int k = 0; k < 4; ++k)
    g.op[k] = (T)gr.op[k];
  for (int d = 0; d < 3; ++d)
    g.f[d] = (T)gr.f[d];
  return g;
}

template <typename T> cw::Points<T> make_points(const cw_points &p) {
  cw::Points<T> q{};
  q.starts = p.starts;
  q.lin = p.lin;
  q.shot = p.shot;
  q.col = p.col;
  q.weight = (const T *)p.weight;
  q.rs = p.row_stride;
  q.cs = p.col_stride;
  q.ss = p.shot_stride;
  return q;
}

template <typename T> cw::Levels<T> make_levels(const cw_levels &l) {
  cw::Levels<T> q{};
  q.base = (T *)l.base;
  q.slot = l.slot_stride;
  q.slots = l.slots;
  q.offset = l.offset;
  return q;
}

template <typename T> cw::Layer<T> make_layer(const cw_pml &p, int ndim) {
  cw::Layer<T> q{};
  for (int d = 0; d < 3; ++d) {
    q.lo[d] = d < ndim ? p.lo[d] : 0;
    q.hi[d] = d < ndim ? p.hi[d] : 0;
    q.profile[d] = (const T *)p.profile[d];
    q.aux[d] = (T *)p.aux[d];
    q.eta[d][0] = (T *)p.eta[d][0];
    q.eta[d][1] = (T *)p.eta[d][1];
  }
  return q;
}

template <typename T> cw::Strip<T> make_strip(const cw_strip &s) {
  cw::Strip<T> q{};
  q.map = s.map;
  q.record = s.record;
  q.count = s.count;
  q.first = s.first;
  q.rows = s.rows;
  q.bf16 = s.bf16;
  return q;
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
                         (const T *)m.minv, (T *)m.inv, (T *)m.cell);
    else
      return Prepare<T, NDIM, 1>::run(gr, m, stream);
    return hip_status("prepare_kernel");
  }
};

// ---------------------------------------------------------- stream box
constexpr int kChunk = 32; // rows a streaming thread marches

// the stream box of a 2D grid: every node order / 2 + 2 off the walls, out
// of reach of the wall mirror, the refill, the held nodes and the adjoint's
// reflected taps; empty on other grids, on grids too small for one and on
// levels past the 4 GiB a buffer resource addresses
cw::Box stream_box(const cw_grid &g) {
  cw::Box box{};
  const int margin = g.order / 2 + 2;
  box.chunk = kChunk;
  if (g.ndim != 2 || g.n[0] - margin <= margin || g.n[1] - margin <= margin)
    return box;
  const int64_t bytes = (int64_t)g.np[0] * g.np[1] * g.lanes * (g.dtype ? 8 : 4);
  if (bytes >= ((int64_t)1 << 32))
    return box;
  box.y0 = margin;
  box.y1 = g.n[0] - margin;
  box.x0 = margin;
  box.x1 = g.n[1] - margin;
  return box;
}

// the stream blocks of a sweep, padded to a multiple of the XCD count so
// the general blocks after them start on XCD 0
int box_blocks(const cw::Box &b, int lanes) {
  if (b.y0 >= b.y1)
    return 0;
  const int cols = cw::kThreads / lanes;
  const int xa = b.x0 / cols * cols;
  const int nchunk = (b.y1 - b.y0 + b.chunk - 1) / b.chunk;
  const int blocks = nchunk * ((b.x1 - xa + cols - 1) / cols);
  return (blocks + cw::kXcds - 1) / cw::kXcds * cw::kXcds;
}

// where the general kernels run: the band around the box, or every node
cw::Region band_of(const cw_grid &g, const cw::Box &b) {
  cw::Region r{};
  if (b.y0 >= b.y1)
    return r;
  r.band = 1;
  r.y0 = b.y0;
  r.y1 = b.y1;
  r.x0 = b.x0;
  r.x1 = b.x1;
  const int W = g.n[1] - 2;
  r.count = (b.y0 - 1) * W + (g.n[0] - 1 - b.y1) * W +
            (b.y1 - b.y0) * ((b.x0 - 1) + (g.n[1] - 1 - b.x1));
  return r;
}

int general_blocks(const cw_grid &gr, const cw::Region &r) {
  const int per = cw::kThreads / gr.lanes; // nodes a block takes
  int64_t nodes = r.band ? r.count : 1;
  if (!r.band)
    for (int d = 0; d < gr.ndim; ++d)
      nodes *= gr.n[d] - 2;
  return (int)((nodes + per - 1) / per);
}

template <typename T, int NDIM, int R> struct Forward {
  static int run(const cw_forward_args &a, int t0, int t1,
                 hipStream_t stream) {
    cw::FwdParams<T> p{};
    p.g = make_geom<T>(a.grid);
    p.inv = (const T *)a.mat.inv;
    p.cell = (const T *)a.mat.cell;
    p.pml = make_layer<T>(a.pml, a.grid.ndim);
    p.u = make_levels<T>(a.u);
    p.src = make_points<T>(a.sources);
    p.signal = (const T *)a.signal;
    p.rec = make_points<T>(a.sensors);
    p.record = (T *)a.record;
    p.strip = make_strip<T>(a.strip);
    const cw::Box box = stream_box(a.grid);
    p.region = band_of(a.grid, box);
    const int streams = NDIM == 2 ? box_blocks(box, a.grid.lanes) : 0;
    const int blocks = streams + general_blocks(a.grid, p.region);
    for (int t = t0; t < t1; ++t)
      hipLaunchKernelGGL((cw::forward_step<T, NDIM, R>), dim3(blocks),
                         dim3(cw::kThreads), 0, stream, p, box, t, streams);
    return hip_status("forward");
  }
};

template <typename T, int NDIM, int R> struct Backward {
  static int run(const cw_backward_args &a, int n0, int n1,
                 hipStream_t stream) {
    cw::BwdParams<T> p{};
    p.g = make_geom<T>(a.grid);
    p.inv = (const T *)a.mat.inv;
    p.cell = (const T *)a.mat.cell;
    p.stiff = (const T *)a.mat.stiff;
    p.pml = make_layer<T>(a.pml, a.grid.ndim);
    p.lam = make_levels<T>(a.lam);
    p.u = make_levels<T>(a.u);
    p.reverse = a.reverse;
    p.rec = make_points<T>(a.sensors);
    p.adjoint = (const T *)a.adjoint_signal;
    p.src = make_points<T>(a.sources);
    p.signal = (const T *)a.signal;
    p.strip = make_strip<T>(a.strip);
    p.gmass = (T *)a.g_mass;
    p.gstiff = (T *)a.g_stiff;
    p.combine = a.combine;
    p.jm = (T)a.jac_mass;
    p.js = (T)a.jac_stiff;
    p.stiff_grad = !a.combine || a.jac_stiff != 0.0;
    p.skip_layer = a.skip_layer;
    p.mf = (T)a.mass_factor;
    p.steps = a.steps;
    const cw::Box box = stream_box(a.grid);
    p.region = band_of(a.grid, box);
    const int streams = NDIM == 2 ? box_blocks(box, a.grid.lanes) : 0;
    const int blocks = streams + general_blocks(a.grid, p.region);
    for (int n = n1 - 1; n >= n0; --n) {
      if (p.stiff_grad)
        hipLaunchKernelGGL((cw::backward_step<T, NDIM, R, true>), dim3(blocks),
                           dim3(cw::kThreads), 0, stream, p, box, n, streams);
      else
        hipLaunchKernelGGL((cw::backward_step<T, NDIM, R, false>),
                           dim3(blocks), dim3(cw::kThreads), 0, stream, p, box,
                           n, streams);
    }
    return hip_status("backward");
  }
};

template <typename T, int NDIM, int R> struct Superposed {
  static int run(const cw_superposed_args &a, hipStream_t stream) {
    const cw::Geom<T> g = make_geom<T>(a.grid);
    const int blocks = general_blocks(a.grid, cw::Region{});
    hipLaunchKernelGGL((cw::superposed_kernel<T, NDIM, R>), dim3(blocks),
                       dim3(cw::kThreads), 0, stream, g, (const T *)a.mat.inv,
                       (const T *)a.mat.cell, (const T *)a.u0, (const T *)a.u1,
                       (T *)a.u2, (T *)a.acc_mass, (T *)a.acc_stiff, (T)a.ft,
                       (T)a.fs, a.step);
    return hip_status("superposed_kernel");
  }
};

} // namespace

extern "C" {

int cw_version(void) { return CW_ABI_VERSION; }

const char *cw_last_error(void) { return g_error.c_str(); }

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
  if (!mat || !mat->stiff || !mat->cell || !mat->inv)
    return fail("prepare needs stiff, cell and inv");
  return dispatch<Prepare>(*grid, *grid, *mat, (hipStream_t)stream);
}

int cw_scalar_forward(const cw_forward_args *args, int32_t t0, int32_t t1,
                      void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->mat.inv || !args->mat.cell)
    return fail("forward needs the prepared inv and cell");
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
  if (args->strip.map && t0 < t1 &&
      (t0 < args->strip.first || t1 > args->strip.first + args->strip.rows))
    return fail("the strip record does not hold steps t0 .. t1");
  if (t0 == t1)
    return 0;
  return dispatch<Forward>(args->grid, *args, t0, t1, (hipStream_t)stream);
}

int cw_scalar_backward(const cw_backward_args *args, int32_t n0, int32_t n1,
                       void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->mat.inv || !args->mat.cell || !args->mat.stiff)
    return fail("backward needs stiff and the prepared inv and cell");
  if (int e = check_levels(args->lam, "lambda"))
    return e;
  if (int e = check_levels(args->u, "u"))
    return e;
  if (args->lam.slots != 2)
    return fail("lambda takes two slots");
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

int cw_scalar_superposed(const cw_superposed_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (args->grid.nbatch != 1 || args->grid.lanes != 1)
    return fail("the superposition step takes one shot a call");
  if (!args->mat.inv || !args->mat.cell || !args->u0 || !args->u1 ||
      !args->u2 || !args->acc_mass || !args->acc_stiff)
    return fail("superposed needs inv, cell, u0, u1, u2 and both accumulators");
  return dispatch<Superposed>(args->grid, *args, (hipStream_t)stream);
}

} // extern "C"
