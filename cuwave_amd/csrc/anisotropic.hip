// The C ABI of include/cuwave_amd.h for the cell-assembled elastic family:
// argument checks, dispatch over (precision, dimension, lanes a thread) and
// the launches. The point kernels are the elastic family's.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cuwave_amd.h"
#include "anisotropic.hpp"
#include "elastic.hpp"

namespace cwh {
int fail(const std::string &what);
} // namespace cwh

namespace {

using cwh::fail;

int hip_status(const char *where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fail(std::string(where) + ": " + hipGetErrorString(e));
    return 2;
  }
  return 0;
}

int check_grid(const cw_aniso_grid *g) {
  if (!g)
    return fail("grid is NULL");
  if (g->ndim < 1 || g->ndim > 3)
    return fail("ndim must be 1, 2 or 3");
  if (g->dtype != 0 && g->dtype != 1)
    return fail("dtype must be 0 (float32) or 1 (float64)");
  if (g->nbatch < 1)
    return fail("nbatch must be at least 1");
  if (g->lanes < g->nbatch || g->lanes > cw::kWave || (g->lanes & (g->lanes - 1)))
    return fail("lanes must be a power of two from nbatch to 64");
  int64_t nodes = 1;
  for (int d = 0; d < g->ndim; ++d) {
    if (g->n[d] < 4)
      return fail("every axis needs 4 nodes: two interior ones a cell");
    const int want = d == g->ndim - 1 ? (g->n[d] + 31) / 32 * 32 : g->n[d];
    if (g->np[d] != want)
      return fail("np must be n with the last axis padded to a multiple of 32");
    nodes *= g->np[d];
  }
  if (nodes * g->ndim * g->lanes >= (int64_t)1 << 31)
    return fail("the fields of the batch overflow the int32 element index");
  return 0;
}

template <typename T> cwa::AnGeom<T> make_geom(const cw_aniso_grid &gr) {
  cwa::AnGeom<T> g{};
  int64_t stride = 1;
  for (int d = 2; d >= 0; --d) {
    g.n[d] = d < gr.ndim ? gr.n[d] : 1;
    g.s[d] = (int)stride;
    stride *= d < gr.ndim ? gr.np[d] : 1;
  }
  g.nodes = stride;
  g.lanes = gr.lanes;
  g.interior = 1;
  for (int d = 0; d < gr.ndim; ++d)
    g.interior *= gr.n[d] - 2;
  g.f = (T)gr.step;
  return g;
}

int blocks_for(const cw_aniso_grid &g, int V) {
  int64_t interior = 1;
  for (int d = 0; d < g.ndim; ++d)
    interior *= g.n[d] - 2;
  const int64_t threads = interior * (g.lanes / V);
  return (int)((threads + cwa::kThreads - 1) / cwa::kThreads);
}

template <typename T, int NDIM, int V> int step_lanes(const cw_aniso_step_args &a, hipStream_t s) {
  hipLaunchKernelGGL((cwa::an_step<T, NDIM, V>), dim3(blocks_for(a.grid, V)),
                     dim3(cwa::kThreads), 0, s, make_geom<T>(a.grid), (const T *)a.table,
                     (const T *)a.minv, (const T *)a.cell, (const T *)a.u0, (const T *)a.u1,
                     (T *)a.u2, (const T *)a.damping, (T)a.dt);
  return hip_status("anisotropic step");
}

template <typename T, int NDIM, int V>
int gradient_lanes(const cw_aniso_gradient_args &a, hipStream_t s) {
  hipLaunchKernelGGL((cwa::an_gradient<T, NDIM, V>), dim3(blocks_for(a.grid, V)),
                     dim3(cwa::kThreads), 0, s, make_geom<T>(a.grid), (const T *)a.table,
                     (T *)a.g_mass, (T *)a.g_cell, (const T *)a.u0, (const T *)a.u1,
                     (const T *)a.u2, (const T *)a.l1, (T)a.mass_factor);
  return hip_status("anisotropic gradient");
}

// F<T, NDIM, V>(args, stream) for the grid's precision, dimension and lanes
#define CW_AN_DISPATCH(F, a, stream)                                           \
  do {                                                                         \
    const cw_aniso_grid &g = (a).grid;                                         \
    if (g.dtype == 1) {                                                        \
      const int V = cw::lanes_per_thread<double>(g.lanes);                     \
      CW_AN_NDIM(F, double, V, a, stream)                                      \
    } else {                                                                   \
      const int V = cw::lanes_per_thread<float>(g.lanes);                      \
      CW_AN_NDIM(F, float, V, a, stream)                                       \
      if (V == 4)                                                              \
        CW_AN_V(F, float, 4, a, stream)                                        \
    }                                                                          \
    return fail("no anisotropic kernel for this precision, dimension and lane count"); \
  } while (0)
#define CW_AN_NDIM(F, T, V, a, stream)                                         \
  if (V == 1)                                                                  \
    CW_AN_V(F, T, 1, a, stream)                                                \
  if (V == 2)                                                                  \
    CW_AN_V(F, T, 2, a, stream)
#define CW_AN_V(F, T, VV, a, stream)                                           \
  {                                                                            \
    if (g.ndim == 1)                                                           \
      return F<T, 1, VV>(a, stream);                                           \
    if (g.ndim == 2)                                                           \
      return F<T, 2, VV>(a, stream);                                           \
    if (g.ndim == 3)                                                           \
      return F<T, 3, VV>(a, stream);                                           \
  }

int dispatch_step(const cw_aniso_step_args &a, hipStream_t stream) {
  CW_AN_DISPATCH(step_lanes, a, stream);
}

int dispatch_gradient(const cw_aniso_gradient_args &a, hipStream_t stream) {
  CW_AN_DISPATCH(gradient_lanes, a, stream);
}

#undef CW_AN_V
#undef CW_AN_NDIM
#undef CW_AN_DISPATCH

template <typename T>
int launch_points(const cw_aniso_grid &g, const cw_elastic_point_args &a, hipStream_t stream) {
  cwe::ElPoints p{a.starts, a.lin, a.shot, a.col, a.runs, a.row_stride, a.col_stride,
                  a.shot_stride};
  const int64_t threads = a.kind == 3 ? (int64_t)a.runs : (int64_t)a.runs * g.lanes;
  const int blocks = (int)((threads + cwe::kThreads - 1) / cwe::kThreads);
  hipLaunchKernelGGL((cwe::el_points<T>), dim3(blocks), dim3(cwe::kThreads), 0, stream, p,
                     a.kind, g.lanes, g.nbatch, a.row, (T *)a.field, (T *)a.table,
                     (const T *)a.weight, (T *)a.g_mass, (const T *)a.u1, (T)a.mass_factor);
  return hip_status("anisotropic points");
}

} // namespace

extern "C" {

int cw_aniso_step(const cw_aniso_step_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->table || !args->minv || !args->cell || !args->u0 || !args->u1 || !args->u2)
    return fail("an anisotropic step needs the table, minv, cell, u0, u1 and u2");
  if (args->u1 == args->u2)
    return fail("u2 must not alias u1");
  return dispatch_step(*args, (hipStream_t)stream);
}

int cw_aniso_gradient(const cw_aniso_gradient_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->table || !args->g_mass || !args->g_cell || !args->u0 || !args->u1 || !args->u2 ||
      !args->l1)
    return fail("the gradient needs the table, both accumulators, u0, u1, u2 and l1");
  return dispatch_gradient(*args, (hipStream_t)stream);
}

int cw_aniso_points(const cw_aniso_grid *grid, const cw_elastic_point_args *args, void *stream) {
  if (int e = check_grid(grid))
    return e;
  if (!args)
    return fail("args is NULL");
  if (args->runs == 0)
    return 0;
  if (args->runs < 0 || !args->starts || !args->lin || !args->shot || !args->col ||
      !args->field || !args->table)
    return fail("points need starts, lin, shot, col, the field and the table");
  if (args->kind < 0 || args->kind > 2)
    return fail("kind is 0 (inject), 1 (record) or 2 (replay)");
  if (args->kind == 0 && !args->weight)
    return fail("an injection needs its weights");
  if (args->row < 0 || args->row_stride < 0 || args->col_stride < 0 || args->shot_stride < 0)
    return fail("row and strides must be non-negative");
  if (grid->dtype)
    return launch_points<double>(*grid, *args, (hipStream_t)stream);
  return launch_points<float>(*grid, *args, (hipStream_t)stream);
}

} // extern "C"
