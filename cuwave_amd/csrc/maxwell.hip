// The C ABI of include/cuwave_amd.h for the vector Maxwell family: argument
// checks, dispatch over (precision, dimension, radius, lanes a thread) and
// the launches of the pair. The point kernels are the elastic family's.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cuwave_amd.h"
#include "elastic.hpp"
#include "maxwell.hpp"

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

int check_grid(const cw_maxwell_grid *g) {
  if (!g)
    return fail("grid is NULL");
  if (g->ndim != 2 && g->ndim != 3)
    return fail("ndim must be 2 or 3: a single in-plane component has no curl");
  if (g->order != 2 && g->order != 4 && g->order != 6 && g->order != 8)
    return fail("space order must be 2, 4, 6 or 8");
  if (g->dtype != 0 && g->dtype != 1)
    return fail("dtype must be 0 (float32) or 1 (float64)");
  if (g->nbatch < 1)
    return fail("nbatch must be at least 1");
  if (g->lanes < g->nbatch || g->lanes > cw::kWave || (g->lanes & (g->lanes - 1)))
    return fail("lanes must be a power of two from nbatch to 64");
  if (g->conductor < 0 || g->conductor >= (1 << (2 * g->ndim)))
    return fail("conductor holds one bit a face");
  const int R = g->order / 2;
  int64_t nodes = 1;
  for (int d = 0; d < g->ndim; ++d) {
    if (g->n[d] < R + 3)
      return fail("every axis needs order / 2 + 3 nodes to image its taps");
    const int want = d == g->ndim - 1 ? (g->n[d] + 31) / 32 * 32 : g->n[d];
    if (g->np[d] != want)
      return fail("np must be n with the last axis padded to a multiple of 32");
    nodes *= g->np[d];
  }
  if (nodes * g->ndim * g->lanes >= (int64_t)1 << 31)
    return fail("the fields of the batch overflow the int32 element index");
  return 0;
}

template <typename T> cwm::MxGeom<T> make_geom(const cw_maxwell_grid &gr, const double *f) {
  cwm::MxGeom<T> g{};
  int64_t stride = 1;
  for (int d = 2; d >= 0; --d) {
    g.n[d] = d < gr.ndim ? gr.n[d] : 1;
    g.np[d] = d < gr.ndim ? gr.np[d] : 1;
    g.s[d] = (int)stride;
    stride *= g.np[d];
  }
  g.nodes = stride;
  g.lanes = gr.lanes;
  g.conductor = gr.conductor;
  g.interior = 1;
  for (int d = 0; d < gr.ndim; ++d)
    g.interior *= gr.n[d] - 2;
  for (int k = 0; k < 4; ++k)
    g.sg[k] = (T)gr.stag[k];
  g.nu = (T)gr.nu;
  for (int d = 0; d < 3; ++d)
    g.f[d] = d < gr.ndim ? (T)f[d] : (T)0;
  return g;
}

template <typename T, int NDIM, int R, int V>
int step_lanes(const cw_maxwell_step_args &a, hipStream_t stream) {
  const cwm::MxGeom<T> gc = make_geom<T>(a.grid, a.grid.inv_dx);
  const cwm::MxGeom<T> gu = make_geom<T>(a.grid, a.grid.step);
  const int64_t threads = (int64_t)gc.interior * (a.grid.lanes / V);
  const int blocks = (int)((threads + cwm::kThreads - 1) / cwm::kThreads);
  T *h = (T *)a.curl;
  const T *nu_pair = (const T *)a.nu_pair;
  if (a.g_nu)
    hipLaunchKernelGGL((cwm::mx_curl<T, NDIM, R, V, true, true>), dim3(blocks),
                       dim3(cwm::kThreads), 0, stream, gc, (const T *)a.u1, h, nu_pair,
                       (T *)a.g_nu, (const T *)a.v1);
  else if (nu_pair)
    hipLaunchKernelGGL((cwm::mx_curl<T, NDIM, R, V, true, false>), dim3(blocks),
                       dim3(cwm::kThreads), 0, stream, gc, (const T *)a.u1, h, nu_pair,
                       (T *)nullptr, (const T *)nullptr);
  else
    hipLaunchKernelGGL((cwm::mx_curl<T, NDIM, R, V, false, false>), dim3(blocks),
                       dim3(cwm::kThreads), 0, stream, gc, (const T *)a.u1, h, (const T *)nullptr,
                       (T *)nullptr, (const T *)nullptr);
  if (a.g_mass)
    hipLaunchKernelGGL((cwm::mx_update<T, NDIM, R, V, true>), dim3(blocks), dim3(cwm::kThreads),
                       0, stream, gu, (const T *)a.minv, (const T *)h, (const T *)a.u0,
                       (const T *)a.u1, (T *)a.u2, (T *)a.g_mass, (const T *)a.v1,
                       (T)a.mass_factor, (const T *)a.damping, (T)a.dt);
  else
    hipLaunchKernelGGL((cwm::mx_update<T, NDIM, R, V, false>), dim3(blocks), dim3(cwm::kThreads),
                       0, stream, gu, (const T *)a.minv, (const T *)h, (const T *)a.u0,
                       (const T *)a.u1, (T *)a.u2, (T *)nullptr, (const T *)nullptr, (T)0,
                       (const T *)a.damping, (T)a.dt);
  return hip_status("maxwell step");
}

template <typename T, int NDIM, int R>
int step_radius(const cw_maxwell_step_args &a, hipStream_t stream) {
  switch (cw::lanes_per_thread<T>(a.grid.lanes)) {
  case 1:
    return step_lanes<T, NDIM, R, 1>(a, stream);
  case 2:
    return step_lanes<T, NDIM, R, 2>(a, stream);
  default:
    if constexpr (sizeof(T) == 4)
      return step_lanes<T, NDIM, R, 4>(a, stream);
  }
  return fail("no maxwell kernel for this lane count");
}

int dispatch_step(const cw_maxwell_step_args &a, hipStream_t stream) {
  const cw_maxwell_grid &g = a.grid;
  const int R = g.order / 2;
#define CW_CASE(T, NDIM, RR)                                                   \
  if ((g.dtype == 1) == (sizeof(T) == 8) && g.ndim == NDIM && R == RR)         \
    return step_radius<T, NDIM, RR>(a, stream);
#define CW_RADII(T, NDIM)                                                      \
  CW_CASE(T, NDIM, 1) CW_CASE(T, NDIM, 2) CW_CASE(T, NDIM, 3) CW_CASE(T, NDIM, 4)
  CW_RADII(float, 2)
  CW_RADII(float, 3)
  CW_RADII(double, 2)
  CW_RADII(double, 3)
#undef CW_RADII
#undef CW_CASE
  return fail("no maxwell kernel for this precision, dimension and order");
}

template <typename T>
int launch_points(const cw_maxwell_grid &g, const cw_elastic_point_args &a, hipStream_t stream) {
  cwe::ElPoints p{a.starts, a.lin, a.shot, a.col, a.runs, a.row_stride, a.col_stride,
                  a.shot_stride};
  const int64_t threads = a.kind == 3 ? (int64_t)a.runs : (int64_t)a.runs * g.lanes;
  const int blocks = (int)((threads + cwe::kThreads - 1) / cwe::kThreads);
  hipLaunchKernelGGL((cwe::el_points<T>), dim3(blocks), dim3(cwe::kThreads), 0, stream, p,
                     a.kind, g.lanes, g.nbatch, a.row, (T *)a.field, (T *)a.table,
                     (const T *)a.weight, (T *)a.g_mass, (const T *)a.u1, (T)a.mass_factor);
  return hip_status("maxwell points");
}

} // namespace

extern "C" {

int cw_maxwell_step(const cw_maxwell_step_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->minv || !args->curl || !args->u0 || !args->u1 || !args->u2)
    return fail("a maxwell step needs minv, the curl scratch, u0, u1 and u2");
  if (args->u1 == args->u2 || args->u1 == args->curl || args->u2 == args->curl)
    return fail("u2 and the curl scratch must not alias u1 or each other");
  if (args->g_mass && !args->v1)
    return fail("the adjoint step needs the forward field v1");
  if (args->g_nu && (!args->g_mass || !args->nu_pair))
    return fail("the permeability density needs the adjoint step and nu_pair");
  return dispatch_step(*args, (hipStream_t)stream);
}

int cw_maxwell_points(const cw_maxwell_grid *grid, const cw_elastic_point_args *args,
                      void *stream) {
  if (int e = check_grid(grid))
    return e;
  if (!args)
    return fail("args is NULL");
  if (args->runs == 0)
    return 0;
  if (args->runs < 0 || !args->starts || !args->lin || !args->shot || !args->col ||
      !args->field || !args->table)
    return fail("points need starts, lin, shot, col, the field and the table");
  if (args->kind < 0 || args->kind > 3)
    return fail("kind is 0 (inject), 1 (record), 2 (replay) or 3 (adjoint load)");
  if ((args->kind == 0 || args->kind == 3) && !args->weight)
    return fail("an injection needs its weights");
  if (args->kind == 3 && (!args->g_mass || !args->u1))
    return fail("the adjoint load needs g_mass and u1");
  if (args->row < 0 || args->row_stride < 0 || args->col_stride < 0 || args->shot_stride < 0)
    return fail("row and strides must be non-negative");
  if (grid->dtype)
    return launch_points<double>(*grid, *args, (hipStream_t)stream);
  return launch_points<float>(*grid, *args, (hipStream_t)stream);
}

} // extern "C"
