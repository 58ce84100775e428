// The C ABI of include/cuwave_amd.h for the elastic family: argument checks,
// dispatch over (precision, dimension, radius), the material preparation and
// the point kernels. The step and gradient launchers live in
// elastic_*.hip, one translation unit per (precision, radius).
#include <hip/hip_runtime.h>

#include <string>

#include "elastic_host.hpp"

namespace {

using cweh::fail;
using cweh::hip_status;

int check_grid(const cw_elastic_grid *g) {
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
  if (g->lanes < g->nbatch || g->lanes > cw::kWave || (g->lanes & (g->lanes - 1)))
    return fail("lanes must be a power of two from nbatch to 64");
  const int *c = g->closure;
  if (c[0] < 1 || c[1] < 1 || c[0] > 12 || c[1] > 12 || c[2] < 1 || c[3] < 1 || c[2] > 16 ||
      c[3] > 16)
    return fail("closure counts: 1 to 12 rows, spans of 1 to 16 points");
  int64_t nodes = 1;
  for (int d = 0; d < g->ndim; ++d) {
    const int band = c[0] > c[1] ? c[0] : c[1];
    const int span = c[2] > c[3] ? c[2] : c[3];
    if (g->n[d] < 2 * band + 3 || g->n[d] < span + 3)
      return fail("every axis needs room for both walls' closures");
    const int want = d == g->ndim - 1 ? (g->n[d] + 31) / 32 * 32 : g->n[d];
    if (g->np[d] != want)
      return fail("np must be n with the last axis padded to a multiple of 32");
    nodes *= g->np[d];
  }
  const int nvoigt = g->ndim * (g->ndim + 1) / 2;
  if (nodes * nvoigt * g->lanes >= (int64_t)1 << 31)
    return fail("the stresses of the batch overflow the int32 element index");
  return 0;
}

int check_layer(const cw_elastic_grid &g, const cw_elastic_layer &p) {
  if (p.phase < 0 || p.phase > 2)
    return fail("phase is 0 (forward), 1 or 2 (adjoint)");
  const int R = g.order / 2;
  for (int d = 0; d < g.ndim; ++d) {
    if (p.lo[d] < 0 || p.hi[d] < 0)
      return fail("pml widths must be non-negative");
    if (!(p.lo[d] + p.hi[d]))
      continue;
    if ((p.lo[d] ? p.lo[d] + R : 0) + (p.hi[d] ? p.hi[d] + R : 0) >= g.n[d] - 2)
      return fail("the pml layers and their reach leave no interior");
    if (!p.profile[d] || !p.memory[d])
      return fail("a lined axis needs its profile and memory");
  }
  return 0;
}

template <template <typename, int, int> class F, typename... A>
int dispatch(const cw_elastic_grid &g, A &&...args) {
  const int R = g.order / 2;
#define CW_CASE(T, NDIM, RR)                                                   \
  if ((g.dtype == 1) == (sizeof(T) == 8) && g.ndim == NDIM && R == RR)         \
    return F<T, NDIM, RR>::run(args...);
#define CW_RADII(T, NDIM)                                                      \
  CW_CASE(T, NDIM, 1) CW_CASE(T, NDIM, 2) CW_CASE(T, NDIM, 3) CW_CASE(T, NDIM, 4)
  CW_RADII(float, 1)
  CW_RADII(float, 2)
  CW_RADII(float, 3)
  CW_RADII(double, 1)
  CW_RADII(double, 2)
  CW_RADII(double, 3)
#undef CW_RADII
#undef CW_CASE
  return fail("no elastic kernel for this precision, dimension and order");
}

template <typename T, int NDIM, int R> struct Prepare {
  static int run(const cw_elastic_grid &gr, const cw_elastic_material &m, hipStream_t stream) {
    const cwe::ElGeom<T> g = cweh::make_elgeom<T>(gr, (const T *)m.rows, gr.inv_dx);
    const int blocks = cweh::node_blocks(gr, gr.lanes);
    hipLaunchKernelGGL((cwe::el_prepare<T, NDIM>), dim3(blocks), dim3(cwe::kThreads), 0, stream,
                       g, (const T *)m.gamma, (T)gr.density, (T *)m.prep);
    return hip_status("el_prepare");
  }
};

template <typename T, int NDIM, int R> struct Step {
  static int run(const cw_elastic_step_args &a, hipStream_t stream) {
    return cweh::elastic_step<T, NDIM, R>(a, stream);
  }
};

template <typename T, int NDIM, int R> struct Gradient {
  static int run(const cw_elastic_gradient_args &a, hipStream_t stream) {
    return cweh::elastic_gradient<T, NDIM, R>(a, stream);
  }
};

template <typename T>
int launch_points(const cw_elastic_grid &g, const cw_elastic_point_args &a, hipStream_t stream) {
  cwe::ElPoints p{a.starts, a.lin, a.shot, a.col, a.runs, a.row_stride, a.col_stride,
                  a.shot_stride};
  const int64_t threads = a.kind == 3 ? (int64_t)a.runs : (int64_t)a.runs * g.lanes;
  const int blocks = (int)((threads + cwe::kThreads - 1) / cwe::kThreads);
  hipLaunchKernelGGL((cwe::el_points<T>), dim3(blocks), dim3(cwe::kThreads), 0, stream, p,
                     a.kind, g.lanes, g.nbatch, a.row, (T *)a.field, (T *)a.table,
                     (const T *)a.weight, (T *)a.g_mass, (const T *)a.u1, (T)a.mass_factor);
  return hip_status("el_points");
}

} // namespace

extern "C" {

int cw_elastic_prepare(const cw_elastic_grid *grid, const cw_elastic_material *mat,
                       void *stream) {
  if (int e = check_grid(grid))
    return e;
  if (!mat || !mat->gamma || !mat->prep)
    return fail("prepare needs gamma and prep");
  if (!(grid->density > 0.0))
    return fail("the background density must be positive");
  return dispatch<Prepare>(*grid, *grid, *mat, (hipStream_t)stream);
}

int cw_elastic_step(const cw_elastic_step_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->mat.gamma || !args->mat.prep || !args->mat.rows)
    return fail("a step needs gamma, the prepared material and the closure rows");
  if (!args->sigma || !args->u0 || !args->u1 || !args->u2)
    return fail("a step needs sigma, u0, u1 and u2");
  if (int e = check_layer(args->grid, args->pml))
    return e;
  if (args->g_mass && (!args->g_normal || !args->v1 || (args->grid.ndim > 1 && !args->g_shear)))
    return fail("the adjoint pair needs every accumulator and the forward field");
  if (args->sigma == args->u1 || args->sigma == args->u2 || args->u1 == args->u2)
    return fail("sigma, u1 and u2 must be distinct");
  return dispatch<Step>(args->grid, *args, (hipStream_t)stream);
}

int cw_elastic_gradient(const cw_elastic_gradient_args *args, void *stream) {
  if (!args)
    return fail("args is NULL");
  if (int e = check_grid(&args->grid))
    return e;
  if (!args->rows || !args->g_mass || !args->g_normal ||
      (args->grid.ndim > 1 && !args->g_shear))
    return fail("the gradient needs the closure rows and every accumulator");
  if (!args->u0 || !args->u1 || !args->u2 || !args->l1)
    return fail("the gradient needs u0, u1, u2 and l1");
  return dispatch<Gradient>(args->grid, *args, (hipStream_t)stream);
}

int cw_elastic_points(const cw_elastic_grid *grid, const cw_elastic_point_args *args,
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
