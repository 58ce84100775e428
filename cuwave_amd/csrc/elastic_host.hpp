// Host-side pieces of the elastic C ABI shared by its translation units: the
// ABI structs turned into the kernels' parameter blocks, and the launchers
// one translation unit per (precision, radius) instantiates.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cuwave_amd.h"
#include "elastic.hpp"

namespace cwh {
int fail(const std::string &what);
} // namespace cwh

namespace cweh {

using cwh::fail;

inline int hip_status(const char *where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fail(std::string(where) + ": " + hipGetErrorString(e));
    return 2;
  }
  return 0;
}

inline int interior_nodes(const cw_elastic_grid &g) {
  int count = 1;
  for (int d = 0; d < g.ndim; ++d)
    count *= g.n[d] - 2;
  return count;
}

// blocks of one thread a node and V lanes
inline int node_blocks(const cw_elastic_grid &g, int V) {
  const int64_t threads = (int64_t)interior_nodes(g) * (g.lanes / V);
  return (int)((threads + cwe::kThreads - 1) / cwe::kThreads);
}

// the kernels' geometry with the per-axis factors `f` (1 / dx for the
// strains, dt^2 / dx for the update)
template <typename T>
cwe::ElGeom<T> make_elgeom(const cw_elastic_grid &gr, const T *rows, const double *f) {
  cwe::ElGeom<T> g{};
  const int nd = gr.ndim;
  int64_t stride = 1;
  for (int d = 2; d >= 0; --d) {
    g.n[d] = d < nd ? gr.n[d] : 1;
    g.np[d] = d < nd ? gr.np[d] : 1;
  }
  for (int d = nd - 1; d >= 0; --d) {
    g.s[d] = (int)stride;
    stride *= g.np[d];
  }
  for (int d = nd; d < 3; ++d)
    g.s[d] = 0;
  g.nodes = stride;
  g.lanes = gr.lanes;
  g.nbatch = gr.nbatch;
  g.clamped = gr.clamped;
  g.nc = gr.closure[0];
  g.hc = gr.closure[1];
  g.nspan = gr.closure[2];
  g.hspan = gr.closure[3];
  g.interior = interior_nodes(gr);
  for (int k = 0; k < 4; ++k)
    g.sg[k] = (T)gr.stag[k];
  g.lam = (T)gr.lame;
  g.mu = (T)gr.shear;
  for (int d = 0; d < 3; ++d)
    g.f[d] = d < nd ? (T)f[d] : (T)0;
  g.rows = rows;
  return g;
}

template <typename T> cwe::ElLayer<T> make_ellayer(const cw_elastic_layer &l, int ndim) {
  cwe::ElLayer<T> p{};
  for (int d = 0; d < 3; ++d) {
    const bool lined = d < ndim && l.lo[d] + l.hi[d] > 0;
    p.lo[d] = lined ? l.lo[d] : 0;
    p.hi[d] = lined ? l.hi[d] : 0;
    p.aux[d] = lined ? (T *)l.memory[d] : nullptr;
    p.profile[d] = lined ? (const T *)l.profile[d] : nullptr;
  }
  p.phase = l.phase;
  return p;
}

template <typename T, int NDIM, int R>
int elastic_step(const cw_elastic_step_args &a, hipStream_t stream);
template <typename T, int NDIM, int R>
int elastic_gradient(const cw_elastic_gradient_args &a, hipStream_t stream);

} // namespace cweh
