// Host-side pieces shared by the translation units of the library: the C
// ABI structs turned into the kernels' parameter blocks, the split of a 2D
// grid into the stream box and the wall band, and the step launchers, one
// translation unit per (precision, radius) so the library builds in
// parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cuwave_amd.h"
#include "scalar_kernels.hpp"
#include "stream.hpp"

namespace cwh {

// records the error text cw_last_error() returns; always returns 1
int fail(const std::string &what);

inline int hip_status(const char *where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fail(std::string(where) + ": " + hipGetErrorString(e));
    return 2;
  }
  return 0;
}

// The second stream a 2D sweep runs its edge launches on, beside the core
// launches on the caller's stream, and the two events that order them. A
// step's core and edge launches write disjoint nodes and read only levels
// the step before finished, so each waits only for the other's previous
// step. The stream is non-blocking, so the legacy null stream does not
// serialise it.
//
// Sets live for the process in a pool per device (scalar.hip): a sweep
// leases one for the time it takes to enqueue its launches and hands it
// back, so a process holds as many sets as it ever had sweeps enqueueing at
// once, whatever its host threads. Work a set's stream still holds from a
// previous lease is ordered before the next lease's, which only delays it.
// `cw_set_side_stream(0)` runs every launch on the caller's stream instead.
struct Side {
  hipStream_t stream = nullptr;
  hipEvent_t core = nullptr, edge = nullptr;
  int device = -1;
};

// a set of the current device (nullptr on error, the text in cw_last_error),
// or nullptr with *off set where side streams are switched off
Side *lease_side(bool *off);
void return_side(Side *s);

struct SideLease {
  Side *side = nullptr;
  bool off = false;
  explicit SideLease(bool want) {
    if (want)
      side = lease_side(&off);
  }
  ~SideLease() {
    if (side)
      return_side(side);
  }
  SideLease(const SideLease &) = delete;
  SideLease &operator=(const SideLease &) = delete;
};

// the end of a step: each stream's next launch waits for the other's last;
// after the final step the caller's stream holds every launch
inline void meet(Side &s, hipStream_t main) {
  hipEventRecord(s.core, main);
  hipEventRecord(s.edge, s.stream);
  hipStreamWaitEvent(main, s.edge, 0);
  hipStreamWaitEvent(s.stream, s.core, 0);
}

inline int lanes_shift(int lanes) {
  int k = 0;
  while ((1 << k) < lanes)
    ++k;
  return k;
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
  g.lshift = lanes_shift(gr.lanes);
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

// ---------------------------------------------------------- stream box
constexpr int kChunk = 32; // rows a streaming thread marches
static_assert(kChunk <= 32, "a chunk's rows with points are one 32-bit mask");

// the fewest blocks a stream box must have for the stream kernels to run
// (cw_set_stream_threshold); below it the general kernels take every node
int stream_threshold();

// the stream box of a 2D grid: the nodes order / 2 + 2 off the walls (out of
// reach of the wall mirror, the refill, the held nodes and the adjoint's
// reflected taps) that neither lie in a layer slab nor reach one with their
// transposed taps (order / 2 past the slab); empty on other grids, on grids
// too small for one and where two levels pass the 4 GiB a buffer resource
// addresses. The general kernels take every other node: the layers are
// thin, and there one independent thread a node hides the latency of the
// layer memory that a marching thread would wait on row after row.
inline cw::Box stream_box(const cw_grid &g, const cw_pml &pml) {
  cw::Box box{};
  const int R = g.order / 2, margin = R + 2;
  box.chunk = kChunk;
  if (g.ndim != 2)
    return box;
  const int64_t level = (int64_t)g.np[0] * g.np[1] * g.lanes * (g.dtype ? 8 : 4);
  if (2 * level >= ((int64_t)1 << 32))
    return box;
  int lo[2], hi[2];
  for (int d = 0; d < 2; ++d) {
    lo[d] = std::max(margin, pml.lo[d] ? pml.lo[d] + R + 1 : 0);
    hi[d] = std::min(g.n[d] - margin, pml.hi[d] ? g.n[d] - 2 - pml.hi[d] - R : g.n[d]);
  }
  if (hi[0] <= lo[0] || hi[1] <= lo[1])
    return box;
  box.y0 = lo[0];
  box.y1 = hi[0];
  box.x0 = lo[1];
  box.x1 = hi[1];
  // every block of the box is a core block: the second launch's stream
  // blocks are those holding sources or sensors
  const int C = cw::kThreads / g.lanes;
  const int xa = box.x0 / C * C;
  box.c0 = 0;
  box.c1 = (box.y1 - box.y0 + box.chunk - 1) / box.chunk;
  box.k0 = 0;
  box.k1 = (box.x1 - xa + C - 1) / C;
  // a box of few blocks marches few long chunks, each one row at a time: a
  // small grid steps faster with one independent thread a node everywhere
  if ((box.c1 - box.c0) * (box.k1 - box.k0) < stream_threshold())
    return cw::Box{0, 0, 0, 0, kChunk};
  return box;
}

// the stream blocks of a sweep, and of its core, each padded to a multiple
// of the XCD count (the second launch's general blocks then start on XCD 0)
inline int pad_xcds(int blocks) {
  return (blocks + cw::kXcds - 1) / cw::kXcds * cw::kXcds;
}
inline int box_blocks(const cw::Box &b, int lanes) {
  if (b.y0 >= b.y1)
    return 0;
  const int cols = cw::kThreads / lanes;
  const int xa = b.x0 / cols * cols;
  const int nchunk = (b.y1 - b.y0 + b.chunk - 1) / b.chunk;
  return pad_xcds(nchunk * ((b.x1 - xa + cols - 1) / cols));
}
inline int core_blocks(const cw::Box &b) {
  return pad_xcds((b.c1 - b.c0) * (b.k1 - b.k0));
}

// where the general kernels run: the band around the box, or every node
inline cw::Region band_of(const cw_grid &g, const cw::Box &b) {
  cw::Region r{};
  if (b.y0 >= b.y1)
    return r;
  r.band = 1;
  r.y0 = b.y0;
  r.y1 = b.y1;
  r.x0 = b.x0;
  r.x1 = b.x1;
  r.q = g.order / 2 + 1; // the box keeps R + 2 from every wall (stream_box)
  const int W = g.n[1] - 2;
  r.count = (b.y0 - 1) * W + (g.n[0] - 1 - b.y1) * W +
            (b.y1 - b.y0) * ((b.x0 - 1) + (g.n[1] - 1 - b.x1));
  return r;
}

// the general blocks of a sweep at V shot lanes a thread
inline int general_blocks(const cw_grid &gr, const cw::Region &r, int V) {
  const int per = cw::kThreads / (gr.lanes / V); // nodes a block takes
  int64_t nodes = r.band ? r.count : 1;
  if (!r.band)
    for (int d = 0; d < gr.ndim; ++d)
      nodes *= gr.n[d] - 2;
  int blocks = (int)((nodes + per - 1) / per);
  if (r.band && gr.ndim == 2) // whole rounds of band chunks, one chunk an XCD (general_node)
    blocks = (blocks + cw::kXcds * cw::kBandChunk - 1) /
             (cw::kXcds * cw::kBandChunk) * (cw::kXcds * cw::kBandChunk);
  return blocks;
}

// ------------------------------------------------------------ launchers
// forward steps t0 .. t1 - 1 and adjoint levels n1 - 1 .. n0, each the
// core launch (2D) and the frame launch; defined in steps.inc, one
// translation unit per (T, R)
template <typename T, int NDIM, int R>
int forward_steps(const cw_forward_args &a, int t0, int t1,
                  hipStream_t stream, const int *clock = nullptr,
                  int t_end = 0);
template <typename T, int NDIM, int R>
int backward_steps(const cw_backward_args &a, int n0, int n1,
                   hipStream_t stream);

} // namespace cwh
