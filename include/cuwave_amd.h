/* cuwave_amd: the C ABI of the MI355X (gfx950) finite-difference path.
 *
 * This library replaces the reference's kernel ABI for the pressure family:
 * `compile_kernels(sim)` returning a module whose `get_function(name)` is
 * launched as `kernel(grid, block, args)` with raw device pointers and
 * by-value scalars (cuwave/wave.py:637-647, cuwave/cpu.py:85-153; the CPU
 * fallback's C form is `launch_<name>(gx, gy, gz, bx, by, bz, void **args)`,
 * cuwave/kernels/cpu.hpp:93-98). One entry point here stands for a whole
 * step sequence of several of those kernels, fused, over a batch of shots.
 *
 * Conventions:
 * - Every pointer is device memory owned by the caller; nothing allocates.
 * - Per-node arrays (materials, cell means, gradients, maps) are
 *   C-contiguous over the padded grid, the fastest axis padded to a multiple
 *   of 32 (cuwave/wave.py:120-122).
 * - Per-shot arrays carry the shots innermost: a field level is
 *   (*np, lanes), `lanes` a power of two from nbatch to 64, the lanes past
 *   nbatch padding the caller zeroes; the layer memories, the strip record,
 *   the sensor record and the adjoint signal end in a lanes axis too. One
 *   node's shots are then one contiguous run, which the kernels read and
 *   write whole.
 * - `dtype` 0 is float32, 1 is float64; every field of a call has it.
 * - Launches are asynchronous on `stream` (a hipStream_t; 0 the null stream).
 *   A 2D adjoint sweep also launches its frame kernels on a non-blocking
 *   stream of the library's own, ordered after `stream`'s earlier work;
 *   `stream` waits for them before its next work, so to the caller the call
 *   stays one stream's. Those streams and their events come from a pool per
 *   device that lives for the process: a call leases one set while it
 *   enqueues and returns it, so the pool holds as many sets as calls ever
 *   enqueued at once. A call made while `stream` is being captured into a
 *   graph should find a set in the pool (make one uncaptured call first) or
 *   switch side streams off with cw_set_side_stream(0).
 * - A return of 0 is success; anything else is an error whose text
 *   `cw_last_error()` returns (thread-local). Arguments are checked before
 *   any launch, so an error leaves the device untouched.
 * - Calls on one stream are ordered; calls on different streams that share
 *   an output race.
 */
#ifndef CUWAVE_AMD_H
#define CUWAVE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CW_ABI_VERSION 15

/* the grid and the operator of one pressure simulation (cuwave/wave.py:479-631,
 * cuwave/scalar.py:140-160): the per-axis step factors and the full-radius
 * cell-flux row (cuwave/stencils.py:186-213) */
typedef struct cw_grid {
  int32_t ndim;      /* 1, 2 or 3 */
  int32_t order;     /* space order: 2, 4, 6 or 8 */
  int32_t n[3];      /* logical extents, ghosts included (Simulation.Nx) */
  int32_t np[3];     /* padded extents (Simulation.Nx_padded) */
  int32_t nbatch;    /* shots one call advances */
  int32_t lanes;     /* innermost extent of the per-shot arrays: a power of
                        two, nbatch <= lanes <= 64 */
  int32_t dtype;     /* 0 float32, 1 float64 */
  int32_t dirichlet; /* faces (bit 2 axis + side) held odd, Dirichlet */
  int32_t neumann;   /* faces refilled even, Neumann */
  double op[4];      /* op[k - 1]: the cell-flux weight k of the full radius */
  double f[3];       /* per-axis step factors (Simulation.step_factors) */
} cw_grid;

/* the material the step reads (cuwave/scalar.py:407-420, 739-745) */
typedef struct cw_materials {
  const void *stiff; /* nodal stiffness, its ghosts mirrored */
  const void *minv;  /* nodal inverse inertia, or NULL to derive 1 / stiff */
  void *prep;        /* (ndim + 1, field), from prepare: the harmonic cell
                        stiffness of every axis, then the inverse inertia */
  /* a nodal damping field d, or NULL for a lossless step: the update becomes
   * (2 u1 - u0 (1 - beta) + minv L u1) / (1 + beta), beta = minv d dt / 2
   * (cuwave/kernels/scalar.cu:455-459), the adjoint likewise; the 2D stream
   * kernels are left out (every node runs the general kernels) */
  const void *damping;
  /* the field the reverse march steps with (the damping negated where the
   * march runs backwards, cuwave/sensitivity.py:592-598), or NULL */
  const void *damping_reverse;
  double dt;
  /* a domain mask (cuwave/wave.py:661-726), or NULL for every interior node
   * stepped. Per node: 0 where the step leaves it as it is (outside
   * `domain`, or held at zero by `dirichlet`); 1 << 30 where every cell
   * keeps the full radius; else, per axis d, the radii rp and rm of its cells
   * to the next and to the previous node packed as (rp | rm << 3) << 6 d,
   * 0 for a closed cell (no flux) and 7 for a cell open onto a held node.
   * Forward only: the adjoint, gradient and superposition entry points
   * refuse it (the reference's gradients refuse a mask too,
   * cuwave/sensitivity.py:121-126), and so does a pml or a strip. The 2D
   * stream kernels are left out (every node runs the general kernels). */
  const int32_t *cells;
  /* rows[r - 1][k - 1]: weight k of a radius-r cell (cuwave/stencils.py:
   * 186-200), the rows a graded cell of `cells` reads */
  double rows[4][4];
} cw_materials;

/* the convolutional PML behind the faces (cuwave/boundary.py:272-376) */
typedef struct cw_pml {
  int32_t lo[3], hi[3];     /* layer rows behind the low / high face, 0 none */
  const void *profile[3];   /* (4, n[d]): a, b at nodes and at half points */
  void *aux[3];             /* forward memory: (3, slab nodes, lanes) */
  void *eta[3][2];          /* adjoint memory, two buffers alternating, each
                               (5, slab nodes, lanes): the three recursion
                               variables, then the two tap fields the
                               neighbours' transposed taps read; zeroed */
} cw_pml;

/* points a step injects at or records from, bucketed by the library's tiles
 * (cw_point_lists); entry e at step row r for shot b reads or writes element
 * r row_stride + col[e] col_stride + b shot_stride of the signal / record */
typedef struct cw_points {
  const int32_t *starts; /* (tiles + 1) CSR offsets, NULL for no points */
  const int32_t *lin;    /* flat node index within one shot's field */
  const int32_t *shot;   /* the shot of the entry, -1 for every shot */
  const int32_t *col;    /* column in the signal / record row */
  const void *weight;    /* injection weight, NULL for a record */
  int64_t row_stride;
  int64_t col_stride;
  int64_t shot_stride;   /* 0 when one table serves every shot */
} cw_points;

/* the boundary strip a reverse march replays (cuwave/sensitivity.py:160-195) */
typedef struct cw_strip {
  const int32_t *map; /* (field) strip column of a node, -1 off the strip */
  void *record;       /* (rows, count, lanes): row t - first holds step t */
  int32_t count;      /* strip nodes */
  int32_t first;      /* the step of record row 0 */
  int32_t rows;       /* record rows a shot holds */
  int32_t bf16;       /* 1: record stored as bfloat16 bits (float32 only) */
} cw_strip;

/* time levels in slots: level k sits in slot (k + offset) mod slots, each
 * slot a (*np, lanes) level; slot_stride in elements */
typedef struct cw_levels {
  void *base;
  int64_t slot_stride;
  int32_t slots;
  int32_t offset;
} cw_levels;

/* forward steps t0 .. t1 - 1 of a batch (cuwave/wave.py:1071-1163): per
 * step the flux-divergence update (cuwave/kernels/scalar.cu:346-466), the
 * CPML memories of every lined axis (:499-532), the source injection
 * (cuwave/kernels/common.cuh:456-464), the ghost refill (scalar.cu:535-555),
 * the sensor record (common.cuh:483-490) and the strip record
 * (cuwave/sensitivity.py:608-633), fused into one kernel. */
typedef struct cw_forward_args {
  cw_grid grid;
  cw_materials mat;
  cw_pml pml;            /* lo = hi = 0 on every axis for no layer */
  cw_levels u;           /* the field levels */
  cw_points sources;     /* signal rows are steps */
  const void *signal;
  cw_points sensors;     /* record rows are steps */
  void *record;
  cw_strip strip;        /* strip.map NULL for none */
  int32_t steps;         /* N, the rows of the signal and record tables */
  /* the illumination map, NULL for none: per node, each step adds the sum
   * over the shot lanes of u^2, u the step's final field (no reference
   * counterpart; the closed form sum_t sum_s u^2) */
  void *illum;
} cw_forward_args;

/* adjoint steps over levels n = n1 - 1 down to n0 (cuwave/sensitivity.py:
 * 384-497, 500-750): per level the exact transpose W^-1 A^T W of the step
 * with both gradient densities, the inertia one by parts in time
 * (cuwave/kernels/scalar_sensitivity.cu:443-515), the transposed CPML
 * (:517-597), the adjoint load with its share of the inertia gradient
 * (cuwave/kernels/common.cuh:467-480), and, when `reverse` is set, the
 * reverse march of the forward field one level ahead with the source
 * re-injected, the strip replayed and the ghosts refilled
 * (cuwave/sensitivity.py:678-689), fused into one kernel. The reverse runs
 * at levels n <= steps - 2, writing u^(n - 1) over u^(n + 1) and replaying
 * the strip row of step n - 1 (zero rows stand for steps before 0). */
typedef struct cw_backward_args {
  cw_grid grid;
  cw_materials mat;
  cw_pml pml;
  cw_levels lam;           /* adjoint levels, two slots */
  cw_levels u;             /* forward levels: two slots reversed, or history */
  int32_t reverse;         /* 1: rebuild u^(n - 1) from u^n, u^(n + 1) */
  cw_points sensors;       /* adjoint load */
  const void *adjoint_signal; /* row N - 1 - n drives level n */
  cw_points sources;       /* re-injected by the reverse march */
  const void *signal;
  cw_strip strip;
  void *g_mass, *g_stiff;  /* per-node accumulators, the shots summed
                              (cuwave/scalar.py:751-756) */
  int32_t combine;         /* 1: g_mass += jm gm + js gs (g_stiff unused) */
  double jac_mass, jac_stiff;
  int32_t skip_layer;      /* 1: no gradient inside the layer (zeroed after) */
  double mass_factor;      /* 1 / dt^2 (cuwave/wave.py:597-599) */
  int32_t steps;           /* N */
} cw_backward_args;

/* the superposition step (cuwave/kernels/scalar_sensitivity.cu:599-668) and
 * the unfused Frechet densities (:352-385), one call a step, one shot */
typedef struct cw_superposed_args {
  cw_grid grid;
  cw_materials mat;
  const void *u0, *u1;
  void *u2;
  void *acc_mass, *acc_stiff;
  double ft, fs;
  int32_t step; /* 1: superposed step; 0: Frechet densities of (u0, u1, u2) */
} cw_superposed_args;

/* the gradient densities of a forward triplet (u0, u1, u2) against an adjoint
 * field l1 (cuwave/kernels/scalar_sensitivity.cu:303-350 `gradient_kernel`),
 * subtracted from the accumulators; one shot (grid.lanes 1). `factor` are the
 * step factors over dt^2 (cuwave/scalar.py:768-784), `inv_dt2` 1 / dt^2 */
typedef struct cw_gradient_args {
  cw_grid grid;
  cw_materials mat;
  const void *u0, *u1, *u2, *l1;
  void *g_mass, *g_stiff;
  double inv_dt2;
  double factor[3];
} cw_gradient_args;

int cw_version(void);
const char *cw_last_error(void);

/* 0: every launch of a sweep goes on the caller's stream (the frame launches
 * of a 2D adjoint level then follow its core launch instead of running beside
 * it; the results are the same bits); 1 (the default): side streams as
 * above. Process-wide; returns the previous setting. No reference
 * counterpart: the reference's graphs and side stream are its own host
 * code's (cuwave/wave.py:929-936, cuwave/sensitivity.py:670-689). */
int cw_set_side_stream(int32_t enabled);

/* The fewest blocks the stream box of a 2D sweep must hold for the stream
 * kernels to march it (default 512, two a CU); a smaller box, as on small
 * grids or with few lanes, leaves every node to the general kernels, one
 * independent thread a node (the same bits either way). Process-wide;
 * returns the previous setting; 0 marches every box. No reference
 * counterpart: the reference picks its tiers from device attributes
 * (cuwave/scalar.py:422-737). */
int cw_set_stream_threshold(int32_t blocks);

/* the tiles a step kernel walks, for bucketing points */
int cw_tile_count(const cw_grid *grid, int32_t *tiles);

/* bucket `count` points by tile into CSR: `starts` (tiles + 1), `order`
 * (count) the point index in bucketed order, stable; host memory */
int cw_point_lists(const cw_grid *grid, const int32_t *lin, int32_t count,
                   int32_t *starts, int32_t *order);

/* the inverse inertia and the harmonic cell stiffness of every axis
 * (cuwave/kernels/scalar.cu:48-64 `sc sp / (sc + sp)`, :454) */
int cw_scalar_prepare(const cw_grid *grid, const cw_materials *mat,
                      void *stream);

int cw_scalar_forward(const cw_forward_args *args, int32_t t0, int32_t t1,
                      void *stream);

/* A window of `steps` forward steps whose step index comes from device
 * memory: step k of the window is *clock + k, and after the window *clock
 * grows by `steps`. A window's launches are then the same on every call, so
 * a caller can capture one call in a graph and replay it for every window of
 * a march (what the reference does with its 24-step graphs,
 * cuwave/wave.py:890-952; here without re-capturing). Steps the clock puts
 * outside 0 .. t_end - 1 are skipped by the kernels; the buffers are checked
 * for the whole range 0 .. t_end as cw_scalar_forward checks t0 .. t1.
 * Nothing here allocates or sets memory. */
int cw_scalar_forward_clocked(const cw_forward_args *args, int32_t *clock,
                              int32_t steps, int32_t t_end, void *stream);

int cw_scalar_backward(const cw_backward_args *args, int32_t n0, int32_t n1,
                       void *stream);

/* No counterpart in the reference: cw_scalar_forward for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The CPML memories, the only workspace of
 * a forward sweep, come from a scratch array the library holds on every GPU
 * it runs on, CW_STATIC_SCRATCH_BYTES of it, until the process ends;
 * `args->pml.aux` is ignored. A call with t0 == 0 clears the memories first;
 * a call with t0 > 0 continues from what the previous call on that GPU left
 * in them. Calls on one GPU share the array, so calls that overlap on one GPU
 * (two streams, two host threads, or one sweep split into calls with another
 * simulation's calls between them) race. A sweep whose memories need more
 * than CW_STATIC_SCRATCH_BYTES (the sum over lined axes of 3 x slab nodes x
 * lanes x element size, each axis rounded up to 256 bytes) is refused before
 * any launch. Nothing here allocates: no hipMalloc, no library allocation. */
#define CW_STATIC_SCRATCH_BYTES (64u << 20)
int cw_scalar_forward_static(const cw_forward_args *args, int32_t t0,
                             int32_t t1, void *stream);

int cw_scalar_superposed(const cw_superposed_args *args, void *stream);

int cw_scalar_gradient(const cw_gradient_args *args, void *stream);

/* ------------------------------------------------------------ elastic
 * The staggered isotropic elastic family (cuwave/elastic.py:154-794): the
 * displacement component c half a node up axis c, gamma scaling density and
 * stiffness alike. Fields are (ncomp, *np, lanes), ncomp = ndim; the stress
 * scratch (nvoigt, *np, lanes), nvoigt = ndim (ndim + 1) / 2; the gradient
 * accumulators per node: mass (ncomp, *np), normal (*np), shear
 * (nvoigt - ndim, *np). */

/* the grid, operator and material constants (cuwave/elastic.py:154-345;
 * the closure of the staggered SBP pair, cuwave/stencils.py:60-213) */
typedef struct cw_elastic_grid {
  int32_t ndim;       /* 1, 2 or 3 */
  int32_t order;      /* space order: 2, 4, 6 or 8 */
  int32_t n[3];       /* logical extents, ghosts included */
  int32_t np[3];      /* padded extents */
  int32_t nbatch, lanes, dtype;
  int32_t clamped;    /* faces held at zero (Clamped), bit 2 axis + side; the
                         others are traction-free */
  int32_t closure[4]; /* node rows, half rows, node span, half span */
  double stag[4];     /* the full-radius staggered weights, stag[k - 1] */
  double lame, shear, density;
  double inv_dx[3];   /* the strain factors 1 / dx */
  double step[3];     /* the update factors dt^2 / dx */
} cw_elastic_grid;

/* the material: gamma and what cw_elastic_prepare makes of it */
typedef struct cw_elastic_material {
  const void *gamma; /* nodal design field */
  void *prep;        /* (nvoigt, field): the inverse inertia of each
                        component point (rows 0 .. ndim - 1), then the shear
                        modulus scale of each pair point (the pair's stress
                        row); zeroed before prepare */
  const void *rows;  /* the closure rows in the grid's dtype: node rows
                        (closure[0] x closure[2]) then half rows
                        (closure[1] x closure[3]) */
} cw_elastic_material;

/* the PML folded into both launches of the pair (cuwave/elastic.py:597-612) */
typedef struct cw_elastic_layer {
  int32_t lo[3], hi[3];   /* layer rows behind the low / high face, 0 none */
  const void *profile[3]; /* (4, n[d]): a, b at nodes and at half points */
  void *memory[3];        /* (rows, slab nodes, lanes), zeroed: 2 ndim rows
                             forward (strain, then divergence memory), 3 ndim
                             in the adjoint (the strain memory twice) */
  int32_t phase;          /* 0 forward; 1 or 2 adjoint, reading its strain
                             memory off row 0 or 2 ndim, alternating a step */
} cw_elastic_layer;

/* one step of the pair over a batch (cuwave/kernels/elastic.cu:754-853):
 * the stresses of u1 into sigma, then u2 = -u0 + 2 u1 + minv div(sigma);
 * u2 may be u0. With g_mass set it is the adjoint pair: both gradient
 * densities of u1 against the forward field v1, the shots summed into the
 * accumulators (cuwave/elastic.py:436-488). */
typedef struct cw_elastic_step_args {
  cw_elastic_grid grid;
  cw_elastic_material mat;
  cw_elastic_layer pml;
  void *sigma;
  const void *u0, *u1;
  void *u2;
  void *g_mass, *g_normal, *g_shear; /* NULL for a forward step */
  const void *v1;
  double mass_factor;                /* 1 / dt^2 (cuwave/wave.py:597-599) */
} cw_elastic_step_args;

/* the densities of a forward triplet against the adjoint l1
 * (cuwave/kernels/elastic_sensitivity.cu:115-158) */
typedef struct cw_elastic_gradient_args {
  cw_elastic_grid grid;
  const void *rows;
  void *g_mass, *g_normal, *g_shear;
  const void *u0, *u1, *u2, *l1;
  double mass_factor;
} cw_elastic_gradient_args;

/* points of a field: `starts` (runs + 1) CSR groups the entries into runs of
 * one node each (a run's entries keep their order); entry e touches element
 * lin[e] lanes + b of the field for shot lane b (every lane when shot[e] < 0)
 * and element row row_stride + col[e] col_stride + b shot_stride of the
 * table. kind 0 injects (field += weight table, cuwave/kernels/common.cuh:
 * 456-464), 1 records (table = field, :483-490), 2 replays (field = table,
 * :493-503), 3 injects the adjoint load and its inertia density
 * (g_mass[lin] += -mass_factor u1 load, :467-480). */
typedef struct cw_elastic_point_args {
  const int32_t *starts, *lin, *shot, *col;
  int32_t runs;
  int32_t kind;
  int64_t row_stride, col_stride, shot_stride;
  int64_t row;
  void *field;
  void *table;
  const void *weight;
  void *g_mass;
  const void *u1;
  double mass_factor;
} cw_elastic_point_args;

/* the prepared material (point inverse inertia, cuwave/kernels/elastic.cu:
 * 200-211; pair shear scale, :189-197) of every interior node */
int cw_elastic_prepare(const cw_elastic_grid *grid, const cw_elastic_material *mat,
                       void *stream);
int cw_elastic_step(const cw_elastic_step_args *args, void *stream);
int cw_elastic_gradient(const cw_elastic_gradient_args *args, void *stream);
int cw_elastic_points(const cw_elastic_grid *grid, const cw_elastic_point_args *args,
                      void *stream);

/* ------------------------------------------------------------ Maxwell
 * The vector Maxwell family: the Yee lattice in second-order curl-curl form
 * (cuwave/maxwell.py:162-260, cuwave/kernels/maxwell.cu:44-275). Electric
 * component c sits half a node up axis c, the curl of the pair (k, l) half a
 * node up both. Per-shot arrays are (component, *np, lanes); the curl
 * scratch (npairs, *np, lanes), npairs 1 in 2D and 3 in 3D (pair (k, l) at
 * row 3 - k - l). */

/* the grid and the operator constants (MaxwellWave's compile flags and
 * `_pair_args`, cuwave/maxwell.py:185-190, 322-340) */
typedef struct cw_maxwell_grid {
  int32_t ndim;      /* 2 or 3 */
  int32_t order;     /* space order: 2, 4, 6 or 8 */
  int32_t n[3];      /* logical extents, ghosts included */
  int32_t np[3];     /* padded extents */
  int32_t nbatch, lanes, dtype;
  int32_t conductor; /* Conductor faces, bit 2 axis + side; the others are
                        Magnetic */
  double stag[4];    /* the full-radius staggered weights, stag[k - 1] */
  double nu;         /* 1 / permeability */
  double inv_dx[3];  /* the curl factors 1 / dx */
  double step[3];    /* the update factors dt^2 / dx */
} cw_maxwell_grid;

/* one step over a batch (curl_kernel then fd_kernel, cuwave/kernels/
 * maxwell.cu:209-274): nu W curl(u1) into `curl`, then u2 = -u0 + 2 u1 +
 * minv force; u2 may be u0. With g_mass set it is the adjoint step: the
 * inertia density of u1 against the forward v1, the shots summed into the
 * per-point accumulator (ncomp, *np) (the GRADIENT build, :268-272). */
typedef struct cw_maxwell_step_args {
  cw_maxwell_grid grid;
  const void *minv; /* (ncomp, *np) point inverse inertia */
  /* (npairs, *np) nu W of each pair point for a designed permeability (the
   * USE_MAGNETIC build, cuwave/maxwell.py:264-270), or NULL for grid.nu and
   * the wall weights */
  const void *nu_pair;
  /* (npairs, *np) with g_mass and nu_pair set: the permeability density, the
   * adjoint curl against the forward one (maxwell.cu:229-231), or NULL */
  void *g_nu;
  void *curl;
  const void *u0, *u1;
  void *u2;
  void *g_mass;     /* NULL for a forward step */
  const void *v1;
  double mass_factor; /* 1 / dt^2 (cuwave/wave.py:597-599) */
  /* a nodal damping field d, or NULL: the update becomes (2 u1 - u0 (1 - beta)
   * + minv force) / (1 + beta), beta = minv d dt / 2 (the USE_DAMPING build,
   * maxwell.cu:257-264); the adjoint step is the same damped step */
  const void *damping;
  double dt;
} cw_maxwell_step_args;

int cw_maxwell_step(const cw_maxwell_step_args *args, void *stream);
/* points of a Maxwell field, as cw_elastic_points */
int cw_maxwell_points(const cw_maxwell_grid *grid, const cw_elastic_point_args *args,
                      void *stream);

/* ------------------------------------------------------------ cell-assembled elastic
 * The anisotropic elastic family: collocated displacement components, the
 * operator gathered cell by cell from one cell's stencil K = -B^T C B
 * (cuwave/anisotropic.py:1-160, cuwave/kernels/anisotropic.cu:72-397),
 * order 2. Per-shot arrays are (ndim, *np, lanes); K is (nloc, nloc),
 * nloc = 2^ndim ndim, row corner ndim + component, corner c having bit d set
 * on the plus side of axis d. */

typedef struct cw_aniso_grid {
  int32_t ndim;   /* 1, 2 or 3 */
  int32_t n[3];   /* logical extents, ghosts included */
  int32_t np[3];  /* padded extents */
  int32_t nbatch, lanes, dtype;
  double step;    /* the update factor dt^2 (the spacing sits in K) */
} cw_aniso_grid;

/* one step over a batch (fd_kernel, anisotropic.cu:367-397): the
 * cell-gathered force on every component, u2 = -u0 + 2 u1 + dt^2 minv force;
 * u2 may be u0 */
typedef struct cw_aniso_step_args {
  cw_aniso_grid grid;
  const void *table; /* K in the grid's dtype */
  const void *minv;  /* (*np) lumped inverse inertia, zero on a clamped wall */
  const void *cell;  /* (*np) the cell material at each cell's low corner */
  const void *u0, *u1;
  void *u2;
  /* a nodal damping field d, or NULL: u2 = (2 u1 - u0 (1 - beta) + dt^2 minv
   * force) / (1 + beta), beta = minv d dt / 2 (the USE_DAMPING build,
   * anisotropic.cu:380-389) */
  const void *damping;
  double dt;
} cw_aniso_step_args;

/* the densities of a forward triplet against the adjoint l1
 * (gradient_kernel, anisotropic_sensitivity.cu:111-156), the shots summed:
 * g_mass -= mass_factor sum l1 (u2 - 2 u1 + u0) at every node, g_cell -= l1
 * K u1 over the cell whose low corner the node is */
typedef struct cw_aniso_gradient_args {
  cw_aniso_grid grid;
  const void *table;
  void *g_mass, *g_cell; /* (*np) each */
  const void *u0, *u1, *u2, *l1;
  double mass_factor;    /* rho0 V / dt^2 */
} cw_aniso_gradient_args;

int cw_aniso_step(const cw_aniso_step_args *args, void *stream);
int cw_aniso_gradient(const cw_aniso_gradient_args *args, void *stream);
/* points of a field (kinds 0 inject, 1 record, 2 replay), as cw_elastic_points */
int cw_aniso_points(const cw_aniso_grid *grid, const cw_elastic_point_args *args,
                    void *stream);

/* ------------------------------------------------------------ U-Net
 * The learned preconditioner's kernels (no reference counterpart: the
 * reference has no learned preconditioner). 3x3 convolutions, stride 1, zero
 * padding 1, as implicit GEMMs on the bfloat16 matrix cores with float32
 * accumulation. Activations are NHWC bfloat16; weights (cout, 9 cin), row
 * k = tap cin + ci, tap = 3 (dh + 1) + (dw + 1). */
typedef struct cw_conv3x3_args {
  int32_t n, h, w;             /* batch, height, width */
  int32_t cin, cout;           /* cin a multiple of 16; cout 32, 64 or 128 */
  const void *x;               /* (n, h, w, cin) bfloat16 */
  const void *weight;          /* (cout, 9 cin) bfloat16 */
  const void *weight_flipped;  /* (cin, 9 cout) bfloat16: weight flipped in space,
                                  transposed in channels (cw_conv3x3_weights) */
  const void *bias;            /* (cout) float32, or NULL */
  int32_t relu;                /* 1: max(0, .) after the bias */
  void *y;                     /* (n, h, w, cout) bfloat16 */
  const void *dy;              /* (n, h, w, cout) bfloat16: the gradient at y
                                  before the activation */
  void *dx;                    /* (n, h, w, cin) bfloat16 */
  void *dweight;               /* (cout, 9 cin) float32 */
  void *workspace;             /* (splits, 9 cin, cout) float32 partial sums */
  int32_t splits;              /* pixel runs the weight gradient sums apart */
} cw_conv3x3_args;

/* y = act(bias + conv(x, weight)) */
int cw_conv3x3_bf16_fwd(const cw_conv3x3_args *args, void *stream);
/* dx = conv(dy, weight_flipped): the gradient at x (cin 32, 64 or 128) */
int cw_conv3x3_bf16_bwd_data(const cw_conv3x3_args *args, void *stream);
/* dweight = sum over pixels of dy (x) x shifted by each tap; deterministic:
 * each of `splits` runs of pixels sums into its own workspace slice, then
 * the slices add in order */
int cw_conv3x3_bf16_bwd_weight(const cw_conv3x3_args *args, void *stream);
/* the kernels' bfloat16 weights from a float32 (cout, 9 cin) master: wt as
 * is (round to nearest even), wd flipped and transposed (NULL to skip) */
int cw_conv3x3_weights(int32_t cout, int32_t cin, const void *w, void *wt, void *wd,
                       void *stream);

/* one AdamW step over `count` float32 parameters (decoupled weight decay,
 * bias-corrected moments; the update torch.optim.AdamW makes) */
typedef struct cw_adamw_args {
  int64_t count;
  void *param;
  const void *grad;
  void *m, *v;      /* first and second moments, zeroed before step 1 */
  float lr, beta1, beta2, eps, weight_decay;
  int32_t step;     /* 1 for the first step */
} cw_adamw_args;
int cw_adamw(const cw_adamw_args *args, void *stream);

#ifdef __cplusplus
}
#endif

#endif
