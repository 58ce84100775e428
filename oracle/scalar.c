/* CPU oracle of the scalar (pressure) finite-difference path: TEST
 * INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this; the product never does.
 *
 * A plain-C restatement, node by node, of the reference's scalar kernels as
 * its CPU fallback computes them: IEEE arithmetic, no contraction
 * (-ffp-contract=off, cuwave/cpu.py:22), an explicit fma exactly where the
 * reference pins FMA_RN (cuwave/kernels/common.cuh:11-19), blocks of rows
 * dealt over OpenMP threads (cuwave/kernels/cpu.hpp:73-91).
 *
 * Built twice, REAL=float (suffix _f32) and REAL=double (suffix _f64).
 *
 * Parity pins (tests/test_oracle.py): every function below is checked
 * against the golden vectors the reference itself produced
 * (tests/golden/ *.npz) through the host driver in oracle/driver.py, and the
 * adjoint ones also against the transpose identity <W A^T p, u> = <W p, A u>
 * and central finite differences in float64.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
#ifndef SUF
#define SUF _f64
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

typedef REAL real;

/* the grid: logical extents n (ghosts included), padded extents np, element
 * strides s (s[ndim - 1] == 1), cell-stencil table op[r - 1][k - 1] of the
 * radius-r rows (cuwave/stencils.py:186-213) */
typedef struct {
  int ndim;
  int R;
  int n[3];
  int np[3];
  long s[3];
  double op[4][4];
} grid_t;

#define OPW(g, r, k) ((real)(g)->op[(r)-1][(k)-1])

static inline real fmar(real a, real b, real c) {
#if defined(REAL_IS_FLOAT)
  return fmaf(a, b, c);
#else
  return fma(a, b, c);
#endif
}

static long total_nodes(const grid_t *g) {
  long t = 1;
  for (int d = 0; d < g->ndim; ++d)
    t *= g->np[d];
  return t;
}

/* node index a[] of flat idx */
static void node_axes(const grid_t *g, long idx, int *a) {
  long rest = idx;
  for (int d = 0; d < g->ndim; ++d) {
    a[d] = (int)(rest / g->s[d]);
    rest -= (long)a[d] * g->s[d];
  }
}

static int interior(const grid_t *g, const int *a) {
  for (int d = 0; d < g->ndim; ++d)
    if (!(a[d] > 0 && a[d] < g->n[d] - 1))
      return 0;
  return 1;
}

/* cuwave/kernels/common.cuh:42-46 */
static inline int held_wall(int a, int N, int odd) {
  return ((odd & 1) && a == 1) || ((odd & 2) && a == N - 2);
}

static int held_any(const grid_t *g, const int *a, int faces) {
  int held = 0;
  for (int d = 0; d < g->ndim; ++d)
    held |= held_wall(a[d], g->n[d], faces >> (2 * d));
  return held;
}

/* cuwave/kernels/common.cuh:48-69: tap j of the line through idx, past the
 * stored ghost read off its mirror about the wall node, odd on `odd` */
static inline real mirrored(const real *u, long idx, long s, int a, int N,
                            int odd, int j) {
  int b = a + j;
  int sign = ((b < 0 && (odd & 1)) || (b > N - 1 && (odd & 2))) ? -1 : 1;
  b = b < 0 ? 2 - b : b > N - 1 ? 2 * (N - 2) - b : b;
  real v = u[idx + (long)(b - a) * s];
  return sign > 0 ? v : -v;
}

/* cuwave/kernels/common.cuh:72-77 node_image + scalar_sensitivity.cu:176-184
 * `reflected`: the image about the wall node, zero on a held wall node */
static inline real reflected(const real *u, long idx, long s, int a, int N,
                             int odd, int j) {
  int b = a + j;
  int sign = ((b < 1 && (odd & 1)) || (b > N - 2 && (odd & 2))) ? -1 : 1;
  b = b < 1 ? 2 - b : b > N - 2 ? 2 * (N - 2) - b : b;
  if (held_wall(b, N, odd))
    return (real)0;
  real v = u[idx + (long)(b - a) * s];
  return sign > 0 ? v : -v;
}

/* cuwave/kernels/scalar.cu:14-33 flux_divergence_taps with r == R, factor 1,
 * over the mirrored taps of flux_divergence_axis (:43-65) */
static real flux_axis(const grid_t *g, const real *u1, const real *stiff,
                      long idx, long s, real uc, real sc, int a, int N,
                      int odd) {
  const int R = g->R;
  real t[9];
  for (int j = -R; j <= R; ++j)
    t[j + R] = mirrored(u1, idx, s, a, N, odd, j);
#define TAP(j) t[(j) + R]
  const real sp = stiff[idx + s], sm = stiff[idx - s];
  const real gp = sc * sp / (sc + sp), gm = sc * sm / (sc + sm);
  real Dp = TAP(1) - uc, Dm = uc - TAP(-1);
  if (R > 1) {
    Dp = OPW(g, R, 1) * Dp;
    Dm = OPW(g, R, 1) * Dm;
  }
  for (int k = 2; k <= R; ++k) {
    Dp = fmar(OPW(g, R, k), TAP(k) - TAP(1 - k), Dp);
    Dm = fmar(OPW(g, R, k), TAP(k - 1) - TAP(-k), Dm);
  }
  const real mm = Dm * gm;
  return (real)1 * fmar(Dp, gp, -mm);
#undef TAP
}

/* the axes combined as cuwave/kernels/scalar.cu:416-436 rounds them */
static real laplacian(const grid_t *g, const real *u1, const real *stiff,
                      long idx, const int *a, real uc, real sc, const real *f,
                      int faces) {
  real d[3] = {0, 0, 0};
  for (int e = 0; e < g->ndim; ++e)
    d[e] = flux_axis(g, u1, stiff, idx, g->s[e], uc, sc, a[e], g->n[e],
                     faces >> (2 * e));
  if (g->ndim == 1)
    return f[0] * d[0];
  if (g->ndim == 2) {
    if (g->R == 1) {
      const real p = f[1] * d[1];
      return fmar(f[0], d[0], p);
    }
    const real p = f[0] * d[0];
    return fmar(f[1], d[1], p);
  }
  const real p = f[1] * d[1];
  return fmar(f[2], d[2], fmar(f[0], d[0], p));
}

/* cuwave/wave.py:661-662: the code of a node whose cells all keep the full
 * radius, and the radius code of a cell open onto a held node */
#define DEEP_CELLS (1 << 30)
#define HELD_CELL 7

/* cuwave/kernels/scalar.cu:85-120 flux_divergence_cells, one axis of a node
 * at a domain wall: the cell to the next node at radius rp, to the previous
 * at rm; 0 closed, HELD_CELL open onto the zero halfway to a held node at the
 * node's own stiffness; plain sums (no fma) */
static real graded_axis(const grid_t *g, const real *u1, const real *stiff,
                        long idx, long s, real uc, real sc, real factor,
                        int rp, int rm) {
  real Dp = 0, Dm = 0;
  if (rp == HELD_CELL) {
    Dp = -uc * sc;
  } else if (rp) {
    const real sp = stiff[idx + s];
    Dp = OPW(g, rp, 1) * (u1[idx + s] - uc);
    for (int k = 2; k <= rp; ++k)
      Dp = Dp + OPW(g, rp, k) * (u1[idx + k * s] - u1[idx - (k - 1) * s]);
    Dp = Dp * (sc * sp / (sc + sp));
  }
  if (rm == HELD_CELL) {
    Dm = uc * sc;
  } else if (rm) {
    const real sm = stiff[idx - s];
    Dm = OPW(g, rm, 1) * (uc - u1[idx - s]);
    for (int k = 2; k <= rm; ++k)
      Dm = Dm + OPW(g, rm, k) * (u1[idx + (k - 1) * s] - u1[idx - k * s]);
    Dm = Dm * (sc * sm / (sc + sm));
  }
  return factor * (Dp - Dm);
}

/* cuwave/kernels/scalar.cu:132-158 wall_laplacian: the axes in order, summed */
static real graded_laplacian(const grid_t *g, const real *u1,
                             const real *stiff, long idx, real uc, real sc,
                             const real *f, int code) {
  real sum = 0;
  for (int d = 0; d < g->ndim; ++d) {
    const real term =
        graded_axis(g, u1, stiff, idx, g->s[d], uc, sc, f[d],
                    (code >> (6 * d)) & 7, (code >> (6 * d + 3)) & 7);
    sum = d == 0 ? term : sum + term;
  }
  return sum;
}

/* cuwave/kernels/scalar.cu:346-466 fd_kernel: u2 = -u0 + 2 u1 + minv L u1,
 * damped variant :455-459, Dirichlet wall nodes held at zero. u2 may alias u0
 * (the step writes in place, cuwave/wave.py:1119-1120). `cells`, when given,
 * is a domain mask's per-node code (cuwave/wave.py:675-726): a node of code 0
 * is left as it is, a node at a domain wall (bit 30 clear) takes the graded
 * laplacian (scalar.cu:434-452). */
void FN(cwo_fd_step)(const grid_t *g, const real *u0, const real *u1, real *u2,
                     const real *stiff, const real *minv, int derive_inertia,
                     const real *damping, double dt, const double *fd,
                     int dirichlet, const int *cells) {
  const long total = total_nodes(g);
  real f[3];
  for (int d = 0; d < g->ndim; ++d)
    f[d] = (real)fd[d];
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int a[3];
    node_axes(g, idx, a);
    if (!interior(g, a))
      continue;
    const int code = cells ? cells[idx] : DEEP_CELLS;
    if (!code)
      continue;
    const real uc = u1[idx], sc = stiff[idx];
    const real lap =
        (code & DEEP_CELLS)
            ? laplacian(g, u1, stiff, idx, a, uc, sc, f, dirichlet)
            : graded_laplacian(g, u1, stiff, idx, uc, sc, f, code);
    const real mi = derive_inertia ? (real)1 / sc : minv[idx];
    real v;
    if (damping) {
      const real beta = (real)0.5 * mi * damping[idx] * (real)dt;
      v = ((real)2 * uc - u0[idx] * ((real)1 - beta) + mi * lap) /
          ((real)1 + beta);
    } else {
      v = -u0[idx] + (real)2 * uc + mi * lap;
    }
    if (held_any(g, a, dirichlet))
      v = 0;
    u2[idx] = v;
  }
}

/* --------------------------------------------------------------- pml slabs
 * cuwave/kernels/common.cuh:318-387: slab thread t of axis d, rows lo + hi
 * off its walls; the fast axis keeps its slab rows last */
static long slab_plane(const grid_t *g, int d) {
  long plane = 1;
  for (int e = 0; e < g->ndim; ++e)
    if (e != d)
      plane *= g->np[e];
  return plane;
}

static int slab_node(const grid_t *g, long t, int d, int lo, int hi,
                     long *idx, int *a, long *off, real *w) {
  const long plane = slab_plane(g, d), rows = lo + hi;
  const long row = d < g->ndim - 1 ? t / plane : t % rows;
  *off = d < g->ndim - 1 ? t % plane : t / rows;
  if (row >= rows || *off >= plane)
    return 0;
  *a = row < lo ? 1 + (int)row : g->n[d] - 2 - (int)(row - lo);
  *idx = (long)(*a) * g->s[d];
  *w = (*a == 1 || *a == g->n[d] - 2) ? (real)0.5 : (real)1;
  long rest = *off;
  for (int e = g->ndim - 1; e >= 0; --e)
    if (e != d) {
      const int p = g->np[e], ae = (int)(rest % p);
      rest /= p;
      if (ae < 1 || ae > g->n[e] - 2)
        return 0;
      *idx += (long)ae * g->s[e];
      *w *= (ae == 1 || ae == g->n[e] - 2) ? (real)0.5 : (real)1;
    }
  return 1;
}

static long slab_index(const grid_t *g, int a, long off, int d, int lo, int hi,
                       long plane) {
  const int N = g->n[d];
  const long row = (a >= 1 && a - 1 < lo)           ? a - 1
                   : (a <= N - 2 && N - 2 - a < hi) ? lo + N - 2 - a
                                                    : -1;
  if (row < 0)
    return -1;
  return d < g->ndim - 1 ? row * plane + off : off * (lo + hi) + row;
}

/* cuwave/kernels/scalar.cu:68-82 cell_differences: plain sums, mirrored */
static void cell_differences(const grid_t *g, const real *u1, long idx, long s,
                             real uc, int a, int N, int odd, real *Dp,
                             real *Dm) {
  const int R = g->R;
#define U(j) mirrored(u1, idx, s, a, N, odd, (j))
  real p = OPW(g, R, 1) * (U(1) - uc);
  real m = OPW(g, R, 1) * (uc - U(-1));
  for (int k = 2; k <= R; ++k) {
    p += OPW(g, R, k) * (U(k) - U(1 - k));
    m += OPW(g, R, k) * (U(k - 1) - U(-k));
  }
#undef U
  *Dp = p;
  *Dm = m;
}

/* cuwave/kernels/scalar.cu:499-532 pml_kernel: the CPML memories of axis d
 * (plus cell, minus cell, divergence) and their share of u2 */
void FN(cwo_pml_step)(const grid_t *g, const real *u1, real *u2,
                      const real *stiff, const real *minv, int derive_inertia,
                      real *aux, const real *profile, int d, int lo, int hi,
                      double fd, int dirichlet) {
  const long plane = slab_plane(g, d), size = (lo + hi) * plane;
  const int N = g->n[d];
  const long o = g->s[d];
  const real f = (real)fd;
  const real *an = profile, *bn = profile + N, *ah = profile + 2 * N,
             *bh = profile + 3 * N;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < size; ++t) {
    long idx, off;
    int a;
    real w;
    if (!slab_node(g, t, d, lo, hi, &idx, &a, &off, &w))
      continue;
    const real uc = u1[idx], sc = stiff[idx];
    real Dp, Dm;
    cell_differences(g, u1, idx, o, uc, a, N, dirichlet >> (2 * d), &Dp, &Dm);
    const real sp = stiff[idx + o], sm = stiff[idx - o];
    const real gp = sc * sp / (sc + sp), gm = sc * sm / (sc + sm);
    const real pp = bh[a] * aux[t] + ah[a] * Dp;
    const real pm = bh[a - 1] * aux[size + t] + ah[a - 1] * Dm;
    const real z = bn[a] * aux[2 * size + t] +
                   an[a] * f * (gp * (Dp + pp) - gm * (Dm + pm));
    aux[t] = pp;
    aux[size + t] = pm;
    aux[2 * size + t] = z;
    const real mi = derive_inertia ? (real)1 / sc : minv[idx];
    int A[3];
    node_axes(g, idx, A);
    if (!held_any(g, A, dirichlet))
      u2[idx] += mi * (f * (gp * pp - gm * pm) + z);
  }
}

/* cuwave/kernels/scalar.cu:535-555 the ghost refill of the faces in the
 * masks: Neumann copies the mirror, Dirichlet negates it and zeroes the wall
 * node; edges and corners are left as they are */
void FN(cwo_refill)(const grid_t *g, real *u, int neumann, int dirichlet) {
  for (int pass = 0; pass < 2; ++pass) {
    const int faces = pass ? dirichlet : neumann;
    for (int d = 0; d < g->ndim; ++d)
      for (int side = 0; side < 2; ++side) {
        if (!((faces >> (2 * d + side)) & 1))
          continue;
        const long normal = side ? -g->s[d] : g->s[d];
        const int ag = side ? g->n[d] - 1 : 0;
        /* every node of the face with the other axes interior */
        int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
        for (int e = 0; e < g->ndim; ++e) {
          lo[e] = 1;
          hi[e] = g->n[e] - 2;
        }
        lo[d] = hi[d] = ag;
        int a[3];
        for (a[0] = lo[0]; a[0] <= hi[0]; ++a[0])
          for (a[1] = (g->ndim > 1 ? lo[1] : 0);
               a[1] <= (g->ndim > 1 ? hi[1] : 0); ++a[1])
            for (a[2] = (g->ndim > 2 ? lo[2] : 0);
                 a[2] <= (g->ndim > 2 ? hi[2] : 0); ++a[2]) {
              long ghost = 0;
              for (int e = 0; e < g->ndim; ++e)
                ghost += (long)a[e] * g->s[e];
              if (pass == 0) {
                u[ghost] = u[ghost + 2 * normal];
              } else {
                u[ghost] = -u[ghost + 2 * normal];
                u[ghost + normal] = 0;
              }
            }
      }
  }
}

/* cuwave/kernels/common.cuh:456-464 and :483-503 */
void FN(cwo_excite)(real *u, const real *row, const int *lin, int num,
                    const real *weight) {
  for (int j = 0; j < num; ++j)
    u[lin[j]] += weight[j] * row[j];
}

void FN(cwo_record)(const real *u, real *row, const int *lin, int num) {
  for (int j = 0; j < num; ++j)
    row[j] = u[lin[j]];
}

void FN(cwo_replay)(real *u, const real *row, const int *lin, int num) {
  for (int j = 0; j < num; ++j)
    u[lin[j]] = row[j];
}

/* cuwave/kernels/common.cuh:467-480 adjoint_excitation_kernel */
void FN(cwo_adjoint_excite)(real *l2, const real *row, const int *lin, int num,
                            const real *weight, real *g_mass, const real *u1,
                            double mf) {
  for (int j = 0; j < num; ++j) {
    const int n = lin[j];
    const real load = weight[j] * row[j];
    l2[n] += load;
    g_mass[n] += -(real)mf * u1[n] * load;
  }
}

/* ------------------------------------------------------------ the adjoint
 * cuwave/kernels/scalar_sensitivity.cu:42-48: a row whose taps reach a wall
 * node or past it takes the mirrored path */
static inline int near_wall(const grid_t *g, int a, int N) {
  const int reach = g->R + 2;
  return a < reach || a > N - 1 - reach;
}

typedef real (*tap_fn)(const real *, long, long, int, int, int, int);

static inline real direct(const real *u, long idx, long s, int a, int N,
                          int odd, int j) {
  (void)a;
  (void)N;
  (void)odd;
  return u[idx + (long)j * s];
}

/* scalar_sensitivity.cu:116-135 stiffness_gradient_taps */
static real stiffness_gradient(const grid_t *g, tap_fn tu, tap_fn tk,
                               const real *u1, const real *l1,
                               const real *stiff, long idx, long s, int a,
                               int N, int odd, real uc, real lc, real sc,
                               real factor) {
  const int R = g->R;
  const real sp = tk(stiff, idx, s, a, N, 0, 1);
  const real sm = tk(stiff, idx, s, a, N, 0, -1);
  const real dgp = sp * sp / ((sc + sp) * (sc + sp));
  const real dgm = sm * sm / ((sc + sm) * (sc + sm));
  real Dp = OPW(g, R, 1) * (tu(u1, idx, s, a, N, odd, 1) - uc);
  real Dm = OPW(g, R, 1) * (uc - tu(u1, idx, s, a, N, odd, -1));
  for (int j = 2; j <= R; ++j) {
    Dp += OPW(g, R, j) *
          (tu(u1, idx, s, a, N, odd, j) - tu(u1, idx, s, a, N, odd, 1 - j));
    Dm += OPW(g, R, j) *
          (tu(u1, idx, s, a, N, odd, j - 1) - tu(u1, idx, s, a, N, odd, -j));
  }
  return factor * (dgp * Dp * (tu(l1, idx, s, a, N, odd, 1) - lc) +
                   dgm * Dm * (lc - tu(l1, idx, s, a, N, odd, -1)));
}

/* scalar_sensitivity.cu:137-154 transposed_divergence_taps */
static real transposed_divergence(const grid_t *g, tap_fn tl, tap_fn tk,
                                  const real *l1, const real *stiff, long idx,
                                  long s, int a, int N, int odd, real lc,
                                  real sc) {
  real div = 0, sh = sc, sl = sc, lh = lc, ll = lc;
  for (int j = 1; j <= g->R; ++j) {
    const real sh1 = tk(stiff, idx, s, a, N, 0, j);
    const real sl1 = tk(stiff, idx, s, a, N, 0, -j);
    const real lh1 = tl(l1, idx, s, a, N, odd, j);
    const real ll1 = tl(l1, idx, s, a, N, odd, -j);
    div += OPW(g, g->R, j) * (sh * sh1 / (sh + sh1) * (lh1 - lh) -
                              sl * sl1 / (sl + sl1) * (ll - ll1));
    sh = sh1, sl = sl1, lh = lh1, ll = ll1;
  }
  return div;
}

/* scalar_sensitivity.cu:387-441 adjoint_kernel: the exact transpose
 * W^-1 A^T W as a gather; l2 may alias l0 */
void FN(cwo_adjoint_step)(const grid_t *g, const real *l0, const real *l1,
                          real *l2, const real *stiff, const real *minv,
                          int derive_inertia, const real *damping, double dt,
                          const double *fd, int dirichlet) {
  const long total = total_nodes(g);
  real f[3];
  for (int d = 0; d < g->ndim; ++d)
    f[d] = (real)fd[d];
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int a[3];
    node_axes(g, idx, a);
    if (!interior(g, a))
      continue;
    const real lc = l1[idx], sc = stiff[idx], lo = l0[idx];
    const real mi = derive_inertia ? (real)1 / sc : minv[idx];
    real div_l = 0;
    const int held = held_any(g, a, dirichlet);
    for (int d = 0; d < g->ndim; ++d) {
      const int odd = dirichlet >> (2 * d);
      const int wall = near_wall(g, a[d], g->n[d]);
      div_l += f[d] * transposed_divergence(
                          g, wall ? reflected : direct, wall ? reflected : direct,
                          l1, stiff, idx, g->s[d], a[d], g->n[d], odd, lc, sc);
    }
    real ln;
    if (damping) {
      const real beta = (real)0.5 * mi * damping[idx] * (real)dt;
      ln = ((real)2 * lc - lo * ((real)1 - beta) + mi * div_l) /
           ((real)1 + beta);
    } else {
      ln = -lo + (real)2 * lc + mi * div_l;
    }
    l2[idx] = held ? (real)0 : ln;
  }
}

/* scalar_sensitivity.cu:443-515 adjoint_gradient_kernel: one adjoint step
 * over l0 (in place) and both gradients of l1, the inertia term by parts */
void FN(cwo_adjoint_gradient)(const grid_t *g, real *l0, const real *l1,
                              real *g_mass, real *g_stiff, const real *u1,
                              const real *stiff, const real *minv,
                              int derive_inertia, const real *damping,
                              double dt, int dirichlet, double mf_,
                              const double *fd) {
  const long total = total_nodes(g);
  real f[3];
  for (int d = 0; d < g->ndim; ++d)
    f[d] = (real)fd[d];
  const real mf = (real)mf_;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int a[3];
    node_axes(g, idx, a);
    if (!interior(g, a))
      continue;
    const real uc = u1[idx], lc = l1[idx], sc = stiff[idx], lo = l0[idx];
    const real gm = g_mass[idx], gs = g_stiff[idx];
    const real mi = derive_inertia ? (real)1 / sc : minv[idx];
    real div_l = 0, gsum = 0;
    const int held = held_any(g, a, dirichlet);
    const real lw = held ? (real)0 : lc;
    for (int d = 0; d < g->ndim; ++d) {
      const int odd = dirichlet >> (2 * d);
      if (near_wall(g, a[d], g->n[d])) {
        div_l += f[d] * transposed_divergence(g, reflected, reflected, l1,
                                              stiff, idx, g->s[d], a[d],
                                              g->n[d], odd, lc, sc);
        gsum += stiffness_gradient(g, reflected, reflected, u1, l1, stiff, idx,
                                   g->s[d], a[d], g->n[d], odd, uc, lw, sc,
                                   f[d]);
      } else {
        div_l += f[d] * transposed_divergence(g, direct, direct, l1, stiff,
                                              idx, g->s[d], a[d], g->n[d], odd,
                                              lc, sc);
        gsum += stiffness_gradient(g, direct, direct, u1, l1, stiff, idx,
                                   g->s[d], a[d], g->n[d], odd, uc, lc, sc,
                                   f[d]);
      }
    }
    real ln;
    if (damping) {
      const real beta = (real)0.5 * mi * damping[idx] * (real)dt;
      ln = ((real)2 * lc - lo * ((real)1 - beta) + mi * div_l) /
           ((real)1 + beta);
    } else {
      ln = -lo + (real)2 * lc + mi * div_l;
    }
    if (held)
      ln = 0;
    l0[idx] = ln;
    g_mass[idx] = gm - mf * uc * (ln - (real)2 * lc + lo);
    g_stiff[idx] = gs - mf * gsum;
  }
}

/* scalar_sensitivity.cu:302-350 gradient_kernel: both densities of the
 * forward triplet (u0, u1, u2) against l1 */
void FN(cwo_gradient)(const grid_t *g, real *g_mass, real *g_stiff,
                      const real *u0, const real *u1, const real *u2,
                      const real *l1, const real *stiff, double inv_dt2_,
                      const double *Fd, int dirichlet) {
  const long total = total_nodes(g);
  real F[3];
  for (int d = 0; d < g->ndim; ++d)
    F[d] = (real)Fd[d];
  const real inv_dt2 = (real)inv_dt2_;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int a[3];
    node_axes(g, idx, a);
    if (!interior(g, a))
      continue;
    const real uc = u1[idx], lc = l1[idx];
    g_mass[idx] -= inv_dt2 * lc * (u2[idx] - (real)2 * uc + u0[idx]);
    const real sc = stiff[idx];
    real gsum = 0;
    const int held = held_any(g, a, dirichlet);
    const real lw = held ? (real)0 : lc;
    for (int d = 0; d < g->ndim; ++d) {
      const int odd = dirichlet >> (2 * d);
      if (near_wall(g, a[d], g->n[d]))
        gsum += stiffness_gradient(g, reflected, reflected, u1, l1, stiff, idx,
                                   g->s[d], a[d], g->n[d], odd, uc, lw, sc,
                                   F[d]);
      else
        gsum += stiffness_gradient(g, direct, direct, u1, l1, stiff, idx,
                                   g->s[d], a[d], g->n[d], odd, uc, lc, sc,
                                   F[d]);
    }
    g_stiff[idx] -= gsum;
  }
}

/* scalar_sensitivity.cu:219-251 d(Dp)/d(u at q) and d(Dm)/d(u at q) */
static real plus_tap(const grid_t *g, int q) {
  const int r = g->R;
  if (q >= 1)
    return q <= r ? OPW(g, r, q) : (real)0;
  return 1 - q <= r ? -OPW(g, r, 1 - q) : (real)0;
}

static real minus_tap(const grid_t *g, int q) {
  const int r = g->R;
  if (q >= 0)
    return q + 1 <= r ? OPW(g, r, q + 1) : (real)0;
  return -q <= r ? -OPW(g, r, -q) : (real)0;
}

static real pml_tap(const grid_t *g, const real *aux, const real *stiff,
                    const real *profile, real f, int N, long o, long size,
                    long i, long iidx, int ai, int q) {
  const real *an = profile, *ah = profile + 2 * N;
  const real sc = stiff[iidx], sp = stiff[iidx + o], sm = stiff[iidx - o];
  const real gp = sc * sp / (sc + sp), gm = sc * sm / (sc + sm);
  const real ez = an[ai] * f * aux[2 * size + i];
  const real cp = ah[ai] * aux[i] + gp * (1 + ah[ai]) * ez;
  const real cm = ah[ai - 1] * aux[size + i] - gm * (1 + ah[ai - 1]) * ez;
  return cp * plus_tap(g, q) + cm * minus_tap(g, q);
}

/* scalar_sensitivity.cu:517-542 pml_aux_kernel */
void FN(cwo_pml_aux)(const grid_t *g, const real *l1, const real *stiff,
                     real *aux, const real *profile, int d, int lo, int hi,
                     double fd, int dirichlet) {
  const long plane = slab_plane(g, d), size = (lo + hi) * plane;
  const int N = g->n[d];
  const long o = g->s[d];
  const real f = (real)fd;
  const real *an = profile, *bn = profile + N, *bh = profile + 3 * N;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < size; ++t) {
    long idx, off;
    int a;
    real w;
    if (!slab_node(g, t, d, lo, hi, &idx, &a, &off, &w))
      continue;
    const real sc = stiff[idx], sp = stiff[idx + o], sm = stiff[idx - o];
    const real gp = sc * sp / (sc + sp), gm = sc * sm / (sc + sm);
    int A[3];
    node_axes(g, idx, A);
    const real q = held_any(g, A, dirichlet) ? (real)0 : w * l1[idx];
    const real ez = aux[2 * size + t];
    aux[t] = bh[a] * (aux[t] + an[a] * f * gp * ez) + f * gp * q;
    aux[size + t] = bh[a - 1] * (aux[size + t] - an[a] * f * gm * ez) -
                    f * gm * q;
    aux[2 * size + t] = bn[a] * ez + q;
  }
}

/* scalar_sensitivity.cu:545-597 pml_adjoint_kernel: the transposed memory
 * terms gathered onto every node the layer's taps reach */
void FN(cwo_pml_adjoint)(const grid_t *g, real *l2, real *g_mass,
                         const real *u1, double mf_, int with_mass,
                         const real *stiff, const real *minv,
                         int derive_inertia, const real *aux,
                         const real *profile, int d, int lo, int hi, double fd,
                         int dirichlet) {
  const int R = g->R;
  const int elo = lo ? lo + R : 0, ehi = hi ? hi + R : 0;
  const long plane = slab_plane(g, d), size = (lo + hi) * plane;
  const long esize = (long)(elo + ehi) * plane;
  const int N = g->n[d];
  const long o = g->s[d];
  const real f = (real)fd, mf = (real)mf_;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < esize; ++t) {
    long idx, off;
    int a;
    real w;
    if (!slab_node(g, t, d, elo, ehi, &idx, &a, &off, &w))
      continue;
    real sum = 0;
    for (int m = -R; m <= R; ++m) {
      const long i = slab_index(g, a + m, off, d, lo, hi, plane);
      if (i >= 0)
        sum += pml_tap(g, aux, stiff, profile, f, N, o, size, i, idx + m * o,
                       a + m, -m);
    }
    for (int high = 0; high < 2; ++high) {
      const int ghost = high ? 2 * (N - 2) - a : 2 - a;
      if (high ? ghost < N - 1 : ghost > 0)
        continue;
      real fold = 0;
      for (int k = 1; k <= R; ++k) {
        const int ai = high ? N - 1 - k : k;
        const long i = slab_index(g, ai, off, d, lo, hi, plane);
        if (i >= 0 && ghost - ai >= -R && ghost - ai <= R)
          fold += pml_tap(g, aux, stiff, profile, f, N, o, size, i,
                          idx + (long)(ai - a) * o, ai, ghost - ai);
      }
      sum += ((dirichlet >> (2 * d + high)) & 1) ? -fold : fold;
    }
    const real mi = derive_inertia ? (real)1 / stiff[idx] : minv[idx];
    const real delta = mi * sum / w;
    l2[idx] += delta;
    if (with_mass)
      g_mass[idx] -= mf * u1[idx] * delta;
  }
}

/* scalar_sensitivity.cu:599-668 superposed_kernel: the step plus the
 * Frechet densities, mass one triplet late (u2 still holding it) */
void FN(cwo_superposed)(const grid_t *g, const real *u0, const real *u1,
                        real *u2, real *acc_mass, real *acc_stiff,
                        const real *stiff, const real *minv,
                        int derive_inertia, double ft_, double fs_,
                        const double *fd, int dirichlet) {
  const long total = total_nodes(g);
  real f[3];
  for (int d = 0; d < g->ndim; ++d)
    f[d] = (real)fd[d];
  const real ft = (real)ft_, fs = (real)fs_;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int a[3];
    node_axes(g, idx, a);
    if (!interior(g, a))
      continue;
    const real uc = u1[idx], sc = stiff[idx];
    const real lap = laplacian(g, u1, stiff, idx, a, uc, sc, f, dirichlet);
    const int held = held_any(g, a, dirichlet);
    const real dudt = uc - u2[idx];
    acc_mass[idx] += ft * dudt * dudt;
    real sum = 0;
    for (int d = 0; d < g->ndim; ++d) {
      const real gd = u1[idx + g->s[d]] - u1[idx - g->s[d]];
      if (d == 0)
        sum = f[0] * gd * gd;
      else
        sum += f[d] * gd * gd;
    }
    acc_stiff[idx] += fs * sum;
    const real mi = derive_inertia ? (real)1 / sc : minv[idx];
    u2[idx] = held ? (real)0 : -u0[idx] + (real)2 * uc + mi * lap;
  }
}

/* scalar_sensitivity.cu:352-385 frechet_kernel */
void FN(cwo_frechet)(const grid_t *g, real *acc_mass, real *acc_stiff,
                     const real *u0, const real *u1, const real *u2,
                     double ft_, const double *fd) {
  const long total = total_nodes(g);
  real f[3];
  for (int d = 0; d < g->ndim; ++d)
    f[d] = (real)fd[d];
  const real ft = (real)ft_;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int a[3];
    node_axes(g, idx, a);
    if (!interior(g, a))
      continue;
    const real dudt = u2[idx] - u0[idx];
    acc_mass[idx] += ft * dudt * dudt;
    real sum = 0;
    for (int d = 0; d < g->ndim; ++d) {
      const real gd = u1[idx + g->s[d]] - u1[idx - g->s[d]];
      if (d == 0)
        sum = f[0] * gd * gd;
      else
        sum += f[d] * gd * gd;
    }
    acc_stiff[idx] += sum;
  }
}
