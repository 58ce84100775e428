/* CPU oracle of the elastic (staggered displacement) finite-difference path:
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this; the product never does.
 *
 * A plain-C restatement, node by node, of the reference's elastic kernels as
 * its CPU fallback computes them: IEEE arithmetic, no contraction
 * (-ffp-contract=off, cuwave/cpu.py:22), an explicit fma exactly where the
 * reference pins FMA_RN and a plain product where it pins MUL_RN
 * (cuwave/kernels/common.cuh:11-19):
 * - the staggered SBP derivatives and their closure rows
 *   (cuwave/kernels/common.cuh:152-298);
 * - stress_kernel and fd_kernel with the folded pml and the gradient
 *   densities (cuwave/kernels/elastic.cu:121-600, 754-853);
 * - gradient_kernel (cuwave/kernels/elastic_sensitivity.cu:115-158).
 *
 * Built twice, REAL=float (suffix _f32) and REAL=double (suffix _f64).
 *
 * Parity pins (tests/test_elastic_oracle.py): the golden vectors the
 * reference's own CPU backend wrote (tests/golden/el_*.npz) bit for bit, the
 * symmetry of the lossless operator in its norm, the one-dimensional
 * reduction to the scalar solver, and central finite differences in float64.
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

#if defined(REAL_IS_FLOAT)
#define TINY 1.17549435e-38f
static inline real fmar(real a, real b, real c) { return fmaf(a, b, c); }
static inline real fmaxr(real a, real b) { return fmaxf(a, b); }
#else
#define TINY 2.2250738585072014e-308
static inline real fmar(real a, real b, real c) { return fma(a, b, c); }
static inline real fmaxr(real a, real b) { return fmax(a, b); }
#endif

/* the grid, the staggered operator and its closure rows
 * (cuwave/stencils.py:186-213 preamble), the material constants */
typedef struct {
  int ndim, R;
  int n[3], np[3];
  long s[3];       /* node strides, s[ndim - 1] == 1 */
  long cs;         /* elements between components */
  double sg[4];    /* full-radius staggered weights, sg[k - 1] */
  int nc, hc;      /* node / half closure rows */
  int nspan, hspan;
  double node_rows[12][16];
  double half_rows[12][16];
  double lam, mu, rho; /* Lame parameters and the background density */
  int clamped;         /* faces 2 axis + side held at zero */
} elgeom_t;

#define SGW(g, k) ((real)(g)->sg[(k)-1])

static int clamped_face(const elgeom_t *g, int d, int side) {
  return (g->clamped >> (2 * d + side)) & 1;
}

/* -------------------------------------------------------- line accessors */
/* kinds of line a derivative reads: a field component along an axis, a
 * stress row along an axis, or either through the pml memory (the Memory
 * accessor, cuwave/kernels/common.cuh:421-437) */
typedef struct {
  const real *p; /* field or stresses */
  long idx, cs, step;
  int e;                         /* the stress accessor's axis */
  int stress;                    /* 1: p holds stress rows */
  /* memory: coef * (base + aux[(row0 + k) size + i]) off the slab index */
  int memory;
  const real *aux, *profile;
  long size, plane, off;
  int row0, a, lo, hi, own_half;
} line_t;

static int stress_row(const elgeom_t *g, int c, int e) {
  if (c == e)
    return c;
  int k = c < e ? c : e, l = c < e ? e : c;
  return g->ndim == 3 ? g->ndim + 3 - k - l : 2;
}

/* the slab index of row a beside `off` (cuwave/kernels/common.cuh:366-376) */
static long slab_index(const elgeom_t *g, int a, long off, int d, int lo, int hi,
                       long plane) {
  const int N = g->n[d];
  const int row = (a >= 1 && a - 1 < lo)               ? a - 1
                  : (a <= N - 2 && N - 2 - a < hi) ? lo + N - 2 - a
                                                       : -1;
  if (row < 0)
    return -1;
  return d < g->ndim - 1 ? (long)row * plane + off : off * (lo + hi) + row;
}

/* component (or stress row) k at j points along the line */
static real line_at(const elgeom_t *g, const line_t *L, int k, int j) {
  real base;
  if (L->stress)
    base = L->p[stress_row(g, k, L->e) * L->cs + L->idx + j * L->step];
  else
    base = L->p[k * L->cs + L->idx + j * L->step];
  if (!L->memory)
    return base;
  const int at = L->a + j;
  const long i = slab_index(g, at, L->off, L->e, L->lo, L->hi, L->plane);
  if (i < 0)
    return (real)0;
  const int half = (k == L->e) == L->own_half;
  const real coef = L->profile[(half ? 2 * g->n[L->e] : 0) + at];
  return coef * (base + L->aux[(L->row0 + k) * L->size + i]);
}

/* --------------------------------------------- staggered SBP derivatives */
/* cuwave/kernels/common.cuh:214-242 */
static real node_rows(const elgeom_t *g, const line_t *L, int k, int lo, int hi) {
  real acc = 0;
  if (lo < g->nc) {
    for (int m = 0; m < g->nspan; ++m)
      acc += (real)g->node_rows[lo][m] * line_at(g, L, k, m - lo);
  } else {
    for (int m = 0; m < g->nspan; ++m)
      acc -= (real)g->node_rows[hi][m] * line_at(g, L, k, hi - 1 - m);
  }
  return acc;
}

static real half_rows(const elgeom_t *g, const line_t *L, int k, int lo, int hi) {
  real acc = 0;
  if (lo < g->hc) {
    for (int i = 0; i < g->hspan; ++i)
      acc += (real)g->half_rows[lo][i] * line_at(g, L, k, i - lo);
  } else {
    for (int i = 0; i < g->hspan; ++i)
      acc -= (real)g->half_rows[hi][i] * line_at(g, L, k, hi + 1 - i);
  }
  return acc;
}

/* d/dx at node a of a half-point field (common.cuh:247-258) */
static real node_derivative(const elgeom_t *g, const line_t *L, int k, int a, int N) {
  const int lo = a - 1, hi = N - 2 - a;
  if (lo < g->nc || hi < g->nc)
    return node_rows(g, L, k, lo, hi);
  real acc = 0;
  for (int q = 1; q <= g->R; ++q)
    acc += SGW(g, q) * (line_at(g, L, k, q - 1) - line_at(g, L, k, -q));
  return acc;
}

/* d/dx at half point a of a nodal field (common.cuh:261-271) */
static real half_derivative(const elgeom_t *g, const line_t *L, int k, int a, int N) {
  const int lo = a - 1, hi = N - 3 - a;
  if (lo < g->hc || hi < g->hc)
    return half_rows(g, L, k, lo, hi);
  real acc = 0;
  for (int q = 1; q <= g->R; ++q)
    acc += SGW(g, q) * (line_at(g, L, k, q) - line_at(g, L, k, 1 - q));
  return acc;
}

/* the divergence of a nodal stress at half point a (common.cuh:276-288) */
static real node_divergence(const elgeom_t *g, const line_t *L, int k, int a, int N) {
  const int lo = a - 1, hi = N - 3 - a;
  if (lo < g->hc || hi < g->hc)
    return half_rows(g, L, k, lo, hi);
  real acc = 0;
  for (int q = 1; q <= g->R; ++q) {
    acc += SGW(g, q) * line_at(g, L, k, q);
    acc -= SGW(g, q) * line_at(g, L, k, 1 - q);
  }
  return acc;
}

/* the divergence of a half-point stress at node a (common.cuh:292-304) */
static real half_divergence(const elgeom_t *g, const line_t *L, int k, int a, int N) {
  const int lo = a - 1, hi = N - 2 - a;
  if (lo < g->nc || hi < g->nc)
    return node_rows(g, L, k, lo, hi);
  real acc = 0;
  for (int q = 1; q <= g->R; ++q) {
    acc += SGW(g, q) * line_at(g, L, k, q - 1);
    acc -= SGW(g, q) * line_at(g, L, k, -q);
  }
  return acc;
}

static line_t field_line(const real *p, long idx, long cs, long step) {
  line_t L;
  memset(&L, 0, sizeof L);
  L.p = p;
  L.idx = idx;
  L.cs = cs;
  L.step = step;
  return L;
}

/* ------------------------------------------------------ strains, stresses */
static int traction_zero(const elgeom_t *g, const int *A, int d) {
  return (A[d] == 1 && !clamped_face(g, d, 0)) ||
         (A[d] == g->n[d] - 2 && !clamped_face(g, d, 1));
}

/* elastic.cu:121-141 */
static int normal_strains(const elgeom_t *g, const real *u, long idx, const int *A,
                          const real *F, real *eps) {
  int zeroed = 0;
  for (int d = 0; d < g->ndim; ++d) {
    if (traction_zero(g, A, d)) {
      eps[d] = 0;
      zeroed |= 1 << d;
    } else {
      line_t L = field_line(u, idx, g->cs, g->s[d]);
      eps[d] = F[d] * node_derivative(g, &L, d, A[d], g->n[d]);
    }
  }
  return zeroed;
}

/* elastic.cu:145-154 */
static real condensed_lame(const elgeom_t *g, real lam, real mu, int zeroed) {
  real lam_eff = lam;
  for (int d = 0; d < g->ndim; ++d)
    if ((zeroed >> d) & 1)
      lam_eff = (real)2.f * lam_eff * mu / (lam_eff + (real)2.f * mu);
  return lam_eff;
}

/* elastic.cu:159-170 */
static real shear_strain(const elgeom_t *g, const real *u, long idx, int k, int l,
                         const int *A, const real *F) {
  line_t Lk = field_line(u, idx, g->cs, g->s[k]);
  line_t Ll = field_line(u, idx, g->cs, g->s[l]);
  const real dk = half_derivative(g, &Lk, l, A[k], g->n[k]);
  const real dl = half_derivative(g, &Ll, k, A[l], g->n[l]);
  return fmar(F[k], dk, F[l] * dl);
}

static int shear_point(const elgeom_t *g, int k, int l, const int *A) {
  return A[k] <= g->n[k] - 3 && A[l] <= g->n[l] - 3;
}

/* elastic.cu:189-197. The reference's CPU build resolves `fmax` to the C
 * library's double function (its kernels include <cmath> only, which
 * declares no float overload in the global namespace), so each reciprocal
 * is a double quotient, added into the float sum before rounding: the sum
 * rounds once per term from the exact double quotient. */
static real shear_gamma(const elgeom_t *g, const real *gamma, long idx, int k, int l) {
  const long at[4] = {idx, idx + g->s[l], idx + g->s[k], idx + g->s[k] + g->s[l]};
  real sum = (real)(1.0 / fmax((double)gamma[at[0]], (double)TINY));
  for (int i = 1; i < 4; ++i)
    sum = (real)((double)sum + 1.0 / fmax((double)gamma[at[i]], (double)TINY));
  return (real)4 / sum;
}

/* elastic.cu:200-211 */
static real point_minv(const elgeom_t *g, const real *gamma, long idx, int c,
                       const int *A) {
  for (int d = 0; d < g->ndim; ++d)
    if (d != c && ((A[d] == 1 && clamped_face(g, d, 0)) ||
                   (A[d] == g->n[d] - 2 && clamped_face(g, d, 1))))
      return (real)0;
  const real mean = (real)0.5 * (gamma[idx] + gamma[idx + g->s[c]]);
  return (real)1 / fmaxr((real)g->rho * mean, TINY);
}

static int nvoigt(const elgeom_t *g) { return g->ndim + g->ndim * (g->ndim - 1) / 2; }

/* elastic.cu:216-233 */
static int strain_point(const elgeom_t *g, const real *u, long idx, const int *A,
                        const real *F, real *eps) {
  const int zeroed = normal_strains(g, u, idx, A, F, eps);
  for (int k = 0; k < g->ndim - 1; ++k)
    for (int l = k + 1; l < g->ndim; ++l)
      eps[stress_row(g, k, l)] =
          shear_point(g, k, l, A) ? shear_strain(g, u, idx, k, l, A, F) : (real)0;
  return zeroed;
}

/* elastic.cu:236-263 */
static void stress_of(const elgeom_t *g, const real *eps, int zeroed, const real *gamma,
                      long idx, const int *A, real *sig) {
  const real lam = (real)g->lam, mu = (real)g->mu;
  const real lam_eff = condensed_lame(g, lam, mu, zeroed);
  real tr = 0;
  for (int d = 0; d < g->ndim; ++d)
    tr += eps[d];
  const real gn = gamma[idx];
  for (int d = 0; d < g->ndim; ++d)
    sig[d] = ((zeroed >> d) & 1) ? (real)0
                                 : gn * ((real)2.f * mu * eps[d] + lam_eff * tr);
  for (int k = 0; k < g->ndim - 1; ++k)
    for (int l = k + 1; l < g->ndim; ++l) {
      const int v = stress_row(g, k, l);
      sig[v] = shear_point(g, k, l, A) ? shear_gamma(g, gamma, idx, k, l) * mu * eps[v]
                                       : (real)0;
    }
}

/* elastic.cu:281-303 */
static real nodal_force(const elgeom_t *g, const real *sigma, long idx, int c,
                        const int *A, const real *F) {
  line_t Lc = field_line(sigma, idx, g->cs, g->s[c]);
  real force = F[c] * node_divergence(g, &Lc, c, A[c], g->n[c]);
  for (int l = 0; l < g->ndim; ++l) {
    if (l == c)
      continue;
    const int v = stress_row(g, c, l);
    line_t Lv = field_line(sigma, idx, g->cs, g->s[l]);
    force = fmar(F[l], half_divergence(g, &Lv, v, A[l], g->n[l]), force);
  }
  return force;
}

/* elastic.cu:339-351 */
static real normal_form(const elgeom_t *g, const real *ea, const real *eb, int zeroed) {
  const real lam = (real)g->lam, mu = (real)g->mu;
  const real lam_eff = condensed_lame(g, lam, mu, zeroed);
  real tra = 0, trb = 0, dot = 0;
  for (int d = 0; d < g->ndim; ++d) {
    tra += ea[d];
    trb += eb[d];
    dot += ea[d] * eb[d];
  }
  return (real)2.f * mu * dot + lam_eff * tra * trb;
}

/* ----------------------------------------------------------- the pml fold */
/* one lined axis: memory, profile (4, n[e]) and the rows behind each face */
typedef struct {
  real *aux;
  const real *profile;
  int lo, hi;
} elpml_t;

static long padded_extent(const elgeom_t *g, int e) { return g->np[e]; }

static long slab_plane(const elgeom_t *g, int d) {
  long plane = 1;
  for (int e = 0; e < g->ndim; ++e)
    if (e != d)
      plane *= padded_extent(g, e);
  return plane;
}

static long slab_offset(const elgeom_t *g, int d, const int *A) {
  long off = 0;
  for (int e = 0; e < g->ndim; ++e)
    if (e != d)
      off = off * padded_extent(g, e) + A[e];
  return off;
}

static int in_slab(const elgeom_t *g, int a, int e, int lo, int hi, int ext) {
  return (lo && a - 1 < lo + ext) || (hi && g->n[e] - 2 - a < hi + ext);
}

static int zeroed_axes(const elgeom_t *g, const int *A) {
  int zeroed = 0;
  for (int d = 0; d < g->ndim; ++d)
    if (traction_zero(g, A, d))
      zeroed |= 1 << d;
  return zeroed;
}

/* d u_k / d x_e at the strain point (k, e) (elastic.cu:417-431) */
static real strain_term(const elgeom_t *g, const line_t *L, int k, int e, const int *A,
                        real fe) {
  if (k == e) {
    if (traction_zero(g, A, e))
      return (real)0;
    return fe * node_derivative(g, L, e, A[e], g->n[e]);
  }
  if (A[k] > g->n[k] - 3 || A[e] > g->n[e] - 3)
    return (real)0;
  return fe * half_derivative(g, L, k, A[e], g->n[e]);
}

/* elastic.cu:434-442 */
static real divergence_term(const elgeom_t *g, const line_t *L, int c, int e,
                            const int *A, real fe) {
  if (c == e)
    return fe * node_divergence(g, L, c, A[c], g->n[c]);
  return fe * half_divergence(g, L, c, A[e], g->n[e]);
}

/* elastic.cu:445-467 */
static void add_stress(const elgeom_t *g, real *sig, const real *gamma, long idx, int e,
                       const real *t, const int *A) {
  const real lam = (real)g->lam, mu = (real)g->mu;
  const int zeroed = zeroed_axes(g, A);
  if (!((zeroed >> e) & 1)) {
    const real lam_eff = condensed_lame(g, lam, mu, zeroed);
    const real gg = gamma[idx] * t[e];
    for (int d = 0; d < g->ndim; ++d)
      if (!((zeroed >> d) & 1))
        sig[d] += gg * ((d == e ? (real)2.f * mu : (real)0) + lam_eff);
  }
  for (int k = 0; k < g->ndim; ++k)
    if (k != e && A[k] <= g->n[k] - 3 && A[e] <= g->n[e] - 3) {
      const int lo = k < e ? k : e, hi = k < e ? e : k;
      sig[stress_row(g, k, e)] += shear_gamma(g, gamma, idx, lo, hi) * mu * t[k];
    }
}

/* elastic.cu:486-528 */
static void stress_memory(const elgeom_t *g, const real *u, const real *gamma, long idx,
                          int e, const elpml_t *P, int phase, const int *A,
                          const real *F, real *sig) {
  const int a = A[e];
  if (!in_slab(g, a, e, P->lo, P->hi, phase ? g->R : 0))
    return;
  const long plane = slab_plane(g, e), size = (long)(P->lo + P->hi) * plane;
  const long off = slab_offset(g, e, A);
  real q[3];
  if (phase) {
    line_t h = field_line(u, idx, g->cs, g->s[e]);
    h.memory = 1;
    h.aux = P->aux;
    h.profile = P->profile;
    h.size = size;
    h.row0 = g->ndim;
    h.a = a;
    h.off = off;
    h.e = e;
    h.lo = P->lo;
    h.hi = P->hi;
    h.plane = plane;
    h.own_half = 1;
    for (int k = 0; k < g->ndim; ++k)
      q[k] = strain_term(g, &h, k, e, A, F[e]);
  } else {
    line_t f = field_line(u, idx, g->cs, g->s[e]);
    const long t = slab_index(g, a, off, e, P->lo, P->hi, plane);
    const int N = g->n[e];
    for (int k = 0; k < g->ndim; ++k) {
      const int at = (k == e ? 0 : 2 * N) + a;
      q[k] = P->profile[N + at] * P->aux[k * size + t] +
             P->profile[at] * strain_term(g, &f, k, e, A, F[e]);
      P->aux[k * size + t] = q[k];
    }
  }
  add_stress(g, sig, gamma, idx, e, q, A);
}

/* elastic.cu:532-573 */
static real force_memory(const elgeom_t *g, const real *sigma, const real *u1, long idx,
                         int c, int e, const elpml_t *P, int phase, const int *A,
                         const real *F) {
  const int a = A[e];
  if (!in_slab(g, a, e, P->lo, P->hi, phase ? g->R : 0))
    return (real)0;
  const long plane = slab_plane(g, e), size = (long)(P->lo + P->hi) * plane;
  const long off = slab_offset(g, e, A);
  const long t = slab_index(g, a, off, e, P->lo, P->hi, plane);
  const int N = g->n[e];
  const int at = (c == e ? 2 * N : 0) + a;
  real *z = P->aux + (long)(g->ndim + c) * size + t;
  line_t k = field_line(sigma, idx, g->cs, g->s[e]);
  k.stress = 1;
  k.e = e;
  if (phase) {
    k.memory = 1;
    k.aux = P->aux;
    k.profile = P->profile;
    k.size = size;
    k.row0 = phase == 1 ? 0 : 2 * g->ndim;
    k.a = a;
    k.off = off;
    k.lo = P->lo;
    k.hi = P->hi;
    k.plane = plane;
    k.own_half = 0;
    const real term = divergence_term(g, &k, c, e, A, F[e]);
    if (t >= 0)
      *z = P->profile[N + at] * (*z + u1[c * g->cs + idx]);
    return term;
  }
  *z = P->profile[N + at] * *z + P->profile[at] * divergence_term(g, &k, c, e, A, F[e]);
  return *z;
}

/* elastic.cu:577-595 */
static void settle_memory(const elgeom_t *g, const real *sigma, long idx, int e,
                          const elpml_t *P, int phase, const int *A) {
  const int a = A[e];
  if (!in_slab(g, a, e, P->lo, P->hi, 0))
    return;
  const long plane = slab_plane(g, e), size = (long)(P->lo + P->hi) * plane;
  const long t = slab_index(g, a, slab_offset(g, e, A), e, P->lo, P->hi, plane);
  const int from = phase == 1 ? 0 : 2 * g->ndim, to = 2 * g->ndim - from;
  const int N = g->n[e];
  for (int k = 0; k < g->ndim; ++k) {
    const int at = (k == e ? 0 : 2 * N) + a;
    P->aux[(to + k) * size + t] =
        P->profile[N + at] *
        (P->aux[(from + k) * size + t] + sigma[stress_row(g, k, e) * g->cs + idx]);
  }
}

/* ------------------------------------------------------------- the sweeps */
static long nodes_of(const elgeom_t *g) {
  long t = 1;
  for (int d = 0; d < g->ndim; ++d)
    t *= g->np[d];
  return t;
}

static int node_of(const elgeom_t *g, long idx, int *A) {
  long rest = idx;
  for (int d = 0; d < g->ndim; ++d) {
    A[d] = (int)(rest / g->s[d]);
    rest -= (long)A[d] * g->s[d];
  }
  for (int d = 0; d < g->ndim; ++d)
    if (!(A[d] > 0 && A[d] < g->n[d] - 1))
      return 0;
  return 1;
}

/* stress_kernel (elastic.cu:756-795): the stresses of u1 into sigma, the
 * pml fold of every lined axis (pml[d].lo + pml[d].hi > 0), and with
 * `gradient` set the stiffness densities of u1 against the forward field v1 */
void FN(el_stress)(const elgeom_t *g, const real *u1, real *sigma, const real *gamma,
                   const double *F_, const elpml_t *pml, int phase, int gradient,
                   real *g_normal, real *g_shear, const real *v1) {
  real F[3];
  for (int d = 0; d < 3; ++d)
    F[d] = (real)F_[d];
  const long total = nodes_of(g);
  const int nv = nvoigt(g);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int A[3];
    if (!node_of(g, idx, A))
      continue;
    real sig[6], eps[6];
    if (gradient) {
      const int zeroed = strain_point(g, u1, idx, A, F, eps);
      /* stiffness_density (elastic.cu:356-383) */
      real eu[3];
      const int z2 = normal_strains(g, v1, idx, A, F, eu);
      g_normal[idx] -= normal_form(g, eps, eu, z2);
      for (int k = 0; k < g->ndim - 1; ++k)
        for (int l = k + 1; l < g->ndim; ++l) {
          if (!shear_point(g, k, l, A))
            continue;
          const int v = stress_row(g, k, l);
          g_shear[(v - g->ndim) * g->cs + idx] -=
              (real)g->mu * eps[v] * shear_strain(g, v1, idx, k, l, A, F);
        }
      stress_of(g, eps, zeroed, gamma, idx, A, sig);
    } else {
      const int zeroed = strain_point(g, u1, idx, A, F, eps);
      stress_of(g, eps, zeroed, gamma, idx, A, sig);
    }
    if (pml)
      for (int e = 0; e < g->ndim; ++e)
        if (pml[e].lo + pml[e].hi)
          stress_memory(g, u1, gamma, idx, e, &pml[e], phase, A, F, sig);
    for (int v = 0; v < nv; ++v)
      sigma[v * g->cs + idx] = sig[v];
  }
}

/* fd_kernel (elastic.cu:798-853): u2 from u0, u1 and the stresses; with
 * `gradient` set the inertia density against v1 into g_mass */
void FN(el_update)(const elgeom_t *g, const real *u0, const real *u1, real *u2,
                   const real *sigma, const real *gamma, const double *F_,
                   const elpml_t *pml, int phase, int gradient, real *g_mass,
                   const real *v1, double mf_) {
  real F[3];
  for (int d = 0; d < 3; ++d)
    F[d] = (real)F_[d];
  const real mf = (real)mf_;
  const long total = nodes_of(g);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int A[3];
    if (!node_of(g, idx, A))
      continue;
    for (int c = 0; c < g->ndim; ++c) {
      if (A[c] > g->n[c] - 3)
        continue;
      const real force = nodal_force(g, sigma, idx, c, A, F);
      const long n = c * g->cs + idx;
      const real mi = point_minv(g, gamma, idx, c, A);
      real next = -u0[n] + (real)2.f * u1[n] + mi * force;
      if (pml)
        for (int e = 0; e < g->ndim; ++e)
          if (pml[e].lo + pml[e].hi)
            next += mi * force_memory(g, sigma, u1, idx, c, e, &pml[e], phase, A, F);
      if (gradient)
        g_mass[n] -= mf * v1[n] * (next - (real)2.f * u1[n] + u0[n]);
      u2[n] = next;
    }
    if (pml && phase)
      for (int e = 0; e < g->ndim; ++e)
        if (pml[e].lo + pml[e].hi)
          settle_memory(g, sigma, idx, e, &pml[e], phase, A);
  }
}

/* gradient_kernel (elastic_sensitivity.cu:115-158) */
void FN(el_gradient)(const elgeom_t *g, real *g_mass, real *g_normal, real *g_shear,
                     const real *u0, const real *u1, const real *u2, const real *l1,
                     double mf_, const double *F_) {
  real F[3];
  for (int d = 0; d < 3; ++d)
    F[d] = (real)F_[d];
  const real mf = (real)mf_;
  const long total = nodes_of(g);
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < total; ++idx) {
    int A[3];
    if (!node_of(g, idx, A))
      continue;
    for (int c = 0; c < g->ndim; ++c) {
      if (A[c] > g->n[c] - 3)
        continue;
      const long n = c * g->cs + idx;
      g_mass[n] -= mf * l1[n] * (u2[n] - (real)2.f * u1[n] + u0[n]);
    }
    real eu[3], el[3];
    const int zeroed = normal_strains(g, u1, idx, A, F, eu);
    normal_strains(g, l1, idx, A, F, el);
    g_normal[idx] -= normal_form(g, el, eu, zeroed);
    for (int k = 0; k < g->ndim - 1; ++k)
      for (int l = k + 1; l < g->ndim; ++l) {
        if (A[k] > g->n[k] - 3 || A[l] > g->n[l] - 3)
          continue;
        const int v = stress_row(g, k, l);
        g_shear[(v - g->ndim) * g->cs + idx] -=
            (real)g->mu * shear_strain(g, l1, idx, k, l, A, F) *
            shear_strain(g, u1, idx, k, l, A, F);
      }
  }
}

/* adjoint_excitation_kernel (cuwave/kernels/common.cuh:467-480), serial */
void FN(el_adjoint_excite)(real *l2, const real *row, const int32_t *lin, int count,
                           const real *weight, real *g_mass, const real *u1, double mf_) {
  const real mf = (real)mf_;
  for (int k = 0; k < count; ++k) {
    const long n = lin[k];
    const real load = weight[k] * row[k];
    l2[n] += load;
    g_mass[n] += -mf * u1[n] * load;
  }
}
