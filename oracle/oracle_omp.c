/* "Efficient CPU" timing baseline: the FP64 restatement's static PCG with its
 * loops spread over OpenMP threads.  TEST INFRASTRUCTURE ONLY (see oracle.h):
 * bench.py times it beside the single-core restatement; nothing checks
 * results against it and the product never calls it.
 *
 * Same algorithm and data layout as orc_impl.h's pcg (gpuPCG_EbE_vec,
 * femfunc.cpp:25-128): stored K_e, element-local vectors, per-DOF sums.  The
 * per-DOF sums run one DOF per iteration over the DOF's slots in element
 * order (the order of the sequential assemble, so bitwise the same sums);
 * the dot products are OpenMP reductions (their rounding differs from the
 * sequential sums). */
#define _POSIX_C_SOURCE 199309L
#define R double
#define FN(x) x##_omp64
#include "orc_impl.h"

#include <omp.h>

typedef struct {
  size_t *perm;   /* slot indices sorted (stably) by DOF */
  size_t *start;  /* ndof + 1 offsets into perm          */
} omp_dofmap;

static int omp_dofmap_build(const mesh_omp64 *m, omp_dofmap *d) {
  size_t n = (size_t)m->ne * m->nde;
  d->perm = (size_t *)malloc(sizeof(size_t) * n);
  d->start = (size_t *)calloc((size_t)m->ndof + 1, sizeof(size_t));
  size_t *next = (size_t *)malloc(sizeof(size_t) * ((size_t)m->ndof + 1));
  if (!d->perm || !d->start || !next) {
    free(d->perm); free(d->start); free(next);
    return ORC_E_ARG;
  }
  for (size_t i = 0; i < n; ++i) d->start[m->mask[i] + 1]++;
  for (int k = 0; k < m->ndof; ++k) d->start[k + 1] += d->start[k];
  memcpy(next, d->start, sizeof(size_t) * ((size_t)m->ndof + 1));
  for (size_t i = 0; i < n; ++i) d->perm[next[m->mask[i]]++] = i;
  free(next);
  return ORC_OK;
}

static void omp_assemble(const mesh_omp64 *m, const omp_dofmap *d, const R *loc, R *glob) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < m->ndof; ++k) {
    R s = 0.0;
    for (size_t j = d->start[k]; j < d->start[k + 1]; ++j) s += loc[d->perm[j]];
    glob[k] = s;
  }
}

static void omp_gather(const mesh_omp64 *m, const R *glob, R *loc) {
  const long n = (long)m->ne * m->nde;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) loc[i] = glob[m->mask[i]];
}

static R omp_dot(const R *a, const R *b, long n) {
  R s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* u_e = p_e' K_e per element (the column-major block product of ebe_mv) */
static void omp_ebe_mv(const mesh_omp64 *m, const R *K, const R *p, R *u) {
  const int nde = m->nde;
  const size_t blk = (size_t)nde * nde;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m->ne; ++e) {
    const R *Ke = K + blk * e;
    const R *pe = p + (size_t)e * nde;
    R *ue = u + (size_t)e * nde;
    for (int c = 0; c < nde; ++c) {
      R acc = 0.0;
      for (int k = 0; k < nde; ++k) acc += pe[k] * Ke[k + c * nde];
      ue[c] = acc;
    }
  }
}

/* The iteration of orc_impl.h's pcg with eps = 0: exactly n_iter steps. */
static void omp_pcg(const mesh_omp64 *m, const omp_dofmap *d, const R *K, const R *diagK,
                    cg_omp64 *c, int n_iter) {
  const long n = (long)m->ne * m->nde;
  omp_assemble(m, d, diagK, c->tmpg);
  omp_gather(m, c->tmpg, c->m);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) c->z[i] = c->r[i] / c->m[i];
  omp_assemble(m, d, c->z, c->tmpg);
  omp_gather(m, c->tmpg, c->s);
  R gamma = omp_dot(c->r, c->s, n);
  memcpy(c->p, c->s, sizeof(R) * (size_t)n);
  for (int it = 0; it < n_iter; ++it) {
    omp_ebe_mv(m, K, c->p, c->u);
    const R alpha = gamma / omp_dot(c->p, c->u, n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      c->x[i] += alpha * c->p[i];
      c->r[i] -= alpha * c->u[i];
      c->z[i] = c->r[i] / c->m[i];
    }
    omp_assemble(m, d, c->z, c->tmpg);
    omp_gather(m, c->tmpg, c->s);
    const R gamma_new = omp_dot(c->r, c->s, n);
    const R beta = gamma_new / gamma;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) c->p[i] = beta * c->p[i] + c->s[i];
    gamma = gamma_new;
  }
}

static double omp_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_static_time_omp_f64(const orc_problem *pb, int n_iter, int n_threads, double *setup_s,
                            double *iter_s) {
  if (n_threads > 0) omp_set_num_threads(n_threads);
  double t0 = omp_now();
  mesh_omp64 m;
  int st = mesh_load_omp64(pb, &m);
  if (st) return st;
  const size_t n = (size_t)m.ne * m.nde, blk = (size_t)m.nde * m.nde;
  R *K = (R *)malloc(sizeof(R) * blk * (size_t)m.ne);
  R *diagK = (R *)malloc(sizeof(R) * n);
  cg_omp64 c;
  rhs_omp64 rh;
  omp_dofmap d = {NULL, NULL};
  memset(&c, 0, sizeof(c));
  memset(&rh, 0, sizeof(rh));
  if (!K || !diagK || cg_alloc_omp64(&m, &c) || rhs_init_omp64(&m, &rh) ||
      omp_dofmap_build(&m, &d)) {
    st = ORC_E_ARG;
    goto done;
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m.ne; ++e) {
    elem_K_omp64(&m, e, K + blk * e, NULL);
    for (int j = 0; j < m.nde; ++j) diagK[(size_t)e * m.nde + j] = K[blk * e + m.nde * j + j];
  }
  {
    orc_wavelet none = {0, 15.0, 0.0, 1.0};
    rhs_calc_omp64(&m, &rh, 0.0, &none);
  }
  memcpy(c.r, rh.F, sizeof(R) * n);
  double t1 = omp_now();
  omp_pcg(&m, &d, K, diagK, &c, n_iter);
  double t2 = omp_now();
  if (setup_s) *setup_s = t1 - t0;
  if (iter_s) *iter_s = t2 - t1;
done:
  free(d.perm);
  free(d.start);
  cg_free_omp64(&c);
  rhs_free_omp64(&rh);
  free(K);
  free(diagK);
  mesh_free_omp64(&m);
  return st;
}
