/*
 * orc_impl.h -- body of the oracle, compiled twice (R = float, R = double).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Included by oracle_f32.c and
 * oracle_f64.c with R (the real type) and FN(name) (the suffixing macro)
 * defined.  The code restates the reference's CPU algorithm:
 *   element-local vectors of n_elems x ndof_e entries, the "mask" DOF map,
 *   per-DOF sums taken in element-major order (what std::multimap-based
 *   sort_by_key + reduce_by_key produce, cpu_matrix.cpp:189-223),
 *   sequential dot products (cpu_matrix.cpp:97-105),
 *   batched products in the column-major block convention of bmm
 *   (cpu_matrix.cpp:74-95).
 * The reference's arrays of float indices are int here (exact below 2^24).
 */

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

typedef struct {
  int dim, ne, nn, npe, nde, nv, ndof;
  R *nodes;              /* nn x dim                                       */
  int *elems;            /* ne x npe                                       */
  int *mask;             /* ne x nde: DOF of each element slot (genMask)   */
  R *adj;                /* ndof: number of element slots per DOF          */
  int *ctype;            /* per node constraint code (last line wins)      */
  unsigned char *fixed;  /* per DOF: listed in constraintsIds              */
  int nf;
  int *fnodes, *fadj, *flocal;  /* flocal: local node index in adj elem    */
  R *fnormal, *fp, *flen;
  int nl;
  int *lnode;
  R *lvec;
  R D[36];
  R E, nu;
  const double *escale;
  int convention;        /* 0 reference, 1 physical (oracle_common.c)      */
} FN(mesh);

static void FN(mesh_free)(FN(mesh) *m) {
  free(m->nodes); free(m->elems); free(m->mask); free(m->adj);
  free(m->ctype); free(m->fixed); free(m->fnodes); free(m->fadj);
  free(m->flocal); free(m->fnormal); free(m->fp); free(m->flen);
  free(m->lnode); free(m->lvec);
  memset(m, 0, sizeof(*m));
}

/* datakeeper.cpp:96-114 (CreateMatrixD).  The 2D scale factor goes through
 * double because the reference evaluates pow(nu, 2) in double. */
static void FN(make_D)(int dim, R E, R nu, R *D) {
  if (dim == 2) {
    R s = (R)((double)E / (1.0 - (double)nu * (double)nu));
    D[0] = (R)1; D[1] = nu;     D[2] = (R)0;
    D[3] = nu;   D[4] = (R)1;   D[5] = (R)0;
    D[6] = (R)0; D[7] = (R)0;   D[8] = ((R)1 - nu) / (R)2;
    for (int i = 0; i < 9; ++i) D[i] *= s;
  } else {
    for (int i = 0; i < 36; ++i) D[i] = (R)0;
    R off = nu / ((R)1 - nu);
    R sh = ((R)1 - (R)2 * nu) / ((R)2 * ((R)1 - nu));
    D[0] = D[7] = D[14] = (R)1;
    D[1] = D[6] = D[2] = D[12] = D[13] = D[8] = off;
    D[21] = D[28] = D[35] = sh;
    R s = (E * ((R)1 - nu)) / (((R)1 + nu) * ((R)1 - (R)2 * nu));
    for (int i = 0; i < 36; ++i) D[i] *= s;
  }
}

void FN(orc_D)(int dim, double E, double nu, R *D) {
  FN(make_D)(dim, (R)E, (R)nu, D);
}

/* Load and validate a problem (the dataKeeper reader, datakeeper.cpp:116-213,
 * plus genMask/genAdjElements, cpu_fem_utils.cpp:182-219). */
static int FN(mesh_load)(const orc_problem *pb, FN(mesh) *m) {
  memset(m, 0, sizeof(*m));
  if (!pb || (pb->dim != 2 && pb->dim != 3) || pb->n_nodes <= 0 || pb->n_elems <= 0)
    return ORC_E_ARG;
  int dim = pb->dim;
  m->dim = dim; m->ne = pb->n_elems; m->nn = pb->n_nodes;
  m->npe = dim + 1; m->nde = dim * (dim + 1); m->nv = 3 * (dim - 1);
  m->ndof = dim * pb->n_nodes;
  m->E = (R)pb->E; m->nu = (R)pb->nu;
  m->escale = pb->elem_scale;
  m->convention = 0;
  FN(make_D)(dim, m->E, m->nu, m->D);

  m->nodes = (R *)malloc(sizeof(R) * (size_t)m->nn * dim);
  m->elems = (int *)malloc(sizeof(int) * (size_t)m->ne * m->npe);
  m->mask = (int *)malloc(sizeof(int) * (size_t)m->ne * m->nde);
  m->adj = (R *)calloc((size_t)m->ndof, sizeof(R));
  m->ctype = (int *)calloc((size_t)m->nn, sizeof(int));
  m->fixed = (unsigned char *)calloc((size_t)m->ndof, 1);
  if (!m->nodes || !m->elems || !m->mask || !m->adj || !m->ctype || !m->fixed) {
    FN(mesh_free)(m); return ORC_E_ARG;
  }
  for (size_t i = 0; i < (size_t)m->nn * dim; ++i) m->nodes[i] = (R)pb->nodes[i];
  for (size_t i = 0; i < (size_t)m->ne * m->npe; ++i) {
    int v = pb->elems[i];
    if (v < 0 || v >= m->nn) { FN(mesh_free)(m); return ORC_E_ARG; }
    m->elems[i] = v;
  }
  for (int e = 0; e < m->ne; ++e)
    for (int a = 0; a < m->npe; ++a)
      for (int d = 0; d < dim; ++d) {
        int dof = m->elems[e * m->npe + a] * dim + d;
        m->mask[e * m->nde + a * dim + d] = dof;
        m->adj[dof] += (R)1;
      }
  /* reduce_by_key compacts over the keys present (cpu_matrix.cpp:205-223):
   * a node that no element references would shift every later DOF.  The
   * reference gives garbage there; the oracle refuses such meshes. */
  for (int i = 0; i < m->ndof; ++i)
    if (m->adj[i] == (R)0) { FN(mesh_free)(m); return ORC_E_UNSUPPORTED; }

  /* Constraints: bitwise decode of Constraint::Type (datakeeper.cpp:148-164). */
  for (int c = 0; c < pb->n_constr; ++c) {
    int node = pb->constr_node[c], type = pb->constr_code[c];
    if (node < 0 || node >= m->nn) { FN(mesh_free)(m); return ORC_E_ARG; }
    m->ctype[node] = type;
    if (type & 1) m->fixed[dim * node + 0] = 1;
    if (type & 2) m->fixed[dim * node + 1] = 1;
    if (dim == 3 && (type & 3)) m->fixed[dim * node + 2] = 1;
  }

  /* Boundary faces (datakeeper.cpp:187-213): normals normalised on read. */
  m->nf = pb->n_faces;
  if (m->nf > 0) {
    m->fnodes = (int *)malloc(sizeof(int) * (size_t)m->nf * dim);
    m->fadj = (int *)malloc(sizeof(int) * (size_t)m->nf);
    m->flocal = (int *)malloc(sizeof(int) * (size_t)m->nf * dim);
    m->fnormal = (R *)malloc(sizeof(R) * (size_t)m->nf * dim);
    m->fp = (R *)malloc(sizeof(R) * (size_t)m->nf);
    m->flen = (R *)malloc(sizeof(R) * (size_t)m->nf);
    for (int f = 0; f < m->nf; ++f) {
      int el = pb->face_adj[f];
      if (el < 0 || el >= m->ne) { FN(mesh_free)(m); return ORC_E_ARG; }
      m->fadj[f] = el;
      R len2 = (R)0;
      R nrm[3];
      for (int j = 0; j < dim; ++j) {
        nrm[j] = (R)pb->face_normal[f * dim + j];
        len2 += nrm[j] * nrm[j];
      }
      for (int j = 0; j < dim; ++j) m->fnormal[f * dim + j] = nrm[j] / (R)sqrt(len2);
      m->fp[f] = (R)pb->face_p[f];
      for (int i = 0; i < dim; ++i) {
        int node = pb->face_nodes[f * dim + i];
        m->fnodes[f * dim + i] = node;
        /* getLocalId (cpu_fem_utils.cpp:555-562): first match, else the
         * reference writes out of bounds -> refused here. */
        int loc = -1;
        for (int a = 0; a < m->npe; ++a)
          if (m->elems[el * m->npe + a] == node) { loc = a; break; }
        if (loc < 0) { FN(mesh_free)(m); return ORC_E_UNSUPPORTED; }
        m->flocal[f * dim + i] = loc;
      }
      /* calculateLength / calculateLength3D (cpu_fem_utils.cpp:247-272). */
      const R *p0 = &m->nodes[(size_t)m->fnodes[f * dim + 0] * dim];
      const R *p1 = &m->nodes[(size_t)m->fnodes[f * dim + 1] * dim];
      if (dim == 2) {
        R dx = p1[0] - p0[0], dy = p1[1] - p0[1];
        m->flen[f] = (R)sqrt(dx * dx + dy * dy);
      } else {
        const R *p2 = &m->nodes[(size_t)m->fnodes[f * dim + 2] * dim];
        R ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
        R bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
        R a1 = ay * bz - az * by, a2 = az * bx - ax * bz, a3 = ax * by - ay * bx;
        m->flen[f] = (R)0.5 * (a1 * a1 + a2 * a2 + a3 * a3);
      }
    }
  }
  m->nl = pb->n_loads;
  if (m->nl > 0) {
    m->lnode = (int *)malloc(sizeof(int) * (size_t)m->nl);
    m->lvec = (R *)malloc(sizeof(R) * (size_t)m->nl * dim);
    for (int l = 0; l < m->nl; ++l) {
      int node = pb->load_node[l];
      if (node < 0 || node >= m->nn) { FN(mesh_free)(m); return ORC_E_ARG; }
      m->lnode[l] = node;
      for (int d = 0; d < dim; ++d) m->lvec[l * dim + d] = (R)pb->load_vec[l * dim + d];
    }
  }
  return ORC_OK;
}

/* Cofactor-expansion determinants in the reference's evaluation order
 * (cpu_matrix.cpp:237-267, cpu_fem_utils.cpp:290-299). */
static R FN(det3)(R a0, R a1, R a2, R a3, R a4, R a5, R a6, R a7, R a8) {
  return a0 * a4 * a8 + a1 * a6 * a5 + a2 * a3 * a7 - a6 * a4 * a2 - a0 * a5 * a7 - a1 * a3 * a8;
}
static R FN(det4)(const R *m) {
  R v1 = FN(det3)(m[5], m[6], m[7], m[9], m[10], m[11], m[13], m[14], m[15]);
  R v2 = FN(det3)(m[1], m[2], m[3], m[9], m[10], m[11], m[13], m[14], m[15]);
  R v3 = FN(det3)(m[1], m[2], m[3], m[5], m[6], m[7], m[13], m[14], m[15]);
  R v4 = FN(det3)(m[1], m[2], m[3], m[5], m[6], m[7], m[9], m[10], m[11]);
  return v1 - v2 + v3 - v4;
}

/* genCoordinates + calculateArea + genGradientMatrix{,3D} for one element
 * (cpu_fem_utils.cpp:221-231, 274-288, 338-435).  B: nv x nde row-major,
 * already multiplied by 1/|det C|; returns |det C| (= 2A or 6V) and the
 * signed determinant in *sdet. */
static R FN(elem_B)(const FN(mesh) *m, int el, R *B, R *sdet) {
  int dim = m->dim, npe = m->npe, nde = m->nde;
  R coords[12];  /* dim x npe row-major: coords[i*npe + j] = coord i of node j */
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < npe; ++j)
      coords[i * npe + j] = m->nodes[(size_t)m->elems[el * npe + j] * dim + i];
  R C[16];
  for (int i = 0; i < npe; ++i)
    for (int j = 0; j < npe; ++j)
      C[j + i * npe] = j == 0 ? (R)1 : coords[(j - 1) * npe + i];
  R det = dim == 2 ? FN(det3)(C[0], C[1], C[2], C[3], C[4], C[5], C[6], C[7], C[8])
                   : FN(det4)(C);
  if (sdet) *sdet = det;
  R area = (R)fabs(det);
  for (int i = 0; i < m->nv * nde; ++i) B[i] = (R)0;
  if (dim == 2) {
#define CO(r, c) coords[(r) * 3 + (c)]
    B[0 * 6 + 0] = CO(1, 1) - CO(1, 2);
    B[0 * 6 + 2] = CO(1, 2) - CO(1, 0);
    B[0 * 6 + 4] = CO(1, 0) - CO(1, 1);
    B[1 * 6 + 1] = CO(0, 2) - CO(0, 1);
    B[1 * 6 + 3] = CO(0, 0) - CO(0, 2);
    B[1 * 6 + 5] = CO(0, 1) - CO(0, 0);
    B[2 * 6 + 0] = CO(0, 2) - CO(0, 1);
    B[2 * 6 + 1] = CO(1, 1) - CO(1, 2);
    B[2 * 6 + 2] = CO(0, 0) - CO(0, 2);
    B[2 * 6 + 3] = CO(1, 2) - CO(1, 0);
    B[2 * 6 + 4] = CO(0, 1) - CO(0, 0);
    B[2 * 6 + 5] = CO(1, 0) - CO(1, 1);
#undef CO
  } else {
#define X(k) coords[((k) % 4) + 0]
#define Y(k) coords[((k) % 4) + 4]
#define Z(k) coords[((k) % 4) + 8]
    for (int i = 0; i < 4; ++i) {
      R sg = (i % 2 == 0) ? (R)1 : (R)-1;
      R b = sg * FN(det3)((R)1, Y(i + 3), Z(i + 3), (R)1, Y(i + 2), Z(i + 2), (R)1, Y(i + 1), Z(i + 1));
      R c = sg * FN(det3)(X(i + 3), (R)1, Z(i + 3), X(i + 2), (R)1, Z(i + 2), X(i + 1), (R)1, Z(i + 1));
      R d = sg * FN(det3)(X(i + 3), Y(i + 3), (R)1, X(i + 2), Y(i + 2), (R)1, X(i + 1), Y(i + 1), (R)1);
      int s = 3 * i;
      B[0 * 12 + s + 0] = b; B[3 * 12 + s + 1] = b; B[5 * 12 + s + 2] = b;
      B[3 * 12 + s + 0] = c; B[1 * 12 + s + 1] = c; B[4 * 12 + s + 2] = c;
      B[5 * 12 + s + 0] = d; B[4 * 12 + s + 1] = d; B[2 * 12 + s + 2] = d;
    }
#undef X
#undef Y
#undef Z
  }
  R inv = (R)1 / area;   /* divideElementwise(areas, X): multiply by 1/a */
  for (int i = 0; i < m->nv * nde; ++i) B[i] *= inv;
  return area;
}

/* calculateKlocal + applyConstraints for one element (cpu_fem_utils.cpp:
 * 461-503).  K: nde x nde column-major. */
static void FN(elem_K)(const FN(mesh) *m, int el, R *K, R *area_out) {
  int nde = m->nde, nv = m->nv;
  R B[72], tB[72];
  R area = FN(elem_B)(m, el, B, NULL);
  /* tB = D * B^T-read (bmm(D, nv, nv, false, B, nde, true)): column-major */
  for (int mm = 0; mm < nv; ++mm)
    for (int n = 0; n < nde; ++n) {
      R acc = (R)0;
      for (int k = 0; k < nv; ++k) acc += m->D[mm + k * nv] * B[n + k * nde];
      tB[mm + n * nv] = (R)1 * acc;
    }
  R coeff = m->dim == 2 ? (R)0.5 : (R)((R)1 / (R)6);
  /* K = coeff * B^T tB (bmm(B, nde, nv, false, tB, nde, false, coeff)) */
  for (int mm = 0; mm < nde; ++mm)
    for (int n = 0; n < nde; ++n) {
      R acc = (R)0;
      for (int k = 0; k < nv; ++k) acc += B[mm + k * nde] * tB[k + n * nv];
      K[mm + n * nde] = coeff * acc;
    }
  for (int i = 0; i < nde * nde; ++i) K[i] *= area;     /* scale(areas, X) */
  /* applyConstraints: zero (r,c), r != c as global DOFs, when either DOF is
   * in constraintsIds.  A per-DOF flag gives the same zero pattern as the
   * reference's loop over the constraint list. */
  const int *mk = &m->mask[(size_t)el * nde];
  for (int r = 0; r < nde; ++r)
    for (int c = 0; c < nde; ++c)
      if ((m->fixed[mk[r]] || m->fixed[mk[c]]) && mk[r] != mk[c]) K[c + r * nde] = (R)0;
  if (m->escale) {
    R s = (R)m->escale[el];
    for (int i = 0; i < nde * nde; ++i) K[i] *= s;
  }
  if (area_out) *area_out = area;
}

static R *FN(all_K)(const FN(mesh) *m, R *areas) {
  size_t blk = (size_t)m->nde * m->nde;
  R *K = (R *)malloc(sizeof(R) * blk * m->ne);
  if (!K) return NULL;
  for (int e = 0; e < m->ne; ++e) FN(elem_K)(m, e, K + blk * e, areas ? &areas[e] : NULL);
  return K;
}

int FN(orc_klocals)(const orc_problem *pb, R *K_out, R *areas_out) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  size_t blk = (size_t)m.nde * m.nde;
  for (int e = 0; e < m.ne; ++e) FN(elem_K)(&m, e, K_out + blk * e, areas_out ? &areas_out[e] : NULL);
  FN(mesh_free)(&m);
  return ORC_OK;
}

/* Aᵀ: per-DOF sums of an element-local array, element-major order
 * (reductionWithMask, cpu_fem_utils.cpp:437-445). */
static void FN(assemble)(const FN(mesh) *m, const R *loc, R *glob) {
  for (int i = 0; i < m->ndof; ++i) glob[i] = (R)0;
  size_t n = (size_t)m->ne * m->nde;
  for (size_t i = 0; i < n; ++i) glob[m->mask[i]] += loc[i];
}
/* A: gather (transformWithMask, cpu_fem_utils.cpp:447-452). */
static void FN(gather)(const FN(mesh) *m, const R *glob, R *loc) {
  size_t n = (size_t)m->ne * m->nde;
  for (size_t i = 0; i < n; ++i) loc[i] = glob[m->mask[i]];
}
/* The in-place side effect of sort_by_key on the reduced array
 * (cpu_matrix.cpp:189-203): a stable sort of the values by DOF key. */
static void FN(sort_by_mask)(const FN(mesh) *m, R *loc, R *tmp, const size_t *perm) {
  size_t n = (size_t)m->ne * m->nde;
  for (size_t i = 0; i < n; ++i) tmp[i] = loc[perm[i]];
  memcpy(loc, tmp, sizeof(R) * n);
}
static size_t *FN(mask_perm)(const FN(mesh) *m) {
  /* stable counting sort of slot indices by DOF key */
  size_t n = (size_t)m->ne * m->nde;
  size_t *perm = (size_t *)malloc(sizeof(size_t) * n);
  size_t *start = (size_t *)calloc((size_t)m->ndof + 1, sizeof(size_t));
  if (!perm || !start) { free(perm); free(start); return NULL; }
  for (size_t i = 0; i < n; ++i) start[m->mask[i] + 1]++;
  for (int d = 0; d < m->ndof; ++d) start[d + 1] += start[d];
  for (size_t i = 0; i < n; ++i) perm[start[m->mask[i]]++] = i;
  free(start);
  return perm;
}
static R FN(dot)(const R *a, const R *b, size_t n) {
  R s = (R)0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
/* u_e = p_eᵀ K_e (bmm(p, 1, nde, true, K, nde, false), cpu_matrix.cpp:74-95). */
static void FN(ebe_mv)(const FN(mesh) *m, const R *K, const R *p, R *u) {
  int nde = m->nde;
  size_t blk = (size_t)nde * nde;
  for (int e = 0; e < m->ne; ++e) {
    const R *Ke = K + blk * e;
    const R *pe = p + (size_t)e * nde;
    R *ue = u + (size_t)e * nde;
    for (int n = 0; n < nde; ++n) {
      R acc = (R)0;
      for (int k = 0; k < nde; ++k) acc += pe[k] * Ke[k + n * nde];
      ue[n] = (R)1 * acc;
    }
  }
}

/* updateWavelet (fem_utils.cpp:53-82), in float (as the reference) for the
 * FP32 restatement and in double for the FP64 one. */
static R FN(wavelet)(double t, const orc_wavelet *w) {
  return sizeof(R) == sizeof(float) ? (R)orc_wavelet_value(t, w) : (R)orc_wavelet_value_f64(t, w);
}

/* Element-local RHS state: pressures can be overwritten by the wavelet value
 * and stay overwritten (calcPressures, cpu_fem_utils.cpp:564-582). */
typedef struct {
  R *F;      /* ne x nde */
  R *press;  /* nf */
  R *loads;  /* ndof (only when nl > 0) */
} FN(rhs);

static int FN(rhs_init)(const FN(mesh) *m, FN(rhs) *r) {
  r->F = (R *)calloc((size_t)m->ne * m->nde, sizeof(R));
  r->press = m->nf > 0 ? (R *)malloc(sizeof(R) * (size_t)m->nf) : NULL;
  r->loads = m->nl > 0 ? (R *)calloc((size_t)m->ndof, sizeof(R)) : NULL;
  if (!r->F || (m->nf > 0 && !r->press) || (m->nl > 0 && !r->loads)) return ORC_E_ARG;
  if (m->nf > 0) memcpy(r->press, m->fp, sizeof(R) * (size_t)m->nf);
  return ORC_OK;
}
static void FN(rhs_free)(FN(rhs) *r) { free(r->F); free(r->press); free(r->loads); }

/* calcFlocals (cpu_fem_utils.cpp:532-596). */
static void FN(rhs_calc)(const FN(mesh) *m, FN(rhs) *r, double t, const orc_wavelet *w) {
  int dim = m->dim, nde = m->nde;
  if (m->nf > 0) {
    R v = FN(wavelet)(t, w);
    if (v != (R)0)
      for (int f = 0; f < m->nf; ++f) r->press[f] = v;
    R coeff = dim == 2 ? (R)-0.5 : (R)((R)-1 / (R)12);
    for (int f = 0; f < m->nf; ++f) {
      int el = m->fadj[f];
      for (int i = 0; i < dim; ++i) {
        int node = m->fnodes[f * dim + i];
        int loc = m->flocal[f * dim + i];
        for (int j = 0; j < dim; ++j)
          r->F[(size_t)el * nde + dim * loc + j] =
              (m->ctype[node] & j) ? (R)0
                                   : coeff * r->press[f] * m->flen[f] * m->fnormal[f * dim + j];
      }
    }
  }
  if (m->nl > 0) {
    for (int l = 0; l < m->nl; ++l)
      for (int d = 0; d < dim; ++d) r->loads[dim * m->lnode[l] + d] = m->lvec[l * dim + d];
    FN(gather)(m, r->loads, r->F);
  }
}

int FN(orc_flocals)(const orc_problem *pb, double t, const orc_wavelet *w, R *F_out) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  FN(rhs) r;
  if ((st = FN(rhs_init)(&m, &r)) == ORC_OK) {
    FN(rhs_calc)(&m, &r, t, w);
    memcpy(F_out, r.F, sizeof(R) * (size_t)m.ne * m.nde);
  }
  FN(rhs_free)(&r);
  FN(mesh_free)(&m);
  return st;
}

int FN(orc_rhs)(const orc_problem *pb, R *f_out) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  FN(rhs) r;
  if ((st = FN(rhs_init)(&m, &r)) == ORC_OK) {
    orc_wavelet none = {0, 15.0, 0.0, 1.0};
    FN(rhs_calc)(&m, &r, 0.0, &none);
    FN(assemble)(&m, r.F, f_out);
  }
  FN(rhs_free)(&r);
  FN(mesh_free)(&m);
  return st;
}

int FN(orc_diag)(const orc_problem *pb, R *d_out) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  int nde = m.nde;
  R Ke[144];
  R *dl = (R *)malloc(sizeof(R) * (size_t)m.ne * nde);
  if (!dl) { FN(mesh_free)(&m); return ORC_E_ARG; }
  for (int e = 0; e < m.ne; ++e) {
    FN(elem_K)(&m, e, Ke, NULL);
    for (int j = 0; j < nde; ++j) dl[(size_t)e * nde + j] = Ke[nde * j + j];
  }
  FN(assemble)(&m, dl, d_out);
  free(dl);
  FN(mesh_free)(&m);
  return ORC_OK;
}

int FN(orc_matvec)(const orc_problem *pb, const R *x, R *y) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  int nde = m.nde;
  R Ke[144], xe[12];
  for (int i = 0; i < m.ndof; ++i) y[i] = (R)0;
  for (int e = 0; e < m.ne; ++e) {
    FN(elem_K)(&m, e, Ke, NULL);
    const int *mk = &m.mask[(size_t)e * nde];
    for (int k = 0; k < nde; ++k) xe[k] = x[mk[k]];
    for (int n = 0; n < nde; ++n) {
      R acc = (R)0;
      for (int k = 0; k < nde; ++k) acc += xe[k] * Ke[k + n * nde];
      y[mk[n]] += acc;
    }
  }
  FN(mesh_free)(&m);
  return ORC_OK;
}

/* The PCG vectors of CGMData (fem_utils.cpp:3-29). */
typedef struct { R *r, *m, *z, *s, *p, *u, *x, *tmpg, *tmpl; } FN(cg);

static int FN(cg_alloc)(const FN(mesh) *m, FN(cg) *c) {
  size_t n = (size_t)m->ne * m->nde;
  c->r = (R *)calloc(n, sizeof(R)); c->m = (R *)calloc(n, sizeof(R));
  c->z = (R *)calloc(n, sizeof(R)); c->s = (R *)calloc(n, sizeof(R));
  c->p = (R *)calloc(n, sizeof(R)); c->u = (R *)calloc(n, sizeof(R));
  c->x = (R *)calloc(n, sizeof(R)); c->tmpl = (R *)calloc(n, sizeof(R));
  c->tmpg = (R *)calloc((size_t)m->ndof, sizeof(R));
  return (c->r && c->m && c->z && c->s && c->p && c->u && c->x && c->tmpg && c->tmpl)
             ? ORC_OK : ORC_E_ARG;
}
static void FN(cg_free)(FN(cg) *c) {
  free(c->r); free(c->m); free(c->z); free(c->s); free(c->p); free(c->u);
  free(c->x); free(c->tmpg); free(c->tmpl);
}

/* gpuPCG_EbE_vec (femfunc.cpp:25-128).  r holds the initial residual, x is
 * zero on entry.  diagK is the element-local diagonal; when perm is given,
 * diagK is permuted in place as the reference's sort_by_key does.  Returns
 * ORC_OK / ORC_E_NONFINITE / ORC_E_MAXIT. */
static int FN(pcg)(const FN(mesh) *m, const R *K, R *diagK, FN(cg) *c, R eps,
                   int max_iter, const size_t *perm, int *n_iter_out) {
  size_t n = (size_t)m->ne * m->nde;
  /* (0b) m = A Aᵀ diagK */
  FN(assemble)(m, diagK, c->tmpg);
  if (perm) FN(sort_by_mask)(m, diagK, c->tmpl, perm);
  FN(gather)(m, c->tmpg, c->m);
  /* (0c) z = r / m */
  for (size_t i = 0; i < n; ++i) c->z[i] = c->r[i] / c->m[i];
  /* (0d) s = A Aᵀ z */
  FN(assemble)(m, c->z, c->tmpg);
  FN(gather)(m, c->tmpg, c->s);
  R gamma0 = FN(dot)(c->r, c->s, n);
  R gamma = gamma0, gamma_new = (R)0;
  memcpy(c->p, c->s, sizeof(R) * n);   /* (0e) */
  int it = 0;
  int st = ORC_OK;
  for (;;) {
    ++it;
    FN(ebe_mv)(m, K, c->p, c->u);                      /* (1a) */
    R sumElem = FN(dot)(c->p, c->u, n);                /* (1b) */
    R alpha = gamma / sumElem;                         /* (1c) */
    for (size_t i = 0; i < n; ++i) c->x[i] = (R)1 * c->x[i] + alpha * c->p[i];
    R malpha = (R)-1 * alpha;
    for (size_t i = 0; i < n; ++i) c->r[i] = (R)1 * c->r[i] + malpha * c->u[i];
    for (size_t i = 0; i < n; ++i) c->z[i] = c->r[i] / c->m[i];
    FN(assemble)(m, c->z, c->tmpg);
    FN(gather)(m, c->tmpg, c->s);
    gamma_new = FN(dot)(c->r, c->s, n);
    if (gamma_new < eps * gamma0) break;               /* (5) */
    R beta = gamma_new / gamma;
    for (size_t i = 0; i < n; ++i) c->p[i] = beta * c->p[i] + (R)1 * c->s[i];
    gamma = gamma_new;
    if (!isfinite(gamma)) { st = ORC_E_NONFINITE; break; }
    if (max_iter > 0 && it >= max_iter) { st = ORC_E_MAXIT; break; }
  }
  if (n_iter_out) *n_iter_out = it;
  return st;
}

/* gpuCalculateFEM_EbE_vec (femfunc.cpp:3-23). */
int FN(orc_static_solve)(const orc_problem *pb, double eps, int max_iter, R *u_out, int *n_iter) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  size_t n = (size_t)m.ne * m.nde, blk = (size_t)m.nde * m.nde;
  R *K = FN(all_K)(&m, NULL);
  R *diagK = (R *)malloc(sizeof(R) * n);
  FN(cg) c;
  FN(rhs) rh;
  memset(&rh, 0, sizeof(rh));
  if (!K || !diagK || FN(cg_alloc)(&m, &c) || FN(rhs_init)(&m, &rh)) {
    st = ORC_E_ARG; goto done;
  }
  for (int e = 0; e < m.ne; ++e)
    for (int j = 0; j < m.nde; ++j) diagK[(size_t)e * m.nde + j] = K[blk * e + m.nde * j + j];
  {
    orc_wavelet none = {0, 15.0, 0.0, 1.0};
    FN(rhs_calc)(&m, &rh, 0.0, &none);
  }
  memcpy(c.r, rh.F, sizeof(R) * n);                   /* (0a) */
  st = FN(pcg)(&m, K, diagK, &c, (R)eps, max_iter, NULL, n_iter);
  /* assemble(x) / adjElements (femfunc.cpp:113-117) */
  FN(assemble)(&m, c.x, c.tmpg);
  for (int i = 0; i < m.ndof; ++i) u_out[i] = c.tmpg[i] / m.adj[i];
done:
  FN(cg_free)(&c);
  FN(rhs_free)(&rh);
  free(K); free(diagK);
  FN(mesh_free)(&m);
  return st;
}

/* calculateMlocals (cpu_fem_utils.cpp:598-734): lumped (ne x nde) or the
 * diagonal of the consistent matrix (the only part the reference uses). */
static void FN(mass_diag)(const FN(mesh) *m, R rho, int lumped, int physical, R *M) {
  int dim = m->dim, nde = m->nde, npe = m->npe;
  for (int e = 0; e < m->ne; ++e) {
    R val;
    if (dim == 2) {
      R c[6];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 3; ++j)
          c[i * 3 + j] = m->nodes[(size_t)m->elems[e * npe + j] * dim + i];
      R area = (R)0.5 * (R)fabs((c[0] - c[2]) * (c[4] - c[5]) - (c[1] - c[2]) * (c[3] - c[5]));
      R mass = rho * area;
      val = lumped ? mass / (R)3 : mass / (R)6;
    } else {
      R B[72];
      R area = FN(elem_B)(m, e, B, NULL);
      R mass = rho * area / (R)6;
      if (physical) val = lumped ? mass / (R)4 : mass / (R)10;
      else val = lumped ? mass / (R)12 : mass / (R)24;
    }
    for (int j = 0; j < nde; ++j) M[(size_t)e * nde + j] = val;
  }
}

/* gpuCalculateFEM_DYN (femfunc.cpp:130-381) with damping off, as in the
 * reference (isDamping = false, femfunc.cpp:151). */
int FN(orc_dyn_solve)(const orc_problem *pb, const orc_dyn_params *dp, R *u_out, int *n_steps) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  size_t n = (size_t)m.ne * m.nde, blk = (size_t)m.nde * m.nde;
  R beta1 = (R)dp->beta1, beta2 = (R)dp->beta2, dt = (R)dp->dt, endtime = (R)dp->endtime;
  R rho = (R)dp->rho;
  int isExplicit = beta2 == (R)0;
  R *K = FN(all_K)(&m, NULL);
  R *M = (R *)malloc(sizeof(R) * n);
  R *diagK = (R *)malloc(sizeof(R) * n);
  R *displ = (R *)calloc(n, sizeof(R));
  R *vel = (R *)calloc(n, sizeof(R));
  R *Mass_g = (R *)malloc(sizeof(R) * (size_t)m.ndof);
  size_t *perm = dp->compat_scramble ? FN(mask_perm)(&m) : NULL;
  FN(cg) c;
  FN(rhs) rh;
  memset(&c, 0, sizeof(c));
  memset(&rh, 0, sizeof(rh));
  if (!K || !M || !diagK || !displ || !vel || !Mass_g || (dp->compat_scramble && !perm) ||
      FN(cg_alloc)(&m, &c) || FN(rhs_init)(&m, &rh)) {
    st = ORC_E_ARG; goto done;
  }
  FN(mass_diag)(&m, rho, isExplicit, 0, M);
  FN(rhs_calc)(&m, &rh, 0.0, &dp->wavelet);            /* initFlocals(0) */
  if (!isExplicit) {
    R cK = (R)0.5 * beta2 * dt * dt, cM = (R)1;
    for (int e = 0; e < m.ne; ++e)
      for (int j = 0; j < m.nde; ++j) {
        size_t i = (size_t)e * m.nde + j;
        diagK[i] = cK * K[blk * e + m.nde * j + j] + cM * M[i];
      }
  }
  int endnt = endtime < (R)0 ? 1 : (int)(endtime / dt);
  int nt = 1;
  R *x_dyn = c.x;
  for (;;) {
    R t = (R)nt * dt;
    R c1 = (R)0.5 * ((R)1 - beta2) * dt * dt;
    R c2 = ((R)1 - beta1) * dt;
    for (size_t i = 0; i < n; ++i) displ[i] = (R)1 * displ[i] + dt * vel[i];
    for (size_t i = 0; i < n; ++i) displ[i] = (R)1 * displ[i] + c1 * x_dyn[i];
    for (size_t i = 0; i < n; ++i) vel[i] = (R)1 * vel[i] + c2 * x_dyn[i];
    for (size_t i = 0; i < n; ++i) x_dyn[i] = (R)0;
    FN(ebe_mv)(&m, K, displ, c.r);                      /* r = d_eᵀ K_e */
    FN(rhs_calc)(&m, &rh, (double)t, &dp->wavelet);
    for (size_t i = 0; i < n; ++i) c.r[i] = (R)-1 * c.r[i] + (R)1 * rh.F[i];
    if (isExplicit) {
      /* solveDiagSystem(Mlocals, r, x_dyn, false) (cpu_fem_utils.cpp:736-750) */
      FN(assemble)(&m, M, Mass_g);
      if (perm) FN(sort_by_mask)(&m, M, c.tmpl, perm);
      FN(assemble)(&m, c.r, c.tmpg);
      for (int i = 0; i < m.ndof; ++i) c.tmpg[i] = c.tmpg[i] / Mass_g[i];
      FN(gather)(&m, c.tmpg, x_dyn);
    } else {
      int it = 0;
      const R peps = dp->pcg_eps > 0 ? (R)dp->pcg_eps : (R)1e-4f;
      st = FN(pcg)(&m, K, diagK, &c, peps, 100000, perm, &it);
      if (st) goto done;
    }
    R c3 = beta1 * dt;
    for (size_t i = 0; i < n; ++i) vel[i] = (R)1 * vel[i] + c3 * x_dyn[i];
    if (!isExplicit) {
      R c4 = (R)0.5 * beta2 * dt * dt;
      for (size_t i = 0; i < n; ++i) displ[i] = (R)1 * displ[i] + c4 * x_dyn[i];
    }
    ++nt;
    if (nt > endnt) break;
  }
  if (n_steps) *n_steps = nt - 1;
  FN(assemble)(&m, displ, c.tmpg);
  for (int i = 0; i < m.ndof; ++i) u_out[i] = c.tmpg[i] / m.adj[i];
done:
  FN(cg_free)(&c);
  FN(rhs_free)(&rh);
  free(K); free(M); free(diagK); free(displ); free(vel); free(Mass_g); free(perm);
  FN(mesh_free)(&m);
  return st;
}

/* Per-element stress sigma = D B u_e with the signed element determinant
 * (the legacy fixtures' TENSORS stress; deprecated/femfunc/femfunc.cpp:1311-1400
 * computed it from B and D the same way). */
int FN(orc_stress)(const orc_problem *pb, const R *u, R *sigma_out) {
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  int nde = m.nde, nv = m.nv;
  R B[72], eps[6];
  for (int e = 0; e < m.ne; ++e) {
    R sdet;
    FN(elem_B)(&m, e, B, &sdet);
    R sgn = sdet < (R)0 ? (R)-1 : (R)1;
    const int *mk = &m.mask[(size_t)e * nde];
    for (int r = 0; r < nv; ++r) {
      R acc = (R)0;
      for (int k = 0; k < nde; ++k) acc += B[r * nde + k] * u[mk[k]];
      eps[r] = sgn * acc;
    }
    R sc = m.escale ? (R)m.escale[e] : (R)1;
    for (int r = 0; r < nv; ++r) {
      R acc = (R)0;
      for (int k = 0; k < nv; ++k) acc += m.D[r * nv + k] * eps[k];
      sigma_out[(size_t)e * nv + r] = sc * acc;
    }
  }
  FN(mesh_free)(&m);
  return ORC_OK;
}

#ifdef ORC_WITH_TIMER
static double FN(now)(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int FN(orc_static_time)(const orc_problem *pb, int n_iter, double *setup_s, double *iter_s) {
  double t0 = FN(now)();
  FN(mesh) m;
  int st = FN(mesh_load)(pb, &m);
  if (st) return st;
  size_t n = (size_t)m.ne * m.nde, blk = (size_t)m.nde * m.nde;
  R *K = FN(all_K)(&m, NULL);
  R *diagK = (R *)malloc(sizeof(R) * n);
  FN(cg) c;
  FN(rhs) rh;
  memset(&c, 0, sizeof(c));
  memset(&rh, 0, sizeof(rh));
  if (!K || !diagK || FN(cg_alloc)(&m, &c) || FN(rhs_init)(&m, &rh)) {
    st = ORC_E_ARG; goto done;
  }
  for (int e = 0; e < m.ne; ++e)
    for (int j = 0; j < m.nde; ++j) diagK[(size_t)e * m.nde + j] = K[blk * e + m.nde * j + j];
  {
    orc_wavelet none = {0, 15.0, 0.0, 1.0};
    FN(rhs_calc)(&m, &rh, 0.0, &none);
  }
  memcpy(c.r, rh.F, sizeof(R) * n);
  double t1 = FN(now)();
  int it = 0;
  /* eps = 0: never stops early; exactly n_iter iterations */
  st = FN(pcg)(&m, K, diagK, &c, (R)0, n_iter, NULL, &it);
  if (st == ORC_E_MAXIT) st = ORC_OK;
  double t2 = FN(now)();
  if (setup_s) *setup_s = t1 - t0;
  if (iter_s) *iter_s = t2 - t1;
done:
  FN(cg_free)(&c);
  FN(rhs_free)(&rh);
  free(K); free(diagK);
  FN(mesh_free)(&m);
  return st;
}
#endif
