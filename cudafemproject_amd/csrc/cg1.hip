// cg1.hip -- single-reduction Jacobi-PCG: one launch per iteration.
//
// The reference's PCG loop (femfunc.cpp:59-111) needs two global reductions
// per iteration (p'Ap for alpha, r'z for beta), i.e. two grid-wide
// dependencies and two launches here (k_ebe_patch MODE 1 + k_pcg_update).
// This is the same Jacobi-PCG in the Chronopoulos-Gear form (A. T.
// Chronopoulos, C. W. Gear, "s-step iterative methods for symmetric linear
// systems", J. Comput. Appl. Math. 25 (1989)), which carries s = A p and
// w = A u (u = M^-1 r) as recurrences so that both inner products of an
// iteration come from one reduction:
//   p_j = u_j + b_j p_{j-1},  s_j = w_j + b_j s_{j-1},
//   x_{j+1} = x_j + a_j p_j,  r_{j+1} = r_j - a_j s_j,  u_{j+1} = M^-1 r_{j+1},
//   w_{j+1} = A u_{j+1},  g_{j+1} = r_{j+1}'u_{j+1},  d_{j+1} = w_{j+1}'u_{j+1},
//   b_{j+1} = g_{j+1} / g_j,  a_{j+1} = g_{j+1} / (d_{j+1} - b_{j+1} g_{j+1} / a_j).
// In exact arithmetic the iterates equal the reference's.  That step length
// rests on the orthogonality of successive residuals, which FP32 vectors lose
// within a few hundred iterations (the mixed solve needed 1102 iterations at
// the 64^3 cube against the standard recurrence's 870).  The FP32 kernel
// therefore forms the denominator p_{j+1}'A p_{j+1} itself, over terms the same
// reduction delivers: d_{j+1} + 2 b_{j+1} e_{j+1} + b_{j+1}^2 z_{j+1} with
// e_{j+1} = u_{j+1}'s_j and z_{j+1} = p_j's_j (s_j'u_{j+1} = p_j'w_{j+1} by the
// symmetry of A), which assumes only s = A p (938 iterations there).  FP64
// keeps the paper's form, which matches the standard recurrence's count.
//
// One launch = one iteration, one workgroup per patch (patch_common.h):
//  * preamble: every block reduces the previous launch's per-patch partials
//    of g, d and r'r (FP32 also e and z) in the same fixed order, runs the
//    stop test and forms a_j, b_j;
//  * phase A, one patch node per thread: w_j of the node (owned: the vector;
//    interface: the fixed-order sum of its slots), then s_j, r_{j+1},
//    u_{j+1}.  Every patch that touches an interface node computes these
//    (bitwise the same values: same inputs, same order); the node's first
//    patch also forms p_j, x_{j+1} and writes the node's state;
//  * phases B and C: w_{j+1} = A u_{j+1} (owned nodes in place, interface
//    nodes into slots) and the patch's partial of d_{j+1} = sum_e u_e'K_e u_e
//    (+ constrained diagonal terms); g and r'r (FP32: e, z) partials over the
//    nodes the patch writes.
// The node state (x, r, p, s) and the slots ping-pong by the launch's parity,
// so no block reads a value another block of the same launch writes.  The
// parity follows the launches (every graph replay starts at parity 0), not
// the iteration count: after a residual replacement the two differ.
#include <hip/hip_runtime.h>

#include <cmath>

#include "cg1.h"
#include "element.h"
#include "patch_common.h"

namespace femgpu {

namespace {

using namespace elem;
using namespace patch;

template <typename T>
struct Cg1Args {
  const PatchInfo *patches;
  const uint32_t *erec;
  const int4 *sh_info;  // per patch interface node: id, slot, node's first slot, end slot
  const uint16_t *node_end;
  const uint8_t *fixmask;
  const T *coords, *escale, *diagK, *invD;
  T lam, mu;
  int np, p8, ent, pn;
  Cg1State<T> st;
  CgScalars *sc;
  int parity;
};

// INIT: w = A u, u = M^-1 r[0], into w / slot[0], and the partials of g, d,
// r'r (FP32: e = u's[0], z = p[0]'s[0]) into part[0], the state the first
// iteration launch reads (p[0] = s[0] = 0 at a start or restart; the kept
// direction and s[0] = A p[0] after a residual replacement).
template <typename T, int DIM, int PT, bool INIT>
__global__ __launch_bounds__(PT) void k_cg1(Cg1Args<T> a) {
  constexpr unsigned FMASK = (1u << DIM) - 1u;
  extern __shared__ __align__(16) unsigned char lds[];
  const int b = blockIdx.x;
  // XCD-aware: consecutive patches share interface nodes and an L2
  const int pid = (b & 7) * a.p8 + (b >> 3);
  if (pid >= a.np) return;
  const LdsLayout L = lds_layout(DIM, int(sizeof(T)), a.pn, a.ent, RED_CG1);
  T *s_c = reinterpret_cast<T *>(lds + L.off_c);
  T *s_x = reinterpret_cast<T *>(lds + L.off_x);      // u, constrained components zeroed
  T *s_res = reinterpret_cast<T *>(lds + L.off_res);  // one component of the entries
  double *s_red = reinterpret_cast<double *>(lds + L.off_red);
  const int in = INIT ? 0 : a.parity;
  const int out = INIT ? 0 : a.parity ^ 1;
  const Cg1State<T> &S = a.st;
  // FP32: step length from p'A p itself (e and z partials); FP64: the
  // paper's recurrence
  constexpr bool EXACT = sizeof(T) == 4;
  constexpr int NPART = EXACT ? 5 : 3;

  T alpha = T(0), beta = T(0);
  if constexpr (!INIT) {
    CgScalars *sc = a.sc;
    if (sc->done) return;
    const double *pin = S.part + size_t(in) * CG1_PARTS * a.np;
    double v[NPART];
#pragma unroll
    for (int m = 0; m < NPART; ++m) v[m] = 0.0;
    for (int i = threadIdx.x; i < a.np; i += PT) {
#pragma unroll
      for (int m = 0; m < NPART; ++m) v[m] += pin[m * a.np + i];
    }
    block_sum_n<PT, NPART>(v, s_red);
    const double g = v[0], d = v[1], q = v[2];
    const int q1 = in ^ 1;
    const int j = sc->it[q1] + 1;
    int stop = 0;
    if (j >= 1) {
      const bool conv = sc->kind == 0 ? (g < sc->thresh) : (q <= sc->thresh);
      if (conv) stop = 1;
    } else if (q == 0.0) {
      stop = 1;
    }
    if (!stop && !isfinite(g)) stop = 2;
    if (!stop && j >= sc->max_iter) stop = 4;
    if (!stop && q < sc->pause) stop = 5;
    double bt = 0.0, al = 0.0;
    if (!stop) {
      // b_j and a_j; g[q1] = a[q1] = +inf at a (re)start gives b = 0, a = g/d
      bt = g / sc->g[q1];
      double den;
      if constexpr (EXACT)
        den = d + bt * (2.0 * v[3] + bt * v[4]);
      else
        den = d - bt * g / sc->al[q1];
      al = g / den;
      if (!(den > 0.0) || !isfinite(al)) stop = isfinite(den) ? 3 : 2;
    }
    if (pid == 0 && threadIdx.x == 0) {
      sc->g_final = g;
      sc->rr_final = q;
      sc->iter_final = j;
      if (stop) {
        sc->done = stop;
        sc->par_final = in;
      } else {
        sc->g[in] = g;
        sc->al[in] = al;
        sc->b[in] = bt;
        sc->it[in] = j;
        sc->pq = d;
      }
    }
    if (stop) return;
    alpha = T(al);
    beta = T(bt);
  }

  const PatchInfo P = a.patches[pid];
  const int nloc = P.n_own + P.n_sh;
  uint3 rec[EPT];
  T esc[EPT];
  // per-patch bases (uniform) and small unsigned thread offsets: scalar
  // base + one offset register per load
  const uint3 *prec = reinterpret_cast<const uint3 *>(a.erec) + P.elem_start;
  const T *pesc = a.escale ? a.escale + P.elem_start : nullptr;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int i = threadIdx.x + e * PT;
    rec[e] = i < P.n_elem ? prec[unsigned(i)] : make_uint3(0u, 0u, 0u);
    esc[e] = pesc && i < P.n_elem ? pesc[unsigned(i)] : T(1);
  }
  clear_entries<T, PT>(s_res, a.ent);

  // Phase A
  const int k = threadIdx.x;
  const bool has_node = k < nloc;
  const bool own = k < P.n_own;
  bool first = own;
  int64_t gid = 0, slot = 0;
  int b0 = 0, b1 = 0;
  unsigned fm = 0;
  T u[DIM];
  constexpr bool REREAD = sizeof(T) == 4;
#pragma unroll
  for (int c = 0; c < DIM; ++c) u[c] = T(0);
  double acc_g = 0.0, acc_d = 0.0, acc_q = 0.0, acc_e = 0.0, acc_z = 0.0;
  if (has_node) {
    int s0 = 0, s1 = 0;
    if (own) {
      gid = int64_t(P.own_start + k);
    } else {
      const int4 si = a.sh_info[P.sh_start + k - P.n_own];
      gid = si.x;
      slot = si.y;
      s0 = si.z;
      s1 = si.w;
      first = si.y == si.z;
    }
    b1 = int(a.node_end[P.pnode_start + k]);
    b0 = k ? int(a.node_end[P.pnode_start + k - 1]) : 0;
#pragma unroll
    for (int c = 0; c < DIM; ++c) s_c[k * DIM + c] = a.coords[gid * DIM + c];
    fm = a.fixmask[gid] & FMASK;
    const int64_t o = gid * DIM;
    T rn[DIM];
    if constexpr (INIT) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        rn[c] = S.r[0][o + c];
        u[c] = a.invD[o + c] * rn[c];
      }
      if (EXACT && first) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const T sv = S.s[0][o + c];
          acc_e += double(u[c]) * double(sv);
          acc_z += double(S.p[0][o + c]) * double(sv);
        }
      }
    } else {
      T rv[DIM], sv[DIM], dv[DIM], wv[DIM], pv[DIM], xv[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        rv[c] = S.r[in][o + c];
        sv[c] = S.s[in][o + c];
        dv[c] = a.invD[o + c];
        pv[c] = T(0);
        xv[c] = T(0);
      }
      if (own) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) wv[c] = S.w[o + c];
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) wv[c] = T(0);
        add_slots<T, DIM>(S.slot[in], s0, s1, wv);
      }
      if (first) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          pv[c] = S.p[in][o + c];
          xv[c] = S.x[in][o + c];
        }
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const T sn = wv[c] + beta * sv[c];
        rn[c] = rv[c] - alpha * sn;
        u[c] = dv[c] * rn[c];
        const T pn = dv[c] * rv[c] + beta * pv[c];
        // this node's new state, stored by its first patch
        if (first) {
          S.p[out][o + c] = pn;
          S.s[out][o + c] = sn;
          S.x[out][o + c] = xv[c] + alpha * pn;
          S.r[out][o + c] = rn[c];
          if (EXACT) {
            acc_e += double(u[c]) * double(sn);
            acc_z += double(pn) * double(sn);
          }
        }
      }
    }
    if (!REREAD && first) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        acc_g += double(rn[c]) * double(u[c]);
        acc_q += double(rn[c]) * double(rn[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) s_x[k * DIM + c] = ((fm >> c) & 1u) ? T(0) : u[c];
  }
  // wave sums of e and z now: wave-uniform, they stay in scalar registers
  // through the assembly
  if constexpr (EXACT) {
    acc_e = wave_sum(acc_e);
    acc_z = wave_sum(acc_z);
  }
  __syncthreads();

  // Phases B and C: w_{j+1} = A u_{j+1}; d = u'w is summed node by node
  // below from what the patch writes (its rows of w and its slots)
  T acc[DIM];
  assemble<T, DIM, PT, false, false>(s_c, s_x, s_res, rec, esc, P.n_elem, a.lam, a.mu, b0, b1, acc,
                                     acc_d, 1.0);
  if (has_node) {
    // FP32: u and the new residual again (not held in registers through
    // the assembly, which keeps eight workgroups per CU): u from the LDS
    // copy; the residual (needed by the node's first patch, for r'u and r'r,
    // and for u = M^-1 r in constrained components) from where phase A
    // stored it.  FP64 keeps them in registers.
    const T *rnew = INIT ? S.r[0] : S.r[out];
    if (REREAD) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) u[c] = s_x[k * DIM + c];
    }
    if (REREAD && first) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const T rn = rnew[gid * DIM + c];
        if ((fm >> c) & 1u) u[c] = a.invD[gid * DIM + c] * rn;
        acc_g += double(rn) * double(u[c]);
        acc_q += double(rn) * double(rn);
      }
    }
    // constrained rows: the assembled diagonal only (applyConstraints,
    // cpu_fem_utils.cpp:461-476), written by the node's first patch
    if (fm) {
#pragma unroll
      for (int c = 0; c < DIM; ++c)
        if ((fm >> c) & 1u) acc[c] = first ? a.diagK[gid * DIM + c] * u[c] : T(0);
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc_d += double(u[c]) * double(acc[c]);
    if (own) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) S.w[gid * DIM + c] = acc[c];
    } else {
#pragma unroll
      for (int c = 0; c < DIM; ++c) S.slot[out][slot * DIM + c] = acc[c];
    }
  }
  double v[NPART];
  v[0] = wave_sum(acc_g);
  v[1] = wave_sum(acc_d);
  v[2] = wave_sum(acc_q);
  if constexpr (EXACT) {
    v[3] = acc_e;
    v[4] = acc_z;
  }
  waves_sum_n<PT, NPART>(v, s_red);
  if (threadIdx.x == 0) {
    double *po = S.part + size_t(out) * CG1_PARTS * a.np;
#pragma unroll
    for (int m = 0; m < NPART; ++m) po[m * a.np + pid] = v[m];
  }
}

// Scalars after an INIT launch.  mode 0: fresh start (g0, thresholds,
// counters); 1: residual replacement keeping the search direction p, state
// moved from parity q (FP32 state only: its step length is p'A p of the new
// direction p = u + b p_prev, exact for a replaced residual too, from the
// INIT partials e = u's_prev and z = p_prev's_prev); 2: replacement with a
// restart (b = 0 next), also after a breakdown stop (done = 3).
__global__ void k_cg1_fin(CgScalars *sc, const double *part, int np, int mode, int q,
                          double pause_drop) {
  __shared__ double s_red[3 * 256 / 64];
  double g = 0.0, d = 0.0, rr = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    g += part[i];
    d += part[np + i];
    rr += part[2 * np + i];
  }
  block_sum3<256>(g, d, rr, s_red);
  if (threadIdx.x != 0) return;
  const double inf = INFINITY;
  sc->g_final = g;
  sc->rr_final = rr;
  sc->pause = pause_drop > 0.0 ? pause_drop * pause_drop * rr : 0.0;
  if (mode == 0) {
    sc->gamma0 = g;
    sc->rr0 = rr;
    sc->g[0] = sc->g[1] = inf;
    sc->al[0] = sc->al[1] = inf;
    sc->b[0] = sc->b[1] = 0.0;
    sc->it[0] = sc->it[1] = -1;
    sc->iter_final = 0;
    sc->par_final = 0;
    sc->thresh = sc->kind == 0 ? sc->eps * g : sc->eps * sc->eps * rr;
    sc->done = isfinite(g) && isfinite(d) ? 0 : 2;
    return;
  }
  if (q) {
    sc->g[1] = sc->g[0];
    sc->al[1] = sc->al[0];
    sc->b[1] = sc->b[0];
    sc->it[1] = sc->it[0];
  }
  if (mode == 2) sc->g[1] = sc->al[1] = inf;
  // a restart also clears a breakdown stop (the host restarts after one)
  if (sc->done == 1 || sc->done == 5 || (mode == 2 && sc->done == 3)) sc->done = 0;
}

template <typename T>
Cg1Args<T> cg1_args(const DevMesh &dm, const Cg1State<T> &st, const T *invD, CgScalars *sc,
                    int parity) {
  Cg1Args<T> a{};
  a.patches = dm.patches;
  a.erec = dm.erec;
  a.sh_info = dm.sh_info;
  a.node_end = dm.node_end;
  a.fixmask = dm.fixmask;
  if constexpr (sizeof(T) == 8) {
    a.coords = reinterpret_cast<const T *>(dm.coords64);
    a.escale = reinterpret_cast<const T *>(dm.escale64);
    a.diagK = reinterpret_cast<const T *>(dm.diag64);
  } else {
    a.coords = reinterpret_cast<const T *>(dm.coords32);
    a.escale = reinterpret_cast<const T *>(dm.escale32);
    a.diagK = reinterpret_cast<const T *>(dm.diag32);
  }
  a.invD = invD;
  a.lam = T(dm.lam);
  a.mu = T(dm.mu);
  a.np = dm.np;
  a.p8 = (dm.np + 7) / 8;
  a.ent = dm.ent;
  a.pn = dm.pn;
  a.st = st;
  a.sc = sc;
  a.parity = parity;
  return a;
}

template <typename T, int DIM, int PT, bool INIT>
void launch_pt(const DevMesh &dm, const Cg1Args<T> &a, hipStream_t s) {
  const LdsLayout L = lds_layout(DIM, int(sizeof(T)), dm.pn, dm.ent, RED_CG1);
  hipLaunchKernelGGL((k_cg1<T, DIM, PT, INIT>), dim3(8 * a.p8), dim3(PT), L.total, s, a);
}

template <typename T, bool INIT>
void launch(const DevMesh &dm, const Cg1Args<T> &a, hipStream_t s) {
  if (dm.dim == 3) {
    if (dm.patch_threads == 1024)
      launch_pt<T, 3, 1024, INIT>(dm, a, s);
    else if (dm.patch_threads == 512)
      launch_pt<T, 3, 512, INIT>(dm, a, s);
    else
      launch_pt<T, 3, 256, INIT>(dm, a, s);
  } else {
    if (dm.patch_threads == 1024)
      launch_pt<T, 2, 1024, INIT>(dm, a, s);
    else if (dm.patch_threads == 512)
      launch_pt<T, 2, 512, INIT>(dm, a, s);
    else
      launch_pt<T, 2, 256, INIT>(dm, a, s);
  }
}

template <typename T, int DIM, int PT, bool INIT>
hipError_t attr() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_cg1<T, DIM, PT, INIT>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

}  // namespace

hipError_t prepare_cg1_kernels() {
  hipError_t e = hipSuccess, r;
#define FEMGPU_CG1_ATTR(T, D)                                                         \
  if ((r = attr<T, D, 256, false>()) != hipSuccess) e = r;                            \
  if ((r = attr<T, D, 512, false>()) != hipSuccess) e = r;                            \
  if ((r = attr<T, D, 1024, false>()) != hipSuccess) e = r;                           \
  if ((r = attr<T, D, 256, true>()) != hipSuccess) e = r;                             \
  if ((r = attr<T, D, 512, true>()) != hipSuccess) e = r;                             \
  if ((r = attr<T, D, 1024, true>()) != hipSuccess) e = r;
  FEMGPU_CG1_ATTR(double, 3)
  FEMGPU_CG1_ATTR(double, 2)
  FEMGPU_CG1_ATTR(float, 3)
  FEMGPU_CG1_ATTR(float, 2)
#undef FEMGPU_CG1_ATTR
  return e;
}

template <typename T>
void launch_cg1_init(const DevMesh &dm, const Cg1State<T> &st, const T *invD, hipStream_t s) {
  launch<T, true>(dm, cg1_args<T>(dm, st, invD, nullptr, 0), s);
}

template <typename T>
void launch_cg1_iter(const DevMesh &dm, const Cg1State<T> &st, const T *invD, CgScalars *sc,
                     int parity, hipStream_t s) {
  launch<T, false>(dm, cg1_args<T>(dm, st, invD, sc, parity), s);
}

void launch_cg1_fin(CgScalars *sc, const double *part, int np, int mode, int q,
                    double pause_drop, hipStream_t s) {
  hipLaunchKernelGGL(k_cg1_fin, dim3(1), dim3(256), 0, s, sc, part, np, mode, q, pause_drop);
}

template void launch_cg1_init<double>(const DevMesh &, const Cg1State<double> &, const double *,
                                      hipStream_t);
template void launch_cg1_init<float>(const DevMesh &, const Cg1State<float> &, const float *,
                                     hipStream_t);
template void launch_cg1_iter<double>(const DevMesh &, const Cg1State<double> &, const double *,
                                      CgScalars *, int, hipStream_t);
template void launch_cg1_iter<float>(const DevMesh &, const Cg1State<float> &, const float *,
                                     CgScalars *, int, hipStream_t);

}  // namespace femgpu
