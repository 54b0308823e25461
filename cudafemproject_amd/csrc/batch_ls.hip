// batch_ls.hip -- lockstep schedule of the batched static solves (layout in
// batch_ls.h).  Per iteration, two launches over all instances:
//  * k_bl_patch<MODE 1>: per patch, the stop test and beta of its instance's
//    NR right-hand sides from the previous update's per-block partials
//    (reduced by one wave in a fixed order), p = z + beta p, q = K p for
//    every active right-hand side (patch_common.h assembly, one right-hand
//    side at a time through the single-component LDS entry list), per-patch
//    p'q partials;
//  * k_bl_update: per block of an instance's nodes, alpha from the
//    instance's per-patch p'q partials, interface q from the slots,
//    x += alpha p (FP64), r -= alpha q, z = D^-1 r, per-block r'z / r'r.
// A system whose recurrence residual meets the tolerance, or (mixed
// precision: FP32 inner vectors and matvec, x in FP64) has dropped by
// `drop`, pauses and requests a residual replacement, done for all
// requesting systems at the end of each graph replay (FP64 matvec of their
// instances, r = f - K x); a system finishes on its FP64 residual.  All sums
// have a fixed order.
//
// The reference has no batched solve; each system is the static problem of
// gpuPCG_EbE_vec (femfunc.cpp:25-128) with the relative-residual rule.
#include <hip/hip_runtime.h>

#include <cmath>

#include "batch_ls.h"
#include "element.h"
#include "femgpu.h"
#include "patch_common.h"

namespace femgpu {

namespace {

using namespace elem;
using namespace patch;

constexpr int PT = BL_THREADS;

struct BLLds {
  int off_c, off_x, off_res, off_red, off_sc, total;
};

__host__ __device__ inline BLLds bl_lds(int dim, int nv, int tb, int pn, int ent) {
  BLLds L;
  L.off_c = 0;
  L.off_x = al32(pn * dim * tb);
  L.off_res = L.off_x + al32(pn * dim * nv * tb);
  L.off_red = L.off_res + al32(ent * tb);
  L.off_sc = L.off_red + 48 * 8;
  L.total = L.off_sc + 64;
  return L;
}

// Sum of vals[i * NV + v] over i in [i0, i1) for every v, by the 64 lanes of
// one wave (fixed order: strided lanes, then wave_sum, which gives every
// lane the same bits).  Call from one full wave.
template <int NV>
__device__ __forceinline__ void wave_sums(const double *vals, int i0, int i1, double *out) {
  const int lane = threadIdx.x & 63;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int i = i0 + lane; i < i1; i += 64)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += vals[int64_t(i) * NV + v];
#pragma unroll
  for (int v = 0; v < NV; ++v) out[v] = wave_sum(acc[v]);
}

// Block sums of NV values per thread; sh holds NV * PT / 64 doubles.  The
// totals are valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sums(double *v, double *sh) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (l == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[w * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < PT / 64; ++i) s += sh[i * NV + k];
      v[k] = s;
    }
}

// ---------------------------------------------------------------------------
// Patch kernel.  MODE 0: y = K x for the NR vectors (only_flagged: only
// instances with a replacement request).  MODE 1: PCG step.  MODE 2: diag(K)
// (one vector).
template <typename T>
struct BLPatch {
  BLArgs<T> m;
  const T *x;           // MODE 0: x; MODE 1: z
  T *p;                 // MODE 1: search directions (owned nodes written)
  T *y;                 // owned-node output
  T *slots;             // interface output
  const double *prz;    // MODE 1: previous update's partials [nblk][NR]
  const double *prr;
  double *ppq;          // MODE 1: [np][NR]
  int parity;           // MODE 1
  int only_flagged;     // MODE 0
};

template <typename T, int DIM, int NR, int MODE>
__global__ __launch_bounds__(PT) void k_bl_patch(BLPatch<T> a) {
  constexpr unsigned FMASK = (1u << DIM) - 1u;
  constexpr int NV = MODE == 2 ? 1 : NR;
  extern __shared__ __align__(16) unsigned char lds[];
  const int b = blockIdx.x;
  const int pid = (b & 7) * a.m.p8 + (b >> 3);
  if (pid >= a.m.np) return;
  const BLLds L = bl_lds(DIM, NV, int(sizeof(T)), a.m.pn, a.m.ent);
  T *s_c = reinterpret_cast<T *>(lds + L.off_c);
  T *s_x = reinterpret_cast<T *>(lds + L.off_x);
  T *s_res = reinterpret_cast<T *>(lds + L.off_res);
  double *s_red = reinterpret_cast<double *>(lds + L.off_red);
  double *s_beta = reinterpret_cast<double *>(lds + L.off_sc);
  int *s_act = reinterpret_cast<int *>(lds + L.off_sc + 32);

  const PatchInfo P = a.m.patches[pid];
  const int iid = P.pad;
  const BLInst I = a.m.inst[iid];

  if constexpr (MODE == 1) {
    if (threadIdx.x < 64) {
      double g[NR], rr[NR];
      wave_sums<NR>(a.prz, I.u0, I.u1, g);
      wave_sums<NR>(a.prr, I.u0, I.u1, rr);
      const int r = threadIdx.x;
      if (r < NR) {
        BLState &S = a.m.st[int64_t(iid) * NR + r];
        const int q1 = a.parity ^ 1;
        const int j = S.it[q1] + 1;
        const int s = S.state;
        const double rrv = rr[r], gv = g[r];
        double thr = S.thr;
        const bool set_thr = thr < 0.0;
        if (set_thr) thr = a.m.tol2 * rrv;  // first iteration: r = f
        int ns = s, status = S.status;
        bool done = s == 3, pause = false, restart = false;
        if (!done) {
          if (!isfinite(gv) || !isfinite(rrv)) {
            done = true;
            status = FEMGPU_E_NONFINITE;
          } else if (s == 1) {
            pause = true;  // waiting for the replacement at the end of this replay
          } else if (s == 2) {
            // rr is the FP64 residual of the replaced r = f - K x
            if (rrv <= thr) {
              done = true;
              status = FEMGPU_OK;
            } else {
              ns = 0;
              restart = S.restart != 0;  // decided by k_bl_mark
            }
          } else if (rrv <= thr || rrv < a.m.drop2 * S.rr_last) {
            // the recurrence's residual met the tolerance (to be confirmed on
            // the FP64 residual) or dropped by `drop` (mixed): replacement
            // requested; the system pauses until then
            ns = 1;
            pause = true;
          }
          if (!done && !pause && j >= a.m.max_iter) {
            done = true;
            status = FEMGPU_E_MAXIT;
          }
        }
        // Every patch of the instance reaches the same decision from the same
        // partials and state: the fields the first patch rewrites in this
        // launch (state 0 -> 1 or 3, thr) lead to the same outcome whether
        // read before or after, and the others (rr_last, restart) change
        // only in k_bl_mark.  A replaced system (state 2) stays in state 2
        // through this launch, since its restart decision is taken only in
        // that state; the update launch that follows moves it to state 0.
        const bool act = !done && !pause;
        const double bt = act && !restart ? gv / S.g[q1] : 0.0;
        if (pid == I.p0) {
          if (set_thr) {
            S.thr = thr;
            S.ff = rrv;
          }
          if (s != 1) S.rr = rrv;  // paused: keep the recurrence's r'r at the pause
          if (done) {
            if (s != 3) {
              S.state = 3;
              S.status = status;
              S.iters = j;
            }
          } else if (pause) {
            if (s == 0) {
              // the launch after the replacement has parity 0 and reads
              // slot 1: carry the last recurrence scalars there
              S.state = 1;
              S.g[1] = S.g[q1];
              S.b[1] = S.b[q1];
              S.it[1] = S.it[q1];
            }
          } else {
            if (s != 2) S.state = ns;
            S.g[a.parity] = gv;
            S.b[a.parity] = bt;
            S.it[a.parity] = j;
          }
        }
        s_beta[r] = bt;
        s_act[r] = act ? 1 : 0;
      }
    }
    __syncthreads();
    int any = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) any |= s_act[r];
    if (!any) {
      if (threadIdx.x < NR) a.ppq[int64_t(pid) * NR + threadIdx.x] = 0.0;
      return;
    }
  } else if constexpr (MODE == 0) {
    if (a.only_flagged) {
      int any = 0;
#pragma unroll
      for (int r = 0; r < NR; ++r) any |= a.m.st[int64_t(iid) * NR + r].state == 1;
      if (!any) return;
    }
  }

  const int nloc = P.n_own + P.n_sh;
  uint3 rec[EPT];
  T esc[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int i = threadIdx.x + e * PT;
    rec[e] = i < P.n_elem ? reinterpret_cast<const uint3 *>(a.m.erec)[P.elem_start + i]
                          : make_uint3(0u, 0u, 0u);
    esc[e] = T(1);
  }
  clear_entries<T, PT>(s_res, a.m.ent);

  // Phase A
  const int k = threadIdx.x;
  const bool has_node = k < nloc;
  const bool own = k < P.n_own;
  bool first = own;
  int64_t gid = 0, slot = 0;
  int b0 = 0, b1 = 0;
  unsigned fm = 0;
  T xin[NV][DIM];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int c = 0; c < DIM; ++c) xin[v][c] = T(0);
  double acc_en[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc_en[v] = 0.0;
  if (has_node) {
    if (own) {
      gid = int64_t(P.own_start + k);
    } else {
      const int4 si = a.m.sh_info[P.sh_start + k - P.n_own];
      gid = si.x;
      slot = si.y;
      first = si.y == si.z;
    }
    b1 = int(a.m.node_end[P.pnode_start + k]);
    b0 = k ? int(a.m.node_end[P.pnode_start + k - 1]) : 0;
#pragma unroll
    for (int c = 0; c < DIM; ++c) s_c[k * DIM + c] = a.m.coords[gid * DIM + c];
    if constexpr (MODE != 2) {
      fm = a.m.fixmask[gid] & FMASK;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int64_t o = (gid * NV + v) * DIM;
        if constexpr (MODE == 0) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) xin[v][c] = a.x[o + c];
        } else {
          if (s_act[v]) {
            const T bt = T(s_beta[v]);
#pragma unroll
            for (int c = 0; c < DIM; ++c) xin[v][c] = a.x[o + c] + bt * a.p[o + c];
            if (own)
#pragma unroll
              for (int c = 0; c < DIM; ++c) a.p[o + c] = xin[v][c];
          }
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          s_x[(v * a.m.pn + k) * DIM + c] = ((fm >> c) & 1u) ? T(0) : xin[v][c];
        if constexpr (MODE == 1) {
          if (first)
#pragma unroll
            for (int c = 0; c < DIM; ++c)
              if ((fm >> c) & 1u) {
                const double pv = double(xin[v][c]);
                acc_en[v] += double(a.m.diagK[gid * DIM + c]) * pv * pv;
              }
        }
      }
    }
  }
  __syncthreads();

  // Phases B and C, one vector at a time
  const T lam = T(I.lam), mu = T(I.mu);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if constexpr (MODE == 1) {
      if (!s_act[v]) continue;
    }
    T acc[DIM];
    assemble<T, DIM, PT, MODE == 2>(s_c, s_x + v * a.m.pn * DIM, s_res, rec, esc, P.n_elem, lam,
                                    mu, b0, b1, acc, acc_en[v], 1.0);
    if (has_node) {
      if constexpr (MODE != 2) {
        if (fm)
#pragma unroll
          for (int c = 0; c < DIM; ++c)
            if ((fm >> c) & 1u) acc[c] = first ? a.m.diagK[gid * DIM + c] * xin[v][c] : T(0);
      }
      if (own) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) a.y[(gid * NV + v) * DIM + c] = acc[c];
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) a.slots[(slot * NV + v) * DIM + c] = acc[c];
      }
    }
    __syncthreads();  // the entry list is reused by the next vector
  }

  if constexpr (MODE == 1) {
    block_sums<NR>(acc_en, s_red);
    if (threadIdx.x == 0)
#pragma unroll
      for (int r = 0; r < NR; ++r) a.ppq[int64_t(pid) * NR + r] = s_act[r] ? acc_en[r] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Node-parallel kernels: one thread per (node, right-hand side) of a block of
// BL_UPD_NODES nodes, so that a wave reads contiguous [node][rhs][dim] data.

// Per-rhs block totals of a (thread rhs = lane % NR): xor butterfly over the
// lanes of one rhs, then the W waves of the block in order.  sh: 2 * NR * W
// doubles.  Totals valid in threads < NR (thread v: rhs v).
template <int NR, int W>
__device__ __forceinline__ void rhs_sums(double &a, double &b, double *sh) {
#pragma unroll
  for (int o = 32; o >= NR; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l < NR) {
    sh[(w * NR + l) * 2] = a;
    sh[(w * NR + l) * 2 + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < W; ++i) {
      sa += sh[(i * NR + threadIdx.x) * 2];
      sb += sh[(i * NR + threadIdx.x) * 2 + 1];
    }
    a = sa;
    b = sb;
  }
}

// Node update of one PCG iteration.
template <typename T>
struct BLUpdate {
  BLArgs<T> m;
  const T *q, *slots, *invD;
  T *p, *r, *z;
  double *x;           // solution (FP64)
  T *xinc;             // non-null: the increments go here (flushed into x by k_bl_flush)
  const double *ppq;
  double *prz, *prr;   // this iteration's partials [nblk][NR]
  int parity;
};

template <typename T, int DIM, int NR>
__global__ __launch_bounds__(BL_UPD_NODES * NR) void k_bl_update(BLUpdate<T> a) {
  constexpr int W = BL_UPD_NODES * NR / 64;
  __shared__ double s_red[2 * NR * W];
  __shared__ double s_al[NR], s_bt[NR];
  __shared__ int s_act[NR];
  const BLBlock B = a.m.blocks[blockIdx.x];
  const BLInst I = a.m.inst[B.inst];
  if (threadIdx.x < 64) {
    double pq[NR];
    wave_sums<NR>(a.ppq, I.p0, I.p1, pq);
    const int r = threadIdx.x;
    if (r < NR) {
      BLState &S = a.m.st[int64_t(B.inst) * NR + r];
      const int st0 = S.state;
      int act = st0 == 0 || st0 == 2;  // 1: paused for a replacement, 3: finished
      // a replaced system iterates again from here on (state 2 -> 0, set by
      // the instance's first block; the others read 2 or 0, both active)
      if (st0 == 2 && blockIdx.x == I.u0) S.state = 0;
      double al = 0.0;
      if (act) {
        al = S.g[a.parity] / pq[r];
        if (!(pq[r] > 0.0) || !isfinite(al)) {
          if (blockIdx.x == I.u0) {
            S.state = 3;
            S.status = isfinite(pq[r]) ? FEMGPU_E_BREAKDOWN : FEMGPU_E_NONFINITE;
            S.iters = S.it[a.parity];
          }
          act = 0;
          al = 0.0;
        }
      }
      s_al[r] = al;
      s_bt[r] = S.b[a.parity];
      s_act[r] = act;
    }
  }
  __syncthreads();
  int any = 0;
#pragma unroll
  for (int k = 0; k < NR; ++k) any |= s_act[k];
  if (!any) return;  // the partials of inactive systems are never read
  const int v = threadIdx.x % NR;
  const int64_t g = B.n0 + threadIdx.x / NR;
  double rz = 0.0, rr = 0.0;
  if (g < B.n1 && s_act[v]) {
    const int64_t loc = g - I.node_off;
    const int64_t o = (g * NR + v) * DIM;
    T qv[DIM], pv[DIM];
    if (loc < I.n_owned) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        qv[c] = a.q[o + c];
        pv[c] = a.p[o + c];
      }
    } else {
      const int4 fi = a.m.iface[I.if0 + (loc - I.n_owned)];
#pragma unroll
      for (int c = 0; c < DIM; ++c) qv[c] = T(0);
      for (int j = fi.y; j < fi.z; ++j)
#pragma unroll
        for (int c = 0; c < DIM; ++c) qv[c] += a.slots[(int64_t(j) * NR + v) * DIM + c];
      const T bt = T(s_bt[v]);
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        pv[c] = a.z[o + c] + bt * a.p[o + c];
        a.p[o + c] = pv[c];
      }
    }
    const T al = T(s_al[v]);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if (a.xinc)
        a.xinc[o + c] += al * pv[c];
      else
        a.x[o + c] += s_al[v] * double(pv[c]);
      const T rv = a.r[o + c] - al * qv[c];
      a.r[o + c] = rv;
      const T zv = a.invD[g * DIM + c] * rv;
      a.z[o + c] = zv;
      rz += double(rv) * double(zv);
      rr += double(rv) * double(rv);
    }
  }
  rhs_sums<NR, W>(rz, rr, s_red);
  if (threadIdx.x < NR) {
    a.prz[int64_t(blockIdx.x) * NR + threadIdx.x] = rz;
    a.prr[int64_t(blockIdx.x) * NR + threadIdx.x] = rr;
  }
}

// x += xinc, xinc = 0 for the nodes of a block (the increments of the
// replay's iterations, accumulated in the inner precision); blocks whose
// instance has finished every system are skipped (their xinc is zero: the
// flush before the replacement that finished them emptied it).
template <typename T, int DIM, int NR>
__global__ __launch_bounds__(BL_UPD_NODES * NR) void k_bl_flush(BLArgs<T> m, T *xinc, double *x) {
  const BLBlock B = m.blocks[blockIdx.x];
  bool live = false;
#pragma unroll
  for (int r = 0; r < NR; ++r) live |= m.st[int64_t(B.inst) * NR + r].state != 3;
  if (!live) return;
  const int v = threadIdx.x % NR;
  const int64_t g = B.n0 + threadIdx.x / NR;
  if (g >= B.n1) return;
  const int64_t o = (g * NR + v) * DIM;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    x[o + c] += double(xinc[o + c]);
    xinc[o + c] = T(0);
  }
}

// ---------------------------------------------------------------------------
// Residual replacement of the systems with state 1 (all systems at the
// start): r64 = f - y64 with y64 = K x (interface nodes from the FP64
// slots), r = r64, z = D^-1 r, and the partials the next iteration (parity
// 0) reads: r'z and the FP64 r'r.  FINAL: every system, only the FP64 r'r.
template <typename T>
struct BLReplace {
  BLArgs<T> m;
  const double *f, *y64, *slots64;
  const T *invD;
  T *r, *z;
  double *prz, *prr;
};

template <typename T, int DIM, int NR, bool FINAL>
__global__ __launch_bounds__(BL_UPD_NODES * NR) void k_bl_replace(BLReplace<T> a) {
  constexpr int W = BL_UPD_NODES * NR / 64;
  __shared__ double s_red[2 * NR * W];
  const BLBlock B = a.m.blocks[blockIdx.x];
  const BLInst I = a.m.inst[B.inst];
  int any = 0;
#pragma unroll
  for (int k = 0; k < NR; ++k) any |= FINAL || a.m.st[int64_t(B.inst) * NR + k].state == 1;
  if (!any) return;
  const int v = threadIdx.x % NR;
  const bool flag = FINAL || a.m.st[int64_t(B.inst) * NR + v].state == 1;
  const int64_t g = B.n0 + threadIdx.x / NR;
  double rz = 0.0, rr = 0.0;
  if (g < B.n1 && flag) {
    const int64_t loc = g - I.node_off;
    const int64_t o = (g * NR + v) * DIM;
    double kx[DIM];
    if (loc < I.n_owned) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) kx[c] = a.y64[o + c];
    } else {
      const int4 fi = a.m.iface[I.if0 + (loc - I.n_owned)];
#pragma unroll
      for (int c = 0; c < DIM; ++c) kx[c] = 0.0;
      for (int j = fi.y; j < fi.z; ++j)
#pragma unroll
        for (int c = 0; c < DIM; ++c) kx[c] += a.slots64[(int64_t(j) * NR + v) * DIM + c];
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const double r64 = a.f[o + c] - kx[c];
      rr += r64 * r64;
      if constexpr (!FINAL) {
        const T rv = T(r64);
        a.r[o + c] = rv;
        const T zv = a.invD[g * DIM + c] * rv;
        a.z[o + c] = zv;
        rz += double(rv) * double(zv);
      }
    }
  }
  rhs_sums<NR, W>(rz, rr, s_red);
  if (threadIdx.x < NR && (FINAL || a.m.st[int64_t(B.inst) * NR + threadIdx.x].state == 1)) {
    a.prz[int64_t(blockIdx.x) * NR + threadIdx.x] = rz;
    a.prr[int64_t(blockIdx.x) * NR + threadIdx.x] = rr;
  }
}

// After a replacement, between replays (no other kernel runs): for every
// system with state 1 its new FP64 r'r (sum of the replacement partials
// over its instance's blocks, fixed order), the restart decision (the FP64
// residual did not halve since the previous replacement: p = z), state
// 1 -> 2.  count: unfinished systems.
__global__ __launch_bounds__(1024) void k_bl_mark(int n_inst, int nr, const BLInst *inst,
                                                  const double *prr, BLState *st, int *count) {
  __shared__ double s_red[1024 / 64];
  double c = 0.0;
  for (int i = threadIdx.x; i < n_inst * nr; i += 1024) {
    BLState &S = st[i];
    if (S.state == 1) {
      const BLInst &I = inst[i / nr];
      const int r = i % nr;
      double rr = 0.0;
      for (int b = I.u0; b < I.u1; ++b) rr += prr[int64_t(b) * nr + r];
      // restart the directions when the FP64 residual did not at least halve
      // since the previous replacement, or when the recurrence had drifted
      // far from it (FP64 residual more than twice the recurrence's)
      S.restart = rr > 0.25 * S.rr_last || rr > 4.0 * S.rr ? 1 : 0;
      // two replacements in a row without any decrease of the FP64 residual
      // (the tolerance is below the attainable accuracy): stop the system
      // here, where its residual is still near the best reached, instead of
      // iterating on rounding noise
      if (S.nrep >= 1 && rr >= S.rr_last) {
        ++S.stag;
      } else {
        S.stag = 0;
      }
      S.rr_last = rr;
      ++S.nrep;
      if (S.stag >= 2 && !(rr <= S.thr)) {
        S.state = 3;
        S.status = FEMGPU_E_STAGNATION;
        S.iters = S.it[1];
      } else {
        S.state = 2;
      }
    }
    c += S.state != 3 ? 1.0 : 0.0;
  }
  c = block_sum<1024>(c, s_red);
  if (threadIdx.x == 0) *count = int(c);
}

// FP64 r'r of every system from the FINAL replacement partials.
template <int NR>
__global__ __launch_bounds__(64) void k_bl_final_sum(int n_inst, const BLInst *inst,
                                                     const double *prr, BLState *st) {
  const int i = blockIdx.x;
  if (i >= n_inst) return;
  double rr[NR];
  wave_sums<NR>(prr, inst[i].u0, inst[i].u1, rr);
  if (threadIdx.x < NR) st[int64_t(i) * NR + threadIdx.x].rr = rr[threadIdx.x];
}

__global__ void k_bl_init_state(int n, BLState *st) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  BLState S{};
  S.g[0] = S.g[1] = INFINITY;
  S.b[0] = S.b[1] = 0.0;
  S.thr = -1.0;
  S.it[0] = S.it[1] = -1;
  S.state = 1;
  S.status = FEMGPU_OK;
  S.nrep = -1;  // the initial r = f does not count as a replacement
  st[i] = S;
}

// Interface nodes of a single vector: y[node] = sum of its slots.
__global__ void k_bl_gather(int n_if, const int4 *iface, const double *slots, double *y,
                            int dim) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_if) return;
  const int4 fi = iface[i];
  for (int c = 0; c < dim; ++c) {
    double s = 0.0;
    for (int j = fi.y; j < fi.z; ++j) s += slots[int64_t(j) * dim + c];
    y[int64_t(fi.x) * dim + c] = s;
  }
}

template <typename T>
__global__ void k_bl_inv(int64_t n, const double *d, T *inv) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) inv[i] = d[i] != 0.0 ? T(1.0 / d[i]) : T(0);
}

template <typename T, int DIM, int NR, int MODE>
void launch_patch(const BLPatch<T> &a, hipStream_t s) {
  constexpr int NV = MODE == 2 ? 1 : NR;
  const BLLds L = bl_lds(DIM, NV, int(sizeof(T)), a.m.pn, a.m.ent);
  hipLaunchKernelGGL((k_bl_patch<T, DIM, NR, MODE>), dim3(8 * a.m.p8), dim3(PT), L.total, s, a);
}

template <typename T, int MODE>
void patch_dispatch(int dim, int nr, const BLPatch<T> &a, hipStream_t s) {
  if (dim == 2) {
    if (nr == 1) launch_patch<T, 2, 1, MODE>(a, s);
    else if (nr == 2) launch_patch<T, 2, 2, MODE>(a, s);
    else launch_patch<T, 2, 4, MODE>(a, s);
  } else {
    if (nr == 1) launch_patch<T, 3, 1, MODE>(a, s);
    else if (nr == 2) launch_patch<T, 3, 2, MODE>(a, s);
    else launch_patch<T, 3, 4, MODE>(a, s);
  }
}

template <typename T, int DIM, int MODE>
hipError_t attr3() {
  hipError_t e = hipSuccess, r;
  const void *f1 = reinterpret_cast<const void *>(&k_bl_patch<T, DIM, 1, MODE>);
  const void *f2 = reinterpret_cast<const void *>(&k_bl_patch<T, DIM, 2, MODE>);
  const void *f4 = reinterpret_cast<const void *>(&k_bl_patch<T, DIM, 4, MODE>);
  for (const void *f : {f1, f2, f4})
    if ((r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) !=
        hipSuccess)
      e = r;
  return e;
}

}  // namespace

hipError_t prepare_bl_kernels() {
  hipError_t e = hipSuccess, r;
#define FEMGPU_BL_ATTR(T, D)                                         \
  if ((r = attr3<T, D, 0>()) != hipSuccess) e = r;                   \
  if ((r = attr3<T, D, 1>()) != hipSuccess) e = r;                   \
  if ((r = attr3<T, D, 2>()) != hipSuccess) e = r;
  FEMGPU_BL_ATTR(double, 2)
  FEMGPU_BL_ATTR(double, 3)
  FEMGPU_BL_ATTR(float, 2)
  FEMGPU_BL_ATTR(float, 3)
#undef FEMGPU_BL_ATTR
  return e;
}

void bl_diag(int dim, const BLArgs<double> &m, int n_if, double *diag, double *slots64,
             hipStream_t s) {
  BLPatch<double> a{};
  a.m = m;
  a.y = diag;
  a.slots = slots64;
  patch_dispatch<double, 2>(dim, 1, a, s);
  if (n_if > 0)
    hipLaunchKernelGGL(k_bl_gather, dim3((n_if + 255) / 256), dim3(256), 0, s, n_if, m.iface,
                       slots64, diag, dim);
}

template <typename T>
void bl_inv_diag(int64_t n, const double *diag, T *invD, hipStream_t s) {
  hipLaunchKernelGGL((k_bl_inv<T>), dim3(int((n + 255) / 256)), dim3(256), 0, s, n, diag, invD);
}

void bl_init_state(int n_state, BLState *st, hipStream_t s) {
  hipLaunchKernelGGL(k_bl_init_state, dim3((n_state + 255) / 256), dim3(256), 0, s, n_state, st);
}

template <typename T>
void bl_iteration(int dim, int nr, const BLArgs<T> &m, const T *invD, T *z, T *p, T *q, T *r,
                  T *slots, double *x64, T *xinc, double *ppq, double *pnode, int parity,
                  hipStream_t s) {
  BLPatch<T> a{};
  a.m = m;
  a.x = z;
  a.p = p;
  a.y = q;
  a.slots = slots;
  a.prz = pnode + size_t(2 * (parity ^ 1) + 0) * m.nblk * nr;
  a.prr = pnode + size_t(2 * (parity ^ 1) + 1) * m.nblk * nr;
  a.ppq = ppq;
  a.parity = parity;
  patch_dispatch<T, 1>(dim, nr, a, s);
  BLUpdate<T> u{};
  u.m = m;
  u.q = q;
  u.slots = slots;
  u.invD = invD;
  u.p = p;
  u.r = r;
  u.z = z;
  u.x = x64;
  u.xinc = xinc;
  u.ppq = ppq;
  u.prz = pnode + size_t(2 * parity + 0) * m.nblk * nr;
  u.prr = pnode + size_t(2 * parity + 1) * m.nblk * nr;
  u.parity = parity;
#define FEMGPU_BL_UPD(D, R) \
  hipLaunchKernelGGL((k_bl_update<T, D, R>), dim3(m.nblk), dim3(BL_UPD_NODES * R), 0, s, u)
  if (dim == 2) {
    if (nr == 1) FEMGPU_BL_UPD(2, 1);
    else if (nr == 2) FEMGPU_BL_UPD(2, 2);
    else FEMGPU_BL_UPD(2, 4);
  } else {
    if (nr == 1) FEMGPU_BL_UPD(3, 1);
    else if (nr == 2) FEMGPU_BL_UPD(3, 2);
    else FEMGPU_BL_UPD(3, 4);
  }
#undef FEMGPU_BL_UPD
}

template <typename T>
void bl_flush(int dim, int nr, const BLArgs<T> &m, T *xinc, double *x64, hipStream_t s) {
#define FEMGPU_BL_FL(D, R) \
  hipLaunchKernelGGL((k_bl_flush<T, D, R>), dim3(m.nblk), dim3(BL_UPD_NODES * R), 0, s, m, xinc, x64)
  if (dim == 2) {
    if (nr == 1) FEMGPU_BL_FL(2, 1);
    else if (nr == 2) FEMGPU_BL_FL(2, 2);
    else FEMGPU_BL_FL(2, 4);
  } else {
    if (nr == 1) FEMGPU_BL_FL(3, 1);
    else if (nr == 2) FEMGPU_BL_FL(3, 2);
    else FEMGPU_BL_FL(3, 4);
  }
#undef FEMGPU_BL_FL
}

template <typename T, bool FINAL>
static void replace_dispatch(int dim, int nr, const BLReplace<T> &a, hipStream_t s) {
#define FEMGPU_BL_REP(D, R) \
  hipLaunchKernelGGL((k_bl_replace<T, D, R, FINAL>), dim3(a.m.nblk), dim3(BL_UPD_NODES * R), 0, s, a)
  if (dim == 2) {
    if (nr == 1) FEMGPU_BL_REP(2, 1);
    else if (nr == 2) FEMGPU_BL_REP(2, 2);
    else FEMGPU_BL_REP(2, 4);
  } else {
    if (nr == 1) FEMGPU_BL_REP(3, 1);
    else if (nr == 2) FEMGPU_BL_REP(3, 2);
    else FEMGPU_BL_REP(3, 4);
  }
#undef FEMGPU_BL_REP
}

template <typename T>
void bl_replace(int dim, int nr, const BLArgs<T> &m, const BLArgs<double> &m64, const double *f,
                const double *x64, double *y64, double *slots64, const T *invD, T *r, T *z,
                double *pnode, int n_state, int *count, hipStream_t s) {
  BLPatch<double> mv{};
  mv.m = m64;
  mv.x = x64;
  mv.y = y64;
  mv.slots = slots64;
  mv.only_flagged = 1;
  patch_dispatch<double, 0>(dim, nr, mv, s);
  BLReplace<T> a{};
  a.m = m;
  a.f = f;
  a.y64 = y64;
  a.slots64 = slots64;
  a.invD = invD;
  a.r = r;
  a.z = z;
  a.prz = pnode + size_t(2 * 1 + 0) * m.nblk * nr;  // read by the next parity-0 iteration
  a.prr = pnode + size_t(2 * 1 + 1) * m.nblk * nr;
  replace_dispatch<T, false>(dim, nr, a, s);
  hipLaunchKernelGGL(k_bl_mark, dim3(1), dim3(1024), 0, s, n_state / nr, nr, m.inst, a.prr, m.st,
                     count);
}

void bl_final(int dim, int nr, const BLArgs<double> &m64, const double *f, const double *x64,
              double *y64, double *slots64, double *pnode, int n_inst, hipStream_t s) {
  BLPatch<double> mv{};
  mv.m = m64;
  mv.x = x64;
  mv.y = y64;
  mv.slots = slots64;
  mv.only_flagged = 0;
  patch_dispatch<double, 0>(dim, nr, mv, s);
  BLReplace<double> a{};
  a.m = m64;
  a.f = f;
  a.y64 = y64;
  a.slots64 = slots64;
  a.prz = pnode;
  a.prr = pnode + size_t(m64.nblk) * nr;
  replace_dispatch<double, true>(dim, nr, a, s);
  if (nr == 1)
    hipLaunchKernelGGL((k_bl_final_sum<1>), dim3(n_inst), dim3(64), 0, s, n_inst, m64.inst, a.prr, m64.st);
  else if (nr == 2)
    hipLaunchKernelGGL((k_bl_final_sum<2>), dim3(n_inst), dim3(64), 0, s, n_inst, m64.inst, a.prr, m64.st);
  else
    hipLaunchKernelGGL((k_bl_final_sum<4>), dim3(n_inst), dim3(64), 0, s, n_inst, m64.inst, a.prr, m64.st);
}

template void bl_inv_diag<float>(int64_t, const double *, float *, hipStream_t);
template void bl_inv_diag<double>(int64_t, const double *, double *, hipStream_t);
template void bl_iteration<float>(int, int, const BLArgs<float> &, const float *, float *, float *,
                                  float *, float *, float *, double *, float *, double *, double *,
                                  int, hipStream_t);
template void bl_iteration<double>(int, int, const BLArgs<double> &, const double *, double *,
                                   double *, double *, double *, double *, double *, double *,
                                   double *, double *, int, hipStream_t);
template void bl_flush<float>(int, int, const BLArgs<float> &, float *, double *, hipStream_t);
template void bl_replace<float>(int, int, const BLArgs<float> &, const BLArgs<double> &,
                                const double *, const double *, double *, double *, const float *,
                                float *, float *, double *, int, int *, hipStream_t);
template void bl_replace<double>(int, int, const BLArgs<double> &, const BLArgs<double> &,
                                 const double *, const double *, double *, double *,
                                 const double *, double *, double *, double *, int, int *,
                                 hipStream_t);

}  // namespace femgpu
