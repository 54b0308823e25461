// kernels.hip -- hand-written HIP kernels (gfx950) of the EbE elasticity path.
//
// What replaces what in the reference (moarseniy/CudaFEMproject):
//  * k_ebe_patch replaces, per PCG iteration, Matrix::bmm with the stored
//    per-element K_e (cuBLAS SgemmStridedBatched, cuda_matrix.cpp:252-275) plus
//    the sort-based assembly reductionWithMask{,AndTransform}
//    (cuda_fem_utils.cpp:196-216, thrust sort_by_key/reduce_by_key,
//    matrix_kernels.cu:37-51).  K_e = V B'DB is never stored: each element's
//    force K_e u_e is formed on the fly from the node coordinates, one element
//    per thread, with the patch's node data staged in LDS, and assembled in
//    LDS through fixed-order incidence lists.  Interface nodes are finished by
//    a fixed-order sum over per-patch partial slots.
//  * k_pcg_update fuses the PCG vector updates and both dot products of an
//    iteration (femfunc.cpp:59-111: addWeighted x3, divideElementwise,
//    dotProduct x2) into one pass; the dot products reduce with wave64
//    shuffles into fixed-order per-block partials that the next kernel
//    reduces, and the scalars never leave the device.
//  * The element-force math is the reference's K_e = B'DB * area * coeff
//    (cpu_fem_utils.cpp:338-503) rewritten with the barycentric gradients
//    g_a = c_a / det J (c_a: cofactor vectors), so that
//    f_a = (1 / (2|det J|) or 1 / (6|det J|)) * S(H) c_a, H = sum_a du_a (x) c_a,
//    S(H) = lam tr(H) I + mu (H + H').
//  * Constrained DOFs follow applyConstraints (cpu_fem_utils.cpp:461-476):
//    off-diagonal couplings of a fixed DOF are dropped and its diagonal kept.
// All reductions have a fixed order, so results are bitwise reproducible
// run to run for a given mesh plan.
#include <hip/hip_runtime.h>

#include "element.h"
#include "kernels.h"
#include "patch_common.h"

namespace femgpu {

namespace {

constexpr int NT = 256;          // threads per block for every kernel
constexpr int NODE_GRID_MAX = 2048;

using namespace elem;
using namespace patch;

// ---------------------------------------------------------------------------
// Patch kernel.  MODE 0: y = K x.  MODE 1: fused PCG step (stop test and beta
// from the previous update's partial sums, p = z + beta p, q = cK K p + cM m p,
// per-patch p'q partial).  MODE 2: diag(K).  MODE 3: explicit GN22 step, the
// predictor d~ = d + dt v + c1 a fused with y = K d~.  MODE 4: the whole
// explicit step in one launch: MODE 3, then the corrector of the owned nodes
// from their complete K d~; an interface node is corrected one launch late,
// by every patch that holds it, from the previous launch's slots (the first
// patch stores the result).

// FP64 PCG: x += alpha p in the next iteration's patch kernel (k_ebe_patch
// MODE 1); FP32: in the update kernel.
template <typename T>
constexpr bool PCG_DEFER_X = sizeof(T) == 8;

template <typename T>
struct PatchArgs {
  const PatchInfo *patches;
  const uint32_t *erec_s;      // 3 words per element, patch p's at p * emax (pack_records)
  int emax;                    // record stride per patch
  const int32_t *pgid;     // per patch node at stride PT: node id
  const int32_t *pslot;    // per patch node at stride PT: slot | first-slot bit
  const uint16_t *pend;    // per patch node at stride PT: end of its entry run
  const uint8_t *fixmask;
  const int32_t *slot_start;
  const T *coords;
  const T *escale;
  const T *diagK;     // assembled diagonal (rows of constrained DOFs), MODE 0/1
  T lam, mu;
  int np, p8, ent, pn;
  int64_t n_owned;
  const T *x;   // MODE 0: x; MODE 1: r (z = invD r is formed here)
  const T *invD;      // MODE 1: Jacobi preconditioner
  T *xs;              // MODE 1: the PCG iterate (x += alpha p of the previous update)
  T *p;         // MODE 1: search direction (old in, new out for owned nodes)
  T *y;         // owned-node output
  T *partial;   // interface slots
  OpShift<T> sh;
  CgScalars *sc;
  double *part;       // MODE 1: per-patch p'q partials
  const double *prz;  // MODE 1: r'z partials of the previous update (NODE_PARTS)
  const double *prr;  // MODE 1: r'r partials of the previous update
  int parity;         // MODE 1: j & 1
  const T *vel, *acc; // MODE 3/4: velocity and acceleration of the previous step
  T *xout;            // MODE 3/4: predicted displacement (owned and first-slot nodes)
  T dt, c1;           // MODE 3/4: GN22 predictor coefficients
  // MODE 4 (for interface nodes vel/acc and vout/aout lag one step)
  T *vout, *aout;           // corrected velocity and acceleration
  const T *prev_part;       // the previous launch's interface slots
  const T *f_file, *f_unit; // load vectors and their per-step coefficients
  const double *coef_file, *coef_unit;
  const int *step_base;     // step = *step_base + step_off (1-based)
  int step_off;
  const T *invM;            // lumped inverse mass
  T c2, c3, alpha, scale;   // GN22 corrector and mass-damping coefficients
};

// GN22 corrector of one DOF (the operations of k_dyn_correct): r = F - K d~,
// v~ = v + c2 a, (1 + alpha c3) M a' = r - alpha M v~, v' = v~ + c3 a'.
template <typename T>
__device__ __forceinline__ void gn22_correct(T cf, T cu, T ff, T fu, T kd, T im, T c2, T c3,
                                             T alpha, T scale, T &a, T &v) {
  const T r = cf * ff + cu * fu - kd;
  const T vp = v + c2 * a;
  const T av = alpha != T(0) ? (r * im - alpha * vp) * scale : r * im;
  a = av;
  v = vp + c3 * av;
}

// One workgroup per patch, one patch node per thread (phases A and C), up to
// EPT elements per thread (phase B).  The element forces stay in registers
// and are assembled one component at a time through a single-component LDS
// entry list, which keeps the LDS per element at npe words.
// Waves per SIMD asked of the register allocator.  The FP64 matvec and PCG
// step of 256-thread patches fit 128 VGPRs with a few spilled words: four
// resident workgroups per CU instead of three (FP64 matvec 29.1 -> 28.3 us,
// PCG iteration -0.3 us at cube64).  The FP32 PCG step at six waves spills
// and is slower; everything else keeps the allocator's choice.
template <typename T, int MODE, int PT>
struct PatchWaves {
  static constexpr int value = PT == 256 && sizeof(T) == 8 && (MODE == 0 || MODE == 1) ? 4 : 1;
};

template <typename T, int DIM, int MODE, int PT>
__global__ __launch_bounds__(PT) __attribute__((amdgpu_waves_per_eu(PatchWaves<T, MODE, PT>::value)))
void k_ebe_patch(PatchArgs<T> a) {
  constexpr unsigned FMASK = (1u << DIM) - 1u;
  extern __shared__ __align__(16) unsigned char lds[];
  const int b = blockIdx.x;
  // XCD-aware: consecutive patches (neighbours on the Morton curve, sharing
  // interface nodes) go to the same XCD and its L2.
  const int pid = (b & 7) * a.p8 + (b >> 3);
  if (pid >= a.np) return;
  const LdsLayout L = lds_layout(DIM, int(sizeof(T)), a.pn, a.ent);
  T *s_c = reinterpret_cast<T *>(lds + L.off_c);
  T *s_x = reinterpret_cast<T *>(lds + L.off_x);      // constrained components zeroed
  T *s_res = reinterpret_cast<T *>(lds + L.off_res);  // one component of the entries
  double *s_red = reinterpret_cast<double *>(lds + L.off_red);

  T beta = T(0), xstep = T(0);
  if constexpr (MODE == 1) {
    CgScalars *sc = a.sc;
    if (sc->done) return;
    // Reduce the previous update's r'z and r'r (fixed order, every block
    // gets the same values), run the stop test, form beta.
    double g = 0.0, q = 0.0;
    if constexpr (sizeof(T) == 8) {
      // FP64 (four workgroups per CU, two rounds of patches): every wave
      // sums all partials itself, in the same order in every wave, with no
      // block barrier (measured 0.8 us faster per launch at the 64^3 cube;
      // no gain in FP32, where eight workgroups per CU hide the barrier)
      const int lane = threadIdx.x & 63;
      constexpr int NL = NODE_PARTS / 64;
#pragma unroll
      for (int c0 = 0; c0 < NL; c0 += 4) {
        double gv[4], qv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          gv[j] = a.prz[lane + 64 * (c0 + j)];
          qv[j] = a.prr[lane + 64 * (c0 + j)];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          g += gv[j];
          q += qv[j];
        }
      }
      g = wave_sum(g);
      q = wave_sum(q);
    } else {
      for (int i = threadIdx.x; i < NODE_PARTS; i += PT) {
        g += a.prz[i];
        q += a.prr[i];
      }
      block_sum2<PT>(g, q, s_red);
    }
    const int q1 = a.parity ^ 1;
    const int n_upd = sc->it[q1] + 1;
    int stop = 0;
    if (n_upd >= 1) {
      const bool conv = sc->kind == 0 ? (g < sc->thresh) : (q <= sc->thresh);
      if (conv) stop = 1;
    } else if (q == 0.0) {
      stop = 1;
    }
    if (!stop && !isfinite(g)) stop = 2;
    if (!stop && n_upd >= sc->max_iter) stop = 4;
    const double bt = g / sc->g[q1];
    if (pid == 0 && threadIdx.x == 0) {
      sc->g_final = g;
      sc->rr_final = q;
      sc->iter_final = n_upd;
      if (stop) {
        sc->done = stop;
        sc->par_final = a.parity;
      } else {
        sc->g[a.parity] = g;
        sc->b[a.parity] = bt;
        sc->it[a.parity] = n_upd;
      }
    }
    if (stop) return;
    beta = T(bt);
    xstep = T(sc->x_pend);
  }
  // this thread's node id and run end from the patch node tables, at a
  // fixed stride: they do not wait for the PatchInfo load, so both are in
  // flight together and the node data below is one dependent load away
  // (threads past the plan's largest patch have no node in any patch)
  const size_t pbase = size_t(pid) * PT;
  int gid_t = 0, e1 = 0, e0 = 0;
  if (threadIdx.x < unsigned(a.pn)) {
    gid_t = a.pgid[pbase + threadIdx.x];
    e1 = int(a.pend[pbase + threadIdx.x]);
    e0 = threadIdx.x ? int(a.pend[pbase + threadIdx.x - 1]) : 0;
  }
  const PatchInfo P = a.patches[pid];
  const int nloc = P.n_own + P.n_sh;
  double acc_en = 0.0;
  const T cK = MODE == 1 ? a.sh.cK : T(1);
  const T cM = MODE == 1 ? a.sh.cM : T(0);

  // Element records of this thread.  They lie at a fixed stride per patch,
  // so they are in flight before the PatchInfo arrives; indices are clamped
  // to the stride in the matvec and the PCG step (records past the patch's
  // elements are never decoded), so the loads are unconditional.  These two
  // kernels load the first
  // element's record here and the others after the node loads: every
  // workgroup stages its nodes at the same time, bandwidth-bound, and the
  // later records arrive while the first element is computed.
  constexpr bool LATE_REC = MODE == 0 || MODE == 1;
  uint3 rec[EPT];
  T esc[EPT];
  // per-patch bases (uniform) and small unsigned thread offsets: scalar
  // base + one offset register per load
  const uint3 *prec = reinterpret_cast<const uint3 *>(a.erec_s) + size_t(pid) * unsigned(a.emax);
  const T *pesc = a.escale ? a.escale + P.elem_start : nullptr;
#pragma unroll
  for (int e = 0; e < (LATE_REC ? 1 : EPT); ++e) {
    const int i = threadIdx.x + e * PT;
    if constexpr (LATE_REC)
      rec[e] = prec[unsigned(min(i, a.emax - 1))];
    else
      rec[e] = i < a.emax ? prec[unsigned(i)] : make_uint3(0u, 0u, 0u);
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int i = threadIdx.x + e * PT;
    esc[e] = pesc && i < P.n_elem ? pesc[unsigned(i)] : T(1);
  }

  clear_entries<T, PT>(s_res, a.ent);

  // Phase A: this thread's patch node.  Constrained components enter the
  // element products as zero (their couplings are dropped); their rows are
  // set from the assembled diagonal below.
  const int k = threadIdx.x;
  const bool has_node = k < nloc;
  const bool own = k < P.n_own;
  int gid = 0;
  // element offsets of the node and its slot, 32-bit (few live registers);
  // the node's signed, so that its components' accesses merge into one
  // vector access (the slot's stay apart: one register fewer at the peak)
  int go = 0;
  unsigned so = 0;
  int b0 = 0, b1 = 0;
  unsigned fm = 0;
  T xin[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) xin[c] = T(0);
  // interface nodes: their slot, and whether this patch holds the node's
  // first slot (owned nodes: their only patch); MODE 4: the node's
  // acceleration and velocity entering this step
  const int slot_t = has_node && !own ? a.pslot[pbase + k] : 0;
  const bool first = own || slot_t < 0;
  T av4[DIM], vv4[DIM];
  int step4 = 0;
  if constexpr (MODE == 4) step4 = *a.step_base + a.step_off;
  if (has_node) {
    gid = gid_t;
    go = gid * DIM;
    if (!own) so = unsigned(slot_t & 0x7fffffff) * unsigned(DIM);
    b1 = e1;
    b0 = e0;
#pragma unroll
    for (int c = 0; c < DIM; ++c) s_c[k * DIM + c] = a.coords[go + c];
    if constexpr (MODE != 2) fm = a.fixmask[gid] & FMASK;
    if constexpr (MODE == 0) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) xin[c] = a.x[go + c];
    } else if constexpr (MODE == 1) {
      T pv[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) pv[c] = a.p[go + c];
#pragma unroll
      for (int c = 0; c < DIM; ++c) xin[c] = pcg_dir(a.invD[go + c], a.x[go + c], beta, pv[c]);
      if (own) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) a.p[go + c] = xin[c];
      }
      // FP64: the previous update's x += alpha p, by the node's first patch
      // (the update kernel then neither reads nor writes x).  Two rounds of
      // FP64 workgroups per CU hide these accesses; FP32 workgroups all
      // stage their nodes at once, bandwidth-bound, and keep x in the update.
      if (PCG_DEFER_X<T> && first) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) a.xs[go + c] += xstep * pv[c];
      }
    } else if constexpr (MODE == 3) {
      // the order of k_dyn_predict: (d + dt v) + c1 a
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const T dv = a.x[go + c] + a.dt * a.vel[go + c];
        xin[c] = dv + a.c1 * a.acc[go + c];
      }
    } else if constexpr (MODE == 4) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        av4[c] = a.acc[go + c];
        vv4[c] = a.vel[go + c];
      }
      if (!own) {
        const int64_t s = int64_t(gid) - a.n_owned;
        const int s0 = a.slot_start[s], s1 = a.slot_start[s + 1];
        if (step4 > 1) {
          // the previous step of this interface node, from its slots
          T kd[DIM];
#pragma unroll
          for (int c = 0; c < DIM; ++c) kd[c] = T(0);
          add_slots<T, DIM>(a.prev_part, s0, s1, kd);
          const T cf = T(a.coef_file[step4 - 1]), cu = T(a.coef_unit[step4 - 1]);
#pragma unroll
          for (int c = 0; c < DIM; ++c)
            gn22_correct<T>(cf, cu, a.f_file[go + c], a.f_unit[go + c], kd[c],
                            a.invM[go + c], a.c2, a.c3, a.alpha, a.scale, av4[c], vv4[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const T dv = a.x[go + c] + a.dt * vv4[c];
        xin[c] = dv + a.c1 * av4[c];
      }
    }
    if constexpr (MODE != 2) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) s_x[k * DIM + c] = ((fm >> c) & 1u) ? T(0) : xin[c];
    }
  }
  // the fixed components and the first-slot flag travel through the
  // assembly in one register (an interface node's first patch writes its
  // constrained rows and its predicted state and, in the PCG step, adds its
  // mass term of p'(cK K + cM M) p)
  unsigned fmf = fm | (first ? 0x10u : 0u);
  asm volatile("" : "+v"(fmf));  // opaque: not re-derived from its sources later
  if constexpr (LATE_REC) {
#pragma unroll
    for (int e = 1; e < EPT; ++e)
      rec[e] = prec[unsigned(min(int(threadIdx.x) + e * PT, a.emax - 1))];
    // (the staged data is in LDS; the records stay in flight)
    lds_barrier();
  } else {
    __syncthreads();
  }

  // Phases B and C: element forces, node-sorted LDS assembly, node sums.
  T acc[DIM];
  // The PCG step forms p'q node by node below (sum over the patch's nodes of
  // p times what the patch writes for the node: its q row or its slot), not
  // element by element: sum_e p_e'K_e p_e regrouped, with fewer operations.
  assemble<T, DIM, PT, MODE == 2, false>(s_c, s_x, s_res, rec, esc, P.n_elem, a.lam, a.mu, b0,
                                         b1, acc, acc_en, double(cK));
  {
  const unsigned fm = fmf & FMASK;
  const bool first = (fmf & 0x10u) != 0u;

  if constexpr (MODE == 0 || MODE == 1) {
    if (has_node) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) xin[c] = s_x[k * DIM + c];
      if (fm) {
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          if ((fm >> c) & 1u) {
            if constexpr (MODE == 0)
              xin[c] = a.x[go + c];
            else if (own)
              xin[c] = a.p[go + c];
            else
              xin[c] = pcg_dir(a.invD[go + c], a.x[go + c], beta, a.p[go + c]);
          }
      }
    }
    if constexpr (MODE == 1) {
      if (has_node && !own && first && a.sh.m) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const double pv = double(xin[c]);
          acc_en += double(cM) * double(a.sh.m[go + c]) * pv * pv;
        }
      }
    }
  }

  if (has_node) {
    // rows of constrained components: the assembled diagonal only
    // (applyConstraints, cpu_fem_utils.cpp:461-476); for an interface node
    // its first patch writes the row and the others contribute zero
    if constexpr (MODE != 2) {
      if (fm) {
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          if ((fm >> c) & 1u) acc[c] = first ? a.diagK[go + c] * xin[c] : T(0);
      }
    }
    if constexpr (MODE == 4) {
      if (own) {
        const T cf = T(a.coef_file[step4]), cu = T(a.coef_unit[step4]);
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          gn22_correct<T>(cf, cu, a.f_file[go + c], a.f_unit[go + c], acc[c],
                          a.invM[go + c], a.c2, a.c3, a.alpha, a.scale, av4[c], vv4[c]);
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) a.partial[so + c] = acc[c];
      }
      if (first) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          a.aout[go + c] = av4[c];
          a.vout[go + c] = vv4[c];
          a.xout[go + c] = xin[c];
        }
      }
    } else if (own) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        T v = acc[c];
        if constexpr (MODE == 1) {
          v = cK * v;
          if (a.sh.m) v += cM * a.sh.m[go + c] * xin[c];
          acc_en += double(xin[c]) * double(v);
        }
        a.y[go + c] = v;
      }
    } else {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const T v = MODE == 1 ? cK * acc[c] : acc[c];
        if constexpr (MODE == 1) acc_en += double(xin[c]) * double(v);
        a.partial[so + c] = v;
      }
    }
    if constexpr (MODE == 3) {
      if (first)
#pragma unroll
        for (int c = 0; c < DIM; ++c) a.xout[go + c] = xin[c];
    }
  }
  }

  if constexpr (MODE == 1) {
    const double tot = block_sum<PT>(acc_en, s_red);
    if (threadIdx.x == 0) a.part[pid] = tot;
  }
}

// Interface gather: y[iface] = sum of slots (fixed order).
template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_iface_gather(int64_t n_owned, int64_t n_iface,
                                                     const int32_t *slot_start, const T *partial,
                                                     T *y) {
  const int64_t s = int64_t(blockIdx.x) * NT + threadIdx.x;
  if (s >= n_iface) return;
  T acc[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) acc[c] = T(0);
  add_slots<T, DIM>(partial, slot_start[s], slot_start[s + 1], acc);
#pragma unroll
  for (int c = 0; c < DIM; ++c) y[(n_owned + s) * DIM + c] = acc[c];
}

// ---------------------------------------------------------------------------
// PCG vector kernels.

template <typename T>
struct UpdArgs {
  int64_t n_owned, n_active;
  const int32_t *slot_start;
  const T *partial;
  const T *q;
  T *p, *r, *x;
  const T *invD;
  OpShift<T> sh;
  CgScalars *sc;
  const double *ppq;  // per-patch p'q partials of this iteration
  int np;
  double *prz, *prr;  // this iteration's r'z, r'r partials (NODE_PARTS each)
  int parity;
};
constexpr int UPT = 512;  // threads per block of the node-parallel PCG kernels

// One node of the PCG update, loaded before alpha is known.
template <typename T, int DIM>
struct UpdNode {
  T q[DIM], p[DIM], r[DIM], d[DIM], x[DIM];
};

template <typename T, int DIM>
__device__ __forceinline__ void upd_load(const UpdArgs<T> &a, int64_t i, UpdNode<T, DIM> &n) {
  const int64_t o = i * DIM;
  if (i < a.n_owned) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) n.q[c] = a.q[o + c];
  } else {
    const int64_t s = i - a.n_owned;
#pragma unroll
    for (int c = 0; c < DIM; ++c) n.q[c] = T(0);
    add_slots<T, DIM>(a.partial, a.slot_start[s], a.slot_start[s + 1], n.q);
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    n.p[c] = a.p[o + c];
    if constexpr (!PCG_DEFER_X<T>) n.x[c] = a.x[o + c];
    n.r[c] = a.r[o + c];
    n.d[c] = a.invD[o + c];
  }
}

template <typename T, int DIM>
__device__ __forceinline__ void upd_finish(const UpdArgs<T> &a, int64_t i, UpdNode<T, DIM> &n,
                                           T alpha, T beta, double &rz, double &rr) {
  const int64_t o = i * DIM;
  if (i >= a.n_owned) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      n.p[c] = pcg_dir(n.d[c], n.r[c], beta, n.p[c]);
      a.p[o + c] = n.p[c];
      if (a.sh.m) n.q[c] += a.sh.cM * a.sh.m[o + c] * n.p[c];
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    if constexpr (!PCG_DEFER_X<T>) a.x[o + c] = n.x[c] + alpha * n.p[c];
    const T rv = n.r[c] - alpha * n.q[c];
    a.r[o + c] = rv;
    const T zv = mul_rn(n.d[c], rv);
    rz += double(rv) * double(zv);
    rr += double(rv) * double(rv);
  }
}

// Fused PCG update (femfunc.cpp:66-81): alpha from the per-patch p'q
// partials (every block reduces them in the same fixed order), finish the
// interface nodes of q (and p = invD r + beta p), x += alpha p (FP64: left
// to the next iteration's patch kernel, which loads p anyway, with alpha in
// sc->x_pend), r -= alpha q, and per-block r'z, r'r partials (z = invD r)
// for the next iteration's stop test; z itself is formed again where it is
// needed.  A
// thread's first two nodes are loaded before the partials are reduced:
// those loads do not depend on alpha, so their latency overlaps the
// reduction.
template <typename T, int DIM>
__global__ __launch_bounds__(UPT) void k_pcg_update(UpdArgs<T> a) {
  __shared__ double s_red[2 * UPT / 64];
  CgScalars *sc = a.sc;
  const int64_t stride = int64_t(gridDim.x) * UPT;
  const int64_t i0 = int64_t(blockIdx.x) * UPT + threadIdx.x;
  // The per-patch p'q partials are loaded first (up to PQ_REG per thread
  // into registers), then this thread's first two nodes: the partials return
  // first and their reduction does not wait for the node loads.
  constexpr int PQ_REG = 8;
  double pqv[PQ_REG];
#pragma unroll
  for (int j = 0; j < PQ_REG; ++j) {
    const int i = threadIdx.x + j * UPT;
    pqv[j] = i < a.np ? a.ppq[i] : 0.0;
  }
  // A thread's first two nodes are loaded before the reduction: with one
  // node per thread short of covering the mesh, the second node's load
  // chain (an interface node's slot offsets, then its slots) would
  // otherwise start only after the first node is finished.
  const int64_t i1 = i0 + stride;
  UpdNode<T, DIM> n0, n1;
  if (i0 < a.n_active) upd_load<T, DIM>(a, i0, n0);
  if (i1 < a.n_active) upd_load<T, DIM>(a, i1, n1);
  double pq = 0.0;
#pragma unroll
  for (int j = 0; j < PQ_REG; ++j) pq += pqv[j];
  for (int i = threadIdx.x + PQ_REG * UPT; i < a.np; i += UPT) pq += a.ppq[i];
  if (sc->done) return;
  pq = block_sum<UPT>(pq, s_red);
  const double gam = sc->g[a.parity];
  const double alpha_d = gam / pq;
  if (!(pq > 0.0) || !isfinite(alpha_d)) {
    // no block writes x or r: the state is the last completed iteration's,
    // as after a stop in the patch kernel's preamble at this parity
    // (this launch's patch kernel already applied the previous x step)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pq = pq;
      sc->x_pend = 0.0;
      sc->done = isfinite(pq) ? 3 : 2;
      sc->par_final = a.parity;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->pq = pq;
    if constexpr (PCG_DEFER_X<T>) sc->x_pend = alpha_d;
  }
  const T alpha = T(alpha_d), beta = T(sc->b[a.parity]);
  double rz = 0.0, rr = 0.0;
  if (i0 < a.n_active) upd_finish<T, DIM>(a, i0, n0, alpha, beta, rz, rr);
  if (i1 < a.n_active) upd_finish<T, DIM>(a, i1, n1, alpha, beta, rz, rr);
  for (int64_t i = i1 + stride; i < a.n_active; i += stride) {
    UpdNode<T, DIM> n;
    upd_load<T, DIM>(a, i, n);
    upd_finish<T, DIM>(a, i, n, alpha, beta, rz, rr);
  }
  block_sum2<UPT>(rz, rr, s_red);
  if (threadIdx.x == 0) {
    a.prz[blockIdx.x] = rz;
    a.prr[blockIdx.x] = rr;
  }
}
// The x += alpha p still pending after the last update of a stopped or
// paused standard PCG (sc->x_pend; 0: none), then none pending.
template <typename T>
__global__ __launch_bounds__(NT) void k_pcg_apply_x(int64_t n, T *x, const T *p,
                                                    const CgScalars *sc) {
  const T al = T(sc->x_pend);
  if (al == T(0)) return;
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) x[i] += al * p[i];
}
__global__ void k_clear_x_pend(CgScalars *sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc->x_pend = 0.0;
}

// PCG initialisation: the r'z, r'r partials (z = invD r) that the first
// fused iteration (parity 0) reads from slot 1; with x0 != nullptr also
// x0 = p0 = 0 over n0 entries.
template <typename T>
__global__ __launch_bounds__(UPT) void k_pcg_init(int64_t n, const T *r, const T *invD,
                                                  double *prz, double *prr, T *x0, T *p0,
                                                  int64_t n0) {
  __shared__ double s_red[UPT / 64];
  double rz = 0.0, rr = 0.0;
  const int64_t stride = int64_t(gridDim.x) * UPT;
  if (x0) {
    for (int64_t i = int64_t(blockIdx.x) * UPT + threadIdx.x; i < n0; i += stride) {
      x0[i] = T(0);
      p0[i] = T(0);
    }
  }
  for (int64_t i = int64_t(blockIdx.x) * UPT + threadIdx.x; i < n; i += stride) {
    const T rv = r[i];
    const T zv = mul_rn(invD[i], rv);
    rz += double(rv) * double(zv);
    rr += double(rv) * double(rv);
  }
  rz = block_sum<UPT>(rz, s_red);
  rr = block_sum<UPT>(rr, s_red);
  if (threadIdx.x == 0) {
    prz[blockIdx.x] = rz;
    prr[blockIdx.x] = rr;
  }
}

__device__ void reset_scalars(CgScalars *sc, int kind, double eps, int max_iter) {
  sc->gamma0 = sc->rr0 = sc->thresh = 0.0;
  sc->g[0] = sc->g[1] = 1.0;
  sc->b[0] = sc->b[1] = 0.0;
  sc->al[0] = sc->al[1] = 1.0;
  sc->g_final = sc->rr_final = sc->pq = sc->rr = sc->scale = sc->pause = sc->x_pend = 0.0;
  sc->it[0] = sc->it[1] = -1;
  sc->iter_final = sc->par_final = 0;
  sc->eps = eps;
  sc->kind = kind;
  sc->done = 0;
  sc->max_iter = max_iter;
}

// Scalars of the initial residual (one block): gamma0, thresholds, and the
// slot-1 values the first iteration reads (beta_0 is irrelevant: p = 0);
// reset != 0 first resets every scalar as k_reset_scalars does.
__global__ __launch_bounds__(NT) void k_fin_init(CgScalars *sc, const double *pa, const double *pb,
                                                 int n, int reset, int kind, double eps,
                                                 int max_iter) {
  __shared__ double s_red[NT / 64];
  double g = 0.0, q = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    g += pa[i];
    q += pb[i];
  }
  g = block_sum(g, s_red);
  q = block_sum(q, s_red);
  if (threadIdx.x == 0) {
    if (reset) reset_scalars(sc, kind, eps, max_iter);
    sc->gamma0 = g;
    sc->rr0 = q;
    sc->g_final = g;
    sc->rr_final = q;
    sc->g[0] = sc->g[1] = 1.0;
    sc->b[0] = sc->b[1] = 0.0;
    sc->it[0] = sc->it[1] = -1;
    sc->iter_final = 0;
    sc->thresh = sc->kind == 0 ? sc->eps * g : sc->eps * sc->eps * q;
    sc->done = isfinite(g) ? 0 : 2;
  }
}

// Finalize of the residual norm (one block): rr and its square root.
__global__ __launch_bounds__(NT) void k_fin_norm(CgScalars *sc, const double *part, int n) {
  __shared__ double s_red[NT / 64];
  double q = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) q += part[i];
  q = block_sum(q, s_red);
  if (threadIdx.x == 0) {
    sc->rr = q;
    sc->scale = sqrt(q);
  }
}

__global__ void k_reset_scalars(CgScalars *sc, int kind, double eps, int max_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) reset_scalars(sc, kind, eps, max_iter);
}

// r = f - K x for the FP64 refinement loop; rr -> sc->rr, sqrt -> sc->scale.
template <int DIM>
__global__ __launch_bounds__(NT) void k_residual64(int64_t n_owned, int64_t n_active,
                                                   const int32_t *slot_start, const double *f,
                                                   const double *y, const double *partial,
                                                   double *r, CgScalars *sc, double *part) {
  __shared__ double s_red[NT / 64];
  double rr = 0.0;
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n_active; i += stride) {
    double kv[DIM];
    if (i < n_owned) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) kv[c] = y[i * DIM + c];
    } else {
      const int64_t s = i - n_owned;
#pragma unroll
      for (int c = 0; c < DIM; ++c) kv[c] = 0.0;
      for (int j = slot_start[s]; j < slot_start[s + 1]; ++j)
#pragma unroll
        for (int c = 0; c < DIM; ++c) kv[c] += partial[int64_t(j) * DIM + c];
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const double rv = f[i * DIM + c] - kv[c];
      r[i * DIM + c] = rv;
      rr += rv * rv;
    }
  }
  rr = block_sum(rr, s_red);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// Reliable-update step of the mixed solve: x64 += x32, x32 = 0.
__global__ __launch_bounds__(NT) void k_flush_x(int64_t n, double *x64, float *x32) {
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) {
    x64[i] += double(x32[i]);
    x32[i] = 0.0f;
  }
}

// Residual replacement: r32 = r64 (the FP64 residual f - K x64) and the
// r'z, r'r partials (z = invD r32) the next fused iteration (parity 0) reads
// from slot 1.
__global__ __launch_bounds__(UPT) void k_pcg_replace(int64_t n, const double *r64, float *r32,
                                                     const float *invD, double *prz,
                                                     double *prr) {
  __shared__ double s_red[UPT / 64];
  double rz = 0.0, rr = 0.0;
  const int64_t stride = int64_t(gridDim.x) * UPT;
  for (int64_t i = int64_t(blockIdx.x) * UPT + threadIdx.x; i < n; i += stride) {
    const float rv = float(r64[i]);
    const float zv = mul_rn(invD[i], rv);
    r32[i] = rv;
    rz += double(rv) * double(zv);
    rr += double(rv) * double(rv);
  }
  rz = block_sum<UPT>(rz, s_red);
  rr = block_sum<UPT>(rr, s_red);
  if (threadIdx.x == 0) {
    prz[blockIdx.x] = rz;
    prr[blockIdx.x] = rr;
  }
}

// Resume a stopped PCG at the next even iteration: the iteration that set
// the stop flag had parity iter_final & 1 and left the last completed
// iteration's gamma and count in the other slot; the next replay reads slot 1.
// Resume the fused PCG after a stop (the residual replaced, its partials in
// slot 1): the next replay starts at parity 0 and reads the scalars of slot
// 1, which must be the last iteration's -- in slot 0 when the stopping
// launch had parity 1 (recorded at the stop: after a resume at an odd
// iteration the launch parity no longer follows the iteration count).
__global__ void k_pcg_resume(CgScalars *sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (sc->par_final == 1) {
      sc->g[1] = sc->g[0];
      sc->b[1] = sc->b[0];
      sc->it[1] = sc->it[0];
    }
    sc->done = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_inv_diag(int64_t n, const T *diag, T cK, T cM, const T *m,
                                                 T *invD) {
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) {
    const T v = cK * diag[i] + (m ? cM * m[i] : T(0));
    invD[i] = v != T(0) ? T(1) / v : T(0);
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_fill(T *p, int64_t n, T v) {
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(NT) void k_convert(const TA *a, TB *b, int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) b[i] = TB(a[i]);
}

// ---------------------------------------------------------------------------
// Per-element outputs (internal element order).

template <typename T, int DIM>
__device__ __forceinline__ void load_elem(const int32_t *elems, const T *coords, const T *u,
                                          int64_t e, T *X, T *U) {
  constexpr int NPE = DIM + 1;
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    const int64_t g = elems[e * NPE + n];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      X[n * DIM + c] = coords[g * DIM + c];
      U[n * DIM + c] = u[g * DIM + c];
    }
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_stress(int64_t ne, const int32_t *elems, const T *coords,
                                               const T *escale, T lam, T mu, const T *u, T *sigma,
                                               bool strain) {
  constexpr int NDE = DIM * (DIM + 1), NV = 3 * (DIM - 1);
  const int64_t e = int64_t(blockIdx.x) * NT + threadIdx.x;
  if (e >= ne) return;
  T X[NDE], U[NDE];
  load_elem<T, DIM>(elems, coords, u, e, X, U);
  const T es = escale ? escale[e] : T(1);
  // sigma = S(H)/det with H = sum_a du_a (x) c_a (signed det: true gradients).
  if constexpr (DIM == 3) {
    const T e1x = X[3] - X[0], e1y = X[4] - X[1], e1z = X[5] - X[2];
    const T e2x = X[6] - X[0], e2y = X[7] - X[1], e2z = X[8] - X[2];
    const T e3x = X[9] - X[0], e3y = X[10] - X[1], e3z = X[11] - X[2];
    T c[4][3];
    c[1][0] = e2y * e3z - e2z * e3y;
    c[1][1] = e2z * e3x - e2x * e3z;
    c[1][2] = e2x * e3y - e2y * e3x;
    c[2][0] = e3y * e1z - e3z * e1y;
    c[2][1] = e3z * e1x - e3x * e1z;
    c[2][2] = e3x * e1y - e3y * e1x;
    c[3][0] = e1y * e2z - e1z * e2y;
    c[3][1] = e1z * e2x - e1x * e2z;
    c[3][2] = e1x * e2y - e1y * e2x;
    const T det = e1x * c[1][0] + e1y * c[1][1] + e1z * c[1][2];
    T H[3][3] = {};
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const T du = U[3 * a + i] - U[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) H[i][j] += du * c[a][j];
      }
    T *o = sigma + e * NV;
    if (strain) {  // engineering strain B u_e: xx, yy, zz, 2 xy, 2 yz, 2 xz
      const T k = T(1) / det;
      o[0] = k * H[0][0];
      o[1] = k * H[1][1];
      o[2] = k * H[2][2];
      o[3] = k * (H[0][1] + H[1][0]);
      o[4] = k * (H[1][2] + H[2][1]);
      o[5] = k * (H[0][2] + H[2][0]);
      return;
    }
    const T k = es / det;
    const T tr = H[0][0] + H[1][1] + H[2][2];
    o[0] = k * (lam * tr + T(2) * mu * H[0][0]);
    o[1] = k * (lam * tr + T(2) * mu * H[1][1]);
    o[2] = k * (lam * tr + T(2) * mu * H[2][2]);
    o[3] = k * mu * (H[0][1] + H[1][0]);
    o[4] = k * mu * (H[1][2] + H[2][1]);
    o[5] = k * mu * (H[0][2] + H[2][0]);
  } else {
    const T e1x = X[2] - X[0], e1y = X[3] - X[1];
    const T e2x = X[4] - X[0], e2y = X[5] - X[1];
    T c[3][2];
    c[1][0] = e2y;
    c[1][1] = -e2x;
    c[2][0] = -e1y;
    c[2][1] = e1x;
    const T det = e1x * e2y - e1y * e2x;
    T H00 = T(0), H01 = T(0), H10 = T(0), H11 = T(0);
#pragma unroll
    for (int a = 1; a < 3; ++a) {
      const T dx = U[2 * a] - U[0], dy = U[2 * a + 1] - U[1];
      H00 += dx * c[a][0];
      H01 += dx * c[a][1];
      H10 += dy * c[a][0];
      H11 += dy * c[a][1];
    }
    T *o = sigma + e * NV;
    if (strain) {  // engineering strain B u_e: xx, yy, 2 xy
      const T k = T(1) / det;
      o[0] = k * H00;
      o[1] = k * H11;
      o[2] = k * (H01 + H10);
      return;
    }
    const T k = es / det;
    const T tr = H00 + H11;
    o[0] = k * (lam * tr + T(2) * mu * H00);
    o[1] = k * (lam * tr + T(2) * mu * H11);
    o[2] = k * mu * (H01 + H10);
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_energy(int64_t ne, const int32_t *elems, const T *coords,
                                               const uint8_t *fixmask, const T *escale, T lam,
                                               T mu, const T *u, T *en) {
  constexpr int NPE = DIM + 1, NDE = DIM * NPE;
  const int64_t e = int64_t(blockIdx.x) * NT + threadIdx.x;
  if (e >= ne) return;
  T X[NDE], U[NDE], F[NDE];
  load_elem<T, DIM>(elems, coords, u, e, X, U);
  unsigned fix = 0;
#pragma unroll
  for (int n = 0; n < NPE; ++n)
    fix |= unsigned(fixmask[elems[e * NPE + n]] & ((1 << DIM) - 1)) << (n * DIM);
  const T es = escale ? escale[e] : T(1);
  en[e] = elem_force<T, DIM>(X, U, lam, mu, es, fix, F);
}

// out_e = factor * w_e' K~_e^0 u_e: the element stiffness at unit modulus
// scale with the constrained-DOF treatment of K (off-diagonal couplings of a
// constrained DOF dropped, its diagonal kept).  K = sum_e s_e K~_e^0, so with
// factor -1 and w the adjoint solution this is dL/ds_e.
template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_bilinear(int64_t ne, const int32_t *elems, const T *coords,
                                                 const uint8_t *fixmask, T lam, T mu, const T *u,
                                                 const T *w, T factor, T *out) {
  constexpr int NPE = DIM + 1, NDE = DIM * NPE;
  const int64_t e = int64_t(blockIdx.x) * NT + threadIdx.x;
  if (e >= ne) return;
  T X[NDE], U[NDE], W[NDE], F[NDE];
  load_elem<T, DIM>(elems, coords, u, e, X, U);
  unsigned fix = 0;
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    const int64_t g = elems[e * NPE + n];
    fix |= unsigned(fixmask[g] & ((1 << DIM) - 1)) << (n * DIM);
#pragma unroll
    for (int c = 0; c < DIM; ++c) W[n * DIM + c] = w[g * DIM + c];
  }
  (void)elem_force<T, DIM>(X, U, lam, mu, T(1), fix, F);
  T acc = T(0);
#pragma unroll
  for (int i = 0; i < NDE; ++i) acc += W[i] * F[i];
  out[e] = factor * acc;
}

// ---------------------------------------------------------------------------
// Dynamics (GN22, femfunc.cpp:247-343) on assembled vectors.

template <typename T>
__global__ __launch_bounds__(NT) void k_dyn_predict(int64_t n, T *d, T *v, T *a, T dt, T c1, T c2,
                                                    int zero_a, int *step) {
  if (step && blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) {
    const T av = a[i];
    d[i] = d[i] + dt * v[i];
    d[i] = d[i] + c1 * av;
    v[i] = v[i] + c2 * av;
    if (zero_a) a[i] = T(0);
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(NT) void k_dyn_correct(int64_t n_owned, int64_t n_active,
                                                    const int32_t *slot_start, const T *y,
                                                    const T *partial, const T *f_file,
                                                    const T *f_unit, const double *coef_file,
                                                    const double *coef_unit, const int *step_base,
                                                    int step_off, const T *invM, T *a, T *v,
                                                    T *r_out, T c3, const T *mlump, T alpha,
                                                    T scale, int pred_v, T c2) {
  const int step = *step_base + step_off;
  const T cf = T(coef_file[step]), cu = T(coef_unit[step]);
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n_active; i += stride) {
    T kd[DIM];
    if (i < n_owned) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) kd[c] = y[i * DIM + c];
    } else {
      const int64_t s = i - n_owned;
#pragma unroll
      for (int c = 0; c < DIM; ++c) kd[c] = T(0);
      add_slots<T, DIM>(partial, slot_start[s], slot_start[s + 1], kd);
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const int64_t o = i * DIM + c;
      const T r = cf * f_file[o] + cu * f_unit[o] - kd[c];
      if (invM) {
        // explicit; with mass-proportional damping alpha M:
        // (1 + alpha c3) M a = F - K d - alpha M v_pred.  pred_v: v holds the
        // previous step's velocity and a its acceleration, v_pred = v + c2 a
        // is formed here (the predictor of k_dyn_predict, fused away)
        const T vp = pred_v ? v[o] + c2 * a[o] : v[o];
        const T av = alpha != T(0) ? (r * invM[o] - alpha * vp) * scale : r * invM[o];
        a[o] = av;
        v[o] = vp + c3 * av;
      } else {
        r_out[o] = alpha != T(0) ? r - alpha * mlump[o] * v[o] : r;
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_dyn_implicit_post(int64_t n, const T *a, T *v, T *d, T c3,
                                                          T c4) {
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) {
    v[i] = v[i] + c3 * a[i];
    d[i] = d[i] + c4 * a[i];
  }
}

// out = x + b y
template <typename T>
__global__ __launch_bounds__(NT) void k_xpby(int64_t n, const T *x, const T *y, T b, T *out) {
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride)
    out[i] = x[i] + b * y[i];
}

// Receiver traces: at steps that are multiples of `every`, copy the state
// vector of the receiver nodes into sample (step / every - 1).
template <typename T>
__global__ __launch_bounds__(NT) void k_record(const int *step_base, int step_off, int every,
                                               const T *src, const int32_t *recv, int nrec,
                                               int dim, int max_samples, T *traces) {
  const int step = *step_base + step_off;
  if (step % every != 0) return;
  const int smp = step / every - 1;
  if (smp < 0 || smp >= max_samples) return;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= nrec * dim) return;
  const int r = i / dim, c = i - r * dim;
  traces[(int64_t(smp) * nrec + r) * dim + c] = src[int64_t(recv[r]) * dim + c];
}

// max |x_i| in double: per-block partials, then one block.
template <typename T>
__global__ __launch_bounds__(NT) void k_absmax_part(int64_t n, const T *x, double *part) {
  __shared__ double sh[NT / 64];
  double m = 0.0;
  const int64_t stride = int64_t(gridDim.x) * NT;
  for (int64_t i = int64_t(blockIdx.x) * NT + threadIdx.x; i < n; i += stride)
    m = fmax(m, fabs(double(x[i])));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < NT / 64; ++w) r = fmax(r, sh[w]);
    part[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(NT) void k_absmax_fin(const double *part, int np, double *out) {
  __shared__ double sh[NT / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += NT) m = fmax(m, part[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < NT / 64; ++w) r = fmax(r, sh[w]);
    *out = r;
  }
}

__global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

inline int grid_for(int64_t n) {
  int64_t g = (n + NT - 1) / NT;
  if (g < 1) g = 1;
  if (g > NODE_GRID_MAX) g = NODE_GRID_MAX;
  return int(g);
}

template <typename T>
PatchArgs<T> patch_args(const DevMesh &dm) {
  PatchArgs<T> a{};
  a.patches = dm.patches;
  a.erec_s = dm.erec_s;
  a.emax = dm.emax;
  a.pgid = dm.pgid;
  a.pslot = dm.pslot;
  a.pend = dm.pend;
  a.fixmask = dm.fixmask;
  a.slot_start = dm.slot_start;
  if constexpr (sizeof(T) == 8) {
    a.coords = reinterpret_cast<const T *>(dm.coords64);
    a.escale = reinterpret_cast<const T *>(dm.escale64);
    a.diagK = reinterpret_cast<const T *>(dm.diag64);
  } else {
    a.coords = reinterpret_cast<const T *>(dm.coords32);
    a.escale = reinterpret_cast<const T *>(dm.escale32);
    a.diagK = reinterpret_cast<const T *>(dm.diag32);
  }
  a.lam = T(dm.lam);
  a.mu = T(dm.mu);
  a.np = dm.np;
  a.p8 = (dm.np + 7) / 8;
  a.ent = dm.ent;
  a.pn = dm.pn;
  a.n_owned = dm.n_owned;
  return a;
}

template <typename T, int DIM, int MODE, int PT>
hipError_t patch_attr() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ebe_patch<T, DIM, MODE, PT>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <typename T, int DIM, int MODE, int PT>
void launch_patch_pt(const DevMesh &dm, const PatchArgs<T> &a, hipStream_t s) {
  const LdsLayout L = lds_layout(DIM, int(sizeof(T)), dm.pn, dm.ent);
  const int grid = 8 * a.p8;
  hipLaunchKernelGGL((k_ebe_patch<T, DIM, MODE, PT>), dim3(grid), dim3(PT), L.total, s, a);
}

template <typename T, int DIM, int MODE>
void launch_patch(const DevMesh &dm, const PatchArgs<T> &a, hipStream_t s) {
  if (dm.patch_threads == 1024)
    launch_patch_pt<T, DIM, MODE, 1024>(dm, a, s);
  else if (dm.patch_threads == 512)
    launch_patch_pt<T, DIM, MODE, 512>(dm, a, s);
  else
    launch_patch_pt<T, DIM, MODE, 256>(dm, a, s);
}

}  // namespace

hipError_t prepare_kernels() {
  // dynamic LDS above 64 KiB must be allowed per kernel, once, outside any
  // stream capture
  hipError_t e = hipSuccess, r;
#define FEMGPU_ATTR(T, D, M, P) \
  if ((r = patch_attr<T, D, M, P>()) != hipSuccess) e = r;
#define FEMGPU_ATTR_TD(T, D)                                                              \
  FEMGPU_ATTR(T, D, 0, 256) FEMGPU_ATTR(T, D, 1, 256) FEMGPU_ATTR(T, D, 2, 256)                \
  FEMGPU_ATTR(T, D, 3, 256) FEMGPU_ATTR(T, D, 0, 512) FEMGPU_ATTR(T, D, 1, 512)                \
  FEMGPU_ATTR(T, D, 2, 512) FEMGPU_ATTR(T, D, 3, 512) FEMGPU_ATTR(T, D, 0, 1024)               \
  FEMGPU_ATTR(T, D, 1, 1024) FEMGPU_ATTR(T, D, 2, 1024) FEMGPU_ATTR(T, D, 3, 1024)          \
  FEMGPU_ATTR(T, D, 4, 256) FEMGPU_ATTR(T, D, 4, 512) FEMGPU_ATTR(T, D, 4, 1024)
  FEMGPU_ATTR_TD(double, 3)
  FEMGPU_ATTR_TD(double, 2)
  FEMGPU_ATTR_TD(float, 3)
  FEMGPU_ATTR_TD(float, 2)
#undef FEMGPU_ATTR_TD
#undef FEMGPU_ATTR
  return e;
}

int grid_blocks(int64_t n) { return grid_for(n); }

// the largest of the patch kernels' layouts (the single-reduction kernel's
// reduction area)
int lds_bytes(int dim, int elem_bytes, int pn, int ent) {
  return lds_layout(dim, elem_bytes, pn, ent, RED_CG1).total;
}

template <typename T>
void launch_matvec_patch(const DevMesh &dm, const T *x, T *y, T *partial, hipStream_t s) {
  PatchArgs<T> a = patch_args<T>(dm);
  a.x = x;
  a.y = y;
  a.partial = partial;
  if (dm.dim == 3)
    launch_patch<T, 3, 0>(dm, a, s);
  else
    launch_patch<T, 2, 0>(dm, a, s);
}

template <typename T>
void launch_matvec(const DevMesh &dm, const T *x, T *y, T *partial, hipStream_t s) {
  launch_matvec_patch<T>(dm, x, y, partial, s);
  if (dm.n_iface > 0) {
    const int g = int((dm.n_iface + NT - 1) / NT);
    if (dm.dim == 3)
      hipLaunchKernelGGL((k_iface_gather<T, 3>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_iface,
                         dm.slot_start, partial, y);
    else
      hipLaunchKernelGGL((k_iface_gather<T, 2>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_iface,
                         dm.slot_start, partial, y);
  }
}

template <typename T>
void launch_diag(const DevMesh &dm, T *d, T *partial, hipStream_t s) {
  PatchArgs<T> a = patch_args<T>(dm);
  a.y = d;
  a.partial = partial;
  if (dm.dim == 3)
    launch_patch<T, 3, 2>(dm, a, s);
  else
    launch_patch<T, 2, 2>(dm, a, s);
  if (dm.n_iface > 0) {
    const int g = int((dm.n_iface + NT - 1) / NT);
    if (dm.dim == 3)
      hipLaunchKernelGGL((k_iface_gather<T, 3>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_iface,
                         dm.slot_start, partial, d);
    else
      hipLaunchKernelGGL((k_iface_gather<T, 2>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_iface,
                         dm.slot_start, partial, d);
  }
}

template <typename T>
void launch_inv_diag(const DevMesh &dm, const T *diag, T cK, T cM, const T *m, T *invD,
                     hipStream_t s) {
  const int64_t n = dm.nn * dm.dim;
  hipLaunchKernelGGL((k_inv_diag<T>), dim3(grid_for(n)), dim3(NT), 0, s, n, diag, cK, cM, m, invD);
}

template <typename T>
void launch_pcg_start(const DevMesh &dm, T *x, T *p, int64_t n_all, const T *r, const T *invD,
                      CgScalars *sc, int kind, double eps, int max_iter, double *part_node,
                      hipStream_t s) {
  const int64_t n = dm.n_active * dm.dim;
  double *prz = part_node + 2 * NODE_PARTS, *prr = part_node + 3 * NODE_PARTS;  // slot 1
  hipLaunchKernelGGL((k_pcg_init<T>), dim3(NODE_PARTS), dim3(UPT), 0, s, n, r, invD, prz, prr, x,
                     p, n_all);
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(NT), 0, s, sc, prz, prr, NODE_PARTS, 1, kind, eps,
                     max_iter);
}

template <typename T>
void launch_pcg_iter(const DevMesh &dm, T *x, T *r, T *p, T *q, T *partial, const T *invD,
                     OpShift<T> sh, CgScalars *sc, double *part_patch, double *part_node,
                     int parity, hipStream_t s) {
  PatchArgs<T> a = patch_args<T>(dm);
  a.x = r;
  a.invD = invD;
  a.xs = x;
  a.p = p;
  a.y = q;
  a.partial = partial;
  a.sh = sh;
  a.sc = sc;
  a.part = part_patch;
  a.prz = part_node + (2 * (parity ^ 1) + 0) * NODE_PARTS;
  a.prr = part_node + (2 * (parity ^ 1) + 1) * NODE_PARTS;
  a.parity = parity;
  if (dm.dim == 3)
    launch_patch<T, 3, 1>(dm, a, s);
  else
    launch_patch<T, 2, 1>(dm, a, s);
  UpdArgs<T> u{};
  u.n_owned = dm.n_owned;
  u.n_active = dm.n_active;
  u.slot_start = dm.slot_start;
  u.partial = partial;
  u.q = q;
  u.p = p;
  u.r = r;
  u.x = x;
  u.invD = invD;
  u.sh = sh;
  u.sc = sc;
  u.ppq = part_patch;
  u.np = dm.np;
  u.prz = part_node + (2 * parity + 0) * NODE_PARTS;
  u.prr = part_node + (2 * parity + 1) * NODE_PARTS;
  u.parity = parity;
  if (dm.dim == 3)
    hipLaunchKernelGGL((k_pcg_update<T, 3>), dim3(NODE_PARTS), dim3(UPT), 0, s, u);
  else
    hipLaunchKernelGGL((k_pcg_update<T, 2>), dim3(NODE_PARTS), dim3(UPT), 0, s, u);
}

template <typename T>
void launch_pcg_apply_x(const DevMesh &dm, T *x, const T *p, CgScalars *sc, hipStream_t s,
                        bool clear_pending) {
  const int64_t n = dm.n_active * dm.dim;
  hipLaunchKernelGGL((k_pcg_apply_x<T>), dim3(grid_for(n)), dim3(NT), 0, s, n, x, p, sc);
  if (clear_pending) hipLaunchKernelGGL(k_clear_x_pend, dim3(1), dim3(64), 0, s, sc);
}

void launch_residual64(const DevMesh &dm, const double *f, const double *y, const double *partial,
                       double *r, CgScalars *sc, double *part, hipStream_t s) {
  const int g = grid_for(dm.n_active);
  if (dm.dim == 3)
    hipLaunchKernelGGL((k_residual64<3>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_active,
                       dm.slot_start, f, y, partial, r, sc, part);
  else
    hipLaunchKernelGGL((k_residual64<2>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_active,
                       dm.slot_start, f, y, partial, r, sc, part);
  hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(NT), 0, s, sc, part, g);
}

void launch_flush_x(const DevMesh &dm, double *x64, float *x32, hipStream_t s) {
  const int64_t n = dm.n_active * dm.dim;
  hipLaunchKernelGGL(k_flush_x, dim3(grid_for(n)), dim3(NT), 0, s, n, x64, x32);
}

void launch_pcg_replace(const DevMesh &dm, const double *r64, float *r32, const float *invD,
                        double *part_node, hipStream_t s) {
  const int64_t n = dm.n_active * dm.dim;
  hipLaunchKernelGGL(k_pcg_replace, dim3(NODE_PARTS), dim3(UPT), 0, s, n, r64, r32, invD,
                     part_node + 2 * NODE_PARTS, part_node + 3 * NODE_PARTS);
}

void launch_fin_init(CgScalars *sc, double *part_node, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(NT), 0, s, sc, part_node + 2 * NODE_PARTS,
                     part_node + 3 * NODE_PARTS, NODE_PARTS, 0, 0, 0.0, 0);
}

void launch_pcg_resume(CgScalars *sc, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_resume, dim3(1), dim3(64), 0, s, sc);
}

template <typename T>
void launch_fill(T *p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n)), dim3(NT), 0, s, p, n, v);
}

template <typename T>
void launch_copy(const T *a, T *b, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_convert<T, T>), dim3(grid_for(n)), dim3(NT), 0, s, a, b, n);
}

template <typename TA, typename TB>
void launch_convert(const TA *a, TB *b, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_convert<TA, TB>), dim3(grid_for(n)), dim3(NT), 0, s, a, b, n);
}

template <typename T>
void launch_stress(const DevMesh &dm, const T *u, T *sigma, hipStream_t s, bool strain) {
  const T *coords = sizeof(T) == 8 ? reinterpret_cast<const T *>(dm.coords64)
                                   : reinterpret_cast<const T *>(dm.coords32);
  const T *es = sizeof(T) == 8 ? reinterpret_cast<const T *>(dm.escale64)
                               : reinterpret_cast<const T *>(dm.escale32);
  const int g = int((dm.ne + NT - 1) / NT);
  if (dm.dim == 3)
    hipLaunchKernelGGL((k_stress<T, 3>), dim3(g), dim3(NT), 0, s, dm.ne, dm.elems_int, coords, es,
                       T(dm.lam), T(dm.mu), u, sigma, strain);
  else
    hipLaunchKernelGGL((k_stress<T, 2>), dim3(g), dim3(NT), 0, s, dm.ne, dm.elems_int, coords, es,
                       T(dm.lam), T(dm.mu), u, sigma, strain);
}

template <typename T>
void launch_energy(const DevMesh &dm, const T *u, T *en, hipStream_t s) {
  const T *coords = sizeof(T) == 8 ? reinterpret_cast<const T *>(dm.coords64)
                                   : reinterpret_cast<const T *>(dm.coords32);
  const T *es = sizeof(T) == 8 ? reinterpret_cast<const T *>(dm.escale64)
                               : reinterpret_cast<const T *>(dm.escale32);
  const int g = int((dm.ne + NT - 1) / NT);
  if (dm.dim == 3)
    hipLaunchKernelGGL((k_energy<T, 3>), dim3(g), dim3(NT), 0, s, dm.ne, dm.elems_int, coords,
                       dm.fixmask, es, T(dm.lam), T(dm.mu), u, en);
  else
    hipLaunchKernelGGL((k_energy<T, 2>), dim3(g), dim3(NT), 0, s, dm.ne, dm.elems_int, coords,
                       dm.fixmask, es, T(dm.lam), T(dm.mu), u, en);
}

template <typename T>
void launch_bilinear(const DevMesh &dm, const T *u, const T *w, double factor, T *out,
                     hipStream_t s) {
  const T *coords = sizeof(T) == 8 ? reinterpret_cast<const T *>(dm.coords64)
                                   : reinterpret_cast<const T *>(dm.coords32);
  const int g = int((dm.ne + NT - 1) / NT);
  if (dm.dim == 3)
    hipLaunchKernelGGL((k_bilinear<T, 3>), dim3(g), dim3(NT), 0, s, dm.ne, dm.elems_int, coords,
                       dm.fixmask, T(dm.lam), T(dm.mu), u, w, T(factor), out);
  else
    hipLaunchKernelGGL((k_bilinear<T, 2>), dim3(g), dim3(NT), 0, s, dm.ne, dm.elems_int, coords,
                       dm.fixmask, T(dm.lam), T(dm.mu), u, w, T(factor), out);
}

template <typename T>
void launch_dyn_predict(const DevMesh &dm, T *d, T *v, T *a, DynCoeffs c, bool zero_a,
                        int *step_inc, hipStream_t s) {
  const int64_t n = dm.n_active * dm.dim;
  hipLaunchKernelGGL((k_dyn_predict<T>), dim3(grid_for(n)), dim3(NT), 0, s, n, d, v, a, T(c.dt),
                     T(c.c1), T(c.c2), int(zero_a), step_inc);
}

template <typename T>
void launch_dyn_correct(const DevMesh &dm, const T *y, const T *partial, const T *f_file,
                        const T *f_unit, const double *coef_file, const double *coef_unit,
                        const int *step_base, int step_off, const T *invM, T *a, T *v, T *r_out,
                        DynCoeffs c, hipStream_t s, const T *mlump, bool pred_v) {
  const int g = grid_for(dm.n_active);
  if (dm.dim == 3)
    hipLaunchKernelGGL((k_dyn_correct<T, 3>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_active,
                       dm.slot_start, y, partial, f_file, f_unit, coef_file, coef_unit, step_base,
                       step_off, invM, a, v, r_out, T(c.c3), mlump, T(c.alpha), T(c.expl_scale),
                       pred_v ? 1 : 0, T(c.c2));
  else
    hipLaunchKernelGGL((k_dyn_correct<T, 2>), dim3(g), dim3(NT), 0, s, dm.n_owned, dm.n_active,
                       dm.slot_start, y, partial, f_file, f_unit, coef_file, coef_unit, step_base,
                       step_off, invM, a, v, r_out, T(c.c3), mlump, T(c.alpha), T(c.expl_scale),
                       pred_v ? 1 : 0, T(c.c2));
}

template <typename T>
void launch_dyn_step(const DevMesh &dm, const DynStep<T> &st, DynCoeffs c, hipStream_t s) {
  PatchArgs<T> a = patch_args<T>(dm);
  a.x = st.d_in;
  a.vel = st.v_in;
  a.acc = st.a_in;
  a.xout = st.d_out;
  a.vout = st.v_out;
  a.aout = st.a_out;
  a.partial = st.part_out;
  a.prev_part = st.part_in;
  a.f_file = st.f_file;
  a.f_unit = st.f_unit;
  a.coef_file = st.coef_file;
  a.coef_unit = st.coef_unit;
  a.step_base = st.step_base;
  a.step_off = st.step_off;
  a.invM = st.invM;
  a.dt = T(c.dt);
  a.c1 = T(c.c1);
  a.c2 = T(c.c2);
  a.c3 = T(c.c3);
  a.alpha = T(c.alpha);
  a.scale = T(c.expl_scale);
  if (dm.dim == 3)
    launch_patch<T, 3, 4>(dm, a, s);
  else
    launch_patch<T, 2, 4>(dm, a, s);
}

template <typename T>
void launch_dyn_predict_matvec(const DevMesh &dm, const T *d_in, const T *v, const T *acc,
                               T *d_out, DynCoeffs c, T *y, T *partial, hipStream_t s) {
  PatchArgs<T> a = patch_args<T>(dm);
  a.x = d_in;
  a.vel = v;
  a.acc = acc;
  a.xout = d_out;
  a.dt = T(c.dt);
  a.c1 = T(c.c1);
  a.y = y;
  a.partial = partial;
  if (dm.dim == 3)
    launch_patch<T, 3, 3>(dm, a, s);
  else
    launch_patch<T, 2, 3>(dm, a, s);
}

template <typename T>
void launch_xpby(int64_t n, const T *x, const T *y, double b, T *out, hipStream_t s) {
  hipLaunchKernelGGL((k_xpby<T>), dim3(grid_for(n)), dim3(NT), 0, s, n, x, y, T(b), out);
}

template <typename T>
void launch_record(const int *step_base, int step_off, int every, const T *src, const int32_t *recv,
                   int nrec, int dim, int max_samples, T *traces, hipStream_t s) {
  const int g = (nrec * dim + NT - 1) / NT;
  hipLaunchKernelGGL((k_record<T>), dim3(g), dim3(NT), 0, s, step_base, step_off, every, src,
                     recv, nrec, dim, max_samples, traces);
}

template <typename T>
void launch_absmax(int64_t n, const T *x, double *part, double *out, hipStream_t s) {
  hipLaunchKernelGGL((k_absmax_part<T>), dim3(256), dim3(NT), 0, s, n, x, part);
  hipLaunchKernelGGL(k_absmax_fin, dim3(1), dim3(NT), 0, s, part, 256, out);
}

template <typename T>
void launch_dyn_implicit_post(const DevMesh &dm, const T *a, T *v, T *d, DynCoeffs c,
                              hipStream_t s) {
  const int64_t n = dm.n_active * dm.dim;
  hipLaunchKernelGGL((k_dyn_implicit_post<T>), dim3(grid_for(n)), dim3(NT), 0, s, n, a, v, d,
                     T(c.c3), T(c.c4));
}

void launch_set_int(int *p, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, p, v);
}

void launch_reset_scalars(CgScalars *sc, int kind, double eps, int max_iter, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(64), 0, s, sc, kind, eps, max_iter);
}

#define FEMGPU_INST(T)                                                                          \
  template void launch_matvec<T>(const DevMesh &, const T *, T *, T *, hipStream_t);           \
  template void launch_matvec_patch<T>(const DevMesh &, const T *, T *, T *, hipStream_t);     \
  template void launch_diag<T>(const DevMesh &, T *, T *, hipStream_t);                        \
  template void launch_inv_diag<T>(const DevMesh &, const T *, T, T, const T *, T *,           \
                                   hipStream_t);                                               \
  template void launch_pcg_start<T>(const DevMesh &, T *, T *, int64_t, const T *, const T *,   \
                                    CgScalars *, int, double, int, double *, hipStream_t);     \
  template void launch_pcg_apply_x<T>(const DevMesh &, T *, const T *, CgScalars *,            \
                                      hipStream_t, bool);                                      \
  template void launch_pcg_iter<T>(const DevMesh &, T *, T *, T *, T *, T *, const T *,        \
                                   OpShift<T>, CgScalars *, double *, double *, int,           \
                                   hipStream_t);                                               \
  template void launch_fill<T>(T *, int64_t, T, hipStream_t);                                  \
  template void launch_copy<T>(const T *, T *, int64_t, hipStream_t);                          \
  template void launch_stress<T>(const DevMesh &, const T *, T *, hipStream_t, bool);                \
  template void launch_energy<T>(const DevMesh &, const T *, T *, hipStream_t);                \
  template void launch_bilinear<T>(const DevMesh &, const T *, const T *, double, T *, hipStream_t); \
  template void launch_dyn_predict<T>(const DevMesh &, T *, T *, T *, DynCoeffs, bool, int *,  \
                                      hipStream_t);                                            \
  template void launch_dyn_correct<T>(const DevMesh &, const T *, const T *, const T *,        \
                                      const T *, const double *, const double *, const int *,  \
                                      int, const T *, T *, T *, T *, DynCoeffs, hipStream_t,   \
                                      const T *, bool);                                        \
  template void launch_dyn_step<T>(const DevMesh &, const DynStep<T> &, DynCoeffs, hipStream_t); \
  template void launch_dyn_predict_matvec<T>(const DevMesh &, const T *, const T *, const T *, \
                                             T *, DynCoeffs, T *, T *, hipStream_t);          \
  template void launch_dyn_implicit_post<T>(const DevMesh &, const T *, T *, T *, DynCoeffs,   \
                                            hipStream_t);                                      \
  template void launch_xpby<T>(int64_t, const T *, const T *, double, T *, hipStream_t);       \
  template void launch_record<T>(const int *, int, int, const T *, const int32_t *, int, int, \
                                 int, T *, hipStream_t);                                       \
  template void launch_absmax<T>(int64_t, const T *, double *, double *, hipStream_t);

FEMGPU_INST(double)
FEMGPU_INST(float)
template void launch_convert<double, float>(const double *, float *, int64_t, hipStream_t);
template void launch_convert<float, double>(const float *, double *, int64_t, hipStream_t);
template void launch_fill<int>(int *, int64_t, int, hipStream_t);

}  // namespace femgpu
