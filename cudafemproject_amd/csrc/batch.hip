// batch.hip -- batched independent static solves in one persistent launch.
//
// One workgroup runs the complete Jacobi-PCG of one instance (a small mesh
// with NR right-hand sides) from diagonal to converged solution: no host
// round trip, no inter-workgroup communication, every loop bounded by
// max_iter.  Workgroups take instances largest first.
//
// Per PCG iteration of an instance: q = K p for all NR right-hand sides in
// three workgroup-wide passes (element forces, geometry formed once per
// element and applied to the NR vectors, written to their node-sorted entry
// positions; fixed-order sums of each patch node's entry run; fixed-order sums
// of the interface slots), per-rhs p'q, alpha, x += alpha p, r -= alpha q,
// z = D^-1 r, r'z and r'r, beta, p = z + beta p.  Dot products reduce in FP64
// with a fixed order, so results are bitwise reproducible.
//
// Mixed precision: the inner vectors and the matvec are FP32; x is
// accumulated in FP64 and the residual of a right-hand side is replaced by
// the FP64 residual f - K x (an FP64 matvec in the same workgroup) whenever
// its FP32 residual has dropped by `drop` since the previous replacement, and
// when it meets the tolerance; a right-hand side converges on its FP64
// residual.  If a replacement shows less than a halving of the FP64 residual
// the search direction is restarted (p = z).
//
// The reference has no batched solve; each instance is the static problem of
// gpuPCG_EbE_vec (femfunc.cpp:25-128) with the relative-residual stopping rule.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "element.h"
#include "femgpu.h"

namespace femgpu {

namespace {

using namespace elem;

constexpr int BT = 1024;  // threads per workgroup

// Thread t holds a partial of right-hand side t % NR; dst[r] (LDS) gets the
// total of right-hand side r (fixed order: wave shuffles, then waves in
// order).  Ends with a barrier, so every thread can read dst.
template <int NR>
__device__ __forceinline__ void reduce_rhs(double v, double *sh, double *dst) {
#pragma unroll
  for (int o = 32; o >= NR; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l < NR) sh[w * NR + l] = v;
  __syncthreads();
  if (threadIdx.x < NR) {
    double s = 0.0;
    for (int i = 0; i < BT / 64; ++i) s += sh[i * NR + threadIdx.x];
    dst[threadIdx.x] = s;
  }
  __syncthreads();
}

// MODE 0: y = K x (NR vectors).  MODE 2: y = diag(K) (one vector).
// x, y, slots, ent point at the instance's first entry.  Starts and ends
// with a workgroup barrier.
template <typename TT, int DIM, int NR, int MODE, typename A>
__device__ void bmatvec(const A &a, const BInst &I, const TT *coords, const TT *x, TT *y,
                        TT *slots, TT *ent) {
  constexpr int NPE = DIM + 1, NDE = DIM * NPE;
  constexpr int NV = MODE == 2 ? 1 : NR;
  const TT lam = TT(I.lam), mu = TT(I.mu);
  const int64_t n0 = I.node_off;
  __syncthreads();
  // pass 1: element forces to their entry positions
  for (int e = threadIdx.x; e < I.n_elems; e += BT) {
    const int64_t ge = int64_t(I.elem_off) + e;
    int nd[NPE], ps[NPE];
    TT X[NDE];
    unsigned fix = 0;
#pragma unroll
    for (int n = 0; n < NPE; ++n) {
      nd[n] = a.elems[ge * NPE + n];
      ps[n] = a.gpos[ge * NPE + n];
#pragma unroll
      for (int c = 0; c < DIM; ++c) X[n * DIM + c] = coords[(n0 + nd[n]) * DIM + c];
      fix |= unsigned(a.fixmask[n0 + nd[n]] & ((1 << DIM) - 1)) << (n * DIM);
    }
    if constexpr (MODE == 2) {
      TT F[NDE];
      elem_diag<TT, DIM>(X, lam, mu, TT(1), F);
#pragma unroll
      for (int n = 0; n < NPE; ++n)
#pragma unroll
        for (int c = 0; c < DIM; ++c) ent[int64_t(ps[n]) * DIM + c] = F[n * DIM + c];
    } else {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        TT U[NDE], F[NDE];
#pragma unroll
        for (int n = 0; n < NPE; ++n)
#pragma unroll
          for (int c = 0; c < DIM; ++c) U[n * DIM + c] = x[(int64_t(nd[n]) * NR + r) * DIM + c];
        (void)elem_force<TT, DIM>(X, U, lam, mu, TT(1), fix, F);
#pragma unroll
        for (int n = 0; n < NPE; ++n)
#pragma unroll
          for (int c = 0; c < DIM; ++c) ent[(int64_t(ps[n]) * NR + r) * DIM + c] = F[n * DIM + c];
      }
    }
  }
  __syncthreads();
  // pass 2: fixed-order sums of each patch node's entry run
  for (int it = threadIdx.x; it < I.n_pn * NV; it += BT) {
    const int k = it / NV, r = it % NV;
    const int64_t gk = int64_t(I.pn_off) + k;
    const int b0 = a.pn_beg[gk], b1 = a.pn_end[gk], tg = a.pn_tgt[gk];
    TT acc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = TT(0);
    for (int j = b0; j < b1; ++j)
#pragma unroll
      for (int c = 0; c < DIM; ++c) acc[c] += ent[(int64_t(j) * NV + r) * DIM + c];
    TT *dst = tg >= 0 ? y + (int64_t(tg) * NV + r) * DIM : slots + (int64_t(-1 - tg) * NV + r) * DIM;
#pragma unroll
    for (int c = 0; c < DIM; ++c) dst[c] = acc[c];
  }
  __syncthreads();
  // pass 3: interface nodes, fixed-order sums of their slots
  const int n_if = I.n_active - I.n_owned;
  for (int it = threadIdx.x; it < n_if * NV; it += BT) {
    const int j = it / NV, r = it % NV;
    const int s0 = a.slot_start[I.iface_off + j] - I.slot_off;
    const int s1 = a.slot_start[I.iface_off + j + 1] - I.slot_off;
    TT acc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = TT(0);
    for (int s = s0; s < s1; ++s)
#pragma unroll
      for (int c = 0; c < DIM; ++c) acc[c] += slots[(int64_t(s) * NV + r) * DIM + c];
    const int64_t lg = int64_t(I.n_owned + j);
#pragma unroll
    for (int c = 0; c < DIM; ++c) y[(lg * NV + r) * DIM + c] = acc[c];
  }
  __syncthreads();
}

// Per right-hand side scalars of one instance, in LDS.  Thread r < NR runs
// the scalar logic of right-hand side r; the others read what they need.
struct RhsState {
  double gam[4], ff[4], thr[4], rr_last[4];
  double pq[4], g[4], rr[4], g_rep[4], rr_rep[4];
  double alpha[4], beta[4];
  int active[4], status[4], iters[4], nrep[4], trig[4], restart[4], stag[4];
  int any;
};

template <typename T, int DIM, int NR>
__global__ __launch_bounds__(BT) void k_batch_pcg(BatchArgs<T> a) {
  __shared__ double s_red[(BT / 64) * 4];
  __shared__ RhsState S;
  constexpr bool MIXED = sizeof(T) == sizeof(float);
  const int inst = a.order[blockIdx.x];
  const BInst I = a.inst[inst];
  const int64_t vb = int64_t(I.node_off) * NR * DIM;   // first vector entry
  const int64_t sb = int64_t(I.slot_off) * NR * DIM;   // first slot entry
  // entries of this instance, FP64 view and inner-precision view of the same bytes
  double *e64 = a.ent + int64_t(I.elem_off) * (DIM + 1) * NR * DIM;
  T *eT = reinterpret_cast<T *>(e64);
  const double *f = a.f + vb;
  double *x = a.x + vb, *y64 = a.y64 + vb, *sl64 = a.slots64 + sb;
  T *r = a.r + vb, *p = a.p + vb, *q = a.q + vb, *slT = a.slotsT + sb;
  T *iD = a.invD + int64_t(I.node_off) * DIM;
  const int tid = threadIdx.x;
  const int rt = tid % NR;            // right-hand side of this thread's entries
  const int n_it = I.n_active * NR;   // (node, rhs) items with an equation
  const double drop2 = MIXED ? a.drop * a.drop : 0.0;

  // diag(K) in FP64 -> D^-1 in the inner precision
  bmatvec<double, DIM, NR, 2>(a, I, a.coords64, static_cast<const double *>(nullptr), y64, sl64, e64);
  for (int i = tid; i < I.n_active * DIM; i += BT) iD[i] = T(1.0 / y64[i]);
  __syncthreads();

  // r = f, z = D^-1 r, p = z, x = 0
  for (int i = tid; i < I.n_nodes * NR * DIM; i += BT) x[i] = 0.0;
  double gp = 0.0, fp = 0.0;
  for (int i = tid; i < n_it; i += BT) {
    const int node = i / NR;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const double fv = f[i * DIM + c];
      const T rv = T(fv);
      const T z = iD[node * DIM + c] * rv;
      r[i * DIM + c] = rv;
      p[i * DIM + c] = z;
      gp += double(rv) * double(z);
      fp += fv * fv;
    }
  }
  reduce_rhs<NR>(gp, s_red, S.gam);
  reduce_rhs<NR>(fp, s_red, S.ff);
  if (tid < NR) {
    const double ff = S.ff[tid];
    S.active[tid] = ff > 0.0 && isfinite(ff);
    S.status[tid] = isfinite(ff) ? FEMGPU_OK : FEMGPU_E_NONFINITE;
    S.thr[tid] = a.tol * a.tol * ff;
    S.rr_last[tid] = ff;
    S.iters[tid] = 0;
    S.nrep[tid] = 0;
    S.stag[tid] = 0;
  }
  __syncthreads();
  if (tid == 0) {
    int any = 0;
    for (int k = 0; k < NR; ++k) any |= S.active[k];
    S.any = any;
  }
  __syncthreads();

  for (int it = 0; it < a.max_iter && S.any; ++it) {
    // q = K p, p'q
    bmatvec<T, DIM, NR, 0>(a, I, a.coordsT, p, q, slT, eT);
    double pqp = 0.0;
    for (int i = tid; i < n_it; i += BT)
#pragma unroll
      for (int c = 0; c < DIM; ++c) pqp += double(p[i * DIM + c]) * double(q[i * DIM + c]);
    reduce_rhs<NR>(pqp, s_red, S.pq);
    if (tid < NR) {
      S.alpha[tid] = 0.0;
      if (S.active[tid]) {
        const double pq = S.pq[tid], al = S.gam[tid] / pq;
        if (!(pq > 0.0) || !isfinite(al)) {
          S.status[tid] = isfinite(pq) ? FEMGPU_E_BREAKDOWN : FEMGPU_E_NONFINITE;
          S.active[tid] = 0;
        } else {
          S.alpha[tid] = al;
          S.iters[tid] = it + 1;
        }
      }
    }
    __syncthreads();
    // x += alpha p, r -= alpha q, z = D^-1 r, partial r'z and r'r
    const double my_al = S.alpha[rt];
    double g2 = 0.0, rr2 = 0.0;
    if (my_al != 0.0) {
      const T alT = T(my_al);
      for (int i = tid; i < n_it; i += BT) {
        const int node = i / NR;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const int j = i * DIM + c;
          x[j] += my_al * double(p[j]);
          const T rv = r[j] - alT * q[j];
          r[j] = rv;
          const T z = iD[node * DIM + c] * rv;
          g2 += double(rv) * double(z);
          rr2 += double(rv) * double(rv);
        }
      }
    }
    reduce_rhs<NR>(g2, s_red, S.g);
    reduce_rhs<NR>(rr2, s_red, S.rr);
    // residual replacement / convergence check on the FP64 residual
    if (tid < NR) {
      const double rr = S.rr[tid];
      S.trig[tid] = S.active[tid] &&
                    (rr <= S.thr[tid] || rr < drop2 * S.rr_last[tid] || !isfinite(rr));
      S.restart[tid] = 0;
    }
    __syncthreads();
    int trig = 0;
#pragma unroll
    for (int k = 0; k < NR; ++k) trig |= S.trig[k];
    if (trig) {
      bmatvec<double, DIM, NR, 0>(a, I, a.coords64, x, y64, sl64, e64);
      double g3 = 0.0, rr3 = 0.0;
      if (S.trig[rt]) {
        for (int i = tid; i < n_it; i += BT) {
          const int node = i / NR;
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            const int j = i * DIM + c;
            const double rv64 = f[j] - y64[j];
            const T rv = T(rv64);
            r[j] = rv;
            const T z = iD[node * DIM + c] * rv;
            g3 += double(rv) * double(z);
            rr3 += rv64 * rv64;
          }
        }
      }
      reduce_rhs<NR>(g3, s_red, S.g_rep);
      reduce_rhs<NR>(rr3, s_red, S.rr_rep);
      if (tid < NR && S.trig[tid]) {
        const double rr = S.rr_rep[tid];
        ++S.nrep[tid];
        if (!isfinite(rr)) {
          S.status[tid] = FEMGPU_E_NONFINITE;
          S.active[tid] = 0;
        } else if (rr <= S.thr[tid]) {
          S.active[tid] = 0;
        } else {
          // two replacements in a row without a decrease of the FP64
          // residual: the tolerance is below the attainable accuracy
          S.stag[tid] = rr >= S.rr_last[tid] ? S.stag[tid] + 1 : 0;
          if (S.stag[tid] >= 2) {
            S.status[tid] = FEMGPU_E_STAGNATION;
            S.active[tid] = 0;
          }
          S.restart[tid] = rr > 0.25 * S.rr_last[tid];
          S.rr_last[tid] = rr;
          S.g[tid] = S.g_rep[tid];
        }
      }
    }
    // beta, p = z + beta p
    if (tid < NR) {
      S.beta[tid] = S.restart[tid] ? 0.0 : S.g[tid] / S.gam[tid];
      if (S.active[tid]) S.gam[tid] = S.g[tid];
    }
    __syncthreads();
    if (tid == 0) {
      int any = 0;
      for (int k = 0; k < NR; ++k) any |= S.active[k];
      S.any = any;
    }
    if (S.active[rt]) {
      const T bt = T(S.beta[rt]);
      for (int i = tid; i < n_it; i += BT) {
        const int node = i / NR;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const int j = i * DIM + c;
          p[j] = iD[node * DIM + c] * r[j] + bt * p[j];
        }
      }
    }
    __syncthreads();
  }

  // final FP64 residual of every right-hand side
  bmatvec<double, DIM, NR, 0>(a, I, a.coords64, x, y64, sl64, e64);
  double rr4 = 0.0;
  for (int i = tid; i < n_it; i += BT)
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const double rv = f[i * DIM + c] - y64[i * DIM + c];
      rr4 += rv * rv;
    }
  reduce_rhs<NR>(rr4, s_red, S.rr);
  if (tid < NR) {
    BStat st;
    st.status = S.status[tid] == FEMGPU_OK && S.active[tid] ? FEMGPU_E_MAXIT : S.status[tid];
    st.iterations = S.iters[tid];
    st.replacements = S.nrep[tid];
    st.pad = 0;
    st.rel_res = S.ff[tid] > 0.0 ? sqrt(S.rr[tid] / S.ff[tid]) : 0.0;
    a.stat[int64_t(inst) * NR + tid] = st;
  }
}

template <typename T, int DIM, int NR>
void launch_one(const BatchArgs<T> &a, int n_inst, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_pcg<T, DIM, NR>), dim3(n_inst), dim3(BT), 0, s, a);
}

}  // namespace

int launch_batch_pcg(int dim, int nr, bool mixed, const BatchArgs<float> &af,
                     const BatchArgs<double> &ad, int n_inst, hipStream_t s) {
#define FEMGPU_BATCH_CASE(D, R)                                  \
  if (dim == D && nr == R) {                                     \
    if (mixed)                                                   \
      launch_one<float, D, R>(af, n_inst, s);               \
    else                                                         \
      launch_one<double, D, R>(ad, n_inst, s);              \
    return 0;                                                    \
  }
  FEMGPU_BATCH_CASE(2, 1)
  FEMGPU_BATCH_CASE(2, 2)
  FEMGPU_BATCH_CASE(2, 4)
  FEMGPU_BATCH_CASE(3, 1)
  FEMGPU_BATCH_CASE(3, 2)
  FEMGPU_BATCH_CASE(3, 4)
#undef FEMGPU_BATCH_CASE
  return -1;
}

}  // namespace femgpu
