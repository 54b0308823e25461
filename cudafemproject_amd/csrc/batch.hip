// batch.hip -- batched independent static solves in one persistent launch.
//
// One workgroup runs the complete Jacobi-PCG of one instance (a small mesh
// with NR right-hand sides) from diagonal to converged solution: no host
// round trip, no inter-workgroup communication, every loop bounded by
// max_iter.  Workgroups take instances largest first.
//
// Per PCG iteration of an instance: q = K p for all NR right-hand sides (the
// instance's element patches one after another, each staged in LDS; the
// element geometry is formed once per element and applied to the NR
// vectors), per-rhs p'q, alpha, x += alpha p, r -= alpha q, z = D^-1 r,
// r'z and r'r, beta, p = z + beta p.  Dot products reduce in FP64 with a
// fixed order, so results are bitwise reproducible.
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

constexpr int BT = 256;  // threads per workgroup

struct BLds {
  int off_c, off_x, off_res, off_end, off_fix, off_red, total;
};

__host__ __device__ inline int bal16(int v) { return (v + 15) & ~15; }

// LDS regions sized for double so the FP64 passes share them.
__host__ __device__ inline BLds blds(int dim, int nr, int pn, int emax) {
  BLds L;
  L.off_c = 0;
  L.off_x = bal16(pn * dim * 8);
  L.off_res = L.off_x + bal16(pn * nr * dim * 8);
  L.off_end = L.off_res + bal16(emax * (dim + 1) * nr * dim * 8);
  L.off_fix = L.off_end + bal16(pn * 2);
  L.off_red = L.off_fix + bal16(pn);
  L.total = L.off_red + (BT / 64) * 4 * 8;
  return L;
}

// Thread t holds a partial of right-hand side t % NR; every thread gets the
// NR totals (fixed order: wave shuffles, then waves in order).
template <int NR>
__device__ __forceinline__ void block_sum_rhs(double v, double *sh, double *out) {
#pragma unroll
  for (int o = 32; o >= NR; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l < NR) sh[w * NR + l] = v;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BT / 64; ++i) s += sh[i * NR + r];
    out[r] = s;
  }
}

template <int NR, typename V>
__device__ __forceinline__ V pick(const V *v, int r) {
  V o = v[0];
#pragma unroll
  for (int i = 1; i < NR; ++i)
    if (r == i) o = v[i];
  return o;
}

struct Plan {
  const PatchInfo *patches;
  const uint32_t *elem_lid;
  const uint16_t *epos;
  const uint16_t *node_end;
  const int32_t *sh_gid;
  const int32_t *sh_slot;
  const int32_t *slot_start;
  const uint8_t *fixmask;
};

// MODE 0: y = K x (NR vectors).  MODE 2: y = diag(K) (one vector).
// x, y, slots point at the instance's first entry; node and slot ids are
// made instance-local.  Starts and ends with a workgroup barrier.
template <typename TT, int DIM, int NR, int MODE>
__device__ void bmatvec(const Plan &pl, const BInst &I, const TT *coords, const TT *x, TT *y,
                        TT *slots, unsigned char *lds, const BLds &L) {
  constexpr int NPE = DIM + 1, NDE = DIM * NPE;
  constexpr int NV = MODE == 2 ? 1 : NR;
  TT *s_c = reinterpret_cast<TT *>(lds + L.off_c);
  TT *s_x = reinterpret_cast<TT *>(lds + L.off_x);
  TT *s_res = reinterpret_cast<TT *>(lds + L.off_res);
  uint16_t *s_end = reinterpret_cast<uint16_t *>(lds + L.off_end);
  uint8_t *s_fix = reinterpret_cast<uint8_t *>(lds + L.off_fix);
  const TT lam = TT(I.lam), mu = TT(I.mu);
  __syncthreads();
  for (int pi = I.patch_begin; pi < I.patch_end; ++pi) {
    const PatchInfo P = pl.patches[pi];
    const int nloc = P.n_own + P.n_sh;
    // Phase A: stage node coordinates, vectors, constraint bits, run ends.
    for (int k = threadIdx.x; k < nloc; k += BT) {
      const int64_t gid = k < P.n_own ? int64_t(P.own_start + k)
                                      : int64_t(pl.sh_gid[P.sh_start + k - P.n_own]);
      const int64_t lg = gid - I.node_off;
#pragma unroll
      for (int c = 0; c < DIM; ++c) s_c[k * DIM + c] = coords[gid * DIM + c];
      if constexpr (MODE == 0) {
#pragma unroll
        for (int i = 0; i < NR * DIM; ++i) s_x[k * NR * DIM + i] = x[lg * NR * DIM + i];
      }
      s_fix[k] = pl.fixmask[gid];
      s_end[k] = pl.node_end[P.pnode_start + k];
    }
    __syncthreads();
    // Phase B: element forces (geometry once, NR vectors) to their
    // node-sorted entry positions.
    for (int i = threadIdx.x; i < P.n_elem; i += BT) {
      const uint32_t lid = pl.elem_lid[P.elem_start + i];
      const uint2 pp = reinterpret_cast<const uint2 *>(pl.epos)[P.elem_start + i];
      const unsigned pos[4] = {pp.x & 0xffffu, pp.x >> 16, pp.y & 0xffffu, pp.y >> 16};
      int l[NPE];
      TT X[NDE];
      unsigned fix = 0;
#pragma unroll
      for (int n = 0; n < NPE; ++n) {
        l[n] = int((lid >> (8 * n)) & 255u);
#pragma unroll
        for (int c = 0; c < DIM; ++c) X[n * DIM + c] = s_c[l[n] * DIM + c];
        fix |= unsigned(s_fix[l[n]] & ((1 << DIM) - 1)) << (n * DIM);
      }
      if constexpr (MODE == 2) {
        TT F[NDE];
        elem_diag<TT, DIM>(X, lam, mu, TT(1), F);
#pragma unroll
        for (int n = 0; n < NPE; ++n)
#pragma unroll
          for (int c = 0; c < DIM; ++c) s_res[pos[n] * DIM + c] = F[n * DIM + c];
      } else {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          TT U[NDE], F[NDE];
#pragma unroll
          for (int n = 0; n < NPE; ++n)
#pragma unroll
            for (int c = 0; c < DIM; ++c) U[n * DIM + c] = s_x[(l[n] * NR + r) * DIM + c];
          (void)elem_force<TT, DIM>(X, U, lam, mu, TT(1), fix, F);
#pragma unroll
          for (int n = 0; n < NPE; ++n)
#pragma unroll
            for (int c = 0; c < DIM; ++c) s_res[(pos[n] * NR + r) * DIM + c] = F[n * DIM + c];
        }
      }
    }
    __syncthreads();
    // Phase C: fixed-order sums of each node's entry run.
    for (int it = threadIdx.x; it < nloc * NV; it += BT) {
      const int k = it / NV, r = it % NV;
      const int b0 = k ? int(s_end[k - 1]) : 0, b1 = int(s_end[k]);
      TT acc[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) acc[c] = TT(0);
      for (int j = b0; j < b1; ++j)
#pragma unroll
        for (int c = 0; c < DIM; ++c) acc[c] += s_res[(j * NV + r) * DIM + c];
      if (k < P.n_own) {
        const int64_t lg = int64_t(P.own_start + k) - I.node_off;
#pragma unroll
        for (int c = 0; c < DIM; ++c) y[(lg * NV + r) * DIM + c] = acc[c];
      } else {
        const int64_t ls = int64_t(pl.sh_slot[P.sh_start + k - P.n_own]) - I.slot_off;
#pragma unroll
        for (int c = 0; c < DIM; ++c) slots[(ls * NV + r) * DIM + c] = acc[c];
      }
    }
    __syncthreads();
  }
  // Interface nodes: fixed-order sums of their slots.
  const int n_if = I.n_active - I.n_owned;
  for (int it = threadIdx.x; it < n_if * NV; it += BT) {
    const int j = it / NV, r = it % NV;
    const int s0 = pl.slot_start[I.iface_off + j] - I.slot_off;
    const int s1 = pl.slot_start[I.iface_off + j + 1] - I.slot_off;
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

template <typename T, int DIM, int NR>
__global__ __launch_bounds__(BT) void k_batch_pcg(BatchArgs<T> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr bool MIXED = sizeof(T) == sizeof(float);
  const int inst = a.order[blockIdx.x];
  const BInst I = a.inst[inst];
  const BLds L = blds(DIM, NR, a.pn, a.emax);
  double *s_red = reinterpret_cast<double *>(lds + L.off_red);
  const Plan pl{a.patches, a.elem_lid, a.epos, a.node_end, a.sh_gid, a.sh_slot, a.slot_start,
                a.fixmask};
  const int64_t vb = int64_t(I.node_off) * NR * DIM;   // first vector entry
  const int64_t sb = int64_t(I.slot_off) * NR * DIM;   // first slot entry
  const double *f = a.f + vb;
  double *x = a.x + vb, *y64 = a.y64 + vb, *sl64 = a.slots64 + sb;
  T *r = a.r + vb, *p = a.p + vb, *q = a.q + vb, *slT = a.slotsT + sb;
  T *iD = a.invD + int64_t(I.node_off) * DIM;
  const int rt = threadIdx.x % NR;        // right-hand side of this thread's entries
  const int n_it = I.n_active * NR;       // (node, rhs) items with an equation

  // diag(K) in FP64 -> D^-1 in the inner precision
  bmatvec<double, DIM, NR, 2>(pl, I, a.coords64, nullptr, y64, sl64, lds, L);
  for (int i = threadIdx.x; i < I.n_active * DIM; i += BT) iD[i] = T(1.0 / y64[i]);
  __syncthreads();

  // r = f, z = D^-1 r, p = z, x = 0
  for (int i = threadIdx.x; i < I.n_nodes * NR * DIM; i += BT) x[i] = 0.0;
  double gp = 0.0, fp = 0.0;
  for (int i = threadIdx.x; i < n_it; i += BT) {
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
  double gam[NR], ff[NR], thr[NR], rr_last[NR];
  bool active[NR];
  int status[NR], iters[NR], nrep[NR];
  block_sum_rhs<NR>(gp, s_red, gam);
  block_sum_rhs<NR>(fp, s_red, ff);
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    active[k] = ff[k] > 0.0;
    status[k] = isfinite(ff[k]) ? FEMGPU_OK : FEMGPU_E_NONFINITE;
    if (!isfinite(ff[k])) active[k] = false;
    thr[k] = a.tol * a.tol * ff[k];
    rr_last[k] = ff[k];
    iters[k] = 0;
    nrep[k] = 0;
  }
  const double drop2 = MIXED ? a.drop * a.drop : 0.0;

  for (int it = 0; it < a.max_iter; ++it) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < NR; ++k) any = any || active[k];
    if (!any) break;

    // q = K p
    bmatvec<T, DIM, NR, 0>(pl, I, a.coordsT, p, q, slT, lds, L);
    double pqp = 0.0;
    for (int i = threadIdx.x; i < n_it; i += BT)
#pragma unroll
      for (int c = 0; c < DIM; ++c) pqp += double(p[i * DIM + c]) * double(q[i * DIM + c]);
    double pq[NR], alpha[NR];
    block_sum_rhs<NR>(pqp, s_red, pq);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      alpha[k] = 0.0;
      if (!active[k]) continue;
      const double al = gam[k] / pq[k];
      if (!(pq[k] > 0.0) || !isfinite(al)) {
        status[k] = isfinite(pq[k]) ? FEMGPU_E_BREAKDOWN : FEMGPU_E_NONFINITE;
        active[k] = false;
      } else {
        alpha[k] = al;
        iters[k] = it + 1;
      }
    }
    // x += alpha p, r -= alpha q, z = D^-1 r, partial r'z and r'r
    const double my_al = pick<NR>(alpha, rt);
    double g2 = 0.0, rr2 = 0.0;
    if (my_al != 0.0) {
      const T alT = T(my_al);
      for (int i = threadIdx.x; i < n_it; i += BT) {
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
    double gnew[NR], rr[NR];
    block_sum_rhs<NR>(g2, s_red, gnew);
    block_sum_rhs<NR>(rr2, s_red, rr);

    // residual replacement / convergence check on the FP64 residual
    unsigned trig = 0;
#pragma unroll
    for (int k = 0; k < NR; ++k)
      if (active[k] && (rr[k] <= thr[k] || rr[k] < drop2 * rr_last[k] || !isfinite(rr[k])))
        trig |= 1u << k;
    bool restart[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) restart[k] = false;
    if (trig) {
      bmatvec<double, DIM, NR, 0>(pl, I, a.coords64, x, y64, sl64, lds, L);
      double g3 = 0.0, rr3 = 0.0;
      if ((trig >> rt) & 1u) {
        for (int i = threadIdx.x; i < n_it; i += BT) {
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
      double gR[NR], rR[NR];
      block_sum_rhs<NR>(g3, s_red, gR);
      block_sum_rhs<NR>(rr3, s_red, rR);
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if (!((trig >> k) & 1u)) continue;
        ++nrep[k];
        if (!isfinite(rR[k])) {
          status[k] = FEMGPU_E_NONFINITE;
          active[k] = false;
        } else if (rR[k] <= thr[k]) {
          active[k] = false;
        } else {
          restart[k] = rR[k] > 0.25 * rr_last[k];
          rr_last[k] = rR[k];
          gnew[k] = gR[k];
        }
      }
    }
    // p = z + beta p
    double beta[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      beta[k] = restart[k] ? 0.0 : gnew[k] / gam[k];
      if (active[k]) gam[k] = gnew[k];
    }
    bool my_active = false;
#pragma unroll
    for (int k = 0; k < NR; ++k)
      if (k == rt) my_active = active[k];
    if (my_active) {
      const T bt = T(pick<NR>(beta, rt));
      for (int i = threadIdx.x; i < n_it; i += BT) {
        const int node = i / NR;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const int j = i * DIM + c;
          p[j] = iD[node * DIM + c] * r[j] + bt * p[j];
        }
      }
    }
  }

  // final FP64 residual of every right-hand side
  bmatvec<double, DIM, NR, 0>(pl, I, a.coords64, x, y64, sl64, lds, L);
  double rr4 = 0.0;
  for (int i = threadIdx.x; i < n_it; i += BT)
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const double rv = f[i * DIM + c] - y64[i * DIM + c];
      rr4 += rv * rv;
    }
  double rF[NR];
  block_sum_rhs<NR>(rr4, s_red, rF);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      BStat s;
      s.status = status[k] == FEMGPU_OK && active[k] ? FEMGPU_E_MAXIT : status[k];
      s.iterations = iters[k];
      s.replacements = nrep[k];
      s.pad = 0;
      s.rel_res = ff[k] > 0.0 ? sqrt(rF[k] / ff[k]) : 0.0;
      a.stat[int64_t(inst) * NR + k] = s;
    }
  }
}

template <typename T, int DIM, int NR>
void launch_one(const BatchArgs<T> &a, int n_inst, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_pcg<T, DIM, NR>), dim3(n_inst), dim3(BT), lds, s, a);
}

template <typename T, int DIM, int NR>
void set_attr() {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_pcg<T, DIM, NR>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

}  // namespace

size_t batch_lds_bytes(int dim, int nr, int pn, int emax) {
  return size_t(blds(dim, nr, pn, emax).total);
}

void batch_prepare() {
  set_attr<float, 2, 1>();
  set_attr<float, 2, 2>();
  set_attr<float, 2, 4>();
  set_attr<float, 3, 1>();
  set_attr<float, 3, 2>();
  set_attr<float, 3, 4>();
  set_attr<double, 2, 1>();
  set_attr<double, 2, 2>();
  set_attr<double, 2, 4>();
  set_attr<double, 3, 1>();
  set_attr<double, 3, 2>();
  set_attr<double, 3, 4>();
}

int launch_batch_pcg(int dim, int nr, bool mixed, const BatchArgs<float> &af,
                     const BatchArgs<double> &ad, int n_inst, size_t lds, hipStream_t s) {
#define FEMGPU_BATCH_CASE(D, R)                                  \
  if (dim == D && nr == R) {                                     \
    if (mixed)                                                   \
      launch_one<float, D, R>(af, n_inst, lds, s);               \
    else                                                         \
      launch_one<double, D, R>(ad, n_inst, lds, s);              \
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
