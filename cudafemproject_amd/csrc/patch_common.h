// patch_common.h -- device pieces shared by the patch kernels (kernels.hip:
// matvec, fused PCG step, diagonal; cg1.hip: single-reduction PCG step).
//
// A patch is a Morton-contiguous run of elements touching at most PT nodes,
// processed by one PT-thread workgroup: one patch node per thread, at most EPT
// elements per thread.  Element records and the LDS layout are described in
// femgpu_internal.h (pack_records).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "element.h"

namespace femgpu {
namespace patch {

constexpr int EPT = 3;  // elements per thread: emax <= EPT * PT, patch nodes <= PT

struct LdsLayout {
  int off_c, off_x, off_res, off_red, total;
};

__host__ __device__ inline int al32(int v) { return (v + 31) & ~31; }  // 32-byte aligned regions

// Reduction area of the single-reduction PCG kernel (cg1.hip): five block
// sums of up to 1024 threads, the largest of the patch kernels'.
constexpr int RED_CG1 = 5 * 16;

// Patch kernel LDS: node coordinates and vector (pn x dim), one component of
// the node-sorted element entries (ent, padded runs), reduction scratch of
// red doubles (48: up to three block sums of 1024 threads).
__host__ __device__ inline LdsLayout lds_layout(int dim, int tb, int pn, int ent, int red = 48) {
  LdsLayout L;
  L.off_c = 0;
  L.off_x = al32(pn * dim * tb);
  L.off_res = L.off_x + al32(pn * dim * tb);
  L.off_red = L.off_res + al32(ent * tb);
  L.total = L.off_red + red * 8;
  return L;
}

// Four consecutive values of T as one 16-byte (float) or 32-byte (double)
// LDS access.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
  __device__ static type zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
  __device__ static float get(const type &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
  }
};
template <>
struct Vec4<double> {
  using type = double4;
  __device__ static type zero() { return make_double4(0.0, 0.0, 0.0, 0.0); }
  __device__ static double get(const type &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
  }
};

// Decode a 12-byte element record (pack_records): local node ids and the
// positions of the element's corner entries in the patch's node-sorted list.
__device__ __forceinline__ void decode_rec(uint3 r, int *l, int *q) {
  l[0] = int(r.x & 1023u);
  l[1] = int((r.x >> 10) & 1023u);
  l[2] = int((r.x >> 20) & 1023u);
  l[3] = int((r.x >> 30) | ((r.y & 255u) << 2));
  q[0] = int((r.y >> 8) & 16383u);
  q[1] = int((r.y >> 22) | ((r.z & 15u) << 10));
  q[2] = int((r.z >> 4) & 16383u);
  q[3] = int(r.z >> 18);
}

// Zero the LDS entry list: the padding entries of the node runs (never
// written by an element) must read as zero.
template <typename T, int PT>
__device__ __forceinline__ void clear_entries(T *s_res, int ent) {
  using V4 = typename Vec4<T>::type;
  V4 *r4 = reinterpret_cast<V4 *>(s_res);
  const V4 z = Vec4<T>::zero();
  for (int i = threadIdx.x; i < (ent >> 2); i += PT) r4[i] = z;
}

// Phases B and C of a patch workgroup, after the node data (s_c coordinates,
// s_x vector with constrained components zeroed) is staged and the entry list
// cleared.  B: the forces of this thread's elements, kept in registers.
// C, one component at a time: entries to their node-sorted LDS positions,
// then every node sums its contiguous run [b0, b1) in order (runs start and
// end on multiples of 4 entries: 16-byte reads, two in flight per step).
// acc receives this thread's node sums; with EN, en_acc += en_scale *
// u_e'K_e u_e for each of this thread's elements (DIAG false; DIAG: the
// element diagonals).
template <typename T, int DIM, int PT, bool DIAG, bool EN = true>
__device__ __forceinline__ void assemble(const T *s_c, const T *s_x, T *s_res, const uint3 *rec,
                                         const T *esc, int n_elem, T lam, T mu, int b0, int b1,
                                         T *acc, double &en_acc, double en_scale) {
  constexpr int NPE = DIM + 1, NDE = DIM * NPE;
  // Forces of corners 1..DIM per element (corner 0 of the force-free forms
  // is -(sum of the others) and is formed again in phase C; the diagonal
  // keeps all corners).  Few live registers here keep eight 256-thread
  // workgroups resident per CU in FP32.
  constexpr int NK = DIAG ? NDE : NDE - DIM;
  T F[EPT][NK];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int i = threadIdx.x + e * PT;
    if (i >= n_elem) continue;
    int l[4], qd[4];
    decode_rec(rec[e], l, qd);
    T X[NDE], U[NDE], Fe[NDE];
    // byte offset of the corner's node row, shared by s_c and s_x
#pragma unroll
    for (int n = 0; n < NPE; ++n) {
      const unsigned bo = unsigned(l[n]) * unsigned(DIM * sizeof(T));
      const T *xc = reinterpret_cast<const T *>(reinterpret_cast<const char *>(s_c) + bo);
      const T *xu = reinterpret_cast<const T *>(reinterpret_cast<const char *>(s_x) + bo);
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        X[n * DIM + c] = xc[c];
        if constexpr (!DIAG) U[n * DIM + c] = xu[c];
      }
    }
    if constexpr (DIAG) {
      elem::elem_diag<T, DIM>(X, lam, mu, esc[e], Fe);
    } else {
      const T en = elem::elem_force_free<T, DIM>(X, U, lam, mu, esc[e], Fe);
      if constexpr (EN) en_acc += en_scale * double(en);
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) F[e][k] = Fe[NDE - NK + k];
  }
  int q[EPT][4];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    int l[4];
    decode_rec(rec[e], l, q[e]);
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
#pragma unroll
    for (int e = 0; e < EPT; ++e)
      if (threadIdx.x + e * PT < n_elem) {
        if constexpr (DIAG) {
#pragma unroll
          for (int n = 0; n < NPE; ++n) s_res[q[e][n]] = F[e][n * DIM + c];
        } else {
          // f_0 = -(f_1 + f_2 + ...), the order of elem_force_free
          T f0 = F[e][c];
#pragma unroll
          for (int n = 2; n < NPE; ++n) f0 += F[e][(n - 1) * DIM + c];
          s_res[q[e][0]] = -f0;
#pragma unroll
          for (int n = 1; n < NPE; ++n) s_res[q[e][n]] = F[e][(n - 1) * DIM + c];
        }
      }
    __syncthreads();
    using V4 = typename Vec4<T>::type;
    const V4 *r4 = reinterpret_cast<const V4 *>(s_res);
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
    const V4 *rp = r4 + (b0 >> 2);
    const int n4 = (b1 - b0) >> 2;
    for (int m = n4 >> 1; m > 0; --m, rp += 2) {
      const V4 u = rp[0], v = rp[1];
      s0 += Vec4<T>::get(u, 0) + Vec4<T>::get(v, 0);
      s1 += Vec4<T>::get(u, 1) + Vec4<T>::get(v, 1);
      s2 += Vec4<T>::get(u, 2) + Vec4<T>::get(v, 2);
      s3 += Vec4<T>::get(u, 3) + Vec4<T>::get(v, 3);
    }
    if (n4 & 1) {
      const V4 u = rp[0];
      s0 += Vec4<T>::get(u, 0);
      s1 += Vec4<T>::get(u, 1);
      s2 += Vec4<T>::get(u, 2);
      s3 += Vec4<T>::get(u, 3);
    }
    acc[c] = (s0 + s1) + (s2 + s3);
    if (c + 1 < DIM) __syncthreads();
  }
}

// acc += the slots [j0, j1) of an interface node, in slot order, two loads
// in flight per step.
template <typename T, int DIM>
__device__ __forceinline__ void add_slots(const T *partial, int j0, int j1, T *acc) {
  int j = j0;
  for (; j + 2 <= j1; j += 2) {
    T u[DIM], w[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      u[c] = partial[int64_t(j) * DIM + c];
      w[c] = partial[int64_t(j + 1) * DIM + c];
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = (acc[c] + u[c]) + w[c];
  }
  if (j < j1) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] += partial[int64_t(j) * DIM + c];
  }
}

}  // namespace patch
}  // namespace femgpu
