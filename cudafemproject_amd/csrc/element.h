// element.h -- device math shared by the EbE kernels (kernels.hip) and the
// batched solver (batch.hip): wave64 / block reductions and the on-the-fly
// element force of linear triangles (plane stress) and tetrahedra.
//
// The element force is the reference's K_e = B'DB * area * coeff
// (cpu_fem_utils.cpp:338-503) rewritten with the barycentric gradients
// g_a = c_a / det J (c_a: cofactor vectors), so that
//   f_a = (1 / (2|det J|) or 1 / (6|det J|)) * S(H) c_a,
//   H = sum_a du_a (x) c_a,  S(H) = lam tr(H) I + mu (H + H').
// Constrained DOFs follow applyConstraints (cpu_fem_utils.cpp:461-476):
// off-diagonal couplings of a fixed DOF are dropped and its diagonal kept.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace femgpu {
namespace elem {

// v from another lane by a DPP pattern (both 32-bit halves); rows outside
// ROW_MASK read 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Sum over the 64 lanes of a full wave, the same bits in every lane (fixed
// order).  DPP steps inside the VALU: lane pairs and quads (quad_perm), 8 and
// 16 lanes (half-row and row mirrors), then rows 1 and 3 add the last lane of
// the row before them (row_bcast:15) and rows 2 and 3 add lane 31
// (row_bcast:31), which leaves the total in lane 63, broadcast by readlane.
// No LDS traffic and no lane-address registers, unlike a shuffle butterfly.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);         // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);         // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);        // row_half_mirror
  v += dpp_f64<0x140>(v);        // row_mirror
  v += dpp_f64<0x142, 0xA>(v);   // row_bcast:15
  v += dpp_f64<0x143, 0xC>(v);   // row_bcast:31
  const long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(b), 63);
  const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(b >> 32), 63);
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Block-wide sum, same value returned in every thread; fixed order.
template <int BT = 256>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
#pragma unroll
  for (int i = 0; i < BT / 64; ++i) r += sh[i];
  return r;
}

// Workgroup barrier that orders LDS accesses only: global loads and stores
// issued before it stay in flight (each load is waited for at its first
// use).  For kernels whose data guarded by the barrier is in LDS and that
// keep loads for later work in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Two block-wide sums with one pair of barriers; sh holds 2 * BT / 64 doubles.
template <int BT>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[2 * w] = a;
    sh[2 * w + 1] = b;
  }
  __syncthreads();
  double ra = 0, rb = 0;
#pragma unroll
  for (int i = 0; i < BT / 64; ++i) {
    ra += sh[2 * i];
    rb += sh[2 * i + 1];
  }
  a = ra;
  b = rb;
}

// Three block-wide sums with one pair of barriers; sh holds 3 * BT / 64 doubles.
template <int BT>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[3 * w] = a;
    sh[3 * w + 1] = b;
    sh[3 * w + 2] = c;
  }
  __syncthreads();
  double ra = 0, rb = 0, rc = 0;
#pragma unroll
  for (int i = 0; i < BT / 64; ++i) {
    ra += sh[3 * i];
    rb += sh[3 * i + 1];
    rc += sh[3 * i + 2];
  }
  a = ra;
  b = rb;
  c = rc;
}

// Block-wide sums of N values that are already wave-uniform (each wave's
// total in every lane): one pair of barriers, sh holds N * BT / 64 doubles,
// the totals in every thread, summed over the waves in order.
template <int BT, int N>
__device__ __forceinline__ void waves_sum_n(double (&v)[N], double *sh) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) sh[N * w + i] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = 0.0;
#pragma unroll
  for (int k = 0; k < BT / 64; ++k) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += sh[N * k + i];
  }
}

// N block-wide sums with one pair of barriers; sh holds N * BT / 64 doubles.
template <int BT, int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double *sh) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
  waves_sum_n<BT, N>(v, sh);
}

// Product rounded on its own (never fused into a following add).
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }

// Jacobi-PCG search direction p' = M^-1 r + beta p (invD = M^-1), formed
// the same way wherever it is needed, so that every copy has the same bits:
// z = invD r rounded, then one fused multiply-add.
template <typename T>
__device__ __forceinline__ T pcg_dir(T invd, T r, T beta, T p) {
  return fma(beta, p, mul_rn(invd, r));
}

// ---------------------------------------------------------------------------
// Element kernels (registers only).  X, U: node-major (npe x dim).
// fix: bit (dim*a + c) set when that local DOF is constrained.
// Returns u'K u of the (constraint-modified) element matrix.

template <typename T>
__device__ __forceinline__ T tet_force(const T *X, const T *U, T lam, T mu, T es, unsigned fix,
                                       T *F) {
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
#pragma unroll
  for (int j = 0; j < 3; ++j) c[0][j] = -(c[1][j] + c[2][j] + c[3][j]);
  const T det = e1x * c[1][0] + e1y * c[1][1] + e1z * c[1][2];
  const T s = es / (T(6) * fabs(det));
  T um[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) um[i] = ((fix >> i) & 1u) ? T(0) : U[i];
  T H[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i][j] = T(0);
#pragma unroll
  for (int a = 1; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T du = um[3 * a + i] - um[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) H[i][j] += du * c[a][j];
    }
  const T tr = H[0][0] + H[1][1] + H[2][2];
  const T S00 = lam * tr + T(2) * mu * H[0][0];
  const T S11 = lam * tr + T(2) * mu * H[1][1];
  const T S22 = lam * tr + T(2) * mu * H[2][2];
  const T S01 = mu * (H[0][1] + H[1][0]);
  const T S02 = mu * (H[0][2] + H[2][0]);
  const T S12 = mu * (H[1][2] + H[2][1]);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    F[3 * a + 0] = s * (S00 * c[a][0] + S01 * c[a][1] + S02 * c[a][2]);
    F[3 * a + 1] = s * (S01 * c[a][0] + S11 * c[a][1] + S12 * c[a][2]);
    F[3 * a + 2] = s * (S02 * c[a][0] + S12 * c[a][1] + S22 * c[a][2]);
  }
  if (fix) {
#pragma unroll
    for (int i = 0; i < 12; ++i)
      if ((fix >> i) & 1u) {
        const int a = i / 3, cc = i % 3;
        const T n2 = c[a][0] * c[a][0] + c[a][1] * c[a][1] + c[a][2] * c[a][2];
        F[i] = s * ((lam + mu) * c[a][cc] * c[a][cc] + mu * n2) * U[i];
      }
  }
  T en = T(0);
#pragma unroll
  for (int i = 0; i < 12; ++i) en += U[i] * F[i];
  return en;
}

template <typename T>
__device__ __forceinline__ T tri_force(const T *X, const T *U, T lam, T mu, T es, unsigned fix,
                                       T *F) {
  const T e1x = X[2] - X[0], e1y = X[3] - X[1];
  const T e2x = X[4] - X[0], e2y = X[5] - X[1];
  T c[3][2];
  c[1][0] = e2y;
  c[1][1] = -e2x;
  c[2][0] = -e1y;
  c[2][1] = e1x;
  c[0][0] = -(c[1][0] + c[2][0]);
  c[0][1] = -(c[1][1] + c[2][1]);
  const T det = e1x * e2y - e1y * e2x;
  const T s = es / (T(2) * fabs(det));
  T um[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) um[i] = ((fix >> i) & 1u) ? T(0) : U[i];
  T H00 = T(0), H01 = T(0), H10 = T(0), H11 = T(0);
#pragma unroll
  for (int a = 1; a < 3; ++a) {
    const T dx = um[2 * a] - um[0], dy = um[2 * a + 1] - um[1];
    H00 += dx * c[a][0];
    H01 += dx * c[a][1];
    H10 += dy * c[a][0];
    H11 += dy * c[a][1];
  }
  const T tr = H00 + H11;
  const T S00 = lam * tr + T(2) * mu * H00;
  const T S11 = lam * tr + T(2) * mu * H11;
  const T S01 = mu * (H01 + H10);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    F[2 * a + 0] = s * (S00 * c[a][0] + S01 * c[a][1]);
    F[2 * a + 1] = s * (S01 * c[a][0] + S11 * c[a][1]);
  }
  if (fix) {
#pragma unroll
    for (int i = 0; i < 6; ++i)
      if ((fix >> i) & 1u) {
        const int a = i / 2, cc = i % 2;
        const T n2 = c[a][0] * c[a][0] + c[a][1] * c[a][1];
        F[i] = s * ((lam + mu) * c[a][cc] * c[a][cc] + mu * n2) * U[i];
      }
  }
  T en = T(0);
#pragma unroll
  for (int i = 0; i < 6; ++i) en += U[i] * F[i];
  return en;
}

// 1/x: the hardware reciprocal (1 ulp) in FP32, a correctly rounded
// division in FP64.
template <typename T>
__device__ __forceinline__ T recip(T x) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_rcpf(x);
  else
    return T(1) / x;
}

// Element force of an unconstrained element block (the patch kernels zero
// constrained components of u per node and set the constrained rows from
// the assembled diagonal): f_a = S(H)' c_a with S(H) scaled by
// es / (6|det J|) once, and f_0 = -(f_1 + f_2 + f_3) since c_0 = -sum c_a.
// Returns u'K u of the element.
template <typename T>
__device__ __forceinline__ T tet_force_free(const T *X, const T *U, T lam, T mu, T es, T *F) {
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
  const T s = es * recip<T>(T(6) * fabs(det));
  T H[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T d1 = U[3 + i] - U[i], d2 = U[6 + i] - U[i], d3 = U[9 + i] - U[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i][j] = d1 * c[1][j] + d2 * c[2][j] + d3 * c[3][j];
  }
  const T ls = lam * s, ms = mu * s;
  const T tr = ls * (H[0][0] + H[1][1] + H[2][2]);
  const T S00 = tr + T(2) * ms * H[0][0];
  const T S11 = tr + T(2) * ms * H[1][1];
  const T S22 = tr + T(2) * ms * H[2][2];
  const T S01 = ms * (H[0][1] + H[1][0]);
  const T S02 = ms * (H[0][2] + H[2][0]);
  const T S12 = ms * (H[1][2] + H[2][1]);
#pragma unroll
  for (int a = 1; a < 4; ++a) {
    F[3 * a + 0] = S00 * c[a][0] + S01 * c[a][1] + S02 * c[a][2];
    F[3 * a + 1] = S01 * c[a][0] + S11 * c[a][1] + S12 * c[a][2];
    F[3 * a + 2] = S02 * c[a][0] + S12 * c[a][1] + S22 * c[a][2];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) F[k] = -(F[3 + k] + F[6 + k] + F[9 + k]);
  T en = T(0);
#pragma unroll
  for (int i = 0; i < 12; ++i) en += U[i] * F[i];
  return en;
}

template <typename T>
__device__ __forceinline__ T tri_force_free(const T *X, const T *U, T lam, T mu, T es, T *F) {
  const T e1x = X[2] - X[0], e1y = X[3] - X[1];
  const T e2x = X[4] - X[0], e2y = X[5] - X[1];
  // c_1 = (e2y, -e2x), c_2 = (-e1y, e1x)
  const T det = e1x * e2y - e1y * e2x;
  const T s = es * recip<T>(T(2) * fabs(det));
  const T dx1 = U[2] - U[0], dy1 = U[3] - U[1], dx2 = U[4] - U[0], dy2 = U[5] - U[1];
  const T H00 = dx1 * e2y - dx2 * e1y, H01 = -dx1 * e2x + dx2 * e1x;
  const T H10 = dy1 * e2y - dy2 * e1y, H11 = -dy1 * e2x + dy2 * e1x;
  const T ls = lam * s, ms = mu * s;
  const T tr = ls * (H00 + H11);
  const T S00 = tr + T(2) * ms * H00;
  const T S11 = tr + T(2) * ms * H11;
  const T S01 = ms * (H01 + H10);
  F[2] = S00 * e2y - S01 * e2x;
  F[3] = S01 * e2y - S11 * e2x;
  F[4] = -S00 * e1y + S01 * e1x;
  F[5] = -S01 * e1y + S11 * e1x;
  F[0] = -(F[2] + F[4]);
  F[1] = -(F[3] + F[5]);
  T en = T(0);
#pragma unroll
  for (int i = 0; i < 6; ++i) en += U[i] * F[i];
  return en;
}

template <typename T, int DIM>
__device__ __forceinline__ T elem_force_free(const T *X, const T *U, T lam, T mu, T es, T *F) {
  if constexpr (DIM == 3)
    return tet_force_free<T>(X, U, lam, mu, es, F);
  else
    return tri_force_free<T>(X, U, lam, mu, es, F);
}

template <typename T, int DIM>
__device__ __forceinline__ T elem_force(const T *X, const T *U, T lam, T mu, T es, unsigned fix,
                                        T *F) {
  if constexpr (DIM == 3)
    return tet_force<T>(X, U, lam, mu, es, fix, F);
  else
    return tri_force<T>(X, U, lam, mu, es, fix, F);
}

// Diagonal entries K_ii = s((lam+mu) c_ai^2 + mu |c_a|^2).
template <typename T, int DIM>
__device__ __forceinline__ void elem_diag(const T *X, T lam, T mu, T es, T *F) {
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
#pragma unroll
    for (int j = 0; j < 3; ++j) c[0][j] = -(c[1][j] + c[2][j] + c[3][j]);
    const T det = e1x * c[1][0] + e1y * c[1][1] + e1z * c[1][2];
    const T s = es / (T(6) * fabs(det));
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const T n2 = c[a][0] * c[a][0] + c[a][1] * c[a][1] + c[a][2] * c[a][2];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) F[3 * a + cc] = s * ((lam + mu) * c[a][cc] * c[a][cc] + mu * n2);
    }
  } else {
    const T e1x = X[2] - X[0], e1y = X[3] - X[1];
    const T e2x = X[4] - X[0], e2y = X[5] - X[1];
    T c[3][2];
    c[1][0] = e2y;
    c[1][1] = -e2x;
    c[2][0] = -e1y;
    c[2][1] = e1x;
    c[0][0] = -(c[1][0] + c[2][0]);
    c[0][1] = -(c[1][1] + c[2][1]);
    const T det = e1x * e2y - e1y * e2x;
    const T s = es / (T(2) * fabs(det));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T n2 = c[a][0] * c[a][0] + c[a][1] * c[a][1];
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) F[2 * a + cc] = s * ((lam + mu) * c[a][cc] * c[a][cc] + mu * n2);
    }
  }
}


}  // namespace elem
}  // namespace femgpu
