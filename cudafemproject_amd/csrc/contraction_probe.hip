// contraction_probe.hip -- measurement kernels, built as libfemgpu_probe.so
// beside the solver library and not part of it.  They time the element
// contraction of the batched plate solves, F_e = K_e U_e for four load cases
// per linear triangle, in two forms that produce the same forces:
//
//  * k_contract_valu: one lane per element, forces on the fly: the geometry
//    once, then the four load cases through the solver's tri_force_free
//    (the kept form).
//  * k_contract_mfma: four lanes per element (16 elements per wave).  Lane r
//    forms rows r and r + 4 of K_e,
//      K[(a,i),(b,j)] = s (lam c_ai c_bj + mu c_bi c_aj + [i = j] mu c_a.c_b)
//    with the cofactor vectors c_a and s = 1 / (2 |det J|) of tri_force_free,
//    and v_mfma_f32_4x4x1f32 (16 blocks, one element per block) accumulates
//    F = K_e U_e over the six DOFs: lane r supplies row r of the A operand
//    and load case r of the B operand, and receives load case r of F.
//
// Both read the connectivity, the coordinates and the [node][case][2] load
// vectors from HBM.  Output mode 0 writes F as [element][case][6] (for the
// comparison of the two forms); mode 1 writes each load case's element
// energy u_case' F_case as [element][case], the force work of the batched
// PCG step beside its LDS scatter, so that the time is that of the
// contraction rather than of 96 bytes of stores per element.
#include <hip/hip_runtime.h>

#include "element.h"

namespace femgpu {
namespace probe {
namespace {

constexpr int NC = 4;     // load cases
constexpr int BT = 256;   // threads per block
typedef float v4f __attribute__((ext_vector_type(4)));

struct Args {
  int ne;
  const int *elems;      // 3 per element
  const float *coords;   // 2 per node
  const float *u;        // [node][case][2]
  float lam, mu;
  int mode;              // 0: forces, 1: energies
  float *f;              // mode 0: [element][case][6]; mode 1: [element][case]
};

__global__ __launch_bounds__(BT) void k_contract_valu(Args a) {
  const int e = blockIdx.x * BT + threadIdx.x;
  if (e >= a.ne) return;
  int n[3];
  float X[6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = a.elems[3 * e + i];
    X[2 * i] = a.coords[2 * n[i]];
    X[2 * i + 1] = a.coords[2 * n[i] + 1];
  }
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    float U[6], F[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      U[2 * i] = a.u[(n[i] * NC + v) * 2];
      U[2 * i + 1] = a.u[(n[i] * NC + v) * 2 + 1];
    }
    const float en = elem::tri_force_free<float>(X, U, a.lam, a.mu, 1.f, F);
    if (a.mode == 0) {
#pragma unroll
      for (int i = 0; i < 6; ++i) a.f[(e * NC + v) * 6 + i] = F[i];
    } else {
      a.f[e * NC + v] = en;
    }
  }
}

// Row R = 2 a + i of K_e (i = R & 1; ca the row node's cofactor vector).
__device__ __forceinline__ void k_row(float ca_x, float ca_y, int i, const float (&c)[3][2],
                                      float ls, float ms, float *kr) {
  const float ca_i = i ? ca_y : ca_x;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const float cb_i = i ? c[b][1] : c[b][0];
    const float dot = ms * (ca_x * c[b][0] + ca_y * c[b][1]);
    // j = 0, 1
    kr[2 * b] = ls * ca_i * c[b][0] + ms * cb_i * ca_x + (i == 0 ? dot : 0.f);
    kr[2 * b + 1] = ls * ca_i * c[b][1] + ms * cb_i * ca_y + (i == 1 ? dot : 0.f);
  }
}

__global__ __launch_bounds__(BT) void k_contract_mfma(Args a) {
  const int t = blockIdx.x * BT + threadIdx.x;
  const int r = threadIdx.x & 3;
  const int e_t = t >> 2;
  const bool valid = e_t < a.ne;
  // every lane takes part in the matrix instruction: lanes past the last
  // element repeat it and do not write
  const int e = valid ? e_t : a.ne - 1;
  int n[3];
  float X[6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = a.elems[3 * e + i];
    X[2 * i] = a.coords[2 * n[i]];
    X[2 * i + 1] = a.coords[2 * n[i] + 1];
  }
  // load case r of this element
  float ur[6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ur[2 * i] = a.u[(n[i] * NC + r) * 2];
    ur[2 * i + 1] = a.u[(n[i] * NC + r) * 2 + 1];
  }
  // geometry of tri_force_free
  const float e1x = X[2] - X[0], e1y = X[3] - X[1];
  const float e2x = X[4] - X[0], e2y = X[5] - X[1];
  float c[3][2];
  c[1][0] = e2y;
  c[1][1] = -e2x;
  c[2][0] = -e1y;
  c[2][1] = e1x;
  c[0][0] = -(c[1][0] + c[2][0]);
  c[0][1] = -(c[1][1] + c[2][1]);
  const float det = e1x * e2y - e1y * e2x;
  const float s = elem::recip<float>(2.f * fabsf(det));
  const float ls = a.lam * s, ms = a.mu * s;
  // rows r (node r >> 1) and r + 4 (node 2; lanes 0, 1 only)
  float k0[6], k1[6];
  const int i = r & 1;
  k_row(r < 2 ? c[0][0] : c[1][0], r < 2 ? c[0][1] : c[1][1], i, c, ls, ms, k0);
  k_row(c[2][0], c[2][1], i, c, ls, ms, k1);
  if (r >= 2) {
#pragma unroll
    for (int C = 0; C < 6; ++C) k1[C] = 0.f;
  }
  v4f d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int C = 0; C < 6; ++C) {
    d0 = __builtin_amdgcn_mfma_f32_4x4x1f32(k0[C], ur[C], d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_4x4x1f32(k1[C], ur[C], d1, 0, 0, 0);
  }
  if (valid) {
    if (a.mode == 0) {
      float *fo = a.f + (int64_t(e) * NC + r) * 6;
#pragma unroll
      for (int q = 0; q < 4; ++q) fo[q] = d0[q];
      fo[4] = d1[0];
      fo[5] = d1[1];
    } else {
      float en = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) en += ur[q] * d0[q];
      en += ur[4] * d1[0];
      en += ur[5] * d1[1];
      a.f[int64_t(e) * NC + r] = en;
    }
  }
}

// One wave: A = lane index, B = 100 + lane index, C = 0; D[lane][reg] out.
__global__ __launch_bounds__(64) void k_mfma_layout(float *out) {
  const int l = threadIdx.x;
  v4f d = {0.f, 0.f, 0.f, 0.f};
  d = __builtin_amdgcn_mfma_f32_4x4x1f32(float(l), float(100 + l), d, 0, 0, 0);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[l * 4 + q] = d[q];
}

}  // namespace
}  // namespace probe
}  // namespace femgpu

using namespace femgpu::probe;

// variant 0: on the fly (VALU), 1: K_e and the matrix instruction; mode 0:
// forces, 1: energies (see Args).  Launches the kernel `reps` times on
// `stream`.  Returns a hipError_t value (0: ok).
extern "C" int femgpu_probe_contraction(int variant, int mode, int ne, const int *elems,
                                        const float *coords, const float *u, float lam,
                                        float mu, float *f, int reps, void *stream) {
  if ((variant != 0 && variant != 1) || (mode != 0 && mode != 1) || ne <= 0 || !elems ||
      !coords || !u || !f || reps < 1)
    return int(hipErrorInvalidValue);
  const Args a{ne, elems, coords, u, lam, mu, mode, f};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t threads = variant == 0 ? int64_t(ne) : 4 * int64_t(ne);
  const int64_t blocks = (threads + BT - 1) / BT;
  if (blocks > 0x7fffffff) return int(hipErrorInvalidValue);
  for (int k = 0; k < reps; ++k) {
    if (variant == 0)
      hipLaunchKernelGGL(k_contract_valu, dim3(unsigned(blocks)), dim3(BT), 0, s, a);
    else
      hipLaunchKernelGGL(k_contract_mfma, dim3(unsigned(blocks)), dim3(BT), 0, s, a);
  }
  return int(hipGetLastError());
}

// out: 256 floats of device memory (64 lanes x 4 registers).
extern "C" int femgpu_probe_mfma_layout(float *out, void *stream) {
  if (!out) return int(hipErrorInvalidValue);
  hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), out);
  return int(hipGetLastError());
}
