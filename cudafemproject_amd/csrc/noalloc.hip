// noalloc.hip -- fixed device scratch of femgpu_matvec_noalloc and the
// user-order gather / scatter kernels around its EbE matvec.
//
// femgpu_matvec_noalloc runs the standalone EbE matvec on caller-owned device
// vectors without allocating: its internal-order copies of x and y and the
// interface partial-sum slots live in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel: its device address, taken once
// per device with hipGetSymbolAddress (noalloc_scratch_address, called by
// femgpu_create with the device current), is what the launches receive.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"
#include "noalloc.h"

static __device__ __align__(256) unsigned char scratch[femgpu::NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
constexpr int GT = 256;

inline int blocks_for(int64_t n) {
  const int64_t b = (n + GT - 1) / GT;
  return int(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

// xi[i * dim + c] = x_user[perm[i] * dim + c] (internal node i, user node perm[i])
template <typename T>
__global__ __launch_bounds__(GT) void k_gather_user(int64_t n, int dim, const int32_t *perm,
                                                    const double *xu, T *xi) {
  const int64_t total = n * dim;
  for (int64_t k = int64_t(blockIdx.x) * GT + threadIdx.x; k < total;
       k += int64_t(gridDim.x) * GT) {
    const int64_t i = k / dim;
    const int c = int(k - i * dim);
    xi[k] = T(xu[int64_t(perm[i]) * dim + c]);
  }
}

// y_user[perm[i] * dim + c] = yi[i * dim + c] for the nodes the operator
// covers (internal ids < n_active), 0 for unreferenced nodes.
template <typename T>
__global__ __launch_bounds__(GT) void k_scatter_user(int64_t n, int64_t n_active, int dim,
                                                     const int32_t *perm, const T *yi,
                                                     double *yu) {
  const int64_t total = n * dim;
  for (int64_t k = int64_t(blockIdx.x) * GT + threadIdx.x; k < total;
       k += int64_t(gridDim.x) * GT) {
    const int64_t i = k / dim;
    const int c = int(k - i * dim);
    yu[int64_t(perm[i]) * dim + c] = i < n_active ? double(yi[k]) : 0.0;
  }
}

}  // namespace

namespace femgpu {

hipError_t noalloc_scratch_address(int device, void **ptr) {
  static std::mutex mu;
  static void *addr[MAX_DEVICES] = {};
  if (!ptr || device < 0 || device >= MAX_DEVICES) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

template <typename T>
void launch_gather_user(int64_t n, int dim, const int32_t *perm, const double *xu, T *xi,
                        hipStream_t s) {
  hipLaunchKernelGGL((k_gather_user<T>), dim3(blocks_for(n * dim)), dim3(GT), 0, s, n, dim, perm,
                     xu, xi);
}

template <typename T>
void launch_scatter_user(int64_t n, int64_t n_active, int dim, const int32_t *perm, const T *yi,
                         double *yu, hipStream_t s) {
  hipLaunchKernelGGL((k_scatter_user<T>), dim3(blocks_for(n * dim)), dim3(GT), 0, s, n, n_active,
                     dim, perm, yi, yu);
}

template void launch_gather_user<double>(int64_t, int, const int32_t *, const double *, double *,
                                         hipStream_t);
template void launch_gather_user<float>(int64_t, int, const int32_t *, const double *, float *,
                                        hipStream_t);
template void launch_scatter_user<double>(int64_t, int64_t, int, const int32_t *, const double *,
                                          double *, hipStream_t);
template void launch_scatter_user<float>(int64_t, int64_t, int, const int32_t *, const float *,
                                         double *, hipStream_t);

}  // namespace femgpu
