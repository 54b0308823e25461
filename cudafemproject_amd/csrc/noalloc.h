// noalloc.h -- the fixed device scratch of femgpu_matvec_noalloc (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace femgpu {

// Bytes of the per-device scratch array: internal-order x and y (each
// rounded up to 256 bytes) plus the interface partial-sum slots of one
// matvec.  32 MiB holds the 64^3-hex tet cube (274,625 nodes, 315,800 slots)
// in FP64 (20.8 MB).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;

// Device address of this device's scratch array (the current device must be
// `device`); taken with hipGetSymbolAddress on the first call per device and
// cached.
hipError_t noalloc_scratch_address(int device, void **ptr);

template <typename T>
void launch_gather_user(int64_t n, int dim, const int32_t *perm, const double *xu, T *xi,
                        hipStream_t s);
template <typename T>
void launch_scatter_user(int64_t n, int64_t n_active, int dim, const int32_t *perm, const T *yi,
                         double *yu, hipStream_t s);

}  // namespace femgpu
