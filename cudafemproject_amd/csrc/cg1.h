// cg1.h -- launch interface of the single-reduction PCG (cg1.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace femgpu {

// Per-patch partial sums of one single-reduction launch.
constexpr int CG1_PARTS = 5;

// Device state of the single-reduction PCG.  [2]: ping-pong by launch
// parity; the first iteration launch (parity 0) reads index 0.
template <typename T>
struct Cg1State {
  T *x[2], *r[2], *p[2], *s[2];  // n_nodes x dim each
  T *w;                          // w = A u, owned nodes (interface nodes: slots)
  T *slot[2];                    // interface partial slots of w (n_slots x dim)
  // [2][CG1_PARTS][np]: per-patch r'u, u'A u, r'r, u's_prev, p_prev's_prev
  // (the last two in FP32 only)
  double *part;
};

hipError_t prepare_cg1_kernels();

// w = A u with u = M^-1 r[0] (owned nodes into w, interface nodes into
// slot[0]) and the partials part[0]; reads r[0] (FP32 also p[0] and s[0]).
template <typename T>
void launch_cg1_init(const DevMesh &dm, const Cg1State<T> &st, const T *invD, hipStream_t s);
// One iteration: reads index `parity` of the state, writes the other.
template <typename T>
void launch_cg1_iter(const DevMesh &dm, const Cg1State<T> &st, const T *invD, CgScalars *sc,
                     int parity, hipStream_t s);
// Scalars from part[0] after launch_cg1_init.  mode 0: fresh start
// (g0, thresholds from sc->kind / sc->eps, counters); 1: residual replacement
// that keeps the search direction p[0] (s[0] = A p[0]; FP32 state only), the
// previous scalars taken from parity q; 2: replacement with a restart (b = 0
// at the next iteration).  pause_drop > 0: the iterations stop with done = 5
// once the residual norm has dropped by that factor (the next residual
// replacement).
void launch_cg1_fin(CgScalars *sc, const double *part, int np, int mode, int q,
                    double pause_drop, hipStream_t s);

}  // namespace femgpu
