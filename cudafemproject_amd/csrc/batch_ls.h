// batch_ls.h -- lockstep schedule of the batched static solves: every PCG
// iteration of every instance runs in two grid-wide launches (patch step,
// node update) over the union of the instances' patch plans, with
// per-(instance, right-hand side) scalars, so each instance is spread over
// the whole GPU instead of one workgroup.  Kernels in batch_ls.hip, host
// side in batch.cpp.
//
// Layout: every instance is planned on its own (build_plan with the
// single-problem patch limits) and the plans are concatenated: patch
// records, element records and node runs as in the single-problem solver,
// node ids offset by the instance's first node (an instance's nodes:
// owned, interface, unreferenced), slots by its first slot; PatchInfo.pad
// holds the instance.  Node vectors are [node][rhs][dim].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "femgpu_internal.h"

namespace femgpu {

constexpr int BL_THREADS = 256;    // patch workgroup (= patch node limit)
constexpr int BL_UPD_NODES = 256;  // nodes per node-parallel block (BL_UPD_NODES x NR threads)

struct BLInst {
  int32_t node_off, n_owned, n_active, n_nodes;
  int32_t p0, p1;      // patches
  int32_t u0, u1;      // update blocks
  int32_t if0;         // first BLIface entry (interface nodes in order)
  int32_t pad[3];
  double lam, mu;      // Lame constants (2D: plane stress)
};

struct BLBlock {       // update block: nodes [n0, n1) of one instance
  int32_t inst, n0, n1, pad;
};

// Per (instance, rhs) solver state.  state: 0 iterating, 1 residual
// replacement requested, 2 residual just replaced, 3 finished.
struct BLState {
  double g[2], b[2];   // r'z and beta by launch parity
  double thr;          // tol^2 ||f||^2 (< 0: not yet known)
  double ff;           // ||f||^2 (FP64)
  double rr_last;      // FP64 r'r at the last replacement
  double rr;           // last r'r seen by the stop test (while paused: the one at the pause)
  int32_t it[2];
  int32_t state, status, iters, nrep;
  int32_t restart;     // set with state 2: restart the directions (p = z)
  int32_t stag;        // consecutive replacements without a decrease of the FP64 residual
};

template <typename T>
struct BLArgs {
  const PatchInfo *patches;
  const uint32_t *erec;
  const int4 *sh_info;       // per patch interface node: node, slot, node's slot range
  const uint16_t *node_end;
  const uint8_t *fixmask;
  const T *coords;
  const T *diagK;            // assembled diagonal (constrained rows)
  const BLInst *inst;
  const BLBlock *blocks;
  const int4 *iface;         // per interface node (instance order): node, first slot, end slot
  int np, p8, ent, pn, nblk;
  BLState *st;
  int max_iter;
  double tol2, drop2;   // drop2 = 0: replacements only to confirm convergence
};

hipError_t prepare_bl_kernels();

// diag(K) (FP64) into diag (referenced nodes; slots64 scratch); n_if: entries
// of a.iface
void bl_diag(int dim, const BLArgs<double> &a, int n_if, double *diag, double *slots64,
             hipStream_t s);
// invD = 1 / diag in the inner precision (0 where diag is 0)
template <typename T>
void bl_inv_diag(int64_t n, const double *diag, T *invD, hipStream_t s);
// per (instance, rhs): state 1 (replacement requested), counters reset
void bl_init_state(int n_state, BLState *st, hipStream_t s);
// One PCG iteration (patch step + node update) at launch parity `parity`;
// the solution increments go to xinc when given (inner precision, flushed
// into x64 by bl_flush), else straight into the FP64 solution x64.
template <typename T>
void bl_iteration(int dim, int nr, const BLArgs<T> &a, const T *invD, T *z, T *p, T *q, T *r,
                  T *slots, double *x64, T *xinc, double *ppq, double *pnode, int parity,
                  hipStream_t s);
// x64 += xinc, xinc = 0 (every node of an instance with an unfinished system)
template <typename T>
void bl_flush(int dim, int nr, const BLArgs<T> &a, T *xinc, double *x64, hipStream_t s);
// Residual replacement of the systems with state 1: y64 = K x64 (flagged
// instances), r = f - y64, z = invD r, partials for the next iteration; then
// state 1 -> 2.  count: unfinished systems.
template <typename T>
void bl_replace(int dim, int nr, const BLArgs<T> &a, const BLArgs<double> &a64, const double *f,
                const double *x64, double *y64, double *slots64, const T *invD, T *r, T *z,
                double *pnode, int n_state, int *count, hipStream_t s);
// Final FP64 residual of every system: r'r into the state's rr field.
void bl_final(int dim, int nr, const BLArgs<double> &a64, const double *f, const double *x64,
              double *y64, double *slots64, double *pnode, int n_inst, hipStream_t s);

}  // namespace femgpu
