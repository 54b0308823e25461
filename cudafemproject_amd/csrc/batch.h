// batch.h -- batched independent static solves: device layout shared by the
// host code (batch.cpp) and the persistent PCG kernel (batch.hip).
//
// A batch is many small meshes ("instances"), each with its own material and
// constraints and NR right-hand sides.  Every instance is planned on its own
// (build_plan: Morton-sorted element patches, owned-then-interface node
// numbering, partial-sum slots) and the per-instance plans are concatenated:
//  * global node id = node_off + instance-internal id; an instance's nodes
//    are [0, n_owned) owned, [n_owned, n_active) interface, then unreferenced;
//  * per element (internal order, from elem_off): its instance-internal node
//    ids and, per corner, the position of its force entry in the instance's
//    node-sorted entry list (entries of one patch node contiguous);
//  * per patch node (from pn_off): its entry run [beg, end) and its target,
//    an owned node (>= 0) or an interface slot (-1 - slot);
//  * slot_start holds n_iface + 1 entries per instance, from iface_off.
// Node vectors are [node][rhs][dim]: v[(g * NR + r) * DIM + c]; entries are
// [entry][rhs][dim] from elem_off * (dim + 1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "femgpu_internal.h"

struct femgpu_ctx;

namespace femgpu {

struct BInst {
  int32_t node_off;                // first global node of the instance
  int32_t n_nodes;                 // all nodes of the instance
  int32_t n_owned, n_active;       // [0, n_owned) owned, [n_owned, n_active) interface
  int32_t elem_off, n_elems;       // elements, internal order
  int32_t pn_off, n_pn;            // patch-node runs
  int32_t iface_off;               // first slot_start entry of the instance
  int32_t slot_off;                // first interface slot of the instance
  int32_t pad[2];
  double lam, mu;                  // Lame constants (2D: plane stress)
};

// Per (instance, rhs) result, device and host.
struct BStat {
  int32_t status;      // FEMGPU_OK, FEMGPU_E_MAXIT, FEMGPU_E_BREAKDOWN, FEMGPU_E_NONFINITE
  int32_t iterations;  // PCG iterations done
  int32_t replacements;
  int32_t pad;
  double rel_res;      // ||f - K u|| / ||f|| (FP64, last replacement)
};

template <typename T>
struct BatchArgs {
  const BInst *inst;
  const int32_t *order;      // workgroup -> instance
  const int32_t *elems;      // per element: dim+1 instance-internal node ids
  const int32_t *gpos;       // per element: dim+1 entry positions (instance-local)
  const int32_t *pn_beg;     // per patch node: entry run begin (instance-local)
  const int32_t *pn_end;     // per patch node: entry run end
  const int32_t *pn_tgt;     // per patch node: owned node (>= 0) or -1 - slot
  const int32_t *slot_start;
  const uint8_t *fixmask;
  const double *coords64;
  const T *coordsT;
  const double *f;           // right-hand sides
  double *x;                 // solution
  double *y64;               // FP64 matvec output
  double *slots64;           // FP64 interface slots
  double *ent;               // element force entries (sized for FP64)
  T *r, *p, *q;              // PCG vectors (inner precision)
  T *slotsT;                 // interface slots (inner precision)
  T *invD;                   // per node and component
  BStat *stat;               // n_inst x NR
  double tol;                // relative residual target
  double drop;               // reliable-update residual drop (0: off)
  int max_iter;
};

int launch_batch_pcg(int dim, int nr, bool mixed, const BatchArgs<float> &af,
                     const BatchArgs<double> &ad, int n_inst, hipStream_t s);

struct BatchState;
void batch_free(BatchState *b);

// accessors of the context (femgpu_api.cpp)
hipStream_t ctx_stream(femgpu_ctx *c);
int ctx_device(femgpu_ctx *c);
int ctx_fail(femgpu_ctx *c, int code, const char *msg);
BatchState *&ctx_batch(femgpu_ctx *c);

}  // namespace femgpu
