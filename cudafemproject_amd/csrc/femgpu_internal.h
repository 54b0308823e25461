// femgpu_internal.h -- data layout shared by the host planner and the HIP
// kernels of the EbE solve path.
//
// Layout in HBM (all device arrays are in *internal* order):
//  * Elements are sorted along a Morton curve of their centroids and cut into
//    patches of at most `emax` consecutive elements touching at most PN_MAX
//    distinct nodes.  One workgroup processes one patch.
//  * Nodes are renumbered: first the nodes *owned* by exactly one patch (all
//    incident elements inside it), patch by patch, so every patch owns one
//    contiguous run [own_start, own_start + n_own); then the *interface*
//    nodes touched by several patches, in order of first touch.
//  * Per element: its patch-local node ids packed as uint8 in one uint32
//    (npe <= 4, patch nodes <= 256).
//  * Per patch node (owned first, then interface): the end offset of its
//    run in the patch's node-sorted list of (element, corner) entries; per
//    element and corner (stride 4): the uint16 position of that entry in the
//    list, so an element writes its nodal forces straight to their sorted
//    place and each node's sum is a contiguous fixed-order run.
//  * Interface nodes: a partial-sum slot per (patch, node), slots of one node
//    contiguous in patch order, so the final sum is a fixed-order loop.
// Node vectors are AoS by DOF: v[node*dim + c].
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace femgpu {

constexpr int PN_MAX = 256;          // nodes per patch (uint8 local ids)
constexpr int EMAX_LIMIT = 1024;     // elements per patch upper bound

struct PatchInfo {                   // 32 bytes, one per workgroup
  int32_t elem_start;                // first element (internal order)
  int32_t n_elem;
  int32_t own_start;                 // first owned node (internal id)
  int32_t n_own;
  int32_t sh_start;                  // first entry in sh_gid / sh_slot
  int32_t n_sh;
  int32_t pnode_start;               // first entry in node_end
  int32_t pad;
};

struct MeshPlan {
  int dim = 0, npe = 0, nde = 0;
  int64_t nn = 0, ne = 0;
  int emax = 0;                      // max elements per patch used to build
  int pn_used = 0;                   // max nodes in any patch
  int emax_used = 0;                 // max elements in any patch
  int64_t n_owned = 0;               // internal nodes [0, n_owned) are owned
  int64_t n_iface = 0;               // interface nodes [n_owned, n_owned+n_iface)
  int64_t n_isolated = 0;            // nodes no element references (at the end)
  int64_t n_slots = 0;               // interface partial-sum slots
  std::vector<int32_t> elem_perm;    // internal element -> user element
  std::vector<int32_t> node_perm;    // internal node -> user node
  std::vector<int32_t> node_inv;     // user node -> internal node
  std::vector<PatchInfo> patches;
  std::vector<uint32_t> elem_lid;    // per internal element
  std::vector<int32_t> sh_gid;       // per patch interface node: internal id
  std::vector<int32_t> sh_slot;      // per patch interface node: slot
  std::vector<uint16_t> node_end;    // per patch node
  std::vector<uint16_t> epos;        // 4 per element: sorted entry position
  std::vector<int32_t> slot_start;   // n_iface + 1
  std::vector<int32_t> elems_int;    // ne x npe internal node ids, internal order
};

// Builds the plan.  Returns "" on success, else an error message.
std::string build_plan(int dim, int64_t nn, const double *nodes, int64_t ne,
                       const int32_t *elems, int emax, MeshPlan &plan);

}  // namespace femgpu
