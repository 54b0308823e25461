// femgpu_internal.h -- data layout shared by the host planner and the HIP
// kernels of the EbE solve path.
//
// Layout in HBM (all device arrays are in *internal* order):
//  * Elements are sorted along a Morton curve of their centroids and cut into
//    patches of at most `emax` consecutive elements touching at most `pn_max`
//    distinct nodes.  One workgroup processes one patch.
//  * Nodes are renumbered: first the nodes *owned* by exactly one patch (all
//    incident elements inside it), patch by patch, so every patch owns one
//    contiguous run [own_start, own_start + n_own); then the *interface*
//    nodes touched by several patches, in order of first touch.
//  * Per element: its patch-local node ids (< 1024) and the positions of its
//    corner entries (< 16384), packed into one 12-byte record on the device
//    (pack_records).
//  * Per patch node (owned first, then interface): the end offset of its
//    run in the patch's node-sorted list of (element, corner) entries; per
//    element and corner (stride 4): the uint16 position of that entry in the
//    list, so an element writes its nodal forces straight to their sorted
//    place and each node's sum is a contiguous fixed-order run.  With
//    `pad_runs` every run starts at a multiple of 4 entries and ends on one
//    (the padding entries are zero in the kernels), so runs are read as
//    16-byte vectors.
//  * Interface nodes: a partial-sum slot per (patch, node), slots of one node
//    contiguous in patch order, so the final sum is a fixed-order loop.
// Node vectors are AoS by DOF: v[node*dim + c].
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace femgpu {

constexpr int PN_LIMIT = 1024;       // nodes per patch (10-bit local ids)
constexpr int EMAX_LIMIT = 3072;     // elements per patch (14-bit padded entry positions)

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
  int ent_used = 0;                  // max entry-list length of a patch (with run padding)
  bool padded = false;               // node runs padded to multiples of 4 entries
  int64_t n_owned = 0;               // internal nodes [0, n_owned) are owned
  int64_t n_iface = 0;               // interface nodes [n_owned, n_owned+n_iface)
  int64_t n_isolated = 0;            // nodes no element references (at the end)
  int64_t n_slots = 0;               // interface partial-sum slots
  std::vector<int32_t> elem_perm;    // internal element -> user element
  std::vector<int32_t> node_perm;    // internal node -> user node
  std::vector<int32_t> node_inv;     // user node -> internal node
  std::vector<PatchInfo> patches;
  std::vector<uint16_t> elem_lid;    // 4 per internal element: patch-local node ids
  std::vector<int32_t> sh_gid;       // per patch interface node: internal id
  std::vector<int32_t> sh_slot;      // per patch interface node: slot
  std::vector<uint16_t> node_end;    // per patch node
  std::vector<uint16_t> epos;        // 4 per element: sorted entry position
  std::vector<int32_t> slot_start;   // n_iface + 1
  std::vector<int32_t> elems_int;    // ne x npe internal node ids, internal order
};

// Builds the plan.  Returns "" on success, else an error message.
std::string build_plan(int dim, int64_t nn, const double *nodes, int64_t ne,
                       const int32_t *elems, int emax, int pn_max, bool pad_runs,
                       MeshPlan &plan);

// Device element records, 3 x uint32 per element: local ids l0..l3 (10 bits
// each) and entry positions p0..p3 (14 bits each),
//   w0 = l0 | l1 << 10 | l2 << 20 | (l3 & 3) << 30
//   w1 = l3 >> 2 | p0 << 8 | (p1 & 1023) << 22
//   w2 = p1 >> 10 | p2 << 4 | p3 << 18
void pack_records(const MeshPlan &plan, std::vector<uint32_t> &rec);

// The records (pack_records) at a fixed stride per patch, patch p's at
// p * emax_used with zero padding, for kernels that address them without
// the PatchInfo.  When every patch starts at p * emax_used, rec itself has
// that layout: it gets one patch of zero padding at the end and the function
// returns true; otherwise rec_s receives the strided copy (false).
bool stride_records(const MeshPlan &plan, std::vector<uint32_t> &rec,
                    std::vector<uint32_t> &rec_s);

// Per-patch node tables at a fixed stride (entry p * stride + k for node k
// of patch p, stride >= the patch's node count), so that a workgroup reads
// them without first reading its PatchInfo: pgid the internal node id,
// pslot for an interface node its slot with bit 31 set when the slot is the
// node's first (0 for owned nodes), pend the end of the node's run in the
// patch's entry list (node_end).  Entries beyond a patch's nodes are zero.
void pack_patch_nodes(const MeshPlan &plan, int stride, std::vector<int32_t> &pgid,
                      std::vector<int32_t> &pslot, std::vector<uint16_t> &pend);

}  // namespace femgpu
