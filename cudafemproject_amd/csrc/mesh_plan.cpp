// mesh_plan.cpp -- host-side preparation of the device mesh layout.
//
// Replaces the reference's per-iteration sort-based assembly set-up
// (genMask + sort + reduce_by_key, fem_utils/src/cpu_fem_utils.cpp:182-219 and
// fem_utils/src/cu/cuda_fem_utils.cpp:178-216): instead of sorting all
// element-local entries by DOF every PCG iteration, the element -> patch and
// node -> owner maps are built once here, and the kernels assemble in LDS and
// through fixed-order interface sums.  See femgpu_internal.h for the layout.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "femgpu_internal.h"

namespace femgpu {

namespace {

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

uint64_t spread2(uint64_t v) {  // 31 bits -> every second bit
  v &= 0x7fffffffULL;
  v = (v | v << 16) & 0x0000ffff0000ffffULL;
  v = (v | v << 8) & 0x00ff00ff00ff00ffULL;
  v = (v | v << 4) & 0x0f0f0f0f0f0f0f0fULL;
  v = (v | v << 2) & 0x3333333333333333ULL;
  v = (v | v << 1) & 0x5555555555555555ULL;
  return v;
}

// Entry positions within the padded node runs of one patch.  The patch
// kernels write the entries with one instruction per (element slot e,
// corner a): thread t writes the entry of its element t + e * pn_max.  Any
// position of a node's run may hold any of the node's entries (the padding
// positions are the unused ones, which stay zero), so the positions are
// chosen to spread the lanes of each write instruction over the LDS banks
// (4-byte words: 64 banks per wave, 32 per half wave): a greedy pass in
// the kernel's write order, then two sweeps of pairwise swaps inside each run
// that lower sum_W (2 * sum_b n64(W,b)^2 + sum_h,b n32(W,h,b)^2), n = lanes of
// write instruction W on a bank.  The order of the entries inside a run is
// the summation order of the node's sum; it is fixed by the plan.
struct BankScratch {
  std::vector<int32_t> occ, rstart, wid;
  std::vector<uint8_t> half;
  std::vector<int16_t> c64, c32;
};

std::string spread_entry_banks(MeshPlan &plan, const PatchInfo &P, int npe, int pn_max,
                               int32_t nloc, int32_t run, BankScratch &B) {
  const int32_t nwave = (pn_max + 63) / 64;
  const int32_t slots_e = (P.n_elem + pn_max - 1) / pn_max;
  const int32_t n_w = slots_e * npe * nwave;
  B.occ.assign(size_t(run), -1);
  B.rstart.resize(nloc);
  for (int32_t k = 0; k < nloc; ++k) B.rstart[k] = k ? int32_t(plan.node_end[P.pnode_start + k - 1]) : 0;
  B.wid.assign(size_t(P.n_elem) * 4, 0);
  B.half.assign(size_t(P.n_elem) * 4, 0);
  B.c64.assign(size_t(n_w) * 64, 0);
  B.c32.assign(size_t(n_w) * 64, 0);
  uint16_t *epos = &plan.epos[size_t(P.elem_start) * 4];
  const uint16_t *lid = &plan.elem_lid[size_t(P.elem_start) * 4];
  for (int32_t e = 0; e < slots_e; ++e)
    for (int a = 0; a < npe; ++a)
      for (int32_t t = 0; t < pn_max; ++t) {
        const int32_t i = e * pn_max + t;
        if (i >= P.n_elem) break;
        const int32_t j = i * 4 + a, l = lid[j];
        const int32_t W = (e * npe + a) * nwave + t / 64, h = (t & 63) >> 5;
        int16_t *u64 = &B.c64[size_t(W) * 64], *u32 = &B.c32[size_t(W) * 64 + h * 32];
        const int32_t r1 = int32_t(plan.node_end[P.pnode_start + l]);
        int32_t best = -1, best_cost = INT32_MAX;
        for (int32_t p = B.rstart[l]; p < r1; ++p) {
          if (B.occ[p] >= 0) continue;
          const int32_t cost = 2 * u64[p & 63] + u32[p & 31];
          if (cost < best_cost) {
            best_cost = cost;
            best = p;
            if (cost == 0) break;
          }
        }
        if (best < 0) return "internal error: node run full";
        B.occ[best] = j;
        B.wid[j] = W;
        B.half[j] = uint8_t(h);
        ++u64[best & 63];
        ++u32[best & 31];
      }
  // cost change of moving entry j from position p to q (applied)
  auto move = [&](int32_t j, int32_t p, int32_t q) {
    int16_t *u64 = &B.c64[size_t(B.wid[j]) * 64], *u32 = &B.c32[size_t(B.wid[j]) * 64 + B.half[j] * 32];
    int32_t d = 0;
    if ((p ^ q) & 63) {
      d += 4 * (u64[q & 63] - u64[p & 63] + 1);
      --u64[p & 63];
      ++u64[q & 63];
    }
    if ((p ^ q) & 31) {
      d += 2 * (u32[q & 31] - u32[p & 31] + 1);
      --u32[p & 31];
      ++u32[q & 31];
    }
    return d;
  };
  for (int sweep = 0; sweep < 2; ++sweep) {
    int32_t improved = 0;
    for (int32_t k = 0; k < nloc; ++k) {
      const int32_t r0 = B.rstart[k], r1 = int32_t(plan.node_end[P.pnode_start + k]);
      for (int32_t p = r0; p < r1; ++p)
        for (int32_t q = p + 1; q < r1; ++q) {
          const int32_t x = B.occ[p], y = B.occ[q];
          if (x < 0 && y < 0) continue;
          int32_t d = 0;
          if (x >= 0) d += move(x, p, q);
          if (y >= 0) d += move(y, q, p);
          if (d < 0) {
            B.occ[p] = y;
            B.occ[q] = x;
            ++improved;
          } else {
            if (y >= 0) move(y, p, q);
            if (x >= 0) move(x, q, p);
          }
        }
    }
    if (!improved) break;
  }
  for (int32_t p = 0; p < run; ++p)
    if (B.occ[p] >= 0) epos[B.occ[p]] = uint16_t(p);
  return "";
}

}  // namespace

std::string build_plan(int dim, int64_t nn, const double *nodes, int64_t ne,
                       const int32_t *elems, int emax, int pn_max, bool pad_runs,
                       MeshPlan &plan) {
  if (dim != 2 && dim != 3) return "dim must be 2 or 3";
  if (nn <= 0 || ne <= 0) return "empty mesh";
  if (nn >= (int64_t(1) << 31) / 3 || ne >= (int64_t(1) << 31) / 12)
    return "mesh too large for int32 DOF indexing";
  if (emax < 16 || emax > EMAX_LIMIT) return "elements per patch out of range";
  if (pn_max < 16 || pn_max > PN_LIMIT) return "nodes per patch out of range";
  const int npe = dim + 1;
  plan = MeshPlan();
  plan.dim = dim;
  plan.npe = npe;
  plan.nde = dim * npe;
  plan.nn = nn;
  plan.ne = ne;
  plan.emax = emax;
  plan.padded = pad_runs;

  for (int64_t e = 0; e < ne; ++e) {
    for (int a = 0; a < npe; ++a) {
      int32_t v = elems[e * npe + a];
      if (v < 0 || v >= nn) return "element node index out of range";
      for (int b = 0; b < a; ++b)
        if (elems[e * npe + b] == v) return "element with a repeated node";
    }
  }

  // 1. Morton order of element centroids.
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int c = 0; c < dim; ++c) lo[c] = hi[c] = nodes[c];
  for (int64_t i = 0; i < nn; ++i)
    for (int c = 0; c < dim; ++c) {
      double v = nodes[i * dim + c];
      if (!std::isfinite(v)) return "non-finite node coordinate";
      lo[c] = std::min(lo[c], v);
      hi[c] = std::max(hi[c], v);
    }
  const int bits = dim == 3 ? 21 : 31;
  const double qmax = double((uint64_t(1) << bits) - 1);
  std::vector<std::pair<uint64_t, int32_t>> key(ne);
  for (int64_t e = 0; e < ne; ++e) {
    uint64_t q[3] = {0, 0, 0};
    for (int c = 0; c < dim; ++c) {
      double s = 0;
      for (int a = 0; a < npe; ++a) s += nodes[int64_t(elems[e * npe + a]) * dim + c];
      s /= npe;
      double span = hi[c] - lo[c];
      double t = span > 0 ? (s - lo[c]) / span : 0.0;
      q[c] = uint64_t(std::min(qmax, std::max(0.0, t * qmax)));
    }
    uint64_t k = dim == 3 ? (spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2)
                          : (spread2(q[0]) | spread2(q[1]) << 1);
    key[e] = {k, int32_t(e)};
  }
  std::sort(key.begin(), key.end());
  plan.elem_perm.resize(ne);
  for (int64_t i = 0; i < ne; ++i) plan.elem_perm[i] = key[i].second;
  key.clear();
  key.shrink_to_fit();

  // 2. Greedy patches over the sorted elements.
  std::vector<int32_t> stamp(nn, -1);
  std::vector<int32_t> pstart;  // element start per patch
  {
    int32_t pid = 0, cur_nodes = 0, cur_elems = 0;
    pstart.push_back(0);
    for (int64_t i = 0; i < ne; ++i) {
      const int32_t *en = &elems[int64_t(plan.elem_perm[i]) * npe];
      int add = 0;
      for (int a = 0; a < npe; ++a) add += stamp[en[a]] != pid;
      if (cur_elems > 0 && (cur_elems + 1 > emax || cur_nodes + add > pn_max)) {
        ++pid;
        pstart.push_back(int32_t(i));
        cur_nodes = 0;
        cur_elems = 0;
        add = npe;
      }
      for (int a = 0; a < npe; ++a)
        if (stamp[en[a]] != pid) {
          stamp[en[a]] = pid;
          ++cur_nodes;
        }
      ++cur_elems;
    }
    pstart.push_back(int32_t(ne));
  }
  const int32_t np = int32_t(pstart.size()) - 1;

  // 3. Touch counts: how many patches reference each node.
  std::vector<int32_t> touches(nn, 0);
  std::fill(stamp.begin(), stamp.end(), -1);
  for (int32_t p = 0; p < np; ++p)
    for (int32_t i = pstart[p]; i < pstart[p + 1]; ++i) {
      const int32_t *en = &elems[int64_t(plan.elem_perm[i]) * npe];
      for (int a = 0; a < npe; ++a)
        if (stamp[en[a]] != p) {
          stamp[en[a]] = p;
          ++touches[en[a]];
        }
    }

  // 4. Internal node numbering: owned runs per patch, then interface, then isolated.
  std::vector<int32_t> inv(nn, -1);
  int32_t next = 0;
  for (int32_t p = 0; p < np; ++p)
    for (int32_t i = pstart[p]; i < pstart[p + 1]; ++i) {
      const int32_t *en = &elems[int64_t(plan.elem_perm[i]) * npe];
      for (int a = 0; a < npe; ++a)
        if (touches[en[a]] == 1 && inv[en[a]] < 0) inv[en[a]] = next++;
    }
  plan.n_owned = next;
  for (int32_t p = 0; p < np; ++p)
    for (int32_t i = pstart[p]; i < pstart[p + 1]; ++i) {
      const int32_t *en = &elems[int64_t(plan.elem_perm[i]) * npe];
      for (int a = 0; a < npe; ++a)
        if (touches[en[a]] > 1 && inv[en[a]] < 0) inv[en[a]] = next++;
    }
  plan.n_iface = next - plan.n_owned;
  for (int64_t v = 0; v < nn; ++v)
    if (inv[v] < 0) inv[v] = next++;
  plan.n_isolated = nn - plan.n_owned - plan.n_iface;
  plan.node_inv = inv;
  plan.node_perm.assign(nn, 0);
  for (int64_t v = 0; v < nn; ++v) plan.node_perm[inv[v]] = int32_t(v);

  // 5. Interface slots (contiguous per node, patch order).
  plan.slot_start.assign(plan.n_iface + 1, 0);
  for (int64_t j = 0; j < plan.n_iface; ++j)
    plan.slot_start[j + 1] = plan.slot_start[j] + touches[plan.node_perm[plan.n_owned + j]];
  plan.n_slots = plan.slot_start[plan.n_iface];
  std::vector<int32_t> slot_next(plan.slot_start.begin(), plan.slot_start.end() - 1);

  // 6. Per-patch local numbering, packed ids and node-sorted entry positions.
  plan.patches.resize(np);
  plan.elem_lid.assign(size_t(ne) * 4, 0);
  plan.epos.assign(size_t(ne) * 4, 0);
  plan.elems_int.assign(size_t(ne) * npe, 0);
  std::vector<int32_t> local(nn, -1);
  std::vector<int32_t> loc_nodes;
  std::vector<int32_t> cnt;
  BankScratch bank;
  for (int32_t p = 0; p < np; ++p) {
    PatchInfo &P = plan.patches[p];
    P.elem_start = pstart[p];
    P.n_elem = pstart[p + 1] - pstart[p];
    P.sh_start = int32_t(plan.sh_gid.size());
    P.pnode_start = int32_t(plan.node_end.size());
    P.pad = 0;
    // owned nodes: internal ids of this patch are one contiguous run
    int32_t omin = INT32_MAX, omax = -1, n_own = 0;
    loc_nodes.clear();
    for (int32_t i = pstart[p]; i < pstart[p + 1]; ++i) {
      const int32_t *en = &elems[int64_t(plan.elem_perm[i]) * npe];
      for (int a = 0; a < npe; ++a) {
        int32_t v = en[a];
        if (local[v] != -1) continue;
        local[v] = -2;  // seen
        loc_nodes.push_back(v);
        if (touches[v] == 1) {
          omin = std::min(omin, inv[v]);
          omax = std::max(omax, inv[v]);
          ++n_own;
        }
      }
    }
    if (n_own > 0 && omax - omin + 1 != n_own) return "internal error: owned run not contiguous";
    P.own_start = n_own > 0 ? omin : 0;
    P.n_own = n_own;
    int32_t n_sh = 0;
    for (int32_t v : loc_nodes) {
      if (touches[v] == 1) {
        local[v] = inv[v] - P.own_start;
      } else {
        local[v] = n_own + n_sh++;
        plan.sh_gid.push_back(inv[v]);
        int64_t j = inv[v] - plan.n_owned;
        plan.sh_slot.push_back(slot_next[j]++);
      }
    }
    P.n_sh = n_sh;
    const int32_t nloc = n_own + n_sh;
    if (nloc > pn_max) return "internal error: patch node overflow";
    plan.pn_used = std::max(plan.pn_used, int(nloc));
    plan.emax_used = std::max(plan.emax_used, int(P.n_elem));
    // packed ids + counts for the incidence lists
    cnt.assign(nloc, 0);
    for (int32_t i = pstart[p]; i < pstart[p + 1]; ++i) {
      const int32_t *en = &elems[int64_t(plan.elem_perm[i]) * npe];
      for (int a = 0; a < npe; ++a) {
        int32_t l = local[en[a]];
        plan.elem_lid[size_t(i) * 4 + a] = uint16_t(l);
        ++cnt[l];
        plan.elems_int[size_t(i) * npe + a] = inv[en[a]];
      }
    }
    // end offsets, then each (element, corner) entry's position in the
    // node-sorted order (within a node: by element, then corner)
    std::vector<int32_t> pos(nloc);
    int32_t run = 0;
    for (int32_t k = 0; k < nloc; ++k) {
      pos[k] = run;
      run += cnt[k];
      if (pad_runs) run = (run + 3) & ~3;
      plan.node_end.push_back(uint16_t(run));
    }
    if (run >= 16384) return "internal error: entry list above 14-bit positions";
    plan.ent_used = std::max(plan.ent_used, int(run));
    if (!pad_runs) {
      for (int32_t i = 0; i < P.n_elem; ++i)
        for (int a = 0; a < npe; ++a) {
          int32_t l = plan.elem_lid[size_t(P.elem_start + i) * 4 + a];
          plan.epos[size_t(P.elem_start + i) * 4 + a] = uint16_t(pos[l]++);
        }
    } else {
      const std::string err = spread_entry_banks(plan, P, npe, pn_max, nloc, run, bank);
      if (!err.empty()) return err;
    }
    for (int32_t v : loc_nodes) local[v] = -1;
  }
  return "";
}

void pack_records(const MeshPlan &plan, std::vector<uint32_t> &rec) {
  rec.assign(size_t(plan.ne) * 3, 0u);
  for (int64_t e = 0; e < plan.ne; ++e) {
    uint32_t l[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
    for (int a = 0; a < plan.npe; ++a) {
      l[a] = plan.elem_lid[size_t(e) * 4 + a];
      q[a] = plan.epos[size_t(e) * 4 + a];
    }
    rec[size_t(e) * 3 + 0] = l[0] | l[1] << 10 | l[2] << 20 | (l[3] & 3u) << 30;
    rec[size_t(e) * 3 + 1] = l[3] >> 2 | q[0] << 8 | (q[1] & 1023u) << 22;
    rec[size_t(e) * 3 + 2] = q[1] >> 10 | q[2] << 4 | q[3] << 18;
  }
}

bool stride_records(const MeshPlan &plan, std::vector<uint32_t> &rec,
                    std::vector<uint32_t> &rec_s) {
  const size_t emax = size_t(plan.emax_used);
  bool regular = true;
  for (size_t p = 0; p < plan.patches.size(); ++p)
    regular = regular && size_t(plan.patches[p].elem_start) == p * emax;
  rec_s.clear();
  if (regular) {
    rec.resize(rec.size() + emax * 3, 0u);
    return true;
  }
  rec_s.assign(plan.patches.size() * emax * 3, 0u);
  for (size_t p = 0; p < plan.patches.size(); ++p)
    std::copy(rec.begin() + size_t(plan.patches[p].elem_start) * 3,
              rec.begin() + (size_t(plan.patches[p].elem_start) + plan.patches[p].n_elem) * 3,
              rec_s.begin() + p * emax * 3);
  return false;
}

void pack_patch_nodes(const MeshPlan &plan, int stride, std::vector<int32_t> &pgid,
                      std::vector<int32_t> &pslot, std::vector<uint16_t> &pend) {
  const size_t np = plan.patches.size();
  pgid.assign(np * size_t(stride), 0);
  pslot.assign(np * size_t(stride), 0);
  pend.assign(np * size_t(stride), 0);
  for (size_t p = 0; p < np; ++p) {
    const PatchInfo &Q = plan.patches[p];
    for (int k = 0; k < Q.n_own + Q.n_sh; ++k) {
      const size_t i = p * size_t(stride) + size_t(k);
      if (k < Q.n_own) {
        pgid[i] = Q.own_start + k;
      } else {
        const int32_t gid = plan.sh_gid[size_t(Q.sh_start + k - Q.n_own)];
        const int32_t slot = plan.sh_slot[size_t(Q.sh_start + k - Q.n_own)];
        const bool first = slot == plan.slot_start[size_t(gid - plan.n_owned)];
        pgid[i] = gid;
        pslot[i] = int32_t(uint32_t(slot) | (first ? 0x80000000u : 0u));
      }
      pend[i] = plan.node_end[size_t(Q.pnode_start + k)];
    }
  }
}

}  // namespace femgpu
