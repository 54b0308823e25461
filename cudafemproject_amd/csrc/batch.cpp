// batch.cpp -- host side of the batched solve (femgpu_batch_* in femgpu.h):
// plans every instance, turns the plans into the flat per-element and
// per-patch-node tables of batch.h and launches the persistent PCG kernel
// (batch.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "batch.h"
#include "batch_ls.h"
#include "femgpu.h"
#include "kernels.h"

namespace femgpu {

struct BatchState {
  int dim = 0, n_inst = 0;
  int64_t total_nodes = 0, total_elems = 0;
  std::vector<int64_t> node_off, elem_off;  // user offsets, n_inst + 1
  std::vector<BInst> inst;
  std::vector<int32_t> node_perm;  // internal global node -> user global node
  std::vector<int32_t> node_inv;   // user global node -> internal global node
  std::vector<int32_t> order;
  std::vector<uint8_t> fix;        // per internal global node
  int64_t n_slots = 0;
  // lockstep schedule (batch_ls.h): its own plans and node numbering
  std::vector<int32_t> bl_perm, bl_inv;  // internal <-> user global node
  std::vector<uint8_t> bl_fix;
  std::vector<BLInst> bl_inst;
  int bl_np = 0, bl_nblk = 0, bl_nif = 0, bl_ent = 0, bl_pn = 0;
  int64_t bl_slots = 0;
  int *h_count = nullptr;                // pinned
  std::map<std::string, std::pair<void *, size_t>> bufs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_poll[2] = {nullptr, nullptr};
};

void batch_free(BatchState *b) {
  if (!b) return;
  for (auto &kv : b->bufs)
    if (kv.second.first) (void)hipFree(kv.second.first);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (hipEvent_t e : b->ev_poll)
    if (e) (void)hipEventDestroy(e);
  if (b->h_count) (void)hipHostFree(b->h_count);
  delete b;
}

namespace {

// Named device buffer of at least n elements (reallocated when too small).
template <typename T>
hipError_t bbuf(BatchState *b, const std::string &name, size_t n, T **out) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  auto it = b->bufs.find(name);
  if (it != b->bufs.end() && it->second.second >= bytes) {
    *out = static_cast<T *>(it->second.first);
    return hipSuccess;
  }
  if (it != b->bufs.end()) {
    (void)hipFree(it->second.first);
    b->bufs.erase(it);
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return e;
  b->bufs[name] = {p, bytes};
  *out = static_cast<T *>(p);
  return hipSuccess;
}

template <typename T>
hipError_t bup(BatchState *b, hipStream_t s, const std::string &name, const std::vector<T> &v,
               T **out) {
  hipError_t e = bbuf(b, name, v.size(), out);
  if (e != hipSuccess || v.empty()) return e;
  e = hipMemcpyAsync(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

double elem_det(int dim, const double *nodes, const int32_t *en) {
  if (dim == 2) {
    const double *a = &nodes[size_t(en[0]) * 2], *b = &nodes[size_t(en[1]) * 2],
                 *c = &nodes[size_t(en[2]) * 2];
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
  }
  const double *a = &nodes[size_t(en[0]) * 3], *b = &nodes[size_t(en[1]) * 3],
               *c = &nodes[size_t(en[2]) * 3], *d = &nodes[size_t(en[3]) * 3];
  const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double e3[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
         e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
}

// Lockstep plans (batch_ls.h): every instance planned with the
// single-problem patch limits, plans concatenated, tables uploaded.
std::string build_lockstep(BatchState *b, hipStream_t s, const int64_t *node_off,
                           const double *nodes, const int64_t *elem_off, const int32_t *elems) {
  const int dim = b->dim, npe = dim + 1;
  std::vector<PatchInfo> pat;
  std::vector<uint32_t> rec;
  std::vector<int32_t> shinfo, iface;
  std::vector<uint16_t> nend;
  std::vector<BLBlock> blocks;
  b->bl_perm.assign(b->total_nodes, 0);
  b->bl_inv.assign(b->total_nodes, 0);
  b->bl_fix.assign(b->total_nodes, 0);
  b->bl_inst.clear();
  std::vector<double> c64(size_t(b->total_nodes) * dim);
  int32_t E0 = 0, S0 = 0;
  int ent = 0, pn = 0;
  for (int i = 0; i < b->n_inst; ++i) {
    const int64_t nn = node_off[i + 1] - node_off[i], ne = elem_off[i + 1] - elem_off[i];
    const double *nd = nodes + size_t(node_off[i]) * dim;
    const int32_t *el = elems + size_t(elem_off[i]) * npe;
    MeshPlan Q;
    const std::string msg = build_plan(dim, nn, nd, ne, el, 3 * BL_THREADS, BL_THREADS, true, Q);
    if (!msg.empty()) return "instance " + std::to_string(i) + ": " + msg;
    const int32_t N0 = int32_t(node_off[i]);
    BLInst I{};
    I.node_off = N0;
    I.n_owned = int32_t(Q.n_owned);
    I.n_active = int32_t(Q.n_owned + Q.n_iface);
    I.n_nodes = int32_t(nn);
    I.p0 = int32_t(pat.size());
    I.if0 = int32_t(iface.size() / 4);
    const int32_t SH0 = int32_t(shinfo.size() / 4), PN0 = int32_t(nend.size());
    for (PatchInfo pi : Q.patches) {
      pi.elem_start += E0;
      pi.own_start += N0;
      pi.sh_start += SH0;
      pi.pnode_start += PN0;
      pi.pad = i;
      pat.push_back(pi);
    }
    I.p1 = int32_t(pat.size());
    std::vector<uint32_t> r;
    pack_records(Q, r);
    rec.insert(rec.end(), r.begin(), r.end());
    for (size_t k = 0; k < Q.sh_gid.size(); ++k) {
      const int32_t t = Q.sh_gid[k] - int32_t(Q.n_owned);
      shinfo.insert(shinfo.end(), {N0 + Q.sh_gid[k], S0 + Q.sh_slot[k], S0 + Q.slot_start[t],
                                   S0 + Q.slot_start[t + 1]});
    }
    nend.insert(nend.end(), Q.node_end.begin(), Q.node_end.end());
    for (int64_t t = 0; t < Q.n_iface; ++t)
      iface.insert(iface.end(), {N0 + int32_t(Q.n_owned + t), S0 + Q.slot_start[t],
                                 S0 + Q.slot_start[t + 1], 0});
    I.u0 = int32_t(blocks.size());
    for (int32_t n = 0; n < I.n_active; n += BL_UPD_NODES)
      blocks.push_back({i, N0 + n, N0 + std::min(I.n_active, n + BL_UPD_NODES), 0});
    I.u1 = int32_t(blocks.size());
    const double E = 2e7, nu = 0.3;
    I.mu = E / (2.0 * (1.0 + nu));
    I.lam = dim == 3 ? E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)) : E * nu / (1.0 - nu * nu);
    b->bl_inst.push_back(I);
    for (int64_t j = 0; j < nn; ++j) {
      b->bl_perm[N0 + j] = N0 + Q.node_perm[j];
      b->bl_inv[N0 + Q.node_perm[j]] = int32_t(N0 + j);
      for (int c = 0; c < dim; ++c)
        c64[size_t(N0 + j) * dim + c] = nd[size_t(Q.node_perm[j]) * dim + c];
    }
    E0 += int32_t(ne);
    S0 += int32_t(Q.n_slots);
    ent = std::max(ent, (Q.ent_used + 3) & ~3);
    pn = std::max(pn, Q.pn_used);
  }
  b->bl_np = int(pat.size());
  b->bl_nblk = int(blocks.size());
  b->bl_nif = int(iface.size() / 4);
  b->bl_ent = ent;
  b->bl_pn = pn;
  b->bl_slots = S0;
  std::vector<float> c32(c64.begin(), c64.end());
  PatchInfo *d0;
  uint32_t *d1;
  int32_t *d2, *d4;
  uint16_t *d3;
  BLInst *d5;
  BLBlock *d6;
  double *d7;
  float *d8;
  uint8_t *d9;
  hipError_t e;
  if ((e = bup(b, s, "bl_patches", pat, &d0)) != hipSuccess ||
      (e = bup(b, s, "bl_erec", rec, &d1)) != hipSuccess ||
      (e = bup(b, s, "bl_shinfo", shinfo, &d2)) != hipSuccess ||
      (e = bup(b, s, "bl_nodeend", nend, &d3)) != hipSuccess ||
      (e = bup(b, s, "bl_iface", iface, &d4)) != hipSuccess ||
      (e = bup(b, s, "bl_inst", b->bl_inst, &d5)) != hipSuccess ||
      (e = bup(b, s, "bl_blocks", blocks, &d6)) != hipSuccess ||
      (e = bup(b, s, "bl_c64", c64, &d7)) != hipSuccess ||
      (e = bup(b, s, "bl_c32", c32, &d8)) != hipSuccess ||
      (e = bup(b, s, "bl_fix", b->bl_fix, &d9)) != hipSuccess)
    return std::string("upload failed: ") + hipGetErrorString(e);
  return "";
}

// Lockstep solve (batch_ls.h) of n_rhs right-hand sides padded to nr.
// Graph replays of REPLAY iterations (+ the residual replacement step in
// mixed precision) until every system has finished.
template <typename T>
int solve_lockstep(femgpu_ctx *ctx, BatchState *b, int n_rhs, int nr, const double *f,
                   const femgpu_solve_opts &o, double *u_out, femgpu_batch_stat *stats,
                   double *solve_ms) {
  constexpr bool MIXED = sizeof(T) == sizeof(float);
  constexpr int REPLAY = 32;
  const int dim = b->dim;
  hipStream_t s = ctx_stream(ctx);
  const size_t nv = size_t(b->total_nodes) * nr * dim;
  const size_t nd = size_t(b->total_nodes) * dim;
  const int n_state = b->n_inst * nr;
  std::vector<double> fi(nv, 0.0);
  for (int64_t g = 0; g < b->total_nodes; ++g) {
    const int64_t u = b->bl_perm[g];
    for (int r = 0; r < n_rhs; ++r)
      for (int c = 0; c < dim; ++c)
        fi[(size_t(g) * nr + r) * dim + c] = f[(size_t(u) * n_rhs + r) * dim + c];
  }
  const std::string t = MIXED ? "bl32_" : "bl64_";
  double *d_f, *x64, *y64, *s64, *diag64, *ppq, *pnode;
  T *r, *z, *p, *q, *sT, *invD, *diagT;
  BLState *st;
  int *d_count;
  const size_t ns = size_t(b->bl_slots) * nr * dim + 1;
#define BL_BUF(name, n, ptr) BCHK_(bbuf(b, name, n, ptr))
#define BCHK_(call)                                                                         \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ctx_fail(ctx, FEMGPU_E_HIP, (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)
  BCHK_(bup(b, s, "bl_f", fi, &d_f));
  BL_BUF("bl_x64", nv, &x64);
  BL_BUF("bl_y64", nv, &y64);
  BL_BUF("bl_s64", ns, &s64);
  BL_BUF("bl_diag64", nd, &diag64);
  BL_BUF("bl_ppq", size_t(b->bl_np) * nr, &ppq);
  BL_BUF("bl_pnode", size_t(4) * b->bl_nblk * nr, &pnode);
  BL_BUF("bl_state", size_t(n_state), &st);
  BL_BUF("bl_count", 2, &d_count);
  BL_BUF(t + "r", nv, &r);
  BL_BUF(t + "z", nv, &z);
  BL_BUF(t + "p", nv, &p);
  BL_BUF(t + "q", nv, &q);
  BL_BUF(t + "s", ns, &sT);
  BL_BUF(t + "invD", nd, &invD);
  BL_BUF(t + "diag", nd, &diagT);
  // mixed: the replay's solution increments in FP32, added to the FP64
  // solution before each replacement
  T *xinc = nullptr;
  if constexpr (MIXED) BL_BUF(t + "xinc", nv, &xinc);
  auto fill = [&](auto &a) {
    a.patches = static_cast<const PatchInfo *>(b->bufs["bl_patches"].first);
    a.erec = static_cast<const uint32_t *>(b->bufs["bl_erec"].first);
    a.sh_info = static_cast<const int4 *>(b->bufs["bl_shinfo"].first);
    a.node_end = static_cast<const uint16_t *>(b->bufs["bl_nodeend"].first);
    a.fixmask = static_cast<const uint8_t *>(b->bufs["bl_fix"].first);
    a.inst = static_cast<const BLInst *>(b->bufs["bl_inst"].first);
    a.blocks = static_cast<const BLBlock *>(b->bufs["bl_blocks"].first);
    a.iface = static_cast<const int4 *>(b->bufs["bl_iface"].first);
    a.np = b->bl_np;
    a.p8 = (b->bl_np + 7) / 8;
    a.ent = b->bl_ent;
    a.pn = b->bl_pn;
    a.nblk = b->bl_nblk;
    a.st = st;
    a.max_iter = o.max_iter > 0 ? o.max_iter : 100000;
    a.tol2 = o.tol * o.tol;
    const double drop = o.inner_tol > 0 && o.inner_tol < 1 ? o.inner_tol : 0.1;
    a.drop2 = MIXED ? drop * drop : 0.0;
  };
  BLArgs<double> m64{};
  fill(m64);
  m64.coords = static_cast<const double *>(b->bufs["bl_c64"].first);
  m64.diagK = diag64;
  BLArgs<T> m{};
  fill(m);
  if constexpr (MIXED) {
    m.coords = static_cast<const T *>(b->bufs["bl_c32"].first);
    m.diagK = diagT;
  } else {
    m.coords = m64.coords;
    m.diagK = diag64;
  }
  BCHK_(hipEventRecord(b->ev0, s));
  // diag(K), D^-1, state, r = f
  bl_diag(dim, m64, b->bl_nif, diag64, s64, s);
  bl_inv_diag<T>(int64_t(nd), diag64, invD, s);
  if constexpr (MIXED) launch_convert<double, T>(diag64, diagT, int64_t(nd), s);
  launch_fill<double>(x64, int64_t(nv), 0.0, s);
  launch_fill<T>(p, int64_t(nv), T(0), s);
  if (xinc) launch_fill<T>(xinc, int64_t(nv), T(0), s);
  bl_init_state(n_state, st, s);
  bl_replace<T>(dim, nr, m, m64, d_f, x64, y64, s64, invD, r, z, pnode, n_state, d_count, s);
  BCHK_(hipGetLastError());
  // graph of REPLAY iterations (+ replacement) on these buffers
  hipGraph_t g;
  hipGraphExec_t exec = nullptr;
  BCHK_(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < REPLAY; ++k)
    bl_iteration<T>(dim, nr, m, invD, z, p, q, r, sT, x64, xinc, ppq, pnode, k & 1, s);
  if constexpr (MIXED) bl_flush<T>(dim, nr, m, xinc, x64, s);
  bl_replace<T>(dim, nr, m, m64, d_f, x64, y64, s64, invD, r, z, pnode, n_state, d_count, s);
  BCHK_(hipStreamEndCapture(s, &g));
  const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BCHK_(ie);
  // replays, pipelined one behind the host's look at the count
  const long max_replays = long(m64.max_iter) / REPLAY + 4;
  int cur = 0;
  auto launch = [&](int slot) -> hipError_t {
    hipError_t e;
    if ((e = hipGraphLaunch(exec, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&b->h_count[slot], d_count, sizeof(int), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
    return hipEventRecord(b->ev_poll[slot], s);
  };
  hipError_t le = launch(0);
  for (long rep = 0; le == hipSuccess; ++rep) {
    const int nxt = cur ^ 1;
    if (rep + 1 < max_replays && (le = launch(nxt)) != hipSuccess) break;
    if ((le = hipEventSynchronize(b->ev_poll[cur])) != hipSuccess) break;
    if (b->h_count[cur] == 0 || rep + 1 >= max_replays) break;
    cur = nxt;
  }
  if (le != hipSuccess) {
    (void)hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
    BCHK_(le);
  }
  bl_final(dim, nr, m64, d_f, x64, y64, s64, pnode, b->n_inst, s);
  BCHK_(hipGetLastError());
  BCHK_(hipEventRecord(b->ev1, s));
  // the last queued replay may still run until here: destroy the graph only
  // once the stream has passed it
  const hipError_t se = hipEventSynchronize(b->ev1);
  (void)hipGraphExecDestroy(exec);
  BCHK_(se);
  float ms = 0.f;
  BCHK_(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (solve_ms) *solve_ms = ms;
  std::vector<double> xi(nv);
  std::vector<BLState> hs(n_state);
  BCHK_(hipMemcpyAsync(xi.data(), x64, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  BCHK_(hipMemcpyAsync(hs.data(), st, hs.size() * sizeof(BLState), hipMemcpyDeviceToHost, s));
  BCHK_(hipStreamSynchronize(s));
#undef BL_BUF
#undef BCHK_
  if (u_out)
    for (int64_t gq = 0; gq < b->total_nodes; ++gq) {
      const int64_t u = b->bl_perm[gq];
      for (int rr = 0; rr < n_rhs; ++rr)
        for (int c = 0; c < dim; ++c)
          u_out[(size_t(u) * n_rhs + rr) * dim + c] = xi[(size_t(gq) * nr + rr) * dim + c];
    }
  int worst = FEMGPU_OK;
  for (int i = 0; i < b->n_inst; ++i)
    for (int rr = 0; rr < n_rhs; ++rr) {
      const BLState &x = hs[size_t(i) * nr + rr];
      const int status = x.state == 3 ? x.status : FEMGPU_E_MAXIT;
      if (stats) {
        femgpu_batch_stat &o2 = stats[size_t(i) * n_rhs + rr];
        o2.status = status;
        o2.iterations = x.state == 3 ? x.iters : std::max(x.it[0], x.it[1]) + 1;
        o2.replacements = std::max(x.nrep, 0);
        o2.rel_res = x.ff > 0.0 ? std::sqrt(x.rr / x.ff) : 0.0;
      }
      if (status != FEMGPU_OK && worst == FEMGPU_OK) worst = status;
    }
  if (worst != FEMGPU_OK)
    return ctx_fail(ctx, worst, "some batch systems did not converge (see stats)");
  return FEMGPU_OK;
}

}  // namespace
}  // namespace femgpu

using femgpu::BatchState;
using femgpu::BInst;
using femgpu::BStat;

#define BCHK(call)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return femgpu::ctx_fail(ctx, FEMGPU_E_HIP,                                            \
                              (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

extern "C" {

int femgpu_batch_set_mesh(femgpu_ctx *ctx, int dim, int n_inst, const int64_t *node_off,
                          const double *nodes, const int64_t *elem_off, const int32_t *elems) {
  if (!ctx) return FEMGPU_E_ARG;
  if ((dim != 2 && dim != 3) || n_inst < 1 || !node_off || !nodes || !elem_off || !elems)
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "bad batch mesh arguments");
  BCHK(hipSetDevice(femgpu::ctx_device(ctx)));
  hipStream_t s = femgpu::ctx_stream(ctx);
  BCHK(hipStreamSynchronize(s));
  BatchState *&bp = femgpu::ctx_batch(ctx);
  femgpu::batch_free(bp);
  bp = nullptr;
  auto *b = new BatchState();
  const int npe = dim + 1;
  const int emax_b = 512;
  if (node_off[0] != 0 || elem_off[0] != 0) {
    delete b;
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "batch offsets must start at 0");
  }
  for (int i = 0; i < n_inst; ++i)
    if (node_off[i + 1] < node_off[i] + npe || elem_off[i + 1] <= elem_off[i]) {
      delete b;
      return femgpu::ctx_fail(ctx, FEMGPU_E_BAD_MESH,
                              ("instance " + std::to_string(i) + " has too few nodes or no element").c_str());
    }
  if (node_off[n_inst] >= (int64_t(1) << 31) / 16) {
    delete b;
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "batch too large for 32-bit node ids");
  }
  b->dim = dim;
  b->n_inst = n_inst;
  b->node_off.assign(node_off, node_off + n_inst + 1);
  b->elem_off.assign(elem_off, elem_off + n_inst + 1);
  b->total_nodes = node_off[n_inst];
  b->total_elems = elem_off[n_inst];
  b->node_perm.assign(b->total_nodes, 0);
  b->node_inv.assign(b->total_nodes, 0);
  b->fix.assign(b->total_nodes, 0);
  std::vector<int32_t> bel, bpos, pbeg, pend, ptgt, ss;
  std::vector<double> c64(size_t(b->total_nodes) * dim);
  int64_t e_int = 0;
  for (int i = 0; i < n_inst; ++i) {
    const int64_t nn = node_off[i + 1] - node_off[i], ne = elem_off[i + 1] - elem_off[i];
    const double *nd = nodes + size_t(node_off[i]) * dim;
    const int32_t *el = elems + size_t(elem_off[i]) * npe;
    for (int64_t e = 0; e < ne; ++e) {
      for (int a = 0; a < npe; ++a)
        if (el[e * npe + a] < 0 || el[e * npe + a] >= nn) {
          delete b;
          return femgpu::ctx_fail(ctx, FEMGPU_E_BAD_MESH,
                                  ("instance " + std::to_string(i) + ": node id out of range").c_str());
        }
      const double det = femgpu::elem_det(dim, nd, &el[e * npe]);
      if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
        delete b;
        return femgpu::ctx_fail(ctx, FEMGPU_E_BAD_MESH,
                                ("instance " + std::to_string(i) + ": degenerate element " +
                                 std::to_string(e)).c_str());
      }
    }
    femgpu::MeshPlan P;
    const std::string msg = femgpu::build_plan(dim, nn, nd, ne, el, emax_b, femgpu::PN_LIMIT, false, P);
    if (!msg.empty()) {
      delete b;
      return femgpu::ctx_fail(ctx, FEMGPU_E_BAD_MESH, ("instance " + std::to_string(i) + ": " + msg).c_str());
    }
    const int32_t N0 = int32_t(node_off[i]);
    const int32_t PN0 = int32_t(pbeg.size());
    const int32_t S0 = int32_t(b->n_slots), IF0 = int32_t(ss.size());
    bel.insert(bel.end(), P.elems_int.begin(), P.elems_int.end());
    const size_t pos0 = bpos.size();
    bpos.resize(pos0 + size_t(ne) * npe);
    for (const femgpu::PatchInfo &pi : P.patches) {
      const int32_t base = pi.elem_start * npe;  // the patch's entries follow its elements
      for (int32_t e = pi.elem_start; e < pi.elem_start + pi.n_elem; ++e)
        for (int a = 0; a < npe; ++a) bpos[pos0 + size_t(e) * npe + a] = base + P.epos[size_t(e) * 4 + a];
      for (int32_t k = 0; k < pi.n_own + pi.n_sh; ++k) {
        const int32_t e1 = P.node_end[pi.pnode_start + k];
        const int32_t e0 = k ? int32_t(P.node_end[pi.pnode_start + k - 1]) : 0;
        pbeg.push_back(base + e0);
        pend.push_back(base + e1);
        ptgt.push_back(k < pi.n_own ? pi.own_start + k : -1 - P.sh_slot[pi.sh_start + k - pi.n_own]);
      }
    }
    for (int32_t st : P.slot_start) ss.push_back(st + S0);
    for (int64_t j = 0; j < nn; ++j) {
      b->node_perm[N0 + j] = N0 + P.node_perm[j];
      b->node_inv[N0 + P.node_perm[j]] = int32_t(N0 + j);
      for (int c = 0; c < dim; ++c)
        c64[size_t(N0 + j) * dim + c] = nd[size_t(P.node_perm[j]) * dim + c];
    }
    BInst I{};
    I.node_off = N0;
    I.n_nodes = int32_t(nn);
    I.n_owned = int32_t(P.n_owned);
    I.n_active = int32_t(P.n_owned + P.n_iface);
    I.elem_off = int32_t(e_int);
    I.n_elems = int32_t(ne);
    I.pn_off = PN0;
    I.n_pn = int32_t(pbeg.size()) - PN0;
    I.iface_off = IF0;
    I.slot_off = S0;
    const double E = 2e7, nu = 0.3;
    I.mu = E / (2.0 * (1.0 + nu));
    I.lam = dim == 3 ? E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)) : E * nu / (1.0 - nu * nu);
    b->inst.push_back(I);
    b->n_slots += P.n_slots;
    e_int += ne;
  }
  // largest instances first
  b->order.resize(n_inst);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
    return b->elem_off[a + 1] - b->elem_off[a] > b->elem_off[c + 1] - b->elem_off[c];
  });
  std::vector<float> c32(c64.begin(), c64.end());
  int32_t *p0, *p1, *p2, *p3, *p4, *p5, *p6;
  double *dc64;
  float *dc32;
  BInst *di;
  uint8_t *df;
  BCHK(femgpu::bup(b, s, "elems", bel, &p0));
  BCHK(femgpu::bup(b, s, "gpos", bpos, &p1));
  BCHK(femgpu::bup(b, s, "pbeg", pbeg, &p2));
  BCHK(femgpu::bup(b, s, "pend", pend, &p3));
  BCHK(femgpu::bup(b, s, "ptgt", ptgt, &p4));
  BCHK(femgpu::bup(b, s, "ss", ss, &p5));
  BCHK(femgpu::bup(b, s, "order", b->order, &p6));
  BCHK(femgpu::bup(b, s, "c64", c64, &dc64));
  BCHK(femgpu::bup(b, s, "c32", c32, &dc32));
  BCHK(femgpu::bup(b, s, "inst", b->inst, &di));
  BCHK(femgpu::bup(b, s, "fix", b->fix, &df));
  const std::string bl_msg = femgpu::build_lockstep(b, s, node_off, nodes, elem_off, elems);
  if (!bl_msg.empty()) {
    femgpu::batch_free(b);
    return femgpu::ctx_fail(ctx, FEMGPU_E_BAD_MESH, bl_msg.c_str());
  }
  BCHK(hipEventCreate(&b->ev0));
  BCHK(hipEventCreate(&b->ev1));
  for (hipEvent_t &e : b->ev_poll) BCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  BCHK(hipHostMalloc(reinterpret_cast<void **>(&b->h_count), 2 * sizeof(int), 0));
  bp = b;
  return FEMGPU_OK;
}

int femgpu_batch_set_material(femgpu_ctx *ctx, const double *E, const double *nu) {
  if (!ctx) return FEMGPU_E_ARG;
  BatchState *b = femgpu::ctx_batch(ctx);
  if (!b) return femgpu::ctx_fail(ctx, FEMGPU_E_STATE, "no batch mesh");
  if (!E || !nu) return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "null material arrays");
  for (int i = 0; i < b->n_inst; ++i) {
    if (!(E[i] > 0.0) || !(nu[i] > -1.0) || !(nu[i] < (b->dim == 3 ? 0.5 : 1.0)))
      return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, ("bad material of instance " + std::to_string(i)).c_str());
    BInst &I = b->inst[i];
    I.mu = E[i] / (2.0 * (1.0 + nu[i]));
    I.lam = b->dim == 3 ? E[i] * nu[i] / ((1.0 + nu[i]) * (1.0 - 2.0 * nu[i]))
                        : E[i] * nu[i] / (1.0 - nu[i] * nu[i]);
    b->bl_inst[i].mu = I.mu;
    b->bl_inst[i].lam = I.lam;
  }
  BInst *di;
  femgpu::BLInst *dbi;
  BCHK(femgpu::bup(b, femgpu::ctx_stream(ctx), "inst", b->inst, &di));
  BCHK(femgpu::bup(b, femgpu::ctx_stream(ctx), "bl_inst", b->bl_inst, &dbi));
  return FEMGPU_OK;
}

int femgpu_batch_set_constraints(femgpu_ctx *ctx, const int64_t *con_off, const int32_t *node,
                                 const int32_t *code) {
  if (!ctx) return FEMGPU_E_ARG;
  BatchState *b = femgpu::ctx_batch(ctx);
  if (!b) return femgpu::ctx_fail(ctx, FEMGPU_E_STATE, "no batch mesh");
  if (!con_off) return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "null constraint offsets");
  std::fill(b->fix.begin(), b->fix.end(), 0);
  std::fill(b->bl_fix.begin(), b->bl_fix.end(), 0);
  for (int i = 0; i < b->n_inst; ++i) {
    const int64_t nn = b->node_off[i + 1] - b->node_off[i];
    if (con_off[i + 1] < con_off[i]) return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "bad constraint offsets");
    for (int64_t k = con_off[i]; k < con_off[i + 1]; ++k) {
      if (node[k] < 0 || node[k] >= nn)
        return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "constraint node out of range");
      // bitwise decode of datakeeper.cpp:148-164
      uint8_t m = 0;
      if (code[k] & 1) m |= 1;
      if (code[k] & 2) m |= 2;
      if (b->dim == 3 && (code[k] & 3)) m |= 4;
      b->fix[b->node_inv[b->node_off[i] + node[k]]] |= m;
      b->bl_fix[b->bl_inv[b->node_off[i] + node[k]]] |= m;
    }
  }
  uint8_t *df;
  BCHK(femgpu::bup(b, femgpu::ctx_stream(ctx), "fix", b->fix, &df));
  BCHK(femgpu::bup(b, femgpu::ctx_stream(ctx), "bl_fix", b->bl_fix, &df));
  return FEMGPU_OK;
}

int femgpu_batch_solve(femgpu_ctx *ctx, int n_rhs, const double *f, const femgpu_solve_opts *opts,
                       double *u_out, femgpu_batch_stat *stats, double *solve_ms) {
  if (!ctx) return FEMGPU_E_ARG;
  BatchState *b = femgpu::ctx_batch(ctx);
  if (!b) return femgpu::ctx_fail(ctx, FEMGPU_E_STATE, "no batch mesh");
  if (n_rhs < 1 || n_rhs > 4 || !f)
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "n_rhs must be 1..4 and f non-null");
  femgpu_solve_opts o;
  femgpu_default_solve_opts(&o);
  if (opts) o = *opts;
  if (o.precision != FEMGPU_FP64 && o.precision != FEMGPU_MIXED)
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "batch precision must be FP64 or MIXED");
  if (o.tol_kind != FEMGPU_TOL_REL_RES || !(o.tol >= 0))
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "batch solve needs FEMGPU_TOL_REL_RES and tol >= 0");
  const int nr = n_rhs == 3 ? 4 : n_rhs;  // kernels exist for 1, 2, 4 vectors
  const int dim = b->dim;
  hipStream_t s = femgpu::ctx_stream(ctx);
  BCHK(hipSetDevice(femgpu::ctx_device(ctx)));
  if (o.batch_schedule == FEMGPU_BATCH_LOCKSTEP) {
    if (o.precision == FEMGPU_MIXED)
      return femgpu::solve_lockstep<float>(ctx, b, n_rhs, nr, f, o, u_out, stats, solve_ms);
    return femgpu::solve_lockstep<double>(ctx, b, n_rhs, nr, f, o, u_out, stats, solve_ms);
  }
  if (o.batch_schedule != FEMGPU_BATCH_PERSISTENT)
    return femgpu::ctx_fail(ctx, FEMGPU_E_ARG, "unknown batch schedule");
  const size_t nv = size_t(b->total_nodes) * nr * dim;
  const size_t ns = size_t(b->n_slots) * nr * dim;
  std::vector<double> fi(nv, 0.0);
  for (int64_t g = 0; g < b->total_nodes; ++g) {
    const int64_t u = b->node_perm[g];
    for (int r = 0; r < n_rhs; ++r)
      for (int c = 0; c < dim; ++c)
        fi[(size_t(g) * nr + r) * dim + c] = f[(size_t(u) * n_rhs + r) * dim + c];
  }
  femgpu::BatchArgs<float> af{};
  femgpu::BatchArgs<double> ad{};
  double *d_f, *d_x, *d_y, *d_s64;
  BStat *d_stat;
  BCHK(femgpu::bup(b, s, "f", fi, &d_f));
  BCHK(femgpu::bbuf(b, "x", nv, &d_x));
  BCHK(femgpu::bbuf(b, "y64", nv, &d_y));
  BCHK(femgpu::bbuf(b, "s64", ns, &d_s64));
  BCHK(femgpu::bbuf(b, "stat", size_t(b->n_inst) * nr, &d_stat));
  double *d_ent;
  BCHK(femgpu::bbuf(b, "ent", size_t(b->total_elems) * (dim + 1) * nr * dim, &d_ent));
  auto fill_common = [&](auto &a) {
    a.inst = static_cast<const BInst *>(b->bufs["inst"].first);
    a.order = static_cast<const int32_t *>(b->bufs["order"].first);
    a.elems = static_cast<const int32_t *>(b->bufs["elems"].first);
    a.gpos = static_cast<const int32_t *>(b->bufs["gpos"].first);
    a.pn_beg = static_cast<const int32_t *>(b->bufs["pbeg"].first);
    a.pn_end = static_cast<const int32_t *>(b->bufs["pend"].first);
    a.pn_tgt = static_cast<const int32_t *>(b->bufs["ptgt"].first);
    a.slot_start = static_cast<const int32_t *>(b->bufs["ss"].first);
    a.fixmask = static_cast<const uint8_t *>(b->bufs["fix"].first);
    a.ent = d_ent;
    a.coords64 = static_cast<const double *>(b->bufs["c64"].first);
    a.f = d_f;
    a.x = d_x;
    a.y64 = d_y;
    a.slots64 = d_s64;
    a.stat = d_stat;
    a.tol = o.tol;
    a.drop = o.inner_tol > 0 && o.inner_tol < 1 ? o.inner_tol : 0.1;
    a.max_iter = o.max_iter > 0 ? o.max_iter : 100000;
  };
  const bool mixed = o.precision == FEMGPU_MIXED;
  if (mixed) {
    fill_common(af);
    af.coordsT = static_cast<const float *>(b->bufs["c32"].first);
    BCHK(femgpu::bbuf(b, "r32", nv, &af.r));
    BCHK(femgpu::bbuf(b, "p32", nv, &af.p));
    BCHK(femgpu::bbuf(b, "q32", nv, &af.q));
    BCHK(femgpu::bbuf(b, "s32", ns, &af.slotsT));
    BCHK(femgpu::bbuf(b, "iD32", size_t(b->total_nodes) * dim, &af.invD));
  } else {
    fill_common(ad);
    ad.coordsT = static_cast<const double *>(b->bufs["c64"].first);
    BCHK(femgpu::bbuf(b, "r64", nv, &ad.r));
    BCHK(femgpu::bbuf(b, "p64", nv, &ad.p));
    BCHK(femgpu::bbuf(b, "q64", nv, &ad.q));
    BCHK(femgpu::bbuf(b, "sT64", ns, &ad.slotsT));
    BCHK(femgpu::bbuf(b, "iD64", size_t(b->total_nodes) * dim, &ad.invD));
  }
  BCHK(hipEventRecord(b->ev0, s));
  if (femgpu::launch_batch_pcg(dim, nr, mixed, af, ad, b->n_inst, s) != 0)
    return femgpu::ctx_fail(ctx, FEMGPU_E_UNSUPPORTED, "no batch kernel for this shape");
  BCHK(hipGetLastError());
  BCHK(hipEventRecord(b->ev1, s));
  BCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (solve_ms) *solve_ms = ms;
  std::vector<double> xi(nv);
  std::vector<BStat> st(size_t(b->n_inst) * nr);
  BCHK(hipMemcpyAsync(xi.data(), d_x, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  BCHK(hipMemcpyAsync(st.data(), d_stat, st.size() * sizeof(BStat), hipMemcpyDeviceToHost, s));
  BCHK(hipStreamSynchronize(s));
  if (u_out) {
    for (int64_t g = 0; g < b->total_nodes; ++g) {
      const int64_t u = b->node_perm[g];
      for (int r = 0; r < n_rhs; ++r)
        for (int c = 0; c < dim; ++c)
          u_out[(size_t(u) * n_rhs + r) * dim + c] = xi[(size_t(g) * nr + r) * dim + c];
    }
  }
  int worst = FEMGPU_OK;
  for (int i = 0; i < b->n_inst; ++i)
    for (int r = 0; r < n_rhs; ++r) {
      const BStat &x = st[size_t(i) * nr + r];
      if (stats) {
        femgpu_batch_stat &o2 = stats[size_t(i) * n_rhs + r];
        o2.status = x.status;
        o2.iterations = x.iterations;
        o2.replacements = x.replacements;
        o2.rel_res = x.rel_res;
      }
      if (x.status != FEMGPU_OK && worst == FEMGPU_OK) worst = x.status;
    }
  if (worst != FEMGPU_OK)
    return femgpu::ctx_fail(ctx, worst, "some batch systems did not converge (see stats)");
  return FEMGPU_OK;
}

}  // extern "C"
