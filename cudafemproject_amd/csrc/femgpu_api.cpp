// femgpu_api.cpp -- C-ABI implementation (see include/femgpu.h).
//
// Host side of the solve path: keeps the user's problem (user node/element
// order), the planned device mesh (internal order, femgpu_internal.h), the
// device vectors, and drives the HIP kernels of kernels.hip on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

#include "femgpu.h"
#include "batch.h"
#include "batch_ls.h"
#include "cg1.h"
#include "femgpu_internal.h"
#include "kernels.h"
#include "noalloc.h"

using femgpu::CgScalars;
using femgpu::DevMesh;
using femgpu::MeshPlan;
using femgpu::OpShift;

namespace {

constexpr int NODE_PART = 4 * femgpu::NODE_PARTS;  // partials of node kernels (4 arrays)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// Wavelet value exactly as ElementsData::updateWavelet (fem_utils.cpp:53-82):
// float intermediates, M_PI products formed in double.
float wavelet_value(float t, int type, float freq, float timeshift, float ampl) {
  const double PI = 3.14159265358979323846;
  if (t < timeshift) return 0.f;
  t = t - timeshift;
  if (type == FEMGPU_WAVELET_RICKER) {
    float tmp = float(PI * double(freq) * double(t) - PI);
    return ampl * (1.f - 2.f * tmp * tmp) * std::exp(-1.f * tmp * tmp);
  }
  if (type == FEMGPU_WAVELET_BERLAGE) {
    float w0 = float(2 * PI * double(freq));
    float w1 = w0 / std::sqrt(3.f);
    return ampl * w1 * w1 / 4.f * std::exp(-1.f * w1 * t) *
           (std::sin(w0 * t) * (1.f / (w1 * w1 * w1) + t / (w1 * w1) - t * t / w1) -
            std::cos(w0 * t) * std::sqrt(3.f) * (t * t / w1 + t / (w1 * w1)));
  }
  return 0.f;
}

// The same wavelets in double (FP64 state).
double wavelet_value_f64(double t, int type, double freq, double timeshift, double ampl) {
  const double PI = 3.14159265358979323846;
  if (t < timeshift) return 0.0;
  t = t - timeshift;
  if (type == FEMGPU_WAVELET_RICKER) {
    const double tmp = PI * freq * t - PI;
    return ampl * (1.0 - 2.0 * tmp * tmp) * std::exp(-tmp * tmp);
  }
  if (type == FEMGPU_WAVELET_BERLAGE) {
    const double w0 = 2 * PI * freq, w1 = w0 / std::sqrt(3.0);
    return ampl * w1 * w1 / 4.0 * std::exp(-w1 * t) *
           (std::sin(w0 * t) * (1.0 / (w1 * w1 * w1) + t / (w1 * w1) - t * t / w1) -
            std::cos(w0 * t) * std::sqrt(3.0) * (t * t / w1 + t / (w1 * w1)));
  }
  return 0.0;
}

}  // namespace

struct femgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  // problem, user order
  int dim = 0;
  int64_t nn = 0, ne = 0;
  std::vector<double> nodes;
  std::vector<int32_t> elems;
  double E = 2e7, nu = 0.3;
  std::vector<double> escale;
  std::vector<int32_t> cnode, ccode;
  std::vector<int32_t> fnodes, fadj;
  std::vector<double> fnormal, fp;
  std::vector<int32_t> lnode;
  std::vector<double> lvec;
  std::vector<double> extra_rhs;
  int conv = FEMGPU_LOADS_REFERENCE;
  int emax = 0;  // elements per patch (0: default)

  bool have_mesh = false;
  femgpu::BatchState *batch = nullptr;  // batched solve (batch.cpp)
  std::vector<int32_t> recv;            // receiver nodes (user ids)
  int recv_quantity = 0, recv_every = 1;
  std::vector<double> traces;           // last run: n_samples x n_recv x dim
  int n_samples = 0;
  MeshPlan plan;
  DevMesh dm;
  std::map<std::string, DevBuf> bufs;
  std::vector<uint8_t> fix_user;  // per user node
  std::vector<int32_t> ctype;     // per user node (last line wins)

  bool dirty_mat = true, dirty_fix = true, dirty_diag = true, dirty_rhs = true;

  CgScalars *h_sc = nullptr;  // pinned, h_sc[0..3]: poll ring
  hipEvent_t ev_poll[4] = {nullptr, nullptr, nullptr, nullptr};
  int emax_req = 0, threads_req = 0;
  int tuned_threads = 256;  // threads per patch of the current plan
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // graph cache for PCG iteration blocks, keyed by everything baked into
  // the captured launches (buffers, operator shift, block length)
  std::map<std::string, hipGraphExec_t> graphs;
  uint64_t buf_gen = 0;
  uint64_t graph_gen = 0;

  // femgpu_matvec_noalloc: this device's fixed scratch array (device
  // address from hipGetSymbolAddress) and the internal -> user node map
  void *scratch = nullptr;
  const int32_t *d_node_perm = nullptr;
};

namespace femgpu {
hipStream_t ctx_stream(femgpu_ctx *c) { return c->stream; }
int ctx_device(femgpu_ctx *c) { return c->device; }
int ctx_fail(femgpu_ctx *c, int code, const char *msg) {
  if (c) c->err = msg;
  return code;
}
BatchState *&ctx_batch(femgpu_ctx *c) { return c->batch; }
}  // namespace femgpu

namespace {

int fail(femgpu_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, FEMGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// The device plan of a mesh: patches of <= emax_req elements (0: 3 x
// threads) and <= threads nodes, one workgroup thread per patch node and at
// most three elements per thread.  The FP64 patch kernels must fit a
// workgroup's LDS (160 KiB on gfx950): patches with many nodes per element
// (thin or irregular meshes) are re-planned with fewer elements.
int plan_mesh(int dim, int64_t n_nodes, const double *nodes, int64_t n_elems, const int32_t *elems,
              int emax_req, int threads, MeshPlan &plan, std::string &msg) {
  int emax = emax_req > 0 ? emax_req : 3 * threads;
  if (emax > 3 * threads) {
    msg = "elements per patch above 3 x patch threads";
    return FEMGPU_E_ARG;
  }
  constexpr int LDS_MAX = 160 * 1024;
  for (;;) {
    msg = femgpu::build_plan(dim, n_nodes, nodes, n_elems, elems, emax, threads, true, plan);
    if (!msg.empty()) return FEMGPU_E_BAD_MESH;
    const int ent = (plan.ent_used + 3) & ~3;
    if (femgpu::lds_bytes(dim, 8, plan.pn_used, ent) <= LDS_MAX) return FEMGPU_OK;
    if (emax <= 64) {
      msg = "patch does not fit the LDS";
      return FEMGPU_E_BAD_MESH;
    }
    emax = std::max(64, emax * 7 / 8);
  }
}

void drop_graphs(femgpu_ctx *ctx) {
  for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
  ctx->graphs.clear();
}

template <typename T>
int dbuf(femgpu_ctx *ctx, const std::string &name, size_t count, T **out) {
  DevBuf &b = ctx->bufs[name];
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (b.bytes < bytes) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
      b.p = nullptr;
      return fail(ctx, FEMGPU_E_NOMEM, "hipMalloc failed for " + name);
    }
    b.bytes = bytes;
    ctx->buf_gen++;
  }
  *out = static_cast<T *>(b.p);
  return FEMGPU_OK;
}

template <typename T>
int upload(femgpu_ctx *ctx, const std::string &name, const std::vector<T> &h, T **out) {
  int st = dbuf<T>(ctx, name, h.size(), out);
  if (st) return st;
  if (!h.empty())
    HIPCHK(hipMemcpyAsync(*out, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return FEMGPU_OK;
}

// user node vector (nn x dim) -> internal order
std::vector<double> to_internal(const femgpu_ctx *ctx, const double *u) {
  const int d = ctx->dim;
  std::vector<double> v(size_t(ctx->nn) * d);
  for (int64_t i = 0; i < ctx->nn; ++i) {
    const int64_t s = ctx->plan.node_perm[i];
    for (int c = 0; c < d; ++c) v[size_t(i) * d + c] = u[size_t(s) * d + c];
  }
  return v;
}

void from_internal(const femgpu_ctx *ctx, const std::vector<double> &v, double *u) {
  const int d = ctx->dim;
  for (int64_t i = 0; i < ctx->nn; ++i) {
    const int64_t s = ctx->plan.node_perm[i];
    for (int c = 0; c < d; ++c) u[size_t(s) * d + c] = v[size_t(i) * d + c];
  }
}

void lame(int dim, double E, double nu, double &lam, double &mu) {
  mu = E / (2.0 * (1.0 + nu));
  lam = dim == 3 ? E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)) : E * nu / (1.0 - nu * nu);
}

int ensure_material(femgpu_ctx *ctx) {
  if (!ctx->dirty_mat) return FEMGPU_OK;
  lame(ctx->dim, ctx->E, ctx->nu, ctx->dm.lam, ctx->dm.mu);
  if (!ctx->escale.empty()) {
    std::vector<double> s64(ctx->ne);
    std::vector<float> s32(ctx->ne);
    for (int64_t i = 0; i < ctx->ne; ++i) {
      s64[i] = ctx->escale[ctx->plan.elem_perm[i]];
      s32[i] = float(s64[i]);
    }
    double *p64;
    float *p32;
    int st = upload(ctx, "escale64", s64, &p64);
    if (!st) st = upload(ctx, "escale32", s32, &p32);
    if (st) return st;
    ctx->dm.escale64 = p64;
    ctx->dm.escale32 = p32;
  } else {
    ctx->dm.escale64 = nullptr;
    ctx->dm.escale32 = nullptr;
  }
  ctx->dirty_mat = false;
  ctx->dirty_diag = true;
  drop_graphs(ctx);
  return FEMGPU_OK;
}

int ensure_fix(femgpu_ctx *ctx) {
  if (!ctx->dirty_fix) return FEMGPU_OK;
  std::vector<uint8_t> fm(ctx->nn, 0);
  for (int64_t i = 0; i < ctx->nn; ++i) fm[i] = ctx->fix_user[ctx->plan.node_perm[i]];
  uint8_t *p;
  int st = upload(ctx, "fixmask", fm, &p);
  if (st) return st;
  ctx->dm.fixmask = p;
  ctx->dirty_fix = false;
  return FEMGPU_OK;
}

// Assembled diagonal (FP64) and its inverse in FP64 and FP32.
int ensure_diag(femgpu_ctx *ctx) {
  int st = ensure_material(ctx);
  if (!st) st = ensure_fix(ctx);
  if (st) return st;
  if (!ctx->dirty_diag) return FEMGPU_OK;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  double *diag, *inv, *part;
  float *inv32, *diag32;
  if ((st = dbuf(ctx, "diag64", nd, &diag)) || (st = dbuf(ctx, "invD64", nd, &inv)) ||
      (st = dbuf(ctx, "invD32", nd, &inv32)) || (st = dbuf(ctx, "diag32", nd, &diag32)) ||
      (st = dbuf(ctx, "partial64", size_t(ctx->plan.n_slots) * ctx->dim, &part)))
    return st;
  femgpu::launch_fill<double>(diag, int64_t(nd), 0.0, ctx->stream);
  femgpu::launch_diag<double>(ctx->dm, diag, part, ctx->stream);
  femgpu::launch_inv_diag<double>(ctx->dm, diag, 1.0, 0.0, nullptr, inv, ctx->stream);
  femgpu::launch_convert<double, float>(inv, inv32, int64_t(nd), ctx->stream);
  femgpu::launch_convert<double, float>(diag, diag32, int64_t(nd), ctx->stream);
  HIPCHK(hipGetLastError());
  // the patch kernels read the assembled diagonal for the rows of
  // constrained DOFs
  ctx->dm.diag64 = diag;
  ctx->dm.diag32 = diag32;
  ctx->dirty_diag = false;
  return FEMGPU_OK;
}

// Global RHS in user order (FP64).  unit_pressure: evaluate every face with
// pressure 1 (the time-dependent part of the dynamics RHS).
void host_rhs(const femgpu_ctx *ctx, bool unit_pressure, std::vector<double> &f) {
  const int d = ctx->dim, npe = d + 1, nde = d * npe;
  f.assign(size_t(ctx->nn) * d, 0.0);
  const int64_t nf = int64_t(ctx->fadj.size());
  const int64_t nl = int64_t(ctx->lnode.size());
  if (ctx->conv == FEMGPU_LOADS_REFERENCE) {
    if (nl > 0) {
      // calcLoads: loads gathered into every element slot, overwriting the
      // pressures (cpu_fem_utils.cpp:588-596) -> assembled load x valence.
      std::vector<double> loads(size_t(ctx->nn) * d, 0.0);
      for (int64_t l = 0; l < nl; ++l)
        for (int c = 0; c < d; ++c) loads[size_t(ctx->lnode[l]) * d + c] = ctx->lvec[l * d + c];
      for (int64_t e = 0; e < ctx->ne; ++e)
        for (int a = 0; a < npe; ++a) {
          const int64_t n = ctx->elems[e * npe + a];
          for (int c = 0; c < d; ++c) f[size_t(n) * d + c] += loads[size_t(n) * d + c];
        }
    } else if (nf > 0) {
      // calcPressures (cpu_fem_utils.cpp:564-582): assignment per element slot.
      std::unordered_map<int64_t, double> slot;
      slot.reserve(size_t(nf) * nde);
      std::vector<int64_t> order;
      const double coeff = d == 2 ? -0.5 : -1.0 / 12.0;
      for (int64_t fc = 0; fc < nf; ++fc) {
        const int64_t el = ctx->fadj[fc];
        const double *p0 = &ctx->nodes[size_t(ctx->fnodes[fc * d + 0]) * d];
        const double *p1 = &ctx->nodes[size_t(ctx->fnodes[fc * d + 1]) * d];
        double len;
        if (d == 2) {
          len = std::sqrt((p1[0] - p0[0]) * (p1[0] - p0[0]) + (p1[1] - p0[1]) * (p1[1] - p0[1]));
        } else {
          const double *p2 = &ctx->nodes[size_t(ctx->fnodes[fc * d + 2]) * d];
          const double ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
          const double bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
          const double a1 = ay * bz - az * by, a2 = az * bx - ax * bz, a3 = ax * by - ay * bx;
          len = 0.5 * (a1 * a1 + a2 * a2 + a3 * a3);
        }
        const double pr = unit_pressure ? 1.0 : ctx->fp[fc];
        for (int i = 0; i < d; ++i) {
          const int32_t node = ctx->fnodes[fc * d + i];
          int loc = 0;
          for (int a = 0; a < npe; ++a)
            if (ctx->elems[el * npe + a] == node) {
              loc = a;
              break;
            }
          for (int j = 0; j < d; ++j) {
            const int64_t key = el * nde + d * loc + j;
            const double v =
                (ctx->ctype[node] & j) ? 0.0 : coeff * pr * len * ctx->fnormal[fc * d + j];
            if (!slot.count(key)) order.push_back(key);
            slot[key] = v;
          }
        }
      }
      std::sort(order.begin(), order.end());
      for (int64_t key : order) {
        const int64_t el = key / nde, k = key % nde;
        const int64_t n = ctx->elems[el * npe + k / d];
        f[size_t(n) * d + k % d] += slot[key];
      }
    }
  } else {
    // physical: F_node = -p A / d n per face node, duplicates once, summed.
    std::map<std::vector<int32_t>, int64_t> seen;
    for (int64_t fc = 0; fc < nf; ++fc) {
      std::vector<int32_t> key(ctx->fnodes.begin() + fc * d, ctx->fnodes.begin() + fc * d + d);
      std::sort(key.begin(), key.end());
      if (seen.count(key)) continue;
      seen[key] = fc;
      const double *p0 = &ctx->nodes[size_t(ctx->fnodes[fc * d + 0]) * d];
      const double *p1 = &ctx->nodes[size_t(ctx->fnodes[fc * d + 1]) * d];
      double A;
      if (d == 2) {
        A = std::sqrt((p1[0] - p0[0]) * (p1[0] - p0[0]) + (p1[1] - p0[1]) * (p1[1] - p0[1]));
      } else {
        const double *p2 = &ctx->nodes[size_t(ctx->fnodes[fc * d + 2]) * d];
        const double ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
        const double bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
        const double a1 = ay * bz - az * by, a2 = az * bx - ax * bz, a3 = ax * by - ay * bx;
        A = 0.5 * std::sqrt(a1 * a1 + a2 * a2 + a3 * a3);
      }
      const double pr = unit_pressure ? 1.0 : ctx->fp[fc];
      for (int i = 0; i < d; ++i) {
        const int32_t node = ctx->fnodes[fc * d + i];
        for (int j = 0; j < d; ++j) f[size_t(node) * d + j] += -pr * A / d * ctx->fnormal[fc * d + j];
      }
    }
    for (int64_t l = 0; l < nl; ++l)
      for (int c = 0; c < d; ++c) f[size_t(ctx->lnode[l]) * d + c] += ctx->lvec[l * d + c];
    for (int64_t n = 0; n < ctx->nn; ++n)
      for (int c = 0; c < d; ++c)
        if ((ctx->fix_user[n] >> c) & 1) f[size_t(n) * d + c] = 0.0;
  }
  if (!unit_pressure && !ctx->extra_rhs.empty())
    for (size_t i = 0; i < f.size(); ++i) f[i] += ctx->extra_rhs[i];
  // nodes no element references carry no equation
  for (int64_t i = ctx->plan.n_owned + ctx->plan.n_iface; i < ctx->nn; ++i)
    for (int c = 0; c < d; ++c) f[size_t(ctx->plan.node_perm[i]) * d + c] = 0.0;
}

int poll(femgpu_ctx *ctx, const CgScalars *d_sc) {
  HIPCHK(hipMemcpyAsync(ctx->h_sc, d_sc, sizeof(CgScalars), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return FEMGPU_OK;
}

// Workspace of one PCG precision.
template <typename T>
struct PcgWs {
  T *x, *r, *p, *q, *partial;
  const T *invD;
};

template <typename T>
int pcg_ws(femgpu_ctx *ctx, const char *tag, PcgWs<T> &w) {
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  std::string t(tag);
  int st;
  if ((st = dbuf(ctx, t + "x", nd, &w.x)) || (st = dbuf(ctx, t + "r", nd, &w.r)) ||
      (st = dbuf(ctx, t + "p", nd, &w.p)) ||
      (st = dbuf(ctx, t + "q", nd, &w.q)) ||
      (st = dbuf(ctx, t + "partial", size_t(ctx->plan.n_slots) * ctx->dim + 1, &w.partial)))
    return st;
  return FEMGPU_OK;
}

// One replay of a block of PCG iterations: a captured HIP graph, or, with
// FEMGPU_NO_GRAPHS=1 in the environment (for tracing tools that cannot
// follow graph launches), the same kernel launches issued directly.
struct Replay {
  hipGraphExec_t exec = nullptr;
  std::function<void()> eager;
  hipError_t launch(hipStream_t s) const {
    if (exec) return hipGraphLaunch(exec, s);
    eager();
    return hipGetLastError();
  }
};

bool graphs_disabled() {
  static const bool off = [] {
    const char *v = std::getenv("FEMGPU_NO_GRAPHS");
    return v && *v && *v != '0';
  }();
  return off;
}

// Graph of `block` fused PCG iterations on these buffers, the first at
// parity p0 (block even, so every replay starts at the same parity);
// captured once and cached.
template <typename T>
int pcg_graph(femgpu_ctx *ctx, PcgWs<T> &w, OpShift<T> sh, int block, int prec_tag,
              CgScalars *sc, double *part_patch, double *part_node, Replay *out, int p0 = 0) {
  if (graphs_disabled()) {
    femgpu_ctx *c = ctx;
    PcgWs<T> ws = w;
    out->exec = nullptr;
    out->eager = [=]() {
      for (int k = 0; k < block; ++k)
        femgpu::launch_pcg_iter<T>(c->dm, ws.x, ws.r, ws.p, ws.q, ws.partial, ws.invD, sh,
                                   sc, part_patch, part_node, (k + p0) & 1, c->stream);
    };
    return FEMGPU_OK;
  }
  if (ctx->graph_gen != ctx->buf_gen) {
    drop_graphs(ctx);
    ctx->graph_gen = ctx->buf_gen;
  }
  // the key names every pointer and value the captured launches hold
  // (besides the mesh, whose change drops every graph)
  char keybuf[320];
  std::snprintf(keybuf, sizeof(keybuf),
                "pcg/%d/%d/%d/%p/%p/%p/%p/%p/%p/%p/%p/%p/%p/%.17g/%.17g", prec_tag, block, p0,
                (const void *)w.x, (const void *)w.r, (const void *)w.p,
                (const void *)w.q, (const void *)w.partial, (const void *)w.invD, (const void *)sc,
                (const void *)sh.m, (const void *)part_patch, (const void *)part_node,
                double(sh.cK), double(sh.cM));
  const std::string key(keybuf);
  auto it = ctx->graphs.find(key);
  if (it != ctx->graphs.end()) {
    out->exec = it->second;
    return FEMGPU_OK;
  }
  hipGraph_t g;
  hipGraphExec_t exec = nullptr;
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < block; ++k)
    femgpu::launch_pcg_iter<T>(ctx->dm, w.x, w.r, w.p, w.q, w.partial, w.invD, sh, sc,
                               part_patch, part_node, (k + p0) & 1, ctx->stream);
  HIPCHK(hipStreamEndCapture(ctx->stream, &g));
  HIPCHK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
  (void)hipGraphDestroy(g);
  ctx->graphs[key] = exec;
  out->exec = exec;
  return FEMGPU_OK;
}

// Stop prediction of replay_until: the stop metric (gamma or r'r, whichever
// the stop test uses) and r'r of the last replays seen, their log reduction
// per iteration over a window of at least one full replay (r'r of CG is not
// monotone: shorter windows predict poorly), and the device's thresholds.
// A caller that continues the same iteration counter after a pause (the
// mixed solve after a residual replacement) passes it back in, reseeded
// with the new residual; the rates carry over.
struct ReplayPredict {
  static constexpr int W = 8;
  int n = 0;                        // replays in the window
  int it[W];                        // their iteration counts, oldest first
  double m[W], r[W];
  double lm = 0.0, lr = 0.0;        // log reduction per iteration
  double thr = 0.0, pause = 0.0;
  bool seen() const { return n > 0; }
  int iter() const { return n ? it[n - 1] : -1; }
  double m_last() const { return m[n - 1]; }
  double r_last() const { return r[n - 1]; }
  void seed(int iter, double m0, double r0) {
    n = 1;
    it[0] = iter;
    m[0] = m0;
    r[0] = r0;
  }
  void add(int iter, double mv, double rv, int span) {
    if (n == W) {
      for (int k = 1; k < W; ++k) {
        it[k - 1] = it[k];
        m[k - 1] = m[k];
        r[k - 1] = r[k];
      }
      --n;
    }
    it[n] = iter;
    m[n] = mv;
    r[n] = rv;
    ++n;
    int j = n - 2;
    while (j > 0 && iter - it[j] < span) --j;
    if (j < 0) return;
    const double dn = double(iter - it[j]);
    if (m[j] > 0.0 && mv > 0.0) lm = std::min(0.0, std::log(mv / m[j]) / dn);
    if (r[j] > 0.0 && rv > 0.0) lr = std::min(0.0, std::log(rv / r[j]) / dn);
  }
};

// Replays the iteration graph until the device-side stop flag is set or,
// when pause_host > 0, a replay ends with r'r below pause_host.  Replays are
// pipelined: replay k+1 is queued before the host looks at the
// scalars copied after replay k (the kernels of a finished solve exit at
// once, but each such launch still costs a few microseconds), except when
// the replay in flight is expected to stop -- its stop metric (gamma or r'r
// against the threshold, r'r against the pause level) extrapolated at the
// last replays' rate crosses the threshold, or the iteration cap falls
// inside it: then the next replay is launched only if the host sees it is
// still needed.  A replay expected to contain the device stop runs the
// short graph `tail` (tail_block iterations) when there is one, which cuts
// the no-op launches after the stop from up to block - 1 to about
// tail_block (fewer when the stop is foreseen).  Every replay
// either advances the iteration counter or sets the stop flag; anything
// else means the kernels did not run.  On return the stream is idle and
// h_sc[0] holds the last scalars.
int replay_until(femgpu_ctx *ctx, const Replay &exec, CgScalars *sc, int block, int max_iter,
                 double pause_host = 0.0, const Replay *tail = nullptr, int tail_block = 0,
                 ReplayPredict *pred = nullptr) {
  int st;
  if (!tail) tail_block = block;
  const long max_replays = long(max_iter) / std::min(block, tail_block) + 3;
  ReplayPredict local;
  ReplayPredict &P = pred ? *pred : local;
  int n_slot[2] = {0, 0};  // iterations of the replay in each slot
  auto launch = [&](int slot, int n) -> int {
    HIPCHK((n == block ? exec : *tail).launch(ctx->stream));
    HIPCHK(hipMemcpyAsync(&ctx->h_sc[slot], sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev_poll[slot], ctx->stream));
    n_slot[slot] = n;
    return FEMGPU_OK;
  };
  // the device stop within n more iterations, as far as the host can tell
  auto stop_within = [&](int n) {
    return P.seen() && (P.m_last() * std::exp(P.lm * n) <= P.thr ||
                        (P.pause > 0.0 && P.r_last() * std::exp(P.lr * n) < P.pause) ||
                        P.iter() + n >= max_iter);
  };
  auto pause_within = [&](int n) {
    return P.seen() && pause_host > 0.0 && P.r_last() * std::exp(P.lr * n) < pause_host;
  };
  // short replays from twice the horizon on: they cost a copy of the
  // scalars each, a full replay past the stop up to block - 1 no-op launches;
  // short too before the host has seen a replay of this solve (the first
  // two are launched unseen), so that a short solve does not run two full
  // replays of no-op launches
  auto size_for = [&](int horizon) {
    return !P.seen() || stop_within(2 * horizon) ? tail_block : block;
  };
  if ((st = launch(0, size_for(block)))) return st;
  int cur = 0;
  bool queued = false;  // replay nxt was launched before looking at replay cur
  for (long rep = 0;; ++rep) {
    const int nxt = (cur + 1) & 1;
    const bool near = stop_within(n_slot[cur]) || pause_within(n_slot[cur]);
    queued = !near;
    if (queued && (st = launch(nxt, size_for(n_slot[cur] + block)))) return st;
    HIPCHK(hipEventSynchronize(ctx->ev_poll[cur]));
    const CgScalars &h = ctx->h_sc[cur];
    if (h.done || (pause_host > 0 && h.rr_final < pause_host)) break;
    if (h.iter_final <= P.iter() || rep > max_replays)
      return fail(ctx, FEMGPU_E_HIP, "PCG made no progress (kernels not running)");
    P.add(h.iter_final, h.kind == 0 ? h.g_final : h.rr_final, h.rr_final, block);
    P.thr = h.thresh;
    P.pause = h.pause;
    if (near && (st = launch(nxt, size_for(block)))) return st;
    cur = nxt;
  }
  // the state after the last replay run: the queued one (which iterated on
  // after a host-side pause, or exited at once after a stop) when there is one
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->h_sc[0] = ctx->h_sc[queued ? (cur + 1) & 1 : cur];
  return FEMGPU_OK;
}

// Iterations of the short graph that replays expected to contain the stop run.
inline int tail_block_of(int block) { return std::max(2, (block / 4) & ~1); }

// Runs PCG until the stop flag is set.  r holds the initial residual; x and p
// are cleared here.
template <typename T>
int run_pcg(femgpu_ctx *ctx, PcgWs<T> &w, OpShift<T> sh, int kind, double tol, int max_iter,
            int block, int prec_tag, CgScalars *sc, int *iters, int *done_code) {
  double *part_patch, *part_node;
  int st;
  if ((st = dbuf(ctx, "part_patch", size_t(ctx->dm.np) + 1, &part_patch)) ||
      (st = dbuf(ctx, "part_node", NODE_PART, &part_node)))
    return st;
  const int64_t nd = ctx->nn * ctx->dim;
  femgpu::launch_pcg_start<T>(ctx->dm, w.x, w.p, nd, w.r, w.invD, sc, kind, tol, max_iter,
                              part_node, ctx->stream);
  HIPCHK(hipGetLastError());
  block = std::max(2, block + (block & 1));
  const int tb = tail_block_of(block);
  Replay exec, tail;
  if ((st = pcg_graph<T>(ctx, w, sh, block, prec_tag, sc, part_patch, part_node, &exec)) ||
      (st = pcg_graph<T>(ctx, w, sh, tb, prec_tag, sc, part_patch, part_node, &tail)) ||
      (st = replay_until(ctx, exec, sc, block, max_iter, 0.0, &tail, tb)))
    return st;
  femgpu::launch_pcg_apply_x<T>(ctx->dm, w.x, w.p, sc, ctx->stream);
  HIPCHK(hipGetLastError());
  *iters = ctx->h_sc->iter_final;
  *done_code = ctx->h_sc->done;
  return FEMGPU_OK;
}

// Single-reduction PCG (cg1.hip): device state of one precision.
template <typename T>
int cg1_ws(femgpu_ctx *ctx, const char *tag, femgpu::Cg1State<T> &w) {
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  const size_t ns = size_t(ctx->plan.n_slots) * ctx->dim + 1;
  const std::string t = std::string("cg1") + tag;
  int st;
  for (int k = 0; k < 2; ++k) {
    const std::string i = std::to_string(k);
    if ((st = dbuf(ctx, t + "x" + i, nd, &w.x[k])) || (st = dbuf(ctx, t + "r" + i, nd, &w.r[k])) ||
        (st = dbuf(ctx, t + "p" + i, nd, &w.p[k])) || (st = dbuf(ctx, t + "s" + i, nd, &w.s[k])) ||
        (st = dbuf(ctx, t + "slot" + i, ns, &w.slot[k])))
      return st;
  }
  if ((st = dbuf(ctx, t + "w", nd, &w.w)) ||
      (st = dbuf(ctx, "cg1part", size_t(2 * femgpu::CG1_PARTS) * std::max(ctx->dm.np, 1),
                 &w.part)))
    return st;
  return FEMGPU_OK;
}

// Graph of `block` single-reduction iterations (block even: every replay
// starts at parity 0); captured once per buffer set and cached.
template <typename T>
int cg1_graph(femgpu_ctx *ctx, const femgpu::Cg1State<T> &w, const T *invD, int block,
              int prec_tag, CgScalars *sc, Replay *out) {
  if (graphs_disabled()) {
    femgpu_ctx *c = ctx;
    const femgpu::Cg1State<T> ws = w;
    out->exec = nullptr;
    out->eager = [=]() {
      for (int k = 0; k < block; ++k) femgpu::launch_cg1_iter<T>(c->dm, ws, invD, sc, k & 1, c->stream);
    };
    return FEMGPU_OK;
  }
  if (ctx->graph_gen != ctx->buf_gen) {
    drop_graphs(ctx);
    ctx->graph_gen = ctx->buf_gen;
  }
  // the key names every pointer the captured launches hold (besides the
  // mesh, whose change drops every graph)
  std::string key = "cg1/" + std::to_string(prec_tag) + "/" + std::to_string(block);
  {
    char b[32];
    const void *ptrs[] = {w.x[0], w.x[1], w.r[0], w.r[1], w.p[0], w.p[1], w.s[0], w.s[1],
                          w.w,    w.slot[0], w.slot[1], w.part, invD, sc};
    for (const void *q : ptrs) {
      std::snprintf(b, sizeof(b), "/%p", q);
      key += b;
    }
  }
  auto it = ctx->graphs.find(key);
  if (it != ctx->graphs.end()) {
    out->exec = it->second;
    return FEMGPU_OK;
  }
  hipGraph_t g;
  hipGraphExec_t exec = nullptr;
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < block; ++k)
    femgpu::launch_cg1_iter<T>(ctx->dm, w, invD, sc, k & 1, ctx->stream);
  HIPCHK(hipStreamEndCapture(ctx->stream, &g));
  HIPCHK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
  (void)hipGraphDestroy(g);
  ctx->graphs[key] = exec;
  out->exec = exec;
  return FEMGPU_OK;
}

// Iterations done and the parity of the state once the stop flag is set:
// the launch that sets it does no iteration and its parity is the state's.
struct Cg1Pos {
  int iters, q;
};
Cg1Pos cg1_pos(const CgScalars &h) { return {h.iter_final, h.par_final}; }

// Runs the single-reduction PCG from r[0] with x[0], p[0], s[0] cleared
// here; on return x[q] holds the solution.
template <typename T>
int run_cg1(femgpu_ctx *ctx, femgpu::Cg1State<T> &w, const T *invD, int kind, double tol,
            int max_iter, int block, int prec_tag, CgScalars *sc, Cg1Pos *pos, int *done_code) {
  const int64_t nd = ctx->nn * ctx->dim;
  femgpu::launch_fill<T>(w.x[0], nd, T(0), ctx->stream);
  femgpu::launch_fill<T>(w.p[0], nd, T(0), ctx->stream);
  femgpu::launch_fill<T>(w.s[0], nd, T(0), ctx->stream);
  femgpu::launch_reset_scalars(sc, kind, tol, max_iter, ctx->stream);
  femgpu::launch_cg1_init<T>(ctx->dm, w, invD, ctx->stream);
  femgpu::launch_cg1_fin(sc, w.part, ctx->dm.np, 0, 0, 0.0, ctx->stream);
  HIPCHK(hipGetLastError());
  block = std::max(2, block + (block & 1));
  const int tb = tail_block_of(block);
  Replay exec, tail;
  int st;
  if ((st = cg1_graph<T>(ctx, w, invD, block, prec_tag, sc, &exec)) ||
      (st = cg1_graph<T>(ctx, w, invD, tb, prec_tag, sc, &tail)) ||
      (st = replay_until(ctx, exec, sc, block, max_iter, 0.0, &tail, tb)))
    return st;
  *pos = cg1_pos(*ctx->h_sc);
  *done_code = ctx->h_sc->done;
  return FEMGPU_OK;
}

int done_status(int done) {
  switch (done) {
    case 1: return FEMGPU_OK;
    case 2: return FEMGPU_E_NONFINITE;
    case 3: return FEMGPU_E_BREAKDOWN;
    case 4: return FEMGPU_E_MAXIT;
    default: return FEMGPU_E_STATE;
  }
}

int need_mesh(femgpu_ctx *ctx) {
  if (!ctx) return FEMGPU_E_ARG;
  if (!ctx->have_mesh) return fail(ctx, FEMGPU_E_STATE, "no mesh set");
  HIPCHK(hipSetDevice(ctx->device));
  return FEMGPU_OK;
}

// True relative residual ||f - K x|| / ||f|| in FP64 (x, f internal, device).
int true_rel_res(femgpu_ctx *ctx, const double *f, const double *x, double *out) {
  double *y, *r, *part, *part_node;
  CgScalars *sc;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  int st;
  if ((st = dbuf(ctx, "rr_y", nd, &y)) || (st = dbuf(ctx, "rr_r", nd, &r)) ||
      (st = dbuf(ctx, "partial64", size_t(ctx->plan.n_slots) * ctx->dim + 1, &part)) ||
      (st = dbuf(ctx, "part_node", NODE_PART, &part_node)) ||
      (st = dbuf(ctx, "sc_aux", 1, &sc)))
    return st;
  // ||f||^2 first: the residual of K x replaced by 0 (y and slots zero).
  femgpu::launch_reset_scalars(sc, 1, 0.0, 0, ctx->stream);
  femgpu::launch_fill<double>(y, int64_t(nd), 0.0, ctx->stream);
  femgpu::launch_fill<double>(part, int64_t(ctx->plan.n_slots) * ctx->dim, 0.0, ctx->stream);
  femgpu::launch_residual64(ctx->dm, f, y, part, r, sc, part_node, ctx->stream);
  HIPCHK(hipGetLastError());
  if ((st = poll(ctx, sc))) return st;
  const double ff = ctx->h_sc->rr;
  femgpu::launch_matvec_patch<double>(ctx->dm, x, y, part, ctx->stream);
  femgpu::launch_residual64(ctx->dm, f, y, part, r, sc, part_node, ctx->stream);
  HIPCHK(hipGetLastError());
  if ((st = poll(ctx, sc))) return st;
  *out = ff > 0 ? std::sqrt(ctx->h_sc->rr / ff) : 0.0;
  return FEMGPU_OK;
}

}  // namespace

namespace {

// femgpu_matvec_noalloc in arithmetic T on the scratch array.
template <typename T>
int matvec_noalloc_t(femgpu_ctx *ctx, const double *x, double *y) {
  const int64_t nd = ctx->nn * ctx->dim;
  const size_t vec = (size_t(nd) * sizeof(T) + 255) & ~size_t(255);
  const size_t need = 2 * vec + (size_t(ctx->plan.n_slots) * ctx->dim + 1) * sizeof(T);
  if (need > femgpu::NOALLOC_SCRATCH_BYTES)
    return fail(ctx, FEMGPU_E_ARG,
                "problem too large for femgpu_matvec_noalloc: needs " + std::to_string(need) +
                    " bytes of scratch, the limit is " +
                    std::to_string(femgpu::NOALLOC_SCRATCH_BYTES));
  unsigned char *base = static_cast<unsigned char *>(ctx->scratch);
  T *xi = reinterpret_cast<T *>(base);
  T *yi = reinterpret_cast<T *>(base + vec);
  T *part = reinterpret_cast<T *>(base + 2 * vec);
  femgpu::launch_gather_user<T>(ctx->nn, ctx->dim, ctx->d_node_perm, x, xi, ctx->stream);
  femgpu::launch_matvec<T>(ctx->dm, xi, yi, part, ctx->stream);
  femgpu::launch_scatter_user<T>(ctx->nn, ctx->dm.n_active, ctx->dim, ctx->d_node_perm, yi, y,
                                 ctx->stream);
  HIPCHK(hipGetLastError());
  return FEMGPU_OK;
}
}  // namespace

static int solve_core(femgpu_ctx *ctx, const femgpu_solve_opts *opts, const char *f_name,
                      double *u_out, femgpu_stats *stats);
static bool trace_enabled();

namespace {

// The device mesh of ctx->plan (planned with `ctx->tuned_threads` threads per
// patch): patch records, node runs, interface slots, coordinates; then the
// operator's per-node data is rebuilt lazily from the user-order problem
// (material, constraint mask, diagonal, loads).
int upload_plan(femgpu_ctx *ctx) {
  const int threads = ctx->tuned_threads;
  const int dim = ctx->dim;
  const int64_t n_nodes = ctx->nn;
  const double *nodes = ctx->nodes.data();
  const MeshPlan &P = ctx->plan;
  DevMesh &dm = ctx->dm;
  dm = DevMesh();
  dm.dim = dim;
  dm.np = int(P.patches.size());
  dm.emax = P.emax_used;
  dm.ent = (P.ent_used + 3) & ~3;
  dm.pn = P.pn_used;
  dm.nn = P.nn;
  dm.ne = P.ne;
  dm.n_owned = P.n_owned;
  dm.n_iface = P.n_iface;
  dm.n_active = P.n_owned + P.n_iface;
  dm.patch_threads = threads;
  if (dm.pn > threads || dm.emax > 3 * threads)
    return fail(ctx, FEMGPU_E_ARG, "internal error: patch larger than the workgroup");
  int st;
  femgpu::PatchInfo *pp;
  uint32_t *rec;
  int32_t *shg, *shs, *ss, *ei;
  uint16_t *ne_;
  std::vector<uint32_t> recs;
  femgpu::pack_records(P, recs);
  // records at stride emax per patch for the patch kernels: the same array
  // (padded by one patch) when every patch starts at p * emax
  std::vector<uint32_t> recs_s;
  const bool regular = femgpu::stride_records(P, recs, recs_s);
  if ((st = upload(ctx, "patches", P.patches, &pp)) || (st = upload(ctx, "erec", recs, &rec)) ||
      (st = upload(ctx, "sh_gid", P.sh_gid, &shg)) || (st = upload(ctx, "sh_slot", P.sh_slot, &shs)) ||
      (st = upload(ctx, "node_end", P.node_end, &ne_)) ||
      (st = upload(ctx, "slot_start", P.slot_start, &ss)) ||
      (st = upload(ctx, "elems_int", P.elems_int, &ei)))
    return st;
  {
    int32_t *np_;
    if ((st = upload(ctx, "node_perm", P.node_perm, &np_))) return st;
    ctx->d_node_perm = np_;
  }
  dm.patches = pp;
  dm.erec = rec;
  dm.erec_s = rec;
  if (!regular) {
    uint32_t *rs;
    if ((st = upload(ctx, "erec_s", recs_s, &rs))) return st;
    dm.erec_s = rs;
  }
  dm.sh_gid = shg;
  dm.sh_slot = shs;
  {
    // per patch interface node: id, own slot, the node's slot range
    std::vector<int32_t> info(P.sh_gid.size() * 4);
    for (size_t i = 0; i < P.sh_gid.size(); ++i) {
      const int64_t s = P.sh_gid[i] - P.n_owned;
      info[4 * i + 0] = P.sh_gid[i];
      info[4 * i + 1] = P.sh_slot[i];
      info[4 * i + 2] = P.slot_start[s];
      info[4 * i + 3] = P.slot_start[s + 1];
    }
    int32_t *pi;
    if ((st = upload(ctx, "sh_info", info, &pi))) return st;
    dm.sh_info = reinterpret_cast<const int4 *>(pi);
  }
  dm.node_end = ne_;
  {
    std::vector<int32_t> pgid, pslot;
    std::vector<uint16_t> pend;
    femgpu::pack_patch_nodes(P, threads, pgid, pslot, pend);
    int32_t *pg_, *ps_;
    uint16_t *pe_;
    if ((st = upload(ctx, "pgid", pgid, &pg_)) || (st = upload(ctx, "pslot", pslot, &ps_)) ||
        (st = upload(ctx, "pend", pend, &pe_)))
      return st;
    dm.pgid = pg_;
    dm.pslot = ps_;
    dm.pend = pe_;
  }
  dm.slot_start = ss;
  dm.elems_int = ei;
  std::vector<double> c64(size_t(n_nodes) * dim);
  std::vector<float> c32(c64.size());
  for (int64_t i = 0; i < n_nodes; ++i)
    for (int c = 0; c < dim; ++c) {
      c64[size_t(i) * dim + c] = nodes[size_t(P.node_perm[i]) * dim + c];
      c32[size_t(i) * dim + c] = float(c64[size_t(i) * dim + c]);
    }
  double *pc64;
  float *pc32;
  if ((st = upload(ctx, "coords64", c64, &pc64)) || (st = upload(ctx, "coords32", c32, &pc32)))
    return st;
  dm.coords64 = pc64;
  dm.coords32 = pc32;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->have_mesh = true;
  ctx->dirty_mat = ctx->dirty_fix = ctx->dirty_diag = ctx->dirty_rhs = true;
  return ensure_fix(ctx);
}

}  // namespace

extern "C" {

int femgpu_create(femgpu_ctx **out, int device_ordinal) {
  if (!out) return FEMGPU_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FEMGPU_E_HIP;
  if (device_ordinal < 0 || device_ordinal >= n) return FEMGPU_E_ARG;
  std::unique_ptr<femgpu_ctx> c(new femgpu_ctx());
  femgpu_ctx *ctx = c.get();
  ctx->device = device_ordinal;
  HIPCHK(hipSetDevice(device_ordinal));
  HIPCHK(femgpu::noalloc_scratch_address(device_ordinal, &ctx->scratch));
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  ctx->own_stream = true;
  HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_sc), 4 * sizeof(CgScalars), 0));
  for (auto &e : ctx->ev_poll) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&ctx->ev0));
  HIPCHK(hipEventCreate(&ctx->ev1));
  HIPCHK(femgpu::prepare_kernels());
  HIPCHK(femgpu::prepare_cg1_kernels());
  HIPCHK(femgpu::prepare_bl_kernels());
  *out = c.release();
  return FEMGPU_OK;
}

void femgpu_destroy(femgpu_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graphs(ctx);
  femgpu::batch_free(ctx->batch);
  for (auto &kv : ctx->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  if (ctx->h_sc) (void)hipHostFree(ctx->h_sc);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (auto &e : ctx->ev_poll)
    if (e) (void)hipEventDestroy(e);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *femgpu_last_error(const femgpu_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int femgpu_set_stream(femgpu_ctx *ctx, void *hip_stream) {
  if (!ctx) return FEMGPU_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
  ctx->own_stream = false;
  if (hip_stream) {
    ctx->stream = static_cast<hipStream_t>(hip_stream);
  } else {
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  drop_graphs(ctx);
  return FEMGPU_OK;
}

int femgpu_sync(femgpu_ctx *ctx) {
  if (!ctx) return FEMGPU_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return FEMGPU_OK;
}

int femgpu_set_mesh(femgpu_ctx *ctx, int dim, int64_t n_nodes, const double *nodes,
                    int64_t n_elems, const int32_t *elems) {
  if (!ctx || !nodes || !elems) return FEMGPU_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  ctx->have_mesh = false;
  const int npe = dim + 1;
  // one workgroup thread per patch node and at most EPT = 3 elements per
  // thread: patches of <= 3 * threads elements and <= threads nodes
  // Default 256 threads (768-element patches): on gfx950 the fastest PCG
  // iteration in FP64 (no register spills, three workgroups per CU) and
  // within a few percent of the fastest in FP32 among 256/512/1024.
  const int threads = ctx->threads_req ? ctx->threads_req : 256;
  {
    std::string msg;
    const int st = plan_mesh(dim, n_nodes, nodes, n_elems, elems, ctx->emax, threads, ctx->plan, msg);
    if (st) return fail(ctx, st, msg);
  }
  // degenerate elements
  for (int64_t e = 0; e < n_elems; ++e) {
    const int32_t *en = &elems[e * npe];
    double det;
    if (dim == 2) {
      const double *a = &nodes[size_t(en[0]) * 2], *b = &nodes[size_t(en[1]) * 2],
                   *c = &nodes[size_t(en[2]) * 2];
      det = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
    } else {
      const double *a = &nodes[size_t(en[0]) * 3], *b = &nodes[size_t(en[1]) * 3],
                   *c = &nodes[size_t(en[2]) * 3], *d = &nodes[size_t(en[3]) * 3];
      const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      const double e3[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
      det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
            e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    }
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
      return fail(ctx, FEMGPU_E_BAD_MESH, "degenerate element " + std::to_string(e));
  }
  ctx->dim = dim;
  ctx->nn = n_nodes;
  ctx->ne = n_elems;
  ctx->nodes.assign(nodes, nodes + size_t(n_nodes) * dim);
  ctx->elems.assign(elems, elems + size_t(n_elems) * npe);
  ctx->escale.clear();
  ctx->cnode.clear();
  ctx->ccode.clear();
  ctx->fnodes.clear();
  ctx->fadj.clear();
  ctx->fnormal.clear();
  ctx->fp.clear();
  ctx->lnode.clear();
  ctx->lvec.clear();
  ctx->recv.clear();
  ctx->traces.clear();
  ctx->n_samples = 0;
  ctx->extra_rhs.clear();
  ctx->fix_user.assign(n_nodes, 0);
  ctx->ctype.assign(n_nodes, 0);

  ctx->tuned_threads = threads;
  return upload_plan(ctx);
}

int femgpu_set_tuning(femgpu_ctx *ctx, int elems_per_patch, int patch_threads) {
  if (!ctx) return FEMGPU_E_ARG;
  if (elems_per_patch != 0 && (elems_per_patch < 16 || elems_per_patch > femgpu::EMAX_LIMIT))
    return fail(ctx, FEMGPU_E_ARG, "elems_per_patch out of range");
  if (patch_threads != 0 && patch_threads != 256 && patch_threads != 512 && patch_threads != 1024)
    return fail(ctx, FEMGPU_E_ARG, "patch_threads must be 256, 512 or 1024");
  const int threads = patch_threads ? patch_threads : 256;
  if (elems_per_patch > 3 * threads)
    return fail(ctx, FEMGPU_E_ARG, "elems_per_patch above 3 x patch_threads");
  // takes effect at the next femgpu_set_mesh (the plan depends on both)
  ctx->emax = elems_per_patch;
  ctx->threads_req = patch_threads;
  return FEMGPU_OK;
}

int femgpu_plan_info(int dim, int64_t n_nodes, const double *nodes, int64_t n_elems,
                     const int32_t *elems, int elems_per_patch, int patch_threads, int64_t *info) {
  if (!nodes || !elems || !info || (dim != 2 && dim != 3) || n_nodes <= 0 || n_elems <= 0)
    return FEMGPU_E_ARG;
  if (patch_threads != 0 && patch_threads != 256 && patch_threads != 512 && patch_threads != 1024)
    return FEMGPU_E_ARG;
  if (elems_per_patch != 0 && (elems_per_patch < 16 || elems_per_patch > femgpu::EMAX_LIMIT))
    return FEMGPU_E_ARG;
  MeshPlan plan;
  std::string msg;
  const int st = plan_mesh(dim, n_nodes, nodes, n_elems, elems, elems_per_patch,
                           patch_threads ? patch_threads : 256, plan, msg);
  if (st) return st;
  info[0] = int64_t(plan.patches.size());
  info[1] = plan.n_owned;
  info[2] = plan.n_iface;
  info[3] = plan.n_slots;
  info[4] = plan.emax;
  info[5] = plan.pn_used;
  return FEMGPU_OK;
}

}  // extern "C"

namespace {

// Autotuner: candidate patch tunings (threads per patch, elements per patch),
// the process-wide cache of choices per (mesh signature, arithmetic), and
// the mesh signature (FNV-1a over dim, sizes, coordinates and connectivity).
constexpr int TUNE_CAND[][2] = {{256, 768}, {256, 640}, {256, 512}, {256, 384},
                                {512, 1536}, {512, 1024}, {512, 768}, {1024, 3072},
                                {1024, 2048}};
constexpr int TUNE_N = int(sizeof(TUNE_CAND) / sizeof(TUNE_CAND[0]));

struct TuneChoice {
  int threads, elems;
  double ms;
};
std::mutex g_tune_mu;
std::map<std::pair<uint64_t, int>, TuneChoice> g_tune_cache;

uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

uint64_t mesh_signature(const femgpu_ctx *ctx) {
  uint64_t h = 1469598103934665603ULL;
  const int64_t head[3] = {ctx->dim, ctx->nn, ctx->ne};
  h = fnv1a(h, head, sizeof(head));
  h = fnv1a(h, ctx->nodes.data(), ctx->nodes.size() * sizeof(double));
  return fnv1a(h, ctx->elems.data(), ctx->elems.size() * sizeof(int32_t));
}

// Re-plan the current mesh with a given tuning, keeping the problem.
int replan(femgpu_ctx *ctx, MeshPlan &&plan, int threads) {
  HIPCHK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  ctx->plan = std::move(plan);
  ctx->tuned_threads = threads;
  ctx->have_mesh = false;
  int st = upload_plan(ctx);
  if (st) return st;
  ctx->have_mesh = true;
  return FEMGPU_OK;
}

// Device time per PCG iteration of the current plan on a unit right-hand
// side, stopping rule out of reach: the difference between runs of 3 * iters
// and iters iterations (each the faster of two, after a run that captures
// the iteration graphs) over 2 * iters, so that the fixed cost of a solve
// (initialisation, the last replay's wait) does not count.
int time_pcg_iteration(femgpu_ctx *ctx, int precision, int iters, double *ms) {
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  double *f;
  int st;
  if ((st = dbuf(ctx, "tune_f", nd, &f))) return st;
  femgpu::launch_fill<double>(f, int64_t(nd), 1.0, ctx->stream);
  femgpu_solve_opts o;
  femgpu_default_solve_opts(&o);
  o.precision = precision == FEMGPU_FP64 ? FEMGPU_FP64 : FEMGPU_FP32;
  o.tol_kind = FEMGPU_TOL_REL_RES;
  o.tol = 0.0;
  // (a small problem's PCG may stop early, at a breakdown once its residual
  // is at rounding level: the runs then count the iterations they did)
  auto run = [&](int n, double *t, int *it) -> int {
    o.max_iter = n;
    femgpu_stats S{};
    const int rs = solve_core(ctx, &o, "tune_f", nullptr, &S);
    if (rs != FEMGPU_OK && rs != FEMGPU_E_MAXIT && rs != FEMGPU_E_BREAKDOWN) return rs;
    if (S.iterations < 1) return fail(ctx, FEMGPU_E_BREAKDOWN, "autotune: PCG did not iterate");
    *t = S.solve_ms;
    *it = S.iterations;
    return FEMGPU_OK;
  };
  double t1 = 0.0, t3 = 0.0, t;
  int i1 = 0, i3 = 0, it;
  if ((st = run(iters, &t, &it))) return st;
  for (int rep = 0; rep < 2; ++rep) {
    if ((st = run(iters, &t, &it))) return st;
    if (rep == 0 || t < t1) t1 = t, i1 = it;
    if ((st = run(3 * iters, &t, &it))) return st;
    if (rep == 0 || t < t3) t3 = t, i3 = it;
  }
  // without extra iterations in the longer run: the short run's average
  *ms = i3 > i1 ? std::max(t3 - t1, 1e-9) / (i3 - i1) : t1 / i1;
  return FEMGPU_OK;
}

// Device time per step of the time stepper with the current plan (see the
// definition).
int time_dyn_step(femgpu_ctx *ctx, const femgpu_dyn_opts &opts, int steps, double *us);

// The autotuners' common part: plans every candidate on host threads, times
// each with `timer` (its time in *t; FEMGPU_E_BREAKDOWN or FEMGPU_E_MAXIT
// skip the candidate), caches the fastest under `key` and leaves the
// context planned with it as its tuning.  A cached choice is only re-planned.
template <typename Timer>
int tune_candidates(femgpu_ctx *ctx, std::pair<uint64_t, int> key, Timer timer,
                    const char *unit, TuneChoice *out) {
  TuneChoice best{0, 0, 0.0};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      best = it->second;
      cached = true;
    }
  }
  int st;
  if (!cached) {
    std::vector<MeshPlan> plans(TUNE_N);
    std::vector<int> pst(TUNE_N, FEMGPU_OK);
    {
      std::vector<std::thread> th;
      for (int i = 0; i < TUNE_N; ++i)
        th.emplace_back([&, i] {
          std::string msg;
          pst[i] = plan_mesh(ctx->dim, ctx->nn, ctx->nodes.data(), ctx->ne, ctx->elems.data(),
                             TUNE_CAND[i][1], TUNE_CAND[i][0], plans[i], msg);
        });
      for (auto &t : th) t.join();
    }
    for (int i = 0; i < TUNE_N; ++i) {
      if (pst[i] != FEMGPU_OK) continue;
      if ((st = replan(ctx, std::move(plans[i]), TUNE_CAND[i][0]))) return st;
      double t = 0.0;
      st = timer(&t);
      if (st == FEMGPU_E_BREAKDOWN || st == FEMGPU_E_MAXIT) continue;
      if (st) return st;
      if (trace_enabled())
        std::fprintf(stderr, "femgpu autotune: %d threads, %d elements per patch: %.4f %s\n",
                     TUNE_CAND[i][0], TUNE_CAND[i][1], t, unit);
      if (best.threads == 0 || t < best.ms) best = {TUNE_CAND[i][0], TUNE_CAND[i][1], t};
    }
    if (best.threads == 0) return fail(ctx, FEMGPU_E_STATE, "autotune: no candidate ran");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[key] = best;
  }
  ctx->emax = best.elems;
  ctx->threads_req = best.threads;
  MeshPlan plan;
  std::string msg;
  if ((st = plan_mesh(ctx->dim, ctx->nn, ctx->nodes.data(), ctx->ne, ctx->elems.data(), best.elems,
                      best.threads, plan, msg)))
    return fail(ctx, st, msg);
  if ((st = replan(ctx, std::move(plan), best.threads))) return st;
  *out = best;
  return FEMGPU_OK;
}

}  // namespace

extern "C" {

int femgpu_autotune(femgpu_ctx *ctx, int precision, int *elems_per_patch, int *patch_threads,
                    double *ms_per_iteration) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (precision != FEMGPU_FP64 && precision != FEMGPU_FP32 && precision != FEMGPU_MIXED)
    return fail(ctx, FEMGPU_E_ARG, "unknown precision");
  const int arith = precision == FEMGPU_FP64 ? FEMGPU_FP64 : FEMGPU_FP32;
  TuneChoice best{};
  if ((st = tune_candidates(
           ctx, std::make_pair(mesh_signature(ctx), arith),
           [&](double *t) { return time_pcg_iteration(ctx, arith, 32, t); }, "ms/iteration",
           &best)))
    return st;
  if (elems_per_patch) *elems_per_patch = best.elems;
  if (patch_threads) *patch_threads = best.threads;
  if (ms_per_iteration) *ms_per_iteration = best.ms;
  return FEMGPU_OK;
}

int femgpu_autotune_dyn(femgpu_ctx *ctx, const femgpu_dyn_opts *opts, int n_steps,
                        int *elems_per_patch, int *patch_threads, double *us_per_step) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!opts) return FEMGPU_E_ARG;
  if (!(opts->dt > 0) || !(opts->rho > 0)) return fail(ctx, FEMGPU_E_ARG, "dt and rho must be > 0");
  if (opts->precision != FEMGPU_FP64 && opts->precision != FEMGPU_FP32)
    return fail(ctx, FEMGPU_E_ARG, "dynamics precision must be FP64 or FP32");
  if (n_steps < 0) return fail(ctx, FEMGPU_E_ARG, "n_steps must be >= 0");
  const bool implicit = opts->beta2 != 0.0;
  const int steps = n_steps ? n_steps : implicit ? 16 : 256;
  // cached apart from the PCG choices: per stepper kind and precision
  const int kind = 100 + (implicit ? 10 : 0) + opts->precision;
  TuneChoice best{};
  if ((st = tune_candidates(
           ctx, std::make_pair(mesh_signature(ctx), kind),
           [&](double *t) { return time_dyn_step(ctx, *opts, steps, t); }, "us/step", &best)))
    return st;
  if (elems_per_patch) *elems_per_patch = best.elems;
  if (patch_threads) *patch_threads = best.threads;
  if (us_per_step) *us_per_step = best.ms;
  return FEMGPU_OK;
}

int femgpu_set_material(femgpu_ctx *ctx, double E, double nu, const double *elem_scale) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!(E > 0) || !(nu > -1.0) || !(nu < (ctx->dim == 3 ? 0.5 : 1.0)))
    return fail(ctx, FEMGPU_E_ARG, "invalid E / nu");
  ctx->E = E;
  ctx->nu = nu;
  if (elem_scale) {
    ctx->escale.assign(elem_scale, elem_scale + ctx->ne);
    for (double v : ctx->escale)
      if (!(v > 0) || !std::isfinite(v)) return fail(ctx, FEMGPU_E_ARG, "elem_scale must be > 0");
  } else {
    ctx->escale.clear();
  }
  ctx->dirty_mat = true;
  return FEMGPU_OK;
}

int femgpu_set_constraints(femgpu_ctx *ctx, int64_t n, const int32_t *node, const int32_t *code) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (n < 0 || (n > 0 && (!node || !code))) return FEMGPU_E_ARG;
  ctx->fix_user.assign(ctx->nn, 0);
  ctx->ctype.assign(ctx->nn, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (node[i] < 0 || node[i] >= ctx->nn) return fail(ctx, FEMGPU_E_ARG, "constraint node out of range");
    const int32_t t = code[i];
    ctx->ctype[node[i]] = t;
    uint8_t m = 0;
    if (t & 1) m |= 1;
    if (t & 2) m |= 2;
    if (ctx->dim == 3 && (t & 3)) m |= 4;
    ctx->fix_user[node[i]] |= m;
  }
  ctx->cnode.assign(node, node + n);
  ctx->ccode.assign(code, code + n);
  ctx->dirty_fix = true;
  ctx->dirty_diag = true;
  ctx->dirty_rhs = true;
  return FEMGPU_OK;
}

int femgpu_set_pressure_faces(femgpu_ctx *ctx, int64_t n, const int32_t *face_nodes,
                              const int32_t *adj_elem, const double *normal, const double *pressure) {
  int st = need_mesh(ctx);
  if (st) return st;
  const int d = ctx->dim, npe = d + 1;
  if (n < 0 || (n > 0 && (!face_nodes || !adj_elem || !normal || !pressure))) return FEMGPU_E_ARG;
  ctx->fnodes.assign(face_nodes, face_nodes + n * d);
  ctx->fadj.assign(adj_elem, adj_elem + n);
  ctx->fp.assign(pressure, pressure + n);
  ctx->fnormal.resize(size_t(n) * d);
  for (int64_t f = 0; f < n; ++f) {
    if (adj_elem[f] < 0 || adj_elem[f] >= ctx->ne) return fail(ctx, FEMGPU_E_BAD_MESH, "face element out of range");
    for (int i = 0; i < d; ++i) {
      const int32_t v = face_nodes[f * d + i];
      bool on = false;
      for (int a = 0; a < npe; ++a) on |= ctx->elems[size_t(adj_elem[f]) * npe + a] == v;
      if (!on) return fail(ctx, FEMGPU_E_BAD_MESH, "face node not on its element");
    }
    double l2 = 0;
    for (int j = 0; j < d; ++j) l2 += normal[f * d + j] * normal[f * d + j];
    if (!(l2 > 0)) return fail(ctx, FEMGPU_E_BAD_MESH, "zero face normal");
    for (int j = 0; j < d; ++j) ctx->fnormal[f * d + j] = normal[f * d + j] / std::sqrt(l2);
  }
  ctx->dirty_rhs = true;
  return FEMGPU_OK;
}

int femgpu_set_nodal_loads(femgpu_ctx *ctx, int64_t n, const int32_t *node, const double *vec) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (n < 0 || (n > 0 && (!node || !vec))) return FEMGPU_E_ARG;
  for (int64_t i = 0; i < n; ++i)
    if (node[i] < 0 || node[i] >= ctx->nn) return fail(ctx, FEMGPU_E_ARG, "load node out of range");
  ctx->lnode.assign(node, node + n);
  ctx->lvec.assign(vec, vec + n * ctx->dim);
  ctx->dirty_rhs = true;
  return FEMGPU_OK;
}

int femgpu_set_load_convention(femgpu_ctx *ctx, int convention) {
  if (!ctx) return FEMGPU_E_ARG;
  if (convention != FEMGPU_LOADS_REFERENCE && convention != FEMGPU_LOADS_PHYSICAL)
    return fail(ctx, FEMGPU_E_ARG, "unknown load convention");
  ctx->conv = convention;
  ctx->dirty_rhs = true;
  return FEMGPU_OK;
}

int femgpu_set_extra_rhs(femgpu_ctx *ctx, const double *f) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (f)
    ctx->extra_rhs.assign(f, f + size_t(ctx->nn) * ctx->dim);
  else
    ctx->extra_rhs.clear();
  ctx->dirty_rhs = true;
  return FEMGPU_OK;
}

int femgpu_matvec(femgpu_ctx *ctx, int precision, const double *x, double *y) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!x || !y) return FEMGPU_E_ARG;
  if ((st = ensure_diag(ctx))) return st;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  std::vector<double> xi = to_internal(ctx, x), yi(nd, 0.0);
  double *dx, *dy, *part;
  if ((st = upload(ctx, "mv_x", xi, &dx)) || (st = dbuf(ctx, "mv_y", nd, &dy)) ||
      (st = dbuf(ctx, "partial64", size_t(ctx->plan.n_slots) * ctx->dim + 1, &part)))
    return st;
  femgpu::launch_fill<double>(dy, int64_t(nd), 0.0, ctx->stream);
  if (precision == FEMGPU_FP64) {
    femgpu::launch_matvec<double>(ctx->dm, dx, dy, part, ctx->stream);
  } else if (precision == FEMGPU_FP32) {
    float *fx, *fy, *fpart;
    if ((st = dbuf(ctx, "mv_x32", nd, &fx)) || (st = dbuf(ctx, "mv_y32", nd, &fy)) ||
        (st = dbuf(ctx, "partial32", size_t(ctx->plan.n_slots) * ctx->dim + 1, &fpart)))
      return st;
    femgpu::launch_convert<double, float>(dx, fx, int64_t(nd), ctx->stream);
    femgpu::launch_fill<float>(fy, int64_t(nd), 0.f, ctx->stream);
    femgpu::launch_matvec<float>(ctx->dm, fx, fy, fpart, ctx->stream);
    femgpu::launch_convert<float, double>(fy, dy, int64_t(nd), ctx->stream);
  } else {
    return fail(ctx, FEMGPU_E_ARG, "precision must be FP64 or FP32");
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(yi.data(), dy, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  from_internal(ctx, yi, y);
  return FEMGPU_OK;
}


int femgpu_matvec_noalloc(femgpu_ctx *ctx, int precision, const double *x_dev, double *y_dev) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!x_dev || !y_dev) return fail(ctx, FEMGPU_E_ARG, "null vector");
  if (!ctx->scratch) return fail(ctx, FEMGPU_E_STATE, "no scratch address for this device");
  // the operator (material, constraint mask, assembled diagonal of the
  // constrained rows) is prepared by the allocating entry points
  if (ctx->dirty_mat || ctx->dirty_fix || ctx->dirty_diag)
    return fail(ctx, FEMGPU_E_STATE,
                "operator not prepared since the last change of mesh, material or constraints: "
                "call femgpu_diag, femgpu_matvec or a solve once first");
  HIPCHK(hipSetDevice(ctx->device));
  if (precision == FEMGPU_FP64) return matvec_noalloc_t<double>(ctx, x_dev, y_dev);
  if (precision == FEMGPU_FP32) return matvec_noalloc_t<float>(ctx, x_dev, y_dev);
  return fail(ctx, FEMGPU_E_ARG, "precision must be FP64 or FP32");
}

int femgpu_matvec_repeat(femgpu_ctx *ctx, int precision, int reps) {
  int st = need_mesh(ctx);
  if (st) return st;
  if ((st = ensure_diag(ctx))) return st;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  if (precision == FEMGPU_FP64) {
    double *dx, *dy, *part;
    if ((st = dbuf(ctx, "mv_x", nd, &dx)) || (st = dbuf(ctx, "mv_y", nd, &dy)) ||
        (st = dbuf(ctx, "partial64", size_t(ctx->plan.n_slots) * ctx->dim + 1, &part)))
      return st;
    for (int k = 0; k < reps; ++k) femgpu::launch_matvec<double>(ctx->dm, dx, dy, part, ctx->stream);
  } else {
    float *fx, *fy, *fpart;
    if ((st = dbuf(ctx, "mv_x32", nd, &fx)) || (st = dbuf(ctx, "mv_y32", nd, &fy)) ||
        (st = dbuf(ctx, "partial32", size_t(ctx->plan.n_slots) * ctx->dim + 1, &fpart)))
      return st;
    for (int k = 0; k < reps; ++k) femgpu::launch_matvec<float>(ctx->dm, fx, fy, fpart, ctx->stream);
  }
  HIPCHK(hipGetLastError());
  return FEMGPU_OK;
}

int femgpu_time_matvec(femgpu_ctx *ctx, int precision, int reps, double *avg_ms) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (reps < 1 || !avg_ms) return FEMGPU_E_ARG;
  if ((st = femgpu_matvec_repeat(ctx, precision, 1))) return st;  // buffers + warm-up
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  if ((st = femgpu_matvec_repeat(ctx, precision, reps))) return st;
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = double(ms) / reps;
  return FEMGPU_OK;
}

int femgpu_diag(femgpu_ctx *ctx, double *d) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!d) return FEMGPU_E_ARG;
  if ((st = ensure_diag(ctx))) return st;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  std::vector<double> v(nd);
  double *dd;
  if ((st = dbuf(ctx, "diag64", nd, &dd))) return st;
  HIPCHK(hipMemcpyAsync(v.data(), dd, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  from_internal(ctx, v, d);
  return FEMGPU_OK;
}

int femgpu_rhs(femgpu_ctx *ctx, double *f) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!f) return FEMGPU_E_ARG;
  std::vector<double> v;
  host_rhs(ctx, false, v);
  std::copy(v.begin(), v.end(), f);
  return FEMGPU_OK;
}

int femgpu_default_solve_opts(femgpu_solve_opts *o) {
  if (!o) return FEMGPU_E_ARG;
  o->precision = FEMGPU_FP64;
  o->tol_kind = FEMGPU_TOL_REF_GAMMA;
  o->tol = 1e-8;
  o->max_iter = 0;
  o->inner_tol = 0.1;
  o->check_every = 64;
  o->recurrence = FEMGPU_CG_STANDARD;
  o->batch_schedule = FEMGPU_BATCH_LOCKSTEP;
  return FEMGPU_OK;
}

// FEMGPU_TRACE=1 in the environment: the mixed solves print one line per
// FP64 residual replacement to stderr.
static bool trace_enabled() {
  static const bool on = [] {
    const char *v = std::getenv("FEMGPU_TRACE");
    return v && *v && *v != '0';
  }();
  return on;
}

// solve_core with the single-reduction PCG (FEMGPU_CG_SINGLE_REDUCTION):
// FP64 / FP32 run it directly; MIXED runs it in FP32 inside the same FP64
// reliable-update loop as the standard recurrence.  At a residual
// replacement the state is rebuilt from the FP64 residual: r = f - K x64,
// u = M^-1 r, w = A u, and s = A p for the kept direction p (or p = s = 0
// at a restart).
static int solve_cg1(femgpu_ctx *ctx, const femgpu_solve_opts &o, int max_iter, int block,
                     const double *f64, double *x64, CgScalars *sc, CgScalars *sc2,
                     femgpu_stats *S) {
  const int64_t nd = ctx->nn * ctx->dim;
  int st, done = 0;
  Cg1Pos pos{0, 0};
  if (o.precision == FEMGPU_FP64) {
    femgpu::Cg1State<double> w;
    if ((st = cg1_ws(ctx, "64", w))) return st;
    const double *invD = static_cast<double *>(ctx->bufs["invD64"].p);
    femgpu::launch_copy<double>(f64, w.r[0], nd, ctx->stream);
    if ((st = run_cg1<double>(ctx, w, invD, o.tol_kind, o.tol, max_iter, block, 2, sc, &pos,
                              &done)))
      return st;
    femgpu::launch_copy<double>(w.x[pos.q], x64, nd, ctx->stream);
  } else if (o.precision == FEMGPU_FP32) {
    femgpu::Cg1State<float> w;
    if ((st = cg1_ws(ctx, "32", w))) return st;
    const float *invD = static_cast<float *>(ctx->bufs["invD32"].p);
    femgpu::launch_convert<double, float>(f64, w.r[0], nd, ctx->stream);
    if ((st = run_cg1<float>(ctx, w, invD, o.tol_kind, o.tol, max_iter, block, 3, sc, &pos,
                             &done)))
      return st;
    femgpu::launch_convert<float, double>(w.x[pos.q], x64, nd, ctx->stream);
  } else if (o.precision == FEMGPU_MIXED) {
    femgpu::Cg1State<float> w;
    double *r64, *y64, *part64, *part_node;
    if ((st = cg1_ws(ctx, "32", w)) || (st = dbuf(ctx, "ir_r", size_t(nd), &r64)) ||
        (st = dbuf(ctx, "ir_y", size_t(nd), &y64)) ||
        (st = dbuf(ctx, "partial64", size_t(ctx->plan.n_slots) * ctx->dim + 1, &part64)) ||
        (st = dbuf(ctx, "part_node", NODE_PART, &part_node)))
      return st;
    const float *invD = static_cast<float *>(ctx->bufs["invD32"].p);
    const double drop = o.inner_tol > 0 && o.inner_tol < 1 ? o.inner_tol : 0.1;
    const int block_m = std::max(2, block + (block & 1));
    femgpu::launch_fill<double>(x64, nd, 0.0, ctx->stream);
    femgpu::launch_convert<double, float>(f64, w.r[0], nd, ctx->stream);
    femgpu::launch_fill<float>(w.x[0], nd, 0.f, ctx->stream);
    femgpu::launch_fill<float>(w.p[0], nd, 0.f, ctx->stream);
    femgpu::launch_fill<float>(w.s[0], nd, 0.f, ctx->stream);
    femgpu::launch_reset_scalars(sc, FEMGPU_TOL_REL_RES, o.tol, max_iter, ctx->stream);
    femgpu::launch_cg1_init<float>(ctx->dm, w, invD, ctx->stream);
    femgpu::launch_cg1_fin(sc, w.part, ctx->dm.np, 0, 0, drop, ctx->stream);
    HIPCHK(hipGetLastError());
    if ((st = poll(ctx, sc))) return st;
    const double ff = ctx->h_sc->rr0;
    const int tb = tail_block_of(block_m);
    Replay exec, tail;
    if ((st = cg1_graph<float>(ctx, w, invD, block_m, 3, sc, &exec)) ||
        (st = cg1_graph<float>(ctx, w, invD, tb, 3, sc, &tail)))
      return st;
    // the stop prediction carries over replacements (same iteration
    // counter, the previous rate, the new residual)
    ReplayPredict pred;
    int status = FEMGPU_OK, replacements = 0, stag = 0;
    bool kept_p = false;  // the last replacement kept the search direction
    double rel = ff > 0 ? 1.0 : 0.0, rel_prev = rel;
    // The iterations pause on the device (done = 5) as soon as the FP32
    // residual has dropped by `drop`: run further, the FP32 recurrences of
    // s and w drift from A p and A u.
    while (rel > o.tol) {
      if ((st = replay_until(ctx, exec, sc, block_m, max_iter, 0.0, &tail, tb, &pred)))
        return st;
      done = ctx->h_sc->done;
      pos = cg1_pos(*ctx->h_sc);
      if (trace_enabled())
        std::fprintf(stderr, "femgpu   inner stop %d (parity %d): r'r %.6e thresh %.6e pause %.6e rr0 %.6e it %d/%d\n",
                     ctx->h_sc->done, ctx->h_sc->par_final, ctx->h_sc->rr_final, ctx->h_sc->thresh,
                     ctx->h_sc->pause, ctx->h_sc->rr0, ctx->h_sc->it[0], ctx->h_sc->it[1]);
      femgpu::launch_flush_x(ctx->dm, x64, w.x[pos.q], ctx->stream);
      // a breakdown right after a replacement that kept p: the kept
      // direction did not suit the new residual; restart from the FP64
      // residual (the state is the last completed iteration's)
      const bool retry = done == 3 && kept_p;
      if (done != 0 && done != 1 && done != 5 && !retry) {
        status = done_status(done);
        break;
      }
      // FP64 residual of the accumulated solution
      femgpu::launch_reset_scalars(sc2, 1, 0.0, 0, ctx->stream);
      femgpu::launch_matvec_patch<double>(ctx->dm, x64, y64, part64, ctx->stream);
      femgpu::launch_residual64(ctx->dm, f64, y64, part64, r64, sc2, part_node, ctx->stream);
      HIPCHK(hipGetLastError());
      if ((st = poll(ctx, sc2))) return st;
      const double rr64 = ctx->h_sc->rr;
      rel = ff > 0 ? std::sqrt(rr64 / ff) : 0.0;
      if (trace_enabled())
        std::fprintf(stderr, "femgpu mixed/single-reduction: %d iterations, stop %d, FP64 rel-res %.3e\n",
                     pos.iters, done, rel);
      if (rel <= o.tol) break;
      if (pos.iters >= max_iter || replacements > 1000) {
        status = FEMGPU_E_MAXIT;
        break;
      }
      stag = replacements > 0 && rel >= rel_prev ? stag + 1 : 0;
      if (stag >= 2) {
        status = FEMGPU_E_STAGNATION;
        break;
      }
      // Restart (p = s = 0) when the FP64 residual did not at least halve
      // since the previous replacement: the FP32 directions have lost
      // conjugacy.  Otherwise keep p (moved to index 0) with s = A p.
      const bool restart = rel > 0.5 * rel_prev || retry;
      kept_p = !restart;
      femgpu::launch_fill<float>(w.x[0], nd, 0.f, ctx->stream);
      if (restart) {
        femgpu::launch_fill<float>(w.p[0], nd, 0.f, ctx->stream);
        femgpu::launch_fill<float>(w.s[0], nd, 0.f, ctx->stream);
      } else {
        if (pos.q) femgpu::launch_copy<float>(w.p[1], w.p[0], nd, ctx->stream);
        femgpu::launch_matvec<float>(ctx->dm, w.p[0], w.s[0], w.slot[1], ctx->stream);
      }
      femgpu::launch_convert<double, float>(r64, w.r[0], nd, ctx->stream);
      femgpu::launch_cg1_init<float>(ctx->dm, w, invD, ctx->stream);
      femgpu::launch_cg1_fin(sc, w.part, ctx->dm.np, restart ? 2 : 1, pos.q, drop, ctx->stream);
      HIPCHK(hipGetLastError());
      if (trace_enabled()) {
        if ((st = poll(ctx, sc))) return st;
        std::fprintf(stderr, "femgpu   replaced (%s): r'r %.6e g %.6e done %d it %d/%d\n",
                     restart ? "restart" : "keep p", ctx->h_sc->rr_final, ctx->h_sc->g_final,
                     ctx->h_sc->done, ctx->h_sc->it[0], ctx->h_sc->it[1]);
      }
      rel_prev = rel;
      pred.seed(pos.iters, rr64, rr64);
      ++replacements;
    }
    S->iterations = pos.iters;
    S->outer_iterations = replacements;
    S->status = status;
    return FEMGPU_OK;
  } else {
    return fail(ctx, FEMGPU_E_ARG, "unknown precision");
  }
  S->iterations = pos.iters;
  S->gamma_ratio = ctx->h_sc->gamma0 > 0 ? ctx->h_sc->g_final / ctx->h_sc->gamma0 : 0.0;
  S->status = done_status(done);
  return FEMGPU_OK;
}

// Static solve of K x = f with f the device vector `f_name` (n_nodes x dim,
// internal order); the solution stays in the device buffer "sol64" and is
// copied to u_out (user order) when u_out is given.
static int solve_core(femgpu_ctx *ctx, const femgpu_solve_opts *opts, const char *f_name,
                      double *u_out, femgpu_stats *stats) {
  int st;
  femgpu_solve_opts o;
  femgpu_default_solve_opts(&o);
  if (opts) o = *opts;
  if (!(o.tol >= 0) || (o.tol_kind != FEMGPU_TOL_REF_GAMMA && o.tol_kind != FEMGPU_TOL_REL_RES))
    return fail(ctx, FEMGPU_E_ARG, "bad tolerance");
  if (o.precision == FEMGPU_MIXED && o.tol_kind != FEMGPU_TOL_REL_RES)
    return fail(ctx, FEMGPU_E_ARG, "MIXED precision needs FEMGPU_TOL_REL_RES");
  if (o.recurrence != FEMGPU_CG_STANDARD && o.recurrence != FEMGPU_CG_SINGLE_REDUCTION)
    return fail(ctx, FEMGPU_E_ARG, "unknown PCG recurrence");
  const int max_iter = o.max_iter > 0 ? o.max_iter : 1000000;
  const int block = o.check_every > 0 ? o.check_every : 64;
  if ((st = ensure_diag(ctx))) return st;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  double *f64, *x64;
  CgScalars *sc, *sc2;
  if ((st = dbuf(ctx, f_name, nd, &f64)) || (st = dbuf(ctx, "sol64", nd, &x64)) ||
      (st = dbuf(ctx, "sc_main", 1, &sc)) || (st = dbuf(ctx, "sc_outer", 1, &sc2)))
    return st;
  femgpu_stats S{};
  int iters = 0, done = 0;
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  if (o.recurrence == FEMGPU_CG_SINGLE_REDUCTION) {
    if ((st = solve_cg1(ctx, o, max_iter, block, f64, x64, sc, sc2, &S))) return st;
  } else if (o.precision == FEMGPU_FP64) {
    PcgWs<double> w;
    if ((st = pcg_ws(ctx, "c64", w))) return st;
    w.invD = static_cast<double *>(ctx->bufs["invD64"].p);
    femgpu::launch_copy<double>(f64, w.r, int64_t(nd), ctx->stream);
    if ((st = run_pcg<double>(ctx, w, OpShift<double>(), o.tol_kind, o.tol, max_iter, block, 0,
                              sc, &iters, &done)))
      return st;
    femgpu::launch_copy<double>(w.x, x64, int64_t(nd), ctx->stream);
    S.iterations = iters;
    S.gamma_ratio = ctx->h_sc->gamma0 > 0 ? ctx->h_sc->g_final / ctx->h_sc->gamma0 : 0.0;
    S.status = done_status(done);
  } else if (o.precision == FEMGPU_FP32) {
    PcgWs<float> w;
    if ((st = pcg_ws(ctx, "c32", w))) return st;
    w.invD = static_cast<float *>(ctx->bufs["invD32"].p);
    femgpu::launch_convert<double, float>(f64, w.r, int64_t(nd), ctx->stream);
    if ((st = run_pcg<float>(ctx, w, OpShift<float>(), o.tol_kind, o.tol, max_iter, block, 1, sc,
                             &iters, &done)))
      return st;
    femgpu::launch_convert<float, double>(w.x, x64, int64_t(nd), ctx->stream);
    S.iterations = iters;
    S.gamma_ratio = ctx->h_sc->gamma0 > 0 ? ctx->h_sc->g_final / ctx->h_sc->gamma0 : 0.0;
    S.status = done_status(done);
  } else if (o.precision == FEMGPU_MIXED) {
    // FP32 PCG (matvec, preconditioner, vectors) with FP64 dot products, the
    // solution accumulated in FP64 and the residual replaced by the FP64
    // residual f - K x64 whenever the FP32 residual has dropped by
    // `inner_tol` (reliable updates), and at the end until the FP64 residual
    // meets the tolerance.
    PcgWs<float> w;
    double *r64, *y64, *part64, *part_node, *part_patch, *part_res;
    // part_res: the FP64 residual's partials (apart from the inner
    // iterations' part_node)
    if ((st = pcg_ws(ctx, "c32", w)) || (st = dbuf(ctx, "ir_r", nd, &r64)) ||
        (st = dbuf(ctx, "ir_y", nd, &y64)) ||
        (st = dbuf(ctx, "partial64", size_t(ctx->plan.n_slots) * ctx->dim + 1, &part64)) ||
        (st = dbuf(ctx, "part_node", NODE_PART, &part_node)) ||
        (st = dbuf(ctx, "part_res", size_t(femgpu::grid_blocks(ctx->dm.n_active)) + 1, &part_res)) ||
        (st = dbuf(ctx, "part_patch", size_t(ctx->dm.np) + 1, &part_patch)))
      return st;
    w.invD = static_cast<float *>(ctx->bufs["invD32"].p);
    const double drop = o.inner_tol > 0 && o.inner_tol < 1 ? o.inner_tol : 0.1;
    const int block_m = std::max(2, block + (block & 1));
    femgpu::launch_fill<double>(x64, int64_t(nd), 0.0, ctx->stream);
    femgpu::launch_fill<float>(w.x, int64_t(nd), 0.f, ctx->stream);
    femgpu::launch_fill<float>(w.p, int64_t(nd), 0.f, ctx->stream);
    femgpu::launch_reset_scalars(sc, FEMGPU_TOL_REL_RES, o.tol, max_iter, ctx->stream);
    femgpu::launch_pcg_replace(ctx->dm, f64, w.r, w.invD, part_node, ctx->stream);
    femgpu::launch_fin_init(sc, part_node, ctx->stream);
    HIPCHK(hipGetLastError());
    if ((st = poll(ctx, sc))) return st;
    const double ff = ctx->h_sc->rr0;
    // short replays of an eighth of a full one (a quarter elsewhere): each
    // of the 4-5 drops to a residual replacement ends in them as well as the
    // stop, and their no-op launches after the device pause add up
    const int tb = std::max(2, (block_m / 8) & ~1);
    Replay exec, tail;
    if ((st = pcg_graph<float>(ctx, w, OpShift<float>(), block_m, 1, sc, part_patch, part_node,
                               &exec)) ||
        (st = pcg_graph<float>(ctx, w, OpShift<float>(), tb, 1, sc, part_patch, part_node,
                               &tail)))
      return st;
    // the stop prediction carries over replacements: same iteration
    // counter, the rate of the previous inner solve, the new residual
    ReplayPredict pred;
    int status = FEMGPU_OK, replacements = 0, stag = 0;
    bool kept_p = false;  // the last replacement kept the search direction
    double rel = ff > 0 ? 1.0 : 0.0, rel_prev = rel;
    // The host checks after every replay (of check_every iterations) whether
    // the FP32 residual has dropped by `drop`; unless the drop was foreseen,
    // the replay queued behind it keeps iterating meanwhile and the FP64
    // residual is formed after that one.
    double pause_rr = ff * drop * drop;
    while (rel > o.tol) {
      if ((st = replay_until(ctx, exec, sc, block_m, max_iter, pause_rr, &tail, tb, &pred)))
        return st;
      const int done = ctx->h_sc->done;
      const double rr_rec = ctx->h_sc->rr_final;
      iters = ctx->h_sc->iter_final;
      femgpu::launch_pcg_apply_x<float>(ctx->dm, w.x, w.p, sc, ctx->stream);
      femgpu::launch_flush_x(ctx->dm, x64, w.x, ctx->stream);
      // a breakdown right after a replacement that kept p: restart from the
      // FP64 residual (no block of the failing update wrote x or r)
      const bool retry = done == 3 && kept_p;
      if (done != 0 && done != 1 && !retry) {
        status = done_status(done);
        break;
      }
      // FP64 residual of the accumulated solution
      femgpu::launch_reset_scalars(sc2, 1, 0.0, 0, ctx->stream);
      femgpu::launch_matvec_patch<double>(ctx->dm, x64, y64, part64, ctx->stream);
      femgpu::launch_residual64(ctx->dm, f64, y64, part64, r64, sc2, part_res, ctx->stream);
      HIPCHK(hipGetLastError());
      if ((st = poll(ctx, sc2))) return st;
      rel = ff > 0 ? std::sqrt(ctx->h_sc->rr / ff) : 0.0;
      const double rel_rec = ff > 0 ? std::sqrt(rr_rec / ff) : 0.0;
      if (trace_enabled())
        std::fprintf(stderr, "femgpu mixed: %d iterations, stop %d, FP32 recurrence rel-res %.3e, "
                     "FP64 rel-res %.3e\n", iters, done, rel_rec, rel);
      if (rel <= o.tol) break;
      if (iters >= max_iter || replacements > 1000) {
        status = FEMGPU_E_MAXIT;
        break;
      }
      // two replacements in a row without a decrease of the FP64 residual:
      // the tolerance is below the attainable accuracy (stop near the best)
      stag = replacements > 0 && rel >= rel_prev ? stag + 1 : 0;
      if (stag >= 2) {
        status = FEMGPU_E_STAGNATION;
        break;
      }
      // Reliable update: r is replaced by the FP64 residual, the directions
      // kept.  They are dropped (p = 0: plain refinement) when the recurrence
      // had drifted far from the FP64 residual (more than twice it) or the
      // FP64 residual did not at least halve since the previous replacement:
      // then the directions belong to another residual.
      femgpu::launch_pcg_replace(ctx->dm, r64, w.r, w.invD, part_node, ctx->stream);
      const bool restart = rel > 2.0 * rel_rec || rel > 0.5 * rel_prev || retry;
      if (restart) femgpu::launch_fill<float>(w.p, int64_t(nd), 0.f, ctx->stream);
      kept_p = !restart;
      if (done) femgpu::launch_pcg_resume(sc, ctx->stream);
      rel_prev = rel;
      pause_rr = ctx->h_sc->rr * drop * drop;
      pred.seed(iters, ctx->h_sc->rr, ctx->h_sc->rr);
      ++replacements;
    }
    S.iterations = iters;
    S.outer_iterations = replacements;
    S.status = status;
  } else {
    return fail(ctx, FEMGPU_E_ARG, "unknown precision");
  }
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  S.solve_ms = ms;
  double rr;
  if ((st = true_rel_res(ctx, f64, x64, &rr))) return st;
  S.rel_res = rr;
  if (u_out) {
    std::vector<double> xi(nd);
    HIPCHK(hipMemcpyAsync(xi.data(), x64, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    from_internal(ctx, xi, u_out);
  }
  if (stats) *stats = S;
  if (S.status != FEMGPU_OK) {
    ctx->err = "solve stopped with status " + std::to_string(S.status);
    return S.status;
  }
  return FEMGPU_OK;
}

// The assembled load vector of the current problem on the device ("f64").
static int ensure_rhs(femgpu_ctx *ctx) {
  if (!ctx->dirty_rhs) return FEMGPU_OK;
  std::vector<double> fu;
  host_rhs(ctx, false, fu);
  double *f64;
  int st = upload(ctx, "f64", to_internal(ctx, fu.data()), &f64);
  if (st) return st;
  ctx->dirty_rhs = false;
  return FEMGPU_OK;
}

int femgpu_solve_static(femgpu_ctx *ctx, const femgpu_solve_opts *opts, double *u_out,
                        femgpu_stats *stats) {
  int st = need_mesh(ctx);
  if (st) return st;
  if ((st = ensure_rhs(ctx))) return st;
  return solve_core(ctx, opts, "f64", u_out, stats);
}

int femgpu_solve_rhs(femgpu_ctx *ctx, const femgpu_solve_opts *opts, const double *f,
                     double *u_out, femgpu_stats *stats) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!f) return fail(ctx, FEMGPU_E_ARG, "null right-hand side");
  std::vector<double> fi = to_internal(ctx, f);
  // rows of nodes no element references carry no equation
  for (int64_t i = ctx->plan.n_owned + ctx->plan.n_iface; i < ctx->nn; ++i)
    for (int c = 0; c < ctx->dim; ++c) fi[size_t(i) * ctx->dim + c] = 0.0;
  double *df;
  if ((st = upload(ctx, "f_user", fi, &df))) return st;
  return solve_core(ctx, opts, "f_user", u_out, stats);
}

int femgpu_material_grad(femgpu_ctx *ctx, const double *u, const double *w, double *g_out) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!u || !w || !g_out) return fail(ctx, FEMGPU_E_ARG, "null argument");
  if ((st = ensure_material(ctx)) || (st = ensure_fix(ctx))) return st;
  std::vector<double> ui = to_internal(ctx, u), wi = to_internal(ctx, w), g(ctx->ne);
  double *du, *dw, *dg;
  if ((st = upload(ctx, "st_u", ui, &du)) || (st = upload(ctx, "gr_w", wi, &dw)) ||
      (st = dbuf(ctx, "en_e", g.size(), &dg)))
    return st;
  femgpu::launch_bilinear<double>(ctx->dm, du, dw, -1.0, dg, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g.data(), dg, g.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < ctx->ne; ++i) g_out[ctx->plan.elem_perm[i]] = g[i];
  return FEMGPU_OK;
}

int femgpu_compliance_grad(femgpu_ctx *ctx, const femgpu_solve_opts *opts, double *u_out,
                           double *c_out, double *g_out, femgpu_stats *stats) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!c_out || !g_out) return fail(ctx, FEMGPU_E_ARG, "null argument");
  if ((st = ensure_rhs(ctx))) return st;
  femgpu_stats S{};
  st = solve_core(ctx, opts, "f64", u_out, &S);
  if (stats) *stats = S;
  if (st) return st;
  const size_t nd = size_t(ctx->nn) * ctx->dim;
  double *f64, *x64, *dg;
  if ((st = dbuf(ctx, "f64", nd, &f64)) || (st = dbuf(ctx, "sol64", nd, &x64)) ||
      (st = dbuf(ctx, "en_e", size_t(ctx->ne), &dg)))
    return st;
  // K is symmetric and f does not depend on the element scales, so the
  // adjoint of C = f'u is u itself: dC/ds_e = -u_e' K~_e^0 u_e.
  femgpu::launch_bilinear<double>(ctx->dm, x64, x64, -1.0, dg, ctx->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> fh(nd), xh(nd), g(ctx->ne);
  HIPCHK(hipMemcpyAsync(fh.data(), f64, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(xh.data(), x64, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(g.data(), dg, g.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double c = 0.0;
  for (size_t i = 0; i < nd; ++i) c += fh[i] * xh[i];
  *c_out = c;
  for (int64_t i = 0; i < ctx->ne; ++i) g_out[ctx->plan.elem_perm[i]] = g[i];
  return FEMGPU_OK;
}

namespace {
// Per-element stress (strain = false) or engineering strain (true) of u.
int element_tensor(femgpu_ctx *ctx, const double *u, double *out, bool strain) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!u || !out) return FEMGPU_E_ARG;
  if ((st = ensure_material(ctx))) return st;
  const int nv = 3 * (ctx->dim - 1);
  std::vector<double> ui = to_internal(ctx, u), s(size_t(ctx->ne) * nv);
  double *du, *ds;
  if ((st = upload(ctx, "st_u", ui, &du)) || (st = dbuf(ctx, "st_s", s.size(), &ds))) return st;
  femgpu::launch_stress<double>(ctx->dm, du, ds, ctx->stream, strain);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s.data(), ds, s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < ctx->ne; ++i)
    for (int k = 0; k < nv; ++k) out[size_t(ctx->plan.elem_perm[i]) * nv + k] = s[size_t(i) * nv + k];
  return FEMGPU_OK;
}
}  // namespace

int femgpu_stress(femgpu_ctx *ctx, const double *u, double *sigma_out) {
  return element_tensor(ctx, u, sigma_out, false);
}

int femgpu_strain(femgpu_ctx *ctx, const double *u, double *eps_out) {
  return element_tensor(ctx, u, eps_out, true);
}

int femgpu_element_energy(femgpu_ctx *ctx, const double *u, double *e_out) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (!u || !e_out) return FEMGPU_E_ARG;
  if ((st = ensure_material(ctx)) || (st = ensure_fix(ctx))) return st;
  std::vector<double> ui = to_internal(ctx, u), e(ctx->ne);
  double *du, *de;
  if ((st = upload(ctx, "st_u", ui, &du)) || (st = dbuf(ctx, "en_e", e.size(), &de))) return st;
  femgpu::launch_energy<double>(ctx->dm, du, de, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(e.data(), de, e.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < ctx->ne; ++i) e_out[ctx->plan.elem_perm[i]] = e[i];
  return FEMGPU_OK;
}

int femgpu_default_dyn_opts(femgpu_dyn_opts *o) {
  if (!o) return FEMGPU_E_ARG;
  o->rho = 2400;
  o->beta1 = 0.5;
  o->beta2 = 0.0;
  o->dt = 1e-3;
  o->endtime = 0.1;
  o->wavelet_type = FEMGPU_WAVELET_NONE;
  o->wavelet_freq = 15;
  o->wavelet_timeshift = 0;
  o->wavelet_ampl = 1;
  o->precision = FEMGPU_FP64;
  o->mass_convention = FEMGPU_MASS_REFERENCE;
  o->implicit_mode = FEMGPU_IMPLICIT_REFERENCE;
  o->pcg_eps = 1e-4;
  o->steps_per_graph = 64;
  o->damping_alpha = 0.0;
  o->damping_beta = 0.0;
  o->relax_reference = 1;
  o->max_steps = 1000000;
  return FEMGPU_OK;
}

int femgpu_set_receivers(femgpu_ctx *ctx, int n, const int32_t *nodes, int quantity, int every) {
  int st = need_mesh(ctx);
  if (st) return st;
  if (n < 0 || (n > 0 && !nodes) || every < 1 || quantity < FEMGPU_RECV_DISPLACEMENT ||
      quantity > FEMGPU_RECV_ACCELERATION)
    return fail(ctx, FEMGPU_E_ARG, "bad receiver arguments");
  for (int i = 0; i < n; ++i)
    if (nodes[i] < 0 || nodes[i] >= ctx->nn) return fail(ctx, FEMGPU_E_ARG, "receiver node out of range");
  ctx->recv.assign(nodes, nodes + n);
  ctx->recv_quantity = quantity;
  ctx->recv_every = every;
  ctx->traces.clear();
  ctx->n_samples = 0;
  return FEMGPU_OK;
}

int femgpu_get_traces(femgpu_ctx *ctx, double *out, int64_t capacity, int *n_samples) {
  if (!ctx || !n_samples) return FEMGPU_E_ARG;
  *n_samples = ctx->n_samples;
  if (!out) return FEMGPU_OK;
  if (capacity < int64_t(ctx->traces.size())) return fail(ctx, FEMGPU_E_ARG, "trace buffer too small");
  std::copy(ctx->traces.begin(), ctx->traces.end(), out);
  return FEMGPU_OK;
}

}  // extern "C"

namespace {

// Per-DOF assembled mass (user order).  lumped: 2D rho A/3, 3D rho V/12
// (reference) or rho V/4 (physical); consistent diagonal: 2D rho A/6, 3D rho V/24.
std::vector<double> host_mass(const femgpu_ctx *ctx, double rho, bool lumped, int conv) {
  const int d = ctx->dim, npe = d + 1;
  std::vector<double> m(size_t(ctx->nn) * d, 0.0);
  for (int64_t e = 0; e < ctx->ne; ++e) {
    const int32_t *en = &ctx->elems[size_t(e) * npe];
    double meas;
    if (d == 2) {
      const double *a = &ctx->nodes[size_t(en[0]) * 2], *b = &ctx->nodes[size_t(en[1]) * 2],
                   *c = &ctx->nodes[size_t(en[2]) * 2];
      meas = 0.5 * std::fabs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]));
    } else {
      const double *a = &ctx->nodes[size_t(en[0]) * 3], *b = &ctx->nodes[size_t(en[1]) * 3],
                   *c = &ctx->nodes[size_t(en[2]) * 3], *q = &ctx->nodes[size_t(en[3]) * 3];
      const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      const double e3[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
      meas = std::fabs(e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) -
                       e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                       e1[2] * (e2[0] * e3[1] - e2[1] * e3[0])) / 6.0;
    }
    const double mass = rho * meas;
    double per;
    if (d == 2)
      per = lumped ? mass / 3.0 : mass / 6.0;
    else if (lumped)
      per = conv == FEMGPU_MASS_PHYSICAL ? mass / 4.0 : mass / 12.0;
    else
      per = mass / 24.0;
    for (int a = 0; a < npe; ++a)
      for (int c = 0; c < d; ++c) m[size_t(en[a]) * d + c] += per;
  }
  return m;
}

template <typename T>
int dyn_run_t(femgpu_ctx *ctx, const femgpu_dyn_opts &o, double *u_out, femgpu_dyn_stats *stats) {
  const int d = ctx->dim;
  const size_t nd = size_t(ctx->nn) * d;
  const bool expl = o.beta2 == 0.0;
  const double alpha = o.damping_alpha, beta_d = o.damping_beta;
  const bool damped = alpha != 0.0 || beta_d != 0.0;
  if (!(alpha >= 0.0) || !(beta_d >= 0.0)) return fail(ctx, FEMGPU_E_ARG, "damping must be >= 0");
  if (expl && beta_d != 0.0)
    return fail(ctx, FEMGPU_E_ARG, "explicit dynamics needs damping_beta == 0 (femfunc.cpp:145)");
  if (!expl && damped && o.implicit_mode != FEMGPU_IMPLICIT_PHYSICAL)
    return fail(ctx, FEMGPU_E_UNSUPPORTED, "implicit damping needs FEMGPU_IMPLICIT_PHYSICAL");
  const bool relax = o.endtime < 0 && !o.relax_reference;
  const double eps_relax = std::fabs(o.endtime);
  int st;
  if ((st = ensure_diag(ctx))) return st;
  // Time grid, GN22 coefficients and wavelet values follow the reference
  // (endnt = int(endtime/dt), t = nt*dt, pressures overwritten by the wavelet
  // value once it is non-zero, cpu_fem_utils.cpp:564-569), in float for FP32
  // state as the reference computes them and in double for FP64 state.
  const bool f32 = sizeof(T) == sizeof(float);
  const double dtd = f32 ? double(float(o.dt)) : o.dt;
  int endnt;
  if (relax)
    endnt = o.max_steps > 0 ? o.max_steps : 1000000;
  else
    endnt = o.endtime < 0 ? 1 : f32 ? int(float(o.endtime) / float(o.dt)) : int(o.endtime / o.dt);
  if (endnt < 1) return fail(ctx, FEMGPU_E_ARG, "endtime/dt gives no step");
  std::vector<double> cfile(endnt + 1, 1.0), cunit(endnt + 1, 0.0);
  {
    const bool loads = ctx->conv == FEMGPU_LOADS_REFERENCE && !ctx->lnode.empty();
    auto wav = [&](int nt) -> double {
      if (f32)
        return double(wavelet_value(float(nt) * float(o.dt), o.wavelet_type, float(o.wavelet_freq),
                                    float(o.wavelet_timeshift), float(o.wavelet_ampl)));
      return wavelet_value_f64(double(nt) * o.dt, o.wavelet_type, o.wavelet_freq,
                               o.wavelet_timeshift, o.wavelet_ampl);
    };
    double last = wav(0);
    for (int nt = 1; nt <= endnt; ++nt) {
      const double v = wav(nt);
      if (v != 0.0) last = v;
      if (!loads && last != 0.0) {
        cfile[nt] = 0.0;
        cunit[nt] = last;
      }
    }
  }
  std::vector<double> ffile, funit;
  host_rhs(ctx, false, ffile);
  host_rhs(ctx, true, funit);
  std::vector<double> ff_i = to_internal(ctx, ffile.data()), fu_i = to_internal(ctx, funit.data());
  std::vector<T> ff_t(ff_i.begin(), ff_i.end()), fu_t(fu_i.begin(), fu_i.end());
  const int mconv = o.mass_convention;
  std::vector<double> mass_l = host_mass(ctx, o.rho, true, mconv);
  std::vector<double> ml_i = to_internal(ctx, mass_l.data());
  std::vector<T> invM(nd, T(0)), mlump(nd, T(0));
  for (size_t i = 0; i < nd; ++i) {
    mlump[i] = T(ml_i[i]);
    invM[i] = ml_i[i] > 0 ? T(1.0 / ml_i[i]) : T(0);
  }
  const std::string tg = sizeof(T) == 8 ? "d64" : "d32";
  T *dd, *dd2, *dv, *da, *dy, *dpart, *dff, *dfu, *dinvM, *dml;
  double *dcf, *dcu, *dmaxp, *dmax;
  int *dstep;
  if ((st = dbuf(ctx, tg + "d", nd, &dd)) || (st = dbuf(ctx, tg + "d2", nd, &dd2)) ||
      (st = dbuf(ctx, tg + "v", nd, &dv)) ||
      (st = dbuf(ctx, tg + "a", nd, &da)) || (st = dbuf(ctx, tg + "y", nd, &dy)) ||
      (st = dbuf(ctx, tg + "part", size_t(ctx->plan.n_slots) * d + 1, &dpart)) ||
      (st = upload(ctx, tg + "ff", ff_t, &dff)) || (st = upload(ctx, tg + "fu", fu_t, &dfu)) ||
      (st = upload(ctx, tg + "invM", invM, &dinvM)) || (st = upload(ctx, tg + "ml", mlump, &dml)) ||
      (st = upload(ctx, "dyn_cf", cfile, &dcf)) || (st = upload(ctx, "dyn_cu", cunit, &dcu)) ||
      (st = dbuf(ctx, "dyn_step", 1, &dstep)) || (st = dbuf(ctx, "dyn_maxp", 256, &dmaxp)) ||
      (st = dbuf(ctx, "dyn_max", 1, &dmax)))
    return st;
  // receivers
  const int nrec = int(ctx->recv.size());
  const int every = ctx->recv_every;
  const int max_samples = nrec ? endnt / every : 0;
  int32_t *drecv = nullptr;
  T *dtr = nullptr;
  ctx->traces.clear();
  ctx->n_samples = 0;
  if (nrec) {
    if (int64_t(max_samples) * nrec * d > (int64_t(1) << 27))
      return fail(ctx, FEMGPU_E_ARG, "receiver traces too large: increase `every`");
    std::vector<int32_t> ri(nrec);
    for (int i = 0; i < nrec; ++i) ri[i] = ctx->plan.node_inv[ctx->recv[i]];
    if ((st = upload(ctx, "recv", ri, &drecv)) ||
        (st = dbuf(ctx, tg + "traces", size_t(std::max(max_samples, 1)) * nrec * d, &dtr)))
      return st;
  }
  const T *rsrc = ctx->recv_quantity == FEMGPU_RECV_VELOCITY ? dv
                  : ctx->recv_quantity == FEMGPU_RECV_ACCELERATION ? da : dd;
  // dsrc: the step's displacement buffer (the explicit steps alternate two)
  auto record = [&](int step_off, const T *dsrc = nullptr) {
    const T *src = dsrc && ctx->recv_quantity == FEMGPU_RECV_DISPLACEMENT ? dsrc : rsrc;
    if (nrec && max_samples > 0)
      femgpu::launch_record<T>(dstep, step_off, every, src, drecv, nrec, d, max_samples, dtr,
                               ctx->stream);
  };
  // max |v| for relaxation (device reduction, polled)
  auto vmax = [&](double *out) -> int {
    femgpu::launch_absmax<T>(int64_t(ctx->dm.n_active) * d, dv, dmaxp, dmax, ctx->stream);
    HIPCHK(hipMemcpyAsync(out, dmax, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FEMGPU_OK;
  };
  femgpu::launch_fill<T>(dd, int64_t(nd), T(0), ctx->stream);
  femgpu::launch_fill<T>(dd2, int64_t(nd), T(0), ctx->stream);
  femgpu::launch_fill<T>(dv, int64_t(nd), T(0), ctx->stream);
  femgpu::launch_fill<T>(da, int64_t(nd), T(0), ctx->stream);
  femgpu::launch_fill<T>(dy, int64_t(nd), T(0), ctx->stream);
  // GN22 coefficients (femfunc.cpp:258-310).
  femgpu::DynCoeffs C;
  if (f32) {
    const float b1 = float(o.beta1), b2 = float(o.beta2), dtf = float(o.dt);
    C.dt = double(dtf);
    C.c1 = double(0.5f * (1.f - b2) * dtf * dtf);
    C.c2 = double((1.f - b1) * dtf);
    C.c3 = double(b1 * dtf);
    C.c4 = double(0.5f * b2 * dtf * dtf);
  } else {
    C.dt = dtd;
    C.c1 = 0.5 * (1.0 - o.beta2) * dtd * dtd;
    C.c2 = (1.0 - o.beta1) * dtd;
    C.c3 = o.beta1 * dtd;
    C.c4 = 0.5 * o.beta2 * dtd * dtd;
  }
  C.alpha = alpha;
  C.expl_scale = 1.0 / (1.0 + alpha * C.c3);
  femgpu_dyn_stats S{};
  int steps = 0;
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  // explicit steps: the displacement alternates between two buffers (the
  // fused predictor reads one and writes the other); step n writes dbufs[n & 1 ^ 1]
  T *dbufs[2] = {dd, dd2};
  // Without receivers or relaxation (which read v and a every step) a step
  // is one patch launch (MODE 4); v, a and the interface slots then
  // alternate between two buffers as well, and the interface nodes' v and a
  // lag one step (the displacement does not).
  const bool one_launch = expl && nrec == 0 && !relax;
  T *vbufs[2] = {dv, dv}, *abufs[2] = {da, da}, *pbufs[2] = {dpart, dpart};
  if (one_launch) {
    T *dv2, *da2, *dpart2;
    if ((st = dbuf(ctx, tg + "v2", nd, &dv2)) || (st = dbuf(ctx, tg + "a2", nd, &da2)) ||
        (st = dbuf(ctx, tg + "part2", size_t(ctx->plan.n_slots) * d + 1, &dpart2)))
      return st;
    femgpu::launch_fill<T>(dv2, int64_t(nd), T(0), ctx->stream);
    femgpu::launch_fill<T>(da2, int64_t(nd), T(0), ctx->stream);
    vbufs[1] = dv2;
    abufs[1] = da2;
    pbufs[1] = dpart2;
  }
  // one explicit step at buffer parity par (k: the step's offset from dstep)
  auto explicit_step = [&](int par, int k) {
    if (one_launch) {
      femgpu::DynStep<T> ds{};
      ds.d_in = dbufs[par];
      ds.d_out = dbufs[par ^ 1];
      ds.v_in = vbufs[par];
      ds.v_out = vbufs[par ^ 1];
      ds.a_in = abufs[par];
      ds.a_out = abufs[par ^ 1];
      ds.part_in = pbufs[par ^ 1];
      ds.part_out = pbufs[par];
      ds.f_file = dff;
      ds.f_unit = dfu;
      ds.invM = dinvM;
      ds.coef_file = dcf;
      ds.coef_unit = dcu;
      ds.step_base = dstep;
      ds.step_off = k;
      femgpu::launch_dyn_step<T>(ctx->dm, ds, C, ctx->stream);
    } else {
      femgpu::launch_dyn_predict_matvec<T>(ctx->dm, dbufs[par], dv, da, dbufs[par ^ 1], C, dy,
                                           dpart, ctx->stream);
      femgpu::launch_dyn_correct<T>(ctx->dm, dy, dpart, dff, dfu, dcf, dcu, dstep, k, dinvM,
                                    da, dv, nullptr, C, ctx->stream, nullptr, true);
      record(k, dbufs[par ^ 1]);
    }
  };
  if (expl) {
    // Steps are captured in a HIP graph of G steps (G even, so every replay
    // starts on buffer 0) when at least one full graph fits the run; the
    // remainder (or the whole run, and every step when graphs are disabled)
    // is launched eagerly.
    int G = o.steps_per_graph > 0 ? o.steps_per_graph : 64;
    G += G & 1;
    int done = 0;
    bool settled = false;
    if (G <= endnt && !graphs_disabled()) {
      hipGraphExec_t exec = nullptr;
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < G; ++k) explicit_step(k & 1, k);
      HIPCHK(hipStreamEndCapture(ctx->stream, &g));
      const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(ie);
      while (done + G <= endnt && !settled) {
        femgpu::launch_set_int(dstep, done + 1, ctx->stream);
        const hipError_t le = hipGraphLaunch(exec, ctx->stream);
        if (le != hipSuccess) {
          (void)hipStreamSynchronize(ctx->stream);
          (void)hipGraphExecDestroy(exec);
          HIPCHK(le);
        }
        done += G;
        if (relax) {
          double vm = 0.0;
          if ((st = vmax(&vm))) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipGraphExecDestroy(exec);
            return st;
          }
          settled = vm < eps_relax;
        }
      }
      HIPCHK(hipStreamSynchronize(ctx->stream));
      (void)hipGraphExecDestroy(exec);
    }
    femgpu::launch_set_int(dstep, done + 1, ctx->stream);
    for (int k = 0; done < endnt && !settled; ++k, ++done) {
      explicit_step(done & 1, k);
      if (relax) {
        double vm = 0.0;
        if ((st = vmax(&vm))) return st;
        if (vm < eps_relax) settled = true;
      }
    }
    HIPCHK(hipGetLastError());
    steps = done;
  } else {
    // implicit: nested PCG per step on A a = F - K d (+ damping terms).
    PcgWs<T> w;
    if ((st = pcg_ws(ctx, sizeof(T) == 8 ? "c64" : "c32", w))) return st;
    T *invA, *dt_tmp = nullptr;
    if ((st = dbuf(ctx, tg + "invA", nd, &invA))) return st;
    if (beta_d != 0.0 && (st = dbuf(ctx, tg + "dtmp", nd, &dt_tmp))) return st;
    OpShift<T> sh;
    const T ceff = T(C.c4 + C.c3 * beta_d);  // beta2 dt^2/2 (+ beta1 dt beta)
    const T cmass = T(1.0 + C.c3 * alpha);   // 1 (+ beta1 dt alpha)
    const T *diagK_t = sizeof(T) == 8 ? reinterpret_cast<const T *>(ctx->dm.diag64)
                                      : reinterpret_cast<const T *>(ctx->dm.diag32);
    if (o.implicit_mode == FEMGPU_IMPLICIT_PHYSICAL) {
      sh.cK = ceff;
      sh.cM = cmass;
      sh.m = dml;
      femgpu::launch_inv_diag<T>(ctx->dm, diagK_t, ceff, cmass, dml, invA, ctx->stream);
    } else {
      // reference: operator K, preconditioner cK diag(K) + diag(M_consistent)
      std::vector<double> mc = host_mass(ctx, o.rho, false, mconv);
      std::vector<double> mc_i = to_internal(ctx, mc.data());
      std::vector<T> mc_t(mc_i.begin(), mc_i.end());
      T *dmc;
      if ((st = upload(ctx, tg + "mc", mc_t, &dmc))) return st;
      femgpu::launch_inv_diag<T>(ctx->dm, diagK_t, ceff, T(1), dmc, invA, ctx->stream);
    }
    w.invD = invA;
    CgScalars *sc;
    double *part_patch, *part_node;
    if ((st = dbuf(ctx, "sc_main", 1, &sc)) ||
        (st = dbuf(ctx, "part_patch", size_t(ctx->dm.np) + 1, &part_patch)) ||
        (st = dbuf(ctx, "part_node", NODE_PART, &part_node)))
      return st;
    const int max_iter = 1000000;
    const int prec_tag = sizeof(T) == 8 ? 10 : 11;
    // The PCG solves for the acceleration in place: its x is a (the
    // corrector does not read a, the predictor reads it before the PCG
    // start clears it).
    w.x = da;
    // The end of a step: x += alpha p still pending from the last update
    // (the next PCG start resets the pending step), then v and d, and the
    // receivers.
    auto step_end = [&]() {
      femgpu::launch_pcg_apply_x<T>(ctx->dm, w.x, w.p, sc, ctx->stream, false);
      femgpu::launch_dyn_implicit_post<T>(ctx->dm, da, dv, dd, C, ctx->stream);
      record(0);
    };
    // The start of a step up to the PCG's first `first` iterations: the
    // predictor, which also advances the device step counter, K (d~ + beta
    // v~), corrector into the PCG residual, PCG start.
    auto step_begin = [&](int first) {
      femgpu::launch_dyn_predict<T>(ctx->dm, dd, dv, da, C, true, dstep, ctx->stream);
      if (beta_d != 0.0) {
        femgpu::launch_xpby<T>(int64_t(nd), dd, dv, beta_d, dt_tmp, ctx->stream);
        femgpu::launch_matvec_patch<T>(ctx->dm, dt_tmp, dy, dpart, ctx->stream);
      } else {
        femgpu::launch_matvec_patch<T>(ctx->dm, dd, dy, dpart, ctx->stream);
      }
      femgpu::launch_dyn_correct<T>(ctx->dm, dy, dpart, dff, dfu, dcf, dcu, dstep, 0, nullptr, da,
                                    dv, w.r, C, ctx->stream, dml);
      femgpu::launch_pcg_start<T>(ctx->dm, w.x, w.p, int64_t(nd), w.r, w.invD, sc,
                                  FEMGPU_TOL_REF_GAMMA, o.pcg_eps, max_iter, part_node,
                                  ctx->stream);
      for (int k = 0; k < first; ++k)
        femgpu::launch_pcg_iter<T>(ctx->dm, w.x, w.r, w.p, w.q, w.partial, w.invD, sh, sc,
                                   part_patch, part_node, k & 1, ctx->stream);
    };
    // One launch per step: [end of the previous step,] start of this step
    // and its first PCG replay, captured once per (with_end, first) and kept
    // on the context under a key that names every pointer and value the
    // launches hold besides the mesh and material, whose change drops every
    // graph (the step number is not among them).
    std::string key_base;
    {
      char b[40];
      const void *ptrs[] = {dd, dv, da, dy, dpart, dff, dfu, dcf, dcu, dstep, dml, dt_tmp,
                            w.x, w.r, w.p, w.q, w.partial, w.invD, sc, part_patch, part_node,
                            drecv, dtr, rsrc, sh.m};
      const double vals[] = {C.dt, C.c1, C.c2, C.c3, C.c4, C.alpha, C.expl_scale, beta_d,
                             double(sh.cK), double(sh.cM), o.pcg_eps,
                             double(nrec), double(every), double(max_samples)};
      key_base = "impl/" + std::to_string(prec_tag);
      for (const void *q : ptrs) {
        std::snprintf(b, sizeof(b), "/%p", q);
        key_base += b;
      }
      for (double v : vals) {
        std::snprintf(b, sizeof(b), "/%.17g", v);
        key_base += b;
      }
    }
    // Step graphs of other option sets (another dt, tolerance, damping)
    // are released here once there are many: nothing launched from the
    // cache is in flight or held between API calls.
    {
      const std::string own = key_base + "/";
      auto stale = [&](const std::string &k) {
        return k.compare(0, 5, "impl/") == 0 && k.compare(0, own.size(), own) != 0;
      };
      size_t n_stale = 0;
      for (const auto &kv : ctx->graphs) n_stale += stale(kv.first);
      if (n_stale > 256) {
        for (auto it = ctx->graphs.begin(); it != ctx->graphs.end();) {
          if (stale(it->first)) {
            (void)hipGraphExecDestroy(it->second);
            it = ctx->graphs.erase(it);
          } else {
            ++it;
          }
        }
      }
    }
    auto step_graph = [&](bool with_end, int first, Replay *out) -> int {
      if (graphs_disabled()) {
        out->exec = nullptr;
        out->eager = [=, &step_end, &step_begin]() {
          if (with_end) step_end();
          step_begin(first);
        };
        return FEMGPU_OK;
      }
      if (ctx->graph_gen != ctx->buf_gen) {
        drop_graphs(ctx);
        ctx->graph_gen = ctx->buf_gen;
      }
      const std::string key =
          key_base + "/" + std::to_string(int(with_end)) + "/" + std::to_string(first);
      auto it = ctx->graphs.find(key);
      if (it != ctx->graphs.end()) {
        out->exec = it->second;
        return FEMGPU_OK;
      }
      hipGraph_t g;
      hipGraphExec_t exec = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      if (with_end) step_end();
      step_begin(first);
      HIPCHK(hipStreamEndCapture(ctx->stream, &g));
      const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(ie);
      ctx->graphs[key] = exec;
      out->exec = exec;
      return FEMGPU_OK;
    };
    femgpu::launch_set_int(dstep, 0, ctx->stream);
    int total_it = 0;
    // The first replay covers the larger iteration count of the last two
    // steps (successive steps' solves take about the same number) and the
    // launch after it, whose stop test sees the converged residual (at most
    // 64 launches).  A solve that needs more continues, pipelined, in
    // replays of two iterations when the count was foreseen (the stop is
    // near), of 32 when the last steps took more than 64 (long solves, the
    // stop prediction of replay_until ends them in short replays), of 8 on
    // the first step (count unknown); after a first replay of odd length
    // these start at parity 1.
    int it_prev[2] = {0, 0};
    bool end_pending = false;
    int nt = 1;
    for (; nt <= endnt; ++nt) {
      const int pred = std::max(it_prev[0], it_prev[1]);
      const int first = pred > 0 ? std::min(64, pred + 1) : 8;
      const int cont_block = pred == 0 ? 8 : pred > first ? 32 : 2;
      Replay step;
      if ((st = step_graph(end_pending, first, &step))) return st;
      HIPCHK(step.launch(ctx->stream));
      HIPCHK(hipMemcpyAsync(ctx->h_sc, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      end_pending = false;
      if (!ctx->h_sc->done) {
        const int cb = cont_block, ctb = tail_block_of(cb);
        Replay cont, tail;
        const int p0 = first & 1;
        if ((st = pcg_graph<T>(ctx, w, sh, cb, prec_tag, sc, part_patch, part_node, &cont, p0)) ||
            (st = pcg_graph<T>(ctx, w, sh, ctb, prec_tag, sc, part_patch, part_node, &tail, p0)) ||
            (st = replay_until(ctx, cont, sc, cb, max_iter, 0.0, &tail, ctb)))
          return st;
      }
      const int it = ctx->h_sc->iter_final, done = ctx->h_sc->done;
      total_it += it;
      it_prev[nt & 1] = it;
      if (done != 1) {
        S.status = done_status(done);
        ctx->err = "implicit step " + std::to_string(nt) + ": PCG stopped with " + std::to_string(done);
        if (stats) *stats = S;
        return S.status;
      }
      if (relax) {
        step_end();
        double vm = 0.0;
        if ((st = vmax(&vm))) return st;
        if (vm < eps_relax) {
          ++nt;
          break;
        }
      } else {
        end_pending = true;
      }
    }
    if (end_pending) step_end();
    HIPCHK(hipGetLastError());
    steps = nt - 1;
    S.pcg_iterations = total_it;
  }
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  S.run_ms = ms;
  S.steps = steps;
  if (u_out) {
    std::vector<T> xt(nd);
    const T *dfin = expl ? dbufs[steps & 1] : dd;
    HIPCHK(hipMemcpyAsync(xt.data(), dfin, nd * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<double> xi(xt.begin(), xt.end());
    from_internal(ctx, xi, u_out);
  }
  if (nrec) {
    const int ns = std::min(max_samples, steps / every);
    std::vector<T> tr(size_t(std::max(ns, 0)) * nrec * d);
    if (ns > 0) {
      HIPCHK(hipMemcpyAsync(tr.data(), dtr, tr.size() * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    ctx->traces.assign(tr.begin(), tr.end());
    ctx->n_samples = std::max(ns, 0);
  }
  S.status = FEMGPU_OK;
  if (relax && steps >= endnt) {
    S.status = FEMGPU_E_MAXIT;
    ctx->err = "relaxation did not reach max|v| < eps within max_steps";
    if (stats) *stats = S;
    return S.status;
  }
  if (stats) *stats = S;
  return FEMGPU_OK;
}

}  // namespace

extern "C" {

int femgpu_dyn_run(femgpu_ctx *ctx, const femgpu_dyn_opts *opts, double *u_out,
                   femgpu_dyn_stats *stats) {
  int st = need_mesh(ctx);
  if (st) return st;
  femgpu_dyn_opts o;
  femgpu_default_dyn_opts(&o);
  if (opts) o = *opts;
  if (!(o.dt > 0) || !(o.rho > 0)) return fail(ctx, FEMGPU_E_ARG, "dt and rho must be > 0");
  if (o.precision == FEMGPU_FP64) return dyn_run_t<double>(ctx, o, u_out, stats);
  if (o.precision == FEMGPU_FP32) return dyn_run_t<float>(ctx, o, u_out, stats);
  return fail(ctx, FEMGPU_E_ARG, "dynamics precision must be FP64 or FP32");
}

}  // extern "C"

namespace {

int time_dyn_step(femgpu_ctx *ctx, const femgpu_dyn_opts &opts, int steps, double *us) {
  femgpu_dyn_opts o = opts;
  auto run = [&](int n, double *t, int *k) -> int {
    o.endtime = (n + 0.5) * o.dt;  // a fixed number of steps (no relaxation)
    femgpu_dyn_stats S{};
    const int rs = femgpu_dyn_run(ctx, &o, nullptr, &S);
    if (rs) return rs;
    *t = S.run_ms;
    *k = S.steps;
    return FEMGPU_OK;
  };
  // the difference between runs of `steps` and steps / 4 steps (each the
  // faster of two, after a run that captures the step graphs), so that the
  // fixed cost of a run does not count
  const int s1 = std::max(1, steps / 4);
  double t1 = 0.0, t4 = 0.0, t;
  int k1 = 0, k4 = 0, k;
  int st;
  if ((st = run(s1, &t, &k))) return st;
  for (int rep = 0; rep < 2; ++rep) {
    if ((st = run(s1, &t, &k))) return st;
    if (rep == 0 || t < t1) t1 = t, k1 = k;
    if ((st = run(steps, &t, &k))) return st;
    if (rep == 0 || t < t4) t4 = t, k4 = k;
  }
  if (k4 <= k1) return fail(ctx, FEMGPU_E_STATE, "autotune: no time step ran");
  *us = std::max(t4 - t1, 1e-9) * 1e3 / (k4 - k1);
  return FEMGPU_OK;
}

}  // namespace
