// kernels.h -- launch interface of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "femgpu_internal.h"

namespace femgpu {

// Device view of a planned mesh (all pointers are device pointers).
struct DevMesh {
  int dim = 0;
  int np = 0;              // patches
  int emax = 0;            // elements per patch bound
  int ent = 0;             // LDS entry capacity (longest padded entry list, multiple of 4)
  int pn = 0;              // LDS node capacity
  int patch_threads = 256;  // threads per patch workgroup (256, 512 or 1024)
  int64_t nn = 0, ne = 0, n_owned = 0, n_iface = 0, n_active = 0;
  const PatchInfo *patches = nullptr;
  const uint32_t *erec = nullptr;      // 3 words per element (pack_records)
  // the same records at stride emax per patch (patch p's at p * emax, zero
  // padding): erec itself, padded at the end, when every patch starts there
  const uint32_t *erec_s = nullptr;
  const int32_t *sh_gid = nullptr;
  const int32_t *sh_slot = nullptr;
  const int4 *sh_info = nullptr;      // per sh entry: id, slot, node's first slot, end slot
  const uint16_t *node_end = nullptr;
  // per patch node at stride patch_threads (pack_patch_nodes): node id,
  // slot with its first-slot bit, end of the node's entry run
  const int32_t *pgid = nullptr;
  const int32_t *pslot = nullptr;
  const uint16_t *pend = nullptr;
  const int32_t *slot_start = nullptr;
  const uint8_t *fixmask = nullptr;   // per internal node: bit c = component c fixed
  const int32_t *elems_int = nullptr; // ne x npe internal node ids
  const double *coords64 = nullptr;
  const float *coords32 = nullptr;
  const double *escale64 = nullptr;   // per internal element or null
  const float *escale32 = nullptr;
  const double *diag64 = nullptr;     // assembled diag(K), set before any matvec
  const float *diag32 = nullptr;
  double lam = 0, mu = 0;             // Lame constants (2D: plane stress)
};

// Scalars of one PCG run, device resident.  Iteration j of the fused
// iteration writes its gamma/beta/count into slot j&1 and reads slot
// (j&1)^1, so no kernel reads a field that another block of the same launch
// writes.
struct CgScalars {
  double gamma0;       // r0'z0
  double rr0;          // r0'r0
  double thresh;       // stop threshold (kind dependent)
  double eps;          // tol as given
  double g[2];         // gamma_j = r_j'z_j
  double b[2];         // beta_j
  double al[2];        // alpha_j (single-reduction PCG, cg1.hip)
  double pause;        // single-reduction PCG: stop with done = 5 once r'r < pause (0: never)
  double g_final;      // last gamma seen by the stop test
  double rr_final;     // last r'r seen by the stop test
  double pq;           // last p'q (diagnostic)
  // standard PCG in FP64: the step length alpha of the last update whose
  // x += alpha p has not been applied yet (the next iteration's patch kernel
  // applies it, launch_pcg_apply_x after a stop; 0 when none is pending)
  double x_pend;
  double rr;           // r'r of the FP64 refinement residual
  double scale;        // sqrt(rr) (refinement scaling)
  int it[2];           // updates done before iteration j
  int iter_final;      // updates done when the stop flag was set
  int par_final;       // parity of the launch that set the stop flag
  int kind;            // 0: gamma ratio, 1: relative residual
  int done;            // 0 running, 1 converged, 2 non-finite, 3 breakdown, 4 max-iter,
                       // 5 paused for a residual replacement (single-reduction PCG)
  int max_iter;
};

// Optional mass shift of the operator: y = cK*K x + cM*diag(m) x.
template <typename T>
struct OpShift {
  T cK = T(1);
  T cM = T(0);
  const T *m = nullptr;   // per DOF, internal order
};

int lds_bytes(int dim, int elem_bytes, int pn, int ent);
// blocks of the node-parallel kernels over n items (one partial per block)
int grid_blocks(int64_t n);
// Per-kernel launch attributes (large dynamic LDS); call once per device
// before any launch or capture.
hipError_t prepare_kernels();

// y = K x (standalone: patch kernel + interface gather).
template <typename T>
void launch_matvec(const DevMesh &dm, const T *x, T *y, T *partial, hipStream_t s);
// Patch kernel only (interface results left in `partial`).
template <typename T>
void launch_matvec_patch(const DevMesh &dm, const T *x, T *y, T *partial, hipStream_t s);
// Assembled diagonal of K.
template <typename T>
void launch_diag(const DevMesh &dm, T *d, T *partial, hipStream_t s);

// invD = 1 / (cK*diag + cM*m) (0 where the sum is 0).
template <typename T>
void launch_inv_diag(const DevMesh &dm, const T *diag, T cK, T cM, const T *m, T *invD,
                     hipStream_t s);

// PCG: start (x = p = 0 over n_all entries, every scalar reset as
// launch_reset_scalars does, gamma0 = r'z with z = invD r) in two launches,
// and one fused iteration (z is formed from r where it is needed, never
// stored).
template <typename T>
void launch_pcg_start(const DevMesh &dm, T *x, T *p, int64_t n_all, const T *r, const T *invD,
                      CgScalars *sc, int kind, double eps, int max_iter, double *part_node,
                      hipStream_t s);
// parity: j & 1 of the iteration (ping-pong of scalars and partial buffers).
// part_patch: >= np doubles; part_node: 4 * NODE_PARTS doubles.
template <typename T>
void launch_pcg_iter(const DevMesh &dm, T *x, T *r, T *p, T *q, T *partial, const T *invD,
                     OpShift<T> sh, CgScalars *sc, double *part_patch, double *part_node,
                     int parity, hipStream_t s);
// After the last replay of a standard PCG run: the x += alpha p of its
// last update, which the next iteration would have applied (sc->x_pend).
// clear_pending = false leaves sc->x_pend set, for a caller whose next
// PCG start resets the scalars anyway.
template <typename T>
void launch_pcg_apply_x(const DevMesh &dm, T *x, const T *p, CgScalars *sc, hipStream_t s,
                        bool clear_pending = true);
constexpr int NODE_PARTS = 512;  // blocks of the node-parallel PCG kernels
static_assert(NODE_PARTS % 256 == 0, "the patch kernels read the partials in chunks of 256");

// r = f - (interface-finished) K x, rr -> sc->rr (FP64 outer loop of MIXED).
void launch_residual64(const DevMesh &dm, const double *f, const double *y, const double *partial,
                       double *r, CgScalars *sc, double *part, hipStream_t s);
// Mixed solve (reliable updates): x64 += x32, x32 = 0; residual replacement
// r32 = r64 with the partials the next iteration reads; resume after a stop
// flag.
void launch_flush_x(const DevMesh &dm, double *x64, float *x32, hipStream_t s);
void launch_pcg_replace(const DevMesh &dm, const double *r64, float *r32, const float *invD,
                        double *part_node, hipStream_t s);
// Clear the stop flag for the next replay (parity 0) after a residual
// replacement: the scalars of the stopping launch move to slot 1.
void launch_pcg_resume(CgScalars *sc, hipStream_t s);
// gamma0 / thresholds from the slot-1 partials (after launch_pcg_replace).
void launch_fin_init(CgScalars *sc, double *part_node, hipStream_t s);

template <typename T>
void launch_fill(T *p, int64_t n, T v, hipStream_t s);
template <typename T>
void launch_copy(const T *a, T *b, int64_t n, hipStream_t s);
template <typename TA, typename TB>
void launch_convert(const TA *a, TB *b, int64_t n, hipStream_t s);

// Per-element outputs in internal element order.
template <typename T>
void launch_stress(const DevMesh &dm, const T *u, T *sigma, hipStream_t s, bool strain = false);
template <typename T>
void launch_energy(const DevMesh &dm, const T *u, T *en, hipStream_t s);
template <typename T>
void launch_bilinear(const DevMesh &dm, const T *u, const T *w, double factor, T *out, hipStream_t s);

// Dynamics.
struct DynCoeffs {
  double dt, c1, c2, c3, c4;  // predictor / corrector coefficients
  double alpha = 0.0;         // Rayleigh mass coefficient (damping C = alpha M + beta K)
  double expl_scale = 1.0;    // explicit: 1 / (1 + alpha c3)
};
// Implicit predictor d += dt v + c1 a, v += c2 a (a = 0 with zero_a);
// step_inc != nullptr: *step_inc += 1 (the device step counter) as well.
template <typename T>
void launch_dyn_predict(const DevMesh &dm, T *d, T *v, T *a, DynCoeffs c, bool zero_a,
                        int *step_inc, hipStream_t s);
// r = F(step) - K d (interface finished from partials); explicit: a = r*invM, v += c3 a
// (pred_v: v and a are the previous step's, the predicted v + c2 a is formed here).
template <typename T>
void launch_dyn_correct(const DevMesh &dm, const T *y, const T *partial, const T *f_file,
                        const T *f_unit, const double *coef_file, const double *coef_unit,
                        const int *step_base, int step_off, const T *invM, T *a, T *v,
                        T *r_out, DynCoeffs c, hipStream_t s, const T *mlump = nullptr,
                        bool pred_v = false);

// Buffers of one explicit GN22 step in a single patch launch (MODE 4): d, v,
// a and the interface slots each alternate between two buffers; the
// interface nodes' v and a lag one step (corrected at the next launch; the
// displacement of every node is current after each launch).
template <typename T>
struct DynStep {
  const T *d_in, *v_in, *a_in, *part_in;
  T *d_out, *v_out, *a_out, *part_out;
  const T *f_file, *f_unit, *invM;
  const double *coef_file, *coef_unit;
  const int *step_base;
  int step_off;
};
template <typename T>
void launch_dyn_step(const DevMesh &dm, const DynStep<T> &st, DynCoeffs c, hipStream_t s);
// Explicit GN22 predictor fused with the matvec: d_out = (d_in + dt v) + c1 a
// (owned and first-slot nodes), y / partial = K d_out.
template <typename T>
void launch_dyn_predict_matvec(const DevMesh &dm, const T *d_in, const T *v, const T *acc,
                               T *d_out, DynCoeffs c, T *y, T *partial, hipStream_t s);
template <typename T>
void launch_dyn_implicit_post(const DevMesh &dm, const T *a, T *v, T *d, DynCoeffs c,
                              hipStream_t s);
// out = x + b y
template <typename T>
void launch_xpby(int64_t n, const T *x, const T *y, double b, T *out, hipStream_t s);
// receiver traces (samples at steps that are multiples of `every`)
template <typename T>
void launch_record(const int *step_base, int step_off, int every, const T *src, const int32_t *recv,
                   int nrec, int dim, int max_samples, T *traces, hipStream_t s);
// *out = max |x_i| (part: 256 doubles)
template <typename T>
void launch_absmax(int64_t n, const T *x, double *part, double *out, hipStream_t s);
void launch_set_int(int *p, int v, hipStream_t s);
void launch_reset_scalars(CgScalars *sc, int kind, double eps, int max_iter, hipStream_t s);

}  // namespace femgpu
