/*
 * femgpu.h -- C-ABI of the MI355X-native EbE linear-elasticity solve path.
 *
 * This is the drop-in boundary for the reference's solver path
 * (moarseniy/CudaFEMproject).  The reference exposes that path as C++
 * functions taking its own `Matrix`, `ElementsData` and `dataKeeper` objects:
 *
 *   femfunc/src/femfunc.h:20   gpuCalculateFEM_EbE_vec(DEVICE_NAME, dataKeeper&, bool)
 *   femfunc/src/femfunc.h:21   gpuPCG_EbE_vec(Matrix&, DEVICE_NAME, dataKeeper&, ElementsData&, bool, float, bool)
 *   femfunc/src/femfunc.h:22   gpuCalculateFEM_DYN(DEVICE_NAME, dataKeeper&, bool)
 *   fem_utils/include/fem_utils_cpu.h:24-148  ElementsData (the element loops)
 *   matrix_pack/include/matrix_pack_cpu.h:31-121  Matrix (batched local matrices)
 *
 * Each entry point below names the reference interface it replaces.  All
 * signatures use plain pointers and sizes; host arrays are caller-owned and
 * copied in; results are written to caller-owned host arrays.  One context
 * per host thread; every call on a context is issued on the context's HIP
 * stream (see femgpu_set_stream).  Every function returns an int status
 * (FEMGPU_OK == 0, negative on error) and never exits the process, unlike the
 * reference (Tools/src/Tools.cpp:27-33, femfunc.cpp:107-110).
 */
#ifndef FEMGPU_H
#define FEMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct femgpu_ctx femgpu_ctx;

/* Status codes. */
enum {
  FEMGPU_OK = 0,
  FEMGPU_E_ARG = -1,        /* bad argument                                  */
  FEMGPU_E_HIP = -2,        /* HIP runtime error (message in last_error)     */
  FEMGPU_E_NONFINITE = -3,  /* NaN/Inf in the iteration (femfunc.cpp:107)    */
  FEMGPU_E_MAXIT = -4,      /* iteration cap reached before the tolerance    */
  FEMGPU_E_BREAKDOWN = -5,  /* p'Kp <= 0 (indefinite or singular operator)   */
  FEMGPU_E_BAD_MESH = -6,   /* degenerate element, bad index, face not on elem */
  FEMGPU_E_STATE = -7,      /* call order (e.g. solve before set_mesh)       */
  FEMGPU_E_NOMEM = -8,      /* device allocation failed                      */
  FEMGPU_E_UNSUPPORTED = -9,/* input the reference only handles by accident  */
  FEMGPU_E_STAGNATION = -10 /* MIXED / batch: the FP64 residual stopped
                               decreasing above the tolerance (two residual
                               replacements in a row without progress: the
                               tolerance is below the attainable accuracy)  */
};

/* Arithmetic of a solve. */
enum {
  FEMGPU_FP64 = 0,   /* FP64 vectors, FP64 EbE matvec                        */
  FEMGPU_FP32 = 1,   /* FP32 vectors and matvec, FP64 dot accumulation       */
  FEMGPU_MIXED = 2   /* FP32 PCG inside FP64 iterative refinement: FP64
                        solution, FP64 residual replacement (reliable
                        updates), FP64 dot products                          */
};

/* Stopping rules. */
enum {
  FEMGPU_TOL_REF_GAMMA = 0, /* gamma_new < tol*gamma0, gamma = r'M^-1 r      */
                            /* (the reference's rule, femfunc.cpp:92)         */
  FEMGPU_TOL_REL_RES = 1    /* ||f - K u||_2 <= tol * ||f||_2                */
};

/* Schedules of femgpu_batch_solve (no counterpart in the reference). */
enum {
  FEMGPU_BATCH_LOCKSTEP = 0,   /* every iteration of every system in two
                                  grid-wide launches (graph replays): each
                                  instance spread over the whole GPU        */
  FEMGPU_BATCH_PERSISTENT = 1  /* one workgroup runs one instance's whole
                                  solve; the batch is one launch             */
};

/* PCG recurrences (same Jacobi-PCG, same iterates in exact arithmetic). */
enum {
  FEMGPU_CG_STANDARD = 0,         /* the reference's recurrence
                                     (femfunc.cpp:59-111): two reductions
                                     per iteration, two launches           */
  FEMGPU_CG_SINGLE_REDUCTION = 1  /* Chronopoulos-Gear form: s = A p and
                                     w = A M^-1 r carried as recurrences,
                                     one reduction and one launch per
                                     iteration; no counterpart in the
                                     reference                              */
};

/* Load conventions for pressure faces and nodal loads. */
enum {
  FEMGPU_LOADS_REFERENCE = 0, /* the reference's CPU path bit for bit in
                                 meaning: F = c*p*len*n assigned per element
                                 slot, 3D len = |a x b|^2/2 with c = -1/12,
                                 component j dropped when (code & j), nodal
                                 loads gathered per element (cpu_fem_utils.cpp:
                                 564-596)                                    */
  FEMGPU_LOADS_PHYSICAL = 1   /* F = -p*A/npf*n per face node (A = face
                                 length/area), summed over faces, duplicate
                                 faces once, zero on constrained DOFs; nodal
                                 loads added once                            */
};

/* Mass lumping conventions for dynamics. */
enum {
  FEMGPU_MASS_REFERENCE = 0,  /* 2D rho*A/3, 3D rho*V/12 per DOF
                                 (cpu_fem_utils.cpp:598-734)                 */
  FEMGPU_MASS_PHYSICAL = 1    /* 2D rho*A/3, 3D rho*V/4 per DOF              */
};

/* --- context -------------------------------------------------------------
 * Replaces: the static cuBLAS handle and CUDA_Matrix::_initCUDA
 * (matrix_pack/src/cu/cuda_matrix.cpp:3-8, fem/src/main.cpp:18-20) and the
 * device choice DEVICE_NAME (matrix_pack_cpu.h:19-23). */
int femgpu_create(femgpu_ctx **ctx, int device_ordinal);
void femgpu_destroy(femgpu_ctx *ctx);
const char *femgpu_last_error(const femgpu_ctx *ctx);
/* Use a caller-created hipStream_t (passed as void*); NULL = own stream. */
int femgpu_set_stream(femgpu_ctx *ctx, void *hip_stream);
/* Synchronise the context's stream. */
int femgpu_sync(femgpu_ctx *ctx);

/* --- problem definition ---------------------------------------------------
 * Replaces: the dataKeeper readers (DataKeeper/src/datakeeper.cpp:116-213)
 * and the ElementsData constructor's host->device copies
 * (fem_utils/src/cu/cuda_fem_utils.cpp:13-47).  Index arrays are int32
 * (the reference keeps them as float, datakeeper.cpp:124-137). */

/* nodes: n_nodes x dim row-major; elems: n_elems x (dim+1) 0-based.
 * Builds the device-resident mesh (element patches, node renumbering);
 * resets material, constraints and loads. */
int femgpu_set_mesh(femgpu_ctx *ctx, int dim, int64_t n_nodes, const double *nodes,
                    int64_t n_elems, const int32_t *elems);
/* Launch tuning, applied at the next femgpu_set_mesh: threads per patch
 * workgroup (256, 512 or 1024; 0 = 256), which is also the node limit of a
 * patch, and elements per patch (<= 3 x threads; 0 = 3 x threads).
 * No counterpart in the reference (its kernels use 256 threads and one
 * thread per element, fem_utils_kernels.cu:3). */
int femgpu_set_tuning(femgpu_ctx *ctx, int elems_per_patch, int patch_threads);
/* Autotuner (no counterpart in the reference, whose kernels use fixed
 * 256-thread blocks and one element per thread): for the current problem
 * (mesh, material, constraints set), times the PCG iteration under each
 * candidate patch tuning (256, 512 or 1024 threads per patch with 1.5 to 3
 * elements per thread; the candidates are planned on host threads in
 * parallel) -- the difference between runs of 96 and 32 iterations on a unit
 * right-hand side, each the faster of two, so that a solve's fixed cost does
 * not count; FP64, or FP32 for FEMGPU_FP32 / FEMGPU_MIXED whose inner
 * iterations are FP32 -- re-plans the mesh with the fastest and keeps it as
 * the context's tuning (as femgpu_set_tuning).  The
 * choice is cached for the process per mesh (hash of coordinates and
 * connectivity) and arithmetic, so a later call on the same mesh only
 * re-plans.  Outputs may be NULL. */
int femgpu_autotune(femgpu_ctx *ctx, int precision, int *elems_per_patch, int *patch_threads,
                    double *ms_per_iteration);
/* The patch plan femgpu_set_mesh would build for this mesh and tuning,
 * computed on the host (needs no device): info[0] patches, [1] nodes owned
 * by one patch, [2] interface nodes, [3] interface partial-sum slots,
 * [4] elements per patch planned with, [5] most nodes in one patch.
 * No counterpart in the reference (it has no patches); for tests and
 * tuning tools. */
int femgpu_plan_info(int dim, int64_t n_nodes, const double *nodes, int64_t n_elems,
                     const int32_t *elems, int elems_per_patch, int patch_threads, int64_t *info);
/* Isotropic material (datakeeper.cpp:96-114, CreateMatrixD).  elem_scale:
 * optional per-element modulus factor (n_elems, user element order). */
int femgpu_set_material(femgpu_ctx *ctx, double E, double nu, const double *elem_scale);
/* Constraint lines "node code" (datakeeper.cpp:139-171); code decoded
 * bitwise as there: bit0 -> x, bit1 -> y, (code & 3) -> z in 3D. */
int femgpu_set_constraints(femgpu_ctx *ctx, int64_t n, const int32_t *node,
                           const int32_t *code);
/* Pressure faces (stress.txt, datakeeper.cpp:187-213): face_nodes n x dim,
 * adj_elem n, normal n x dim (normalised here), pressure n. */
int femgpu_set_pressure_faces(femgpu_ctx *ctx, int64_t n, const int32_t *face_nodes,
                              const int32_t *adj_elem, const double *normal,
                              const double *pressure);
/* Nodal loads (loads.txt, datakeeper.cpp:173-185): vec n x dim. */
int femgpu_set_nodal_loads(femgpu_ctx *ctx, int64_t n, const int32_t *node, const double *vec);
int femgpu_set_load_convention(femgpu_ctx *ctx, int convention);
/* Optional explicit global force vector (n_nodes x dim) added to the RHS
 * built from faces/loads (body forces, extra load cases).  NULL clears it. */
int femgpu_set_extra_rhs(femgpu_ctx *ctx, const double *f);

/* --- operators (test and integration hooks) -------------------------------
 * Replaces: Matrix::bmm of the per-element K_e with the element vectors plus
 * reductionWithMask (cpu_matrix.cpp:74-95, cpu_fem_utils.cpp:437-459): here
 * y = K x is one on-the-fly EbE product into assembled global vectors.
 * x, y: n_nodes x dim host arrays in user node order. */
int femgpu_matvec(femgpu_ctx *ctx, int precision, const double *x, double *y);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): y = K x like femgpu_matvec, on caller-owned DEVICE
 * arrays x_dev, y_dev (n_nodes x dim FP64, user node order; y is 0 at nodes
 * no element references), `precision` the arithmetic (FEMGPU_FP64 or
 * FEMGPU_FP32).  Allocates nothing: the internal-order copies of x and y and
 * the interface partial sums live in one fixed scratch array of 32 MiB per
 * GPU, held until the process ends; a problem needing more (2 x n_nodes x
 * dim values rounded up to 256 bytes each plus dim x interface slots + 1
 * values; the 64^3-hex tet cube needs 20.8 MB in FP64) is rejected with
 * FEMGPU_E_ARG.  Every call on one GPU shares that array, so calls on one
 * GPU that overlap (two contexts, two streams) race; calls on one context
 * are ordered by its stream.  Enqueued on the context stream, no
 * synchronisation.  The operator must be prepared by an allocating call
 * (femgpu_diag, femgpu_matvec or a solve) after the last change of mesh,
 * material or constraints, else FEMGPU_E_STATE. */
int femgpu_matvec_noalloc(femgpu_ctx *ctx, int precision, const double *x_dev, double *y_dev);
/* Assembled diagonal of K (Matrix::getDiagonal + reductionWithMask,
 * cpu_matrix.cpp:296-305; femfunc.cpp:10, :41). */
int femgpu_diag(femgpu_ctx *ctx, double *d);
/* Assembled RHS at time t with wavelet scale (ElementsData::initFlocals /
 * calcFlocals + reductionWithMask, cpu_fem_utils.cpp:522-596). */
int femgpu_rhs(femgpu_ctx *ctx, double *f);
/* Repeat the standalone FP32/FP64 EbE matvec `reps` times on device-resident
 * internal vectors, enqueued on the context stream without synchronising
 * (for timing with events on that stream). */
int femgpu_matvec_repeat(femgpu_ctx *ctx, int precision, int reps);
/* Average device time (ms) of one standalone EbE matvec over `reps`
 * back-to-back products, measured with HIP events on the context stream. */
int femgpu_time_matvec(femgpu_ctx *ctx, int precision, int reps, double *avg_ms);

/* --- static solve -----------------------------------------------------------
 * Replaces: gpuCalculateFEM_EbE_vec + gpuPCG_EbE_vec (femfunc.cpp:3-128). */
typedef struct {
  int precision;      /* FEMGPU_FP64 / FEMGPU_FP32 / FEMGPU_MIXED             */
  int tol_kind;       /* FEMGPU_TOL_REF_GAMMA / FEMGPU_TOL_REL_RES            */
  double tol;         /* 1e-8 in the reference's static path (femfunc.cpp:20) */
  int max_iter;       /* total PCG iteration cap; <= 0: 1,000,000             */
  double inner_tol;   /* MIXED: drop of the FP32 residual norm after which it
                         is replaced by the FP64 residual (default 0.1)       */
  int check_every;    /* iterations per graph replay between host polls of the
                         stop flag (default 64); replays the host expects to
                         contain the stop run check_every / 4 (even, >= 2)    */
  int recurrence;     /* FEMGPU_CG_STANDARD (default) / _SINGLE_REDUCTION     */
  int batch_schedule; /* femgpu_batch_solve: FEMGPU_BATCH_LOCKSTEP (default) /
                         FEMGPU_BATCH_PERSISTENT                              */
} femgpu_solve_opts;

typedef struct {
  int status;          /* FEMGPU_OK or an error code                          */
  int iterations;      /* PCG iterations (summed over inner solves in MIXED)  */
  int outer_iterations;/* MIXED: FP64 residual replacements                   */
  double rel_res;      /* true ||f - K u|| / ||f|| recomputed in FP64         */
  double gamma_ratio;  /* final gamma/gamma0 of the (last) PCG                */
  double solve_ms;     /* device time of the solve, setup excluded            */
} femgpu_stats;

int femgpu_default_solve_opts(femgpu_solve_opts *opts);
/* u_out: n_nodes x dim (user order); may be NULL (solution stays on device). */
int femgpu_solve_static(femgpu_ctx *ctx, const femgpu_solve_opts *opts, double *u_out,
                        femgpu_stats *stats);

/* Solve K u = f for an explicit right-hand side f (n_nodes x dim, user node
 * order) instead of the problem's loads; the same solver and options as
 * femgpu_solve_static.  Replaces: gpuPCG_EbE_vec called on a caller-filled
 * Flocals (femfunc.cpp:25-128), which the reference only fills from its
 * loads; here it is also the adjoint solve (K is symmetric). */
int femgpu_solve_rhs(femgpu_ctx *ctx, const femgpu_solve_opts *opts, const double *f,
                     double *u_out, femgpu_stats *stats);

/* --- adjoint material gradient ----------------------------------------------
 * No counterpart in the reference (it has no differentiation).  With element
 * modulus scales s_e (femgpu_set_material's elem_scale), K = sum_e s_e K~_e^0
 * where K~_e^0 is the element stiffness at scale 1 with the constrained-DOF
 * treatment of K.  For a loss L(u) with K u = f and adjoint K w = dL/du:
 *   dL/ds_e = -w_e' K~_e^0 u_e   (written to g_out, n_elems, user order).
 * K~_e^0 is recomputed on the fly, never stored. */
int femgpu_material_grad(femgpu_ctx *ctx, const double *u, const double *w, double *g_out);
/* Compliance C = f'u of the problem's loads and dC/ds_e = -u_e' K~_e^0 u_e
 * from ONE forward solve (the adjoint solution of C is u itself).
 * u_out (n_nodes x dim) may be NULL; c_out: 1 value; g_out: n_elems. */
int femgpu_compliance_grad(femgpu_ctx *ctx, const femgpu_solve_opts *opts, double *u_out,
                           double *c_out, double *g_out, femgpu_stats *stats);

/* --- batched independent static solves ----------------------------------------
 * No counterpart in the reference (it solves one problem per process run).
 * Many small meshes ("instances"), each with its own material, constraints
 * and 1..4 right-hand sides, solved in ONE persistent launch: one workgroup
 * runs an instance's whole Jacobi-PCG (the algorithm of gpuPCG_EbE_vec,
 * femfunc.cpp:25-128, with the relative-residual stopping rule), largest
 * instances first.  Meshes are concatenated: instance i owns user nodes
 * [node_off[i], node_off[i+1]) and elements [elem_off[i], elem_off[i+1]);
 * element node ids and constraint node ids are instance-local. */
typedef struct {
  int status;        /* FEMGPU_OK or the reason the system stopped          */
  int iterations;    /* PCG iterations                                       */
  int replacements;  /* FP64 residual replacements (MIXED)                   */
  int pad;
  double rel_res;    /* ||f - K u||_2 / ||f||_2 in FP64 at the end            */
} femgpu_batch_stat;

/* nodes: node_off[n_inst] x dim; elems: elem_off[n_inst] x (dim+1).
 * Replaces the instance's plan; material defaults to E = 2e7, nu = 0.3. */
int femgpu_batch_set_mesh(femgpu_ctx *ctx, int dim, int n_inst, const int64_t *node_off,
                          const double *nodes, const int64_t *elem_off, const int32_t *elems);
/* Per-instance Young's modulus and Poisson ratio (n_inst each). */
int femgpu_batch_set_material(femgpu_ctx *ctx, const double *E, const double *nu);
/* Constraint lines "node code" per instance (datakeeper.cpp:139-171 bitwise
 * codes): instance i holds entries [con_off[i], con_off[i+1]). */
int femgpu_batch_set_constraints(femgpu_ctx *ctx, const int64_t *con_off, const int32_t *node,
                                 const int32_t *code);
/* Solve every instance for n_rhs (1..4) right-hand sides.
 * f, u_out: node_off[n_inst] x n_rhs x dim (node-major, then rhs, then
 * component).  opts: precision FP64 or MIXED, tol_kind FEMGPU_TOL_REL_RES,
 * tol, max_iter (0: 100000), inner_tol (MIXED residual drop, default 0.1).
 * stats: n_inst x n_rhs (may be NULL); solve_ms: device time (may be NULL).
 * Returns FEMGPU_OK when every system converged, else the first failing
 * system's status (u_out and stats are filled either way). */
int femgpu_batch_solve(femgpu_ctx *ctx, int n_rhs, const double *f, const femgpu_solve_opts *opts,
                       double *u_out, femgpu_batch_stat *stats, double *solve_ms);

/* --- per-element results ----------------------------------------------------
 * No counterpart in the reference's live code; the golden fixtures hold
 * sigma = D B u_e per element (testing/<task>_test/results.vtk, TENSORS stress).
 * u: n_nodes x dim; sigma_out: n_elems x 3(dim-1) Voigt (xx,yy,xy | xx,yy,zz,
 * xy,yz,xz), user element order. */
int femgpu_stress(femgpu_ctx *ctx, const double *u, double *sigma_out);
/* Per-element engineering strain eps = B u_e (xx,yy,2xy | xx,yy,zz,2xy,2yz,
 * 2xz), n_elems x 3(dim-1), user element order.  The reference's live code
 * has none; its earlier post-processing computed this `Deformation = B
 * delta` per element (deprecated/femfunc/femfunc.cpp:42-99) for the legacy
 * results.vtk fields (deformation, XY_DEFORM, mises_deformation). */
int femgpu_strain(femgpu_ctx *ctx, const double *u, double *eps_out);
/* Element strain energies e_e = u_e' K_e u_e (n_elems). */
int femgpu_element_energy(femgpu_ctx *ctx, const double *u, double *e_out);

/* --- dynamics ---------------------------------------------------------------
 * Replaces: gpuCalculateFEM_DYN (femfunc.cpp:130-381): GN22 with beta2 == 0
 * (explicit central difference, lumped mass, solveDiagSystem) or beta2 != 0
 * (implicit, nested PCG per step).  Wavelet: fem_utils.cpp:53-82. */
enum { FEMGPU_WAVELET_NONE = 0, FEMGPU_WAVELET_RICKER = 1, FEMGPU_WAVELET_BERLAGE = 2 };
enum {
  FEMGPU_IMPLICIT_REFERENCE = 0, /* solve K a = F - K d, preconditioned by
                                    cK diag(K) + diag(M) (femfunc.cpp:204-212,302) */
  FEMGPU_IMPLICIT_PHYSICAL = 1   /* solve (M + beta2 dt^2/2 K) a = F - K d     */
};
typedef struct {
  double rho, beta1, beta2, dt, endtime;  /* endtime < 0: relaxation, see below  */
  int wavelet_type;
  double wavelet_freq, wavelet_timeshift, wavelet_ampl;
  int precision;        /* FEMGPU_FP64 or FEMGPU_FP32 state                   */
  int mass_convention;  /* FEMGPU_MASS_REFERENCE / FEMGPU_MASS_PHYSICAL        */
  int implicit_mode;    /* FEMGPU_IMPLICIT_REFERENCE / FEMGPU_IMPLICIT_PHYSICAL */
  double pcg_eps;       /* implicit inner PCG eps (1e-4 in femfunc.cpp:153)    */
  int steps_per_graph;  /* explicit: steps captured per HIP graph (0: 64)      */
  /* Rayleigh damping C = alpha M + beta K (the reference's damping_alpha /
   * damping_beta, femfunc.cpp:137-150,190-200,266-274, disabled there):
   * explicit: beta must be 0, (1 + alpha beta1 dt) M a = F - K d~ - alpha M v~;
   * implicit (FEMGPU_IMPLICIT_PHYSICAL only):
   * ((1 + alpha beta1 dt) M + (beta2 dt^2/2 + beta beta1 dt) K) a
   *     = F - K (d~ + beta v~) - alpha M v~                                  */
  double damping_alpha, damping_beta;
  /* endtime < 0 (relaxation, femfunc.cpp:213-221,349-356): with
   * relax_reference != 0 one step, as the reference (its C-norm is a stub
   * returning 0); else steps until max |v| < |endtime| (checked every
   * steps_per_graph explicit steps / every implicit step), at most max_steps. */
  int relax_reference;
  int max_steps;
} femgpu_dyn_opts;

typedef struct {
  int status;
  int steps;
  int pcg_iterations;  /* implicit: summed over steps                         */
  double run_ms;       /* device time of the time loop                        */
} femgpu_dyn_stats;

int femgpu_default_dyn_opts(femgpu_dyn_opts *opts);
int femgpu_dyn_run(femgpu_ctx *ctx, const femgpu_dyn_opts *opts, double *u_out,
                   femgpu_dyn_stats *stats);
/* Autotuner for the time steppers (no counterpart in the reference): as
 * femgpu_autotune, but every candidate patch tuning is timed on time steps
 * of femgpu_dyn_run with these options (explicit or implicit as opts->beta2
 * selects, endtime replaced): the difference between runs of n_steps and
 * n_steps / 4 steps (0: 256 explicit or 16 implicit steps), each the faster
 * of two, so that the choice fits the stepper's
 * own kernels (the explicit step's patch kernel and launch shape differ
 * from the PCG iteration's).  Re-plans the mesh with the fastest and keeps
 * it as the context's tuning; cached for the process per mesh, stepper kind
 * and precision.  The tuning runs are femgpu_dyn_run calls: receivers
 * record them (femgpu_get_traces then returns the last one).  Outputs may
 * be NULL. */
int femgpu_autotune_dyn(femgpu_ctx *ctx, const femgpu_dyn_opts *opts, int n_steps,
                        int *elems_per_patch, int *patch_threads, double *us_per_step);

/* Receivers (the reference's receiver traces for SEG-Y, writeDisplForSEGY,
 * VTKfile.cpp:72-87, called every segy_step steps in femfunc.cpp:226-243,
 * 357-366 where it is commented out): record the displacement, velocity or
 * acceleration of the given nodes (user ids) at every `every`-th time step
 * of the following femgpu_dyn_run calls; n = 0 clears.  Samples are taken
 * after the step's correction, at times k * every * dt, k = 1, 2, ... */
enum { FEMGPU_RECV_DISPLACEMENT = 0, FEMGPU_RECV_VELOCITY = 1, FEMGPU_RECV_ACCELERATION = 2 };
int femgpu_set_receivers(femgpu_ctx *ctx, int n, const int32_t *nodes, int quantity, int every);
/* Traces of the last femgpu_dyn_run: out[(k * n + r) * dim + c] for sample
 * k < *n_samples, receiver r, component c; capacity in doubles.  out may be
 * NULL to query *n_samples. */
int femgpu_get_traces(femgpu_ctx *ctx, double *out, int64_t capacity, int *n_samples);
/* SEG-Y file of n_traces traces of n_samples samples (traces[t * n_samples +
 * s]) in the layout of the reference's convertToSEGY (VTKfile.cpp:89-138,
 * format 5: big-endian IEEE floats, id in every trace header); host only,
 * needs no device. */
int femgpu_write_segy(const char *path, const float *traces, int n_traces, int n_samples,
                      float sample_interval, int trace_id);

#ifdef __cplusplus
}
#endif
#endif
