/*
 * oracle.h -- CPU restatement of the reference's EbE linear-elasticity path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this library: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Every function exists in an FP32 variant (suffix _f32) that repeats the
 * reference's float arithmetic in the reference's order (element-local
 * vectors, per-DOF sums in element-major order, sequential dot products) and
 * an FP64 variant (suffix _f64) that runs the same algorithm in double.
 * Citations are to files of the reference repository (moarseniy/CudaFEMproject).
 *
 * Pinning (see tests/test_oracle.py):
 *   - orc_static_solve_f32 vs the two legacy golden fixtures
 *     (testing/test_rect_pcg_test, testing/9task_3_test) and vs the
 *     reference's own CPU build (oracle/_ref) on 2D and 3D meshes;
 *   - orc_klocals_*: symmetry, rigid-body null space, |det C| = 2A / 6V,
 *     patch test, brute-force dense assembly vs numpy.linalg.solve;
 *   - orc_matvec_*, orc_diag_*: against dense assembly of orc_klocals_*;
 *   - orc_dyn_solve_f32 vs oracle/_ref (explicit and implicit);
 *   - orc_stress_*: vs the stresses held in the golden fixtures.
 */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define ORC_OK            0
#define ORC_E_ARG        -1   /* bad argument / malformed mesh            */
#define ORC_E_NONFINITE  -2   /* gamma became NaN/Inf (femfunc.cpp:107)   */
#define ORC_E_MAXIT      -3   /* iteration cap reached                     */
#define ORC_E_UNSUPPORTED -4  /* input the reference handles by accident   */

/* A prepared mesh, as read by the reference's dataKeeper
 * (DataKeeper/src/datakeeper.cpp:116-213).  All values are given in double;
 * the _f32 functions round them to float first, as `stream >> float` does. */
typedef struct {
  int dim;                   /* 2 or 3                                      */
  int n_nodes;
  int n_elems;
  const double *nodes;       /* n_nodes x dim, row-major                    */
  const int *elems;          /* n_elems x (dim+1)                           */
  int n_constr;              /* raw constraints.txt lines (dups allowed)    */
  const int *constr_node;
  const int *constr_code;    /* Constraint::Type 1..7, decoded bitwise      */
  int n_faces;               /* stress.txt lines                            */
  const int *face_nodes;     /* n_faces x dim                               */
  const int *face_adj;       /* adjacent element of each face               */
  const double *face_normal; /* n_faces x dim, NOT yet normalised           */
  const double *face_p;      /* pressure per face                           */
  int n_loads;               /* loads.txt lines                             */
  const int *load_node;
  const double *load_vec;    /* n_loads x dim                               */
  double E, nu;              /* youngModulus, poissonRatio                  */
  const double *elem_scale;  /* optional per-element modulus factor (NULL)  */
} orc_problem;

/* Wavelet (fem_utils/src/fem_utils.cpp:53-82). type: 0 none, 1 ricker, 2 berlage */
typedef struct {
  int type;
  double freq, timeshift, ampl;
} orc_wavelet;

/* GN22 dynamics parameters (femfunc/src/femfunc.cpp:130-381). */
typedef struct {
  double rho, beta1, beta2, dt, endtime;
  orc_wavelet wavelet;
  int compat_scramble;  /* 1: repeat the reference's in-place permutation of
                           Mlocals / diagK by sort_by_key (cpu_fem_utils.cpp:437-445) */
  double pcg_eps;       /* implicit inner PCG stopping constant; 0 -> 1e-4 as
                           the reference (femfunc.cpp:153)                      */
} orc_dyn_params;

/* D matrix, row-major 3x3 (2D) or 6x6 (3D) (datakeeper.cpp:96-114). */
void orc_D_f32(int dim, double E, double nu, float *D);
void orc_D_f64(int dim, double E, double nu, double *D);

/* Element stiffness matrices with the reference's constraint zeroing
 * (cpu_fem_utils.cpp:505-520).  K_out: n_elems x (ndof_e*ndof_e), each block
 * stored column-major as the reference's bmm writes it; areas_out: |det C|. */
int orc_klocals_f32(const orc_problem *pb, float *K_out, float *areas_out);
int orc_klocals_f64(const orc_problem *pb, double *K_out, double *areas_out);

/* Element-local right-hand side at time t (cpu_fem_utils.cpp:522-596). */
int orc_flocals_f32(const orc_problem *pb, double t, const orc_wavelet *w, float *F_out);
int orc_flocals_f64(const orc_problem *pb, double t, const orc_wavelet *w, double *F_out);

/* Assembled global vectors (n_nodes*dim). */
int orc_rhs_f32(const orc_problem *pb, float *f_out);
int orc_rhs_f64(const orc_problem *pb, double *f_out);
int orc_diag_f32(const orc_problem *pb, float *d_out);
int orc_diag_f64(const orc_problem *pb, double *d_out);
/* y = K x with K the assembled constrained stiffness (global vectors). */
int orc_matvec_f32(const orc_problem *pb, const float *x, float *y);
int orc_matvec_f64(const orc_problem *pb, const double *x, double *y);

/* Static solve: gpuCalculateFEM_EbE_vec + gpuPCG_EbE_vec (femfunc.cpp:3-128).
 * eps is the reference's stopping constant (gamma_new < eps*gamma0);
 * max_iter <= 0 means no cap (as the reference).  u_out: n_nodes x dim. */
int orc_static_solve_f32(const orc_problem *pb, double eps, int max_iter,
                         float *u_out, int *n_iter);
int orc_static_solve_f64(const orc_problem *pb, double eps, int max_iter,
                         double *u_out, int *n_iter);

/* Time-limited timing helper for the CPU baseline: runs setup, then exactly
 * n_iter PCG iterations (no convergence test) and reports the seconds spent
 * in setup and in the iterations. */
int orc_static_time_f32(const orc_problem *pb, int n_iter,
                        double *setup_s, double *iter_s);
/* The same timing for the FP64 restatement with its loops spread over
 * n_threads OpenMP threads (n_threads <= 0: OpenMP's default). */
int orc_static_time_omp_f64(const orc_problem *pb, int n_iter, int n_threads,
                            double *setup_s, double *iter_s);

/* Dynamics: gpuCalculateFEM_DYN (femfunc.cpp:130-381). u_out: n_nodes x dim. */
int orc_dyn_solve_f32(const orc_problem *pb, const orc_dyn_params *dp,
                      float *u_out, int *n_steps);
int orc_dyn_solve_f64(const orc_problem *pb, const orc_dyn_params *dp,
                      double *u_out, int *n_steps);

/* Per-element stress sigma = D B u_e (Voigt; 3 or 6 values per element). */
int orc_stress_f32(const orc_problem *pb, const float *u, float *sigma_out);
int orc_stress_f64(const orc_problem *pb, const double *u, double *sigma_out);

/* Wavelet value exactly as ElementsData::updateWavelet (float). */
float orc_wavelet_value(double t, const orc_wavelet *w);
double orc_wavelet_value_f64(double t, const orc_wavelet *w);

#ifdef __cplusplus
}
#endif
#endif
