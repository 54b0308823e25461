// femfunc.h -- the reference's solver entry points over the femgpu C-ABI.
//
// Same names, argument meaning and call order as femfunc/src/femfunc.h:20-24:
//   gpuCalculateFEM_EbE_vec(device, dk, verbose)   static solve    (femfunc.cpp:3-23)
//   gpuPCG_EbE_vec(out, device, dk, ..., eps, ...)  the PCG itself   (femfunc.cpp:25-128)
//   gpuCalculateFEM_DYN(device, dk, verbose)        GN22 dynamics   (femfunc.cpp:130-381)
//   WriteResults(dk)                                results.vtk     (femfunc.cpp:384-390)
// Results land in dk.displacements() as in the reference (femfunc.cpp:117,369).
// DEVICE_NAME gains HIP (matrix_pack_cpu.h:19-23 has CPU, CUDA, ONEAPI); this
// build implements HIP only and throws for the others.
#pragma once

#include <string>
#include <vector>

#include "datakeeper.h"
#include "femgpu.h"

namespace femproject {

enum DEVICE_NAME { CPU, CUDA, ONEAPI, HIP };

// Choices the reference hard-codes, made explicit.  Defaults reproduce the
// reference: FP64 arithmetic of its algorithm, its stopping rule
// gamma < 1e-8 gamma0 (femfunc.cpp:20), its load and mass conventions.
struct SolverOptions {
  int precision = FEMGPU_FP64;
  int tol_kind = FEMGPU_TOL_REF_GAMMA;
  double tol = 1e-8;
  int load_convention = FEMGPU_LOADS_REFERENCE;
  int mass_convention = FEMGPU_MASS_REFERENCE;
  int implicit_mode = FEMGPU_IMPLICIT_REFERENCE;
  double pcg_eps_dynamic = 1e-4;  // femfunc.cpp:153
  bool compute_stress = true;     // per-element sigma for results.vtk
  int device = 0;
  // endtime < 0: true = the reference's relaxation stub (one step), false =
  // step until max |v| < |endtime| (femfunc.cpp:213-221,349-356)
  bool relax_reference = true;
  // Rayleigh damping from the run config (damping_alpha, damping_beta): the
  // reference reads them but never applies them (isDamping = false,
  // femfunc.cpp:151); false keeps that behaviour.
  bool apply_damping = false;
  // Receivers (femfunc.cpp:226-243,357-366, commented out there): node ids,
  // recorded quantity, component written to receivers.txt / receivers.sgy,
  // sampling period in steps.
  std::vector<int32_t> receivers;
  int receiver_quantity = FEMGPU_RECV_VELOCITY;
  int receiver_axis = 1;
  int receiver_every = 1;
};

void gpuCalculateFEM_EbE_vec(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                             const SolverOptions &opt = SolverOptions());
// The PCG on the problem held by dk (eps: the reference's stopping constant).
// out: n_nodes x dim; doAssemblyRes is kept for the signature (results are
// always assembled nodal vectors here).
void gpuPCG_EbE_vec(std::vector<double> &out, DEVICE_NAME device, dataKeeper &dk,
                    bool doAssemblyRes, float eps, bool PRINT_DEBUG_INFO,
                    const SolverOptions &opt = SolverOptions());
void gpuCalculateFEM_DYN(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                         const SolverOptions &opt = SolverOptions());
void WriteResults(dataKeeper &dk);

// Legacy-ASCII VTK writers (VTKfile/src/VTKfile.cpp:281-385): displacements
// and |u|; with stresses, also the per-element TENSORS stress block of the
// golden fixtures.
void MakeVTKfile2D(const std::string &path, const dataKeeper &dk);
void MakeVTKfile3D(const std::string &path, const dataKeeper &dk);

}  // namespace femproject
