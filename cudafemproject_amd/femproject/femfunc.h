// femfunc.h -- the reference's solver entry points over the femgpu C-ABI.
//
// Same names, argument meaning and call order as femfunc/src/femfunc.h:20-24:
//   gpuCalculateFEM_EbE_vec(device, dk, verbose)   static solve    (femfunc.cpp:3-23)
//   gpuPCG_EbE_vec(displ, device, dk, elemsData, doAssemblyRes, eps, verbose)
//                                                   the PCG itself   (femfunc.cpp:25-128)
//   gpuCalculateFEM_DYN(device, dk, verbose)        GN22 dynamics   (femfunc.cpp:130-381)
//   WriteResults(dk, RESdata)                       results.vtk     (femfunc.cpp:384-390)
// Results land in dk.get_displacements() as in the reference (femfunc.cpp:117,369).
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

// The element data of a problem on the device (fem_utils/include/
// fem_utils.h ElementsData, made by ElementsData::setElementsData(device, dk),
// femfunc.cpp:5): here a femgpu context holding the planned mesh (patches
// instead of the reference's mask), material, constraints and loads; the
// element matrices are never stored.  genMask / calculateKlocals /
// initFlocals keep the reference's call sequence (femfunc.cpp:7-15); they
// prepare the constraint mask and assembled diagonal and the load vector on
// the device.
class ElementsData {
 public:
  static ElementsData *setElementsData(DEVICE_NAME device, const dataKeeper &dk,
                                       const SolverOptions &opt = SolverOptions());
  ~ElementsData();
  ElementsData(const ElementsData &) = delete;
  ElementsData &operator=(const ElementsData &) = delete;

  void genMask() {}             // the patch plan is built by setElementsData
  void calculateKlocals();      // operator set-up: material, constraint mask, diag(K)
  void initFlocals(float t, const WaveletParams &wavelet);  // load vector (t, wavelet unused
                                                           // by the static path)
  femgpu_ctx *handle() const { return h_; }
  const SolverOptions &options() const { return opt_; }

 private:
  ElementsData() = default;
  femgpu_ctx *h_ = nullptr;
  SolverOptions opt_;
  int64_t nn_ = 0;
  int dim_ = 0;
};

// Static solve (femfunc.cpp:3-23): displacements into dk.get_displacements(),
// per-element stress and strain into dk.stresses() / dk.strains().
void gpuCalculateFEM_EbE_vec(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                             const SolverOptions &opt = SolverOptions());
// The PCG on elemsData's problem (femfunc.cpp:25-128), stopping at
// gamma < eps gamma0 (opt.precision MIXED: ||f - K u|| <= eps ||f||).
// displacements: resized to n_nodes x dim.  doAssemblyRes is kept for the
// signature: the result is always the assembled nodal vector here.
void gpuPCG_EbE_vec(Matrix &displacements, DEVICE_NAME device, dataKeeper &dk,
                    ElementsData &elemsData, bool doAssemblyRes, float eps,
                    bool PRINT_DEBUG_INFO);
void gpuCalculateFEM_DYN(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                         const SolverOptions &opt = SolverOptions());
// results.vtk (femfunc.cpp:384-390): fills RESdata from the last static
// solve's per-element results (strain, stress, their von Mises values and,
// with RESdata.withSmooth, the nodal projection of sigma_xx) and writes the
// legacy field layout of the reference's fixtures (testing/*_test/
// results.vtk; VTKfile.cpp:141-279).  Without per-element results (dynamics,
// compute_stress = false) only displacements and |u| are written, as
// VTKfile.cpp:281-385.
void WriteResults(dataKeeper &dk, ResultsDataKeeper &RESdata);

// Legacy-ASCII VTK writers (VTKfile/src/VTKfile.cpp:141-385).
void MakeVTKfile2D(const std::string &path, const dataKeeper &dk, const ResultsDataKeeper *res);
void MakeVTKfile3D(const std::string &path, const dataKeeper &dk, const ResultsDataKeeper *res);

}  // namespace femproject

// The reference declares these names at global scope (femfunc.h:20-24,
// datakeeper.h, matrix_pack.h): a caller written against it uses them
// unqualified.
using femproject::CPU;
using femproject::CUDA;
using femproject::DEVICE_NAME;
using femproject::ElementsData;
using femproject::HIP;
using femproject::Matrix;
using femproject::ONEAPI;
using femproject::ResultsDataKeeper;
using femproject::WriteResults;
using femproject::dataKeeper;
using femproject::gpuCalculateFEM_DYN;
using femproject::gpuCalculateFEM_EbE_vec;
using femproject::gpuPCG_EbE_vec;
