// main.cpp -- the FEMproject executable (fem/src/main.cpp:10-36) on the HIP
// backend.  Unlike the reference's main, which hard-codes its config path,
// task and device (main.cpp:18-22), this one takes them from argv:
//
//   fem_main <run_config.json> <task> [options]
//   fem_main <task> <dim> <proj_dir> <prep_mesh_dir> <results_dir> <nu> <E> <verbose>
//            [1 <rho> <damping_alpha> <damping_beta> <dt> <endtime> <beta1> <beta2>]
//            (the argv run_fem.py builds, run_fem.py:80-84; note nu before E)
// options: --precision fp64|fp32|mixed  --tol X  --tol-kind gamma|relres
//          --loads reference|physical  --mass reference|physical
//          --implicit reference|physical  --no-stress  --device N
//          --dry-run (read the problem, print its summary, no GPU)
//          --smooth (results.vtk SmoothStress: nodal L2 projection of sigma_xx)
//          --relax (endtime < 0: step until max |v| < |endtime|)
//          --damping (apply the config's damping_alpha / damping_beta)
//          --receivers n1,n2,... --receiver-axis 0|1|2
//          --receiver-quantity displacement|velocity|acceleration --segy-every K
//          (dynamics: receivers.txt and receivers.sgy next to results.vtk)
//
// It reads the prepared mesh, runs the static or dynamic solve on the GPU
// and writes results.vtk, as main.cpp:21-33 does.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "datakeeper.h"
#include "femfunc.h"

using namespace femproject;

namespace {

bool ends_with(const std::string &s, const std::string &t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

bool g_dry_run = false;
bool g_smooth = false;

int parse_opts(int argc, char **argv, int first, SolverOptions &o, bool &tol_set) {
  for (int i = first; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
      return argv[++i];
    };
    if (a == "--precision") {
      const std::string v = val();
      o.precision = v == "fp32" ? FEMGPU_FP32 : v == "mixed" ? FEMGPU_MIXED : FEMGPU_FP64;
      if (o.precision == FEMGPU_MIXED) o.tol_kind = FEMGPU_TOL_REL_RES;
    } else if (a == "--tol") {
      o.tol = std::atof(val().c_str());
      tol_set = true;
    } else if (a == "--tol-kind") {
      o.tol_kind = val() == "relres" ? FEMGPU_TOL_REL_RES : FEMGPU_TOL_REF_GAMMA;
    } else if (a == "--loads") {
      o.load_convention = val() == "physical" ? FEMGPU_LOADS_PHYSICAL : FEMGPU_LOADS_REFERENCE;
    } else if (a == "--mass") {
      o.mass_convention = val() == "physical" ? FEMGPU_MASS_PHYSICAL : FEMGPU_MASS_REFERENCE;
    } else if (a == "--implicit") {
      o.implicit_mode = val() == "physical" ? FEMGPU_IMPLICIT_PHYSICAL : FEMGPU_IMPLICIT_REFERENCE;
    } else if (a == "--no-stress") {
      o.compute_stress = false;
    } else if (a == "--smooth") {
      g_smooth = true;
    } else if (a == "--dry-run") {
      g_dry_run = true;
    } else if (a == "--device") {
      o.device = std::atoi(val().c_str());
    } else if (a == "--relax") {
      o.relax_reference = false;
    } else if (a == "--damping") {
      o.apply_damping = true;
    } else if (a == "--receivers") {
      o.receivers.clear();
      std::string v = val();
      size_t p = 0;
      while (p < v.size()) {
        size_t q = v.find(',', p);
        if (q == std::string::npos) q = v.size();
        o.receivers.push_back(std::atoi(v.substr(p, q - p).c_str()));
        p = q + 1;
      }
    } else if (a == "--receiver-axis") {
      o.receiver_axis = std::atoi(val().c_str());
    } else if (a == "--receiver-quantity") {
      const std::string v = val();
      o.receiver_quantity = v == "displacement" ? FEMGPU_RECV_DISPLACEMENT
                            : v == "acceleration" ? FEMGPU_RECV_ACCELERATION
                                                  : FEMGPU_RECV_VELOCITY;
    } else if (a == "--segy-every") {
      o.receiver_every = std::max(1, std::atoi(val().c_str()));
    } else {
      throw std::runtime_error("unknown option " + a);
    }
  }
  return 0;
}

int run(dataKeeper &dk, const SolverOptions &opt, bool verbose) {
  if (g_dry_run) {
    dk.ShowInfo();
    const auto &m = dk.getMechParams();
    std::cout << "dim " << dk.get_dim() << " task_type " << dk.getTaskType() << " E "
              << m.youngModulus << " nu " << m.poissonRatio << " rho " << m.rho << " dt " << m.dt
              << " endtime " << m.endtime << " beta1 " << m.beta1 << " beta2 " << m.beta2
              << " wavelet " << dk.getWaveletParams().waveletType << "\n";
    std::cout << "prep_mesh_dir " << dk.getDataPaths().prep_mesh_dir << "\noutput_vtk "
              << dk.getDataPaths().output_vtk << "\n";
    return 0;
  }
  if (verbose) dk.ShowInfo();
  if (dk.getTaskType() == "dynamic")
    gpuCalculateFEM_DYN(HIP, dk, verbose, opt);
  else
    gpuCalculateFEM_EbE_vec(HIP, dk, verbose, opt);
  ResultsDataKeeper RESdata(false, false, false, dk.get_nodesCount());  // main.cpp:30-33
  RESdata.withSmooth = g_smooth;
  WriteResults(dk, RESdata);
  if (verbose) std::cout << "wrote " << dk.getDataPaths().output_vtk << "\n";
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc >= 3 && ends_with(argv[1], ".json")) {
      SolverOptions opt;
      bool tol_set = false;
      parse_opts(argc, argv, 3, opt, tol_set);
      dataKeeper dk(argv[1], argv[2]);
      return run(dk, opt, dk.verbose());
    }
    if (argc >= 9) {
      // run_fem.py argv: task dim proj prep res nu E verbose [1 rho da db dt endtime b1 b2]
      const std::string prep = argv[4], res = argv[5];
      const int dim = std::atoi(argv[2]);
      const double nu = std::atof(argv[6]), E = std::atof(argv[7]);
      const bool verbose = std::atoi(argv[8]) != 0;
      dataKeeper dk = dataKeeper::fromMeshDir(prep, dim, E, nu);
      dk.setOutput(res);
      if (argc >= 17 && std::string(argv[9]) == "1") {
        MechanicalParams m;
        m.youngModulus = E;
        m.poissonRatio = nu;
        m.rho = std::atof(argv[10]);
        m.damping_alpha = std::atof(argv[11]);
        m.damping_beta = std::atof(argv[12]);
        m.dt = std::atof(argv[13]);
        m.endtime = std::atof(argv[14]);
        m.beta1 = std::atof(argv[15]);
        m.beta2 = std::atof(argv[16]);
        dk.setDynamics(m);
      }
      SolverOptions opt;
      return run(dk, opt, verbose);
    }
    std::cerr << "usage: fem_main <run_config.json> <task> [options]\n"
                 "       fem_main <task> <dim> <proj> <prep_mesh_dir> <results_dir> <nu> <E> "
                 "<verbose> [dynamics args]\n";
    return 2;
  } catch (const std::exception &e) {
    std::cerr << "fem_main: " << e.what() << "\n";
    return 1;
  }
}
