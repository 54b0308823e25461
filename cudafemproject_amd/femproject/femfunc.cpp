// femfunc.cpp -- see femfunc.h.
#include "femfunc.h"

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>

namespace femproject {

namespace {

struct Ctx {
  femgpu_ctx *h = nullptr;
  explicit Ctx(int device) {
    const int st = femgpu_create(&h, device);
    if (st != FEMGPU_OK) throw std::runtime_error("femgpu_create failed: " + std::to_string(st));
  }
  ~Ctx() { femgpu_destroy(h); }
  void check(int st, const char *what) const {
    if (st != FEMGPU_OK)
      throw std::runtime_error(std::string(what) + ": " + femgpu_last_error(h) + " (" +
                               std::to_string(st) + ")");
  }
};

void require_hip(DEVICE_NAME d) {
  if (d != HIP)
    throw std::runtime_error("this build implements the HIP device only (DEVICE_NAME::HIP)");
}

// ElementsData::setElementsData + the loads of initFlocals (fem_utils.cpp:41-51,
// cpu_fem_utils.cpp:522-530): the whole problem goes to the device once.
void load_problem(const Ctx &c, const dataKeeper &dk, const SolverOptions &opt) {
  const int dim = dk.get_dim();
  c.check(femgpu_set_mesh(c.h, dim, dk.get_nodesCount(), dk.nodes().data(), dk.get_elementsCount(),
                          dk.elements().data()),
          "set_mesh");
  c.check(femgpu_set_material(c.h, dk.getMechParams().youngModulus, dk.getMechParams().poissonRatio,
                              nullptr),
          "set_material");
  if (dk.get_constraintsCount() > 0)
    c.check(femgpu_set_constraints(c.h, dk.get_constraintsCount(), dk.constraintNodes().data(),
                                   dk.constraintCodes().data()),
            "set_constraints");
  if (dk.get_boundaryEdgesCount() > 0)
    c.check(femgpu_set_pressure_faces(c.h, dk.get_boundaryEdgesCount(), dk.faceNodes().data(),
                                      dk.faceAdjElems().data(), dk.faceNormals().data(),
                                      dk.facePressures().data()),
            "set_pressure_faces");
  if (dk.get_loadsCount() > 0)
    c.check(femgpu_set_nodal_loads(c.h, dk.get_loadsCount(), dk.loadNodes().data(),
                                   dk.loadVectors().data()),
            "set_nodal_loads");
  c.check(femgpu_set_load_convention(c.h, opt.load_convention), "set_load_convention");
}

void solve_static(const Ctx &c, dataKeeper &dk, double eps, int tol_kind, const SolverOptions &opt,
                  bool verbose) {
  femgpu_solve_opts so;
  femgpu_default_solve_opts(&so);
  so.precision = opt.precision;
  so.tol_kind = tol_kind;
  so.tol = eps;
  if (so.precision == FEMGPU_MIXED) so.tol_kind = FEMGPU_TOL_REL_RES;
  femgpu_stats st{};
  c.check(femgpu_solve_static(c.h, &so, dk.displacements().data(), &st), "solve_static");
  if (verbose)
    std::cout << "PCG: " << st.iterations << " iterations, rel. residual " << st.rel_res
              << ", gamma/gamma0 " << st.gamma_ratio << ", " << st.solve_ms << " ms\n";
  if (opt.compute_stress) {
    dk.stresses().assign(size_t(dk.get_elementsCount()) * 3 * (dk.get_dim() - 1), 0.0);
    c.check(femgpu_stress(c.h, dk.displacements().data(), dk.stresses().data()), "stress");
  }
}

int wavelet_code(const std::string &t) {
  if (t == "ricker") return FEMGPU_WAVELET_RICKER;
  if (t == "berlage") return FEMGPU_WAVELET_BERLAGE;
  return FEMGPU_WAVELET_NONE;  // fem_utils.cpp:71-82: anything else gives 0
}

}  // namespace

void gpuCalculateFEM_EbE_vec(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                             const SolverOptions &opt) {
  require_hip(device);
  Ctx c(opt.device);
  load_problem(c, dk, opt);
  solve_static(c, dk, opt.tol, opt.tol_kind, opt, PRINT_DEBUG_INFO);
}

void gpuPCG_EbE_vec(std::vector<double> &out, DEVICE_NAME device, dataKeeper &dk,
                    bool /*doAssemblyRes*/, float eps, bool PRINT_DEBUG_INFO,
                    const SolverOptions &opt) {
  require_hip(device);
  Ctx c(opt.device);
  load_problem(c, dk, opt);
  SolverOptions o = opt;
  o.compute_stress = false;
  solve_static(c, dk, double(eps), FEMGPU_TOL_REF_GAMMA, o, PRINT_DEBUG_INFO);
  out = dk.displacements();
}

void gpuCalculateFEM_DYN(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                         const SolverOptions &opt) {
  require_hip(device);
  Ctx c(opt.device);
  load_problem(c, dk, opt);
  const MechanicalParams &m = dk.getMechParams();
  const WaveletParams &w = dk.getWaveletParams();
  femgpu_dyn_opts o;
  femgpu_default_dyn_opts(&o);
  o.rho = m.rho;
  o.beta1 = m.beta1;
  o.beta2 = m.beta2;
  o.dt = m.dt;
  o.endtime = m.endtime;
  o.wavelet_type = wavelet_code(w.waveletType);
  o.wavelet_freq = w.freq;
  o.wavelet_timeshift = w.timeshift;
  o.wavelet_ampl = w.ampl;
  o.precision = opt.precision == FEMGPU_FP32 ? FEMGPU_FP32 : FEMGPU_FP64;
  o.mass_convention = opt.mass_convention;
  o.implicit_mode = opt.implicit_mode;
  o.pcg_eps = opt.pcg_eps_dynamic;
  o.damping_alpha = opt.apply_damping ? m.damping_alpha : 0.0;
  o.damping_beta = opt.apply_damping ? m.damping_beta : 0.0;
  o.relax_reference = opt.relax_reference ? 1 : 0;
  const int nrec = int(opt.receivers.size());
  if (nrec)
    c.check(femgpu_set_receivers(c.h, nrec, opt.receivers.data(), opt.receiver_quantity,
                                 opt.receiver_every),
            "set_receivers");
  femgpu_dyn_stats st{};
  c.check(femgpu_dyn_run(c.h, &o, dk.displacements().data(), &st), "dyn_run");
  if (PRINT_DEBUG_INFO)
    std::cout << "GN22: " << st.steps << " steps" << (m.beta2 == 0 ? " (explicit)" : " (implicit)")
              << ", " << st.pcg_iterations << " PCG iterations, " << st.run_ms << " ms\n";
  if (nrec) {
    // receivers.txt as the reference's receiver file (femfunc.cpp:226-243 header,
    // writeDisplForSEGY rows, VTKfile.cpp:72-87) and receivers.sgy (convertToSEGY)
    int ns = 0;
    c.check(femgpu_get_traces(c.h, nullptr, 0, &ns), "get_traces");
    const int d = dk.get_dim();
    std::vector<double> tr(size_t(ns) * nrec * d);
    c.check(femgpu_get_traces(c.h, tr.data(), int64_t(tr.size()), &ns), "get_traces");
    const int ax = std::min(std::max(opt.receiver_axis, 0), d - 1);
    const double dts = m.dt * opt.receiver_every;
    const std::string dir = dk.getDataPaths().res_dir;
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    std::ofstream out(dir + "/receivers.txt");
    if (!out) throw std::runtime_error("cannot write " + dir + "/receivers.txt");
    out << std::setprecision(9) << nrec << "\t";
    for (int r = 0; r < nrec; ++r) out << opt.receivers[r] << "\t";
    out << "\n";
    std::vector<float> seg(size_t(nrec) * ns);
    for (int k = 0; k < ns; ++k) {
      out << float((k + 1) * dts) << " ";
      for (int r = 0; r < nrec; ++r) {
        const double v = tr[(size_t(k) * nrec + r) * d + ax];
        out << float(v) << " ";
        seg[size_t(r) * ns + k] = float(v);
      }
      out << "\n";
    }
    c.check(femgpu_write_segy((dir + "/receivers.sgy").c_str(), seg.data(), nrec, ns, float(dts), 0),
            "write_segy");
  }
}

void WriteResults(dataKeeper &dk) {
  std::error_code ec;
  std::filesystem::create_directories(
      std::filesystem::path(dk.getDataPaths().output_vtk).parent_path(), ec);
  if (dk.get_dim() == 2)
    MakeVTKfile2D(dk.getDataPaths().output_vtk, dk);
  else
    MakeVTKfile3D(dk.getDataPaths().output_vtk, dk);
}

namespace {

void write_vtk(const std::string &path, const dataKeeper &dk, bool is2d) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write " + path);
  out << std::setprecision(9);
  const int d = dk.get_dim();
  const int64_t nn = dk.get_nodesCount(), ne = dk.get_elementsCount();
  const auto &X = dk.nodes();
  const auto &E = dk.elements();
  const auto &U = dk.displacements();
  out << "# vtk DataFile Version 1.0\nresults.vtk  " << (is2d ? "2D" : "3D")
      << " Unstructured Grid of Linear " << (is2d ? "Triangles" : "Tetrahedra")
      << "\nASCII\n\nDATASET UNSTRUCTURED_GRID\nPOINTS " << nn << " float\n";
  for (int64_t i = 0; i < nn; ++i) {
    for (int k = 0; k < d; ++k) out << X[i * d + k] << " ";
    out << (is2d ? "0.0\n" : "\n");
  }
  out << "\nCELLS " << ne << " " << ne * (d + 2) << "\n";
  for (int64_t e = 0; e < ne; ++e) {
    out << d + 1;
    for (int a = 0; a <= d; ++a) out << " " << E[e * (d + 1) + a];
    out << "\n";
  }
  out << "\nCELL_TYPES " << ne << "\n";
  for (int64_t e = 0; e < ne; ++e) out << (is2d ? 5 : 10) << "\n";
  out << "\nPOINT_DATA " << nn << "\n\nVECTORS displacements float\n";
  for (int64_t i = 0; i < nn; ++i) {
    for (int k = 0; k < d; ++k) out << U[i * d + k] << " ";
    out << (is2d ? "0.0\n" : "\n");
  }
  out << "\nSCALARS summary float\nLOOKUP_TABLE default\n";
  for (int64_t i = 0; i < nn; ++i) {
    double s = 0;
    for (int k = 0; k < d; ++k) s += U[i * d + k] * U[i * d + k];
    out << std::sqrt(s) << "\n";
  }
  const std::vector<double> &S = const_cast<dataKeeper &>(dk).stresses();
  const int nv = 3 * (d - 1);
  if (S.size() == size_t(ne) * nv) {
    // the per-element tensor block of the legacy fixtures (Voigt -> 3x3)
    out << "\nCELL_DATA " << ne << "\nTENSORS stress float\n";
    for (int64_t e = 0; e < ne; ++e) {
      const double *s = &S[e * nv];
      if (is2d) {
        out << s[0] << " " << s[2] << " 0\n" << s[2] << " " << s[1] << " 0\n0 0 0\n";
      } else {
        out << s[0] << " " << s[3] << " " << s[5] << "\n"
            << s[3] << " " << s[1] << " " << s[4] << "\n"
            << s[5] << " " << s[4] << " " << s[2] << "\n";
      }
    }
    out << "\nSCALARS mises_stress float\nLOOKUP_TABLE default\n";
    for (int64_t e = 0; e < ne; ++e) {
      const double *s = &S[e * nv];
      double vm;
      if (is2d)
        vm = std::sqrt(s[0] * s[0] - s[0] * s[1] + s[1] * s[1] + 3 * s[2] * s[2]);
      else
        vm = std::sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                              (s[2] - s[0]) * (s[2] - s[0])) +
                       3 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
      out << vm << "\n";
    }
  }
}

}  // namespace

void MakeVTKfile2D(const std::string &path, const dataKeeper &dk) { write_vtk(path, dk, true); }
void MakeVTKfile3D(const std::string &path, const dataKeeper &dk) { write_vtk(path, dk, false); }

}  // namespace femproject
