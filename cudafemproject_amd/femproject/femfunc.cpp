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
void load_problem(const Ctx &c, const dataKeeper &dk, const SolverOptions &opt);

}  // namespace

ElementsData *ElementsData::setElementsData(DEVICE_NAME device, const dataKeeper &dk,
                                            const SolverOptions &opt) {
  require_hip(device);
  std::unique_ptr<ElementsData> ed(new ElementsData());
  ed->opt_ = opt;
  Ctx c(opt.device);
  load_problem(c, dk, opt);
  ed->h_ = c.h;
  c.h = nullptr;  // owned by ed from here
  ed->nn_ = dk.get_nodesCount();
  ed->dim_ = dk.get_dim();
  return ed.release();
}

ElementsData::~ElementsData() { femgpu_destroy(h_); }

void ElementsData::calculateKlocals() {
  std::vector<double> d(size_t(nn_) * dim_);
  const int st = femgpu_diag(h_, d.data());
  if (st != FEMGPU_OK)
    throw std::runtime_error(std::string("calculateKlocals: ") + femgpu_last_error(h_));
}

void ElementsData::initFlocals(float /*t*/, const WaveletParams & /*wavelet*/) {
  std::vector<double> f(size_t(nn_) * dim_);
  const int st = femgpu_rhs(h_, f.data());
  if (st != FEMGPU_OK)
    throw std::runtime_error(std::string("initFlocals: ") + femgpu_last_error(h_));
}

namespace {

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

// The PCG of femgpu_solve_static on h's problem into out (n_nodes x dim).
void pcg_solve(femgpu_ctx *h, Matrix &out, int64_t nn, int dim, double eps, int tol_kind,
               int precision, bool verbose) {
  femgpu_solve_opts so;
  femgpu_default_solve_opts(&so);
  so.precision = precision;
  so.tol_kind = precision == FEMGPU_MIXED ? FEMGPU_TOL_REL_RES : tol_kind;
  so.tol = eps;
  femgpu_stats st{};
  if (out.get_numRows() != size_t(nn) || out.get_numCols() != size_t(dim))
    out.resize(size_t(nn), size_t(dim));
  const int rc = femgpu_solve_static(h, &so, out.get_data(), &st);
  if (rc != FEMGPU_OK)
    throw std::runtime_error(std::string("solve_static: ") + femgpu_last_error(h) + " (" +
                             std::to_string(rc) + ")");
  if (verbose)
    std::cout << "PCG: " << st.iterations << " iterations, rel. residual " << st.rel_res
              << ", gamma/gamma0 " << st.gamma_ratio << ", " << st.solve_ms << " ms\n";
}

// Per-element stress and strain of the solution (the fields of results.vtk).
void element_results(femgpu_ctx *h, dataKeeper &dk) {
  const size_t n = size_t(dk.get_elementsCount()) * 3 * (dk.get_dim() - 1);
  dk.stresses().assign(n, 0.0);
  dk.strains().assign(n, 0.0);
  int rc = femgpu_stress(h, dk.displacements().data(), dk.stresses().data());
  if (rc == FEMGPU_OK) rc = femgpu_strain(h, dk.displacements().data(), dk.strains().data());
  if (rc != FEMGPU_OK)
    throw std::runtime_error(std::string("element results: ") + femgpu_last_error(h));
}

int wavelet_code(const std::string &t) {
  if (t == "ricker") return FEMGPU_WAVELET_RICKER;
  if (t == "berlage") return FEMGPU_WAVELET_BERLAGE;
  return FEMGPU_WAVELET_NONE;  // fem_utils.cpp:71-82: anything else gives 0
}

}  // namespace

void gpuCalculateFEM_EbE_vec(DEVICE_NAME device, dataKeeper &dk, bool PRINT_DEBUG_INFO,
                             const SolverOptions &opt) {
  std::unique_ptr<ElementsData> ed(ElementsData::setElementsData(device, dk, opt));
  ed->genMask();
  ed->calculateKlocals();
  ed->initFlocals(0.f, dk.getWaveletParams());
  pcg_solve(ed->handle(), *dk.get_displacements(), dk.get_nodesCount(), dk.get_dim(), opt.tol,
            opt.tol_kind, opt.precision, PRINT_DEBUG_INFO);
  dk.stresses().clear();
  dk.strains().clear();
  if (opt.compute_stress) element_results(ed->handle(), dk);
}

void gpuPCG_EbE_vec(Matrix &displacements, DEVICE_NAME device, dataKeeper &dk,
                    ElementsData &elemsData, bool /*doAssemblyRes*/, float eps,
                    bool PRINT_DEBUG_INFO) {
  require_hip(device);
  const SolverOptions &o = elemsData.options();
  pcg_solve(elemsData.handle(), displacements, dk.get_nodesCount(), dk.get_dim(), double(eps),
            o.tol_kind, o.precision, PRINT_DEBUG_INFO);
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
  dk.stresses().clear();
  dk.strains().clear();
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

namespace {

// von Mises value of a Voigt vector (xx,yy,xy | xx,yy,zz,xy,yz,xz) with
// engineering shears for strains as for stresses: the functional
// sqrt(((a-b)^2 + (b-c)^2 + (c-a)^2) / 2 + 3 (s_xy^2 + s_yz^2 + s_xz^2)),
// in 2D with the out-of-plane component zero: sqrt(a^2 - a b + b^2 + 3 s^2)
// (the fixtures' mises_stress and mises_deformation).
double mises(const double *v, int dim) {
  if (dim == 2) return std::sqrt(v[0] * v[0] - v[0] * v[1] + v[1] * v[1] + 3.0 * v[2] * v[2]);
  const double a = v[0] - v[1], b = v[1] - v[2], c = v[2] - v[0];
  return std::sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]));
}

// L2 projection of the element-constant field s_e onto the linear nodal
// basis: M S = b with the consistent mass matrix M = sum_e |e| (1 + d_ij) /
// ((d+1)(d+2)) over the element's nodes and b_i = sum_e s_e |e| / (d+1),
// solved by Jacobi-preconditioned CG (element by element, M is well
// conditioned) to a relative residual of 1e-14.  Nodes no element touches
// get 0.
std::vector<double> project_to_nodes(const dataKeeper &dk, const std::vector<double> &s, int nv,
                                     int comp) {
  const int d = dk.get_dim(), npe = d + 1;
  const int64_t nn = dk.get_nodesCount(), ne = dk.get_elementsCount();
  const auto &X = dk.nodes();
  const auto &E = dk.elements();
  std::vector<double> vol(ne), b(nn, 0.0), diag(nn, 0.0);
  const double den = double(npe) * double(npe + 1);
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t *n = &E[size_t(e) * npe];
    double v;
    if (d == 2) {
      const double *a = &X[size_t(n[0]) * 2], *p = &X[size_t(n[1]) * 2], *q = &X[size_t(n[2]) * 2];
      v = 0.5 * std::fabs((p[0] - a[0]) * (q[1] - a[1]) - (p[1] - a[1]) * (q[0] - a[0]));
    } else {
      const double *a = &X[size_t(n[0]) * 3], *p = &X[size_t(n[1]) * 3], *q = &X[size_t(n[2]) * 3],
                   *r = &X[size_t(n[3]) * 3];
      const double e1[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
      const double e2[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
      const double e3[3] = {r[0] - a[0], r[1] - a[1], r[2] - a[2]};
      v = std::fabs(e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                    e1[2] * (e2[0] * e3[1] - e2[1] * e3[0])) / 6.0;
    }
    vol[e] = v;
    for (int a = 0; a < npe; ++a) {
      b[n[a]] += s[size_t(e) * nv + comp] * v / npe;
      diag[n[a]] += 2.0 * v / den;
    }
  }
  auto apply = [&](const std::vector<double> &x, std::vector<double> &y) {
    std::fill(y.begin(), y.end(), 0.0);
    for (int64_t e = 0; e < ne; ++e) {
      const int32_t *n = &E[size_t(e) * npe];
      double sum = 0.0;
      for (int a = 0; a < npe; ++a) sum += x[n[a]];
      const double w = vol[e] / den;
      for (int a = 0; a < npe; ++a) y[n[a]] += w * (sum + x[n[a]]);
    }
  };
  std::vector<double> x(nn, 0.0), r = b, z(nn), p(nn), q(nn);
  for (int64_t i = 0; i < nn; ++i) z[i] = diag[i] > 0 ? r[i] / diag[i] : 0.0;
  p = z;
  double rz = 0.0, bb = 0.0;
  for (int64_t i = 0; i < nn; ++i) {
    rz += r[i] * z[i];
    bb += b[i] * b[i];
  }
  for (int it = 0; it < 1000 && bb > 0; ++it) {
    apply(p, q);
    double pq = 0.0;
    for (int64_t i = 0; i < nn; ++i) pq += p[i] * q[i];
    if (!(pq > 0)) break;
    const double al = rz / pq;
    double rr = 0.0;
    for (int64_t i = 0; i < nn; ++i) {
      x[i] += al * p[i];
      r[i] -= al * q[i];
      rr += r[i] * r[i];
    }
    if (rr <= 1e-28 * bb) break;
    double rz2 = 0.0;
    for (int64_t i = 0; i < nn; ++i) {
      z[i] = diag[i] > 0 ? r[i] / diag[i] : 0.0;
      rz2 += r[i] * z[i];
    }
    const double be = rz2 / rz;
    rz = rz2;
    for (int64_t i = 0; i < nn; ++i) p[i] = z[i] + be * p[i];
  }
  return x;
}

void make_results(const dataKeeper &dk, ResultsDataKeeper &R) {
  const int d = dk.get_dim(), nv = 3 * (d - 1);
  const int64_t ne = dk.get_elementsCount();
  R.Stress = dk.stresses();
  R.Deformation = dk.strains();
  R.sigma_mises.assign(ne, 0.0);
  R.epsilon_mises.assign(ne, 0.0);
  for (int64_t e = 0; e < ne; ++e) {
    R.sigma_mises[e] = mises(&R.Stress[size_t(e) * nv], d);
    R.epsilon_mises[e] = mises(&R.Deformation[size_t(e) * nv], d);
  }
  if (R.withSmooth)
    R.SmoothStress = project_to_nodes(dk, R.Stress, nv, 0);  // "xx" (femfunc.cpp:1248-1255)
  else
    R.SmoothStress.assign(size_t(dk.get_nodesCount()), 0.0);
}

}  // namespace

void WriteResults(dataKeeper &dk, ResultsDataKeeper &RESdata) {
  std::error_code ec;
  std::filesystem::create_directories(
      std::filesystem::path(dk.getDataPaths().output_vtk).parent_path(), ec);
  const size_t nv = size_t(3 * (dk.get_dim() - 1));
  const bool have = !dk.stresses().empty() &&
                    dk.stresses().size() == size_t(dk.get_elementsCount()) * nv &&
                    dk.strains().size() == dk.stresses().size();
  if (have) make_results(dk, RESdata);
  if (dk.get_dim() == 2)
    MakeVTKfile2D(dk.getDataPaths().output_vtk, dk, have ? &RESdata : nullptr);
  else
    MakeVTKfile3D(dk.getDataPaths().output_vtk, dk, have ? &RESdata : nullptr);
}

namespace {

void write_vtk(const std::string &path, const dataKeeper &dk, const ResultsDataKeeper *R,
               bool is2d) {
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
  if (!R) return;
  // the field layout of the reference's fixtures (testing/*_test/results.vtk)
  const int nv = 3 * (d - 1);
  out << "\nSCALARS SmoothStress float\nLOOKUP_TABLE default\n";
  for (int64_t i = 0; i < nn; ++i) out << R->SmoothStress[i] << "\n";
  // symmetric 3x3 tensor of a Voigt vector; shear factor 1 (stress) or 1/2
  // (engineering strain -> tensor strain)
  auto tensor = [&](const double *s, double h) {
    if (is2d) {
      out << s[0] << " " << h * s[2] << " 0\n" << h * s[2] << " " << s[1] << " 0\n0 0 0\n";
    } else {
      out << s[0] << " " << h * s[3] << " " << h * s[5] << "\n"
          << h * s[3] << " " << s[1] << " " << h * s[4] << "\n"
          << h * s[5] << " " << h * s[4] << " " << s[2] << "\n";
    }
  };
  const int ixy = is2d ? 2 : 3;
  out << "\nCELL_DATA " << ne << "\nTENSORS stress float\n";
  for (int64_t e = 0; e < ne; ++e) tensor(&R->Stress[size_t(e) * nv], 1.0);
  out << "\nSCALARS XY_STRESS float\nLOOKUP_TABLE default\n";
  for (int64_t e = 0; e < ne; ++e) out << R->Stress[size_t(e) * nv + ixy] << "\n";
  out << "\nSCALARS XY_DEFORM float\nLOOKUP_TABLE default\n";
  for (int64_t e = 0; e < ne; ++e) out << R->Deformation[size_t(e) * nv + ixy] << "\n";
  out << "\nSCALARS mises_stress float\nLOOKUP_TABLE default\n";
  for (int64_t e = 0; e < ne; ++e) out << R->sigma_mises[e] << "\n";
  if (is2d) {
    out << "\nVECTORS deformation float\n";
    for (int64_t e = 0; e < ne; ++e) {
      const double *s = &R->Deformation[size_t(e) * nv];
      out << s[0] << " " << s[1] << " " << s[2] << "\n";
    }
  } else {
    out << "\nTENSORS deformation float\n";
    for (int64_t e = 0; e < ne; ++e) tensor(&R->Deformation[size_t(e) * nv], 0.5);
  }
  out << "\nSCALARS mises_deformation float\nLOOKUP_TABLE default\n";
  for (int64_t e = 0; e < ne; ++e) out << R->epsilon_mises[e] << "\n";
}

}  // namespace

void MakeVTKfile2D(const std::string &path, const dataKeeper &dk, const ResultsDataKeeper *res) {
  write_vtk(path, dk, res, true);
}
void MakeVTKfile3D(const std::string &path, const dataKeeper &dk, const ResultsDataKeeper *res) {
  write_vtk(path, dk, res, false);
}

}  // namespace femproject
