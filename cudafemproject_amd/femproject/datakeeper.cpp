// datakeeper.cpp -- see datakeeper.h.
#include "datakeeper.h"

#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "json_min.h"

namespace femproject {

namespace {

std::string join(const std::string &a, const std::string &b) {
  std::string s = a + "/" + b;
  std::string out;
  for (size_t i = 0; i < s.size(); ++i)
    if (!(s[i] == '/' && !out.empty() && out.back() == '/')) out += s[i];
  return out;
}

std::vector<std::vector<std::string>> read_lines(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<std::vector<std::string>> rows;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::vector<std::string> tok;
    std::string t;
    while (is >> t) tok.push_back(t);
    if (!tok.empty()) rows.push_back(tok);
  }
  return rows;
}

// All whitespace-separated numbers of a counted file: first the count, then
// `count * per` values (files may wrap lines arbitrarily).
std::vector<double> read_counted(const std::string &path, int per, int64_t *count) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  double n = 0;
  if (!(f >> n)) throw std::runtime_error("empty file " + path);
  *count = int64_t(n);
  std::vector<double> v(size_t(*count) * per);
  for (auto &x : v)
    if (!(f >> x)) throw std::runtime_error("truncated file " + path);
  return v;
}

}  // namespace

dataKeeper::dataKeeper(const std::string &configPath, const std::string &taskName)
    : taskName_(taskName) {
  parseJsonConfig(configPath);
  parseFiles();
}

dataKeeper dataKeeper::fromMeshDir(const std::string &meshDir, int dim, double E, double nu) {
  dataKeeper dk;
  dk.dim_ = dim;
  dk.mech_.youngModulus = E;
  dk.mech_.poissonRatio = nu;
  dk.paths_.prep_mesh_dir = meshDir;
  dk.paths_.res_dir = meshDir;
  dk.paths_.output_vtk = join(meshDir, "results.vtk");
  dk.parseFiles();
  return dk;
}

// datakeeper.cpp:6-48: the keys, their defaults, and "dynamic" as the only
// task_type that turns the dynamics parameters on.
void dataKeeper::parseJsonConfig(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open config " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const Json cfg = Json::parse(ss.str());
  dim_ = int(cfg.at("DIM").num);
  if (dim_ != 2 && dim_ != 3) throw std::runtime_error("DIM must be 2 or 3");
  paths_.proj_dir = cfg.string_or("CudaFEMproject_path", "undefined_path");
  paths_.work_dir = cfg.string_or("working_path", "undefined_path");
  paths_.res_dir = join(paths_.work_dir, "/fem_results/" + std::to_string(dim_) + "D/" + taskName_);
  paths_.prep_mesh_dir =
      join(paths_.work_dir, "/prepared_meshes/" + std::to_string(dim_) + "D/" + taskName_);
  paths_.output_vtk = join(paths_.res_dir, "/results.vtk");
  mech_.youngModulus = cfg.number_or("youngModulus", 2e7);
  mech_.poissonRatio = cfg.number_or("poissonRatio", 0.3);
  verbose_ = cfg.bool_or("verbose", false);
  taskType_ = cfg.string_or("task_type", "static");
  if (taskType_ == "dynamic") {
    mech_.rho = cfg.number_or("rho", 2400.0);
    mech_.beta1 = cfg.number_or("beta1", 0.5);
    mech_.beta2 = cfg.number_or("beta2", 0.0);
    mech_.damping_alpha = cfg.number_or("damping_alpha", 0.0);
    mech_.damping_beta = cfg.number_or("damping_beta", 0.0);
    mech_.dt = cfg.number_or("dt", 1e-3);
    mech_.endtime = cfg.number_or("endtime", 0.1);
  }
  if (cfg.has("wavelet")) {
    const Json &w = cfg.at("wavelet");
    wave_.waveletType = w.string_or("WaveletType", "default");
    wave_.freq = w.number_or("freq", 15.0);
    wave_.timeshift = w.number_or("timeshift", 0.0);
    wave_.ampl = w.number_or("ampl", 1.0);
  }
}

// datakeeper.cpp:50-213.  stress.txt lines are "n_1..n_dim n_interior adjElem
// normal p" (the current layout) or "n_1..n_dim adjElem normal p" (the layout
// of the legacy golden fixtures); the field count tells them apart.
void dataKeeper::parseFiles() {
  const std::string &d = paths_.prep_mesh_dir;
  int64_t nn = 0, ne = 0, nc = 0, nl = 0;
  nodes_ = read_counted(join(d, "nodes.txt"), dim_, &nn);
  std::vector<double> e = read_counted(join(d, "elements.txt"), dim_ + 1, &ne);
  if (nn <= 0 || ne <= 0) throw std::runtime_error("empty mesh in " + d);
  elems_.assign(e.size(), 0);
  for (size_t i = 0; i < e.size(); ++i) elems_[i] = int32_t(e[i]);
  std::vector<double> c = read_counted(join(d, "constraints.txt"), 2, &nc);
  constrNodes_.resize(nc);
  constrCodes_.resize(nc);
  for (int64_t i = 0; i < nc; ++i) {
    constrNodes_[i] = int32_t(c[2 * i]);
    constrCodes_[i] = int32_t(c[2 * i + 1]);
  }
  std::vector<double> l = read_counted(join(d, "loads.txt"), dim_ + 1, &nl);
  loadNodes_.resize(nl);
  loadVecs_.resize(size_t(nl) * dim_);
  for (int64_t i = 0; i < nl; ++i) {
    loadNodes_[i] = int32_t(l[i * (dim_ + 1)]);
    for (int k = 0; k < dim_; ++k) loadVecs_[i * dim_ + k] = l[i * (dim_ + 1) + 1 + k];
  }
  auto rows = read_lines(join(d, "stress.txt"));
  if (!rows.empty()) {
    const int64_t nb = int64_t(std::stod(rows[0][0]));
    if (int64_t(rows.size()) - 1 < nb) throw std::runtime_error("truncated stress.txt");
    for (int64_t b = 0; b < nb; ++b) {
      const auto &r = rows[b + 1];
      std::vector<double> v;
      for (auto &t : r) v.push_back(std::stod(t));
      int off;
      if (int(v.size()) == 2 * dim_ + 3)
        off = 1;  // interior node present
      else if (int(v.size()) == 2 * dim_ + 2)
        off = 0;
      else
        throw std::runtime_error("stress.txt line with " + std::to_string(v.size()) + " fields");
      for (int k = 0; k < dim_; ++k) faceNodes_.push_back(int32_t(v[k]));
      faceAdj_.push_back(int32_t(v[dim_ + off]));
      for (int k = 0; k < dim_; ++k) faceNormals_.push_back(v[dim_ + off + 1 + k]);
      facePress_.push_back(v[2 * dim_ + off + 1]);
    }
  }
  displ_.resize(size_t(nn), size_t(dim_));
}

void dataKeeper::setOutput(const std::string &resDir) {
  paths_.res_dir = resDir;
  paths_.output_vtk = join(resDir, "results.vtk");
}

void dataKeeper::ShowInfo() const {
  std::cout << "==========INFO==========\nTask name = " << taskName_
            << "\nNodes count = " << get_nodesCount() << "\nElements count = " << get_elementsCount()
            << "\nConstraints count = " << get_constraintsCount()
            << "\nLoads count = " << get_loadsCount()
            << "\nBoundary edges with b.c. count = " << get_boundaryEdgesCount()
            << "\n========================\n\n";
}

}  // namespace femproject
