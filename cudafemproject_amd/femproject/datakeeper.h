// datakeeper.h -- run_config + prepared-mesh reader of the FEMproject-compatible
// host layer.
//
// Same construction and meaning as the reference's dataKeeper
// (DataKeeper/src/datakeeper.h:56-139, datakeeper.cpp:6-252):
// dataKeeper(configPath, taskName) reads the JSON run config, derives
//   prep_mesh_dir = working_path/prepared_meshes/<DIM>D/<task>
//   res_dir       = working_path/fem_results/<DIM>D/<task>
//   output_vtk    = res_dir/results.vtk
// and reads nodes/elements/constraints/loads/stress.txt from prep_mesh_dir.
// Arrays are std::vector (double values, int32 indices) instead of the
// reference's float Matrix; errors throw std::runtime_error instead of exit().
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace femproject {

// Host matrix of doubles, row-major (rows x cols): the role of the
// reference's Matrix for the results the solver hands back
// (matrix_pack/include/matrix_pack.h; dk.get_displacements() is n x dim).
class Matrix {
 public:
  Matrix() = default;
  Matrix(size_t rows, size_t cols) : rows_(rows), cols_(cols), v_(rows * cols, 0.0) {}
  void resize(size_t rows, size_t cols) {
    rows_ = rows;
    cols_ = cols;
    v_.assign(rows * cols, 0.0);
  }
  size_t get_numRows() const { return rows_; }
  size_t get_numCols() const { return cols_; }
  size_t get_numElements() const { return v_.size(); }
  double &operator()(size_t i, size_t j) { return v_[i * cols_ + j]; }
  double operator()(size_t i, size_t j) const { return v_[i * cols_ + j]; }
  double &operator[](size_t k) { return v_[k]; }
  double operator[](size_t k) const { return v_[k]; }
  double *get_data() { return v_.data(); }
  const double *get_data() const { return v_.data(); }
  std::vector<double> &values() { return v_; }
  const std::vector<double> &values() const { return v_; }

 private:
  size_t rows_ = 0, cols_ = 0;
  std::vector<double> v_;
};

struct MechanicalParams {  // datakeeper.h:21-30
  double poissonRatio = 0.3, youngModulus = 2e7, rho = 2400;
  double damping_alpha = 0, damping_beta = 0;
  double dt = 1e-3, endtime = 0.1;
  double beta1 = 0.5, beta2 = 0;
};

struct WaveletParams {  // datakeeper.h:32-37
  std::string waveletType;
  double timeshift = 0, freq = 15, ampl = 1;
};

struct DataPaths {  // datakeeper.h:39-53
  std::string proj_dir, prep_mesh_dir, work_dir, res_dir, output_vtk;
};

class dataKeeper {
 public:
  dataKeeper(const std::string &configPath, const std::string &taskName);
  // Direct construction from a prepared-mesh directory (no run config).
  static dataKeeper fromMeshDir(const std::string &meshDir, int dim, double E, double nu);

  std::string getTaskName() const { return taskName_; }
  std::string getTaskType() const { return taskType_; }
  const MechanicalParams &getMechParams() const { return mech_; }
  const WaveletParams &getWaveletParams() const { return wave_; }
  const DataPaths &getDataPaths() const { return paths_; }
  bool verbose() const { return verbose_; }

  int get_dim() const { return dim_; }
  int64_t get_nodesCount() const { return int64_t(nodes_.size()) / dim_; }
  int64_t get_elementsCount() const { return int64_t(elems_.size()) / (dim_ + 1); }
  int64_t get_boundaryEdgesCount() const { return int64_t(faceAdj_.size()); }
  int64_t get_loadsCount() const { return int64_t(loadNodes_.size()); }
  int64_t get_constraintsCount() const { return int64_t(constrNodes_.size()); }

  const std::vector<double> &nodes() const { return nodes_; }          // n x dim
  const std::vector<int32_t> &elements() const { return elems_; }      // m x (dim+1)
  const std::vector<int32_t> &constraintNodes() const { return constrNodes_; }
  const std::vector<int32_t> &constraintCodes() const { return constrCodes_; }
  const std::vector<int32_t> &faceNodes() const { return faceNodes_; }  // b x dim
  const std::vector<int32_t> &faceAdjElems() const { return faceAdj_; }
  const std::vector<double> &faceNormals() const { return faceNormals_; }  // b x dim, raw
  const std::vector<double> &facePressures() const { return facePress_; }
  const std::vector<int32_t> &loadNodes() const { return loadNodes_; }
  const std::vector<double> &loadVectors() const { return loadVecs_; }  // l x dim

  // set the task type and mechanical parameters when there is no run config
  void setDynamics(const MechanicalParams &m) {
    taskType_ = "dynamic";
    mech_ = m;
  }
  void setOutput(const std::string &resDir);

  std::vector<double> &displacements() { return displ_.values(); }    // n x dim
  const std::vector<double> &displacements() const { return displ_.values(); }
  Matrix *get_displacements() { return &displ_; }                     // datakeeper.h:116
  // per-element results of the last static solve (Voigt, 3(dim-1) each),
  // empty when they were not computed
  std::vector<double> &stresses() { return stress_; }                 // sigma = D B u_e
  const std::vector<double> &stresses() const { return stress_; }
  std::vector<double> &strains() { return strain_; }                  // B u_e (engineering)
  const std::vector<double> &strains() const { return strain_; }

  void ShowInfo() const;

 private:
  dataKeeper() = default;
  void parseJsonConfig(const std::string &path);
  void parseFiles();

  std::string taskName_, taskType_ = "static";
  int dim_ = 0;
  bool verbose_ = false;
  MechanicalParams mech_;
  WaveletParams wave_;
  DataPaths paths_;
  std::vector<double> nodes_;
  std::vector<int32_t> elems_;
  std::vector<int32_t> constrNodes_, constrCodes_;
  std::vector<int32_t> faceNodes_, faceAdj_;
  std::vector<double> faceNormals_, facePress_;
  std::vector<int32_t> loadNodes_;
  std::vector<double> loadVecs_;
  Matrix displ_;
  std::vector<double> stress_, strain_;
};

// Per-element and nodal post-processing results written to results.vtk
// (DataKeeper/src/datakeeper.h:200-226, same constructor).  Stress and
// Deformation hold 3(dim-1) Voigt values per element (Deformation: B u_e,
// engineering shears), sigma_mises / epsilon_mises one value per element,
// SmoothStress one value per node: the L2 projection of sigma_xx onto the
// linear nodal basis when withSmooth, else zeros (the reference's main
// passes withSmooth = false, and its fixtures hold zeros).
// withStressAlongAxis (the reference's line-sampling text files) is kept
// for the signature and writes nothing.
class ResultsDataKeeper {
 public:
  ResultsDataKeeper(bool withStressAlongAxis, bool withSmooth, bool withMises, int64_t nodesCount)
      : withStressAlongAxis(withStressAlongAxis), withSmooth(withSmooth), withMises(withMises),
        SmoothStress(size_t(nodesCount > 0 ? nodesCount : 0), 0.0) {}

  bool withStressAlongAxis;
  bool withSmooth;
  bool withMises;

  std::vector<double> Deformation;
  std::vector<double> Stress;
  std::vector<double> sigma_mises;
  std::vector<double> epsilon_mises;
  std::vector<double> SmoothStress;
};

}  // namespace femproject
