// voomgpu.hpp -- C++ host classes over the C ABI (voomgpu.h) that keep the
// reference's interface, so a driver written against voomEP's CardiacBody /
// EPSolver / IonicModel (Body/CardiacBody.h:82-353, Solvers/EPSolver.h:51-308,
// IonicModel/IonicModel.h:16-50) runs on the GPU with std::vector in place of
// Epetra vectors.  Header-only; link with -lvoomgpu.  Errors throw
// voomgpu::Error carrying vg_last_error().
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "voomgpu.h"

namespace voomgpu {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Mesh arrays owned by the caller's containers (0-based ids).
struct Mesh {
  std::vector<double> xyz;      // 3 per node
  std::vector<int8_t> elem_nn;  // 4 (tet) or 2 (bar)
  std::vector<int32_t> conn;    // 4 slots per element
  std::vector<double> fibers;   // 9 per element, or empty = isotropic
  std::vector<int8_t> tissue;   // per node, or empty
  std::vector<int32_t> pmj_node;
  std::vector<double> pmj_data;  // 4 per PMJ: cosines, length
  double radius = 0.04;

  vg_mesh view() const {
    vg_mesh m{};
    m.n_nodes = (int64_t)(xyz.size() / 3);
    m.xyz = xyz.data();
    m.n_elems = (int64_t)elem_nn.size();
    m.elem_nn = elem_nn.data();
    m.conn = conn.data();
    m.fibers = fibers.empty() ? nullptr : fibers.data();
    m.tissue = tissue.empty() ? nullptr : tissue.data();
    m.n_pmj = (int64_t)pmj_node.size();
    m.pmj_node = pmj_node.empty() ? nullptr : pmj_node.data();
    m.pmj_data = pmj_data.empty() ? nullptr : pmj_data.data();
    m.radius = radius;
    return m;
  }
};

// Mesh::New's files (<model>.node/.ele[/.fibers/.purk/.cell/.angle],
// Mesh/Mesh.cc:142-416) into a Mesh; ids become 0-based.
inline Mesh readMesh(const std::string &model, double radius = 0.0275) {
  char err[1024] = {0};
  vg_meshfile *mf = nullptr;
  const int rc = vg_mesh_read(model.c_str(), radius, &mf, err, sizeof err);
  if (rc != VG_OK) throw Error(rc, std::string("readMesh: ") + err);
  const vg_mesh *v = vg_meshfile_mesh(mf);
  Mesh m;
  m.xyz.assign(v->xyz, v->xyz + 3 * v->n_nodes);
  m.elem_nn.assign(v->elem_nn, v->elem_nn + v->n_elems);
  m.conn.assign(v->conn, v->conn + 4 * v->n_elems);
  if (v->fibers) m.fibers.assign(v->fibers, v->fibers + 9 * v->n_elems);
  if (v->tissue) m.tissue.assign(v->tissue, v->tissue + v->n_nodes);
  if (v->n_pmj) {
    m.pmj_node.assign(v->pmj_node, v->pmj_node + v->n_pmj);
    m.pmj_data.assign(v->pmj_data, v->pmj_data + 4 * v->n_pmj);
  }
  m.radius = v->radius;
  vg_meshfile_free(mf);
  return m;
}

inline vg_model default_model(int cell_model) {
  vg_model m{};
  m.cell_model = cell_model;
  m.precision = VG_PREC_FP64;
  m.integrator = VG_INTEG_REFERENCE;
  // EPSimulator.cc:84-113 / CardiacBody.cc:22-27 defaults.
  for (int i = 0; i < 3; i++) {
    m.D[i] = 0.001;
    m.Dp[i] = 0.0016;
    m.DpLeft[i] = 0.0016;
  }
  m.Cm = 1.0;
  m.dt = 0.1;
  m.dt_ionic = 0.1;
  m.cg_tol = 1e-7;  // EP_CrankNicholson.cc:104
  m.cg_maxit = 1000;
  return m;
}

// CardiacBody + EP_CrankNicholson on one GPU.
class GpuCNSolver {
 public:
  GpuCNSolver(const Mesh &mesh, const vg_model &model) : mesh_(mesh) {
    const vg_mesh mv = mesh_.view();
    const int rc = vg_create(&mv, &model, &h_);
    if (rc != VG_OK) {
      std::string msg = h_ ? vg_last_error(h_) : "vg_create failed";
      if (h_) vg_destroy(h_);
      h_ = nullptr;
      throw Error(rc, "vg_create: " + msg);
    }
    n_ = (size_t)mv.n_nodes;
  }
  ~GpuCNSolver() {
    if (h_) vg_destroy(h_);
  }
  GpuCNSolver(const GpuCNSolver &) = delete;
  GpuCNSolver &operator=(const GpuCNSolver &) = delete;

  // EP_CrankNicholson::Solve (EP_CrankNicholson.cc:78-156), n times.
  void Solve(int nsteps = 1) { check(vg_step(h_, nsteps), "Solve"); }

  // CardiacBody::setStimulusCurrent overloads (CardiacBody.cc:243-324).
  void setStimulusCurrent(const std::vector<int32_t> &nodeSet, double istim = 0.0) {
    check(vg_set_stimulus_nodes(h_, nodeSet.data(), (int64_t)nodeSet.size(), istim), "stimulus");
  }
  void setStimulusCurrent(const double xc[3], double rad, double stim = 0.0) {
    check(vg_set_stimulus_sphere(h_, xc, rad, stim), "stimulus");
  }
  void setStimulusCurrent(int axis, double limit, double stim = 0.0) {
    check(vg_set_stimulus_axis(h_, axis, limit, stim), "stimulus");
  }
  void setStimulusCurrent(const double xo[3], const double xn[3], double stim = 0.0) {
    check(vg_set_stimulus_plane(h_, xo, xn, stim), "stimulus");
  }

  // EPSolver::InitializeVoltage and the voltage view written by WriteOutput.
  void InitializeVoltage(const std::vector<double> &V) {
    if (V.size() != n_) throw Error(VG_ERR_ARG, "InitializeVoltage: one value per node");
    check(vg_set_voltage(h_, V.data()), "InitializeVoltage");
  }
  void InitializeVoltage(double v) { InitializeVoltage(std::vector<double>(n_, v)); }
  std::vector<double> getVoltage() {
    std::vector<double> V(n_);
    check(vg_get_voltage(h_, V.data()), "getVoltage");
    return V;
  }

  // CardiacBody::getStateVariables / setStateVariables (CardiacBody.cc:438-463):
  // per node its own model's states (Purkinje nodes carry 22, tT myocardium
  // 19, ...), V first.
  void getStateVariables(std::vector<int> &numVar, std::vector<std::vector<double>> &data) {
    std::vector<int32_t> nv(n_);
    check(vg_get_states_ragged(h_, nv.data(), nullptr), "getStateVariables");
    size_t tot = 0;
    for (int32_t k : nv) tot += (size_t)k;
    std::vector<double> flat(tot);
    check(vg_get_states_ragged(h_, nv.data(), flat.data()), "getStateVariables");
    numVar.assign(nv.begin(), nv.end());
    data.assign(n_, std::vector<double>());
    size_t o = 0;
    for (size_t i = 0; i < n_; i++) {
      data[i].assign(flat.begin() + o, flat.begin() + o + nv[i]);
      o += (size_t)nv[i];
    }
  }
  void setStateVariables(const std::vector<std::vector<double>> &data) {
    if (data.size() != n_) throw Error(VG_ERR_ARG, "setStateVariables: one state list per node");
    std::vector<int32_t> nv(n_);
    std::vector<double> flat;
    for (size_t i = 0; i < n_; i++) {
      nv[i] = (int32_t)data[i].size();
      flat.insert(flat.end(), data[i].begin(), data[i].end());
    }
    check(vg_set_states_ragged(h_, nv.data(), flat.data()), "setStateVariables");
  }

  std::vector<double> lumpedCapacitance() {
    std::vector<double> c(n_);
    check(vg_get_lumped_capacitance(h_, c.data()), "lumpedCapacitance");
    return c;
  }

  // Pseudo-ECG (ComputeECG.h:30-189), electrodes as 3 numbers each.
  void setupECG(const std::vector<double> &electrodes, const double D[3], double radius) {
    const vg_mesh mv = mesh_.view();
    ne_ = (int)(electrodes.size() / 3);
    check(vg_ecg_setup(h_, &mv, electrodes.data(), ne_, D, radius), "setupECG");
  }
  std::vector<double> computeECG() {
    std::vector<double> e(ne_);
    check(vg_ecg(h_, e.data()), "computeECG");
    return e;
  }

  // EPSolver::WriteOutput / writeRestart / readRestart (EPSolver.cc:116-211):
  // Voltage_time_%08.2f.%02.0f and InterVariables_time_* under path.
  void WriteOutput(double time, const std::string &path, int rank = 0) {
    check(vg_write_voltage(h_, path.c_str(), time, rank), "WriteOutput");
  }
  void writeRestart(double time, const std::string &path, int rank = 0) {
    check(vg_write_restart(h_, path.c_str(), time, rank), "writeRestart");
  }
  void readRestart(double time, const std::string &path, int rank = 0) {
    check(vg_read_restart(h_, path.c_str(), time, rank), "readRestart");
  }

  vg_stats stats() {
    vg_stats s{};
    check(vg_stats_get(h_, &s), "stats");
    return s;
  }
  vg_handle *handle() { return h_; }
  size_t numNodes() const { return n_; }

 private:
  void check(int rc, const char *what) {
    if (rc != VG_OK) throw Error(rc, std::string(what) + ": " + vg_last_error(h_));
  }
  Mesh mesh_;
  vg_handle *h_ = nullptr;
  size_t n_ = 0;
  int ne_ = 0;
};

// IonicModel-shaped single cell (IonicModel.h:33-34): one Compute_Ion per
// call, run on the device.  Meant for unit tests and cell studies; a tissue
// runs all its cells in one launch inside GpuCNSolver.
class GpuIonicModel {
 public:
  GpuIonicModel(int cell_model, int precision = VG_PREC_FP64,
                int integrator = VG_INTEG_REFERENCE) {
    const int rc = vg_cells_create(cell_model, precision, integrator, 1, &c_);
    if (rc != VG_OK) throw Error(rc, "vg_cells_create");
  }
  ~GpuIonicModel() {
    if (c_) vg_cells_destroy(c_);
  }
  GpuIonicModel(const GpuIonicModel &) = delete;
  GpuIonicModel &operator=(const GpuIonicModel &) = delete;

  // Returns dV/dt and advances the internal states by dt at fixed V.  Xi and
  // userXi are accepted for signature compatibility; like the reference
  // driver (which never sets a user Xi) the model's own ratio is used.
  double Compute_Ion(double /*Xi*/, bool /*userXi*/, double C_m, double dt, double Volt,
                     double istim) {
    double dv = 0.0;
    const int rc = vg_cells_compute_ion(c_, C_m, dt, &Volt, &istim, &dv);
    if (rc != VG_OK) throw Error(rc, "Compute_Ion");
    return dv;
  }
  std::vector<double> getInternalParameters(int &nData) {
    vg_cells_get_states(c_, &nData, nullptr);
    std::vector<double> s(nData);
    vg_cells_get_states(c_, &nData, s.data());
    return s;
  }
  void setInternalParameters(const std::vector<double> &data) {
    const int rc = vg_cells_set_states(c_, (int)data.size(), data.data());
    if (rc != VG_OK) throw Error(rc, "setInternalParameters");
  }

 private:
  vg_cells *c_ = nullptr;
};

}  // namespace voomgpu
