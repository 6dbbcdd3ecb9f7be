// voomgpu_ep -- the EPSimulator driver (Applications/CardiacEP/EPSimulator.cc)
// over libvoomgpu: reads the same input deck and mesh files, runs the same
// stimulus schedule with the Crank-Nicolson step on the GPU, and writes the
// same Voltage_time_* frames and InterVariables_time_* restart files.  Written
// against the C++ classes of voomgpu.hpp (readMesh for Mesh::New, GpuCNSolver
// for CardiacBody + EP_CrankNicholson) as EPSimulator is against voomEP's.
//
//   voomgpu_ep deck.input
//
// Deck keywords (case-insensitive, one per line; a line containing #, $ or *
// anywhere is a comment, EPSimulator.cc:142-146): MODEL, OUTPUTDIR, CELLMODEL
// (LUORUDY, LUORUDYWITHGATE, TUSSCHER, MAHAJAN, MAHAJAN_FAIL), MESHTYPE (MIX3D, TET), QUADORDER, DT, IONICDT,
// SIMULATIONTIME, FREQUENCY, TWODRADIUS, TISSUEDIFFUSION, PURKINJEDIFFUSION,
// LVPURKINJEDIFFUSION, SMOOTHING [alpha], INITIALSTATES, READRESTART t,
// WRITERESTART t, STIMULUS NODAL|CIRC|ZONAL|PLANE ... (EPSimulator.cc:147-266).
// Mechanics keywords are read and ignored.  Keywords of this driver only:
// GPUPRECISION FP64|MIXED, IONICINTEGRATOR REFERENCE|RUSHLARSEN, CGX0
// ZERO|PREVIOUS|LAST|EXTRAPOLATE, GRAPH 0|1, QUIRKS <VG_QUIRK_* bits>, CGTOLERANCE,
// REFERENCEDRIVER 0|1 (below), and
// SOLVERTYPE CRANKNICKOLSON|BACKWARDEULER|EXPLICIT (the reference fixes
// CRANKNICKOLSON, EPSimulator.cc:88, and has getSolverType for the others).
//
// Behaviour that differs from the reference on purpose (each would be a
// defect there): a zonal stimulus is stopped without falling through into
// the planar case (EPSimulator.cc:552-556); a planar stimulus is stopped with
// its own centre and normal (:556-558 use the start counter's); LVPURKINJE-
// DIFFUSION is honoured (the reference parses it and passes Dp, :430);
// DIFFUSIONDT sets the explicit scheme's stride (the reference parses it and
// solves with Solve()'s default 0.1); WRITERESTART below one step writes no
// restart (the reference divides by 0); ZONAL axis letters are read in
// either case (the reference's map lookup is case-sensitive).
//
// REFERENCEDRIVER 1 runs the reference's driver instead, where it differs
// observably: LVPURKINJEDIFFUSION is read but the LV Purkinje diffusion is
// PURKINJEDIFFUSION's (EPSimulator.cc:418-420); stopping a zonal stimulus
// falls through into the planar stop, which with zero current zeroes every
// non-Purkinje node whatever its plane (:552-558, CardiacBody.cc:265-281);
// and WRITERESTART below one step stops the run where the reference would
// divide by zero (:489, :585).  The planar stop's use of the start counter's
// plane (:556-558) changes nothing: with zero current both sides of any
// plane are set to zero.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/voomgpu.hpp"

namespace {

struct Stim {
  int type = 0;  // 0 nodal, 1 circ, 2 zonal, 3 plane
  double center[3] = {0, 0, 0}, normal[3] = {0, 0, 0}, radius = 0, limit = 0;
  int axis = 0;
  std::string node_file;
  double t_start = 0, t_stop = 0, i_stim = 50000.0;
};

std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

[[noreturn]] void die(const std::string &m) {
  std::fprintf(stderr, "voomgpu_ep: %s\n", m.c_str());
  std::exit(1);
}

double num(const std::vector<std::string> &t, size_t i) {
  if (i >= t.size()) die("missing value after " + t[0]);
  return std::atof(t[i].c_str());
}

std::vector<int32_t> read_node_set(const std::string &path) {
  std::ifstream in(path);
  if (!in) die("cannot open node set " + path);
  std::vector<int32_t> ids;
  long v;
  while (in >> v) ids.push_back((int32_t)(v - 1));  // 1-based in the file
  return ids;
}

int run(int argc, char **argv);

}  // namespace

// Errors of the library arrive as voomgpu::Error (code and vg_last_error).
int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const voomgpu::Error &e) {
    die(std::string(e.what()) + " (" + std::to_string(e.code) + ")");
  }
}

namespace {

int run(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "Usage : %s inputFile\n", argv[0]);
    return 1;
  }
  std::ifstream deck(argv[1]);
  if (!deck) die(std::string("Error Opening input file: ") + argv[1]);
  // Defaults of EPSimulator.cc:86-118.
  std::string model_name, out_dir = "Output/", cell_model = "MAHAJAN", mesh_type = "MIX3D",
                          initial_states;
  double dt_ionic = 0.1, sim_time = 300.0, dt = 0.1, frequency = 1.0, radius = 0.0275;
  double dt_diffusion = 0.1;
  int scheme = VG_SCHEME_CN;
  bool tol_given = false;
  bool reference_driver = false;
  double D[3] = {0.001, 0.0005, 0.00025}, Dp[3] = {0.0032, 0.0032, 0.0032},
         DpLeft[3] = {0.0032, 0.0032, 0.0032};
  bool read_restart = false, write_restart = false;
  double restart_read_t = 99999.0, restart_write_t = 99999.0, smoothing = 0.0;
  int precision = VG_PREC_FP64, integ = VG_INTEG_REFERENCE, cg_x0 = VG_X0_ZERO, graph = 0;
  unsigned quirks = 0;
  double cg_tol = 1e-7;  // AztecOO Iterate(1000, 1E-7), EP_CrankNicholson.cc:104
  std::vector<Stim> stims;
  std::string line;
  while (std::getline(deck, line)) {
    if (line.find_first_of("#$*") != std::string::npos) continue;
    std::istringstream ls(line);
    std::vector<std::string> t;
    std::string w;
    while (ls >> w) t.push_back(w);
    if (t.empty()) continue;
    const std::string k = upper(t[0]);
    if (k == "MODEL") model_name = t.at(1);
    else if (k == "OUTPUTDIR") out_dir = t.at(1);
    else if (k == "CELLMODEL") cell_model = upper(t.at(1));
    else if (k == "MESHTYPE") mesh_type = upper(t.at(1));
    else if (k == "IONICDT") dt_ionic = num(t, 1);
    else if (k == "DT") dt = num(t, 1);
    else if (k == "READRESTART") read_restart = true, restart_read_t = num(t, 1);
    else if (k == "WRITERESTART") write_restart = true, restart_write_t = num(t, 1);
    else if (k == "SIMULATIONTIME") sim_time = num(t, 1);
    else if (k == "FREQUENCY") frequency = num(t, 1);
    else if (k == "TWODRADIUS") radius = num(t, 1);
    else if (k == "TISSUEDIFFUSION") for (int j = 0; j < 3; j++) D[j] = num(t, j + 1);
    else if (k == "PURKINJEDIFFUSION") for (int j = 0; j < 3; j++) Dp[j] = num(t, j + 1);
    else if (k == "LVPURKINJEDIFFUSION") for (int j = 0; j < 3; j++) DpLeft[j] = num(t, j + 1);
    else if (k == "SMOOTHING") smoothing = t.size() >= 2 ? num(t, 1) : 0.1;
    else if (k == "INITIALSTATES") initial_states = t.at(1);
    else if (k == "GPUPRECISION") precision = upper(t.at(1)) == "MIXED" ? VG_PREC_MIXED : VG_PREC_FP64;
    else if (k == "IONICINTEGRATOR")
      integ = upper(t.at(1)) == "RUSHLARSEN" ? VG_INTEG_RUSH_LARSEN : VG_INTEG_REFERENCE;
    else if (k == "CGX0") {
      const std::string v = upper(t.at(1));
      cg_x0 = v == "LAST"          ? VG_X0_LAST
              : v == "PREVIOUS"    ? VG_X0_PREVIOUS
              : v == "EXTRAPOLATE" ? VG_X0_EXTRAPOLATE
                                   : VG_X0_ZERO;
    } else if (k == "GRAPH") graph = (int)num(t, 1);
    else if (k == "QUIRKS") quirks = (unsigned)num(t, 1);
    else if (k == "CGTOLERANCE") cg_tol = num(t, 1), tol_given = true;
    else if (k == "REFERENCEDRIVER") reference_driver = num(t, 1) != 0.0;
    else if (k == "DIFFUSIONDT") dt_diffusion = num(t, 1);
    else if (k == "SOLVERTYPE") {
      const std::string v = upper(t.at(1));
      if (v == "EXPLICIT") scheme = VG_SCHEME_EXPLICIT;
      else if (v == "BACKWARDEULER") scheme = VG_SCHEME_BE;
      else if (v == "CRANKNICKOLSON") scheme = VG_SCHEME_CN;
      else die("Unknown Solver type " + t[1]);
    }
    else if (k == "STIMULUS") {
      Stim s;
      static const std::map<std::string, int> types = {
          {"NODAL", 0}, {"CIRC", 1}, {"ZONAL", 2}, {"PLANE", 3}};
      static const std::map<std::string, int> axes = {{"X", 0}, {"Y", 1}, {"Z", 2}};
      auto it = types.find(upper(t.at(1)));
      if (it == types.end()) die("Unknown Stimulus " + t[1]);
      s.type = it->second;
      size_t p = 2;
      switch (s.type) {
        case 0: s.node_file = t.at(p++); break;
        case 1:
          for (int j = 0; j < 3; j++) s.center[j] = num(t, p++);
          s.radius = num(t, p++);
          break;
        case 2: {
          auto a = axes.find(upper(t.at(p++)));
          s.axis = a == axes.end() ? 0 : a->second;
          s.limit = num(t, p++);
          break;
        }
        default:
          for (int j = 0; j < 3; j++) s.center[j] = num(t, p++);
          for (int j = 0; j < 3; j++) s.normal[j] = num(t, p++);
      }
      s.t_start = num(t, p++);
      s.t_stop = num(t, p++);
      if (t.size() > p) s.i_stim = num(t, p);
      stims.push_back(s);
    }
    // QUADORDER, DIFFUSIONDT, mechanics and CONSISTENCYCHECK keywords:
    // accepted and ignored (P1 operators are exact; no mechanics here).
  }
  if (model_name.empty()) die("no MODEL in the deck");
  if (mesh_type != "MIX3D" && mesh_type != "TET") die("MESHTYPE " + mesh_type + " is not supported");
  int cm;
  if (cell_model == "LUORUDY") cm = VG_MODEL_LR;
  else if (cell_model == "LUORUDYWITHGATE") cm = VG_MODEL_LR_GATE;
  else if (cell_model == "TUSSCHER") cm = VG_MODEL_TT;
  else if (cell_model == "MAHAJAN") cm = VG_MODEL_MJ;
  else if (cell_model == "MAHAJAN_FAIL") cm = VG_MODEL_MJ_FAIL;
  else die("CELLMODEL " + cell_model + " is not supported");

  const voomgpu::Mesh mesh = voomgpu::readMesh(model_name, radius);
  vg_model model = voomgpu::default_model(cm);
  model.precision = precision;
  model.integrator = integ;
  for (int j = 0; j < 3; j++) {
    model.D[j] = D[j];
    model.Dp[j] = Dp[j];
    model.DpLeft[j] = reference_driver ? Dp[j] : DpLeft[j];
  }
  model.Cm = 1.0;  // EPSimulator.cc:416
  model.dt = dt;
  model.dt_ionic = dt_ionic;
  model.cg_tol = cg_tol;
  model.cg_maxit = 1000;
  model.quirks = quirks;
  model.cg_x0 = cg_x0;
  model.graph = graph;
  model.smoothing = smoothing;
  model.scheme = scheme;
  model.dt_diffusion = dt_diffusion;
  // EP_BackwardEuler iterates to 1e-10 (EP_BackwardEuler.cc:65), CN to 1e-7.
  if (scheme == VG_SCHEME_BE && !tol_given) model.cg_tol = 1e-10;
  voomgpu::GpuCNSolver body(mesh, model);
  const int64_t n = (int64_t)body.numNodes();
  std::printf("Modelname: %s\nNodes: %lld  Elements: %lld  CellModel: %s  dt: %g ms\n",
              model_name.c_str(), (long long)n, (long long)mesh.elem_nn.size(),
              cell_model.c_str(), dt);

  // INITIALSTATES (EPSimulator.cc:379-404): a "CELL" and a "PURKINJE" block
  // of one number per line.  Purkinje nodes take the PURKINJE block;
  // Mahajan myocardium takes the CELL block (26 values, or 9 x 26 by tissue
  // type), CardiacBody.cc:1104-1214.
  if (!initial_states.empty()) {
    std::ifstream in(initial_states);
    if (!in) die("cannot open " + initial_states);
    std::vector<double> pk, cell;
    std::string flag;
    while (std::getline(in, line)) {
      const std::string u = upper(line);
      if (u.find_first_not_of(" \t\r") == std::string::npos) {
        flag.clear();
        continue;
      }
      if (u.find("PURKINJE") != std::string::npos) {
        flag = "PURKINJE";
        continue;
      }
      if (u.find("CELL") != std::string::npos) {
        flag = "CELL";
        continue;
      }
      if (flag == "PURKINJE") pk.push_back(std::atof(line.c_str()));
      else if (flag == "CELL") cell.push_back(std::atof(line.c_str()));
    }
    std::vector<int> nv;
    std::vector<std::vector<double>> data;
    body.getStateVariables(nv, data);
    std::vector<double> V = body.getVoltage();
    for (int64_t i = 0; i < n; i++) {
      const int t = mesh.tissue.empty() ? 0 : mesh.tissue[i];
      const double *src = nullptr;
      if ((t == 11 || t == 12) && !pk.empty()) {
        if ((int)pk.size() != nv[i]) die("PURKINJE block has the wrong number of states");
        src = pk.data();
      } else if ((cm == VG_MODEL_MJ || cm == VG_MODEL_MJ_FAIL) && t != 11 && t != 12 &&
                 !cell.empty()) {
        if (cell.size() == (size_t)nv[i]) src = cell.data();
        else if (cell.size() == (size_t)9 * nv[i] && t >= 0 && t < 9) src = cell.data() + (size_t)t * nv[i];
        else die("CELL block has the wrong number of states");
      }
      if (src) {
        data[i].assign(src, src + nv[i]);
        V[i] = src[0];  // V is the first state of both models
      }
    }
    body.setStateVariables(data);
    body.InitializeVoltage(V);
  }

  // Time loop (EPSimulator.cc:478-572).
  const long steps_per_frame = std::max(1L, (long)(frequency / dt));
  long count = 0;
  double simtime = 0.0;
  size_t start_ctr = 0, stop_ctr = 0;
  long pulse_start = stims.empty() ? -1 : (long)(stims[0].t_start / dt);
  long pulse_stop = stims.empty() ? -1 : (long)(stims[0].t_stop / dt);
  const long restart_write = (long)(restart_write_t / dt);
  if (reference_driver && write_restart && restart_write == 0)
    die("WRITERESTART is below one step: the reference divides by zero here");
  if (read_restart) {
    std::printf("@@ Reading Restart information\n");
    simtime = restart_read_t;
    body.readRestart(restart_read_t, out_dir);
    count = (long)(simtime / dt);
  } else {
    body.WriteOutput(0.0, out_dir);
  }
  auto apply = [&](const Stim &s, double amp) {
    switch (s.type) {
      case 0: {
        const std::vector<int32_t> ids = read_node_set(s.node_file);
        if (amp != 0.0) std::printf("** Will Stimulate %zu Node(s)\n", ids.size());
        body.setStimulusCurrent(ids, amp);
        break;
      }
      case 1: body.setStimulusCurrent(s.center, s.radius, amp); break;
      case 2:
        body.setStimulusCurrent(s.axis, s.limit, amp);
        if (reference_driver && amp == 0.0) {
          // The reference's zonal stop falls through into the planar stop with
          // zero current: every non-Purkinje node gets zero (a sphere of
          // negative radius contains no node, so it zeroes them all).
          const double o[3] = {0.0, 0.0, 0.0};
          body.setStimulusCurrent(o, -1.0, 0.0);
        }
        break;
      default: body.setStimulusCurrent(s.center, s.normal, amp);
    }
  };
  while (simtime < sim_time) {
    if (count == pulse_start) {
      apply(stims[start_ctr], stims[start_ctr].i_stim);
      std::printf("@@ Starting Stimulus %zu\n", start_ctr + 1);
      if (++start_ctr == stims.size()) start_ctr--;
      pulse_start = (long)(stims[start_ctr].t_start / dt);
    }
    if (count == pulse_stop) {
      apply(stims[stop_ctr], 0.0);
      std::printf("@@ Stopping Stimulus %zu\n", stop_ctr + 1);
      if (++stop_ctr == stims.size()) stop_ctr--;
      pulse_stop = (long)(stims[stop_ctr].t_stop / dt);
    }
    body.Solve(1);
    simtime += dt;
    count++;
    if (count % steps_per_frame == 0) {
      body.WriteOutput(simtime, out_dir);
      std::printf("** Writing Output at Time: %g (ms)\n", simtime);
      std::fflush(stdout);
    }
    if (write_restart && restart_write > 0 && count % restart_write == 0) {
      body.writeRestart(simtime, out_dir);
      std::printf("@@ Writing Restart at Time: %g (ms)\n", simtime);
    }
  }
  const vg_stats st = body.stats();
  std::printf("done: %lld steps, %.2f CG iterations per solve, %lld unconverged solves\n",
              (long long)st.steps, st.cg_solves ? (double)st.cg_iterations / st.cg_solves : 0.0,
              (long long)st.cg_not_converged);
  return 0;
}

}  // namespace
