// voomgpu_ecg -- the ComputeECG post-processor (Applications/CardiacEP/
// ComputeECG.cc) over libvoomgpu's lead field: reads the same parameter file
// and Voltage_time_* frames and writes the same ECG_Data.00 table.
//
//   voomgpu_ecg ECG.input
//
// Parameter file (whitespace-separated, ComputeECG.cc:45-53): modelName,
// mesh directory, voltage directory, output directory, quadrature order,
// mesh type, nPoints, dim, nPoints x dim electrode coordinates, number of
// partitions, start step, end step, frequency, ASCII|BINARY.  Frames are
// read for step = start, start + frequency, ... < end (a float counter, as
// in the reference) until one is missing; with P partitions node i's value
// comes from the next entry of partition npart[i]'s file.  The lead field
// uses D = {1e-3, 5e-4, 2.5e-4} and a bar radius of 0.0275 cm like the
// reference (ComputeECG.cc:88-90, 79).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/voomgpu.h"

namespace {
[[noreturn]] void die(const std::string &m) {
  std::fprintf(stderr, "voomgpu_ecg: %s\n", m.c_str());
  std::exit(1);
}
}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "Usage: %s parameters_file\n", argv[0]);
    return 0;
  }
  std::ifstream inp(argv[1]);
  if (!inp) die(std::string("ERROR opening file: ") + argv[1]);
  std::string model, meshdir, vdir, outdir, mesh_type, nparts_s, fmt;
  int quad = 2, np = 0, dim = 0;
  inp >> model >> meshdir >> vdir >> outdir >> quad >> mesh_type >> np >> dim;
  if (!inp || np < 1 || dim != 3) die("bad parameter file header");
  std::vector<double> el((size_t)np * 3);
  for (auto &v : el) inp >> v;
  float start = 0, end = 0, freq = 1;
  inp >> nparts_s >> start >> end >> freq >> fmt;
  if (!inp) die("bad parameter file tail");
  if (mesh_type != "MIX3D" && mesh_type != "TET") die("mesh type " + mesh_type + " is not supported");
  const int nparts = std::atoi(nparts_s.c_str());
  if (nparts < 1) die("bad number of partitions");

  char err[1024] = {0};
  vg_meshfile *mf = nullptr;
  const std::string base = meshdir + "/" + model;
  if (vg_mesh_read(base.c_str(), 0.0275, &mf, err, sizeof err) != VG_OK) die(err);
  const vg_mesh *mesh = vg_meshfile_mesh(mf);
  const int64_t n = mesh->n_nodes;
  const double D[3] = {0.001, 0.0005, 0.00025};
  std::vector<double> H((size_t)np * n);
  if (vg_ecg_leadfield(mesh, 0u, el.data(), np, D, 0.0275, H.data(), err, sizeof err) != VG_OK)
    die(err);
  std::vector<int> part((size_t)n, 0);
  if (nparts > 1) {
    std::ifstream pf(base + ".metis.npart." + nparts_s);
    if (!pf) die("cannot open " + base + ".metis.npart." + nparts_s);
    for (auto &p : part)
      if (!(pf >> p) || p < 0 || p >= nparts) die("bad partition file");
  }
  char oname[4096];
  std::snprintf(oname, sizeof oname, "%s/ECG_Data.%02.0f", outdir.c_str(), 0.0);
  FILE *out = std::fopen(oname, "w");
  if (!out) die(std::string("cannot write ") + oname);
  std::vector<double> V((size_t)n);
  const bool ascii = fmt == "ASCII";
  for (float step = start; step < end; step = step + freq) {
    std::vector<std::ifstream> files((size_t)nparts);
    bool all = true;
    for (int p = 0; p < nparts; p++) {
      char fname[4096];
      std::snprintf(fname, sizeof fname, "%s/Voltage_time_%08.2f.%02.0f", vdir.c_str(), (double)step,
                    (double)p);
      files[p].open(fname, ascii ? std::ios::in : std::ios::in | std::ios::binary);
      if (!files[p]) {
        std::fprintf(stderr, "Error opening %s\n", fname);
        all = false;
      }
    }
    if (!all) break;  // the reference stops at the first missing frame
    for (int64_t i = 0; i < n; i++) {
      std::ifstream &f = files[part[i]];
      if (ascii) f >> V[i];
      else f.read(reinterpret_cast<char *>(&V[i]), sizeof(double));
      if (!f) die("short voltage frame at step " + std::to_string(step));
    }
    std::fprintf(out, "%g\t", (double)step);
    for (int e = 0; e < np; e++) {
      const double *row = H.data() + (size_t)e * n;
      double s = 0.0;
      for (int64_t i = 0; i < n; i++) s += row[i] * V[i];
      std::fprintf(out, "%g\t", s);
    }
    std::fprintf(out, "\n");
  }
  std::fclose(out);
  vg_meshfile_free(mf);
  return 0;
}
