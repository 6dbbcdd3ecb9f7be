// vg_io.cpp -- the reference's file formats on the host side of the library:
// mesh input (<model>.node/.ele/.fibers/.purk/.cell/.angle, Mesh::New in
// Mesh/Mesh.cc:142-416), voltage frames and restart files
// (EPSolver::WriteOutput / writeRestart / readRestart, Solvers/EPSolver.cc:116-211).
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/voomgpu.h"
#include "vg_host.hpp"

namespace {

void set_err(char *err, int len, const std::string &m) {
  if (err && len > 0) {
    std::snprintf(err, (size_t)len, "%s", m.c_str());
  }
}

bool exists(const std::string &p) {
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

// Whitespace-separated tokens of a text file, read as the reference's
// `ifstream >>` does.
struct Tokens {
  FILE *f = nullptr;
  explicit Tokens(const std::string &p) { f = std::fopen(p.c_str(), "r"); }
  ~Tokens() {
    if (f) std::fclose(f);
  }
  bool ok() const { return f != nullptr; }
  bool get(int64_t &v) { return f && std::fscanf(f, "%" SCNd64, &v) == 1; }
  bool get(double &v) { return f && std::fscanf(f, "%lf", &v) == 1; }
};

}  // namespace

struct vg_meshfile {
  std::vector<double> xyz, fibers, pmj_data;
  std::vector<int8_t> elem_nn, tissue;
  std::vector<int32_t> conn, pmj_node;
  vg_mesh mesh{};
};

extern "C" {

int vg_mesh_read(const char *model, double radius, vg_meshfile **out, char *err, int err_len) {
  if (!model || !out) return VG_ERR_ARG;
  *out = nullptr;
  const std::string base(model);
  std::unique_ptr<vg_meshfile> m(new vg_meshfile());
  // .node: "n dim x x", then "id x y z" per node; the i-th line is node i.
  int64_t n = 0, dim = 0, t0 = 0, t1 = 0;
  {
    Tokens in(base + ".node");
    if (!in.ok()) return set_err(err, err_len, "cannot open " + base + ".node"), VG_ERR_ARG;
    if (!in.get(n) || !in.get(dim) || !in.get(t0) || !in.get(t1) || n < 1 || dim != 3)
      return set_err(err, err_len, base + ".node: bad header (need 'n 3 x x')"), VG_ERR_ARG;
    m->xyz.resize((size_t)n * 3);
    for (int64_t i = 0; i < n; i++) {
      int64_t id;
      if (!in.get(id) || !in.get(m->xyz[3 * i]) || !in.get(m->xyz[3 * i + 1]) ||
          !in.get(m->xyz[3 * i + 2]))
        return set_err(err, err_len, base + ".node: short at node " + std::to_string(i + 1)),
               VG_ERR_ARG;
    }
  }
  // .ele: "n x x", then "id k n1..nk" (1-based) per element.
  int64_t ne = 0;
  {
    Tokens in(base + ".ele");
    if (!in.ok()) return set_err(err, err_len, "cannot open " + base + ".ele"), VG_ERR_ARG;
    if (!in.get(ne) || !in.get(t0) || !in.get(t1) || ne < 1)
      return set_err(err, err_len, base + ".ele: bad header"), VG_ERR_ARG;
    m->elem_nn.resize((size_t)ne);
    m->conn.assign((size_t)ne * 4, 0);
    for (int64_t e = 0; e < ne; e++) {
      int64_t id, k;
      if (!in.get(id) || !in.get(k))
        return set_err(err, err_len, base + ".ele: short at element " + std::to_string(e + 1)),
               VG_ERR_ARG;
      if (k != 4 && k != 2)
        return set_err(err, err_len,
                       base + ".ele: element " + std::to_string(e + 1) + " has " +
                           std::to_string(k) + " nodes; tetrahedra (4) and bars (2) are supported"),
               VG_ERR_UNSUPPORTED;
      m->elem_nn[e] = (int8_t)k;
      for (int j = 0; j < k; j++) {
        int64_t v;
        if (!in.get(v) || v < 1 || v > n)
          return set_err(err, err_len, base + ".ele: bad node id in element " + std::to_string(e + 1)),
                 VG_ERR_ARG;
        m->conn[4 * e + j] = (int32_t)(v - 1);
      }
    }
  }
  // .fibers: 9 numbers per element (Mesh.cc:337-342), optional.
  if (exists(base + ".fibers")) {
    Tokens in(base + ".fibers");
    m->fibers.resize((size_t)ne * 9);
    for (auto &v : m->fibers)
      if (!in.get(v))
        return set_err(err, err_len, base + ".fibers: fewer than 9 numbers per element"), VG_ERR_ARG;
  }
  // Tissue types: .purk ("n" then 1-based ids) sets 11 (Mesh.cc:369-381);
  // .cell ("id type" pairs) then overrides (Mesh.cc:383-395).
  m->tissue.assign((size_t)n, 0);
  bool any_tissue = false;
  if (exists(base + ".purk")) {
    Tokens in(base + ".purk");
    int64_t np = 0;
    if (!in.get(np) || np < 0) return set_err(err, err_len, base + ".purk: bad count"), VG_ERR_ARG;
    for (int64_t i = 0; i < np; i++) {
      int64_t id;
      if (!in.get(id) || id < 1 || id > n)
        return set_err(err, err_len, base + ".purk: bad node id"), VG_ERR_ARG;
      m->tissue[id - 1] = 11;
    }
    any_tissue = np > 0;
  }
  if (exists(base + ".cell")) {
    // The reference reads pairs until end of file; a final partial pair
    // (its extraction fails) is ignored here.
    Tokens in(base + ".cell");
    int64_t id, type;
    while (in.get(id) && in.get(type)) {
      if (id < 1 || id > n || type < 0 || type > 127)
        return set_err(err, err_len, base + ".cell: bad pair " + std::to_string(id)), VG_ERR_ARG;
      m->tissue[id - 1] = (int8_t)type;
      any_tissue = true;
    }
  }
  // .angle: "id c0 c1 c2 L" per PMJ node (Mesh.cc:397-409).
  if (exists(base + ".angle")) {
    Tokens in(base + ".angle");
    int64_t id;
    double d[4];
    while (in.get(id)) {
      if (!in.get(d[0]) || !in.get(d[1]) || !in.get(d[2]) || !in.get(d[3])) break;
      if (id < 1 || id > n) return set_err(err, err_len, base + ".angle: bad node id"), VG_ERR_ARG;
      // A repeated id keeps the last line, like the reference's map.
      bool found = false;
      for (size_t k = 0; k < m->pmj_node.size(); k++)
        if (m->pmj_node[k] == (int32_t)(id - 1)) {
          std::memcpy(&m->pmj_data[4 * k], d, sizeof d);
          found = true;
        }
      if (!found) {
        m->pmj_node.push_back((int32_t)(id - 1));
        m->pmj_data.insert(m->pmj_data.end(), d, d + 4);
      }
    }
  }
  vg_mesh &v = m->mesh;
  v.n_nodes = n;
  v.xyz = m->xyz.data();
  v.n_elems = ne;
  v.elem_nn = m->elem_nn.data();
  v.conn = m->conn.data();
  v.fibers = m->fibers.empty() ? nullptr : m->fibers.data();
  v.tissue = any_tissue ? m->tissue.data() : nullptr;
  v.n_pmj = (int64_t)m->pmj_node.size();
  v.pmj_node = m->pmj_node.empty() ? nullptr : m->pmj_node.data();
  v.pmj_data = m->pmj_data.empty() ? nullptr : m->pmj_data.data();
  v.radius = radius;
  *out = m.release();
  return VG_OK;
}

const vg_mesh *vg_meshfile_mesh(const vg_meshfile *m) { return m ? &m->mesh : nullptr; }

void vg_meshfile_free(vg_meshfile *m) { delete m; }

int vg_mesh_write(const char *model, const vg_mesh *mesh, char *err, int err_len) {
  if (!model || !mesh || mesh->n_nodes < 1 || mesh->n_elems < 1 || !mesh->xyz || !mesh->conn ||
      !mesh->elem_nn)
    return VG_ERR_ARG;
  const std::string base(model);
  auto open = [&](const std::string &suffix) {
    FILE *f = std::fopen((base + suffix).c_str(), "w");
    if (!f) set_err(err, err_len, "cannot write " + base + suffix + ": " + std::strerror(errno));
    return f;
  };
  const int64_t n = mesh->n_nodes, ne = mesh->n_elems;
  FILE *f = open(".node");
  if (!f) return VG_ERR_ARG;
  std::fprintf(f, "%" PRId64 " 3 0 0\n", n);
  for (int64_t i = 0; i < n; i++)
    std::fprintf(f, "%" PRId64 " %.17g %.17g %.17g\n", i + 1, mesh->xyz[3 * i], mesh->xyz[3 * i + 1],
                 mesh->xyz[3 * i + 2]);
  std::fclose(f);
  if (!(f = open(".ele"))) return VG_ERR_ARG;
  std::fprintf(f, "%" PRId64 " 0 0\n", ne);
  for (int64_t e = 0; e < ne; e++) {
    const int k = mesh->elem_nn[e];
    std::fprintf(f, "%" PRId64 " %d", e + 1, k);
    for (int j = 0; j < k; j++) std::fprintf(f, " %d", mesh->conn[4 * e + j] + 1);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  if (mesh->fibers) {
    if (!(f = open(".fibers"))) return VG_ERR_ARG;
    for (int64_t e = 0; e < ne; e++) {
      for (int j = 0; j < 9; j++) std::fprintf(f, j ? " %.17g" : "%.17g", mesh->fibers[9 * e + j]);
      std::fprintf(f, "\n");
    }
    std::fclose(f);
  }
  if (mesh->tissue) {
    int64_t np = 0;
    bool other = false;
    for (int64_t i = 0; i < n; i++) {
      np += mesh->tissue[i] == 11 ? 1 : 0;
      other = other || (mesh->tissue[i] != 0 && mesh->tissue[i] != 11);
    }
    if (np > 0) {
      if (!(f = open(".purk"))) return VG_ERR_ARG;
      std::fprintf(f, "%" PRId64 "\n", np);
      for (int64_t i = 0; i < n; i++)
        if (mesh->tissue[i] == 11) std::fprintf(f, "%" PRId64 "\n", i + 1);
      std::fclose(f);
    }
    if (other) {
      if (!(f = open(".cell"))) return VG_ERR_ARG;
      for (int64_t i = 0; i < n; i++)
        if (mesh->tissue[i] != 0 && mesh->tissue[i] != 11)
          std::fprintf(f, "%" PRId64 " %d\n", i + 1, (int)mesh->tissue[i]);
      std::fclose(f);
    }
  }
  if (mesh->n_pmj > 0) {
    if (!(f = open(".angle"))) return VG_ERR_ARG;
    for (int64_t k = 0; k < mesh->n_pmj; k++)
      std::fprintf(f, "%d %.17g %.17g %.17g %.17g\n", mesh->pmj_node[k] + 1, mesh->pmj_data[4 * k],
                   mesh->pmj_data[4 * k + 1], mesh->pmj_data[4 * k + 2], mesh->pmj_data[4 * k + 3]);
    std::fclose(f);
  }
  // One partition, so that the reference's reader (which requires them) runs.
  if (!(f = open(".metis.npart.1"))) return VG_ERR_ARG;
  for (int64_t i = 0; i < n; i++) std::fputs("0\n", f);
  std::fclose(f);
  if (!(f = open(".metis.epart.1"))) return VG_ERR_ARG;
  for (int64_t e = 0; e < ne; e++) std::fputs("0\n", f);
  std::fclose(f);
  return VG_OK;
}

int vg_ecg_leadfield(const vg_mesh *mesh, unsigned quirks, const double *electrodes, int ne,
                     const double D[3], double radius, double *H, char *err, int err_len) {
  if (!mesh || !electrodes || ne < 1 || !D || !H) return VG_ERR_ARG;
  std::string e = vg::validate_mesh(*mesh);
  if (!e.empty()) return set_err(err, err_len, e), VG_ERR_ARG;
  std::vector<double> out;
  e = vg::build_ecg_matrix(*mesh, quirks, electrodes, ne, D, radius, out);
  if (!e.empty()) return set_err(err, err_len, e), VG_ERR_ARG;
  std::memcpy(H, out.data(), out.size() * sizeof(double));
  return VG_OK;
}

int vg_frame_name(char *buf, int len, const char *dir, const char *prefix, double time, int rank) {
  if (!buf || len < 1 || !dir || !prefix) return VG_ERR_ARG;
  const int w = std::snprintf(buf, (size_t)len, "%s/%s_time_%08.2f.%02.0f", dir, prefix, time,
                              (double)rank);
  return (w < 0 || w >= len) ? VG_ERR_ARG : VG_OK;
}

int vg_write_voltage(vg_handle *h, const char *dir, double time, int rank) {
  if (!h || !dir) return VG_ERR_ARG;
  vg_stats st;
  int rc = vg_stats_get(h, &st);
  if (rc) return rc;
  std::vector<double> V((size_t)st.n_nodes);
  if ((rc = vg_get_voltage(h, V.data()))) return rc;
  char name[4096];
  if ((rc = vg_frame_name(name, sizeof name, dir, "Voltage", time, rank))) return rc;
  FILE *f = std::fopen(name, "wb");
  if (!f) return vg::handle_fail(h, VG_ERR_ARG, std::string("cannot write ") + name);
  const size_t w = std::fwrite(V.data(), 8, V.size(), f);
  std::fclose(f);
  return w == V.size() ? VG_OK : vg::handle_fail(h, VG_ERR_ARG, std::string("short write ") + name);
}

int vg_write_restart(vg_handle *h, const char *dir, double time, int rank) {
  if (!h || !dir) return VG_ERR_ARG;
  vg_stats st;
  int rc = vg_stats_get(h, &st);
  if (rc) return rc;
  const int64_t n = st.n_nodes;
  std::vector<int32_t> nv((size_t)n);
  if ((rc = vg_get_states_ragged(h, nv.data(), nullptr))) return rc;
  int64_t tot = 0;
  for (int32_t k : nv) tot += k;
  std::vector<double> data((size_t)tot);
  if ((rc = vg_get_states_ragged(h, nv.data(), data.data()))) return rc;
  char name[4096];
  if ((rc = vg_frame_name(name, sizeof name, dir, "InterVariables", time, rank))) return rc;
  FILE *f = std::fopen(name, "wb");
  if (!f) return vg::handle_fail(h, VG_ERR_ARG, std::string("cannot write ") + name);
  bool ok = true;
  const double *p = data.data();
  for (int64_t i = 0; i < n && ok; i++) {
    ok = std::fwrite(&nv[i], 4, 1, f) == 1 && std::fwrite(p, 8, (size_t)nv[i], f) == (size_t)nv[i];
    p += nv[i];
  }
  std::fclose(f);
  return ok ? VG_OK : vg::handle_fail(h, VG_ERR_ARG, std::string("short write ") + name);
}

int vg_read_restart(vg_handle *h, const char *dir, double time, int rank) {
  if (!h || !dir) return VG_ERR_ARG;
  vg_stats st;
  int rc = vg_stats_get(h, &st);
  if (rc) return rc;
  const int64_t n = st.n_nodes;
  char vname[4096], sname[4096];
  if ((rc = vg_frame_name(vname, sizeof vname, dir, "Voltage", time, rank))) return rc;
  if ((rc = vg_frame_name(sname, sizeof sname, dir, "InterVariables", time, rank))) return rc;
  std::vector<double> V((size_t)n);
  FILE *f = std::fopen(vname, "rb");
  if (!f) return vg::handle_fail(h, VG_ERR_ARG, std::string("cannot open ") + vname);
  const size_t got = std::fread(V.data(), 8, V.size(), f);
  std::fclose(f);
  if (got != V.size()) return vg::handle_fail(h, VG_ERR_ARG, std::string("short file ") + vname);
  std::vector<int32_t> expect((size_t)n), nv((size_t)n);
  if ((rc = vg_get_states_ragged(h, expect.data(), nullptr))) return rc;
  std::vector<double> data;
  if (!(f = std::fopen(sname, "rb")))
    return vg::handle_fail(h, VG_ERR_ARG, std::string("cannot open ") + sname);
  bool ok = true;
  for (int64_t i = 0; i < n && ok; i++) {
    ok = std::fread(&nv[i], 4, 1, f) == 1 && nv[i] == expect[i];
    if (ok) {
      const size_t o = data.size();
      data.resize(o + (size_t)nv[i]);
      ok = std::fread(data.data() + o, 8, (size_t)nv[i], f) == (size_t)nv[i];
    }
  }
  std::fclose(f);
  if (!ok)
    return vg::handle_fail(h, VG_ERR_ARG,
                           std::string(sname) + ": short, or state counts differ from this body's");
  if ((rc = vg_set_states_ragged(h, nv.data(), data.data()))) return rc;
  return vg_set_voltage(h, V.data());
}

}  // extern "C"
