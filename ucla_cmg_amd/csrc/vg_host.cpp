// vg_host.cpp -- P1-tet + Purkinje-bar assembly, SELL-64 operator build and
// the pseudo-ECG lead-field matrix, on the host, in O(elements).
//
// Discretisation follows the reference: CardiacBody::Compute
// (Body/CardiacBody.cc:78-187), computeLumpedCapacitance (:32-67),
// MixedMesh3D::Compute (Mesh/MixedMesh3D.cc:10-158) with the order-2 tet rule
// (Quadrature/TetQuadrature.cc:16-31) and the 2-point bar rule
// (Quadrature/LineQuadrature.cc:30-35).
#include "vg_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

namespace vg {
namespace {

constexpr double kAlpha = 0.58541020, kBeta = 0.13819660;  // 8-digit rule of the reference
constexpr double kGauss = 0.577350269189626;

// Barycentric coordinates of the four tet quadrature points.
void tet_rule(double N[4][4]) {
  const double pts[4][3] = {{kAlpha, kBeta, kBeta}, {kBeta, kAlpha, kBeta},
                            {kBeta, kBeta, kAlpha}, {kBeta, kBeta, kBeta}};
  for (int q = 0; q < 4; q++) {
    N[q][0] = 1.0 - pts[q][0] - pts[q][1] - pts[q][2];
    N[q][1] = pts[q][0];
    N[q][2] = pts[q][1];
    N[q][3] = pts[q][2];
  }
}

// Jacobian columns x_b - x_0, signed determinant and shape-function
// gradients G[a][k] = dN_a/dx_k of a P1 tet.
double tet_gradients(const double x[4][3], double G[4][3]) {
  double J[3][3];  // J[k][b] = d x_k / d s_b
  for (int k = 0; k < 3; k++)
    for (int b = 0; b < 3; b++) J[k][b] = x[b + 1][k] - x[0][k];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double inv[3][3];  // inv[b][k] = d s_b / d x_k
  inv[0][0] = (J[1][1] * J[2][2] - J[2][1] * J[1][2]) / det;
  inv[0][1] = (J[2][1] * J[0][2] - J[0][1] * J[2][2]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) / det;
  inv[1][0] = (J[2][0] * J[1][2] - J[1][0] * J[2][2]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) / det;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) / det;
  for (int k = 0; k < 3; k++) {
    G[1][k] = inv[0][k];
    G[2][k] = inv[1][k];
    G[3][k] = inv[2][k];
    G[0][k] = -(inv[0][k] + inv[1][k] + inv[2][k]);
  }
  return det;
}

void tensor(const double *fib, const double D[3], double Dm[3][3]) {
  std::memset(Dm, 0, 9 * sizeof(double));
  if (!fib) {
    for (int p = 0; p < 3; p++) Dm[p][p] = D[p];
    return;
  }
  for (int p = 0; p < 3; p++)
    for (int m = 0; m < 3; m++)
      for (int k = 0; k < 3; k++) Dm[m][k] += D[p] * fib[3 * p + m] * fib[3 * p + k];
}

struct PmjTable {
  std::unordered_map<int32_t, int64_t> idx;
  const double *data = nullptr;
  explicit PmjTable(const vg_mesh &m) : data(m.pmj_data) {
    for (int64_t k = 0; k < m.n_pmj; k++) idx.emplace(m.pmj_node[k], k);
  }
  const double *find(int32_t node) const {
    auto it = idx.find(node);
    return it == idx.end() ? nullptr : data + 4 * it->second;
  }
};

// Direction cosines and length of a bar (MixedMesh3D.cc:79-88).
void bar_geometry(const double x0[3], const double x1[3], const double *pmj, double ang[3],
                  double &len) {
  double l2 = 0.0;
  for (int j = 0; j < 3; j++) {
    ang[j] = x1[j] - x0[j];
    l2 += ang[j] * ang[j];
  }
  len = std::sqrt(l2);
  for (int j = 0; j < 3; j++) ang[j] /= len;
  if (pmj) {
    for (int j = 0; j < 3; j++) ang[j] = pmj[j];
    len = pmj[3];
  }
}

}  // namespace

std::string validate_mesh(const vg_mesh &m) {
  if (m.n_nodes <= 0 || m.n_nodes > INT32_MAX) return "n_nodes out of range";
  if (!m.xyz) return "xyz is NULL";
  if (m.n_elems < 0 || (m.n_elems > 0 && (!m.elem_nn || !m.conn))) return "elements missing";
  for (int64_t e = 0; e < m.n_elems; e++) {
    const int nn = m.elem_nn[e];
    if (nn != 4 && nn != 2) return "element " + std::to_string(e) + " has " +
                                   std::to_string(nn) + " nodes (only 4 or 2 supported)";
    for (int a = 0; a < nn; a++) {
      const int32_t v = m.conn[4 * e + a];
      if (v < 0 || v >= m.n_nodes) return "element " + std::to_string(e) + " references node " +
                                          std::to_string(v) + " out of range";
    }
  }
  if (m.n_pmj > 0 && (!m.pmj_node || !m.pmj_data)) return "PMJ arrays missing";
  return "";
}

std::string assemble_csr(const vg_mesh &m, const vg_model &mod, HostCsr &out) {
  std::string err = validate_mesh(m);
  if (!err.empty()) return err;
  const int64_t n = m.n_nodes, ne = m.n_elems;
  // Pattern: candidate columns per row, then sort + unique.
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t e = 0; e < ne; e++) {
    const int nn = m.elem_nn[e];
    for (int a = 0; a < nn; a++) start[m.conn[4 * e + a] + 1] += nn;
  }
  std::partial_sum(start.begin(), start.end(), start.begin());
  std::vector<int32_t> cand(start[n]);
  {
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (int64_t e = 0; e < ne; e++) {
      const int nn = m.elem_nn[e];
      const int32_t *c = m.conn + 4 * e;
      for (int a = 0; a < nn; a++)
        for (int b = 0; b < nn; b++) cand[pos[c[a]]++] = c[b];
    }
  }
  out.n = n;
  out.row_ptr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; i++) {
    auto b = cand.begin() + start[i], e = cand.begin() + start[i + 1];
    std::sort(b, e);
    out.row_ptr[i + 1] = out.row_ptr[i] + (std::unique(b, e) - b);
  }
  out.col.resize(out.row_ptr[n]);
  for (int64_t i = 0; i < n; i++)
    std::copy(cand.begin() + start[i], cand.begin() + start[i] + (out.row_ptr[i + 1] - out.row_ptr[i]),
              out.col.begin() + out.row_ptr[i]);
  cand.clear();
  cand.shrink_to_fit();
  out.K.assign(out.col.size(), 0.0);
  out.M.assign(out.col.size(), 0.0);
  out.CL.assign(n, 0.0);

  double NN[4][4];
  tet_rule(NN);
  double Mref[4][4] = {{0}}, Nsum[4] = {0};  // sum over points of N_a N_b and of N_a
  for (int q = 0; q < 4; q++)
    for (int a = 0; a < 4; a++) {
      Nsum[a] += NN[q][a];
      for (int b = 0; b < 4; b++) Mref[a][b] += NN[q][a] * NN[q][b];
    }
  const PmjTable pmj(m);
  const bool sticky = (mod.quirks & VG_QUIRK_PMJ_STICKY) != 0;
  const double bar_vol = (mod.quirks & VG_QUIRK_BAR_1X_VOLUME) ? 0.5 : 1.0;
  const double Cm = mod.Cm;
  const double *cur_pmj = nullptr;

  auto add = [&](int32_t r, int32_t c, double kv, double mv) {
    const int64_t b = out.row_ptr[r], e = out.row_ptr[r + 1];
    const int64_t k = std::lower_bound(out.col.begin() + b, out.col.begin() + e, c) - out.col.begin();
    out.K[k] += kv;
    out.M[k] += mv;
  };

  for (int64_t e = 0; e < ne; e++) {
    const int nn = m.elem_nn[e];
    const int32_t *c = m.conn + 4 * e;
    if (const double *hit = pmj.find(c[0])) cur_pmj = hit;
    else if (!sticky) cur_pmj = nullptr;
    const double *fib = m.fibers ? m.fibers + 9 * e : nullptr;
    double Dv[3];
    if (nn == 2) {
      const bool lv = m.tissue && (m.tissue[c[0]] == 11 || m.tissue[c[1]] == 11);
      for (int p = 0; p < 3; p++) Dv[p] = m.fibers ? (lv ? mod.DpLeft[p] : mod.Dp[p]) : mod.Dp[0];
    } else {
      for (int p = 0; p < 3; p++) Dv[p] = m.fibers ? mod.D[p] : mod.D[0];
    }
    double Dm[3][3];
    tensor(fib, Dv, Dm);
    if (nn == 4) {
      double x[4][3], G[4][3];
      for (int a = 0; a < 4; a++)
        for (int k = 0; k < 3; k++) x[a][k] = m.xyz[3 * (int64_t)c[a] + k];
      const double det = tet_gradients(x, G);
      const double w = det / 24.0;  // each of the 4 points carries 1/4/6 * detJ
      double DG[4][3];
      for (int a = 0; a < 4; a++)
        for (int k = 0; k < 3; k++) DG[a][k] = Dm[k][0] * G[a][0] + Dm[k][1] * G[a][1] + Dm[k][2] * G[a][2];
      for (int a = 0; a < 4; a++) {
        out.CL[c[a]] += Nsum[a] * Cm * w;
        for (int b = 0; b < 4; b++) {
          const double kab = 4.0 * w * Cm * (G[a][0] * DG[b][0] + G[a][1] * DG[b][1] + G[a][2] * DG[b][2]);
          add(c[a], c[b], kab, Mref[a][b] * w);
        }
      }
    } else {
      double x0[3], x1[3], ang[3], len;
      for (int k = 0; k < 3; k++) {
        x0[k] = m.xyz[3 * (int64_t)c[0] + k];
        x1[k] = m.xyz[3 * (int64_t)c[1] + k];
      }
      bar_geometry(x0, x1, cur_pmj, ang, len);
      // Two Gauss points of weight 1, each scaled by pi r^2 L (MixedMesh3D.cc:107).
      const double wq = bar_vol * M_PI * (m.radius * m.radius) * len;
      const double dad = Dm[0][0] * ang[0] * ang[0] + Dm[1][1] * ang[1] * ang[1] +
                         Dm[2][2] * ang[2] * ang[2] +
                         2.0 * (Dm[0][1] * ang[0] * ang[1] + Dm[0][2] * ang[0] * ang[2] +
                                Dm[1][2] * ang[1] * ang[2]);
      const double kd = 2.0 * wq * Cm * dad / (len * len);  // sum over 2 points of (1/L)^2
      const double Nq[2][2] = {{0.5 * (1 + kGauss), 0.5 * (1 - kGauss)},
                               {0.5 * (1 - kGauss), 0.5 * (1 + kGauss)}};
      for (int a = 0; a < 2; a++) {
        out.CL[c[a]] += (Nq[0][a] + Nq[1][a]) * Cm * wq;
        for (int b = 0; b < 2; b++) {
          const double mab = (Nq[0][a] * Nq[0][b] + Nq[1][a] * Nq[1][b]) * wq;
          add(c[a], c[b], (a == b ? kd : -kd), mab);
        }
      }
    }
  }
  return "";
}

std::string build_sell(const HostCsr &c, double kc, double rc, double Cm, bool lumped, int sigma,
                       const std::vector<int32_t> &base, int64_t n_first, HostSell &o) {
  const int64_t n = c.n;
  const int64_t rows = (int64_t)base.size();
  o.n = n;
  o.rows = rows;
  o.n_slices = (rows + SELL_C - 1) / SELL_C;
  const int64_t n_pad = o.n_slices * SELL_C;
  std::vector<double> d(n, 0.0);
  for (int64_t i = 0; i < n; i++) {
    for (int64_t k = c.row_ptr[i]; k < c.row_ptr[i + 1]; k++)
      if (c.col[k] == i) d[i] = kc * c.K[k] + (lumped ? c.CL[i] : Cm * c.M[k]);
    if (!(d[i] > 0.0)) return "non-positive diagonal of the CN operator at node " + std::to_string(i);
  }
  if (rows < n || n_first < 0 || n_first > rows) return "bad row order";
  {
    std::vector<char> seen(n, 0);
    int64_t real = 0;
    for (int32_t b : base) {
      if (b < 0) continue;
      if (b >= n || seen[b]) return "bad row order";
      seen[b] = 1;
      real++;
    }
    if (real != n) return "bad row order";
    if (rows > n && sigma > 1) return "row-length sorting is not available with instance padding";
  }
  o.new2old = base;
  if (sigma > 1) {
    const int64_t seg[3] = {0, n_first, n};
    for (int sgi = 0; sgi < 2; sgi++)
      for (int64_t w0 = seg[sgi]; w0 < seg[sgi + 1]; w0 += sigma) {
        const int64_t w1 = std::min<int64_t>(seg[sgi + 1], w0 + sigma);
        std::stable_sort(o.new2old.begin() + w0, o.new2old.begin() + w1, [&](int32_t a, int32_t b) {
          return (c.row_ptr[a + 1] - c.row_ptr[a]) > (c.row_ptr[b + 1] - c.row_ptr[b]);
        });
      }
  }
  o.old2new.assign(n, 0);
  for (int64_t i = 0; i < rows; i++)
    if (o.new2old[i] >= 0) o.old2new[o.new2old[i]] = (int32_t)i;
  o.width.assign(o.n_slices, 0);
  o.offset.assign(o.n_slices + 1, 0);
  o.ustart.assign(o.n_slices, -1);
  o.uoff.clear();
  // A slice is stored with shared offsets (ustart >= 0) when the union U of
  // its rows' relative column offsets has at most UNIFORM_MAX members and
  // every row + u of the slice (padding lanes included) is a row the kernels
  // write (< n_slices * 64): then entry j of every lane gathers row + U[j]
  // and no per-lane column is stored.  Structured meshes in grid order have
  // one such set for nearly every slice.
  constexpr int UNIFORM_MAX = 20;
  const int64_t kernel_rows = o.n_slices * SELL_C;
  std::vector<std::vector<int64_t>> urel(o.n_slices);
  for (int64_t s = 0; s < o.n_slices; s++) {
    int w = 0;
    std::vector<int64_t> uni;
    for (int l = 0; l < SELL_C; l++) {
      const int64_t r = s * SELL_C + l;
      if (r < rows && o.new2old[r] >= 0) {
        const int32_t old = o.new2old[r];
        // The diagonal is not stored (see below).
        w = std::max<int>(w, (int)(c.row_ptr[old + 1] - c.row_ptr[old]) - 1);
        for (int64_t k = c.row_ptr[old]; k < c.row_ptr[old + 1]; k++)
          if (c.col[k] != old) uni.push_back((int64_t)o.old2new[c.col[k]] - r);
      }
    }
    std::sort(uni.begin(), uni.end());
    uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
    bool uniform = !uni.empty() && (int)uni.size() <= UNIFORM_MAX &&
                   s * SELL_C + uni.front() >= 0 && s * SELL_C + SELL_C - 1 + uni.back() < kernel_rows;
    if (uniform) {
      w = (int)uni.size();
      urel[s] = uni;
      o.ustart[s] = (int64_t)o.uoff.size();
      for (int64_t u : uni) o.uoff.push_back((int32_t)u);
      while (o.uoff.size() % 4) o.uoff.push_back(0);
    }
    o.width[s] = w;
    o.offset[s + 1] = o.offset[s] + (int64_t)w * SELL_C;
  }
  o.padded = o.offset[o.n_slices];
  o.col.assign(o.padded, 0);
  o.A_hat.assign(o.padded, 0.0);
  o.K_rhs.assign(o.padded, 0.0);
  o.K_exp.assign(o.padded, 0.0);
  o.sinv.assign(n_pad, 0.0);
  o.diag.assign(n_pad, 0.0);
  // Padding rows (no node) carry D = 1 and no entries: they stay zero.
  for (int64_t r = 0; r < rows; r++) {
    const int32_t old = o.new2old[r];
    o.sinv[r] = old >= 0 ? 1.0 / std::sqrt(d[old]) : 1.0;
    o.diag[r] = old >= 0 ? d[old] : 1.0;
  }
  // Off-diagonal entries only.  The diagonal of A_hat = D^-1/2 A D^-1/2 is 1
  // (D = diag A), and K, K_exp have zero row sums (K 1 = 0), so the kernels
  // add x_i for A_hat and form the stiffness products as sum_j K_ij (V_j - V_i):
  // one stored entry and one gather fewer per row.
  std::vector<std::pair<int32_t, int64_t>> row;
  for (int64_t s = 0; s < o.n_slices; s++) {
    for (int l = 0; l < SELL_C; l++) {
      const int64_t r = s * SELL_C + l;
      row.clear();
      if (r < rows && o.new2old[r] >= 0) {
        const int32_t old = o.new2old[r];
        for (int64_t k = c.row_ptr[old]; k < c.row_ptr[old + 1]; k++)
          if (c.col[k] != old) row.emplace_back(o.old2new[c.col[k]], k);
        std::sort(row.begin(), row.end());
      }
      if (o.ustart[s] >= 0) {
        // Shared offsets: entry j gathers r + U[j]; rows without that
        // neighbour hold zeros there.  The row's own entries, sorted by
        // column, keep their order among the slots.
        const std::vector<int64_t> &U = urel[s];
        size_t q = 0;
        for (int j = 0; j < o.width[s]; j++) {
          const int64_t slot = sell_slot(o.offset[s], o.width[s], j, l);
          o.col[slot] = (int32_t)(r + U[j]);
          if (q < row.size() && (int64_t)row[q].first - r == U[j]) {
            const int64_t k = row[q].second;
            const int32_t old_r = o.new2old[r], old_c = c.col[k];
            const double mass = lumped ? (old_c == old_r ? c.CL[old_r] : 0.0) : Cm * c.M[k];
            const double a = kc * c.K[k] + mass;
            o.A_hat[slot] = a / std::sqrt(d[old_r] * d[old_c]);
            o.K_rhs[slot] = -rc * c.K[k] / std::sqrt(d[old_r]);
            o.K_exp[slot] = -c.K[k] / c.CL[old_r];
            q++;
          }
        }
        if (q != row.size()) return "shared-offset slice lost an entry (internal error)";
        continue;
      }
      for (int j = 0; j < o.width[s]; j++) {
        const int64_t slot = sell_slot(o.offset[s], o.width[s], j, l);
        if (j < (int)row.size()) {
          const int32_t cn = row[j].first;
          const int64_t k = row[j].second;
          const int32_t old_r = o.new2old[r], old_c = c.col[k];
          const double mass = lumped ? (old_c == old_r ? c.CL[old_r] : 0.0) : Cm * c.M[k];
          const double a = kc * c.K[k] + mass;
          o.col[slot] = cn;
          o.A_hat[slot] = a / std::sqrt(d[old_r] * d[old_c]);
          o.K_rhs[slot] = -rc * c.K[k] / std::sqrt(d[old_r]);
          o.K_exp[slot] = -c.K[k] / c.CL[old_r];
        } else {
          o.col[slot] = (int32_t)std::min<int64_t>(r, n - 1);  // in-bounds pad, value 0
        }
      }
    }
  }
  // Column encoding per slice.
  o.col16.assign(o.padded, 0);
  o.wide.assign(o.n_slices, -1);
  o.col32.clear();
  for (int64_t s = 0; s < o.n_slices; s++) {
    if (o.ustart[s] >= 0) continue;  // no per-lane columns
    bool narrow = true;
    for (int j = 0; j < o.width[s] && narrow; j++)
      for (int l = 0; l < SELL_C; l++) {
        const int64_t d = (int64_t)o.col[sell_slot(o.offset[s], o.width[s], j, l)] - (s * SELL_C + l);
        if (d < -32767 || d > 32767) {
          narrow = false;
          break;
        }
      }
    const int64_t a = o.offset[s], b = o.offset[s + 1];
    if (narrow) {
      for (int j = 0; j < o.width[s]; j++)
        for (int l = 0; l < SELL_C; l++) {
          const int64_t k = sell_slot(a, o.width[s], j, l);
          o.col16[k] = (int16_t)(o.col[k] - (s * SELL_C + l));
        }
    } else {
      o.wide[s] = (int64_t)o.col32.size();
      o.col32.insert(o.col32.end(), o.col.begin() + a, o.col.begin() + b);
    }
  }
  return "";
}

std::string build_ecg_matrix(const vg_mesh &m, unsigned quirks, const double *el, int ne,
                             const double Decg[3], double radius, std::vector<double> &H) {
  const int64_t n = m.n_nodes;
  H.assign((size_t)ne * n, 0.0);
  double NN[4][4];
  tet_rule(NN);
  const PmjTable pmj(m);
  const bool sticky = (quirks & VG_QUIRK_PMJ_STICKY) != 0;
  const double *cur_pmj = nullptr;
  for (int64_t e = 0; e < m.n_elems; e++) {
    const int nn = m.elem_nn[e];
    const int32_t *c = m.conn + 4 * e;
    if (const double *hit = pmj.find(c[0])) cur_pmj = hit;
    else if (!sticky) cur_pmj = nullptr;
    double Dm[3][3];
    double Dv[3] = {Decg[0], Decg[1], Decg[2]};
    if (!m.fibers) Dv[1] = Dv[2] = Decg[0];
    tensor(m.fibers ? m.fibers + 9 * e : nullptr, Dv, Dm);
    double x[4][3];
    for (int a = 0; a < nn; a++)
      for (int k = 0; k < 3; k++) x[a][k] = m.xyz[3 * (int64_t)c[a] + k];
    double G[4][3], Nq[4][4], wq;
    int nq;
    if (nn == 4) {
      const double det = tet_gradients(x, G);
      wq = det / 24.0;
      nq = 4;
      std::memcpy(Nq, NN, sizeof(NN));
    } else {
      double ang[3], len;
      bar_geometry(x[0], x[1], cur_pmj, ang, len);
      for (int k = 0; k < 3; k++) {
        G[0][k] = -ang[k] / len;
        G[1][k] = ang[k] / len;
      }
      wq = ((quirks & VG_QUIRK_BAR_1X_VOLUME) ? 0.5 : 1.0) * M_PI * radius * radius * len;
      nq = 2;
      Nq[0][0] = 0.5 * (1 + kGauss);
      Nq[0][1] = 0.5 * (1 - kGauss);
      Nq[1][0] = 0.5 * (1 - kGauss);
      Nq[1][1] = 0.5 * (1 + kGauss);
    }
    for (int q = 0; q < nq; q++) {
      double pos[3] = {0, 0, 0};
      for (int a = 0; a < nn; a++)
        for (int k = 0; k < 3; k++) pos[k] += Nq[q][a] * x[a][k];
      for (int ie = 0; ie < ne; ie++) {
        double dv[3], d2 = 0;
        for (int k = 0; k < 3; k++) {
          dv[k] = pos[k] - el[3 * ie + k];
          d2 += dv[k] * dv[k];
        }
        const double dist = std::sqrt(d2);
        if (dist < 1e-6) continue;  // ComputeECG.h:77-79
        const double inv3 = 1.0 / (dist * dist * dist);
        double g[3], Dg[3];
        for (int k = 0; k < 3; k++) g[k] = -dv[k] * inv3;
        for (int k = 0; k < 3; k++) Dg[k] = Dm[k][0] * g[0] + Dm[k][1] * g[1] + Dm[k][2] * g[2];
        for (int a = 0; a < nn; a++)
          H[(size_t)ie * n + c[a]] += wq * (G[a][0] * Dg[0] + G[a][1] * Dg[1] + G[a][2] * Dg[2]);
      }
    }
  }
  return "";
}

int64_t far_rows(const HostCsr &c, const std::vector<int32_t> &order) {
  std::vector<int64_t> pos(c.n, -1);
  for (size_t k = 0; k < order.size(); k++) pos[order[k]] = (int64_t)k;
  int64_t far = 0;
  for (size_t k = 0; k < order.size(); k++) {
    const int32_t i = order[k];
    for (int64_t e = c.row_ptr[i]; e < c.row_ptr[i + 1]; e++) {
      const int64_t q = pos[c.col[e]];
      if (q >= 0 && (q - (int64_t)k > 32767 || (int64_t)k - q > 32767)) {
        far++;
        break;
      }
    }
  }
  return far;
}

void geometric_order(const HostCsr &c, const double *xyz, std::vector<int32_t> &order) {
  if (order.size() < 2) return;
  double lo[3], hi[3];
  for (int a = 0; a < 3; a++) lo[a] = hi[a] = xyz[3 * (int64_t)order[0] + a];
  for (int32_t i : order)
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], xyz[3 * (int64_t)i + a]);
      hi[a] = std::max(hi[a], xyz[3 * (int64_t)i + a]);
    }
  // Grid step: half the median over nodes of the shortest edge at the node
  // (on a structured grid of spacing h every node then has its own cell).
  std::vector<double> shortest;
  shortest.reserve(order.size());
  for (int32_t i : order) {
    double m = HUGE_VAL;
    for (int64_t e = c.row_ptr[i]; e < c.row_ptr[i + 1]; e++) {
      const int32_t j = c.col[e];
      if (j == i) continue;
      double d2 = 0.0;
      for (int a = 0; a < 3; a++) {
        const double d = xyz[3 * (int64_t)i + a] - xyz[3 * (int64_t)j + a];
        d2 += d * d;
      }
      if (d2 > 0.0) m = std::min(m, d2);
    }
    if (m < HUGE_VAL) shortest.push_back(std::sqrt(m));
  }
  if (shortest.empty()) return;
  std::nth_element(shortest.begin(), shortest.begin() + shortest.size() / 2, shortest.end());
  const double step = 0.5 * shortest[shortest.size() / 2];
  if (!(step > 0.0) || !std::isfinite(step)) return;
  // Axes by extent: the largest slowest, the smallest fastest (ties keep
  // x before y before z), so the band is about one layer across the two
  // smaller extents.
  int ax[3] = {0, 1, 2};
  std::stable_sort(ax, ax + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
  struct Key {
    int64_t q[3];
    int32_t id;
  };
  std::vector<Key> keys(order.size());
  for (size_t k = 0; k < order.size(); k++) {
    const int32_t i = order[k];
    for (int a = 0; a < 3; a++)
      keys[k].q[a] = std::llround((xyz[3 * (int64_t)i + ax[a]] - lo[ax[a]]) / step);
    keys[k].id = i;
  }
  std::sort(keys.begin(), keys.end(), [](const Key &x, const Key &y) {
    if (x.q[0] != y.q[0]) return x.q[0] < y.q[0];
    if (x.q[1] != y.q[1]) return x.q[1] < y.q[1];
    if (x.q[2] != y.q[2]) return x.q[2] < y.q[2];
    return x.id < y.id;
  });
  for (size_t k = 0; k < order.size(); k++) order[k] = keys[k].id;
}

}  // namespace vg
