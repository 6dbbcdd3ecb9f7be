// vg_host.hpp -- host-side operator construction for libvoomgpu.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/voomgpu.h"

namespace vg {

// CSR of K (conductivity, Cm included), M (consistent mass) and the lumped
// capacitance, all on one pattern, in the caller's node numbering.
struct HostCsr {
  int64_t n = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  std::vector<double> K, M, CL;
};

// Returns an empty string on success, else the reason.
std::string assemble_csr(const vg_mesh &mesh, const vg_model &model, HostCsr &out);
std::string validate_mesh(const vg_mesh &mesh);

// SELL-64 layout: rows in slices of 64 lanes, one lane per row; slice s holds
// width[s] (its longest row) entries per lane from offset[s]: the first
// 4*(width/4) entries of a lane in groups of four consecutive entries
// ([group][lane][4], so one 16-byte load per lane fetches a group), the 0-3
// left over lane-fastest ([entry][lane]).  Columns are stored as 16-bit
// offsets from the row (col16) in every slice whose columns all lie within
// +-32767 rows of their row, else as int32 in col32 from wide[s] (wide[s] = -1
// for a 16-bit slice), in the same layout.
constexpr int SELL_C = 64;

// Storage position of entry j of lane l in a slice of width w at offset off.
inline int64_t sell_slot(int64_t off, int w, int j, int l) {
  const int w4 = w & ~3;
  return j < w4 ? off + (int64_t)(j >> 2) * (4 * SELL_C) + 4 * l + (j & 3)
                : off + (int64_t)w4 * SELL_C + (int64_t)(j - w4) * SELL_C + l;
}
struct HostSell {
  int64_t n = 0, rows = 0, n_slices = 0, padded = 0;  // rows: n plus padding rows
  std::vector<int64_t> offset;   // [n_slices + 1]
  std::vector<int32_t> width;    // [n_slices]
  std::vector<int32_t> col;      // [padded] absolute columns (host copy)
  std::vector<int16_t> col16;    // [padded]
  std::vector<int32_t> col32;    // wide slices only, same lane-fastest layout
  std::vector<int64_t> wide;     // [n_slices]
  // Slices whose 64 rows all take their entries from one sorted set of
  // relative offsets (a row lacking one holds a zero there): ustart[s] >= 0
  // indexes that set in uoff (padded to a multiple of 4); -1 otherwise.
  std::vector<int64_t> ustart;   // [n_slices]
  std::vector<int32_t> uoff;
  std::vector<double> A_hat;     // D^-1/2 (mass + kc K) D^-1/2
  std::vector<double> K_rhs;     // -rc D^-1/2 K
  std::vector<double> K_exp;     // -C_L^-1 K (unscaled; the explicit scheme)
  std::vector<double> sinv;      // [n_pad] D^-1/2
  std::vector<double> diag;      // [n_pad] D = diag(Cm M + kc K)
  std::vector<int32_t> new2old, old2new;
};

// Builds the scaled half-step operators in SELL-64 form: A = mass + kc K
// with the solve for the increment A x = -rc K V; mass = Cm M (consistent)
// or C_L on the diagonal (lumped).  Crank-Nicolson: kc = dt/4 (plus
// 4 alpha/dt with diffusive smoothing, EP_CrankNicholson.cc:22-36),
// rc = 2 kc, consistent; backward Euler: kc = rc = dt/2, lumped
// (EP_BackwardEuler.cc:14-27).  `base` lists the mesh
// nodes in the solver's row order, -1 marking a padding row (no node, no
// entries); rows [0, n_first) and [n_first, rows) form two segments, and rows
// are reordered by descending length inside windows of `sigma` rows of one
// segment when sigma > 1 (not with padding rows).
std::string build_sell(const HostCsr &c, double kc, double rc, double Cm, bool lumped, int sigma,
                       const std::vector<int32_t> &base, int64_t n_first, HostSell &out);

// Rows of `order` (the mesh nodes of one row segment, in row order) with a
// column of the same segment more than 32767 rows away: each puts its slice
// on 32-bit columns.
int64_t far_rows(const HostCsr &c, const std::vector<int32_t> &order);

// Reorders one row segment for a narrow operator band: nodes sorted
// lexicographically on a grid of half the median shortest edge, the axis of
// largest extent slowest and of smallest extent fastest.  A structured mesh
// comes out in grid order with its thinnest direction fastest; an unordered
// (e.g. TetGen) mesh gets neighbours within about one layer of the two
// smaller extents.  Ties (nodes of one cell) go by node id.
void geometric_order(const HostCsr &c, const double *xyz, std::vector<int32_t> &order);

// Lead-field matrix of the pseudo-ECG: ecg_e = sum_n H[e][n] V_n with
// H[e][n] = sum over quadrature points of w (dN_n/dx)^T D grad(1/r_e).
std::string build_ecg_matrix(const vg_mesh &mesh, unsigned quirks, const double *electrodes,
                             int ne, const double Decg[3], double radius,
                             std::vector<double> &H /* [ne * n_nodes] */);

// Records a failure on a handle (vg_last_error) and returns code.
int handle_fail(vg_handle *h, int code, const std::string &msg);

}  // namespace vg
