// vg_api.hip -- the C ABI of libvoomgpu: handle life cycle, the Strang/CN
// time step (EP_CrankNicholson::Solve, Solvers/EP_CrankNicholson.cc:78-156),
// stimulus, state and ECG access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/voomgpu.h"
#include "vg_host.hpp"
#include "vg_kernels.hpp"

using namespace vg;

namespace {

// Rest state of the Luo-Rudy model used as the default initial state: the
// reference's initial gates (LR.cc:19-25) run unstimulated for 10 s from
// -84 mV with dt = 0.01 ms.  Order m, h, j, d, f, X, Ca_i; then V.
constexpr double LR_REST[7] = {0.0016534680667748946, 0.9834762339347045, 0.9896274219694648,
                               0.0029656171926269563, 0.999981405275416, 0.005617388831370626,
                               0.00016111399466747764};
constexpr double LR_REST_V = -84.59413360428096;

const char *hip_str(hipError_t e) { return hipGetErrorString(e); }

// Partial sums of vg_ecg_device, the entry point that allocates nothing: one
// FP64 slot per lead and block of its launch, MAX_ECG leads x ECG_SCRATCH_BLOCKS
// blocks (128 KiB per GPU, held until the process ends).  Every call on one
// GPU shares it.  Host code never passes `scratch` itself to a kernel (on the
// host that name is not a device address): vg_create resolves its address on
// each device once (scratch_of) and kernels take that pointer.
constexpr int ECG_SCRATCH_BLOCKS = 1024;
static __device__ __align__(256) double scratch[MAX_ECG * ECG_SCRATCH_BLOCKS];

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device index; null until resolved

// Device address of `scratch` on `device`, which must be the current device.
hipError_t scratch_of(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0) return hipErrorInvalidDevice;
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t alloc(size_t b) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    bytes = b;
    return hipMalloc(&p, b ? b : 1);
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class X> X *as() const { return static_cast<X *>(p); }
};

struct ProfileInterval {
  hipEvent_t a, b;
  int launches;
};

struct ProfileClass {
  std::vector<ProfileInterval> pending;
  double total_ms = 0.0;
  int64_t launches = 0;
};

// States per node as getInternalParameters reports them, the slot holding V
// (kept on the device as the voltage vector instead, -1 if the model keeps its
// own V copy) and the number of state rows stored on the device.
int nstate_of(int model) {
  return model == VG_MODEL_LR ? LR_NSTATE
         : model == VG_MODEL_PURKINJE ? PK_NSTATE
         : model == VG_MODEL_MJ ? MJ_NSTATE
                                 : TT_NSTATE;
}
int vslot_of(int model) { return model == VG_MODEL_LR ? 7 : model == VG_MODEL_PURKINJE ? -1 : 0; }
// Mahajan tissue type of a mesh node: the type of the .cell data, 0 without
// one (Mesh/Mesh.cc:186,375); types beyond 8 take the reference's fallback,
// whose constants and initial state are those of type 2
// (SetUpIonicConstants.h:219-223, Mahajan.cc:147-158).
int mj_type(const int8_t *tissue, int64_t i) {
  const int t = tissue ? tissue[i] : 0;
  return (t >= 0 && t < 9) ? t : 2;
}
// Row of the per-type constants (G_tof, G_Ks): the type, or row 9 for types
// beyond 8 -- the default branch of the constant setup, which for the
// normal set equals type 2 and for the failing-heart set does not.
int mj_row(const int8_t *tissue, int64_t i) {
  const int t = tissue ? tissue[i] : 0;
  return (t >= 0 && t < 9) ? t : 9;
}
// LuoRudyGateTable (LRGateTable.h:67-76): make_parameter1 at -90 + 0.001 i,
// 14 values per row, in FP64 on the host.
std::vector<double> lr_gate_table() {
  std::vector<double> t((size_t)LR_TABLE_ROWS * 14);
  for (int i = 0; i < LR_TABLE_ROWS; i++) lr_tab<double, false>(-90 + 0.001 * i, &t[(size_t)i * 14]);
  return t;
}

// Per-type scaling rows (G_tof, G_Ks in slots 0, 1) of either constant set.
std::vector<double> mj_scale_rows(bool fail) {
  std::vector<double> sc((size_t)10 * VG_TT_NSCALE, 0.0);
  for (int t = 0; t < 10; t++) {
    sc[(size_t)t * VG_TT_NSCALE + 0] = fail ? MJF_GTOF[t] : MJ_GTOF[t < 9 ? t : 2];
    sc[(size_t)t * VG_TT_NSCALE + 1] = fail ? MJF_GKS[t] : MJ_GKS[t < 9 ? t : 2];
  }
  return sc;
}
int ndev_of(int model) { return nstate_of(model) - (vslot_of(model) >= 0 ? 1 : 0); }
const double *init_table(int model) {
  return model == VG_MODEL_LR      ? LR_INIT
         : model == VG_MODEL_PURKINJE ? PK_INIT
         : model == VG_MODEL_MJ       ? MJ_INIT[0]
                                      : TT_INIT;
}

// Ionic sub-step sequence of EP_CrankNicholson::Solve (:114-143).
bool substeps(double dT, double dt_ion, SubSteps &s) {
  s.n = 0;
  double total = 0.0;
  double dt = (dt_ion < dT) ? dt_ion : dT;
  while (total < dT) {
    if (s.n >= MAX_SUB) return false;
    s.dt[s.n++] = dt;
    total += dt;
    dt = (total + dt < dT) ? dt : dT - total;
  }
  return true;
}

}  // namespace

struct vg_handle {
  std::string err;
  int prec = VG_PREC_FP64, model = VG_MODEL_TT, integ = VG_INTEG_REFERENCE;
  unsigned quirks = 0;
  double dt = 0.01, dt_ion = 0.01, Cm = 1.0, tol = 1e-7;
  int maxit = 1000;
  // n: mesh nodes; rows: device rows (n plus the padding rows that align
  // each batched instance to a slice); n_alloc: rows rounded up to BLK.
  int64_t n = 0, rows = 0, n_alloc = 0, n_slices = 0, nnz = 0, padded = 0;
  int n_inst = 1;
  size_t tsz = 8;  // sizeof(T)
  hipStream_t stream = nullptr;
  int device = 0;
  std::vector<int32_t> new2old, old2new;
  std::vector<double> xyz;        // original order
  std::vector<char> purkinje;     // original order
  std::vector<double> CL;         // original order
  std::vector<double> istim_host; // original order
  bool stim_dirty = false, stim_active = false;
  SubSteps sub{};
  // device
  DevBuf off, width, wide, col16, col32, ustart, uoff, Ahat, Krhs, sinv, diag;
  DevBuf Va, Vb, r, x, states, istim, pidx, scales, at;
  // CG rows {r, w, s, p} interleaved, one array per parity of the iteration
  // (the explicit scheme uses q0 as two plain work vectors).
  DevBuf q0, q1;
  DevBuf acc;  // Mahajan: FP64 Ca-cycling and Na_i states [MJ_NACC][n_alloc]
  // Increment of the second half-step (x is the first's) unless both
  // half-steps share x (cg_x0 = VG_X0_LAST); with guess = 1 each solve starts
  // from what its increment buffer holds.
  DevBuf x2;
  int guess = 0;
  bool shared_x = false;
  // cg_x0 = VG_X0_EXTRAPOLATE: each half-step's two earlier increments
  // (xhist: d_{n-1} and d_{n-2} of half-step 0, then of half-step 1) and the
  // number of its solves so far (capped at 3).
  bool extrap = false;
  int extrap_order = 3;
  // Local solve of the Purkinje rows of a guess (k_pk_presolve): the rows'
  // block and the myocardium rows coupled to it, built at set-up when the
  // tree fits (pk_ok).
  bool pk_ok = false;
  int pk_ncpl = 0;
  // Gershgorin interval of the tree block; the Chebyshev form runs when it
  // is well inside (0, 2) and the direction fits LDS twice.
  bool pk_cheb = false;
  double pk_lmin = 0.0, pk_lmax = 0.0;
  PkCheb pk_coef{};  // Chebyshev coefficients of [pk_lmin, pk_lmax]
  // Operator values read with nontemporal loads (see ld_vals): chosen when a
  // CG iteration's streams exceed the Infinity Cache, or set by the caller.
  bool nt_vals = false;
  DevBuf pk_blk_v, pk_blk_j, pk_cpl_row, pk_cpl_v, pk_cpl_j;
  DevBuf xhist;
  int solves_done[2] = {0, 0};
  // Rows [0, n_myo) run the body's cell model, rows [n_myo, n) the Purkinje
  // model (their states in pk_states, SoA with n_pk_alloc lanes per state).
  int64_t n_myo = 0, n_pk = 0, n_pk_alloc = 0;
  DevBuf pk_states;
  // The Purkinje ionic kernel runs beside the myocardium's on a second
  // stream (forked and joined with events).
  hipStream_t stream_pk = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // CG: grid of the SpMV-type launches (multiple of XCDS), partials, carried
  // scalars and the result record (pinned copy for the host).
  unsigned cg_blocks = 0, node_blocks = 0;
  DevBuf part, carry, res, newton;
  // FP64 true residual after every tres_every-th step's solves (0 = off):
  // tres = {largest so far, checks, last}.
  int tres_every = 0;
  DevBuf tres;
  DevBuf blk_inst, inst_blk, inst_slice, slice_inst;  // instance tables (CgBufs)
  std::vector<int64_t> inst_slice_host;               // [n_inst + 1]
  DevBuf H, ecg_part, ecg_out, ecg_frames;
  double *ecg_scratch = nullptr;  // this device's `scratch` (vg_ecg_device)
  CgResult *h_res = nullptr;  // pinned
  CgResult *h_pub = nullptr;  // coherent host memory the finishing CG launch writes
  CgResult *d_pub = nullptr;  // its device address
  // Captured solves cut short so far, published by k_fold to coherent host
  // memory (h_trunc, device address d_trunc); replays do not publish to h_pub.
  unsigned long long *h_trunc = nullptr, *d_trunc = nullptr;
  bool capturing = false;
  int nsets = 1;
  int cg_guess[2] = {8, 8};  // launches expected per solve, by half-step
  int seen_max[2] = {0, 0};  // most iterations seen per half-step
  // Step counter on the device (the apply kernel's time) and the solve
  // totals of captured steps.
  DevBuf step_ctr, totals;
  CgTotals *h_tot = nullptr;  // pinned copy
  // Captured step (graph mode).
  int graph = 0, warm = 0;
  hipGraphExec_t gexec = nullptr;
  int glaunch[2] = {0, 0};
  bool g_stim = false, g_at = false;
  unsigned long long g_notconv_seen = 0, g_trunc_seen = 0;
  bool mj_fail = false;  // Mahajan with MAHAJAN_FAIL's constants
  DevBuf lr_table;       // LUORUDYWITHGATE's table in T, or empty
  // Step scheme; the explicit scheme's diffusion strides of a half-step.
  int scheme = VG_SCHEME_CN;
  double dt_diff = 0.1;
  std::vector<double> strides;
  // Learned ionic step: the attached network, its device view, the
  // per-row fallback flags and the usage counters (two per wave of the
  // persistent grid).  rec_*: recording of the plain update for training.
  vg_surr *surr = nullptr;
  SurrDevice sdev{};
  DevBuf surr_fb, surr_use;
  unsigned surr_grid = 0;
  float *rec_x = nullptr, *rec_y = nullptr;
  int64_t rec_base = 0, rec_cap = 0;
  int rec_stride = 1;
  // activation
  bool at_enabled = false;
  double at_thr = -40.0;
  // ecg
  int ne = 0;
  int ecg_every = 0, ecg_max = 0, ecg_count = 0;
  // stats
  int64_t steps = 0, cg_iters = 0, cg_solves = 0, cg_notconv = 0;
  int64_t n_wide = 0;             // slices on 32-bit columns
  int64_t n_uniform = 0;          // slices with shared column offsets
  int row_order = VG_ORDER_MESH;  // the myocardium rows' order (VG_ORDER_*)
  double last_relres = 0.0, t = 0.0;
  bool profiling = false;
  ProfileClass prof[VG_KC_COUNT];
  std::vector<hipEvent_t> ev_pool;

  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  int hipfail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hip_str(e);
    return VG_ERR_HIP;
  }
  ~vg_handle() {
    for (auto &c : prof)
      for (auto &pr : c.pending) {
        (void)hipEventDestroy(pr.a);
        (void)hipEventDestroy(pr.b);
      }
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (h_res) (void)hipHostFree(h_res);
    if (h_pub) (void)hipHostFree(h_pub);
    if (h_trunc) (void)hipHostFree(h_trunc);
    if (h_tot) (void)hipHostFree(h_tot);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream_pk) (void)hipStreamDestroy(stream_pk);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct vg_cells {
  int prec, model, integ;
  bool mj_fail = false;
  DevBuf lr_table;  // LUORUDYWITHGATE's table, or empty
  int newton_jt = 0, mj_int_abs = 0;
  int64_t n, n_alloc;
  DevBuf states, acc, scales, pidx, V, stim, dvdt;
  int nsets = 1;
};

#define HIPCHK(h, call)                                   \
  do {                                                    \
    hipError_t _e = (call);                               \
    if (_e != hipSuccess) return (h)->hipfail(_e, #call); \
  } while (0)

namespace {

// Launch bracketing for vg_profile_*: a pair of events around one launch, or
// around the `launches` back-to-back launches recorded in between.
struct ProfScope {
  vg_handle *h;
  int kc;
  int launches = 1;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(vg_handle *hh, int k) : h(hh), kc(k) {
    if (!h->profiling) return;
    if (h->ev_pool.size() < 2) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        if (hipEventCreate(&e) == hipSuccess) h->ev_pool.push_back(e);
      }
    }
    if (h->ev_pool.size() < 2) return;
    a = h->ev_pool.back();
    h->ev_pool.pop_back();
    b = h->ev_pool.back();
    h->ev_pool.pop_back();
    (void)hipEventRecord(a, h->stream);
  }
  ~ProfScope() {
    if (!a) return;
    (void)hipEventRecord(b, h->stream);
    h->prof[kc].pending.push_back(ProfileInterval{a, b, launches});
  }
};

void drain_profile(vg_handle *h) {
  for (auto &c : h->prof) {
    for (auto &pr : c.pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pr.a, pr.b) == hipSuccess) {
        c.total_ms += ms;
        c.launches += pr.launches;
      }
      h->ev_pool.push_back(pr.a);
      h->ev_pool.push_back(pr.b);
    }
    c.pending.clear();
  }
}

Sell sell_of(vg_handle *h) {
  return Sell{h->off.as<int64_t>(), h->width.as<int>(),  h->wide.as<int64_t>(),
              h->col16.as<int16_t>(), h->col32.as<int>(), h->ustart.as<int64_t>(),
              h->uoff.as<int>()};
}

CgBufs cgbufs_of(vg_handle *h) {
  double *part = h->part.as<double>();
  return CgBufs{part,
                part + 3 * (size_t)h->cg_blocks,
                h->carry.as<CgCarry>(),
                h->res.as<CgResult>(),
                h->blk_inst.as<int>(),
                h->inst_blk.as<int>(),
                h->inst_slice.as<int64_t>(),
                h->n_inst,
                h->n_slices,
                h->capturing ? nullptr : h->d_pub};
}

// Slice -> instance table for the node-wise kernels (none with one instance).
const int *slice_inst_of(vg_handle *h) {
  return h->n_inst > 1 ? h->slice_inst.as<int>() : nullptr;
}

// Increment buffer of half-step `which` (0: before the ionic update).
template <class T>
T *xbuf(vg_handle *h, int which) {
  return (which && !h->shared_x) ? h->x2.as<T>() : h->x.as<T>();
}

// Local solve of the Purkinje rows of the guess, between the right-hand
// side and the start of a solve that starts from a guess.
template <class T>
void launch_pk_presolve(vg_handle *h, int which) {
  if (!h->guess || !h->pk_ok) return;
  ProfScope ps(h, VG_KC_PK_PRESOLVE);
  if (h->pk_cheb)
    hipLaunchKernelGGL(k_pk_cheb<T>, dim3(1), dim3(PK_BLK), (unsigned)(2 * h->n_pk * sizeof(T)),
                       h->stream, h->r.as<T>(), xbuf<T>(h, which), h->pk_blk_v.as<T>(),
                       h->pk_blk_j.as<int>(), h->n_myo, (int)h->n_pk,
                       h->pk_cpl_row.as<int64_t>(), h->pk_cpl_v.as<T>(), h->pk_cpl_j.as<int>(),
                       h->pk_ncpl, h->pk_coef);
  else
    hipLaunchKernelGGL(k_pk_presolve<T>, dim3(1), dim3(PK_BLK), (unsigned)(h->n_pk * sizeof(T)),
                       h->stream, h->r.as<T>(), xbuf<T>(h, which), h->pk_blk_v.as<T>(),
                       h->pk_blk_j.as<int>(), h->n_myo, (int)h->n_pk, h->pk_cpl_row.as<int64_t>(),
                       h->pk_cpl_v.as<T>(), h->pk_cpl_j.as<int>(), h->pk_ncpl);
}

// Nontemporal value loads pay when one CG iteration's streams -- the operator's
// stored values, the row records read and written, x, diag -- exceed the
// 256 MB Infinity Cache (the 4.6M-row slab: ~440 MB, 5.6 % faster); below it
// they keep the values from staying resident there (the 1.03M-row LV: ~100 MB,
// 19 % slower).
bool nt_vals_auto(const vg_handle *h) {
  const double bytes = (double)h->tsz * ((double)h->padded + 11.0 * (double)h->rows);
  return bytes > 256.0 * 1024.0 * 1024.0;
}

template <class T>
void launch_rhs(vg_handle *h, const T *V, int which) {
  ProfScope ps(h, VG_KC_RHS);
  auto kern = h->guess ? (h->nt_vals ? k_rhs<T, true, true> : k_rhs<T, true, false>)
                       : (h->nt_vals ? k_rhs<T, false, true> : k_rhs<T, false, false>);
  hipLaunchKernelGGL(kern, dim3(h->cg_blocks), dim3(BLK), 0, h->stream, sell_of(h),
                     h->Krhs.as<T>(), h->Ahat.as<T>(), V, xbuf<T>(h, which), h->r.as<T>(),
                     h->diag.as<T>(), cgbufs_of(h));
}

// The solve of half-step `which` turns the other half-step's history into
// that one's next guess (its increment has been used by then).
template <class T>
void launch_cg_start(vg_handle *h, int which) {
  ProfScope ps(h, VG_KC_CG_START);
  XHist<T> hist{nullptr, nullptr, nullptr, 0};
  if (h->extrap) {
    const int o = 1 - which;
    T *base = h->xhist.as<T>() + (size_t)(2 * o) * h->n_alloc;
    hist = XHist<T>{xbuf<T>(h, o), base, base + h->n_alloc,
                    std::min(h->extrap_order, h->solves_done[o])};
  }
  hipLaunchKernelGGL((h->nt_vals ? k_cg_start<T, true> : k_cg_start<T, false>),
                     dim3(h->cg_blocks), dim3(BLK), 0, h->stream,
                     sell_of(h), h->Ahat.as<T>(),
                     h->r.as<T>(), h->q0.as<typename Vec4Of<T>::type>(), h->diag.as<T>(),
                     cgbufs_of(h), hist);
}

// (Profiled per solve, around all its iteration launches: see cg_solve.)
template <class T>
void launch_cg_iter(vg_handle *h, int j, int which) {
  hipLaunchKernelGGL((h->nt_vals ? k_cg_iter<T, true> : k_cg_iter<T, false>),
                     dim3(h->cg_blocks), dim3(BLK), 0, h->stream,
                     sell_of(h), h->Ahat.as<T>(),
                     h->q0.as<typename Vec4Of<T>::type>(), h->q1.as<typename Vec4Of<T>::type>(),
                     xbuf<T>(h, which), h->diag.as<T>(),
                     cgbufs_of(h), j, h->tol * h->tol, h->maxit, h->guess);
}

// FP64 true residual of the solve of half-step `which` that just ended
// (right-hand side from V, increment in its buffer), folded into tres.
template <class T>
void launch_true_res(vg_handle *h, const T *V, int which) {
  double *part = h->part.as<double>() + 6 * (size_t)h->cg_blocks;
  hipLaunchKernelGGL(k_true_res<T>, dim3(h->cg_blocks), dim3(BLK), 0, h->stream, sell_of(h),
                     h->Krhs.as<T>(), h->Ahat.as<T>(), V, xbuf<T>(h, which), h->diag.as<T>(),
                     cgbufs_of(h), part);
  hipLaunchKernelGGL(k_true_finish, dim3(1), dim3(BLK), 0, h->stream, part, (int)h->cg_blocks,
                     cgbufs_of(h), h->tres.as<double>());
}

// Folds the result records of a finished solve (one per instance) into the
// statistics and the expected launch count.
int cg_account(vg_handle *h, int which) {
  int max_iter = 0;
  bool nonfinite = false;
  h->last_relres = 0.0;
  for (int k = 0; k < h->n_inst; k++) {
    const CgResult &c = h->h_res[k];
    if (!c.done) return h->fail(VG_ERR_HIP, "CG launches ended without a result");
    h->cg_iters += c.iter;
    h->cg_solves += 1;
    h->last_relres = std::max(h->last_relres, c.rho_ref > 0 ? std::sqrt(c.rho / c.rho_ref) : 0.0);
    if (c.notconv) h->cg_notconv += 1;
    if (c.notconv && !(std::isfinite(c.rho) && std::isfinite(c.rho_ref))) nonfinite = true;
    max_iter = std::max(max_iter, c.iter);
  }
  h->seen_max[which] = std::max(h->seen_max[which], max_iter);
  // The launch that finds convergence is the (iter+1)-th; one more spare.
  h->cg_guess[which] = std::min(h->maxit + 1, max_iter + 2);
  if (nonfinite) return h->fail(VG_ERR_NONFINITE, "non-finite CG residual");
  return VG_OK;
}

// One diffusion half-step: r_0 = K_rhs V, then CG to convergence.  The
// expected number of iteration launches (from the previous solve of this
// half-step) is queued, then the result record is read back; launches after
// the solve has ended return at once, and a solve still running gets more.
//
// The launch that ends a solve also writes its result to coherent host memory
// (h_pub), so the host goes on as soon as that record appears instead of
// waiting for the stream; the device record, copied once the stream has
// drained, decides whenever the published one has not appeared by then.
template <class T>
int cg_wait(vg_handle *h, bool &all) {
  using clock = std::chrono::steady_clock;
  auto next_query = clock::now() + std::chrono::microseconds(20);
  for (;;) {
    all = true;
    for (int k = 0; k < h->n_inst && all; k++)
      all = __atomic_load_n(&h->h_pub[k].done, __ATOMIC_ACQUIRE) != 0;
    if (all) {
      std::memcpy(h->h_res, h->h_pub, sizeof(CgResult) * h->n_inst);
      return VG_OK;
    }
    if (clock::now() < next_query) continue;
    const hipError_t q = hipStreamQuery(h->stream);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) HIPCHK(h, q);
    next_query = clock::now() + std::chrono::microseconds(20);
  }
  HIPCHK(h, hipMemcpy(h->h_res, h->res.p, sizeof(CgResult) * h->n_inst, hipMemcpyDeviceToHost));
  all = true;
  for (int k = 0; k < h->n_inst; k++) all = all && h->h_res[k].done;
  return VG_OK;
}

template <class T>
int cg_solve(vg_handle *h, const T *V, int which) {
  for (int k = 0; k < h->n_inst; k++) __atomic_store_n(&h->h_pub[k].done, 0, __ATOMIC_SEQ_CST);
  launch_rhs<T>(h, V, which);
  launch_pk_presolve<T>(h, which);
  launch_cg_start<T>(h, which);
  int j = 0;
  int batch = h->cg_guess[which];
  // One event pair around all iteration launches of the solve: an event pair
  // around each launch would add its own markers to every launch boundary.
  ProfScope ps(h, VG_KC_CG_ITER);
  for (;;) {
    for (int k = 0; k < batch && j <= h->maxit; k++) launch_cg_iter<T>(h, j++, which);
    HIPCHK(h, hipGetLastError());
    bool all = false;
    const int rc = cg_wait<T>(h, all);
    if (rc) return rc;
    if (all || j > h->maxit) break;
    batch = 4;
  }
  ps.launches = j;
  h->solves_done[which] = std::min(3, h->solves_done[which] + 1);
  return cg_account(h, which);
}

template <class T, int MODEL, int INTEG>
void launch_ionic_t(vg_handle *h, const T *Vin, const T *x, T *Vout) {
  IonicArgs a{h->n_myo, h->n_alloc, h->nsets, h->Cm, 500, (h->quirks & VG_QUIRK_NEWTON_JT) ? 1 : 0,
              (h->quirks & VG_QUIRK_MJ_INT_ABS) ? 1 : 0};
  a.rec_x = h->rec_x;
  a.rec_y = h->rec_y;
  a.rec_base = h->rec_base;
  a.rec_cap = h->rec_cap;
  a.rec_stride = h->rec_stride;
  a.log_dt = (float)std::log(h->dt);
  a.only = nullptr;
  a.lr_table = h->lr_table.p;
  if constexpr (MODEL == VG_MODEL_TT) if (h->surr && !h->rec_x) {
    // The network advances the rows inside its envelope and flags the rest,
    // which the plain kernel below then updates.
    hipLaunchKernelGGL(k_ionic_surr<T>, dim3(h->surr_grid), dim3(SBLK), 0, h->stream, Vin, x,
                       h->sinv.as<T>(), Vout, h->states.as<T>(),
                       h->stim_active ? h->istim.as<T>() : (const T *)nullptr,
                       h->pidx.p ? h->pidx.as<uint16_t>() : (const uint16_t *)nullptr,
                       h->scales.as<T>(), h->n_myo, h->n_alloc, a.log_dt, h->sdev,
                       h->surr_fb.as<unsigned char>(), h->surr_use.as<unsigned long long>(),
                       h->res.as<CgResult>(), slice_inst_of(h));
    a.only = h->surr_fb.as<unsigned char>();
  }
  auto kern = h->sub.adaptive ? k_ionic<T, MODEL, INTEG, true> : k_ionic<T, MODEL, INTEG, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)((h->n_myo + BLK - 1) / BLK)),
                     dim3(BLK), 0, h->stream, Vin, x, h->sinv.as<T>(), Vout, h->states.as<T>(),
                     h->acc.p ? h->acc.as<double>() : (double *)nullptr,
                     h->stim_active ? h->istim.as<T>() : (const T *)nullptr,
                     h->pidx.p ? h->pidx.as<uint16_t>() : (const uint16_t *)nullptr,
                     h->scales.as<T>(), h->sub, a, h->newton.as<unsigned long long>(),
                     h->res.as<CgResult>(), slice_inst_of(h));
}

// Ionic update of every node.  The Purkinje rows (few, long per-node work)
// run on the second stream beside the myocardium's kernel.
template <class T>
int launch_ionic(vg_handle *h, const T *Vin, const T *x, T *Vout) {
  ProfScope ps(h, VG_KC_IONIC);
  const bool fork = h->n_pk > 0 && h->n_myo > 0;
  if (h->n_pk > 0) {
    hipStream_t sp = h->stream;
    if (fork) {
      HIPCHK(h, hipEventRecord(h->ev_fork, h->stream));
      HIPCHK(h, hipStreamWaitEvent(h->stream_pk, h->ev_fork, 0));
      sp = h->stream_pk;
    }
    auto kpk = h->sub.adaptive ? k_ionic_pk<T, true> : k_ionic_pk<T, false>;
    hipLaunchKernelGGL(kpk, dim3((unsigned)(h->n_pk_alloc / BLK)), dim3(BLK), 0, sp, Vin,
                       x, h->sinv.as<T>(), Vout, h->pk_states.as<T>(),
                       h->stim_active ? h->istim.as<T>() : (const T *)nullptr, h->n_myo, h->n_pk,
                       h->n_pk_alloc, h->sub, h->res.as<CgResult>(), slice_inst_of(h));
    if (fork) HIPCHK(h, hipEventRecord(h->ev_join, h->stream_pk));
  }
  if (h->n_myo > 0) {
    if (h->model == VG_MODEL_LR) launch_ionic_t<T, VG_MODEL_LR, VG_INTEG_REFERENCE>(h, Vin, x, Vout);
    else if (h->model == VG_MODEL_MJ && h->mj_fail)
      launch_ionic_t<T, VG_MODEL_MJ_FAIL, VG_INTEG_REFERENCE>(h, Vin, x, Vout);
    else if (h->model == VG_MODEL_MJ)
      launch_ionic_t<T, VG_MODEL_MJ, VG_INTEG_REFERENCE>(h, Vin, x, Vout);
    else if (h->integ == VG_INTEG_REFERENCE)
      launch_ionic_t<T, VG_MODEL_TT, VG_INTEG_REFERENCE>(h, Vin, x, Vout);
    else launch_ionic_t<T, VG_MODEL_TT, VG_INTEG_RUSH_LARSEN>(h, Vin, x, Vout);
  }
  if (fork) HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_join, 0));
  return VG_OK;
}

// Folds half-step `which`'s results into the device totals and/or advances
// the device step counter (one thread).
void launch_fold(vg_handle *h, int which, bool fold, bool tick) {
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(64), 0, h->stream, h->res.as<CgResult>(), h->n_inst,
                     h->totals.as<CgTotals>(), which, fold ? 1 : 0, h->step_ctr.as<int64_t>(),
                     tick ? 1 : 0, h->d_trunc);
}

// Sums the ECG partials of the last k_apply / k_ecg launch into dst[ne].
void launch_ecg_finish(vg_handle *h, double *dst) {
  hipLaunchKernelGGL(k_ecg_finish, dim3(1), dim3(BLK), 0, h->stream, h->ecg_part.as<double>(),
                     (int)h->node_blocks, h->ne, dst);
}

// V_a = from (+ D^-1/2 x when x is given and valid), activation, ECG frame.
template <class T>
void apply_from(vg_handle *h, const T *from, const T *x, bool ecg_frame) {
  ProfScope ps(h, VG_KC_APPLY);
  const T thr = (T)h->at_thr;
  const T *Hm = h->H.p ? h->H.as<T>() : nullptr;
  double *part = h->ecg_part.p ? h->ecg_part.as<double>() : nullptr;
#define VG_APPLY(AT, ECG)                                                                        \
  hipLaunchKernelGGL((k_apply<T, AT, ECG>), dim3(h->node_blocks), dim3(BLK), 0, h->stream,       \
                     from, x, h->sinv.as<T>(), h->Va.as<T>(), h->rows,                            \
                     h->at.as<T>(), thr, h->step_ctr.as<int64_t>(), h->dt, Hm, h->ne,           \
                     h->n_alloc, part,                                                            \
                     h->res.as<CgResult>(), slice_inst_of(h))
  if (h->at_enabled) {
    if (ecg_frame) VG_APPLY(true, true);
    else VG_APPLY(true, false);
  } else {
    if (ecg_frame) VG_APPLY(false, true);
    else VG_APPLY(false, false);
  }
#undef VG_APPLY
  if (ecg_frame)
    launch_ecg_finish(h, h->ecg_frames.as<double>() + (size_t)h->ecg_count * h->ne);
}

template <class T>
void launch_apply(vg_handle *h, bool ecg_frame) {
  apply_from<T>(h, h->Vb.as<T>(), xbuf<T>(h, 1), ecg_frame);
}

int upload_stim(vg_handle *h) {
  if (!h->stim_dirty) return VG_OK;
  h->stim_active = false;
  for (double v : h->istim_host)
    if (v != 0.0) {
      h->stim_active = true;
      break;
    }
  const int64_t rows = h->rows;
  if (h->tsz == 8) {
    std::vector<double> buf(h->n_alloc, 0.0);
    for (int64_t r = 0; r < rows; r++)
      if (h->new2old[r] >= 0) buf[r] = h->istim_host[h->new2old[r]];
    HIPCHK(h, hipMemcpyAsync(h->istim.p, buf.data(), buf.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  } else {
    std::vector<float> buf(h->n_alloc, 0.0f);
    for (int64_t r = 0; r < rows; r++)
      if (h->new2old[r] >= 0) buf[r] = (float)h->istim_host[h->new2old[r]];
    HIPCHK(h, hipMemcpyAsync(h->istim.p, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  h->stim_dirty = false;
  return VG_OK;
}

template <class T>
int step_t(vg_handle *h) {
  int rc = upload_stim(h);
  if (rc) return rc;
  const bool check = h->tres_every > 0 && h->steps % h->tres_every == 0;
  // Step 1: half-step diffusion, C_m M + dT/4 K against C_m M - dT/4 K.
  if ((rc = cg_solve<T>(h, h->Va.as<T>(), 0))) return rc;
  if (check) launch_true_res<T>(h, h->Va.as<T>(), 0);
  // Step 2: ionic update over dT in the reference's sub-steps, fused with
  // V = V_a + D^-1/2 x.
  if ((rc = launch_ionic<T>(h, h->Va.as<T>(), xbuf<T>(h, 0), h->Vb.as<T>()))) return rc;
  // Step 3: second half-step.
  if ((rc = cg_solve<T>(h, h->Vb.as<T>(), 1))) return rc;
  if (check) launch_true_res<T>(h, h->Vb.as<T>(), 1);
  h->steps += 1;
  h->t = h->steps * h->dt;
  const bool frame = h->ecg_every > 0 && h->ne > 0 && (h->steps % h->ecg_every) == 0 &&
                     h->ecg_count < h->ecg_max;
  launch_apply<T>(h, frame);
  launch_fold(h, 1, false, true);
  if (frame) h->ecg_count += 1;
  return VG_OK;
}

// One step of EP_ExplicitSolver::Solve (EP_ExplicitSolver.cc:23-118):
// forward-Euler diffusion strides over dT/2, the per-node adaptive ionic
// update, strides over dT/2 again.  Strides ping-pong between two work
// vectors (Va keeps V^n for the activation test); k_apply writes Va.
template <class T>
int step_explicit(vg_handle *h) {
  int rc = upload_stim(h);
  if (rc) return rc;
  const unsigned grid = (unsigned)std::min<int64_t>(h->cg_blocks ? h->cg_blocks : 2048,
                                                    (h->n_slices + WAVES - 1) / WAVES);
  T *W1 = h->q0.as<T>(), *W2 = h->q0.as<T>() + h->n_alloc;
  auto half = [&](const T *from) {
    const T *src = from;
    T *dst = W1;
    for (double s : h->strides) {
      hipLaunchKernelGGL(k_explicit<T>, dim3(grid), dim3(BLK), 0, h->stream, sell_of(h),
                         h->Krhs.as<T>(), src, dst, (T)s, h->n_slices);
      src = dst;
      dst = dst == W1 ? W2 : W1;
    }
    return src;
  };
  const T *v1 = half(h->Va.as<T>());
  if ((rc = launch_ionic<T>(h, v1, (const T *)nullptr, h->Vb.as<T>()))) return rc;
  const T *v2 = half(h->Vb.as<T>());
  h->steps += 1;
  h->t = h->steps * h->dt;
  const bool frame = h->ecg_every > 0 && h->ne > 0 && (h->steps % h->ecg_every) == 0 &&
                     h->ecg_count < h->ecg_max;
  apply_from<T>(h, v2, (const T *)nullptr, frame);
  launch_fold(h, 1, false, true);
  if (frame) h->ecg_count += 1;
  return VG_OK;
}

// Captures one whole step with glaunch[w] CG launches per half-step.
template <class T>
int capture_step(vg_handle *h) {
  if (h->gexec) {
    HIPCHK(h, hipGraphExecDestroy(h->gexec));
    h->gexec = nullptr;
  }
  for (int w = 0; w < 2; w++)
    h->glaunch[w] = std::min(h->maxit + 1, h->seen_max[w] + std::max(4, h->seen_max[w] / 4) + 1);
  // Replays run while the host is elsewhere: their solves must not publish
  // into the record a host-driven solve waits on.
  struct Capturing {
    vg_handle *h;
    explicit Capturing(vg_handle *hh) : h(hh) { h->capturing = true; }
    ~Capturing() { h->capturing = false; }
  } capturing_scope(h);
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const T *V[2] = {h->Va.as<T>(), h->Vb.as<T>()};
  for (int w = 0; w < 2; w++) {
    launch_rhs<T>(h, V[w], w);
    launch_pk_presolve<T>(h, w);
    launch_cg_start<T>(h, w);
    for (int j = 0; j < h->glaunch[w]; j++) launch_cg_iter<T>(h, j, w);
    launch_fold(h, w, true, false);
    if (w == 0) {
      const int rc = launch_ionic<T>(h, h->Va.as<T>(), xbuf<T>(h, 0), h->Vb.as<T>());
      if (rc) {
        hipGraph_t g;
        (void)hipStreamEndCapture(h->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
    }
  }
  launch_apply<T>(h, false);
  launch_fold(h, 1, false, true);
  hipGraph_t g = nullptr;
  HIPCHK(h, hipStreamEndCapture(h->stream, &g));
  const hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(h, e);
  h->g_stim = h->stim_active;
  h->g_at = h->at_enabled;
  return VG_OK;
}

// Graph mode: the first steps run host-driven (learning the launch counts),
// then the captured step is replayed.  ECG frames are taken after a replay
// from the new V.  After every replay the host reads the count of solves cut
// short (published by k_fold into coherent host memory, no synchronisation):
// when it has grown, the host waits for the stream, doubles the launch counts
// and re-captures before the next step.  A cut-short solve passes its last
// iterate on; it is counted in vg_stats.cg_truncated and cg_not_converged.
// Every 32 steps the device totals are read for non-finite residuals.  A step
// that is due an FP64 true-residual check runs host-driven (bitwise the same
// step), so the captured graph carries no check.
template <class T>
int step_graph(vg_handle *h) {
  int rc = upload_stim(h);
  if (rc) return rc;
  if (h->warm < 3) {
    h->warm++;
    return step_t<T>(h);
  }
  if (h->tres_every > 0 && h->steps % h->tres_every == 0) return step_t<T>(h);
  if (!h->gexec || h->g_stim != h->stim_active || h->g_at != h->at_enabled)
    if ((rc = capture_step<T>(h))) return rc;
  HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
  h->steps += 1;
  h->t = h->steps * h->dt;
  if (__atomic_load_n(h->h_trunc, __ATOMIC_ACQUIRE) > h->g_trunc_seen) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(h->h_tot, h->totals.p, sizeof(CgTotals), hipMemcpyDeviceToHost));
    h->g_trunc_seen = __atomic_load_n(h->h_trunc, __ATOMIC_ACQUIRE);
    h->g_notconv_seen = h->h_tot->notconv;
    for (int w = 0; w < 2; w++)
      h->seen_max[w] = std::max(h->seen_max[w], std::max(h->glaunch[w], h->h_tot->max_iter[w])) * 2;
    if ((rc = capture_step<T>(h))) return rc;
  }
  if (h->ecg_every > 0 && h->ne > 0 && (h->steps % h->ecg_every) == 0 && h->ecg_count < h->ecg_max) {
    hipLaunchKernelGGL(k_ecg<T>, dim3(h->node_blocks), dim3(BLK), 0, h->stream, h->Va.as<T>(),
                       h->rows, h->H.as<T>(), h->ne, h->n_alloc, h->ecg_part.as<double>());
    launch_ecg_finish(h, h->ecg_frames.as<double>() + (size_t)h->ecg_count * h->ne);
    h->ecg_count += 1;
  }
  if (h->steps % 32 == 0) {
    HIPCHK(h, hipMemcpyAsync(h->h_tot, h->totals.p, sizeof(CgTotals), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->h_tot->nonfinite) return h->fail(VG_ERR_NONFINITE, "non-finite CG residual");
    if (h->h_tot->notconv > h->g_notconv_seen) {
      // Solves that hit cg_maxit or broke down inside the captured launches
      // (those cut short were handled after their replay).
      h->g_notconv_seen = h->h_tot->notconv;
    }
  }
  return VG_OK;
}

template <class T>
int upload_vec(vg_handle *h, void *dst, const double *src_orig, double fill) {
  std::vector<T> buf(h->n_alloc, (T)fill);
  for (int64_t r = 0; r < h->rows; r++)
    if (h->new2old[r] >= 0) buf[r] = (T)src_orig[h->new2old[r]];
  HIPCHK(h, hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return VG_OK;
}

template <class T>
int download_vec(vg_handle *h, const void *src, double *dst_orig) {
  std::vector<T> buf(h->n_alloc);
  HIPCHK(h, hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(T), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int64_t r = 0; r < h->rows; r++)
    if (h->new2old[r] >= 0) dst_orig[h->new2old[r]] = (double)buf[r];
  return VG_OK;
}

int upload(vg_handle *h, void *dst, const double *src_orig, double fill = 0.0) {
  return h->tsz == 8 ? upload_vec<double>(h, dst, src_orig, fill)
                     : upload_vec<float>(h, dst, src_orig, fill);
}
int download(vg_handle *h, const void *src, double *dst_orig) {
  return h->tsz == 8 ? download_vec<double>(h, src, dst_orig) : download_vec<float>(h, src, dst_orig);
}

template <class T>
int upload_raw(vg_handle *h, DevBuf &b, const std::vector<double> &v) {
  std::vector<T> buf(v.begin(), v.end());
  HIPCHK(h, b.alloc(buf.size() * sizeof(T)));
  HIPCHK(h, hipMemcpy(b.p, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice));
  return VG_OK;
}

// The Purkinje rows' block of A_hat and the myocardium rows coupled to it,
// in the entry-major form k_pk_presolve reads; pk_ok stays false when there
// is no tree, several instances, a tree too large for one workgroup, or a
// row with more entries than the kernel keeps.
int setup_pk_presolve(vg_handle *h, const HostSell &sell) {
  h->pk_ok = false;
  const int64_t r0 = h->n_myo, npk = h->n_pk;
  const int64_t cap = std::min<int64_t>(
      (int64_t)PK_BLK * (h->tsz == 8 ? pk_rows<double>() : pk_rows<float>()), 61440 / (int64_t)h->tsz);
  if (npk == 0 || npk > cap || h->n_inst != 1) return VG_OK;
  // Stored entries of one row.
  auto visit = [&](int64_t row, auto f) {
    const int64_t s = row / SELL_C;
    const int lane = (int)(row % SELL_C);
    for (int j = 0; j < sell.width[s]; j++) {
      const int64_t e = sell_slot(sell.offset[s], sell.width[s], j, lane);
      f((int64_t)sell.col[e], sell.A_hat[e]);
    }
  };
  auto in_tree = [&](int64_t c) { return c >= r0 && c < r0 + npk; };
  std::vector<double> bv(PK_MAXE * npk, 0.0);
  std::vector<int> bj(PK_MAXE * npk, 0);
  std::vector<int64_t> myo;
  for (int64_t i = 0; i < npk; i++) {
    // Slot 0: the unit diagonal of A_hat (not stored in the SELL operator).
    bv[i] = 1.0;
    bj[i] = (int)i;
    int m = 1;
    visit(r0 + i, [&](int64_t c, double v) {
      if (v == 0.0) return;
      if (!in_tree(c)) {
        myo.push_back(c);
        return;
      }
      if (m < PK_MAXE) {
        bv[m * npk + i] = v;
        bj[m * npk + i] = (int)(c - r0);
      }
      m++;
    });
    if (m > PK_MAXE) return VG_OK;
  }
  std::sort(myo.begin(), myo.end());
  myo.erase(std::unique(myo.begin(), myo.end()), myo.end());
  const int64_t nc = (int64_t)myo.size();
  std::vector<double> cv(CPL_MAXE * std::max<int64_t>(nc, 1), 0.0);
  std::vector<int> cj(CPL_MAXE * std::max<int64_t>(nc, 1), 0);
  for (int64_t c = 0; c < nc; c++) {
    int m = 0;
    visit(myo[c], [&](int64_t col, double v) {
      if (v == 0.0 || !in_tree(col)) return;
      if (m < CPL_MAXE) {
        cv[m * nc + c] = v;
        cj[m * nc + c] = (int)(col - r0);
      }
      m++;
    });
    if (m > CPL_MAXE) return VG_OK;
  }
  int rc;
  if (h->tsz == 8) {
    if ((rc = upload_raw<double>(h, h->pk_blk_v, bv))) return rc;
    if ((rc = upload_raw<double>(h, h->pk_cpl_v, cv))) return rc;
  } else {
    if ((rc = upload_raw<float>(h, h->pk_blk_v, bv))) return rc;
    if ((rc = upload_raw<float>(h, h->pk_cpl_v, cv))) return rc;
  }
  HIPCHK(h, h->pk_blk_j.alloc(bj.size() * 4));
  HIPCHK(h, hipMemcpy(h->pk_blk_j.p, bj.data(), bj.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(h, h->pk_cpl_j.alloc(cj.size() * 4));
  HIPCHK(h, hipMemcpy(h->pk_cpl_j.p, cj.data(), cj.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(h, h->pk_cpl_row.alloc(std::max<int64_t>(nc, 1) * 8));
  if (nc) HIPCHK(h, hipMemcpy(h->pk_cpl_row.p, myo.data(), nc * 8, hipMemcpyHostToDevice));
  h->pk_ncpl = (int)nc;
  // Gershgorin: A_pp has a unit diagonal, so its eigenvalues lie within
  // 1 -+ the largest off-diagonal row sum of |a_ij|.
  double smax = 0.0;
  for (int64_t i = 0; i < npk; i++) {
    double si = 0.0;
    for (int m = 1; m < PK_MAXE; m++) si += std::fabs(bv[m * npk + i]);
    smax = std::max(smax, si);
  }
  h->pk_lmin = 1.0 - smax;
  h->pk_lmax = 1.0 + smax;
  h->pk_cheb = h->pk_lmin >= 0.1 && 2 * npk * (int64_t)h->tsz <= 61440;
  if (h->pk_cheb) h->pk_coef = pk_cheb_coef(h->pk_lmin, h->pk_lmax);
  h->pk_ok = true;
  return VG_OK;
}

// The cell model of a node: Purkinje nodes always run the Purkinje model
// (CardiacBody.cc:993-1002), all others the body's model.
int node_model(const vg_handle *h, int64_t orig) {
  return h->purkinje[orig] ? VG_MODEL_PURKINJE : h->model;
}

template <class T>
int put_soa(vg_handle *h, DevBuf &b, const std::vector<double> &soa) {
  if (soa.empty()) return VG_OK;
  std::vector<T> f(soa.begin(), soa.end());
  HIPCHK(h, hipMemcpy(b.p, f.data(), f.size() * sizeof(T), hipMemcpyHostToDevice));
  return VG_OK;
}
template <class T>
int get_soa(vg_handle *h, const DevBuf &b, std::vector<double> &soa) {
  if (soa.empty()) return VG_OK;
  std::vector<T> f(soa.size());
  HIPCHK(h, hipMemcpy(f.data(), b.p, f.size() * sizeof(T), hipMemcpyDeviceToHost));
  std::copy(f.begin(), f.end(), soa.begin());
  return VG_OK;
}

// Writes per-node state vectors (getInternalParameters layout, node order of
// the mesh, offsets[i] = start of node i in data) into the two device SoA
// arrays: the body's model for rows [0, n_myo), Purkinje for [n_myo, n).
// The V slot of LR/tT is held by the voltage vector instead.
int set_states_impl(vg_handle *h, const int64_t *offsets, const double *data) {
  const int nd = ndev_of(h->model), vs = vslot_of(h->model);
  const int64_t na = h->n_alloc, npa = h->n_pk_alloc;
  std::vector<double> soa((size_t)nd * na, 0.0), pk((size_t)PK_NSTATE * npa, 0.0);
  int64_t first_real = -1;
  for (int64_t r = 0; r < h->rows; r++) {
    const int64_t o = h->new2old[r];
    if (o < 0) continue;
    if (first_real < 0) first_real = r;
    const double *st = data + offsets[o];
    if (r < h->n_myo) {
      int k2 = 0;
      for (int k = 0; k < nstate_of(h->model); k++) {
        if (k == vs) continue;
        soa[(size_t)k2++ * na + r] = st[k];
      }
    } else {
      for (int k = 0; k < PK_NSTATE; k++) pk[(size_t)k * npa + (r - h->n_myo)] = st[k];
    }
  }
  // Padding rows and lanes copy a real node's states so every lane computes
  // finite values.
  if (h->n_myo > 0 && first_real >= 0)
    for (int64_t r = 0; r < h->n_myo; r++)
      if (h->new2old[r] < 0)
        for (int k = 0; k < nd; k++) soa[(size_t)k * na + r] = soa[(size_t)k * na + first_real];
  if (h->n_myo > 0)
    for (int k = 0; k < nd; k++)
      for (int64_t r = h->n_myo; r < na; r++) soa[(size_t)k * na + r] = soa[(size_t)k * na];
  if (h->n_pk > 0)
    for (int k = 0; k < PK_NSTATE; k++)
      for (int64_t r = h->n_pk; r < npa; r++) pk[(size_t)k * npa + r] = pk[(size_t)k * npa];
  int rc;
  if (h->model == VG_MODEL_MJ) {
    // Rows in state order 1..25 -> gate rows (T) and accumulator rows (FP64).
    std::vector<double> g((size_t)MJ_NGATE * na), a((size_t)MJ_NACC * na);
    for (int k = 0; k < MJ_NGATE; k++)
      std::copy_n(soa.begin() + (size_t)(MJ_GATE[k] - 1) * na, na, g.begin() + (size_t)k * na);
    for (int k = 0; k < MJ_NACC; k++)
      std::copy_n(soa.begin() + (size_t)(MJ_ACC[k] - 1) * na, na, a.begin() + (size_t)k * na);
    rc = h->tsz == 8 ? put_soa<double>(h, h->states, g) : put_soa<float>(h, h->states, g);
    if (rc) return rc;
    rc = put_soa<double>(h, h->acc, a);
  } else {
    rc = h->tsz == 8 ? put_soa<double>(h, h->states, soa) : put_soa<float>(h, h->states, soa);
  }
  if (rc) return rc;
  if (h->n_pk > 0) rc = h->tsz == 8 ? put_soa<double>(h, h->pk_states, pk) : put_soa<float>(h, h->pk_states, pk);
  return rc;
}

int get_states_impl(vg_handle *h, const int64_t *offsets, double *data, const double *V) {
  const int nd = ndev_of(h->model), vs = vslot_of(h->model);
  const int64_t na = h->n_alloc, npa = h->n_pk_alloc;
  std::vector<double> soa((size_t)nd * na), pk((size_t)PK_NSTATE * (h->n_pk > 0 ? npa : 0));
  int rc;
  if (h->model == VG_MODEL_MJ) {
    std::vector<double> g((size_t)MJ_NGATE * na), a((size_t)MJ_NACC * na);
    rc = h->tsz == 8 ? get_soa<double>(h, h->states, g) : get_soa<float>(h, h->states, g);
    if (rc) return rc;
    if ((rc = get_soa<double>(h, h->acc, a))) return rc;
    for (int k = 0; k < MJ_NGATE; k++)
      std::copy_n(g.begin() + (size_t)k * na, na, soa.begin() + (size_t)(MJ_GATE[k] - 1) * na);
    for (int k = 0; k < MJ_NACC; k++)
      std::copy_n(a.begin() + (size_t)k * na, na, soa.begin() + (size_t)(MJ_ACC[k] - 1) * na);
  } else {
    rc = h->tsz == 8 ? get_soa<double>(h, h->states, soa) : get_soa<float>(h, h->states, soa);
  }
  if (rc) return rc;
  rc = h->tsz == 8 ? get_soa<double>(h, h->pk_states, pk) : get_soa<float>(h, h->pk_states, pk);
  if (rc) return rc;
  for (int64_t r = 0; r < h->rows; r++) {
    const int64_t o = h->new2old[r];
    if (o < 0) continue;
    double *st = data + offsets[o];
    if (r < h->n_myo) {
      int k2 = 0;
      for (int k = 0; k < nstate_of(h->model); k++) st[k] = (k == vs) ? V[o] : soa[(size_t)k2++ * na + r];
    } else {
      for (int k = 0; k < PK_NSTATE; k++) st[k] = pk[(size_t)k * npa + (r - h->n_myo)];
    }
  }
  return VG_OK;
}

// CG grid: g blocks in all (0 = one per 4 slices up to 1024; a multiple of
// XCDS), split between instances in proportion to their slices (at least
// one, at most one per slice); instance k owns blocks [inst_blk[k],
// inst_blk[k+1]).  Reallocates the partials and drops a captured step.
int set_cg_blocks(vg_handle *h, int cg_blocks) {
  int64_t g = cg_blocks > 0 ? cg_blocks : std::min<int64_t>(1024, (h->n_slices + WAVES - 1) / WAVES);
  g = std::max<int64_t>(XCDS, ((g + XCDS - 1) / XCDS) * XCDS);
  if (g > 65536) return h->fail(VG_ERR_ARG, "cg_blocks above 65536");
  const int K = h->n_inst;
  const std::vector<int64_t> &isl = h->inst_slice_host;
  std::vector<int> inst_blk(K + 1, 0), blk_inst;
  for (int k = 0; k < K; k++) {
    const int64_t ns = isl[k + 1] - isl[k];
    int64_t gk = K == 1 ? g : (g * ns + h->n_slices / 2) / h->n_slices;
    gk = std::max<int64_t>(1, K == 1 ? gk : std::min<int64_t>(gk, ns));
    inst_blk[k + 1] = inst_blk[k] + (int)gk;
    blk_inst.insert(blk_inst.end(), gk, k);
  }
  if (inst_blk[K] > 65536) return h->fail(VG_ERR_ARG, "too many CG blocks");
  if (h->stream) HIPCHK(h, hipStreamSynchronize(h->stream));
  h->cg_blocks = (unsigned)inst_blk[K];
  HIPCHK(h, h->blk_inst.alloc(blk_inst.size() * 4));
  HIPCHK(h, hipMemcpy(h->blk_inst.p, blk_inst.data(), blk_inst.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(h, h->inst_blk.alloc(inst_blk.size() * 4));
  HIPCHK(h, hipMemcpy(h->inst_blk.p, inst_blk.data(), inst_blk.size() * 4, hipMemcpyHostToDevice));
  // Two sets of 3 CG partials, then the true-residual check's 2.
  HIPCHK(h, h->part.alloc((size_t)h->cg_blocks * 8 * 8));
  HIPCHK(h, hipMemset(h->part.p, 0, (size_t)h->cg_blocks * 8 * 8));
  if (h->gexec) {
    HIPCHK(h, hipGraphExecDestroy(h->gexec));
    h->gexec = nullptr;
  }
  return VG_OK;
}

// Device rows of independent cells, rows in state order without the V slot;
// Mahajan keeps its gate rows in T and its accumulator rows in FP64.
int cells_put_rows(vg_cells *c, const std::vector<double> &soa) {
  const int64_t na = c->n_alloc;
  std::vector<double> g, a;
  const std::vector<double> *rows = &soa;
  if (c->model == VG_MODEL_MJ) {
    g.resize((size_t)MJ_NGATE * na);
    a.resize((size_t)MJ_NACC * na);
    for (int k = 0; k < MJ_NGATE; k++)
      std::copy_n(soa.begin() + (size_t)(MJ_GATE[k] - 1) * na, na, g.begin() + (size_t)k * na);
    for (int k = 0; k < MJ_NACC; k++)
      std::copy_n(soa.begin() + (size_t)(MJ_ACC[k] - 1) * na, na, a.begin() + (size_t)k * na);
    if (hipMemcpy(c->acc.p, a.data(), a.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return VG_ERR_HIP;
    rows = &g;
  }
  if (c->prec == VG_PREC_FP64) {
    if (hipMemcpy(c->states.p, rows->data(), rows->size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return VG_ERR_HIP;
  } else {
    std::vector<float> f(rows->begin(), rows->end());
    if (hipMemcpy(c->states.p, f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return VG_ERR_HIP;
  }
  return VG_OK;
}

int cells_get_rows(vg_cells *c, std::vector<double> &soa) {
  const int64_t na = c->n_alloc;
  const size_t nrow = c->model == VG_MODEL_MJ ? (size_t)MJ_NGATE * na : soa.size();
  std::vector<double> rows(nrow);
  if (c->prec == VG_PREC_FP64) {
    if (hipMemcpy(rows.data(), c->states.p, nrow * 8, hipMemcpyDeviceToHost) != hipSuccess)
      return VG_ERR_HIP;
  } else {
    std::vector<float> f(nrow);
    if (hipMemcpy(f.data(), c->states.p, nrow * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return VG_ERR_HIP;
    std::copy(f.begin(), f.end(), rows.begin());
  }
  if (c->model != VG_MODEL_MJ) {
    soa = rows;
    return VG_OK;
  }
  std::vector<double> a((size_t)MJ_NACC * na);
  if (hipMemcpy(a.data(), c->acc.p, a.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return VG_ERR_HIP;
  for (int k = 0; k < MJ_NGATE; k++)
    std::copy_n(rows.begin() + (size_t)k * na, na, soa.begin() + (size_t)(MJ_GATE[k] - 1) * na);
  for (int k = 0; k < MJ_NACC; k++)
    std::copy_n(a.begin() + (size_t)k * na, na, soa.begin() + (size_t)(MJ_ACC[k] - 1) * na);
  return VG_OK;
}

std::vector<int64_t> state_offsets(const vg_handle *h) {
  std::vector<int64_t> off(h->n + 1, 0);
  for (int64_t i = 0; i < h->n; i++) off[i + 1] = off[i] + nstate_of(node_model(h, i));
  return off;
}

}  // namespace

namespace vg {
int handle_fail(vg_handle *h, int code, const std::string &msg) { return h->fail(code, msg); }
}  // namespace vg

extern "C" {

const char *vg_version(void) { return "voomgpu 0.1 (gfx950)"; }

const char *vg_last_error(const vg_handle *h) { return h ? h->err.c_str() : "null handle"; }

int vg_assemble_csr(const vg_mesh *mesh, const vg_model *model, int64_t *n, int64_t *nnz,
                    int64_t *row_ptr, int32_t *col, double *K, double *M, double *CL) {
  if (!mesh || !model || !n || !nnz) return VG_ERR_ARG;
  HostCsr c;
  if (!assemble_csr(*mesh, *model, c).empty()) return VG_ERR_ARG;
  *n = c.n;
  *nnz = (int64_t)c.col.size();
  if (row_ptr) std::copy(c.row_ptr.begin(), c.row_ptr.end(), row_ptr);
  if (col) std::copy(c.col.begin(), c.col.end(), col);
  if (K) std::copy(c.K.begin(), c.K.end(), K);
  if (M) std::copy(c.M.begin(), c.M.end(), M);
  if (CL) std::copy(c.CL.begin(), c.CL.end(), CL);
  return VG_OK;
}

int vg_create(const vg_mesh *mesh, const vg_model *model, vg_handle **out) {
  if (!out) return VG_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<vg_handle> h(new vg_handle());
  if (!mesh || !model) return VG_ERR_ARG;
  if (model->cell_model != VG_MODEL_LR && model->cell_model != VG_MODEL_LR_GATE &&
      model->cell_model != VG_MODEL_TT && model->cell_model != VG_MODEL_MJ &&
      model->cell_model != VG_MODEL_MJ_FAIL) {
    *out = h.release();
    return (*out)->fail(VG_ERR_UNSUPPORTED, "cell model not supported");
  }
  const bool mj_fail = model->cell_model == VG_MODEL_MJ_FAIL;
  if ((model->cell_model == VG_MODEL_MJ || mj_fail) && model->n_param_sets > 0) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "Mahajan takes its constants from the tissue types, not "
                                    "from parameter sets");
  }
  if (model->precision != VG_PREC_FP64 && model->precision != VG_PREC_MIXED) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "unknown precision");
  }
  if (model->cg_x0 != VG_X0_ZERO && model->cg_x0 != VG_X0_PREVIOUS && model->cg_x0 != VG_X0_LAST &&
      model->cg_x0 != VG_X0_EXTRAPOLATE && model->cg_x0 != VG_X0_EXTRAPOLATE_LINEAR) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "unknown cg_x0");
  }
  if (model->row_order != VG_ORDER_AUTO && model->row_order != VG_ORDER_MESH &&
      model->row_order != VG_ORDER_GEOMETRIC) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "unknown row_order");
  }
  if (!(model->dt > 0) || !(model->dt_ionic > 0) || !(model->Cm > 0) || !(model->cg_tol > 0) ||
      model->cg_maxit < 1) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "dt, dt_ionic, Cm, cg_tol and cg_maxit must be positive");
  }
  h->prec = model->precision;
  // The failing-heart set is the Mahajan model with other constants; the
  // gate-table Luo-Rudy is Luo-Rudy with a table.
  h->model = mj_fail ? VG_MODEL_MJ
             : model->cell_model == VG_MODEL_LR_GATE ? VG_MODEL_LR
                                                      : model->cell_model;
  h->mj_fail = mj_fail;
  const bool lr_gate = model->cell_model == VG_MODEL_LR_GATE;
  h->integ = model->integrator;
  h->quirks = model->quirks;
  h->dt = model->dt;
  h->dt_ion = model->dt_ionic;
  h->Cm = model->Cm;
  h->tol = model->cg_tol;
  h->maxit = model->cg_maxit;
  h->guess = model->cg_x0 != VG_X0_ZERO ? 1 : 0;
  h->shared_x = model->cg_x0 == VG_X0_LAST;
  h->extrap = model->cg_x0 == VG_X0_EXTRAPOLATE || model->cg_x0 == VG_X0_EXTRAPOLATE_LINEAR;
  h->extrap_order = model->cg_x0 == VG_X0_EXTRAPOLATE_LINEAR ? 2 : 3;
  h->tsz = h->prec == VG_PREC_FP64 ? 8 : 4;
  if (model->scheme != VG_SCHEME_CN && model->scheme != VG_SCHEME_BE &&
      model->scheme != VG_SCHEME_EXPLICIT) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "unknown scheme");
  }
  if (model->scheme != VG_SCHEME_CN && (model->smoothing > 0.0 || model->n_instances > 1)) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "smoothing and batching need the Crank-Nicolson scheme");
  }
  h->scheme = model->scheme;
  h->dt_diff = model->dt_diffusion > 0.0 ? model->dt_diffusion : 0.1;
  if (h->scheme == VG_SCHEME_EXPLICIT) {
    // Strides of min(dtDiff, dT/2) over dT/2 (EP_ExplicitSolver.cc:35-37, 52-85).
    const double TOLER = 1e-6;
    double solve = 0.5 * h->dt, hs = std::min(h->dt_diff, 0.5 * h->dt);
    while (solve > TOLER) {
      if (h->strides.size() >= 100000) {
        *out = h.release();
        return (*out)->fail(VG_ERR_ARG, "too many explicit strides per half-step");
      }
      h->strides.push_back(hs);
      solve -= hs;
      hs = std::min(hs, solve);
    }
  }
  vg_handle *hp = h.get();
  if (h->scheme != VG_SCHEME_CN) {
    h->sub = SubSteps{};
    h->sub.adaptive = 1;
    h->sub.dT = h->dt;
    h->sub.dt_ion = h->dt_ion;
  } else if (!substeps(h->dt, h->dt_ion, h->sub)) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, "too many ionic sub-steps per step");
  }
  HostCsr csr;
  std::string e = assemble_csr(*mesh, *model, csr);
  if (!e.empty()) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, e);
  }
  // Row order: myocardium first, then the Purkinje segment, each in mesh
  // order.  Batched instances: instance after instance, each padded with
  // node-less rows to a whole number of slices.
  std::vector<int32_t> base;
  std::vector<int64_t> inst_row0(1, 0);
  base.reserve(csr.n);
  h->purkinje.assign(csr.n, 0);
  if (mesh->tissue)
    for (int64_t i = 0; i < csr.n; i++) h->purkinje[i] = (mesh->tissue[i] == 11 || mesh->tissue[i] == 12);
  if (model->n_instances > 1) {
    const int K = model->n_instances;
    std::string bad;
    if (!model->instance) bad = "instance is NULL";
    for (int64_t i = 0; i < csr.n && bad.empty(); i++) {
      if (model->instance[i] < 0 || model->instance[i] >= K) bad = "instance id out of range";
      else if (h->purkinje[i]) bad = "batched instances cannot hold Purkinje nodes";
    }
    for (int64_t e = 0; e < mesh->n_elems && bad.empty(); e++) {
      const int32_t *c = mesh->conn + 4 * e;
      for (int q = 1; q < mesh->elem_nn[e]; q++)
        if (model->instance[c[q]] != model->instance[c[0]]) bad = "an element spans two instances";
    }
    if (!bad.empty()) {
      *out = h.release();
      return (*out)->fail(VG_ERR_ARG, bad);
    }
    std::vector<std::vector<int32_t>> members(K);
    for (int64_t i = 0; i < csr.n; i++) members[model->instance[i]].push_back((int32_t)i);
    for (int k = 0; k < K; k++) {
      base.insert(base.end(), members[k].begin(), members[k].end());
      while (base.size() % SELL_C) base.push_back(-1);
      inst_row0.push_back((int64_t)base.size());
    }
    h->n_inst = K;
    h->n_myo = (int64_t)base.size();
  } else {
    for (int64_t i = 0; i < csr.n; i++)
      if (!h->purkinje[i]) base.push_back((int32_t)i);
    // A far column puts its whole slice on 32-bit columns and its gathers
    // outside the L2 the neighbouring rows are in; the geometric order keeps
    // the band to about one layer of the mesh.
    if (model->row_order != VG_ORDER_MESH) {
      const int64_t far_mesh = model->row_order == VG_ORDER_AUTO ? far_rows(csr, base) : 1;
      if (far_mesh > 0) {
        std::vector<int32_t> geo = base;
        geometric_order(csr, mesh->xyz, geo);
        if (model->row_order == VG_ORDER_GEOMETRIC || far_rows(csr, geo) < far_mesh) {
          base.swap(geo);
          h->row_order = VG_ORDER_GEOMETRIC;
        }
      }
    }
    h->n_myo = (int64_t)base.size();
    for (int64_t i = 0; i < csr.n; i++)
      if (h->purkinje[i]) base.push_back((int32_t)i);
    inst_row0.push_back((int64_t)base.size());
  }
  h->n_pk = (int64_t)base.size() - h->n_myo;
  h->n_pk_alloc = ((h->n_pk + BLK - 1) / BLK) * BLK;
  HostSell sell;
  // Diffusive smoothing adds alpha (4/dT) K to A and subtracts it from B
  // (EP_CrankNicholson.cc:31-36).
  const double kc = h->dt / 4.0 + (model->smoothing > 0.0 ? 4.0 * model->smoothing / h->dt : 0.0);
  if (h->scheme == VG_SCHEME_BE)
    e = build_sell(csr, h->dt / 2.0, h->dt / 2.0, h->Cm, true, model->sell_sigma, base, h->n_myo,
                   sell);
  else
    e = build_sell(csr, kc, 2.0 * kc, h->Cm, false, model->sell_sigma, base, h->n_myo, sell);
  if (!e.empty()) {
    *out = h.release();
    return (*out)->fail(VG_ERR_ARG, e);
  }
  h->n = csr.n;
  h->rows = sell.rows;
  h->nnz = (int64_t)csr.col.size();
  h->n_slices = sell.n_slices;
  h->padded = sell.padded;
  h->n_wide = std::count_if(sell.wide.begin(), sell.wide.end(), [](int64_t w) { return w >= 0; });
  h->n_uniform = std::count_if(sell.ustart.begin(), sell.ustart.end(), [](int64_t u) { return u >= 0; });
  h->n_alloc = ((h->rows + BLK - 1) / BLK) * BLK;
  h->nt_vals = nt_vals_auto(h.get());
  h->new2old = sell.new2old;
  h->old2new = sell.old2new;
  h->CL = csr.CL;
  h->xyz.assign(mesh->xyz, mesh->xyz + 3 * mesh->n_nodes);
  h->istim_host.assign(h->n, 0.0);

  int rc;
  {
    hipError_t he = hipGetDevice(&h->device);
    if (he != hipSuccess) {
      *out = h.release();
      return (*out)->hipfail(he, "hipGetDevice");
    }
    he = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
      *out = h.release();
      return (*out)->hipfail(he, "hipStreamCreate");
    }
    he = scratch_of(h->device, &h->ecg_scratch);
    if (he != hipSuccess) {
      *out = h.release();
      return (*out)->hipfail(he, "hipGetSymbolAddress(scratch)");
    }
  }
  *out = h.release();
  vg_handle *H_ = *out;
  (void)hp;
  const size_t T = H_->tsz;
  const int64_t na = H_->n_alloc;
  // Operator.
  HIPCHK(H_, H_->off.alloc(sell.offset.size() * 8));
  HIPCHK(H_, hipMemcpy(H_->off.p, sell.offset.data(), sell.offset.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(H_, H_->width.alloc(sell.width.size() * 4));
  HIPCHK(H_, hipMemcpy(H_->width.p, sell.width.data(), sell.width.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(H_, H_->wide.alloc(sell.wide.size() * 8));
  HIPCHK(H_, hipMemcpy(H_->wide.p, sell.wide.data(), sell.wide.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(H_, H_->col16.alloc(sell.col16.size() * 2));
  HIPCHK(H_, hipMemcpy(H_->col16.p, sell.col16.data(), sell.col16.size() * 2, hipMemcpyHostToDevice));
  HIPCHK(H_, H_->ustart.alloc(sell.ustart.size() * 8));
  HIPCHK(H_, hipMemcpy(H_->ustart.p, sell.ustart.data(), sell.ustart.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(H_, H_->uoff.alloc(std::max<size_t>(1, sell.uoff.size()) * 4));
  if (!sell.uoff.empty())
    HIPCHK(H_, hipMemcpy(H_->uoff.p, sell.uoff.data(), sell.uoff.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(H_, H_->col32.alloc(sell.col32.size() * 4));
  if (!sell.col32.empty())
    HIPCHK(H_, hipMemcpy(H_->col32.p, sell.col32.data(), sell.col32.size() * 4, hipMemcpyHostToDevice));
  std::vector<double> sinv_pad(na, 0.0), diag_pad(na, 0.0);
  std::copy(sell.sinv.begin(), sell.sinv.begin() + std::min<int64_t>(na, sell.sinv.size()), sinv_pad.begin());
  std::copy(sell.diag.begin(), sell.diag.begin() + std::min<int64_t>(na, sell.diag.size()), diag_pad.begin());
  if (T == 8) {
    if ((rc = upload_raw<double>(H_, H_->Ahat, sell.A_hat))) return rc;
    if ((rc = upload_raw<double>(H_, H_->Krhs, H_->scheme == VG_SCHEME_EXPLICIT ? sell.K_exp : sell.K_rhs)))
      return rc;
    if ((rc = upload_raw<double>(H_, H_->sinv, sinv_pad))) return rc;
    if ((rc = upload_raw<double>(H_, H_->diag, diag_pad))) return rc;
  } else {
    if ((rc = upload_raw<float>(H_, H_->Ahat, sell.A_hat))) return rc;
    if ((rc = upload_raw<float>(H_, H_->Krhs, H_->scheme == VG_SCHEME_EXPLICIT ? sell.K_exp : sell.K_rhs)))
      return rc;
    if ((rc = upload_raw<float>(H_, H_->sinv, sinv_pad))) return rc;
    if ((rc = upload_raw<float>(H_, H_->diag, diag_pad))) return rc;
  }
  if ((rc = setup_pk_presolve(H_, sell))) return rc;
  // Grids: the CG block tables (set_cg_blocks) and the node-wise grid.
  {
    H_->inst_slice_host.assign(H_->n_inst + 1, 0);
    for (int k = 0; k <= H_->n_inst; k++) H_->inst_slice_host[k] = inst_row0[k] / SELL_C;
    if (H_->n_inst == 1) H_->inst_slice_host[1] = H_->n_slices;
    if ((rc = set_cg_blocks(H_, model->cg_blocks))) return rc;
    H_->node_blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, na / BLK));
    HIPCHK(H_, H_->inst_slice.alloc(H_->inst_slice_host.size() * 8));
    HIPCHK(H_, hipMemcpy(H_->inst_slice.p, H_->inst_slice_host.data(),
                         H_->inst_slice_host.size() * 8, hipMemcpyHostToDevice));
    if (H_->n_inst > 1) {
      std::vector<int> si(H_->n_slices);
      for (int k = 0; k < H_->n_inst; k++)
        for (int64_t q = H_->inst_slice_host[k]; q < H_->inst_slice_host[k + 1]; q++) si[q] = k;
      HIPCHK(H_, H_->slice_inst.alloc(si.size() * 4));
      HIPCHK(H_, hipMemcpy(H_->slice_inst.p, si.data(), si.size() * 4, hipMemcpyHostToDevice));
    }
  }
  // Vectors.
  for (DevBuf *b : {&H_->Va, &H_->Vb, &H_->r, &H_->x, &H_->istim, &H_->at, &H_->x2}) {
    HIPCHK(H_, b->alloc(na * T));
    HIPCHK(H_, hipMemset(b->p, 0, na * T));
  }
  for (DevBuf *b : {&H_->q0, &H_->q1}) {
    HIPCHK(H_, b->alloc(4 * na * T));
    HIPCHK(H_, hipMemset(b->p, 0, 4 * na * T));
  }
  if (H_->extrap) {
    HIPCHK(H_, H_->xhist.alloc(4 * na * T));
    HIPCHK(H_, hipMemset(H_->xhist.p, 0, 4 * na * T));
  }
  if (H_->model == VG_MODEL_MJ) {
    HIPCHK(H_, H_->states.alloc((size_t)MJ_NGATE * na * T));
    HIPCHK(H_, H_->acc.alloc((size_t)MJ_NACC * na * 8));
  } else {
    HIPCHK(H_, H_->states.alloc((size_t)ndev_of(H_->model) * na * T));
  }
  if (H_->n_pk > 0) {
    HIPCHK(H_, H_->pk_states.alloc((size_t)PK_NSTATE * H_->n_pk_alloc * T));
    HIPCHK(H_, hipStreamCreateWithFlags(&H_->stream_pk, hipStreamNonBlocking));
    HIPCHK(H_, hipEventCreateWithFlags(&H_->ev_fork, hipEventDisableTiming));
    HIPCHK(H_, hipEventCreateWithFlags(&H_->ev_join, hipEventDisableTiming));
  }
  // CG scratch: two sets of 3 FP64 partials per block, the carried scalars
  // and the result record.
  HIPCHK(H_, H_->carry.alloc(2 * sizeof(CgCarry) * H_->n_inst));
  HIPCHK(H_, hipMemset(H_->carry.p, 0, 2 * sizeof(CgCarry) * H_->n_inst));
  HIPCHK(H_, H_->res.alloc(sizeof(CgResult) * H_->n_inst));
  HIPCHK(H_, hipMemset(H_->res.p, 0, sizeof(CgResult) * H_->n_inst));
  HIPCHK(H_, H_->tres.alloc(4 * sizeof(double)));
  HIPCHK(H_, hipMemset(H_->tres.p, 0, 4 * sizeof(double)));
  H_->tres_every = std::max(0, model->true_residual_every);
  HIPCHK(H_, H_->newton.alloc(sizeof(unsigned long long)));
  HIPCHK(H_, hipMemset(H_->newton.p, 0, sizeof(unsigned long long)));
  HIPCHK(H_, H_->step_ctr.alloc(sizeof(int64_t)));
  HIPCHK(H_, hipMemset(H_->step_ctr.p, 0, sizeof(int64_t)));
  HIPCHK(H_, H_->totals.alloc(sizeof(CgTotals)));
  HIPCHK(H_, hipMemset(H_->totals.p, 0, sizeof(CgTotals)));
  HIPCHK(H_, hipHostMalloc((void **)&H_->h_tot, sizeof(CgTotals), hipHostMallocDefault));
  H_->graph = model->graph ? 1 : 0;
  HIPCHK(H_, hipHostMalloc((void **)&H_->h_res, sizeof(CgResult) * H_->n_inst, hipHostMallocDefault));
  HIPCHK(H_, hipHostMalloc((void **)&H_->h_pub, sizeof(CgResult) * H_->n_inst,
                           hipHostMallocCoherent | hipHostMallocMapped));
  std::memset(H_->h_pub, 0, sizeof(CgResult) * H_->n_inst);
  HIPCHK(H_, hipHostGetDevicePointer((void **)&H_->d_pub, H_->h_pub, 0));
  HIPCHK(H_, hipHostMalloc((void **)&H_->h_trunc, sizeof(unsigned long long),
                           hipHostMallocCoherent | hipHostMallocMapped));
  *H_->h_trunc = 0;
  HIPCHK(H_, hipHostGetDevicePointer((void **)&H_->d_trunc, H_->h_trunc, 0));
  if (lr_gate) {
    const std::vector<double> tab = lr_gate_table();
    if (T == 8) {
      if ((rc = upload_raw<double>(H_, H_->lr_table, tab))) return rc;
    } else {
      if ((rc = upload_raw<float>(H_, H_->lr_table, tab))) return rc;
    }
  }
  // Parameter sets (tT scalings; LR uses gNa and gCaL as its Gna and Gsi;
  // Mahajan: one set per tissue type holding G_tof and G_Ks).
  if (H_->model == VG_MODEL_MJ) {
    const std::vector<double> sc = mj_scale_rows(H_->mj_fail);
    H_->nsets = 10;
    if (T == 8) {
      if ((rc = upload_raw<double>(H_, H_->scales, sc))) return rc;
    } else {
      if ((rc = upload_raw<float>(H_, H_->scales, sc))) return rc;
    }
    std::vector<uint16_t> pi(na, 0);
    for (int64_t r = 0; r < H_->rows; r++)
      pi[r] = H_->new2old[r] >= 0 ? (uint16_t)mj_row(mesh->tissue, H_->new2old[r]) : 0;
    HIPCHK(H_, H_->pidx.alloc(na * 2));
    HIPCHK(H_, hipMemcpy(H_->pidx.p, pi.data(), na * 2, hipMemcpyHostToDevice));
  } else {
    int nsets = model->n_param_sets > 0 ? model->n_param_sets : 1;
    std::vector<double> sc((size_t)nsets * VG_TT_NSCALE, 1.0);
    if (model->n_param_sets > 0) {
      if (!model->param_scales) return H_->fail(VG_ERR_ARG, "param_scales is NULL");
      std::copy(model->param_scales, model->param_scales + sc.size(), sc.begin());
    }
    H_->nsets = nsets;
    if (T == 8) {
      if ((rc = upload_raw<double>(H_, H_->scales, sc))) return rc;
    } else {
      if ((rc = upload_raw<float>(H_, H_->scales, sc))) return rc;
    }
    if (model->param_index) {
      std::vector<uint16_t> pi(na, 0);
      for (int64_t r = 0; r < H_->rows; r++) {
        if (H_->new2old[r] < 0) continue;
        const uint16_t v = model->param_index[H_->new2old[r]];
        if (v >= nsets) return H_->fail(VG_ERR_ARG, "param_index out of range");
        pi[r] = v;
      }
      HIPCHK(H_, H_->pidx.alloc(na * 2));
      HIPCHK(H_, hipMemcpy(H_->pidx.p, pi.data(), na * 2, hipMemcpyHostToDevice));
    }
  }
  // Initial states and voltage: each model's own initial table, with V from
  // it (LR: its rest state) unless the driver's V0 = 0 is requested for the
  // body's model; Purkinje nodes start at their own table's V
  // (Purkinje.h:44-50, getVoltage :79-81).
  {
    const int ns = nstate_of(H_->model);
    std::vector<double> one(ns);
    double v0;
    if (H_->model == VG_MODEL_LR) {
      if (H_->quirks & VG_QUIRK_V0_ZERO) {
        for (int k = 0; k < ns; k++) one[k] = LR_INIT[k];
        v0 = 0.0;
      } else {
        for (int k = 0; k < 7; k++) one[k] = LR_REST[k];
        one[7] = LR_REST_V;
        v0 = LR_REST_V;
      }
    } else if (H_->model == VG_MODEL_MJ) {
      for (int k = 0; k < ns; k++) one[k] = MJ_INIT[0][k];
      v0 = MJ_INIT[0][0];
    } else {
      for (int k = 0; k < ns; k++) one[k] = TT_INIT[k];
      v0 = (H_->quirks & VG_QUIRK_V0_ZERO) ? 0.0 : TT_INIT[0];
    }
    const std::vector<int64_t> off = state_offsets(H_);
    std::vector<double> st(off[H_->n]);
    std::vector<double> V0(H_->n);
    for (int64_t i = 0; i < H_->n; i++) {
      if (H_->purkinje[i]) {
        std::copy(PK_INIT, PK_INIT + PK_NSTATE, st.begin() + off[i]);
        V0[i] = PK_INIT[0];
      } else if (H_->model == VG_MODEL_MJ) {
        // Mahajan starts from its type's table with V = its state 0 (the
        // driver's getInitialVoltage, CardiacBody.h:299-306).
        const double *row = MJ_INIT[mj_type(mesh->tissue, i)];
        std::copy(row, row + MJ_NSTATE, st.begin() + off[i]);
        V0[i] = row[0];
      } else {
        std::copy(one.begin(), one.end(), st.begin() + off[i]);
        V0[i] = v0;
      }
    }
    if ((rc = set_states_impl(H_, off.data(), st.data()))) return rc;
    if ((rc = upload(H_, H_->Va.p, V0.data(), v0))) return rc;
    if ((rc = upload(H_, H_->Vb.p, V0.data(), v0))) return rc;
  }
  H_->stim_dirty = true;
  return VG_OK;
}

void vg_destroy(vg_handle *h) { delete h; }

int vg_get_stream(vg_handle *h, void **s) {
  if (!h || !s) return VG_ERR_ARG;
  *s = (void *)h->stream;
  return VG_OK;
}

int vg_set_stimulus_nodes(vg_handle *h, const int32_t *nodes, int64_t n, double istim) {
  if (!h || (n > 0 && !nodes)) return VG_ERR_ARG;
  for (int64_t k = 0; k < n; k++) {
    if (nodes[k] < 0 || nodes[k] >= h->n) return h->fail(VG_ERR_ARG, "stimulus node out of range");
    h->istim_host[nodes[k]] = istim;  // CardiacBody.cc:308-318
  }
  h->stim_dirty = true;
  return VG_OK;
}

int vg_set_stimulus_sphere(vg_handle *h, const double c[3], double radius, double istim) {
  if (!h || !c) return VG_ERR_ARG;
  // CardiacBody.cc:283-305: Purkinje nodes are skipped; others inside get istim,
  // outside get 0.
  for (int64_t i = 0; i < h->n; i++) {
    if (h->purkinje[i]) continue;
    double d2 = 0;
    for (int k = 0; k < 3; k++) d2 += (c[k] - h->xyz[3 * i + k]) * (c[k] - h->xyz[3 * i + k]);
    h->istim_host[i] = (std::sqrt(d2) <= radius) ? istim : 0.0;
  }
  h->stim_dirty = true;
  return VG_OK;
}

int vg_set_stimulus_axis(vg_handle *h, int axis, double limit, double istim) {
  if (!h || axis < 0 || axis > 2) return VG_ERR_ARG;
  // CardiacBody.cc:243-262: nodes with x[axis] <= limit get istim, others unchanged.
  for (int64_t i = 0; i < h->n; i++)
    if (h->xyz[3 * i + axis] <= limit) h->istim_host[i] = istim;
  h->stim_dirty = true;
  return VG_OK;
}

int vg_set_stimulus_plane(vg_handle *h, const double o[3], const double nrm[3], double istim) {
  if (!h || !o || !nrm) return VG_ERR_ARG;
  const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  if (!(len > 0)) return h->fail(VG_ERR_ARG, "zero plane normal");
  // CardiacBody.cc:265-281: below the plane -> istim, above -> 0, Purkinje
  // skipped; the normal is normalised first, then dotted with x - origin.
  const double nh[3] = {nrm[0] / len, nrm[1] / len, nrm[2] / len};
  for (int64_t i = 0; i < h->n; i++) {
    if (h->purkinje[i]) continue;
    double d = 0;
    for (int k = 0; k < 3; k++) d += (h->xyz[3 * i + k] - o[k]) * nh[k];
    h->istim_host[i] = (d < 0.) ? istim : 0.0;
  }
  h->stim_dirty = true;
  return VG_OK;
}

int vg_clear_stimulus(vg_handle *h) {
  if (!h) return VG_ERR_ARG;
  std::fill(h->istim_host.begin(), h->istim_host.end(), 0.0);
  h->stim_dirty = true;
  return VG_OK;
}

int vg_step(vg_handle *h, int nsteps) {
  if (!h || nsteps < 0) return VG_ERR_ARG;
  const bool graph = h->graph && !h->profiling;
  for (int k = 0; k < nsteps; k++) {
    int rc;
    if (h->scheme == VG_SCHEME_EXPLICIT)
      rc = h->tsz == 8 ? step_explicit<double>(h) : step_explicit<float>(h);
    else if (graph) rc = h->tsz == 8 ? step_graph<double>(h) : step_graph<float>(h);
    else rc = h->tsz == 8 ? step_t<double>(h) : step_t<float>(h);
    if (rc) return rc;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hipfail(e, "kernel launch");
  return VG_OK;
}

int vg_set_cg_blocks(vg_handle *h, int cg_blocks) {
  if (!h || cg_blocks < 0) return VG_ERR_ARG;
  return set_cg_blocks(h, cg_blocks);
}

int vg_get_cg_blocks(vg_handle *h, int *cg_blocks) {
  if (!h || !cg_blocks) return VG_ERR_ARG;
  *cg_blocks = (int)h->cg_blocks;
  return VG_OK;
}

int vg_set_graph(vg_handle *h, int enable) {
  if (!h) return VG_ERR_ARG;
  // Queued replays end before a host-driven solve resets the published record.
  if ((enable ? 1 : 0) != h->graph) HIPCHK(h, hipStreamSynchronize(h->stream));
  h->graph = enable ? 1 : 0;
  return VG_OK;
}

int vg_set_value_streaming(vg_handle *h, int mode) {
  if (!h || mode < -1 || mode > 1) return VG_ERR_ARG;
  const bool nt = mode < 0 ? nt_vals_auto(h) : mode == 1;
  if (nt == h->nt_vals) return VG_OK;
  // Launches already queued keep their kernels; a captured step holds the
  // old ones, so it is captured again.
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->nt_vals = nt;
  if (h->gexec) {
    HIPCHK(h, hipGraphExecDestroy(h->gexec));
    h->gexec = nullptr;
  }
  return VG_OK;
}

int vg_get_value_streaming(vg_handle *h, int *on) {
  if (!h || !on) return VG_ERR_ARG;
  *on = h->nt_vals ? 1 : 0;
  return VG_OK;
}

int vg_set_true_residual_check(vg_handle *h, int every) {
  if (!h || every < 0) return VG_ERR_ARG;
  h->tres_every = every;
  return VG_OK;
}

int vg_synchronize(vg_handle *h) {
  if (!h) return VG_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return VG_OK;
}

int vg_get_voltage(vg_handle *h, double *V) {
  if (!h || !V) return VG_ERR_ARG;
  return download(h, h->Va.p, V);
}

int vg_set_voltage(vg_handle *h, const double *V) {
  if (!h || !V) return VG_ERR_ARG;
  int rc = upload(h, h->Va.p, V, V[0]);
  if (rc) return rc;
  return upload(h, h->Vb.p, V, V[0]);
}

int vg_get_states(vg_handle *h, int *nvar, double *st) {
  if (!h || !nvar) return VG_ERR_ARG;
  const int ns = nstate_of(h->model);
  *nvar = ns;
  if (h->n_pk > 0) return h->fail(VG_ERR_ARG, "mixed cell models: use vg_get_states_ragged");
  if (!st) return VG_OK;
  std::vector<int64_t> off(h->n + 1);
  for (int64_t i = 0; i <= h->n; i++) off[i] = i * ns;
  std::vector<double> V(h->n);
  int rc = download(h, h->Va.p, V.data());
  if (rc) return rc;
  return get_states_impl(h, off.data(), st, V.data());
}

int vg_set_states(vg_handle *h, int nvar, const double *st) {
  if (!h || !st) return VG_ERR_ARG;
  if (nvar != nstate_of(h->model)) return h->fail(VG_ERR_ARG, "state count mismatch");
  if (h->n_pk > 0) return h->fail(VG_ERR_ARG, "mixed cell models: use vg_set_states_ragged");
  std::vector<int64_t> off(h->n + 1);
  for (int64_t i = 0; i <= h->n; i++) off[i] = i * nvar;
  return set_states_impl(h, off.data(), st);
}

int vg_get_states_ragged(vg_handle *h, int32_t *num_var, double *data) {
  if (!h || !num_var) return VG_ERR_ARG;
  const std::vector<int64_t> off = state_offsets(h);
  for (int64_t i = 0; i < h->n; i++) num_var[i] = (int32_t)(off[i + 1] - off[i]);
  if (!data) return VG_OK;
  std::vector<double> V(h->n);
  int rc = download(h, h->Va.p, V.data());
  if (rc) return rc;
  return get_states_impl(h, off.data(), data, V.data());
}

int vg_set_states_ragged(vg_handle *h, const int32_t *num_var, const double *data) {
  if (!h || !num_var || !data) return VG_ERR_ARG;
  const std::vector<int64_t> off = state_offsets(h);
  for (int64_t i = 0; i < h->n; i++)
    if (num_var[i] != off[i + 1] - off[i]) return h->fail(VG_ERR_ARG, "state count mismatch at node " + std::to_string(i));
  return set_states_impl(h, off.data(), data);
}

int vg_get_lumped_capacitance(vg_handle *h, double *CL) {
  if (!h || !CL) return VG_ERR_ARG;
  std::copy(h->CL.begin(), h->CL.end(), CL);
  return VG_OK;
}

int vg_activation_enable(vg_handle *h, double thr) {
  if (!h) return VG_ERR_ARG;
  std::vector<double> m1(h->n, -1.0);
  int rc = upload(h, h->at.p, m1.data(), -1.0);
  if (rc) return rc;
  h->at_enabled = true;
  h->at_thr = thr;
  return VG_OK;
}

int vg_get_activation(vg_handle *h, double *t) {
  if (!h || !t) return VG_ERR_ARG;
  if (!h->at_enabled) return h->fail(VG_ERR_ARG, "activation tracking not enabled");
  return download(h, h->at.p, t);
}

}  // extern "C"

// The ECG lead field needs the element list, which the handle does not keep,
// so the mesh is passed again.
extern "C" int vg_ecg_setup(vg_handle *h, const vg_mesh *mesh, const double *el, int ne,
                            const double D[3], double radius) {
  if (!h || !mesh || !el || !D || ne < 1 || ne > MAX_ECG) return VG_ERR_ARG;
  if (mesh->n_nodes != h->n) return h->fail(VG_ERR_ARG, "mesh does not match the handle");
  std::vector<double> Hm;
  std::string e = build_ecg_matrix(*mesh, h->quirks, el, ne, D, radius, Hm);
  if (!e.empty()) return h->fail(VG_ERR_ARG, e);
  const int64_t na = h->n_alloc;
  std::vector<double> Hp((size_t)ne * na, 0.0);
  for (int k = 0; k < ne; k++)
    for (int64_t r = 0; r < h->rows; r++)
      if (h->new2old[r] >= 0) Hp[(size_t)k * na + r] = Hm[(size_t)k * h->n + h->new2old[r]];
  int rc = h->tsz == 8 ? upload_raw<double>(h, h->H, Hp) : upload_raw<float>(h, h->H, Hp);
  if (rc) return rc;
  if (!h->ecg_out.p) HIPCHK(h, h->ecg_out.alloc(MAX_ECG * 8));
  if (!h->ecg_part.p) HIPCHK(h, h->ecg_part.alloc((size_t)MAX_ECG * h->node_blocks * 8));
  h->ne = ne;
  return VG_OK;
}

extern "C" {

int vg_ecg(vg_handle *h, double *out) {
  if (!h || !out) return VG_ERR_ARG;
  if (h->ne == 0) return h->fail(VG_ERR_ARG, "ECG not set up");
  double *part = h->ecg_part.as<double>();
  if (h->tsz == 8)
    hipLaunchKernelGGL(k_ecg<double>, dim3(h->node_blocks), dim3(BLK), 0, h->stream,
                       h->Va.as<double>(), h->rows, h->H.as<double>(), h->ne, h->n_alloc, part);
  else
    hipLaunchKernelGGL(k_ecg<float>, dim3(h->node_blocks), dim3(BLK), 0, h->stream,
                       h->Va.as<float>(), h->rows, h->H.as<float>(), h->ne, h->n_alloc, part);
  launch_ecg_finish(h, h->ecg_out.as<double>());
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, h->ecg_out.p, h->ne * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return VG_OK;
}

int vg_ecg_device(vg_handle *h, double *d_out) {
  if (!h || !d_out) return VG_ERR_ARG;
  if (h->ne == 0) return h->fail(VG_ERR_ARG, "ECG not set up");
  if (h->ne > MAX_ECG || h->node_blocks > (unsigned)ECG_SCRATCH_BLOCKS || !h->ecg_scratch)
    return h->fail(VG_ERR_ARG, "vg_ecg_device: leads or grid beyond its scratch (16 x 1024)");
  // Every block writes all MAX_ECG of its slots, and the finishing block
  // reads only the node_blocks slots of this launch: nothing is cleared.
  double *part = h->ecg_scratch;
  if (h->tsz == 8)
    hipLaunchKernelGGL(k_ecg<double>, dim3(h->node_blocks), dim3(BLK), 0, h->stream,
                       h->Va.as<double>(), h->rows, h->H.as<double>(), h->ne, h->n_alloc, part);
  else
    hipLaunchKernelGGL(k_ecg<float>, dim3(h->node_blocks), dim3(BLK), 0, h->stream,
                       h->Va.as<float>(), h->rows, h->H.as<float>(), h->ne, h->n_alloc, part);
  hipLaunchKernelGGL(k_ecg_finish, dim3(1), dim3(BLK), 0, h->stream, part, (int)h->node_blocks,
                     h->ne, d_out);
  HIPCHK(h, hipGetLastError());
  return VG_OK;
}

int vg_ecg_record(vg_handle *h, int every, int max_frames) {
  if (!h || every < 1 || max_frames < 1) return VG_ERR_ARG;
  if (h->ne == 0) return h->fail(VG_ERR_ARG, "ECG not set up");
  HIPCHK(h, h->ecg_frames.alloc((size_t)max_frames * h->ne * 8));
  h->ecg_every = every;
  h->ecg_max = max_frames;
  h->ecg_count = 0;
  return VG_OK;
}

int vg_ecg_frames(vg_handle *h, double *out, int *n_frames) {
  if (!h || !n_frames) return VG_ERR_ARG;
  *n_frames = h->ecg_count;
  if (!out || h->ecg_count == 0) return VG_OK;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(out, h->ecg_frames.p, (size_t)h->ecg_count * h->ne * 8, hipMemcpyDeviceToHost));
  return VG_OK;
}

int vg_profile_enable(vg_handle *h, int enable) {
  if (!h) return VG_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  drain_profile(h);
  for (auto &c : h->prof) {
    c.total_ms = 0;
    c.launches = 0;
  }
  h->profiling = enable != 0;
  return VG_OK;
}

int vg_profile_get(vg_handle *h, int kc, double *total_ms, int64_t *launches) {
  if (!h || kc < 0 || kc >= VG_KC_COUNT) return VG_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  drain_profile(h);
  if (total_ms) *total_ms = h->prof[kc].total_ms;
  if (launches) *launches = h->prof[kc].launches;
  return VG_OK;
}

int vg_surr_collect(vg_handle *h, vg_surr *s, int nsteps, int every, int node_stride) {
  if (!h || !s) return VG_ERR_ARG;
  if (h->model != VG_MODEL_TT) return h->fail(VG_ERR_ARG, "surrogate data needs a ten Tusscher body");
  if (nsteps < 0 || every < 1 || node_stride < 1) return h->fail(VG_ERR_ARG, "bad collect arguments");
  float *X = nullptr, *Y = nullptr;
  int64_t count = 0, cap = 0;
  if (!surr_dataset(s, &X, &Y, &count, &cap))
    return h->fail(VG_ERR_ARG, "surrogate has no sample storage (vg_surr_reserve)");
  const int64_t per = (h->n_myo + node_stride - 1) / node_stride;
  for (int q = 0; q < nsteps; q++) {
    int rc = upload_stim(h);
    if (rc) return rc;
    // Steps with a stimulus are not recorded: the network has no stimulus input.
    const bool rec = (q % every) == 0 && !h->stim_active && count < cap && h->n_myo > 0;
    if (rec) {
      h->rec_x = X;
      h->rec_y = Y;
      h->rec_base = count;
      h->rec_cap = cap;
      h->rec_stride = node_stride;
    }
    rc = h->tsz == 8 ? step_t<double>(h) : step_t<float>(h);
    h->rec_x = h->rec_y = nullptr;
    if (rc) return rc;
    if (rec) count = std::min(cap, count + per);
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  surr_set_count(s, count);
  return VG_OK;
}

int vg_surr_attach(vg_handle *h, vg_surr *s) {
  if (!h) return VG_ERR_ARG;
  if (s && h->model != VG_MODEL_TT)
    return h->fail(VG_ERR_ARG, "the learned ionic step is a ten Tusscher model");
  if (s) {
    SurrDevice d{};
    if (!surr_device(s, &d)) return h->fail(VG_ERR_ARG, "surrogate has no fitted scaling");
    int cus = 0;
    HIPCHK(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    const int64_t tiles = (h->n_myo + 31) / 32;
    h->surr_grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cus, (tiles + SWAVES - 1) / SWAVES));
    HIPCHK(h, h->surr_fb.alloc((size_t)h->n_alloc));
    HIPCHK(h, h->surr_use.alloc((size_t)h->surr_grid * SWAVES * 2 * 8));
    HIPCHK(h, hipMemsetAsync(h->surr_use.p, 0, (size_t)h->surr_grid * SWAVES * 2 * 8, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->sdev = d;
  }
  h->surr = s;
  // The captured step holds the ionic launches: capture again.
  if (h->gexec) {
    HIPCHK(h, hipGraphExecDestroy(h->gexec));
    h->gexec = nullptr;
  }
  return VG_OK;
}

int vg_surr_usage(vg_handle *h, int64_t *network, int64_t *fallback) {
  if (!h || !network || !fallback) return VG_ERR_ARG;
  *network = *fallback = 0;
  if (!h->surr) return VG_OK;
  std::vector<unsigned long long> u((size_t)h->surr_grid * SWAVES * 2);
  HIPCHK(h, hipMemcpyAsync(u.data(), h->surr_use.p, u.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < u.size(); k += 2) {
    *network += (int64_t)u[k];
    *fallback += (int64_t)u[k + 1];
  }
  return VG_OK;
}

int vg_stats_get(vg_handle *h, vg_stats *o) {
  if (!h || !o) return VG_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  unsigned long long nw = 0;
  HIPCHK(h, hipMemcpy(&nw, h->newton.p, sizeof(nw), hipMemcpyDeviceToHost));
  CgTotals tot;
  HIPCHK(h, hipMemcpy(&tot, h->totals.p, sizeof(CgTotals), hipMemcpyDeviceToHost));
  // Host-driven solves are counted on the host, captured ones on the device.
  o->steps = h->steps;
  o->cg_iterations = h->cg_iters + (int64_t)tot.iters;
  o->cg_solves = h->cg_solves + (int64_t)tot.solves;
  o->cg_not_converged = h->cg_notconv + (int64_t)tot.notconv;
  o->newton_iterations = (int64_t)nw;
  o->n_nodes = h->n;
  o->nnz = h->nnz;
  o->sell_padded_nnz = h->padded;
  o->last_cg_relres = h->last_relres;
  o->sim_time = h->t;
  o->sell_wide_slices = h->n_wide;
  o->sell_shared_offset_slices = h->n_uniform;
  o->row_order = h->row_order;
  o->cg_truncated = (int64_t)tot.truncated;
  double tr[4];
  HIPCHK(h, hipMemcpy(tr, h->tres.p, sizeof(tr), hipMemcpyDeviceToHost));
  o->max_true_relres = tr[0];
  o->true_residual_checks = (int64_t)tr[1];
  o->last_true_relres = tr[2];
  return VG_OK;
}

// ---- independent cells -------------------------------------------------------
int vg_cells_create(int model, int prec, int integ, int64_t n, vg_cells **out) {
  if (!out || n < 1) return VG_ERR_ARG;
  if (model != VG_MODEL_LR && model != VG_MODEL_LR_GATE && model != VG_MODEL_TT &&
      model != VG_MODEL_PURKINJE && model != VG_MODEL_MJ && model != VG_MODEL_MJ_FAIL)
    return VG_ERR_UNSUPPORTED;
  const bool mj_fail = model == VG_MODEL_MJ_FAIL, lr_gate = model == VG_MODEL_LR_GATE;
  if (mj_fail) model = VG_MODEL_MJ;
  if (lr_gate) model = VG_MODEL_LR;
  if (prec != VG_PREC_FP64 && prec != VG_PREC_MIXED) return VG_ERR_ARG;
  std::unique_ptr<vg_cells> c(new vg_cells());
  c->prec = prec;
  c->model = model;
  c->mj_fail = mj_fail;
  if (lr_gate) {
    const std::vector<double> tab = lr_gate_table();
    const size_t bytes = tab.size() * (prec == VG_PREC_FP64 ? 8 : 4);
    if (c->lr_table.alloc(bytes) != hipSuccess) return VG_ERR_HIP;
    if (prec == VG_PREC_FP64) {
      if (hipMemcpy(c->lr_table.p, tab.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
    } else {
      const std::vector<float> f(tab.begin(), tab.end());
      if (hipMemcpy(c->lr_table.p, f.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
    }
  }
  c->integ = integ;
  c->n = n;
  c->n_alloc = ((n + BLK - 1) / BLK) * BLK;
  const size_t T = prec == VG_PREC_FP64 ? 8 : 4;
  const int ns = nstate_of(model);
  if (model == VG_MODEL_MJ) {
    if (c->states.alloc((size_t)MJ_NGATE * c->n_alloc * T) != hipSuccess) return VG_ERR_HIP;
    if (c->acc.alloc((size_t)MJ_NACC * c->n_alloc * 8) != hipSuccess) return VG_ERR_HIP;
  } else if (c->states.alloc((size_t)ndev_of(model) * c->n_alloc * T) != hipSuccess) {
    return VG_ERR_HIP;
  }
  if (c->V.alloc(n * 8) != hipSuccess || c->stim.alloc(n * 8) != hipSuccess ||
      c->dvdt.alloc(n * 8) != hipSuccess)
    return VG_ERR_HIP;
  // Mahajan: one scaling set per tissue type carrying G_tof, G_Ks.
  std::vector<double> one =
      model == VG_MODEL_MJ ? mj_scale_rows(mj_fail) : std::vector<double>(VG_TT_NSCALE, 1.0);
  if (T == 8) {
    if (c->scales.alloc(one.size() * 8) != hipSuccess) return VG_ERR_HIP;
    if (hipMemcpy(c->scales.p, one.data(), one.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  } else {
    std::vector<float> f(one.begin(), one.end());
    if (c->scales.alloc(f.size() * 4) != hipSuccess) return VG_ERR_HIP;
    if (hipMemcpy(c->scales.p, f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  }
  // Initial states: the model's initial table (LR: reference table).
  std::vector<double> st((size_t)n * ns);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < ns; k++) st[i * ns + k] = init_table(model)[k];
  *out = c.release();
  return vg_cells_set_states(*out, ns, st.data());
}

void vg_cells_destroy(vg_cells *c) { delete c; }

int vg_cells_set_quirks(vg_cells *c, unsigned quirks) {
  if (!c) return VG_ERR_ARG;
  c->newton_jt = (quirks & VG_QUIRK_NEWTON_JT) ? 1 : 0;
  c->mj_int_abs = (quirks & VG_QUIRK_MJ_INT_ABS) ? 1 : 0;
  return VG_OK;
}

int vg_cells_set_scales(vg_cells *c, int nsets, const double *sc, const uint16_t *pidx) {
  if (!c) return VG_ERR_ARG;
  if (c->model == VG_MODEL_MJ) {  // pidx = tissue types; no scaling sets
    if (nsets != 0 || sc || !pidx) return VG_ERR_ARG;
    std::vector<uint16_t> t(pidx, pidx + c->n);
    for (auto &v : t) v = (uint16_t)(v < 9 ? v : 9);
    if (!c->pidx.p && c->pidx.alloc(c->n * 2) != hipSuccess) return VG_ERR_HIP;
    if (hipMemcpy(c->pidx.p, t.data(), c->n * 2, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
    return VG_OK;
  }
  if (nsets < 1 || !sc) return VG_ERR_ARG;
  const size_t T = c->prec == VG_PREC_FP64 ? 8 : 4;
  std::vector<double> v(sc, sc + (size_t)nsets * VG_TT_NSCALE);
  if (c->scales.p) {
    (void)hipFree(c->scales.p);
    c->scales.p = nullptr;
  }
  if (T == 8) {
    if (c->scales.alloc(v.size() * 8) != hipSuccess) return VG_ERR_HIP;
    if (hipMemcpy(c->scales.p, v.data(), v.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  } else {
    std::vector<float> f(v.begin(), v.end());
    if (c->scales.alloc(f.size() * 4) != hipSuccess) return VG_ERR_HIP;
    if (hipMemcpy(c->scales.p, f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  }
  c->nsets = nsets;
  if (pidx) {
    for (int64_t i = 0; i < c->n; i++)
      if (pidx[i] >= nsets) return VG_ERR_ARG;
    if (!c->pidx.p && c->pidx.alloc(c->n * 2) != hipSuccess) return VG_ERR_HIP;
    if (hipMemcpy(c->pidx.p, pidx, c->n * 2, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  }
  return VG_OK;
}

int vg_cells_compute_ion(vg_cells *c, double Cm, double dt, const double *V, const double *istim,
                         double *dvdt) {
  if (!c || !V || !dvdt) return VG_ERR_ARG;
  const int64_t n = c->n;
  if (hipMemcpy(c->V.p, V, n * 8, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  if (istim && hipMemcpy(c->stim.p, istim, n * 8, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  const unsigned nb = (unsigned)(c->n_alloc / BLK);
  const uint16_t *pi = c->pidx.p ? c->pidx.as<uint16_t>() : nullptr;
  const double *sp = istim ? c->stim.as<double>() : nullptr;
#define VG_CELLS(T, M, I)                                                                      \
  hipLaunchKernelGGL((k_cells<T, M, I>), dim3(nb), dim3(BLK), 0, 0, c->V.as<double>(), sp,      \
                     c->dvdt.as<double>(), c->states.as<T>(),                                   \
                     c->acc.p ? c->acc.as<double>() : (double *)nullptr, pi, c->scales.as<T>(), \
                     n, c->n_alloc, Cm, dt, 500, c->newton_jt, c->mj_int_abs,              \
                     c->lr_table.p ? c->lr_table.as<T>() : (const T *)nullptr)
  if (c->prec == VG_PREC_FP64) {
    if (c->model == VG_MODEL_LR) VG_CELLS(double, VG_MODEL_LR, VG_INTEG_REFERENCE);
    else if (c->model == VG_MODEL_PURKINJE) VG_CELLS(double, VG_MODEL_PURKINJE, VG_INTEG_REFERENCE);
    else if (c->model == VG_MODEL_MJ && c->mj_fail) VG_CELLS(double, VG_MODEL_MJ_FAIL, VG_INTEG_REFERENCE);
    else if (c->model == VG_MODEL_MJ) VG_CELLS(double, VG_MODEL_MJ, VG_INTEG_REFERENCE);
    else if (c->integ == VG_INTEG_REFERENCE) VG_CELLS(double, VG_MODEL_TT, VG_INTEG_REFERENCE);
    else VG_CELLS(double, VG_MODEL_TT, VG_INTEG_RUSH_LARSEN);
  } else {
    if (c->model == VG_MODEL_LR) VG_CELLS(float, VG_MODEL_LR, VG_INTEG_REFERENCE);
    else if (c->model == VG_MODEL_PURKINJE) VG_CELLS(float, VG_MODEL_PURKINJE, VG_INTEG_REFERENCE);
    else if (c->model == VG_MODEL_MJ && c->mj_fail) VG_CELLS(float, VG_MODEL_MJ_FAIL, VG_INTEG_REFERENCE);
    else if (c->model == VG_MODEL_MJ) VG_CELLS(float, VG_MODEL_MJ, VG_INTEG_REFERENCE);
    else if (c->integ == VG_INTEG_REFERENCE) VG_CELLS(float, VG_MODEL_TT, VG_INTEG_REFERENCE);
    else VG_CELLS(float, VG_MODEL_TT, VG_INTEG_RUSH_LARSEN);
  }
#undef VG_CELLS
  if (hipGetLastError() != hipSuccess) return VG_ERR_HIP;
  if (hipMemcpy(dvdt, c->dvdt.p, n * 8, hipMemcpyDeviceToHost) != hipSuccess) return VG_ERR_HIP;
  return VG_OK;
}

int vg_cells_get_states(vg_cells *c, int *nvar, double *st) {
  if (!c || !nvar) return VG_ERR_ARG;
  const int ns = nstate_of(c->model);
  *nvar = ns;
  if (!st) return VG_OK;
  const int vslot = vslot_of(c->model);
  const int64_t n = c->n, na = c->n_alloc;
  std::vector<double> soa((size_t)ndev_of(c->model) * na);
  if (int rc = cells_get_rows(c, soa)) return rc;
  std::vector<double> V(n);
  if (hipMemcpy(V.data(), c->V.p, n * 8, hipMemcpyDeviceToHost) != hipSuccess) return VG_ERR_HIP;
  for (int64_t i = 0; i < n; i++) {
    int k2 = 0;
    for (int k = 0; k < ns; k++) {
      if (k == vslot) {
        st[i * ns + k] = V[i];
        continue;
      }
      st[i * ns + k] = soa[(size_t)k2 * na + i];
      k2++;
    }
  }
  return VG_OK;
}

int vg_cells_set_states(vg_cells *c, int nvar, const double *st) {
  if (!c || !st) return VG_ERR_ARG;
  const int ns = nstate_of(c->model);
  if (nvar != ns) return VG_ERR_ARG;
  const int vslot = vslot_of(c->model), nd = ndev_of(c->model);
  const int64_t n = c->n, na = c->n_alloc;
  std::vector<double> soa((size_t)nd * na, 0.0);
  std::vector<double> V(n);
  for (int64_t i = 0; i < n; i++) {
    int k2 = 0;
    V[i] = st[i * ns + (vslot >= 0 ? vslot : 0)];
    for (int k = 0; k < ns; k++) {
      if (k == vslot) continue;
      soa[(size_t)k2 * na + i] = st[i * ns + k];
      k2++;
    }
  }
  for (int k = 0; k < nd; k++)
    for (int64_t i = n; i < na; i++) soa[(size_t)k * na + i] = soa[(size_t)k * na];
  if (int rc = cells_put_rows(c, soa)) return rc;
  if (hipMemcpy(c->V.p, V.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  return VG_OK;
}

}  // extern "C"
