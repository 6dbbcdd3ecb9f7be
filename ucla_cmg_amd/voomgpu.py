"""Python host mirror of libvoomgpu (ctypes over the C ABI in include/voomgpu.h).

The classes keep the reference's operator vocabulary so a driver loop reads
like ``EPSimulator.cc:507-596``:

* :class:`CardiacSolver` -- ``CardiacBody`` + ``EP_CrankNicholson`` on one
  device: ``Solve``, ``setStimulusCurrent`` (4 overloads), ``InitializeVoltage``,
  ``getVoltage``, ``getStateVariables`` / ``setStateVariables``, ``computeECG``.
* :class:`IonicCells` -- ``IonicModel::Compute_Ion`` on independent cells.

There is no CPU fallback: every call runs the HIP library, and constructing a
solver without the built library or a visible GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvoomgpu.so")

VG_OK = 0
VG_ERR_ARG, VG_ERR_HIP, VG_ERR_NOT_CONVERGED, VG_ERR_NONFINITE = -1, -2, -3, -4
MODEL_LR = 0
MODEL_LR_GATE = 1  # Luo-Rudy through LuoRudyGateTable (LUORUDYWITHGATE)
MODEL_TT = 2
MODEL_MJ = 3
MODEL_MJ_FAIL = 4  # Mahajan with the failing-heart constants (MAHAJAN_FAIL)
MODEL_PURKINJE = 10
PREC_FP64 = 0
PREC_MIXED = 1
INTEG_REFERENCE = 0
INTEG_RUSH_LARSEN = 1
QUIRK_V0_ZERO = 1
QUIRK_PMJ_STICKY = 2
QUIRK_BAR_1X_VOLUME = 4
QUIRK_NEWTON_JT = 8
QUIRK_MJ_INT_ABS = 16
TT_NSCALE = 5
X0_ZERO, X0_PREVIOUS, X0_LAST, X0_EXTRAPOLATE, X0_EXTRAPOLATE_LINEAR = 0, 1, 2, 3, 4
SCHEME_CN, SCHEME_BE, SCHEME_EXPLICIT = 0, 1, 2
ORDER_AUTO, ORDER_MESH, ORDER_GEOMETRIC = 0, 1, 2
KC_RHS, KC_CG_ITER, KC_CG_START, KC_IONIC, KC_APPLY, KC_PK_PRESOLVE = range(6)
KC_NAMES = ["rhs", "cg_iter", "cg_start", "ionic", "apply", "pk_presolve"]

# Every symbol the header declares (checked by the CPU test suite).
EXPORTED = [
    "vg_version", "vg_create", "vg_destroy", "vg_last_error", "vg_get_stream",
    "vg_set_stimulus_nodes", "vg_set_stimulus_sphere", "vg_set_stimulus_axis",
    "vg_set_stimulus_plane", "vg_clear_stimulus", "vg_step", "vg_synchronize",
    "vg_set_cg_blocks", "vg_get_cg_blocks", "vg_set_graph", "vg_set_true_residual_check",
    "vg_set_value_streaming", "vg_get_value_streaming",
    "vg_get_voltage", "vg_set_voltage", "vg_get_states", "vg_set_states",
    "vg_get_states_ragged", "vg_set_states_ragged",
    "vg_get_lumped_capacitance", "vg_activation_enable", "vg_get_activation",
    "vg_ecg_setup", "vg_ecg", "vg_ecg_device", "vg_ecg_record", "vg_ecg_frames", "vg_profile_enable",
    "vg_profile_get", "vg_stats_get", "vg_assemble_csr", "vg_cells_create",
    "vg_cells_destroy", "vg_cells_set_scales", "vg_cells_set_quirks", "vg_cells_compute_ion",
    "vg_cells_get_states", "vg_cells_set_states",
    "vg_surr_create", "vg_surr_destroy", "vg_surr_last_error", "vg_surr_reserve",
    "vg_surr_add_samples", "vg_surr_num_samples", "vg_surr_get_samples", "vg_surr_collect",
    "vg_surr_fit_scaling", "vg_surr_set_upstroke_fallback", "vg_surr_set_band_fallback",
    "vg_surr_get_scaling",
    "vg_surr_set_scaling", "vg_surr_num_params",
    "vg_surr_get_params", "vg_surr_set_params", "vg_surr_loss_grad", "vg_surr_predict",
    "vg_surr_train", "vg_surr_attach", "vg_surr_usage",
    "vg_mesh_read", "vg_meshfile_mesh", "vg_meshfile_free", "vg_mesh_write", "vg_frame_name",
    "vg_write_voltage", "vg_write_restart", "vg_read_restart", "vg_ecg_leadfield",
]

# Learned ten Tusscher step: inputs V, 18 states, log(dt), 5 scalings;
# outputs the 19 increments; 4 hidden layers of 128.
SURR_NIN, SURR_NOUT, SURR_HID = 25, 19, 128
# Layer 0 reads each of the 25 normalised inputs as two BF16 halves (his, then los).
SURR_LAYERS = [(50, 128), (128, 128), (128, 128), (128, 128), (128, 19)]


class VgMesh(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("xyz", C.c_void_p), ("n_elems", C.c_int64),
        ("elem_nn", C.c_void_p), ("conn", C.c_void_p), ("fibers", C.c_void_p),
        ("tissue", C.c_void_p), ("n_pmj", C.c_int64), ("pmj_node", C.c_void_p),
        ("pmj_data", C.c_void_p), ("radius", C.c_double),
    ]


class VgModel(C.Structure):
    _fields_ = [
        ("cell_model", C.c_int), ("precision", C.c_int), ("integrator", C.c_int),
        ("D", C.c_double * 3), ("Dp", C.c_double * 3), ("DpLeft", C.c_double * 3),
        ("Cm", C.c_double), ("dt", C.c_double), ("dt_ionic", C.c_double),
        ("cg_tol", C.c_double), ("cg_maxit", C.c_int), ("quirks", C.c_uint),
        ("n_param_sets", C.c_int), ("param_scales", C.c_void_p),
        ("param_index", C.c_void_p), ("sell_sigma", C.c_int), ("cg_blocks", C.c_int),
        ("cg_x0", C.c_int), ("n_instances", C.c_int), ("instance", C.c_void_p),
        ("graph", C.c_int), ("smoothing", C.c_double), ("scheme", C.c_int),
        ("dt_diffusion", C.c_double), ("row_order", C.c_int),
        ("true_residual_every", C.c_int),
    ]


class VgStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("cg_iterations", C.c_int64), ("cg_solves", C.c_int64),
        ("cg_not_converged", C.c_int64), ("newton_iterations", C.c_int64),
        ("n_nodes", C.c_int64), ("nnz", C.c_int64), ("sell_padded_nnz", C.c_int64),
        ("last_cg_relres", C.c_double), ("sim_time", C.c_double),
        ("sell_wide_slices", C.c_int64), ("row_order", C.c_int32),
        ("cg_truncated", C.c_int64), ("max_true_relres", C.c_double),
        ("last_true_relres", C.c_double), ("true_residual_checks", C.c_int64),
        ("sell_shared_offset_slices", C.c_int64),
    ]


_lib = None


def load_library(path: str = None):
    """Loads libvoomgpu.so (or the build named by $VOOMGPU_LIB); raises if it
    has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("VOOMGPU_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    H = C.c_void_p
    sig = {
        "vg_version": ([], C.c_char_p),
        "vg_create": ([C.POINTER(VgMesh), C.POINTER(VgModel), C.POINTER(C.c_void_p)], C.c_int),
        "vg_destroy": ([H], None),
        "vg_last_error": ([H], C.c_char_p),
        "vg_get_stream": ([H, C.POINTER(C.c_void_p)], C.c_int),
        "vg_set_stimulus_nodes": ([H, C.c_void_p, C.c_int64, C.c_double], C.c_int),
        "vg_set_stimulus_sphere": ([H, dp, C.c_double, C.c_double], C.c_int),
        "vg_set_stimulus_axis": ([H, C.c_int, C.c_double, C.c_double], C.c_int),
        "vg_set_stimulus_plane": ([H, dp, dp, C.c_double], C.c_int),
        "vg_clear_stimulus": ([H], C.c_int),
        "vg_step": ([H, C.c_int], C.c_int),
        "vg_synchronize": ([H], C.c_int),
        "vg_set_cg_blocks": ([H, C.c_int], C.c_int),
        "vg_get_cg_blocks": ([H, C.POINTER(C.c_int)], C.c_int),
        "vg_set_graph": ([H, C.c_int], C.c_int),
        "vg_set_value_streaming": ([H, C.c_int], C.c_int),
        "vg_get_value_streaming": ([H, C.POINTER(C.c_int)], C.c_int),
        "vg_set_true_residual_check": ([H, C.c_int], C.c_int),
        "vg_get_voltage": ([H, C.c_void_p], C.c_int),
        "vg_set_voltage": ([H, C.c_void_p], C.c_int),
        "vg_get_states": ([H, C.POINTER(C.c_int), C.c_void_p], C.c_int),
        "vg_set_states": ([H, C.c_int, C.c_void_p], C.c_int),
        "vg_get_states_ragged": ([H, C.c_void_p, C.c_void_p], C.c_int),
        "vg_set_states_ragged": ([H, C.c_void_p, C.c_void_p], C.c_int),
        "vg_get_lumped_capacitance": ([H, C.c_void_p], C.c_int),
        "vg_activation_enable": ([H, C.c_double], C.c_int),
        "vg_get_activation": ([H, C.c_void_p], C.c_int),
        "vg_ecg_setup": ([H, C.POINTER(VgMesh), C.c_void_p, C.c_int, dp, C.c_double], C.c_int),
        "vg_ecg": ([H, C.c_void_p], C.c_int),
        "vg_ecg_device": ([H, C.c_void_p], C.c_int),
        "vg_ecg_record": ([H, C.c_int, C.c_int], C.c_int),
        "vg_ecg_frames": ([H, C.c_void_p, C.POINTER(C.c_int)], C.c_int),
        "vg_profile_enable": ([H, C.c_int], C.c_int),
        "vg_profile_get": ([H, C.c_int, dp, C.POINTER(C.c_int64)], C.c_int),
        "vg_stats_get": ([H, C.POINTER(VgStats)], C.c_int),
        "vg_assemble_csr": ([C.POINTER(VgMesh), C.POINTER(VgModel), C.POINTER(C.c_int64),
                             C.POINTER(C.c_int64), C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p], C.c_int),
        "vg_cells_create": ([C.c_int, C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_void_p)], C.c_int),
        "vg_cells_destroy": ([H], None),
        "vg_cells_set_scales": ([H, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
        "vg_cells_set_quirks": ([H, C.c_uint], C.c_int),
        "vg_cells_compute_ion": ([H, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                  C.c_void_p], C.c_int),
        "vg_cells_get_states": ([H, C.POINTER(C.c_int), C.c_void_p], C.c_int),
        "vg_cells_set_states": ([H, C.c_int, C.c_void_p], C.c_int),
        "vg_surr_create": ([C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
        "vg_surr_destroy": ([H], None),
        "vg_surr_last_error": ([H], C.c_char_p),
        "vg_surr_reserve": ([H, C.c_int64], C.c_int),
        "vg_surr_add_samples": ([H, C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
        "vg_surr_num_samples": ([H, C.POINTER(C.c_int64)], C.c_int),
        "vg_surr_get_samples": ([H, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
        "vg_surr_collect": ([H, H, C.c_int, C.c_int, C.c_int], C.c_int),
        "vg_surr_fit_scaling": ([H, C.c_double], C.c_int),
        "vg_surr_set_upstroke_fallback": ([H, C.c_double, C.c_double], C.c_int),
        "vg_surr_set_band_fallback": ([H, C.c_double, C.c_double], C.c_int),
        "vg_surr_get_scaling": ([H, C.c_void_p], C.c_int),
        "vg_surr_set_scaling": ([H, C.c_void_p], C.c_int),
        "vg_surr_num_params": ([C.POINTER(C.c_int64)], C.c_int),
        "vg_surr_get_params": ([H, C.c_void_p], C.c_int),
        "vg_surr_set_params": ([H, C.c_void_p], C.c_int),
        "vg_surr_loss_grad": ([H, C.c_void_p, C.c_void_p, C.c_int, dp, C.c_void_p], C.c_int),
        "vg_surr_predict": ([H, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
        "vg_surr_train": ([H, C.c_int, C.c_int, C.c_double, dp], C.c_int),
        "vg_surr_attach": ([H, H], C.c_int),
        "vg_surr_usage": ([H, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "vg_mesh_read": ([C.c_char_p, C.c_double, C.POINTER(C.c_void_p), C.c_char_p, C.c_int],
                         C.c_int),
        "vg_meshfile_mesh": ([H], C.POINTER(VgMesh)),
        "vg_meshfile_free": ([H], None),
        "vg_mesh_write": ([C.c_char_p, C.POINTER(VgMesh), C.c_char_p, C.c_int], C.c_int),
        "vg_frame_name": ([C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_double, C.c_int],
                          C.c_int),
        "vg_write_voltage": ([H, C.c_char_p, C.c_double, C.c_int], C.c_int),
        "vg_write_restart": ([H, C.c_char_p, C.c_double, C.c_int], C.c_int),
        "vg_read_restart": ([H, C.c_char_p, C.c_double, C.c_int], C.c_int),
        "vg_ecg_leadfield": ([C.POINTER(VgMesh), C.c_uint, C.c_void_p, C.c_int, dp, C.c_double,
                              C.c_void_p, C.c_char_p, C.c_int], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class _MeshView:
    """Keeps contiguous copies of a mesh's arrays alive behind a VgMesh."""

    def __init__(self, mesh):
        self.xyz = np.ascontiguousarray(mesh.xyz, np.float64)
        self.elem_nn = np.ascontiguousarray(mesh.elem_nn, np.int8)
        self.conn = np.ascontiguousarray(mesh.conn, np.int32)
        self.fibers = None if mesh.fibers is None else np.ascontiguousarray(mesh.fibers, np.float64)
        self.tissue = None if mesh.tissue is None else np.ascontiguousarray(mesh.tissue, np.int8)
        pn, pd = getattr(mesh, "pmj_node", None), getattr(mesh, "pmj_data", None)
        self.pmj_node = None if pn is None else np.ascontiguousarray(pn, np.int32)
        self.pmj_data = None if pd is None else np.ascontiguousarray(pd, np.float64)
        m = VgMesh()
        m.n_nodes = self.xyz.shape[0]
        m.xyz = _p(self.xyz)
        m.n_elems = self.elem_nn.shape[0]
        m.elem_nn = _p(self.elem_nn)
        m.conn = _p(self.conn)
        m.fibers = _p(self.fibers)
        m.tissue = _p(self.tissue)
        m.n_pmj = 0 if self.pmj_node is None else self.pmj_node.shape[0]
        m.pmj_node = _p(self.pmj_node)
        m.pmj_data = _p(self.pmj_data)
        m.radius = float(getattr(mesh, "radius", 0.04))
        self.struct = m


def make_model(cell_model=MODEL_TT, precision=PREC_FP64, integrator=INTEG_REFERENCE,
               D=(1e-3, 1e-3, 1e-3), Dp=(1.6e-3,) * 3, DpLeft=None, Cm=1.0, dt=0.01,
               dt_ionic=None, cg_tol=1e-7, cg_maxit=1000, quirks=0, param_scales=None,
               param_index=None, sell_sigma=0, cg_blocks=0, cg_x0=X0_ZERO, instance=None,
               graph=0, smoothing=0.0, scheme=0, dt_diffusion=0.0, row_order=0,
               true_residual_every=0):
    """Builds a VgModel; returns (struct, keep-alive list)."""
    m = VgModel()
    m.cell_model = cell_model
    m.precision = precision
    m.integrator = integrator
    m.D = _d3(D)
    m.Dp = _d3(Dp)
    m.DpLeft = _d3(DpLeft if DpLeft is not None else Dp)
    m.Cm = Cm
    m.dt = dt
    m.dt_ionic = dt if dt_ionic is None else dt_ionic
    m.cg_tol = cg_tol
    m.cg_maxit = cg_maxit
    m.quirks = quirks
    keep = []
    if param_scales is not None:
        ps = np.ascontiguousarray(param_scales, np.float64).reshape(-1, TT_NSCALE)
        keep.append(ps)
        m.n_param_sets = ps.shape[0]
        m.param_scales = _p(ps)
    if param_index is not None:
        pi = np.ascontiguousarray(param_index, np.uint16)
        keep.append(pi)
        m.param_index = _p(pi)
    m.sell_sigma = sell_sigma
    m.cg_blocks = cg_blocks
    m.cg_x0 = cg_x0
    m.graph = graph
    m.smoothing = smoothing
    m.scheme = scheme
    m.dt_diffusion = dt_diffusion
    m.row_order = row_order
    m.true_residual_every = true_residual_every
    if instance is not None:
        ins = np.ascontiguousarray(instance, np.int32)
        keep.append(ins)
        m.n_instances = int(ins.max()) + 1 if ins.size else 0
        m.instance = _p(ins)
    return m, keep


def assemble_csr(mesh, **model_kw):
    """Host-only assembly of K (with Cm), M and C_L through the library."""
    L = load_library()
    mv = _MeshView(mesh)
    mod, keep = make_model(**model_kw)
    n, nnz = C.c_int64(), C.c_int64()
    rc = L.vg_assemble_csr(C.byref(mv.struct), C.byref(mod), C.byref(n), C.byref(nnz),
                           None, None, None, None, None)
    if rc != VG_OK:
        raise ValueError(f"vg_assemble_csr failed ({rc})")
    rp = np.zeros(n.value + 1, np.int64)
    col = np.zeros(nnz.value, np.int32)
    K = np.zeros(nnz.value)
    M = np.zeros(nnz.value)
    CL = np.zeros(n.value)
    rc = L.vg_assemble_csr(C.byref(mv.struct), C.byref(mod), C.byref(n), C.byref(nnz),
                           _p(rp), _p(col), _p(K), _p(M), _p(CL))
    if rc != VG_OK:
        raise ValueError(f"vg_assemble_csr failed ({rc})")
    return rp, col, K, M, CL


class VoomError(RuntimeError):
    pass


def read_mesh(model: str, radius: float = 0.04):
    """``Mesh::New``'s files (<model>.node/.ele[/.fibers/.purk/.cell/.angle],
    Mesh/Mesh.cc:142-416) -> a 0-based :class:`meshgen.Mesh`."""
    from . import meshgen
    L = load_library()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.vg_mesh_read(model.encode(), radius, C.byref(h), err, 512)
    if rc != VG_OK:
        raise VoomError(f"vg_mesh_read: {err.value.decode()} ({rc})")
    try:
        m = L.vg_meshfile_mesh(h).contents
        n, ne = m.n_nodes, m.n_elems

        def arr(ptr, ctype, count, shape):
            if not ptr or count == 0:
                return None
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), (count,)).reshape(shape).copy()

        mesh = meshgen.Mesh(
            xyz=arr(m.xyz, C.c_double, 3 * n, (n, 3)),
            elem_nn=arr(m.elem_nn, C.c_int8, ne, (ne,)),
            conn=arr(m.conn, C.c_int32, 4 * ne, (ne, 4)),
            fibers=arr(m.fibers, C.c_double, 9 * ne, (ne, 9)),
            tissue=arr(m.tissue, C.c_int8, n, (n,)),
            pmj_node=arr(m.pmj_node, C.c_int32, m.n_pmj, (m.n_pmj,)),
            pmj_data=arr(m.pmj_data, C.c_double, 4 * m.n_pmj, (m.n_pmj, 4)),
            radius=m.radius, meta={"kind": "file", "model": model})
    finally:
        L.vg_meshfile_free(h)
    return mesh


def write_mesh(model: str, mesh):
    """Writes ``mesh`` in the reference's files (see :func:`read_mesh`)."""
    L = load_library()
    view = _MeshView(mesh)
    err = C.create_string_buffer(512)
    rc = L.vg_mesh_write(model.encode(), C.byref(view.struct), err, 512)
    if rc != VG_OK:
        raise VoomError(f"vg_mesh_write: {err.value.decode()} ({rc})")


def ecg_leadfield(mesh, electrodes, D=(1e-3, 5e-4, 2.5e-4), radius=0.0275, quirks=0):
    """Host lead field H[ne][n] of the pseudo-ECG (no GPU): ecg = H @ V."""
    L = load_library()
    view = _MeshView(mesh)
    el = np.ascontiguousarray(electrodes, np.float64).reshape(-1, 3)
    H = np.zeros((el.shape[0], view.xyz.shape[0]))
    err = C.create_string_buffer(512)
    rc = L.vg_ecg_leadfield(C.byref(view.struct), quirks, _p(el), el.shape[0], _d3(D), radius,
                            _p(H), err, 512)
    if rc != VG_OK:
        raise VoomError(f"vg_ecg_leadfield: {err.value.decode()} ({rc})")
    return H


def frame_name(directory: str, prefix: str, time: float, rank: int = 0) -> str:
    L = load_library()
    buf = C.create_string_buffer(4096)
    if L.vg_frame_name(buf, 4096, directory.encode(), prefix.encode(), time, rank) != VG_OK:
        raise VoomError("vg_frame_name failed")
    return buf.value.decode()


class CardiacSolver:
    """``CardiacBody`` + ``EP_CrankNicholson`` resident on the current GPU."""

    def __init__(self, mesh, **model_kw):
        self.L = load_library()
        self._mesh = _MeshView(mesh)
        self._model, self._keep = make_model(**model_kw)
        h = C.c_void_p()
        rc = self.L.vg_create(C.byref(self._mesh.struct), C.byref(self._model), C.byref(h))
        self.h = h.value
        if rc != VG_OK:
            msg = self.L.vg_last_error(self.h).decode() if self.h else "vg_create failed"
            if self.h:
                self.L.vg_destroy(self.h)
                self.h = None
            raise VoomError(f"vg_create: {msg} ({rc})")
        self.n = self._mesh.xyz.shape[0]
        self.cell_model = self._model.cell_model

    def close(self):
        if getattr(self, "h", None):
            self.L.vg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != VG_OK:
            raise VoomError(f"{what}: {self.L.vg_last_error(self.h).decode()} ({rc})")

    # -- CardiacBody::setStimulusCurrent (CardiacBody.cc:243-324) -------------
    def setStimulusCurrent(self, *args):
        """Overloads: (nodeSet, istim), (center[3], radius, istim),
        (axis:int, limit, istim), (origin[3], normal[3], istim)."""
        if len(args) == 2:
            nodes = np.ascontiguousarray(args[0], np.int32)
            self._chk(self.L.vg_set_stimulus_nodes(self.h, _p(nodes), nodes.size, float(args[1])),
                      "set_stimulus_nodes")
        elif len(args) == 3 and isinstance(args[0], (int, np.integer)):
            self._chk(self.L.vg_set_stimulus_axis(self.h, int(args[0]), float(args[1]),
                                                  float(args[2])), "set_stimulus_axis")
        elif len(args) == 3 and np.ndim(args[1]) == 0:
            self._chk(self.L.vg_set_stimulus_sphere(self.h, _d3(args[0]), float(args[1]),
                                                    float(args[2])), "set_stimulus_sphere")
        elif len(args) == 3:
            self._chk(self.L.vg_set_stimulus_plane(self.h, _d3(args[0]), _d3(args[1]),
                                                   float(args[2])), "set_stimulus_plane")
        else:
            raise TypeError("unsupported setStimulusCurrent overload")

    def clearStimulus(self):
        self._chk(self.L.vg_clear_stimulus(self.h), "clear_stimulus")

    # -- EP_CrankNicholson::Solve ---------------------------------------------
    def Solve(self, nsteps: int = 1):
        self._chk(self.L.vg_step(self.h, int(nsteps)), "step")

    def synchronize(self):
        self._chk(self.L.vg_synchronize(self.h), "synchronize")

    def setLaunch(self, cg_blocks: Optional[int] = None, graph: Optional[bool] = None):
        """Launch shape between steps: total CG grid and captured-step mode."""
        if cg_blocks is not None:
            self._chk(self.L.vg_set_cg_blocks(self.h, int(cg_blocks)), "set_cg_blocks")
        if graph is not None:
            self._chk(self.L.vg_set_graph(self.h, int(bool(graph))), "set_graph")

    def setTrueResidualCheck(self, every: int):
        """FP64 true residual after both solves of every ``every``-th step
        (0 = off); the largest goes to stats()['max_true_relres']."""
        self._chk(self.L.vg_set_true_residual_check(self.h, int(every)), "set_true_residual_check")

    def setValueStreaming(self, mode: int = -1):
        """How the CG reads the operator's values: -1 the library's choice by
        operator size, 0 plain loads, 1 nontemporal loads (same results)."""
        self._chk(self.L.vg_set_value_streaming(self.h, int(mode)), "set_value_streaming")

    def valueStreaming(self) -> bool:
        on = C.c_int()
        self._chk(self.L.vg_get_value_streaming(self.h, C.byref(on)), "get_value_streaming")
        return bool(on.value)

    def cgBlocks(self) -> int:
        g = C.c_int()
        self._chk(self.L.vg_get_cg_blocks(self.h, C.byref(g)), "get_cg_blocks")
        return g.value

    def getVoltage(self) -> np.ndarray:
        v = np.zeros(self.n)
        self._chk(self.L.vg_get_voltage(self.h, _p(v)), "get_voltage")
        return v

    def InitializeVoltage(self, V):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(V, np.float64), (self.n,)))
        self._chk(self.L.vg_set_voltage(self.h, _p(v)), "set_voltage")

    def getStateVariables(self) -> np.ndarray:
        nv = C.c_int()
        self._chk(self.L.vg_get_states(self.h, C.byref(nv), None), "get_states")
        st = np.zeros((self.n, nv.value))
        self._chk(self.L.vg_get_states(self.h, C.byref(nv), _p(st)), "get_states")
        return st

    def setStateVariables(self, st):
        st = np.ascontiguousarray(st, np.float64)
        self._chk(self.L.vg_set_states(self.h, st.shape[1], _p(st)), "set_states")

    def getStateVariablesRagged(self):
        """(numVar[n], list of per-node state vectors) like
        CardiacBody::getStateVariables (CardiacBody.cc:438-451)."""
        nv = np.zeros(self.n, np.int32)
        self._chk(self.L.vg_get_states_ragged(self.h, _p(nv), None), "get_states_ragged")
        data = np.zeros(int(nv.sum()))
        self._chk(self.L.vg_get_states_ragged(self.h, _p(nv), _p(data)), "get_states_ragged")
        off = np.concatenate([[0], np.cumsum(nv)])
        return nv, [data[off[i]:off[i + 1]] for i in range(self.n)]

    def setStateVariablesRagged(self, num_var, states):
        nv = np.ascontiguousarray(num_var, np.int32)
        data = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in states]))
        self._chk(self.L.vg_set_states_ragged(self.h, _p(nv), _p(data)), "set_states_ragged")

    def lumpedCapacitance(self) -> np.ndarray:
        c = np.zeros(self.n)
        self._chk(self.L.vg_get_lumped_capacitance(self.h, _p(c)), "get_lumped_capacitance")
        return c

    def enableActivation(self, threshold=-40.0):
        self._chk(self.L.vg_activation_enable(self.h, threshold), "activation_enable")

    def activationTimes(self) -> np.ndarray:
        a = np.zeros(self.n)
        self._chk(self.L.vg_get_activation(self.h, _p(a)), "get_activation")
        return a

    def setupECG(self, electrodes, D=(1e-3, 5e-4, 2.5e-4), radius=0.0275):
        el = np.ascontiguousarray(electrodes, np.float64).reshape(-1, 3)
        self.n_electrodes = el.shape[0]
        self._chk(self.L.vg_ecg_setup(self.h, C.byref(self._mesh.struct), _p(el), el.shape[0],
                                      _d3(D), radius), "ecg_setup")

    def computeECG(self) -> np.ndarray:
        out = np.zeros(self.n_electrodes)
        self._chk(self.L.vg_ecg(self.h, _p(out)), "ecg")
        return out

    def recordECG(self, steps_per_frame: int, max_frames: int):
        self._ecg_max = max_frames
        self._chk(self.L.vg_ecg_record(self.h, steps_per_frame, max_frames), "ecg_record")

    def ecgFrames(self) -> np.ndarray:
        nf = C.c_int()
        out = np.zeros((self._ecg_max, self.n_electrodes))
        self._chk(self.L.vg_ecg_frames(self.h, _p(out), C.byref(nf)), "ecg_frames")
        return out[: nf.value]

    def stream(self) -> int:
        s = C.c_void_p()
        self._chk(self.L.vg_get_stream(self.h, C.byref(s)), "get_stream")
        return s.value or 0

    def profile(self, enable: bool):
        self._chk(self.L.vg_profile_enable(self.h, int(enable)), "profile_enable")

    def profileGet(self, kc: int):
        ms, n = C.c_double(), C.c_int64()
        self._chk(self.L.vg_profile_get(self.h, kc, C.byref(ms), C.byref(n)), "profile_get")
        return ms.value, n.value

    # -- EPSolver::WriteOutput / writeRestart / readRestart (EPSolver.cc:116-211)
    def WriteOutput(self, time: float, path: str, rank: int = 0):
        self._chk(self.L.vg_write_voltage(self.h, path.encode(), time, rank), "write_voltage")

    def writeRestart(self, time: float, path: str, rank: int = 0):
        self._chk(self.L.vg_write_restart(self.h, path.encode(), time, rank), "write_restart")

    def readRestart(self, time: float, path: str, rank: int = 0):
        self._chk(self.L.vg_read_restart(self.h, path.encode(), time, rank), "read_restart")

    def stats(self) -> dict:
        s = VgStats()
        self._chk(self.L.vg_stats_get(self.h, C.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in VgStats._fields_}


class IonicCells:
    """``IonicModel::Compute_Ion`` for n independent cells on the device."""

    def __init__(self, cell_model, n, precision=PREC_FP64, integrator=INTEG_REFERENCE):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.vg_cells_create(cell_model, precision, integrator, n, C.byref(h))
        if rc != VG_OK:
            raise VoomError(f"vg_cells_create failed ({rc})")
        self.h = h.value
        self.n = n

    def __del__(self):
        try:
            if self.h:
                self.L.vg_cells_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_scales(self, scales, index=None):
        sc = np.ascontiguousarray(scales, np.float64).reshape(-1, TT_NSCALE)
        pi = None if index is None else np.ascontiguousarray(index, np.uint16)
        rc = self.L.vg_cells_set_scales(self.h, sc.shape[0], _p(sc), _p(pi))
        if rc != VG_OK:
            raise VoomError(f"vg_cells_set_scales failed ({rc})")

    def set_tissue_types(self, types):
        """Mahajan: tissue type 0-8 of each cell (others run type 2)."""
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(types), (self.n,)), np.uint16)
        rc = self.L.vg_cells_set_scales(self.h, 0, None, _p(t))
        if rc != VG_OK:
            raise VoomError(f"vg_cells_set_scales failed ({rc})")

    def set_quirks(self, quirks: int):
        rc = self.L.vg_cells_set_quirks(self.h, quirks)
        if rc != VG_OK:
            raise VoomError(f"vg_cells_set_quirks failed ({rc})")

    def compute_ion(self, Cm, dt, V, istim=None) -> np.ndarray:
        V = np.ascontiguousarray(np.broadcast_to(np.asarray(V, np.float64), (self.n,)))
        st = None if istim is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(istim, np.float64), (self.n,)))
        out = np.zeros(self.n)
        rc = self.L.vg_cells_compute_ion(self.h, Cm, dt, _p(V), _p(st), _p(out))
        if rc != VG_OK:
            raise VoomError(f"vg_cells_compute_ion failed ({rc})")
        return out

    def get_states(self) -> np.ndarray:
        nv = C.c_int()
        self.L.vg_cells_get_states(self.h, C.byref(nv), None)
        st = np.zeros((self.n, nv.value))
        rc = self.L.vg_cells_get_states(self.h, C.byref(nv), _p(st))
        if rc != VG_OK:
            raise VoomError(f"vg_cells_get_states failed ({rc})")
        return st

    def set_states(self, st):
        st = np.ascontiguousarray(st, np.float64)
        rc = self.L.vg_cells_set_states(self.h, st.shape[1], _p(st))
        if rc != VG_OK:
            raise VoomError(f"vg_cells_set_states failed ({rc})")


class VgSurrScaling(C.Structure):
    _fields_ = [(k, C.c_float * (SURR_NIN if k in ("mu_x", "sd_x", "lo", "hi") else SURR_NOUT))
                for k in ("mu_x", "sd_x", "lo", "hi", "mu_y", "sd_y", "c_y", "t_lo", "t_hi")]


class Surrogate:
    """Learned ten Tusscher ionic step (an MLP trained on the device with BF16
    MFMA kernels) and its sample store.  Attached to a solver, it advances
    every myocardium node whose inputs lie inside the training envelope; the
    plain model updates the rest."""


    def __init__(self, seed: int = 1):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.vg_surr_create(seed, C.byref(h))
        if rc != VG_OK:
            raise VoomError(f"vg_surr_create failed ({rc})")
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            self.L.vg_surr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != VG_OK:
            raise VoomError(f"{what}: {self.L.vg_surr_last_error(self.h).decode()} ({rc})")

    def reserve(self, capacity: int):
        self._chk(self.L.vg_surr_reserve(self.h, int(capacity)), "reserve")

    def add_samples(self, x, y):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, SURR_NIN)
        y = np.ascontiguousarray(y, np.float32).reshape(-1, SURR_NOUT)
        assert x.shape[0] == y.shape[0]
        self._chk(self.L.vg_surr_add_samples(self.h, _p(x), _p(y), x.shape[0]), "add_samples")

    def num_samples(self) -> int:
        n = C.c_int64()
        self._chk(self.L.vg_surr_num_samples(self.h, C.byref(n)), "num_samples")
        return n.value

    def get_samples(self, first: int = 0, n: int = None):
        n = self.num_samples() - first if n is None else n
        x = np.zeros((n, SURR_NIN), np.float32)
        y = np.zeros((n, SURR_NOUT), np.float32)
        self._chk(self.L.vg_surr_get_samples(self.h, first, n, _p(x), _p(y)), "get_samples")
        return x, y

    def collect(self, solver, nsteps: int, every: int = 1, node_stride: int = 1):
        """Runs nsteps steps of `solver`, recording the plain ionic update of
        every node_stride-th node on every `every`-th unstimulated step."""
        rc = self.L.vg_surr_collect(solver.h, self.h, nsteps, every, node_stride)
        solver._chk(rc, "surr_collect")

    def fit_scaling(self, margin: float = 0.0):
        self._chk(self.L.vg_surr_fit_scaling(self.h, margin), "fit_scaling")

    def set_upstroke_fallback(self, v_min: float = float("inf"), h_min: float = float("inf")):
        """Leaves rows with V > v_min and the h gate above h_min (the
        upstroke) to the plain model from the next attach; the defaults
        turn it off."""
        self._chk(self.L.vg_surr_set_upstroke_fallback(self.h, float(v_min), float(h_min)),
                  "set_upstroke_fallback")

    def set_band_fallback(self, v_lo: float = float("inf"), v_hi: float = float("-inf")):
        """Leaves rows with v_lo < V < v_hi to the plain model from the next
        attach; the defaults (an empty band) turn it off."""
        self._chk(self.L.vg_surr_set_band_fallback(self.h, float(v_lo), float(v_hi)),
                  "set_band_fallback")

    def get_scaling(self) -> dict:
        sc = VgSurrScaling()
        self._chk(self.L.vg_surr_get_scaling(self.h, C.byref(sc)), "get_scaling")
        return {k: np.array(getattr(sc, k), np.float32) for k, _ in VgSurrScaling._fields_}

    def set_scaling(self, d: dict):
        sc = VgSurrScaling()
        for k, _ in VgSurrScaling._fields_:
            v = np.ascontiguousarray(d[k], np.float32)
            getattr(sc, k)[:] = list(v)
        self._chk(self.L.vg_surr_set_scaling(self.h, C.byref(sc)), "set_scaling")

    @staticmethod
    def num_params() -> int:
        L = load_library()
        n = C.c_int64()
        L.vg_surr_num_params(C.byref(n))
        return n.value

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params(), np.float32)
        self._chk(self.L.vg_surr_get_params(self.h, _p(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params()
        self._chk(self.L.vg_surr_set_params(self.h, _p(p)), "set_params")

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = C.c_double()
        g = np.zeros(self.num_params(), np.float32)
        self._chk(self.L.vg_surr_loss_grad(self.h, _p(x), _p(y), x.shape[0], C.byref(loss), _p(g)),
                  "loss_grad")
        return loss.value, g

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], SURR_NOUT), np.float32)
        self._chk(self.L.vg_surr_predict(self.h, _p(x), x.shape[0], _p(y)), "predict")
        return y

    def train(self, iters: int, batch: int = 4096, lr: float = 1e-3) -> float:
        loss = C.c_double()
        self._chk(self.L.vg_surr_train(self.h, iters, batch, lr, C.byref(loss)), "train")
        return loss.value

    def attach(self, solver):
        solver._chk(self.L.vg_surr_attach(solver.h, self.h), "surr_attach")
        solver._surr = self

    @staticmethod
    def detach(solver):
        solver._chk(solver.L.vg_surr_attach(solver.h, None), "surr_attach")
        solver._surr = None

    @staticmethod
    def usage(solver):
        a, b = C.c_int64(), C.c_int64()
        solver._chk(solver.L.vg_surr_usage(solver.h, C.byref(a), C.byref(b)), "surr_usage")
        return a.value, b.value


def unpack_params(p: np.ndarray):
    """Flat parameters -> [(W[k][n], b[n]) per layer]."""
    out, o = [], 0
    for k, n in SURR_LAYERS:
        W = p[o:o + k * n].reshape(k, n)
        o += k * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    return out
