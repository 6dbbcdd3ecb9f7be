"""Synthetic meshes for the monodomain workloads.

The reference reads meshes from files (``Mesh/Mesh.cc:142-416``: ``.node``,
``.ele``, ``.fibers``, ``.purk``, ``.cell``, ``.angle``).  No mesh ships with it,
so the workloads are generated here, deterministically:

* ``slab``: a structured box of hexahedra, each split into 5 P1 tets with the
  split mirrored on alternate cells (parity of i+j+k) so faces match.
* ``lv_shell``: a truncated prolate-ellipsoid left-ventricle shell on a
  (transmural x circumferential x longitudinal) grid, each hex split into 6
  tets around its main diagonal (Kuhn split, conforming without parity).

Every tet is returned positively oriented, since the reference integrates with
the signed Jacobian determinant (``Mesh/MixedMesh3D.cc:126-136``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Mesh:
    """Mixed P1-tet + 2-node bar mesh, 0-based.

    ``conn`` has 4 slots per element; bars use the first two (``elem_nn`` 2).
    ``fibers`` holds rows f, s, n per element (9 numbers, the layout of the
    reference's ``.fibers`` file) or is ``None`` for an isotropic mesh.
    """

    xyz: np.ndarray
    elem_nn: np.ndarray
    conn: np.ndarray
    fibers: Optional[np.ndarray] = None
    tissue: Optional[np.ndarray] = None
    pmj_node: Optional[np.ndarray] = None
    pmj_data: Optional[np.ndarray] = None
    radius: float = 0.04
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.elem_nn.shape[0])


def _orient(xyz: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Swap the last two nodes of negatively oriented tets."""
    p = xyz[tets]
    d = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    neg = d < 0
    t = tets.copy()
    t[neg, 2], t[neg, 3] = tets[neg, 3], tets[neg, 2]
    return t


# Corner offsets (a,b,c) of a hex, indexed by a + 2b + 4c.
def _v(a, b, c):
    return a + 2 * b + 4 * c


# 5-tet split of an "even" cell: four corner tets and one central tet.
_FIVE_EVEN = [
    (_v(0, 0, 0), _v(1, 0, 0), _v(0, 1, 0), _v(0, 0, 1)),
    (_v(1, 1, 0), _v(1, 0, 0), _v(0, 1, 0), _v(1, 1, 1)),
    (_v(1, 0, 1), _v(1, 0, 0), _v(0, 0, 1), _v(1, 1, 1)),
    (_v(0, 1, 1), _v(0, 1, 0), _v(0, 0, 1), _v(1, 1, 1)),
    (_v(1, 0, 0), _v(0, 1, 0), _v(0, 0, 1), _v(1, 1, 1)),
]
# The mirror image, used on odd cells so shared faces carry the same diagonal.
_FIVE_ODD = [
    (_v(1, 0, 0), _v(0, 0, 0), _v(1, 1, 0), _v(1, 0, 1)),
    (_v(0, 1, 0), _v(0, 0, 0), _v(1, 1, 0), _v(0, 1, 1)),
    (_v(0, 0, 1), _v(0, 0, 0), _v(1, 0, 1), _v(0, 1, 1)),
    (_v(1, 1, 1), _v(1, 1, 0), _v(1, 0, 1), _v(0, 1, 1)),
    (_v(0, 0, 0), _v(1, 1, 0), _v(1, 0, 1), _v(0, 1, 1)),
]
# Kuhn 6-tet split around the 000-111 diagonal.
_SIX = [
    (_v(0, 0, 0), _v(1, 0, 0), _v(1, 1, 0), _v(1, 1, 1)),
    (_v(0, 0, 0), _v(1, 0, 0), _v(1, 0, 1), _v(1, 1, 1)),
    (_v(0, 0, 0), _v(0, 1, 0), _v(1, 1, 0), _v(1, 1, 1)),
    (_v(0, 0, 0), _v(0, 1, 0), _v(0, 1, 1), _v(1, 1, 1)),
    (_v(0, 0, 0), _v(0, 0, 1), _v(1, 0, 1), _v(1, 1, 1)),
    (_v(0, 0, 0), _v(0, 0, 1), _v(0, 1, 1), _v(1, 1, 1)),
]


def slab(nx=41, ny=41, nz=9, h=0.025, fiber_axis=(1.0, 0.0, 0.0)) -> Mesh:
    """Structured slab, node id = i + nx*j + nx*ny*k, 5 tets per hex.

    The default is the 1.0 x 1.0 x 0.2 cm slab at h = 0.025 cm (15,129 nodes,
    64,000 tets) with fibers along x.
    """
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = (i + nx * j + nx * ny * k).astype(np.int64)
    xyz = np.zeros((nx * ny * nz, 3))
    xyz[idx.ravel(), 0] = (i * h).ravel()
    xyz[idx.ravel(), 1] = (j * h).ravel()
    xyz[idx.ravel(), 2] = (k * h).ravel()
    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1),
                             indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    corners = np.stack([
        (ci + a) + nx * (cj + b) + nx * ny * (ck + c)
        for c in (0, 1) for b in (0, 1) for a in (0, 1)
    ], axis=1)  # column order a + 2b + 4c
    even = ((ci + cj + ck) % 2) == 0
    tets = np.empty((ci.size, 5, 4), dtype=np.int64)
    te = np.array(_FIVE_EVEN)
    to = np.array(_FIVE_ODD)
    tets[even] = corners[even][:, te]
    tets[~even] = corners[~even][:, to]
    tets = _orient(xyz, tets.reshape(-1, 4))
    ne = tets.shape[0]
    f = np.asarray(fiber_axis, dtype=np.float64)
    f = f / np.linalg.norm(f)
    # Complete f to an orthonormal frame (s, n).
    trial = np.array([0.0, 1.0, 0.0]) if abs(f[1]) < 0.9 else np.array([0.0, 0.0, 1.0])
    s = trial - f * (trial @ f)
    s /= np.linalg.norm(s)
    n = np.cross(f, s)
    fib = np.tile(np.concatenate([f, s, n]), (ne, 1))
    return Mesh(xyz=xyz, elem_nn=np.full(ne, 4, np.int8), conn=tets.astype(np.int32),
                fibers=fib, meta={"kind": "slab", "shape": (nx, ny, nz), "h": h})


def slab_corner_nodes(nx=41, ny=41, nz=9, m=3) -> np.ndarray:
    """Ids of the m x m x nz corner column (i < m, j < m, all k)."""
    i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(nz), indexing="ij")
    return np.sort((i + nx * j + nx * ny * k).ravel()).astype(np.int32)


def lv_shell(n_trans=26, n_circ=240, n_long=165, base_radius=2.5, apex_depth=7.0,
             wall=1.0, theta_apex_deg=6.0, fiber_endo_deg=60.0, fiber_epi_deg=-60.0) -> Mesh:
    """Truncated prolate-ellipsoid LV shell with a small apical opening.

    The outer surface has equatorial radius ``base_radius`` and long semi-axis
    ``apex_depth``; the inner one is ``wall`` smaller on both axes.  The base
    plane is z = 0 and the apex points to -z.  A structured grid in (depth,
    phi, theta) needs no degenerate cells only if the apical pole is cut out,
    so theta starts at ``theta_apex_deg``.  Fibers rotate linearly from
    ``fiber_endo_deg`` at the endocardium to ``fiber_epi_deg`` at the
    epicardium, measured from the circumferential direction toward the
    longitudinal direction in the wall's tangent plane.

    Defaults give 1,029,600 nodes and 5,904,000 tets.
    ``meta`` carries per-node transmural depth (0 endo .. 1 epi) and
    normalised apicobasal height (0 apex .. 1 base).
    """
    nt, nc, nl = n_trans, n_circ, n_long
    if nc % 2:
        raise ValueError("n_circ must be even")
    d = np.linspace(0.0, 1.0, nt)                     # transmural 0 endo .. 1 epi
    phi = np.arange(nc) * (2 * np.pi / nc)
    th = np.linspace(np.deg2rad(theta_apex_deg), np.pi / 2, nl)  # apex .. base
    a_in, c_in = base_radius - wall, apex_depth - wall
    a = a_in + d * wall
    c = c_in + d * wall
    D, P, T = np.meshgrid(np.arange(nt), np.arange(nc), np.arange(nl), indexing="ij")
    idx = (D + nt * P + nt * nc * T).astype(np.int64)
    ar, cr = a[D], c[D]
    sp, cp = np.sin(phi[P]), np.cos(phi[P])
    st, ct = np.sin(th[T]), np.cos(th[T])
    xyz = np.zeros((nt * nc * nl, 3))
    xyz[idx.ravel(), 0] = (ar * st * cp).ravel()
    xyz[idx.ravel(), 1] = (ar * st * sp).ravel()
    xyz[idx.ravel(), 2] = (-cr * ct).ravel()
    depth = np.zeros(nt * nc * nl)
    depth[idx.ravel()] = d[D].ravel()
    height = np.zeros(nt * nc * nl)
    height[idx.ravel()] = ((th[T] - th[0]) / (th[-1] - th[0])).ravel()

    ci, cj, ck = np.meshgrid(np.arange(nt - 1), np.arange(nc), np.arange(nl - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()

    def node(i, j, k):
        return i + nt * (j % nc) + nt * nc * k

    corners = np.stack([node(ci + a_, cj + b_, ck + c_)
                        for c_ in (0, 1) for b_ in (0, 1) for a_ in (0, 1)], axis=1)
    tets = corners[:, np.array(_SIX)].reshape(-1, 4)
    tets = _orient(xyz, tets)
    # Fibers at each element centroid.
    cen = xyz[tets].mean(axis=1)
    dep = depth[tets].mean(axis=1)
    ang = np.deg2rad(fiber_endo_deg + (fiber_epi_deg - fiber_endo_deg) * dep)
    r_xy = np.hypot(cen[:, 0], cen[:, 1])
    e_c = np.stack([-cen[:, 1] / r_xy, cen[:, 0] / r_xy, np.zeros_like(r_xy)], axis=1)
    # Outward normal of the confocal-ish ellipsoid through the centroid.
    aa = a_in + dep * wall
    cc = c_in + dep * wall
    nrm = np.stack([cen[:, 0] / aa ** 2, cen[:, 1] / aa ** 2, cen[:, 2] / cc ** 2], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    e_l = np.cross(nrm, e_c)
    e_l /= np.linalg.norm(e_l, axis=1, keepdims=True)
    f = np.cos(ang)[:, None] * e_c + np.sin(ang)[:, None] * e_l
    s = nrm
    n = np.cross(f, s)
    fib = np.concatenate([f, s, n], axis=1)
    ne = tets.shape[0]
    return Mesh(xyz=xyz, elem_nn=np.full(ne, 4, np.int8), conn=tets.astype(np.int32),
                fibers=fib,
                meta={"kind": "lv_shell", "shape": (nt, nc, nl), "depth": depth,
                      "height": height})


def electrodes_around(center, radius, n, axis_z=0.0) -> np.ndarray:
    """n electrode positions on a circle of ``radius`` around ``center``."""
    ang = np.arange(n) * (2 * np.pi / n)
    c = np.asarray(center, dtype=np.float64)
    return np.stack([c[0] + radius * np.cos(ang), c[1] + radius * np.sin(ang),
                     np.full(n, c[2] + axis_z)], axis=1)


def pmj_length(resistance_ohm_m: float) -> float:
    """Effective PMJ bar length (cm) standing for a junction resistance.

    The reference describes a junction only by the bar's direction and length
    (``.angle`` file, Mesh/MixedMesh3D.cc:85-88), which set its conductance
    2 D_p pi r^2 / L.  This build maps a resistance linearly onto that length,
    0.4 mm per 1 Ohm m (500 Ohm m -> 0.2 cm), a choice calibrated so a single
    junction still excites the tissue it feeds.
    """
    return 4e-4 * float(resistance_ohm_m)


def purkinje_tree(mesh: Mesh, generations=8, seg_len=0.06, seed=1, offset=0.02,
                  pmj_length=0.1, theta_apex_deg=6.0, base_frac=0.92) -> Mesh:
    """Adds a seeded binary-branching Purkinje network to an ``lv_shell`` mesh.

    The tree runs on a surface ``offset`` cm inside the endocardium: a root
    near the base splits ``generations`` times (2**generations terminals),
    each branch heading apex-ward and spreading in circumferential angle with
    seeded jitter.  Branches are cut into 2-node bars of about ``seg_len`` cm.
    Each terminal gets one PMJ bar to its nearest endocardial node, keyed in
    the PMJ table by the terminal (the element's first node) with the
    direction cosines and the effective length ``pmj_length`` that sets the
    junction's conductance 2 D_p pi r^2 / L (Mesh/MixedMesh3D.cc:85-107).
    Purkinje nodes get tissue type 11 (left ventricle).

    Returns a new Mesh: the myocardium first (unchanged ids), then the
    Purkinje nodes; ``meta['purkinje_root']`` is the root node id and
    ``meta['pmj_terminals']`` the terminal ids.
    """
    if mesh.meta.get("kind") != "lv_shell":
        raise ValueError("purkinje_tree needs an lv_shell mesh")
    rng = np.random.default_rng(seed)
    nt, nc, nl = mesh.meta["shape"]
    depth = mesh.meta["depth"]
    endo = np.nonzero(depth == 0.0)[0]
    xyz_endo = mesh.xyz[endo]
    # Endocardial ellipsoid axes from the generator's geometry.
    a_in = np.max(np.hypot(xyz_endo[:, 0], xyz_endo[:, 1]))
    c_in = -xyz_endo[:, 2].min() / np.cos(np.deg2rad(theta_apex_deg))
    a_p, c_p = a_in - offset, c_in - offset
    th_lo = np.deg2rad(theta_apex_deg) * 1.5
    th_hi = np.pi / 2 * base_frac

    def point(th, ph):
        return np.array([a_p * np.sin(th) * np.cos(ph), a_p * np.sin(th) * np.sin(ph),
                         -c_p * np.cos(th)])

    nodes = []      # (theta, phi)
    bars = []       # (i, j)
    terminals = []

    def add_node(th, ph):
        nodes.append((th, ph))
        return len(nodes) - 1

    def run(i0, th1, ph1):
        th0, ph0 = nodes[i0]
        L = np.linalg.norm(point(th1, ph1) - point(th0, ph0))
        k = max(1, int(round(L / seg_len)))
        prev = i0
        for s in range(1, k + 1):
            f = s / k
            cur = add_node(th0 + f * (th1 - th0), ph0 + f * (ph1 - ph0))
            bars.append((prev, cur))
            prev = cur
        return prev

    root = add_node(th_hi, 0.0)
    frontier = [(root, -np.pi, np.pi)]
    dth = (th_hi - th_lo) / (generations + 1)
    for g in range(generations):
        nxt = []
        for (i, pa, pb) in frontier:
            th, ph = nodes[i]
            mid = 0.5 * (pa + pb)
            for (ca, cb) in ((pa, mid), (mid, pb)):
                tgt_ph = 0.5 * (ca + cb) + rng.uniform(-0.15, 0.15) * (cb - ca)
                tgt_th = max(th_lo, th - dth * rng.uniform(0.7, 1.3))
                end = run(i, tgt_th, tgt_ph)
                nxt.append((end, ca, cb))
        frontier = nxt
    # Terminal twigs towards the apex band so activation reaches all heights.
    for (i, pa, pb) in frontier:
        th, ph = nodes[i]
        tgt_th = th_lo + rng.uniform(0.0, 1.0) * (th - th_lo)
        terminals.append(run(i, tgt_th, ph + rng.uniform(-0.3, 0.3) * (pb - pa)))

    n_myo = mesh.n_nodes
    pk_xyz = np.array([point(th, ph) for th, ph in nodes])
    n_pk = len(nodes)
    # PMJ bars: terminal -> nearest endocardial node.
    term = np.array(terminals)
    d2 = ((pk_xyz[term][:, None, :] - xyz_endo[None, :, :]) ** 2).sum(-1)
    near = endo[np.argmin(d2, axis=1)]
    pmj_bars = [(n_myo + t, m) for t, m in zip(term, near)]
    cable = [(n_myo + i, n_myo + j) for i, j in bars]
    all_bars = np.array(cable + pmj_bars, dtype=np.int32)
    nb = all_bars.shape[0]
    conn_b = np.full((nb, 4), -1, np.int32)
    conn_b[:, :2] = all_bars
    dirs = mesh.xyz[near] - pk_xyz[term]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pmj_data = np.hstack([dirs, np.full((len(term), 1), pmj_length)])
    xyz = np.vstack([mesh.xyz, pk_xyz])
    tissue = np.concatenate([np.zeros(n_myo, np.int8), np.full(n_pk, 11, np.int8)])
    fib_b = np.tile(np.eye(3).ravel(), (nb, 1))
    fibers = None if mesh.fibers is None else np.vstack([mesh.fibers, fib_b])
    meta = dict(mesh.meta)
    meta["depth"] = np.concatenate([depth, np.full(n_pk, -1.0)])
    meta["height"] = np.concatenate([mesh.meta["height"], np.full(n_pk, -1.0)])
    meta["purkinje_root"] = n_myo + root
    meta["pmj_terminals"] = n_myo + term
    meta["n_myocardium"] = n_myo
    return Mesh(xyz=xyz, elem_nn=np.concatenate([mesh.elem_nn, np.full(nb, 2, np.int8)]),
                conn=np.vstack([mesh.conn, conn_b]), fibers=fibers, tissue=tissue,
                pmj_node=(n_myo + term).astype(np.int32), pmj_data=pmj_data,
                radius=mesh.radius, meta=meta)


# Junction resistances of the tree's high and low PMJ variants (Ohm m).
PMJ_RESISTANCE = {"high": 500.0, "low": 125.0}


def lv_workload(scale: float = 1.0, purkinje: str = "high", seg_len: float = 0.06):
    """The ten Tusscher LV workload: ``lv_shell`` (its grid shrunk by
    ``scale`` in node count, the geometry unchanged) with the seeded
    ``purkinje_tree`` (``purkinje`` = "high", "low" or "none" for the PMJ
    resistance; ``seg_len`` its bar length in cm), endo/M/epi cells by transmural depth (30/40/30 %) and an
    apicobasal G_Ks gradient in 8 bands (1.5x at the apex to 1.0x at the
    base).  Returns (mesh, scales [24, 5] in the ten Tusscher scaling order
    gNa gKr gKs gCaL gto, per-node parameter set uint16); Purkinje nodes
    carry a set that their own model ignores."""
    if scale == 1.0:
        mesh = lv_shell()
    else:
        mesh = lv_shell(n_trans=max(4, int(26 * scale ** (1 / 3))),
                        n_circ=2 * max(4, int(120 * scale ** (1 / 3))),
                        n_long=max(4, int(165 * scale ** (1 / 3))))
    if purkinje != "none":
        mesh = purkinje_tree(mesh, seg_len=seg_len,
                             pmj_length=pmj_length(PMJ_RESISTANCE[purkinje]))
    depth = mesh.meta["depth"]
    height = np.clip(mesh.meta["height"], 0.0, 1.0)
    # Cell type by transmural depth: endo [0, .3), M [.3, .7), epi [.7, 1].
    ctype = np.where(depth < 0.3, 0, np.where(depth < 0.7, 1, 2))
    band = np.minimum((height * 8).astype(int), 7)
    ks_grad = 1.5 - 0.5 * (np.arange(8) + 0.5) / 8
    # Published tT2006 layer differences: G_to endo 0.073 (vs 0.294),
    # G_Ks M 0.098 (vs 0.392).
    gto = {0: 0.073 / 0.294, 1: 1.0, 2: 1.0}
    gks = {0: 1.0, 1: 0.098 / 0.392, 2: 1.0}
    scales = []
    for t in range(3):
        for b in range(8):
            scales.append([1.0, 1.0, gks[t] * ks_grad[b], 1.0, gto[t]])
    pidx = (ctype * 8 + band).astype(np.uint16)
    return mesh, np.array(scales), pidx


def his_nodes(mesh: Mesh, n: int = 3) -> np.ndarray:
    """The Purkinje root and the next nodes of its first branch (the His
    stimulus of the LV workload)."""
    r = int(mesh.meta["purkinje_root"])
    return np.arange(r, r + n, dtype=np.int32)


def lv_electrodes() -> np.ndarray:
    """The LV workload's 10 electrodes: 8 on a 5 cm ring around the long
    axis at z = -3.5 cm, one below the apex and one above the base."""
    ring = electrodes_around((0.0, 0.0, -3.5), 5.0, 8)
    return np.vstack([ring, [[0.0, 0.0, -9.0], [0.0, 0.0, 3.0]]])


def scale_cross_fiber(mesh: Mesh, ratio: float) -> Mesh:
    """Copy of ``mesh`` whose sheet and normal fiber vectors are scaled by
    1/sqrt(ratio): with D = (D_l, D_l, D_l) the element tensor
    sum_p D_p f_p f_p^T (CardiacBody.cc:153-168, which uses the nine fiber
    numbers as given) becomes D_l f f^T + (D_l/ratio)(s s^T + n n^T), an
    anisotropy ratio sigma_l/sigma_t = ratio."""
    fib = np.array(mesh.fibers, dtype=np.float64, copy=True).reshape(-1, 9)
    fib[:, 3:] /= np.sqrt(ratio)
    return Mesh(xyz=mesh.xyz, elem_nn=mesh.elem_nn, conn=mesh.conn, fibers=fib,
                tissue=mesh.tissue, meta=dict(mesh.meta, anisotropy=ratio))


def concat(meshes) -> tuple:
    """Disjoint union of tet meshes: (mesh, instance id per node, node offset
    of each part).  Parts are placed 1 cm apart along x; no element is
    shared, so each part is an independent body."""
    xyz, elem_nn, conn, fib, tissue, inst, offs = [], [], [], [], [], [], []
    base, x0 = 0, 0.0
    for k, m in enumerate(meshes):
        p = np.array(m.xyz, dtype=np.float64, copy=True)
        p[:, 0] += x0 - p[:, 0].min()
        x0 = p[:, 0].max() + 1.0
        c = np.array(m.conn, dtype=np.int64, copy=True)
        c[c >= 0] += base
        xyz.append(p)
        elem_nn.append(m.elem_nn)
        conn.append(c.astype(np.int32))
        fib.append(np.asarray(m.fibers, dtype=np.float64).reshape(-1, 9))
        tissue.append(np.zeros(m.n_nodes, np.int8) if m.tissue is None else m.tissue)
        inst.append(np.full(m.n_nodes, k, np.int32))
        offs.append(base)
        base += m.n_nodes
    mesh = Mesh(xyz=np.vstack(xyz), elem_nn=np.concatenate(elem_nn), conn=np.vstack(conn),
                fibers=np.vstack(fib), tissue=np.concatenate(tissue),
                meta={"kind": "batch", "parts": len(meshes)})
    return mesh, np.concatenate(inst), np.array(offs + [base], np.int64)


def slab_sweep(count=96, n_min=5000, n_max=250000, nz=9, h=0.025, seed=4):
    """The batched conductance/anisotropy sweep: ``count`` slabs with node
    counts log-uniform in [n_min, n_max] (nx = ny, nz layers), each with its
    own G_Na, G_Kr, G_CaL scalings ~ U[0.5, 1.5] and anisotropy ratio
    ~ U[2, 10].  Returns (list of meshes, scales [count, 5] in the ten
    Tusscher scaling order gNa gKr gKs gCaL gto, ratios)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count))
    meshes, scales, ratios = [], [], []
    for n in sizes:
        side = max(4, int(round(np.sqrt(n / nz))))
        ratio = rng.uniform(2.0, 10.0)
        meshes.append(scale_cross_fiber(slab(side, side, nz, h), ratio))
        g = rng.uniform(0.5, 1.5, 3)
        scales.append([g[0], g[1], 1.0, g[2], 1.0])
        ratios.append(ratio)
    return meshes, np.array(scales), np.array(ratios)
