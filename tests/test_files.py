"""The reference's file formats on the CPU: mesh files (Mesh::New,
Mesh/Mesh.cc:142-416), frame names (EPSolver.cc:121,159), the host lead
field against the oracle's direct ECG, and the ComputeECG post-processor
(ComputeECG.cc) run on the reference's own parameter file."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from ucla_cmg_amd import meshgen
from ucla_cmg_amd import voomgpu as vg

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
# $VOOMGPU_TOOLS names another build of the two drivers (the AddressSanitizer
# one of the Makefile's asan target, say).
TOOLS = os.environ.get("VOOMGPU_TOOLS") or os.path.join(os.path.dirname(HERE), "ucla_cmg_amd")
ECG_TOOL = os.path.join(TOOLS, "voomgpu_ecg")
EP_TOOL = os.path.join(TOOLS, "voomgpu_ep")


def test_reads_the_hand_written_reference_files():
    m = vg.read_mesh(os.path.join(GOLDEN, "mesh_small", "small"), radius=0.04)
    assert m.xyz.shape == (6, 3) and m.xyz[5, 0] == -0.05
    np.testing.assert_array_equal(m.elem_nn, [4, 4, 2])
    np.testing.assert_array_equal(m.conn[0], [0, 1, 2, 3])
    np.testing.assert_array_equal(m.conn[1], [1, 2, 3, 4])
    np.testing.assert_array_equal(m.conn[2, :2], [5, 0])
    assert m.fibers.shape == (3, 9) and m.fibers[1, 1] == 1.0
    # .purk makes node 6 type 11; .cell then sets 2 -> 4, 3 -> 8, 6 -> 12;
    # the trailing half pair "5" is ignored.
    np.testing.assert_array_equal(m.tissue, [0, 4, 8, 0, 0, 12])
    np.testing.assert_array_equal(m.pmj_node, [5])
    np.testing.assert_array_equal(m.pmj_data, [[0.0, 0.0, 1.0, 0.25]])
    assert m.radius == 0.04


def _same(a, b):
    # Bars use the first two connectivity slots only.
    used = np.arange(4)[None, :] < np.asarray(a.elem_nn)[:, None]
    np.testing.assert_array_equal(np.asarray(a.conn)[used], np.asarray(b.conn)[used])
    for k in ("xyz", "elem_nn", "fibers", "tissue", "pmj_node", "pmj_data"):
        x, y = getattr(a, k), getattr(b, k)
        if x is None or y is None:
            assert x is None and y is None, k
        else:
            np.testing.assert_array_equal(np.asarray(x), np.asarray(y), err_msg=k)


def test_write_read_round_trip_is_exact(tmp_path):
    slab = meshgen.scale_cross_fiber(meshgen.slab(7, 6, 3, 0.025), 3.0)
    vg.write_mesh(str(tmp_path / "slab"), slab)
    _same(slab, vg.read_mesh(str(tmp_path / "slab"), radius=slab.radius))
    assert (tmp_path / "slab.metis.npart.1").read_text().split() == ["0"] * slab.n_nodes
    lv = meshgen.purkinje_tree(meshgen.lv_shell(n_trans=4, n_circ=16, n_long=8),
                               generations=4, pmj_length=meshgen.pmj_length(500.0))
    lv.tissue = np.where(lv.tissue == 11, 11, 3).astype(np.int8)  # some .cell types too
    vg.write_mesh(str(tmp_path / "lv"), lv)
    back = vg.read_mesh(str(tmp_path / "lv"), radius=lv.radius)
    _same(lv, back)
    assert (back.tissue == 11).sum() > 0 and back.pmj_node.size > 0


def test_reader_errors_name_the_problem(tmp_path):
    (tmp_path / "h.node").write_text("8 3 0 0\n" + "".join(f"{i+1} 0 0 {i}\n" for i in range(8)))
    (tmp_path / "h.ele").write_text("1 0 0\n1 8 1 2 3 4 5 6 7 8\n")
    with pytest.raises(vg.VoomError, match="tetrahedra"):
        vg.read_mesh(str(tmp_path / "h"))
    with pytest.raises(vg.VoomError, match="cannot open"):
        vg.read_mesh(str(tmp_path / "missing"))
    (tmp_path / "b.node").write_text("4 3 0 0\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n")
    (tmp_path / "b.ele").write_text("1 0 0\n1 4 1 2 3 9\n")
    with pytest.raises(vg.VoomError, match="bad node id"):
        vg.read_mesh(str(tmp_path / "b"))


@pytest.mark.parametrize("t,rank", [(0.0, 0), (12.345, 3), (1234.5, 31), (99999.0, 0)])
def test_frame_names_follow_the_reference_format(t, rank):
    assert vg.frame_name("out", "Voltage", t, rank) == "out/Voltage_time_%08.2f.%02.0f" % (t, rank)
    assert vg.frame_name("o", "InterVariables", 1.0, 0) == "o/InterVariables_time_00001.00.00"


def test_host_lead_field_equals_the_oracle_direct_ecg():
    mesh = meshgen.purkinje_tree(meshgen.lv_shell(n_trans=4, n_circ=16, n_long=8), generations=3,
                                 pmj_length=meshgen.pmj_length(500.0))
    el = np.array([[4.0, 0.5, -2.0], [0.0, -5.0, 1.0], [0.2, 0.1, -9.0]])
    H = vg.ecg_leadfield(mesh, el)
    rng = np.random.default_rng(3)
    for _ in range(3):
        V = rng.uniform(-90, 30, mesh.n_nodes)
        ref = oracle.ecg_direct(mesh, V, el)
        np.testing.assert_allclose(H @ V, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_compute_ecg_cli_on_the_reference_parameter_file(tmp_path):
    """The reference's own ECG.input (6 electrodes, 32 partitions, binary
    frames) with a mesh and frames laid out where it points."""
    assert os.path.exists(ECG_TOOL), "build the tools first (make -C ucla_cmg_amd/csrc)"
    shutil.copy(os.path.join(GOLDEN, "decks", "ECG.input"), tmp_path / "ECG.input")
    toks = (tmp_path / "ECG.input").read_text().split()
    model, meshdir, vdir, odir = toks[:4]
    el = np.array(toks[8:8 + 18], float).reshape(6, 3)
    nparts = int(toks[26])
    # A slab near the electrodes' region.
    mesh = meshgen.slab(9, 8, 4, 0.1)
    mesh.xyz = mesh.xyz + np.array([0.5, 2.0, 0.5])
    os.makedirs(tmp_path / meshdir)
    os.makedirs(tmp_path / vdir)
    os.makedirs(tmp_path / odir)
    vg.write_mesh(str(tmp_path / meshdir / model), mesh)
    rng = np.random.default_rng(5)
    part = rng.integers(0, nparts, mesh.n_nodes)
    (tmp_path / meshdir / f"{model}.metis.npart.{nparts}").write_text(
        "\n".join(str(p) for p in part) + "\n")
    frames = []
    for step in (0.0, 1.0, 2.0):
        V = rng.uniform(-90, 30, mesh.n_nodes)
        frames.append(V)
        for p in range(nparts):
            V[part == p].astype(np.float64).tofile(
                tmp_path / vdir / ("Voltage_time_%08.2f.%02.0f" % (step, p)))
    r = subprocess.run([ECG_TOOL, "ECG.input"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split("\t") for l in (tmp_path / odir / "ECG_Data.00").read_text().splitlines()]
    assert len(rows) == 3  # stops at the first missing frame (step 3)
    for k, row in enumerate(rows):
        assert float(row[0]) == k
        got = np.array([float(v) for v in row[1:7]])
        ref = oracle.ecg_direct(mesh, frames[k], el)
        np.testing.assert_allclose(got, ref, rtol=1e-5)  # %g: six digits


def test_ep_driver_rejects_bad_decks_before_touching_a_gpu(tmp_path):
    assert os.path.exists(EP_TOOL)
    deck = tmp_path / "d.input"
    deck.write_text("# no model\nCellModel Tusscher\n")
    r = subprocess.run([EP_TOOL, str(deck)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no MODEL" in r.stderr
    deck.write_text("Model m\nCellModel FitzHughNagumo\n")
    r = subprocess.run([EP_TOOL, str(deck)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not supported" in r.stderr
    deck.write_text(f"Model {tmp_path}/nothing\nCellModel Tusscher\n")
    r = subprocess.run([EP_TOOL, str(deck)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
