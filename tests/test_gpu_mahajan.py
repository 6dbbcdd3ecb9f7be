"""GPU parity of the Mahajan 2008 model (IonicModel/Mahajan.cc) against the
CPU oracle: independent cells of every tissue type through an action
potential (adaptive sub-steps included), the int-abs reading of the sub-step
test, and the whole CN step on a slab with per-node tissue types, in FP64 and
in mixed precision (FP32 gates/currents, FP64 Ca-cycling accumulators)."""
import numpy as np
import pytest

import oracle
from ucla_cmg_amd import meshgen
from ucla_cmg_amd import voomgpu as vg

pytestmark = pytest.mark.gpu


def _cells_vs_oracle(prec, quirks=0, steps=3000, dt=0.01, fail=False):
    types = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 11])  # 11: beyond 8, the fallback set
    n = len(types)
    cells = vg.IonicCells(vg.MODEL_MJ_FAIL if fail else vg.MODEL_MJ, n, prec)
    cells.set_tissue_types(types)
    if quirks:
        cells.set_quirks(quirks)
    so = np.array([oracle.mj_init(t) for t in types])
    cells.set_states(so)
    consts = [(oracle.mj_fail_constants if fail else oracle.mj_constants)(t) for t in types]
    Vg = so[:, 0].copy()
    Vo = so[:, 0].copy()
    xi = oracle.mj_xi()
    tg, to = [], []
    for k in range(steps):
        ist = 50.0 * xi if k * dt < 1.0 else 0.0
        dv = cells.compute_ion(1.0, dt, Vg, ist)
        Vg = Vg + dt * dv
        for i in range(n):
            Vo[i] += dt * oracle.mj_compute_ion(consts[i], so[i], 1.0, dt, Vo[i], ist,
                                                int_abs=bool(quirks & vg.QUIRK_MJ_INT_ABS))
        tg.append(Vg.copy())
        to.append(Vo.copy())
    return np.array(tg), np.array(to), cells.get_states(), so


def test_mahajan_cells_fp64_match_oracle(gpu_device):
    tg, to, sg, so = _cells_vs_oracle(vg.PREC_FP64)
    err = np.abs(tg - to).max()
    print(f"Mahajan cells fp64: peak {to.max():.2f} mV, max |dV| {err:.3e} mV")
    assert to.max() > 20.0
    assert err < 1e-8
    np.testing.assert_allclose(sg[:, 1:], so[:, 1:], rtol=1e-9, atol=1e-12)


def test_mahajan_fail_cells_fp64_match_oracle(gpu_device):
    """The failing-heart constant set (MAHAJAN_FAIL), every tissue type and
    the fallback type, through a whole (longer) action potential."""
    tg, to, sg, so = _cells_vs_oracle(vg.PREC_FP64, steps=45000, fail=True)
    err = np.abs(tg - to).max()
    ref, _, _, _ = _cells_vs_oracle(vg.PREC_FP64, steps=45000)
    print(f"Mahajan_fail cells fp64: max |dV| {err:.3e} mV; V at 300 ms fail "
          f"{to[30000].round(1)} normal {ref[30000].round(1)}")
    assert to.max() > 20.0 and err < 1e-7
    assert (to[30000] > ref[30000] + 20.0).all()  # still repolarising where normal cells rest
    np.testing.assert_allclose(sg[:, 1:], so[:, 1:], rtol=1e-8, atol=1e-11)


def test_mahajan_cells_int_abs_quirk(gpu_device):
    tg, to, sg, so = _cells_vs_oracle(vg.PREC_FP64, quirks=vg.QUIRK_MJ_INT_ABS, steps=600)
    assert np.abs(tg - to).max() < 1e-8
    np.testing.assert_allclose(sg[:, 1:], so[:, 1:], rtol=1e-9, atol=1e-12)


def test_mahajan_cells_mixed_track_oracle(gpu_device):
    """FP32 gates and currents with FP64 Ca-cycling/Na_i accumulators over a
    whole action potential: V within 1 mV RMS of the FP64 oracle, upstroke
    time within one step."""
    tg, to, sg, so = _cells_vs_oracle(vg.PREC_MIXED, steps=30000)
    rms = np.sqrt(((tg - to) ** 2).mean(axis=0))
    up_g = (tg > -40.0).argmax(axis=0)
    up_o = (to > -40.0).argmax(axis=0)
    print(f"Mahajan cells mixed: V RMS per type {np.round(rms, 3)}, Ca_i {sg[:, 9]} vs {so[:, 9]}")
    assert rms.max() < 1.0
    assert np.abs(up_g - up_o).max() <= 1
    np.testing.assert_allclose(sg[:, 19], so[:, 19], rtol=1e-4)  # Na_i (FP64 accumulator)


D_ISO = (1e-3 / 1.0, 1e-3, 1e-3)


def _slab(nx=15, ny=9, nz=3):
    mesh = meshgen.slab(nx, ny, nz, 0.025)
    types = (np.arange(mesh.n_nodes) % 9).astype(np.int8)
    mesh.tissue = types
    return mesh, types, meshgen.slab_corner_nodes(nx, ny, nz, 3)


def _tissue_pair(prec, steps, cg_tol, dt=0.01, fail=False):
    mesh, types, stim = _slab()
    gpu = vg.CardiacSolver(mesh, cell_model=vg.MODEL_MJ_FAIL if fail else vg.MODEL_MJ,
                           precision=prec, dt=dt, D=D_ISO, cg_tol=cg_tol)
    ref = oracle.Simulation(mesh, oracle.MODEL_MJ, dt=dt, D=D_ISO, cg_tol=1e-10,
                            param_idx=types.astype(np.uint16),
                            mj_const=oracle.mj_table(fail) if fail else None)
    np.testing.assert_allclose(gpu.getVoltage(), ref.V, rtol=0,
                               atol=0.0 if prec == vg.PREC_FP64 else 1e-4)
    gpu.enableActivation(-40.0)
    ist = 50.0 * oracle.mj_xi()
    gpu.setStimulusCurrent(stim, ist)
    ref.istim[stim] = ist
    for k in range(steps):
        if k == 200:
            gpu.clearStimulus()
            ref.istim[:] = 0.0
        gpu.Solve(1)
        ref.step()
    return gpu, ref, mesh


def test_mahajan_slab_fp64_matches_oracle(gpu_device):
    gpu, ref, mesh = _tissue_pair(vg.PREC_FP64, 600, 1e-10)
    err = np.abs(gpu.getVoltage() - ref.V).max()
    at_g, at_o = gpu.activationTimes(), ref.act
    st = gpu.getStateVariables()
    print(f"Mahajan slab fp64: max|dV| {err:.3e} mV, activated {(at_o >= 0).mean():.2f}")
    assert (at_o >= 0).mean() > 0.5
    assert err < 1e-6
    np.testing.assert_array_equal(at_g >= 0, at_o >= 0)
    both = at_o >= 0
    assert np.abs(at_g[both] - at_o[both]).max() < 1e-6
    np.testing.assert_allclose(st[:, 1:], ref.states[:, 1:], rtol=1e-6, atol=1e-9)


def test_mahajan_fail_slab_fp64_matches_oracle(gpu_device):
    gpu, ref, mesh = _tissue_pair(vg.PREC_FP64, 600, 1e-10, fail=True)
    err = np.abs(gpu.getVoltage() - ref.V).max()
    print(f"Mahajan_fail slab fp64: max|dV| {err:.3e} mV, activated {(ref.act >= 0).mean():.2f}")
    assert (ref.act >= 0).mean() > 0.5 and err < 1e-6


def test_mahajan_slab_mixed_within_tolerance(gpu_device):
    gpu, ref, mesh = _tissue_pair(vg.PREC_MIXED, 600, 1e-7)
    at_g, at_o = gpu.activationTimes(), ref.act
    t_cut = 600 * 0.01 - 0.5
    early = ((at_o >= 0) & (at_o < t_cut)) | ((at_g >= 0) & (at_g < t_cut))
    assert early.mean() > 0.3
    assert (at_o[early] >= 0).all() and (at_g[early] >= 0).all()
    d = np.abs(at_g[early] - at_o[early])
    rms = np.sqrt(((gpu.getVoltage() - ref.V) ** 2).mean())
    print(f"Mahajan slab mixed: max |dAT| {d.max():.4f} ms, V RMS {rms:.4f} mV")
    assert d.max() <= 0.25
    assert rms <= 0.5
