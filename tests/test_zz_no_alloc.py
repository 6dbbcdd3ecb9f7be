"""vg_ecg_device, the ECG entry point that allocates nothing, against vg_ecg.

Both run the same two kernels (per-block partial sums of H V, then one block
summing them in a fixed order) on the same grid; vg_ecg_device keeps its
partials in the library's file-scope device array and writes the leads to
the caller's device memory.  The results must be bit-identical, on two calls
in a row whose second handle is smaller (fewer nodes, blocks and leads), so
slots the first call filled and the second does not use cannot leak in.
"""
import ctypes as C

import numpy as np
import pytest

from ucla_cmg_amd import meshgen
from ucla_cmg_amd import voomgpu as vg

pytestmark = pytest.mark.gpu

D_SLAB = (8.571e-4, 1.429e-4, 1.429e-4)


def _solver(nx, ny, nz, precision, ne):
    mesh = meshgen.slab(nx, ny, nz, 0.025)
    s = vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, precision=precision,
                         integrator=vg.INTEG_RUSH_LARSEN, D=D_SLAB, dt=0.02, cg_tol=1e-8)
    c = mesh.xyz.mean(axis=0)
    s.setupECG(meshgen.electrodes_around(c, 2.0, ne, axis_z=0.5))
    s.setStimulusCurrent(meshgen.slab_corner_nodes(nx, ny, nz, 3), 50.0 * 2000.0)
    s.Solve(60)  # a wave in the slab, so the leads are not all alike
    return s


@pytest.mark.parametrize("precision", [vg.PREC_FP64, vg.PREC_MIXED])
def test_ecg_device_equals_ecg_on_two_calls(gpu_device, precision):
    import torch
    big = _solver(41, 41, 9, precision, 10)
    small = _solver(9, 9, 3, precision, 3)
    out = torch.empty(16, dtype=torch.float64, device="cuda")
    for s, ne in ((big, 10), (small, 3)):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        print(f"vg_ecg_device: precision {precision}, {ne} leads", flush=True)
        rc = s.L.vg_ecg_device(s.h, C.c_void_p(out.data_ptr()))
        assert rc == vg.VG_OK, s.L.vg_last_error(s.h)
        s.synchronize()
        got = out.cpu().numpy()
        ref = s.computeECG()
        assert ref.shape == (ne,) and np.isfinite(ref).all() and np.ptp(ref) > 0
        np.testing.assert_array_equal(got[:ne], ref)
        assert np.isnan(got[ne:]).all()  # only the handle's leads are written
    big.close()
    small.close()


def test_ecg_device_without_ecg_setup_is_an_error(gpu_device):
    import torch
    mesh = meshgen.slab(9, 9, 3, 0.025)
    s = vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, D=D_SLAB, dt=0.02)
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    assert s.L.vg_ecg_device(s.h, C.c_void_p(out.data_ptr())) == vg.VG_ERR_ARG
    assert b"ECG not set up" in s.L.vg_last_error(s.h)
    assert s.L.vg_ecg_device(s.h, None) == vg.VG_ERR_ARG
    s.close()
