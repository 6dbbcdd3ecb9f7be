"""Helpers of the measurement modes, checked on closed forms."""
import numpy as np

import bench


def _spiral_frames(centres, shape=(30, 40), vstar=-40.0):
    y, x = np.mgrid[0:shape[0], 0:shape[1]].astype(float)
    phi = np.zeros(shape)
    for cy, cx, sign in centres:
        phi += sign * np.arctan2(y - cy, x - cx)
    return vstar + np.sin(phi), vstar + np.cos(phi)


def test_phase_singularity_count_of_known_fields():
    Vt, Vlag = _spiral_frames([(10.3, 12.7, 1)])
    assert bench.phase_singularities(Vt, Vlag) == (1, 1)
    Vt, Vlag = _spiral_frames([(10.3, 12.7, 1), (20.6, 30.2, -1)])
    assert bench.phase_singularities(Vt, Vlag) == (2, 0)
    Vt, Vlag = _spiral_frames([(10.3, 12.7, -1), (20.6, 30.2, -1), (5.5, 33.5, -1)])
    assert bench.phase_singularities(Vt, Vlag) == (3, -3)
    flat = np.full((30, 40), -85.0)  # resting tissue: no singularity
    assert bench.phase_singularities(flat, flat) == (0, 0)
