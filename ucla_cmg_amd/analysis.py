"""Analysis of monodomain runs: phase singularities of re-entry.

The reference has none (its ECG and activation post-processing live outside
the solver, `Applications/CardiacEP/ComputeECG.cc`); the S1-S2 workload asks
for a phase-singularity count, defined here.
"""
from __future__ import annotations

import numpy as np


def phase_singularities(Vt, Vlag, vstar=-40.0):
    """Phase singularities of a surface frame on a structured grid: the phase
    comes from the time-delay embedding theta = atan2(V(t) - V*, V(t - tau) -
    V*), and the topological charge of each grid plaquette is the winding of
    theta around it / 2 pi.  Vt and Vlag are 2-D arrays of the same grid.
    Returns (number of plaquettes with nonzero charge, net charge)."""
    th = np.arctan2(np.asarray(Vt) - vstar, np.asarray(Vlag) - vstar)

    def wrap(d):
        return (d + np.pi) % (2.0 * np.pi) - np.pi

    a, b, c, d = th[:-1, :-1], th[:-1, 1:], th[1:, 1:], th[1:, :-1]
    q = np.rint((wrap(b - a) + wrap(c - b) + wrap(d - c) + wrap(a - d)) / (2.0 * np.pi))
    return int(np.count_nonzero(q)), int(q.sum())
