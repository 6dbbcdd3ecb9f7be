"""The oracle's LUORUDYWITHGATE (LuoRudyGateTable, LRGateTable.h; its use,
LR.cc:71-79), pinned against the plain Luo-Rudy restatement: the table holds
make_parameter1 at -90 + 0.001 i, and between grid points each of the 14
values is interpolated linearly; outside (-90, 200) mV the formulas run."""
import numpy as np
import pytest

import oracle


def _grid(i):
    return -90 + 0.001 * i


@pytest.mark.parametrize("i", [1, 5000, 50000, 90000, 95000, 150000, 289999])
def test_table_equals_the_formulas_on_the_grid(i):
    v = _grid(i)
    np.testing.assert_array_equal(oracle.lr_tab(v, gate=True), oracle.lr_tab(v))


@pytest.mark.parametrize("v", [-90.0, -95.0, 200.0, 230.0])
def test_outside_the_table_the_formulas_run(v):
    np.testing.assert_array_equal(oracle.lr_tab(v, gate=True), oracle.lr_tab(v))


@pytest.mark.parametrize("v", [-84.99937, -20.12345, 10.0004, 155.55555])
def test_between_grid_points_the_values_are_linear(v):
    i = int((v + 90) * 1000)
    lo, hi = oracle.lr_tab(_grid(i)), oracle.lr_tab(_grid(i + 1))
    th = 1000 * (v - _grid(i))
    assert 0.0 < th < 1.0
    np.testing.assert_allclose(oracle.lr_tab(v, gate=True), lo + th * (hi - lo), rtol=1e-12,
                               atol=1e-300)


def test_the_table_tracks_the_formulas_within_the_interpolation_error():
    """|f(v) - linear interpolation| <= h^2/8 max|f''| is a few 1e-7 relative
    for these smooth functions (h = 0.001 mV); across the -40 mV switch of the
    h and j rates the interpolation straddles a jump."""
    rng = np.random.default_rng(4)
    for v in rng.uniform(-89.0, 150.0, 400):
        if abs(v + 40.0) < 0.002:
            continue
        a, b = oracle.lr_tab(v), oracle.lr_tab(v, gate=True)
        assert np.all(np.abs(a - b) <= 1e-6 * np.abs(a) + 1e-12), v


def test_gate_variant_action_potential_matches_the_plain_model():
    def run(fn):
        s = oracle.lr_init().copy()
        V, tr = s[7], []
        for k in range(30000):
            dv = fn(s, 1.0, 0.01, V, 50.0 * oracle.lr_xi() if k < 200 else 0.0)
            V += 0.01 * dv
            tr.append(V)
        return np.array(tr)
    a, b = run(oracle.lr_compute_ion), run(oracle.lr_gate_compute_ion)
    assert a.max() > 20.0 and np.abs(a - b).max() < 1e-2
