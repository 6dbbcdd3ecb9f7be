"""Pins of the oracle's cell models (IonicModel/LR.cc, IonicModel/Tusscher.cc)
by invariants, closed forms and the reference's own numeric tables."""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ionic_tables.json")))

# ten Tusscher constant indices (SetUpIonicConstants.h:93-145)
F, CM_CELL, VC, VSR, VSS = 2, 3, 4, 51, 52


def test_tables_match_reference():
    np.testing.assert_array_equal(oracle.tt_constants(), GOLD["tusscher_constants"])
    np.testing.assert_array_equal(oracle.tt_init(), GOLD["tusscher_initial_state"])
    np.testing.assert_array_equal(oracle.lr_init(), GOLD["luorudy_initial_state"])
    r, L = 11e-4, 100e-4
    assert oracle.lr_xi() == 2.0 * (r + L) / (r * L)


def _tt_charge_balance(c, s, alg12):
    """Cm dV/dt = F [Vc (dK_i + dNa_i) + 2 (Vc dCa_i,tot + Vsr dCa_SR,tot +
    Vss dCa_ss,tot)] / Cm_cell, with total = free / buffer factor.  Holds
    exactly for the tT2006 equations (every current enters both sides)."""
    alg, R = oracle.tt_rates(c, s, alg12)
    bi, bsr, bss = alg[63], alg[68], alg[69]
    rhs = c[F] / c[CM_CELL] * (c[VC] * (R[1] + R[2]) +
                                2.0 * (c[VC] * R[3] / bi + c[VSR] * R[17] / bsr + c[VSS] * R[10] / bss))
    return R[0], rhs


@pytest.mark.parametrize("V", [-85.0, -60.0, -20.0, 0.0, 14.3, 30.0])
@pytest.mark.parametrize("stim", [0.0, -52.0])
def test_tt_charge_conservation(V, stim):
    c = oracle.tt_constants()
    rng = np.random.default_rng(int(abs(V) * 10) + 7)
    s = oracle.tt_init() * (1.0 + 0.05 * rng.standard_normal(19))
    s[0] = V
    lhs, rhs = _tt_charge_balance(c, s, stim)
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(lhs))


def test_tt_charge_conservation_catches_a_dropped_current():
    """The balance is sensitive: removing one current from dV/dt breaks it."""
    c = oracle.tt_constants()
    s = oracle.tt_init()
    s[0] = -20.0
    alg, R = oracle.tt_rates(c, s, 0.0)
    lhs, rhs = _tt_charge_balance(c, s, 0.0)
    assert abs((lhs + alg[53]) - rhs) > 1e-3  # I_bCa dropped from dV/dt


def test_tt_initial_state_is_near_rest():
    """The reference's initial table is a paced resting state: every rate is
    small compared with the state's own scale per ms."""
    c = oracle.tt_constants()
    s = oracle.tt_init()
    alg, R = oracle.tt_rates(c, s, 0.0)
    assert abs(R[0]) < 0.1  # mV/ms; an action potential upstroke is ~300 mV/ms
    conc = [1, 2, 3, 10, 17, 18]
    assert np.all(np.abs(R[conc]) / np.abs(s[conc]) < 0.05)
    # Each gate sits near its steady state at the resting V; the slow L-type
    # inactivation f is still recovering in this paced state (0.79 vs 1.0).
    gates = [4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16]
    assert np.all(np.abs(s[gates] - alg[:12]) < 0.25)
    fast = [0, 1, 2, 3, 4, 5, 6, 9, 10, 11]
    assert np.all(np.abs(s[gates][fast] - alg[:12][fast]) < 0.05)


def test_tt_backward_euler_gates_closed_form():
    """At fixed V the 12 gates obey x1 = (x0 + dt x_inf/tau)/(1 + dt/tau)."""
    c = oracle.tt_constants()
    s = oracle.tt_init()
    s[0] = -30.0
    alg, _ = oracle.tt_rates(c, s, 0.0)
    gates = [4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16]
    dind = [34, 35, 36, 37, 38, 39, 42, 20, 21, 22, 23, 24]
    dt = 0.02
    expect = np.array([(s[g] + dt * alg[i] / alg[d]) / (1 + dt / alg[d])
                       for i, (g, d) in enumerate(zip(gates, dind))])
    s1 = s.copy()
    oracle.tt_compute_ion(c, s1, 1.0, dt, -30.0, 0.0)
    np.testing.assert_allclose(s1[gates], expect, rtol=1e-13)


def test_tt_newton_solves_backward_euler_system():
    """After Compute_Ion the six concentrations satisfy x - x^n - dt f(x) = 0
    (to the Newton step tolerance), with f evaluated at the new gates."""
    c = oracle.tt_constants()
    s0 = oracle.tt_init()
    for V, stim in [(-85.423, 0.0), (-10.0, 0.0), (20.0, 104000.0)]:
        s = s0.copy()
        dt = 0.02
        dv, it = oracle.tt_compute_ion(c, s, 1.0, dt, V, stim)
        assert it >= 1
        _, R = oracle.tt_rates(c, s, -stim / (1.0 * 2000.0))
        nr = [1, 2, 3, 10, 17, 18]
        g = s[nr] - s0[nr] - dt * R[nr]
        assert np.abs(g).max() < 1e-7, (V, g)
        assert dv == pytest.approx(R[0], rel=1e-12, abs=1e-12)


def test_tt_newton_step_solves_the_transposed_jacobian():
    """The reference hands a row-major Jacobian to column-major dgesv, so each
    Newton step solves J^T d = g (Tusscher.cc:172-181)."""
    c = oracle.tt_constants()
    s = oracle.tt_init()
    s[0] = -40.0
    nr = [1, 2, 3, 10, 17, 18]
    stateN = s[nr].copy()
    s[nr] *= 1.01
    _, R = oracle.tt_rates(c, s, 0.0)
    J, B, _ = oracle.tt_newton_system(c, s, stateN, R[nr], 0.02)
    Jm = J.reshape(6, 6)  # Jm[j, i] = d g_j / d x_i
    d, info = oracle.solve6_lapack_layout(J, B)
    assert info == 0
    np.testing.assert_allclose(d, np.linalg.solve(Jm.T, B), rtol=1e-10, atol=1e-14)
    # and it is not the untransposed solve
    assert np.abs(np.linalg.solve(Jm, B) - d).max() > 1e-6 * np.abs(d).max()
    dn, info = oracle.solve6(J, B, transpose=False)
    assert info == 0
    np.testing.assert_allclose(dn, np.linalg.solve(Jm, B), rtol=1e-10, atol=1e-14)


def test_tt_transposed_newton_leaves_physical_range_in_upstroke():
    """The reference's J^T Newton drives Ca_i negative a few ms into a paced
    action potential at dt = 0.02 ms (then log(Ca_o/Ca_i) is NaN), while the
    true Newton step converges to the backward-Euler solution.  This is why
    the whole-step oracle and the GPU default to J and keep J^T as a quirk."""
    c = oracle.tt_constants()
    res = {}
    for jt in (False, True):
        s = oracle.tt_init()
        V = s[0]
        for k in range(1000):
            dv, _ = oracle.tt_compute_ion(c, s, 1.0, 0.02, V, 52.0 * 2000.0 if k < 50 else 0.0, jt)
            V += 0.02 * dv
        res[jt] = V
    assert np.isfinite(res[False]) and res[False] > 0.0  # plateau at 20 ms
    assert not np.isfinite(res[True])


def test_tt_action_potential_markers():
    """Paced from the reference's initial state with a 1 ms, 52 pA/pF
    stimulus: an overshooting upstroke and a plateau lasting hundreds of ms
    (tT2006 epicardial APD90 is about 270-310 ms)."""
    c = oracle.tt_constants()
    s = oracle.tt_init()
    V = s[0]
    dt = 0.02
    trace = []
    for k in range(30000):
        st = 52.0 * 2000.0 if k < 50 else 0.0
        dv, _ = oracle.tt_compute_ion(c, s, 1.0, dt, V, st)
        V += dt * dv
        trace.append(V)
    trace = np.array(trace)
    t = np.arange(1, 30001) * dt
    assert 20.0 < trace.max() < 50.0
    peak = trace.max()
    rest = -85.423
    thr = rest + 0.1 * (peak - rest)
    up = np.argmax(trace > -40)
    down = up + np.argmax(trace[up + 100:] < thr) + 100
    apd = t[down] - t[up]
    assert 200.0 < apd < 400.0, apd
    assert abs(trace[-1] - rest) < 1.0


def test_lr_rush_larsen_and_calcium_closed_form():
    """LR gates follow x1 = x_inf - (x_inf - x0) exp(-dt (alpha+beta)) at fixed
    V; Ca_i follows its forward-Euler update with i_si at the new gates."""
    s = oracle.lr_init()
    V, dt = -50.0, 0.01
    s0 = s.copy()
    oracle.lr_compute_ion(s, 1.0, dt, V, 0.0)
    # After many steps at fixed V the gates converge to their steady state:
    s_inf = s0.copy()
    for _ in range(200000):
        oracle.lr_compute_ion(s_inf, 1.0, dt, V, 0.0)
    # Closed form: x1 - x_inf = (x0 - x_inf) * exp(-dt k) with k independent
    # of x, so two steps give a geometric ratio.
    s1, s2 = s.copy(), s.copy()
    oracle.lr_compute_ion(s2, 1.0, dt, V, 0.0)
    for i in range(6):
        r1 = (s1[i] - s_inf[i]) / (s0[i] - s_inf[i])
        r2 = (s2[i] - s_inf[i]) / (s1[i] - s_inf[i])
        assert r1 == pytest.approx(r2, rel=1e-6, abs=1e-9)
        assert 0.0 < r1 < 1.0


def test_lr_upstroke_and_rest():
    s = oracle.lr_init()
    V = -84.0
    dt = 0.01
    for _ in range(100000):
        V += dt * oracle.lr_compute_ion(s, 1.0, dt, V, 0.0)
    rest = V
    assert -86.0 < rest < -83.0  # LR1991 rests near -84 mV
    peak = rest
    for k in range(50000):
        st = 50.0 * oracle.lr_xi() if k < 200 else 0.0
        V += dt * oracle.lr_compute_ion(s, 1.0, dt, V, st)
        peak = max(peak, V)
    assert peak > 20.0
    assert abs(V - rest) < 1.0
