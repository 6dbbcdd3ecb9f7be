"""Pins of the oracle's Purkinje cell model (Corrias rabbit Purkinje fibre as
exported in IonicModel/Purkinje.cc; constants SetUpIonicConstants.h:16-85;
initial state, Xi and cell area Purkinje.h:44-53).

What fixes each part independently of how the oracle writes it:
* the constant table, initial state, Xi and cell area against the numbers
  tests/golden/make_ionic_tables.py reads from the reference text, and the
  derived constants (F/RT, RT/F, compartment volumes) against their stated
  definitions;
* each logistic steady state x_inf = 1/(1 + exp((V + a)/b)) (centre a and
  slope b read from Purkinje.cc by the same script) by two properties that
  hold for every V: x_inf(-a) = 1/2 and d/dV logit(x_inf) = -1/b;
* the voltage-dependent time constants at the points where they reduce to
  constants, their limits, and the two guarded removable singularities of
  the I_Kr gate (V = -7, -10 mV) and the guarded one of the I_Ks gate
  (V = -30 mV) against the formula's limit there;
* charge conservation: the membrane current equals the charge that the ion
  concentration rates move (every current enters V and its ion's balance,
  the exchanger and pump with their stoichiometry, the SR fluxes cancel);
* forward Euler on all 22 states with ten sub-steps of dt/10 above
  |dV/dt| = 1.5 mV/ms (Purkinje.cc:179-192), the stimulus conversion
  -1e6 istim area / Xi (:200-205);
* action-potential markers of a stimulated cell.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ionic_tables.json")))
VGRID = np.linspace(-110.0, 50.0, 161)


def _rates(V, **over):
    c = oracle.pk_constants()
    s = oracle.pk_init()
    s[0] = V
    for k, v in over.items():
        s[int(k[1:])] = v
    return oracle.pk_rates(c, s)


def inf_tau(V, g, **over):
    """Steady state and time constant of gate g at V, read off the rate at
    x = 0 and x = 1 (every gate's rate is (x_inf - x)/tau)."""
    r0 = _rates(V, **dict(over, **{f"s{g}": 0.0}))[1][g]
    r1 = _rates(V, **dict(over, **{f"s{g}": 1.0}))[1][g]
    tau = 1.0 / (r0 - r1)
    return r0 * tau, tau


def test_tables_match_reference():
    c = oracle.pk_constants()
    lit = GOLD["purkinje_constants_literal"]
    np.testing.assert_array_equal(c[:len(lit)], lit)
    T, F, R = lit[0], lit[1], lit[2]
    vol = lit[7]
    np.testing.assert_allclose(c[57:63], [F / (R * T), R * T / F, lit[9] * vol, lit[8] * vol,
                                          (lit[8] + lit[9]) * vol, lit[10] * vol], rtol=1e-15)
    np.testing.assert_array_equal(oracle.pk_init(), GOLD["purkinje_initial_state"])


@pytest.mark.parametrize("g", sorted(int(k) for k in GOLD["purkinje_logistic_gates"]))
def test_logistic_steady_states(g):
    a, b = GOLD["purkinje_logistic_gates"][str(g)]
    x_half, _ = inf_tau(-a, g)
    assert x_half == pytest.approx(0.5, abs=1e-12)
    for V in (-95.0, -60.0, -20.0, 10.0):
        h = 1e-3
        lo, hi = inf_tau(V - h, g)[0], inf_tau(V + h, g)[0]
        slope = (np.log(hi / (1 - hi)) - np.log(lo / (1 - lo))) / (2 * h)
        assert slope == pytest.approx(-1.0 / b, rel=1e-6), (g, V)


def test_other_steady_states():
    # State 16: 0.1 + 0.9 / (1 + exp((V + 75.6)/6.3)) (Purkinje.cc:25).
    assert inf_tau(-75.6, 16)[0] == pytest.approx(0.55, abs=1e-12)
    assert inf_tau(-300.0, 16)[0] == pytest.approx(1.0, abs=1e-12)
    assert inf_tau(300.0, 16)[0] == pytest.approx(0.1, abs=1e-12)
    # State 10 (Ca-dependent inactivation): 0.4 + 0.6/(1 + (Ca/1e-4)^2) (:22).
    assert inf_tau(-80.0, 10, s1=1e-4)[0] == pytest.approx(0.7, abs=1e-12)
    assert inf_tau(-80.0, 10, s1=0.0)[0] == pytest.approx(1.0, abs=1e-12)
    assert inf_tau(-80.0, 10, s1=1.0)[0] == pytest.approx(0.4, abs=1e-7)
    for g in range(6, 22):
        xi = np.array([inf_tau(V, g)[0] for V in VGRID])
        assert np.all(xi >= 0.0) and np.all(xi <= 1.0), g


def test_time_constants():
    c = oracle.pk_constants()
    # Constant time constants (C[16], C[18], C[20], C[21], C[24], C[25], C[27]).
    for g, k in ((6, 16), (8, 18), (11, 20), (12, 21), (13, 24), (14, 25), (15, 27)):
        for V in (-90.0, -30.0, 20.0):
            assert inf_tau(V, g)[1] == pytest.approx(c[k], rel=1e-12), g
    # State 7: 1.42271 exp(-0.05119 V) (:16): 1.42271 at 0, log-slope -0.05119.
    assert inf_tau(0.0, 7)[1] == pytest.approx(1.42271, rel=1e-12)
    t1, t2 = inf_tau(-50.0, 7)[1], inf_tau(-40.0, 7)[1]
    assert np.log(t2 / t1) / 10.0 == pytest.approx(-0.05119, rel=1e-9)
    # State 9: 25.1 / (0.04 + 0.7 exp(-(0.028 (V + 14.5))^2)) (:20).
    assert inf_tau(-14.5, 9)[1] == pytest.approx(25.1 / 0.74, rel=1e-12)
    assert inf_tau(-900.0, 9)[1] == pytest.approx(25.1 / 0.04, rel=1e-9)
    # State 10: 2 + 80 / (1 + (Ca/1e-4)^2) (:23).
    assert inf_tau(-80.0, 10, s1=1e-4)[1] == pytest.approx(42.0, rel=1e-12)
    # State 16: 120 + exp((V + 100)/25) (:26).
    assert inf_tau(-100.0, 16)[1] == pytest.approx(121.0, rel=1e-12)
    # States 20 and 21: 6000 / (exp(-(2.9 + 0.04 V)) + exp(3.6 + 0.11 V)),
    # equal terms at V = -6.5/0.15 mV.
    v0 = -6.5 / 0.15
    for g in (20, 21):
        assert inf_tau(v0, g)[1] == pytest.approx(3000.0 / np.exp(-(2.9 + 0.04 * v0)), rel=1e-12)
    # State 19's time constant is 4x state 18's (:47-48); both share x_inf.
    for V in (-80.0, -20.0, 15.0):
        x18, t18 = inf_tau(V, 18)
        x19, t19 = inf_tau(V, 19)
        assert x19 == pytest.approx(x18, rel=1e-14) and t19 == pytest.approx(4 * t18, rel=1e-12)
    for g in range(6, 22):
        taus = np.array([inf_tau(V, g)[1] for V in VGRID])
        assert np.all(np.isfinite(taus)) and np.all(taus > 0.0), g


@pytest.mark.parametrize("g, V0", [(18, -30.0), (17, -7.0)])
def test_guarded_singularities_equal_their_limits(g, V0):
    """The guarded branches replace 0/0 by the formula's limit."""
    h = 2e-2
    at = 1.0 / inf_tau(V0, g)[1]
    around = 0.5 * (1.0 / inf_tau(V0 + h, g)[1] + 1.0 / inf_tau(V0 - h, g)[1])
    assert at == pytest.approx(around, rel=2e-4), (g, V0, at, around)


def test_reference_guard_at_minus_10_is_ten_times_its_limit():
    """The I_Ks gate's closing rate 6.1e-5 u / (exp(0.145 u) - 1), u = V + 10,
    tends to 6.1e-5/0.145 as u -> 0, but the reference's guard for |u| <=
    0.001 returns 0.00061/0.145 (Purkinje.cc:44), ten times that.  The oracle
    (like the GPU) keeps the reference's value; this pins it: the opening
    rate is continuous there, so the jump in 1/tau is the guard's excess."""
    h = 2e-2
    at = 1.0 / inf_tau(-10.0, 17)[1]
    around = 0.5 * (1.0 / inf_tau(-10.0 + h, 17)[1] + 1.0 / inf_tau(-10.0 - h, 17)[1])
    assert at - around == pytest.approx((0.00061 - 6.1e-5) / 0.145, rel=1e-3)


def _charge(s, alg0=0.0):
    c = oracle.pk_constants()
    _, R = oracle.pk_rates(c, s, alg0)
    F = c[1]
    moved = F * (c[61] * (R[4] + R[5]) + 2.0 * (c[59] * R[1] + c[60] * R[2] + c[62] * R[3]))
    return 1000.0 * c[6] * R[0], moved


@pytest.mark.parametrize("V", [-90.0, -60.0, -20.0, 0.0, 25.0])
@pytest.mark.parametrize("stim", [0.0, -2000.0])
def test_charge_conservation(V, stim):
    """C dV/dt (pA, x1000 for the ms/s of the concentration equations) equals
    F times the charge the concentration rates move:
    F [V_myo (dNa + dK) + 2 (V_1 dCa_1 + V_2 dCa_2 + V_3 dCa_3)]."""
    rng = np.random.default_rng(int(abs(V)) + 3)
    for _ in range(5):
        s = oracle.pk_init()
        s[0] = V
        s[6:] = rng.uniform(0.0, 1.0, 16)
        s[1] *= rng.uniform(0.5, 20.0)
        s[3] *= rng.uniform(0.5, 2.0)
        lhs, rhs = _charge(s, stim)
        assert lhs == pytest.approx(rhs, rel=1e-11, abs=1e-9 * (abs(lhs) + 1.0))


def test_compute_ion_forward_euler_and_adaptive_substeps():
    c = oracle.pk_constants()
    xi, area = GOLD["purkinje_xi"], GOLD["purkinje_area_cm2"]
    # Quiet cell: one forward-Euler step of every state, V included.
    s = oracle.pk_init()
    _, R = oracle.pk_rates(c, s)
    assert abs(R[0]) < 1.5
    s1 = s.copy()
    dv = oracle.pk_compute_ion(c, s1, 1.0, 0.02, s[0], 0.0)
    assert dv == R[0]
    np.testing.assert_allclose(s1, s + 0.02 * R, rtol=1e-15, atol=1e-18)
    # Stimulated: the stimulus enters as -1e6 istim area / Xi pA, and with
    # |dV/dt| > 1.5 the step is ten forward-Euler steps of dt/10.
    istim = 50.0 * 2500.0
    alg0 = -1e6 * istim * area / xi
    _, R = oracle.pk_rates(c, s, alg0)
    assert abs(R[0]) > 1.5
    want = s.copy()
    for _ in range(10):
        _, Rk = oracle.pk_rates(c, want, alg0)
        want = want + Rk * 0.002
    got = s.copy()
    dv = oracle.pk_compute_ion(c, got, 1.0, 0.02, s[0], istim)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-18)
    assert dv == pytest.approx(Rk[0], rel=1e-14)


def test_action_potential_markers():
    """A 0.5 ms stimulus takes the cell from its initial state to about -55
    mV; the cell then fires by itself (its peak comes after the stimulus
    has ended), overshoots, and repolarises to near its diastolic level
    within 1 s."""
    c = oracle.pk_constants()
    s = oracle.pk_init()
    V, dt = s[0], 0.01
    trace = []
    for k in range(100000):
        ist = 2.0 * 50.0 * 2500.0 if k < 50 else 0.0
        dv = oracle.pk_compute_ion(c, s, 1.0, dt, V, ist)
        V += dt * dv
        trace.append(V)
    tr = np.array(trace)
    peak = tr.max()
    t_peak = tr.argmax() * dt
    above = np.nonzero(tr > -60.0)[0]
    apd = (above[-1] - above[0]) * dt
    print(f"Purkinje AP: peak {peak:.1f} mV at {t_peak:.2f} ms, time above -60 mV {apd:.0f} ms, "
          f"V at 1 s {tr[-1]:.1f} mV")
    assert tr[49] < -40.0
    assert peak > 20.0 and 0.5 < t_peak < 10.0
    assert 100.0 < apd < 800.0
    assert abs(tr[-1] - GOLD["purkinje_initial_state"][0]) < 10.0


# Currents of ComputeRates (the reference's algebraic slots) by the ion whose
# balance each enters (Purkinje.cc:57-90): Na (Rates[4]), K (Rates[5]), Ca
# (Rates[1]).  Every one of them with a driving force vanishes at its ion's
# Nernst potential RT/F ln(o/i) (half of it for Ca).
NA_CURRENTS = (38, 40, 56, 66)
K_CURRENTS = (33, 36, 43, 46, 48, 54, 62)
CA_CURRENTS = (25, 31, 58)


@pytest.mark.parametrize("ion, slots", [("na", NA_CURRENTS), ("k", K_CURRENTS),
                                        ("ca", CA_CURRENTS)])
def test_currents_vanish_at_their_reversal_potentials(ion, slots):
    c = oracle.pk_constants()
    s = oracle.pk_init()
    s[6:] = 0.5
    rt_f = c[2] * c[0] / c[1]
    e = {"na": rt_f * np.log(c[4] / s[4]), "k": rt_f * np.log(c[5] / s[5]),
         "ca": 0.5 * rt_f * np.log(c[3] / s[1])}[ion]
    s[0] = e
    alg, _ = oracle.pk_rates(c, s)
    for k in slots:
        assert abs(alg[k]) < 1e-12, (ion, k, alg[k])
    s[0] = e + 10.0
    alg, _ = oracle.pk_rates(c, s)
    for k in slots:
        assert alg[k] > 0.0, (ion, k)  # outward above the reversal potential


def test_adaptive_substeps_start_just_above_the_threshold():
    """|dV/dt| = 5 mV/ms is above the 1.5 mV/ms threshold (Purkinje.cc:183):
    the step is ten sub-steps, not one."""
    c = oracle.pk_constants()
    xi, area = GOLD["purkinje_xi"], GOLD["purkinje_area_cm2"]
    s = oracle.pk_init()
    istim = 5.0 * xi * c[6] / (1e6 * area)
    alg0 = -1e6 * istim * area / xi
    _, R = oracle.pk_rates(c, s, alg0)
    assert 3.0 < abs(R[0]) < 8.0
    want = s.copy()
    for _ in range(10):
        _, Rk = oracle.pk_rates(c, want, alg0)
        want = want + Rk * 0.002
    got = s.copy()
    oracle.pk_compute_ion(c, got, 1.0, 0.02, s[0], istim)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-18)


def test_sodium_calcium_exchanger():
    """I_NaCa (Purkinje.cc:86): k 512.009 (e^{g(n-2)VF/RT} Na_i^n Ca_o -
    Ca_i e^{(g-1)(n-2)VF/RT} Na_o^n) / ((1 + d (Ca_i Na_o^n + Ca_o Na_i^n))
    (1 + Ca_i/0.0069)) with n = 3, g = 0.5: it reverses at 3 E_Na - 2 E_Ca
    and has the stated value at 0 mV."""
    c = oracle.pk_constants()
    s = oracle.pk_init()
    rt_f = c[2] * c[0] / c[1]
    nai, cai, nao, cao = s[4], s[1], c[4], c[3]
    s[0] = rt_f * np.log(cai * nao ** 3 / (cao * nai ** 3))
    s[0] = 0.0
    scale = abs(oracle.pk_rates(c, s)[0][49])
    s[0] = rt_f * np.log(cai * nao ** 3 / (cao * nai ** 3))
    assert abs(oracle.pk_rates(c, s)[0][49]) < 1e-12 * scale
    s[0] = 0.0
    want = (512.009 * c[33] * (nai ** 3 * cao - cai * nao ** 3) /
            ((1.0 + c[34] * (cai * nao ** 3 + cao * nai ** 3)) * (1.0 + cai / 0.0069)))
    assert oracle.pk_rates(c, s)[0][49] == pytest.approx(want, rel=1e-12)
