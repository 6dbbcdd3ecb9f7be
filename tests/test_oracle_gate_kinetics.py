"""Pins of the oracle's gate kinetics against what the published models fix,
independently of how the oracle writes them.

ten Tusscher 2006 (IonicModel/Tusscher.cc:4-96, the CellML export of
ten Tusscher & Panfilov, Am J Physiol Heart 291:H1088, 2006): every gate's
rate is (x_inf(V) - x) / tau(V), so two rate evaluations at x = 0 and x = 1
give x_inf and tau as the oracle computes them.  They are checked against the
paper's half-activation points (where the Boltzmann factor is exactly 1), the
limits of each steady state, its direction of change with V, tau at the
points where its exponentials reduce to constants, and linearity in x.

Luo-Rudy 1991 (IonicModel/LR.cc:102-182, make_parameter1): the two guarded
removable singularities (alpha_m at V = -47.13 mV, the I_K factor at V = -77
mV) must equal the formula's limit there, h and j have no opening rate above
-40 mV, and the steady states move in the published directions.
"""
import numpy as np
import pytest

import oracle

# State indices of the tT2006 vector (Tusscher.h:51-69, CellML order).
XR1, XR2, XS, M, H, J, CASS, D, F, F2, FCASS, S, R = 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16
ACTIVATION = (XR1, XS, M, D, R)
INACTIVATION = (XR2, H, J, F, F2, S)
VGRID = np.linspace(-100.0, 60.0, 321)


def tt_inf_tau(V, g, ca_ss=None):
    c = oracle.tt_constants()
    s = oracle.tt_init()
    s[0] = V
    if ca_ss is not None:
        s[CASS] = ca_ss
    s[g] = 0.0
    r0 = oracle.tt_rates(c, s)[1][g]
    s[g] = 1.0
    r1 = oracle.tt_rates(c, s)[1][g]
    tau = 1.0 / (r0 - r1)
    return r0 * tau, tau


# Half points of the published steady states: m_inf, h_inf and j_inf are
# squared Boltzmann functions (1/4 at their centre), f2_inf = 0.67 B + 0.33
# and fCass_inf = 0.6 / (1 + (Ca_ss / 0.05)^2) + 0.4.
@pytest.mark.parametrize("g, V, want", [
    (M, -56.86, 0.25), (H, -71.55, 0.25), (J, -71.55, 0.25), (D, -8.0, 0.5), (F, -20.0, 0.5),
    (F2, -35.0, 0.665), (XR1, -26.0, 0.5), (XR2, -88.0, 0.5), (XS, -5.0, 0.5), (R, 20.0, 0.5),
    (S, -20.0, 0.5)])
def test_tt_steady_state_half_points(g, V, want):
    x_inf, tau = tt_inf_tau(V, g)
    assert x_inf == pytest.approx(want, abs=1e-12)
    assert tau > 0.0


def test_tt_fcass_steady_state_and_time_constant():
    x_inf, tau = tt_inf_tau(-85.0, FCASS, ca_ss=0.05)
    assert x_inf == pytest.approx(0.7, abs=1e-12)
    assert tau == pytest.approx(80.0 / 2.0 + 2.0, rel=1e-12)
    # Limits in Ca_ss: 1 at no calcium, 0.4 when saturated.
    assert tt_inf_tau(-85.0, FCASS, ca_ss=0.0)[0] == pytest.approx(1.0, abs=1e-12)
    assert tt_inf_tau(-85.0, FCASS, ca_ss=50.0)[0] == pytest.approx(0.4, abs=1e-6)


def test_tt_steady_states_limits_and_direction():
    for g in ACTIVATION + INACTIVATION:
        xi = np.array([tt_inf_tau(V, g)[0] for V in VGRID])
        assert np.all(xi >= 0.0) and np.all(xi <= 1.0), g
        d = np.diff(xi)
        if g in ACTIVATION:
            assert np.all(d >= -1e-15), g
            assert xi[-1] - xi[0] > 0.5, (g, xi[0], xi[-1])
        else:
            assert np.all(d <= 1e-15), g
            assert xi[0] - xi[-1] > 0.5, (g, xi[0], xi[-1])
    # f2 saturates at 0.33 rather than 0 at positive potentials.
    assert tt_inf_tau(200.0, F2)[0] == pytest.approx(0.33, abs=1e-9)


@pytest.mark.parametrize("g, V, want", [
    (R, -40.0, 9.5 + 0.8),                       # 9.5 exp(-(V+40)^2/1800) + 0.8
    (S, -45.0, 85.0 + 5.0 / (1.0 + np.exp(-13.0)) + 3.0),  # epicardial s gate
    (H, -10.66, 0.13 / 0.77 * 2.0),              # above -40 mV: alpha_h = 0, beta_h = 0.77/(0.13 (1+1))
])
def test_tt_time_constants_at_exact_points(g, V, want):
    assert tt_inf_tau(V, g)[1] == pytest.approx(want, rel=1e-12)


def test_tt_time_constants_positive_and_rates_linear_in_the_gate():
    c = oracle.tt_constants()
    for g in ACTIVATION + INACTIVATION + (FCASS,):
        for V in VGRID[::8]:
            x_inf, tau = tt_inf_tau(V, g)
            assert np.isfinite(tau) and tau > 0.0, (g, V)
            s = oracle.tt_init()
            s[0] = V
            s[g] = 0.3
            assert oracle.tt_rates(c, s)[1][g] == pytest.approx((x_inf - 0.3) / tau, rel=1e-9, abs=1e-14)


# ---- Luo-Rudy 1991 ----------------------------------------------------------
# oracle.lr_tab: [m_inf, a+b (m), h_inf, a+b (h), j_inf, a+b (j), d_inf, a+b (d),
#                 f_inf, a+b (f), X_inf, a+b (X), X_i factor, I_K1 + I_Kp + I_b]

def test_lr_guarded_singularities_equal_their_limits():
    # alpha_m = 0.32 u / (1 - exp(-0.1 u)), u = V + 47.13 -> 3.2 as u -> 0.
    V0, h = -47.13, 2e-3
    t0, tp, tm = oracle.lr_tab(V0), oracle.lr_tab(V0 + h), oracle.lr_tab(V0 - h)
    for k in (0, 1):
        assert t0[k] == pytest.approx(0.5 * (tp[k] + tm[k]), rel=1e-6)
    # X_i = 2.837 (exp(0.04 u) - 1) / (u exp(0.04 (V + 35))), u = V + 77:
    # limit 2.837 * 0.04 * exp(1.68) = 0.608884; the reference stores 0.608889.
    V0 = -77.0
    t0, tp, tm = oracle.lr_tab(V0), oracle.lr_tab(V0 + h), oracle.lr_tab(V0 - h)
    assert t0[12] == 0.608889
    assert 0.5 * (tp[12] + tm[12]) == pytest.approx(2.837 * 0.04 * np.exp(1.68), rel=1e-6)
    assert t0[12] == pytest.approx(0.5 * (tp[12] + tm[12]), abs=1e-5)


def test_lr_steady_states_direction_and_range():
    tabs = np.array([oracle.lr_tab(V) for V in VGRID])
    for k in (0, 2, 4, 6, 8, 10):
        assert np.all(tabs[:, k] >= 0.0) and np.all(tabs[:, k] <= 1.0), k
    for k in (0, 6, 10):  # m, d, X open with depolarisation
        assert np.all(np.diff(tabs[:, k]) >= -1e-15), k
    for k in (2, 4, 8):   # h, j, f close with depolarisation
        assert np.all(np.diff(tabs[:, k]) <= 1e-15), k
    above = VGRID > -40.0
    assert np.all(tabs[above, 2] == 0.0) and np.all(tabs[above, 4] == 0.0)
    assert np.all(tabs[:, [1, 3, 5, 7, 9, 11]] > 0.0)


# The twelve LR1991 rate functions (Luo & Rudy, Circ Res 68:1501, 1991,
# appendix; LR.cc:110-133), read back from the oracle's table as
# alpha = x_inf (alpha + beta), beta = (1 - x_inf)(alpha + beta).  Each has
# the form A exp(-a u) / (1 + exp(-b u)) (u = V - c) or a plain exponential,
# so it is fixed by its value where the denominator's exponential is 1 (A/2)
# and by its two asymptotic log-slopes (-a for u -> +inf, b - a for
# u -> -inf).  Gate k of the table: m 0, h 1, j 2, d 3, f 4, X 5.
def lr_rates(V, k):
    t = oracle.lr_tab(V)
    x, s = t[2 * k], t[2 * k + 1]
    return x * s, (1.0 - x) * s


def _logslope(f, V, h=0.05):
    return (np.log(f(V + h)) - np.log(f(V - h))) / (2 * h)


@pytest.mark.parametrize("k, which, V, value, slope_hi, slope_lo, b", [
    (3, 0, 5.0, 0.095 / 2, -0.01, 0.072 - 0.01, 0.072),      # alpha_d
    (3, 1, -44.0, 0.07 / 2, -0.017 - 0.05, -0.017, 0.05),    # beta_d
    (4, 0, -28.0, 0.012 / 2, -0.008 - 0.15, -0.008, 0.15),   # alpha_f
    (4, 1, -30.0, 0.0065 / 2, -0.02, 0.2 - 0.02, 0.2),       # beta_f
    (5, 0, -50.0, 0.0005 / 2, 0.083 - 0.057, 0.083, 0.057),  # alpha_X
    (5, 1, -20.0, 0.0013 / 2, -0.06, 0.04 - 0.06, 0.04),     # beta_X
])
def test_lr_rate_centres_and_asymptotic_slopes(k, which, V, value, slope_hi, slope_lo, b):
    f = lambda v: lr_rates(v, k)[which]  # noqa: E731
    assert f(V) == pytest.approx(value, rel=1e-12)
    # 9/b from the centre the denominator's exponential is exp(-9) of 1 (or
    # of itself); much further the rate can no longer be read back from
    # alpha/(alpha + beta), which rounds to 0 or 1.
    D = 9.0 / b
    tol = 3.0 * b * np.exp(-9.0)
    assert _logslope(f, V + D) == pytest.approx(slope_hi, abs=tol)
    assert _logslope(f, V - D) == pytest.approx(slope_lo, abs=tol)


def test_lr_sodium_gate_rates():
    # beta_m = 0.08 exp(-V/11): 0.08 at 0 mV, log-slope -1/11 everywhere.
    bm = lambda v: lr_rates(v, 0)[1]  # noqa: E731
    assert bm(0.0) == pytest.approx(0.08, rel=1e-12)
    for V in (-90.0, -30.0, 40.0):
        assert _logslope(bm, V) == pytest.approx(-1.0 / 11.0, rel=1e-8)
    # alpha_m = 0.32 u / (1 - exp(-0.1 u)), u = V + 47.13: -> 0.32 u far above.
    assert lr_rates(300.0, 0)[0] == pytest.approx(0.32 * 347.13, rel=1e-12)
    # V > -40: h and j have no opening rate; beta_h = 1/(0.13 (1 + e)) with
    # e = exp(-(V + 10.66)/11.1), beta_j = 0.3 exp(-2.535e-7 V)/(1 + exp(-0.1 (V + 32))).
    assert lr_rates(-10.66, 1)[1] == pytest.approx(1.0 / 0.26, rel=1e-12)
    assert lr_rates(400.0, 1)[1] == pytest.approx(1.0 / 0.13, rel=1e-12)
    g = lambda v: 1.0 / lr_rates(v, 1)[1] - 0.13  # noqa: E731
    assert _logslope(g, -20.0) == pytest.approx(-1.0 / 11.1, rel=1e-8)
    assert lr_rates(-32.0, 2)[1] == pytest.approx(0.15 * np.exp(2.535e-7 * 32.0), rel=1e-12)
    bj = lambda v: lr_rates(v, 2)[1]  # noqa: E731
    assert _logslope(bj, 300.0, h=5.0) == pytest.approx(-2.535e-7, rel=1e-3)
    # V < -40: alpha_h = 0.135 exp(-(80 + V)/6.8); beta_j = 0.1212
    # exp(-0.01052 V)/(1 + exp(-0.1378 (V + 40.14))).
    ah = lambda v: lr_rates(v, 1)[0]  # noqa: E731
    assert ah(-80.0) == pytest.approx(0.135, rel=1e-12)
    assert _logslope(ah, -70.0) == pytest.approx(-1.0 / 6.8, rel=1e-8)
    assert lr_rates(-40.14, 2)[1] == pytest.approx(0.0606 * np.exp(0.01052 * 40.14), rel=1e-12)
    assert _logslope(bj, -40.14 - 9.0 / 0.1378) == pytest.approx(0.1378 - 0.01052, rel=1e-3)
    # beta_h = 3.56 exp(0.079 V) + 3.1e5 exp(0.35 V): the first term rules far
    # below, the sum's log-slope runs from 0.079 towards 0.35 as V rises.
    bh = lambda v: lr_rates(v, 1)[1]  # noqa: E731
    assert _logslope(bh, -150.0) == pytest.approx(0.079, rel=1e-5)
    assert 0.079 < _logslope(bh, -60.0) < _logslope(bh, -45.0) < 0.35
    # alpha_j vanishes with (V + 37.78) only above -40, where the branch
    # sets it to 0; below -40 it is positive and falls off at both ends.
    aj = np.array([lr_rates(v, 2)[0] for v in np.linspace(-120.0, -40.5, 60)])
    assert np.all(aj > 0.0)
    assert lr_rates(-39.0, 2)[0] == 0.0


def test_lr_background_potassium_currents_vanish_at_ek1():
    """tab[13] = I_K1 + I_Kp + I_b; the two K currents are proportional to
    V - E_K1 (E_K1 = RT/F ln(K_o/K_i), LR.cc:55), so at E_K1 only the
    background current 0.0392 (V + 59.87) remains."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ionic_tables.json")))
    c = g["luorudy_constants"]
    ek1 = c["RT_F"] * np.log(c["c_K0"] / c["c_Ki"])
    assert oracle.lr_tab(ek1)[13] == pytest.approx(0.0392 * (ek1 + 59.87), rel=1e-12)
    # I_b alone is linear with slope 0.0392: the total's slope exceeds it.
    t1, t2 = oracle.lr_tab(ek1 + 1e-3)[13], oracle.lr_tab(ek1 - 1e-3)[13]
    assert (t1 - t2) / 2e-3 > 0.0392


# Slopes of the same published steady states (tT2006 equations): each is a
# logistic B(V) = 1/(1 + exp(-(V - V_half)/k)) -- m, h, j its square, f2
# 0.67 B + 0.33 -- so logit(B) has the constant slope 1/k everywhere (its sign
# says whether the gate opens or closes with depolarisation).
@pytest.mark.parametrize("g, k, form", [
    (XR1, 7.0, "b"), (XR2, -24.0, "b"), (XS, 14.0, "b"), (D, 7.5, "b"), (F, -7.0, "b"),
    (R, 6.0, "b"), (S, -5.0, "b"), (M, 9.03, "sq"), (H, -7.43, "sq"), (J, -7.43, "sq"),
    (F2, -7.0, "f2")])
def test_tt_steady_state_slopes(g, k, form):
    def logit(V):
        x = tt_inf_tau(V, g)[0]
        b = np.sqrt(x) if form == "sq" else (x - 0.33) / 0.67 if form == "f2" else x
        return np.log(b / (1.0 - b))
    for V in (-80.0, -40.0, -10.0, 20.0):
        assert (logit(V + 0.01) - logit(V - 0.01)) / 0.02 == pytest.approx(1.0 / k, rel=1e-6), V


def test_lr_alpha_j_where_its_denominator_is_two():
    """alpha_j (V < -40; LR1991 appendix, LR.cc:124-125) at V = -79.23 mV,
    where 1 + exp(0.311 (V + 79.23)) = 2, from the published expression."""
    V = -79.23
    want = (-1.2714e5 * np.exp(0.2444 * V) - 3.474e-5 * np.exp(-0.04391 * V)) * (V + 37.78) / 2.0
    assert lr_rates(V, 2)[0] == pytest.approx(want, rel=1e-12)


def test_lr_fast_sodium_current_reversal_and_conductance():
    """I_Na = G_Na m^3 h j (V - 54.4) (LR.cc:87, E_Na fixed): with dt = 0 the
    gates and Ca_i stay put, so toggling h from 0 to 1 (m = j = 1) changes
    the returned dV/dt by exactly -G_Na (V - 54.4)."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ionic_tables.json")))
    gna = g["luorudy_constants"]["Gna"]
    for V in (54.4, 0.0, -60.0):
        out = []
        for h in (0.0, 1.0):
            s = oracle.lr_init()
            s[0], s[1], s[2] = 1.0, h, 1.0
            out.append(oracle.lr_compute_ion(s, 1.0, 0.0, V, 0.0))
        assert out[1] - out[0] == pytest.approx(-gna * (V - 54.4), rel=1e-12, abs=1e-12)


def test_lr_calcium_update_with_the_slow_inward_current():
    """With the gates at their steady state (left in place by Rush-Larsen),
    Ca_i takes one forward-Euler step of -1e-4 I_si + 0.07 (1e-4 - Ca_i)
    with I_si = G_si d f (V - E_si), E_si = 7.7 - 13.0287 ln Ca_i (LR1991;
    LR.cc:86-93, G_si modified to 0.07)."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ionic_tables.json")))
    gsi = g["luorudy_constants"]["Gsi"]
    V, dt, ca = 10.0, 0.01, 2e-4
    t = oracle.lr_tab(V)
    s = oracle.lr_init()
    s[:6] = t[[0, 2, 4, 6, 8, 10]]
    s[6] = ca
    oracle.lr_compute_ion(s, 1.0, dt, V, 0.0)
    i_si = gsi * t[6] * t[8] * (V - (7.7 - 13.0287 * np.log(ca)))
    assert s[6] == pytest.approx(ca + (-1e-4 * i_si + 0.07 * (1e-4 - ca)) * dt, rel=1e-12)
    assert abs(i_si) > 1e-3


def test_tt_ryanodine_receptor_recovery():
    """R' (tT2006: dR'/dt = -k2 Ca_ss R' + k4 (1 - R'), Tusscher.cc:64-67):
    no change at R' = 1 without subspace Ca, k4 = C[36] at R' = 0."""
    c = oracle.tt_constants()
    s = oracle.tt_init()
    s[CASS] = 0.0
    s[18] = 1.0
    assert oracle.tt_rates(c, s)[1][18] == pytest.approx(0.0, abs=1e-15)
    s[18] = 0.0
    assert oracle.tt_rates(c, s)[1][18] == pytest.approx(c[36], rel=1e-14)
    s[CASS] = 1e-3
    s[18] = 1.0
    assert oracle.tt_rates(c, s)[1][18] < 0.0
