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
