"""The NumPy network that checks the device's surrogate kernels, pinned on
the CPU: its gradient against central finite differences in float64, its
bfloat16 rounding against torch's, its layout against the library's
parameter count."""
import numpy as np
import torch

import surrogate_ref as sr
from ucla_cmg_amd import voomgpu as vg


def _data(seed, B=16):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(B, 25)), rng.normal(size=(B, 19))


def test_gradient_matches_finite_differences():
    p = sr.random_params(3)
    xn, yn = _data(4)
    loss, g = sr.loss_grad(p, xn, yn)
    rng = np.random.default_rng(5)
    # A spread of entries over every layer's weights and biases.
    offs = np.cumsum([0] + [k * n + n for k, n in sr.LAYERS])
    idx = []
    for l, (k, n) in enumerate(sr.LAYERS):
        idx += list(offs[l] + rng.integers(0, k * n, 6))   # weights
        idx += list(offs[l] + k * n + rng.integers(0, n, 3))  # biases
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (sr.loss_grad(pp, xn, yn)[0] - sr.loss_grad(pm, xn, yn)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert loss > 0


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-6, 6, 4000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(sr.bf16(x), ref)


def test_layout_matches_library_parameter_count():
    n = sum(k * m + m for k, m in sr.LAYERS)
    assert n == vg.Surrogate.num_params()
    assert [(k, m) for k, m in sr.LAYERS] == vg.SURR_LAYERS


def test_input_halves_carry_16_bits():
    """Layer 0's two BF16 halves of an input: hi = bf16(x), lo = bf16(x - hi);
    hi alone is within 2^-8 of x relative (8 significant bits), hi + lo
    within 2^-16 (lo rounds a remainder at most 2^-8 of x); without rounding
    the halves are (x, 0)."""
    rng = np.random.default_rng(7)
    x = (rng.normal(size=(2000, 25)) * 10.0 ** rng.integers(-3, 3, (2000, 25))).astype(np.float32)
    h = sr.split(x, bf=True)
    hi, lo = h[:, :25], h[:, 25:]
    np.testing.assert_array_equal(hi, sr.bf16(x))
    xd = x.astype(np.float64)
    rel_hi = np.abs(hi - xd) / np.abs(xd)
    rel = np.abs(hi + lo - xd) / np.abs(xd)
    assert rel_hi.max() <= 2.0 ** -8 and rel.max() <= 2.0 ** -16
    assert rel_hi.max() > 2.0 ** -10  # hi alone is visibly coarser
    np.testing.assert_array_equal(sr.split(x, bf=False), np.concatenate([xd, 0 * xd], axis=1))
