"""The learned ten Tusscher step on the GPU: the device's training kernels
(forward, backward) against the NumPy network, the recording of the plain
update, the fused inference kernel inside the time step against the training
forward, the envelope fallback, and training on recorded data."""
import numpy as np
import pytest

import surrogate_ref as sr
from ucla_cmg_amd import meshgen
from ucla_cmg_amd import voomgpu as vg

pytestmark = pytest.mark.gpu

DL = 1.2 / 1400.0


def _features(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 25)).astype(np.float32) * np.linspace(0.5, 20.0, 25).astype(np.float32)
    y = rng.normal(size=(n, 19)).astype(np.float32) * np.linspace(0.01, 3.0, 19).astype(np.float32)
    return x, y


def _scaled(seed, x, y):
    s = vg.Surrogate(seed)
    s.add_samples(x, y)
    s.fit_scaling(0.0)
    return s


def _norm32(x, y, sc):
    """Network inputs ((x - mu) times the correctly rounded 1 / sd) and
    targets as the device forms them (float32)."""
    inv = np.float32(1.0) / sc["sd_x"].astype(np.float32)
    xn = ((x.astype(np.float32) - sc["mu_x"]) * inv).astype(np.float64)
    t = np.arcsinh(y.astype(np.float32) / sc["c_y"])
    yn = ((t - sc["mu_y"]) / sc["sd_y"]).astype(np.float64)
    return xn, yn


def test_loss_and_gradient_match_the_numpy_network(gpu_device):
    """Device loss and gradient against the NumPy network with bfloat16
    rounding at the same places; the float64 network without that rounding
    is further off, which shows the rounding model is the one the kernels
    use."""
    x, y = _features(512, 1)
    s = _scaled(7, x, y)
    p = sr.random_params(11).astype(np.float32)
    s.set_params(p)
    np.testing.assert_array_equal(s.get_params(), p)
    sc = s.get_scaling()
    xn, yn = _norm32(x, y, sc)
    loss, g = s.loss_grad(x, y)
    loss_bf, g_bf = sr.loss_grad(p.astype(np.float64), xn, yn, bf=True)
    loss_64, g_64 = sr.loss_grad(p.astype(np.float64), xn, yn, bf=False)
    assert abs(loss - loss_bf) <= 1e-4 * loss_bf
    offs = np.cumsum([0] + [k * n + n for k, n in sr.LAYERS])
    errs = []
    for l, (k, n) in enumerate(sr.LAYERS):
        for part in (slice(offs[l], offs[l] + k * n), slice(offs[l] + k * n, offs[l + 1])):
            e = np.linalg.norm(g[part] - g_bf[part]) / np.linalg.norm(g_bf[part])
            e64 = np.linalg.norm(g[part] - g_64[part]) / np.linalg.norm(g_64[part])
            errs.append((e, e64))
            assert e < 5e-3, (l, e)
    print("relative gradient error per block (bf16 model, float64 model):",
          [(f"{a:.1e}", f"{b:.1e}") for a, b in errs])
    assert np.mean([a for a, _ in errs]) < np.mean([b for _, b in errs])


def test_predict_matches_the_numpy_network(gpu_device):
    x, y = _features(256, 2)
    s = _scaled(3, x, y)
    p = sr.random_params(5).astype(np.float32)
    s.set_params(p)
    sc = s.get_scaling()
    got = s.predict(x)
    xn, _ = _norm32(x, y, sc)
    Z, _ = sr.forward(p.astype(np.float64), xn, bf=True)
    # Compared on the network's output (the transformed increment).
    z_got = (np.arcsinh(got.astype(np.float64) / sc["c_y"]) - sc["mu_y"]) / sc["sd_y"]
    # Sums in float32 against float64 flip the bfloat16 rounding of a rare
    # activation sitting on a rounding boundary, which moves that sample's
    # outputs: bound most samples tightly and the rare one loosely.
    err = np.abs(z_got - Z[-1]).max(axis=1)
    print(f"samples off by > 1e-5: {(err > 1e-5).sum()} of {err.size}, max {err.max():.1e}")
    assert (err > 1e-5).mean() < 0.03 and err.max() < 1e-2, err.max()


def test_training_fits_a_smooth_map(gpu_device):
    """Adam on the device drives the loss of a learnable map well down."""
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(1 << 15, 25)).astype(np.float32)
    R = rng.normal(size=(25, 19)) / 3.0
    y = np.sin(x @ R).astype(np.float32)
    s = _scaled(13, x, y)
    l0, _ = s.loss_grad(x[:4096], y[:4096])
    s.train(600, batch=1024, lr=2e-3)
    s.train(300, batch=1024, lr=5e-4)
    l1, _ = s.loss_grad(x[:4096], y[:4096])
    print(f"loss {l0:.4f} -> {l1:.5f}")
    assert l1 < 0.05 * l0


def _varied_states(seed=0):
    """States of a slab 6 ms after a corner stimulus (resting, upstroke and
    plateau nodes), FP64 Rush-Larsen."""
    mesh = meshgen.slab(21, 21, 5, 0.025)
    s = vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, precision=vg.PREC_FP64,
                         integrator=vg.INTEG_RUSH_LARSEN, dt=0.02, D=(DL, DL, DL), cg_tol=1e-10)
    s.setStimulusCurrent(meshgen.slab_corner_nodes(21, 21, 5, 4), 5e4)
    s.Solve(100)
    s.clearStimulus()
    s.Solve(200)
    st = s.getStateVariables()
    s.close()
    return mesh, st


def _isolated(mesh, precision):
    """The same nodes with (almost) no coupling: each node's update is its
    ionic update alone."""
    return vg.CardiacSolver(mesh, cell_model=vg.MODEL_TT, precision=precision,
                            integrator=vg.INTEG_RUSH_LARSEN, dt=0.02, D=(1e-14,) * 3,
                            cg_tol=1e-10)


def _load(h, st):
    """States (slot 0 = V) into a solver: the ionic states and the voltage,
    which the solver keeps apart (InitializeVoltage)."""
    h.setStateVariables(st)
    h.InitializeVoltage(st[:, 0])


def test_collect_records_the_plain_update(gpu_device):
    mesh, st = _varied_states()
    assert (st[:, 0] > 0).any() and (st[:, 0] < -80).any()
    s = vg.Surrogate(1)
    s.reserve(4 * mesh.n_nodes)
    g = _isolated(mesh, vg.PREC_FP64)
    _load(g, st)
    s.collect(g, 1, every=1, node_stride=1)
    after = g.getStateVariables()
    x, y = s.get_samples()
    assert x.shape[0] == mesh.n_nodes
    np.testing.assert_allclose(x[:, 19], np.log(0.02), rtol=1e-6)
    np.testing.assert_array_equal(x[:, 20:], 1.0)
    # Match samples to nodes through their (float32) state vectors.
    key = {tuple(v): i for i, v in enumerate(st.astype(np.float32))}
    idx = np.array([key[tuple(v)] for v in x[:, :19]])
    assert len({tuple(v) for v in x[:, :19]}) == len(key) > 100  # every distinct state
    inc = (after - st)[idx]
    tol = 1e-5 * np.abs(inc).max(axis=0) + 1e-7
    assert (np.abs(y - inc) <= tol).all()
    # Collecting with a stride and a skip records every stride-th row.
    s2 = vg.Surrogate(2)
    s2.reserve(4 * mesh.n_nodes)
    s2.collect(g, 6, every=3, node_stride=5)
    assert s2.num_samples() == 2 * ((mesh.n_nodes + 4) // 5)
    g.close()


@pytest.mark.parametrize("precision,envelope", [(vg.PREC_FP64, "input"), (vg.PREC_MIXED, "input"),
                                                (vg.PREC_MIXED, "output")])
def test_step_kernel_equals_training_forward_and_falls_back(gpu_device, precision, envelope):
    """Attached, the step's fused MFMA kernel advances each in-envelope node
    by what the training forward predicts; nodes whose input (here V) or
    output (here the transformed V increment) leaves its envelope take the
    plain update unchanged."""
    mesh, st = _varied_states()
    n = mesh.n_nodes
    s = vg.Surrogate(4)
    s.reserve(4 * n)
    g = _isolated(mesh, precision)
    _load(g, st)
    s.collect(g, 1)
    s.fit_scaling(0.05)
    s.set_params(sr.random_params(21, scale=0.5).astype(np.float32))
    x, y = s.get_samples()
    y_sd = y.std(axis=0) + 1e-30
    pad = (-x.shape[0]) % 128
    xp = np.concatenate([x, np.repeat(x[:1], pad, 0)])
    pred = s.predict(xp)[: x.shape[0]]
    key = {tuple(v): i for i, v in enumerate(x[:, :19])}

    sc = s.get_scaling()
    sc["t_lo"][:], sc["t_hi"][:] = -1e30, 1e30
    st2 = st.copy()
    if envelope == "input":
        # Half the nodes leave the input envelope through V.
        out = np.zeros(n, bool)
        out[::2] = True
        st2[out, 0] = sc["hi"][0] + 5.0
    else:
        # Nodes whose network output for V lies above a cut leave the output
        # envelope (the cut sits in a gap of the outputs, clear of rounding).
        tv = (np.arcsinh(pred[:, 0].astype(np.float64) / sc["c_y"][0]) - sc["mu_y"][0]) / sc["sd_y"][0]
        node_t = np.array([tv[key[tuple(v)]] for v in st.astype(np.float32)])
        u = np.unique(node_t)
        mid = len(u) // 2
        gap = mid - 10 + np.argmax(np.diff(u[mid - 10: mid + 10]))
        cut = 0.5 * (u[gap] + u[gap + 1])
        assert u[gap + 1] - u[gap] > 1e-4
        sc["t_hi"][0] = cut
        out = node_t > cut
        assert 0.1 < out.mean() < 0.9
    s.set_scaling(sc)
    plain = _isolated(mesh, precision)
    _load(plain, st2)
    plain.Solve(1)
    ref_plain = plain.getStateVariables()

    net = _isolated(mesh, precision)
    _load(net, st2)
    s.attach(net)
    net.Solve(1)
    got = net.getStateVariables()
    used, fb = vg.Surrogate.usage(net)
    assert used == (~out).sum() and fb == out.sum()
    np.testing.assert_array_equal(got[out, 1:], ref_plain[out, 1:])
    x0 = st2.astype(np.float32 if precision == vg.PREC_MIXED else np.float64)
    inside = np.flatnonzero(~out)
    k = np.array([key[tuple(v)] for v in st[inside].astype(np.float32)])
    want = x0[inside].astype(np.float64) + pred[k]
    gates = [4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16, 18]
    want[:, gates] = np.clip(want[:, gates], 0.0, 1.0)
    # Up to the rounding of the stored sum (float32 in mixed precision: Ki's
    # increments lie below its float32 spacing).
    ulp = 2.0 ** -23 if precision == vg.PREC_MIXED else 2.0 ** -52
    err = np.abs(got[inside] - want) / (2e-3 * y_sd + ulp * np.abs(want))
    assert err.max() < 1.0, err.max()
    # Detached, the plain model runs everything again.
    vg.Surrogate.detach(net)
    net.Solve(1)
    for h in (g, plain, net):
        h.close()


def test_stimulated_nodes_fall_back(gpu_device):
    mesh, st = _varied_states()
    s = vg.Surrogate(5)
    s.reserve(2 * mesh.n_nodes)
    g = _isolated(mesh, vg.PREC_MIXED)
    _load(g, st)
    s.collect(g, 1)
    s.fit_scaling(0.5)
    sc = s.get_scaling()
    sc["t_lo"][:], sc["t_hi"][:] = -1e30, 1e30  # an untrained network: outputs unchecked
    s.set_scaling(sc)
    s.attach(g)
    stim = np.arange(10, dtype=np.int32)
    g.setStimulusCurrent(stim, 1e3)
    g.Solve(1)
    used, fb = vg.Surrogate.usage(g)
    assert fb == 10 and used == mesh.n_nodes - 10
    g.close()


def test_upstroke_and_band_rows_fall_back(gpu_device):
    """With the upstroke fallback set, exactly the rows with V above v_min
    and the h gate above h_min take the plain update (bit for bit), the
    others keep the network's (unset, the network takes every row); a
    voltage band adds its rows to the fallback."""
    mesh, st = _varied_states()
    v_min, h_min = -70.0, 0.05
    up = (st[:, 0] > v_min) & (st[:, 8] > h_min)
    assert 0 < up.sum() < mesh.n_nodes // 2
    s = vg.Surrogate(6)
    s.reserve(2 * mesh.n_nodes)
    g = _isolated(mesh, vg.PREC_MIXED)
    _load(g, st)
    s.collect(g, 1)
    s.fit_scaling(0.5)
    sc = s.get_scaling()
    sc["t_lo"][:], sc["t_hi"][:] = -1e30, 1e30
    s.set_scaling(sc)
    s.set_params(sr.random_params(23, scale=0.5).astype(np.float32))
    plain = _isolated(mesh, vg.PREC_MIXED)
    _load(plain, st)
    plain.Solve(1)
    ref = plain.getStateVariables()
    off = _isolated(mesh, vg.PREC_MIXED)
    _load(off, st)
    s.attach(off)
    off.Solve(1)
    assert vg.Surrogate.usage(off) == (mesh.n_nodes, 0)
    net_all = off.getStateVariables()
    s.set_upstroke_fallback(v_min, h_min)
    on = _isolated(mesh, vg.PREC_MIXED)
    _load(on, st)
    s.attach(on)
    on.Solve(1)
    used, fb = vg.Surrogate.usage(on)
    assert fb == up.sum() and used == (~up).sum()
    got = on.getStateVariables()
    # States bit for bit; V up to the ~1e-14 cm^2/ms coupling of the
    # otherwise isolated nodes (the two runs' fallback rows differ).
    np.testing.assert_array_equal(got[up, 1:], ref[up, 1:])
    np.testing.assert_array_equal(got[~up, 1:], net_all[~up, 1:])
    np.testing.assert_allclose(got[up, 0], ref[up, 0], rtol=0, atol=1e-6)
    np.testing.assert_allclose(got[~up, 0], net_all[~up, 0], rtol=0, atol=1e-6)
    # With a voltage band as well, the band's rows fall back too.
    lo, hi = -10.0, 40.0  # plateau rows (h ~ 0), outside the upstroke test
    band = (st[:, 0] > lo) & (st[:, 0] < hi)
    assert (band & ~up).any()
    s.set_band_fallback(lo, hi)
    both = _isolated(mesh, vg.PREC_MIXED)
    _load(both, st)
    s.attach(both)
    both.Solve(1)
    fbk = up | band
    assert vg.Surrogate.usage(both) == ((~fbk).sum(), fbk.sum())
    got = both.getStateVariables()
    np.testing.assert_array_equal(got[fbk, 1:], ref[fbk, 1:])
    np.testing.assert_array_equal(got[~fbk, 1:], net_all[~fbk, 1:])
    with pytest.raises(vg.VoomError):
        s.set_upstroke_fallback(float("nan"), 0.0)
    with pytest.raises(vg.VoomError):
        s.set_band_fallback(0.0, float("nan"))
    for h in (g, plain, off, on, both):
        h.close()
