"""NumPy restatement of the learned ten Tusscher step's network, used by the
tests as the checker of the device's training and inference kernels.

Layers 50 -> 128 -> 128 -> 128 -> 128 -> 19 (layer 0 reading the 25 normalised
inputs as two halves each), SiLU between them, loss the mean
over batch and outputs of the squared error of the normalised increments.
With ``bf16=True`` the values are rounded to bfloat16 (nearest even) exactly
where the device kernels store bfloat16: the normalised inputs (as two
halves, hi = bf16(x) and lo = bf16(x - hi)), the weights,
the hidden activations and the output gradients fed to the backward GEMMs;
sums stay in float64 here (float32 on the device).  With ``bf16=False`` the
whole computation is float64, which the finite-difference test pins."""
import numpy as np

LAYERS = [(50, 128), (128, 128), (128, 128), (128, 128), (128, 19)]


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = [], 0
    for k, n in LAYERS:
        W = np.asarray(p[o:o + k * n], np.float64).reshape(k, n)
        o += k * n
        out.append((W, np.asarray(p[o:o + n], np.float64)))
        o += n
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def silu(z):
    return z / (1.0 + np.exp(-z))


def normalise(x, y, sc):
    xn = (np.asarray(x, np.float64) - sc["mu_x"]) / sc["sd_x"]
    yn = None if y is None else (np.asarray(y, np.float64) - sc["mu_y"]) / sc["sd_y"]
    return xn, yn


def split(xn, bf=False):
    """Layer 0's inputs: each normalised input as two halves, his then los.
    With bf16 rounding, hi = bf16(x) and lo = bf16(x - hi) (x - hi in
    float32, where it is exact); without it, hi = x and lo = 0."""
    if not bf:
        x = np.asarray(xn, np.float64)
        return np.concatenate([x, np.zeros_like(x)], axis=1)
    x = np.asarray(xn, np.float32)
    hi = bf16(x)
    lo = bf16((x - hi.astype(np.float32)).astype(np.float32))
    return np.concatenate([hi, lo], axis=1)


def forward(p, xn, bf=False):
    """Pre-activations Z[l] and layer inputs A[l] of normalised inputs xn."""
    r = bf16 if bf else (lambda a: a)
    A, Z = [split(xn, bf)], []
    layers = unpack(p)
    for l, (W, b) in enumerate(layers):
        z = A[-1] @ r(W) + b
        Z.append(z)
        if l < len(layers) - 1:
            A.append(r(silu(z)))
    return Z, A


def loss_grad(p, xn, yn, bf=False):
    """Mean squared error of the normalised outputs and its gradient."""
    r = bf16 if bf else (lambda a: a)
    layers = unpack(p)
    Z, A = forward(p, xn, bf)
    B, nout = yn.shape
    d = Z[-1] - yn
    loss = float((d * d).sum() / (B * nout))
    dzf = 2.0 * d / (B * nout)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dz = r(dzf)
        grads[l] = (A[l].T @ dz, dzf.sum(axis=0))
        if l > 0:
            da = dz @ r(W).T
            s = 1.0 / (1.0 + np.exp(-Z[l - 1]))
            dzf = da * s * (1.0 + Z[l - 1] * (1.0 - s))
    return loss, pack(grads)


def predict(p, x, sc, bf=False):
    xn, _ = normalise(x, None, sc)
    Z, _ = forward(p, xn, bf)
    return Z[-1] * sc["sd_y"] + sc["mu_y"]


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    layers = []
    for k, n in LAYERS:
        layers.append((rng.uniform(-1, 1, (k, n)) * np.sqrt(6.0 / k) * scale,
                       rng.normal(0, 0.1, n)))
    return pack(layers)
