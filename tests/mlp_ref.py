"""NumPy model of the radiance-cache MLP, used by the tests as the checker of the device kernels.

The network has no counterpart in the reference (it is an addition on top of its path
tracer), so this module is pinned by mathematics instead: the analytic fp64 gradient is
checked against central finite differences, and the BF16 emulation against the fp64 model
(tests/test_mlp_ref.py).

Parameter vector (pt_cache_get_params): per layer W[out][in] row-major, then b[out]; layers
96->64, four 64->64, 64->3; ReLU after every layer but the last.

Two evaluation modes:
- exact: fp64 throughout;
- bf16: the rounding points of the device kernels -- inputs, weights, hidden activations
  and back-propagated gradients rounded to BF16 (round to nearest even), products
  accumulated in high precision, biases in fp32.
"""
import numpy as np

IN, H, OUT, LAYERS = 96, 64, 3, 6
DIMS = [(IN, H)] + [(H, H)] * 4 + [(H, OUT)]


def layout():
    """[(w_offset, b_offset, n_in, n_out)] per layer and the total parameter count."""
    out, o = [], 0
    for n_in, n_out in DIMS:
        out.append((o, o + n_in * n_out, n_in, n_out))
        o += n_in * n_out + n_out
    return out, o


def unpack(params):
    lay, _ = layout()
    return [(params[w:w + ni * no].reshape(no, ni), params[b:b + no]) for w, b, ni, no in lay]


def bf16(x):
    """Round float values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) >> 16, u >> 16)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def forward(params, X, mode="exact"):
    """Output y (n x 3) and the input of every layer (list of 6 arrays)."""
    q = bf16 if mode == "bf16" else (lambda v: v)
    ft = np.float64
    acts = []
    x = q(np.asarray(X, np.float32)).astype(ft) if mode == "bf16" else np.asarray(X, ft)
    for l, (W, b) in enumerate(unpack(np.asarray(params, np.float32 if mode == "bf16" else ft))):
        acts.append(x)
        Wq = q(W).astype(ft)
        z = x @ Wq.T + b.astype(ft)
        if l < LAYERS - 1:
            z = np.maximum(z, 0.0)
            x = q(z.astype(np.float32)).astype(ft) if mode == "bf16" else z
        else:
            return z, acts
    raise AssertionError


def loss(params, X, T, mode="exact", den=None):
    """mean over samples and channels of (y - t)^2 / den, den = y^2 + 0.01 unless given."""
    y, _ = forward(params, X, mode)
    if den is None:
        den = y * y + 0.01
    return float(np.mean((y - T) ** 2 / den)), y


def loss_and_grad(params, X, T, mode="exact"):
    """Loss and its gradient with the denominator held constant (stop-gradient)."""
    q = bf16 if mode == "bf16" else (lambda v: v)
    ft = np.float64
    n = len(X)
    y, acts = forward(params, X, mode)
    T = np.asarray(T, ft)
    den = y * y + 0.01
    L = float(np.mean((y - T) ** 2 / den))
    g = 2.0 * (y - T) / den / (3.0 * n)
    if mode == "bf16":
        g = q(g.astype(np.float32)).astype(ft)
    layers = unpack(np.asarray(params, np.float32 if mode == "bf16" else ft))
    grads = []
    for l in range(LAYERS - 1, -1, -1):
        W, _ = layers[l]
        x = acts[l]
        gW = g.T @ x
        gb = g.sum(axis=0)
        grads.append((gW, gb))
        if l == 0:
            break
        gx = g @ q(W).astype(ft)
        gx = gx * (x > 0)
        if mode == "bf16":
            gx = q(gx.astype(np.float32)).astype(ft)
        g = gx
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, flat


def he_uniform(seed):
    """The device initialisation: He-uniform weights from a Splitmix32 stream, zero biases."""
    lay, count = layout()
    p = np.zeros(count, np.float32)
    state = np.uint32(seed)
    with np.errstate(over="ignore"):
        for w, _, ni, no in lay:
            lim = np.float32(np.sqrt(np.float32(6.0) / np.float32(ni)))
            for i in range(ni * no):
                state = np.uint32(state + np.uint32(0x9E3779B9))
                z = state
                z = np.uint32((z ^ (z >> np.uint32(16))) * np.uint32(0x85EBCA6B))
                z = np.uint32((z ^ (z >> np.uint32(13))) * np.uint32(0xC2B2AE35))
                z = np.uint32(z ^ (z >> np.uint32(16)))
                p[w + i] = (np.float32(z) * np.float32(2.0 ** -32) * np.float32(2) - np.float32(1)) * lim
    return p


def adam(params, grads, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba 2015) with bias correction; step counts from 1."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def encode(records, lo, hi):
    """Input features of 13-float records (position, direction, normal, roughness, albedo)."""
    r = np.asarray(records, np.float64).reshape(-1, 13)
    lo = np.asarray(lo, np.float64)
    ext = np.asarray(hi, np.float64) - lo
    inv = np.where(ext > 0, 1.0 / np.where(ext > 0, ext, 1.0), 1.0)
    p = np.clip((r[:, :3] - lo) * inv, 0.0, 1.0)
    cols = []
    for b in range(12):
        s = (2.0 ** b) * np.pi
        for c in range(3):
            cols.append(np.sin(s * p[:, c]))
            cols.append(np.cos(s * p[:, c]))
    out = np.zeros((len(r), IN))
    out[:, :72] = np.stack(cols, axis=1)
    out[:, 72:75] = p
    out[:, 75:85] = r[:, 3:13]
    return out
