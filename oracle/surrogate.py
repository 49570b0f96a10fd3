"""TEST INFRASTRUCTURE ONLY — CPU oracle of the interval-wrench surrogate kernels.

Used by tests/ as the checker of the HIP MLP and knife-edge patch kernels; the product
package never imports it.  The reference has no learned component, so nothing here follows a
reference file: the network is the textbook dense ReLU network with mean squared error, and
the optimiser is Adam (Kingma and Ba, 2015, Algorithm 1).

Pins (tests/test_surrogate_oracle.py):
- `forward` and `backward` against torch autograd on the CPU (an independent implementation)
  and `backward` against central finite differences of `loss` in float64;
- `adam` against its closed form on the first step (each weight moves by -lr sign(g)) and on
  a constant gradient (the bias-corrected moments equal g and g^2 at every step), and its
  weight decay against the zero-gradient case (w shrinks by the factor 1 - lr decay);
- `bf16` against torch's float32 -> bfloat16 conversion (round to nearest, ties to even);
- `patch` against a brute-force count over the cells on tiny inputs.

Layout: layer l is the augmented matrix Wa_l [rows + 1][cols], last row the bias, and params
is Wa_1 | Wa_2 | Wa_3 flattened, as the C-ABI stores it.
"""

import numpy as np


def shapes(inputs, hidden, outputs):
    return [(inputs + 1, hidden), (hidden + 1, hidden), (hidden + 1, outputs)]


def split(params, inputs, hidden, outputs):
    out, o = [], 0
    for r, c in shapes(inputs, hidden, outputs):
        out.append(np.asarray(params[o : o + r * c]).reshape(r, c))
        o += r * c
    return out


def init(inputs, hidden, outputs, seed=0):
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for r, c in shapes(inputs, hidden, outputs):
        w = np.zeros((r, c), dtype=np.float32)
        bound = np.sqrt(6.0 / (r - 1))
        w[:-1] = rng.uniform(-bound, bound, size=(r - 1, c))
        parts.append(w.ravel())
    return np.concatenate(parts)


def _aug(x):
    return np.concatenate([x, np.ones((x.shape[0], 1), dtype=x.dtype)], axis=1)


def forward(params, x, inputs, hidden, outputs, dtype=np.float64):
    """Activations (h1, h2, out) of [x,1] W1 -> ReLU -> [h1,1] W2 -> ReLU -> [h2,1] W3."""
    W = [w.astype(dtype) for w in split(params, inputs, hidden, outputs)]
    x = np.asarray(x, dtype=dtype)
    h1 = np.maximum(_aug(x) @ W[0], 0)
    h2 = np.maximum(_aug(h1) @ W[1], 0)
    return h1, h2, _aug(h2) @ W[2]


def loss(params, x, target, inputs, hidden, outputs):
    out = forward(params, x, inputs, hidden, outputs)[2]
    return float(np.mean((out - np.asarray(target, dtype=np.float64)) ** 2))


def backward(params, x, target, inputs, hidden, outputs):
    """(loss, gradient of the mean squared error with respect to params), float64."""
    W = [w.astype(np.float64) for w in split(params, inputs, hidden, outputs)]
    x = np.asarray(x, dtype=np.float64)
    h1, h2, out = forward(params, x, inputs, hidden, outputs)
    d = out - np.asarray(target, dtype=np.float64)
    value = float(np.mean(d ** 2))
    d_out = 2.0 * d / d.size
    g3 = _aug(h2).T @ d_out
    d_h2 = (d_out @ W[2][:-1].T) * (h2 > 0)
    g2 = _aug(h1).T @ d_h2
    d_h1 = (d_h2 @ W[1][:-1].T) * (h1 > 0)
    g1 = _aug(x).T @ d_h1
    return value, np.concatenate([g1.ravel(), g2.ravel(), g3.ravel()])


def adam(params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1.0e-8, weight_decay=0.0):
    """One Adam update with decoupled weight decay (AdamW, Loshchilov and Hutter 2019,
    Algorithm 2 with schedule multiplier 1); step is 1-based.  Returns (params, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mhat = m / (1 - beta1 ** step)
    vhat = v / (1 - beta2 ** step)
    return params - lr * (mhat / (np.sqrt(vhat) + eps) + weight_decay * params), m, v


def bf16(values):
    """float32 -> bfloat16 (round to nearest, ties to even) -> float32."""
    u = np.asarray(values, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def infer_bf16(params, x, inputs, hidden, outputs, x_shift=None, x_scale=None, y_shift=None, y_scale=None):
    """The inference kernel's arithmetic: inputs standardised in float32, then weights, inputs
    and hidden activations rounded to bfloat16, products summed exactly (float64 here, float32
    on the device), outputs mapped back by y_scale and y_shift."""
    W = [bf16(w).astype(np.float64) for w in split(params, inputs, hidden, outputs)]
    x = np.asarray(x, dtype=np.float32)
    if x_shift is not None:
        x = x - np.asarray(x_shift, dtype=np.float32)
    if x_scale is not None:
        x = x * np.asarray(x_scale, dtype=np.float32)
    a = bf16(x).astype(np.float64)
    h1 = bf16(np.maximum(_aug(a) @ W[0], 0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(_aug(h1) @ W[1], 0).astype(np.float32)).astype(np.float64)
    y = _aug(h2) @ W[2]
    if y_scale is not None:
        y = y * np.asarray(y_scale, dtype=np.float64)
    if y_shift is not None:
        y = y + np.asarray(y_shift, dtype=np.float64)
    return y


def rotate_inv(q, v):
    """Rotate v by the inverse of the unit quaternion q = (x, y, z, w)."""
    u, w = np.asarray(q[:3], dtype=np.float64), float(q[3])
    v = np.asarray(v, dtype=np.float64)
    return (2 * w * w - 1) * v + 2 * (v @ u)[:, None] * u - 2 * w * np.cross(u, v)


def patch(x, damage, instance, tool_pos, tool_rot, spacing, half):
    """Per instance, particle counts and summed damage in the 8 x 8 x 8 cells around the knife
    edge: local x in [-4h, 4h], y in [-hy, hy], z in [-hz - 4h, -hz + 4h].

    Returns (out [instances][2][8][8][8], margin [n]): margin is each particle's distance, in
    cell widths, to the nearest cell face it straddles, so a test can allow for float32
    rounding at the faces."""
    x = np.asarray(x, dtype=np.float64)
    instances = len(tool_pos)
    out = np.zeros((instances, 2, 8, 8, 8))
    margin = np.full(len(x), np.inf)
    h, hy, hz = spacing, half[1], half[2]
    for i in range(instances):
        sel = np.nonzero(instance == i)[0]
        local = rotate_inv(tool_rot[i], x[sel] - np.asarray(tool_pos[i], dtype=np.float64))
        f = np.stack([(local[:, 0] + 4 * h) / h, (local[:, 1] + hy) * (4.0 / hy), (local[:, 2] + hz + 4 * h) / h], 1)
        cell = np.floor(f).astype(np.int64)
        frac = f - np.floor(f)
        margin[sel] = np.minimum(frac, 1 - frac).min(1)
        inside = ((cell >= 0) & (cell < 8)).all(1)
        for p, c in zip(sel[inside], cell[inside]):
            out[i, 0, c[0], c[1], c[2]] += 1
            out[i, 1, c[0], c[1], c[2]] += damage[p]
    return out, margin
