"""CPU oracle of the proposal network (test infrastructure only: imported by
tests and never by the product path).

The network is the engine's own design -- the reference has no learned
component -- so this oracle restates the documented computation in numpy
FP64:

    H1 = relu(X  W1 + b1)    X  [B][64],  W1 [64][256]
    H2 = relu(H1 W2 + b2)    W2 [256][256]
    H3 = relu(H2 W3 + b3)    W3 [256][256]
    y  = H3 w4 + b4          w4 [256]
    loss = mean((y - t)^2)

Parameters are one flat FP32 vector in the order W1, W2, W3, w4, b1, b2, b3,
b4 (row-major matrices).  With ``emulate=True`` the oracle rounds to bfloat16
exactly where the GPU kernels store bf16: X, the matrix weights W1..W3 used
by the GEMMs, every hidden activation H1..H3 and every back-propagated
activation gradient; w4, the biases, y and all accumulations stay wide.
With ``emulate=False`` it is the plain real-valued network, whose gradient
the tests pin with torch autograd and central finite differences.
"""
import numpy as np

IN, HID = 64, 256
SHAPES = [("W1", (IN, HID)), ("W2", (HID, HID)), ("W3", (HID, HID)), ("w4", (HID,)),
          ("b1", (HID,)), ("b2", (HID,)), ("b3", (HID,)), ("b4", (1,))]
OFFSETS = {}
_o = 0
for _name, _shape in SHAPES:
    OFFSETS[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
N_PARAMS = _o
N_PARAMS_BF16 = OFFSETS["w4"][0]


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64.
    Values are first rounded to float32, as the kernels hold them."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def bf16_bits(x):
    """bfloat16 bit patterns (uint16) of x, rounded as bf16_round."""
    f = np.asarray(bf16_round(x), dtype=np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params):
    params = np.asarray(params, dtype=np.float64)
    out = {}
    for name, (o, shape) in OFFSETS.items():
        out[name] = params[o:o + int(np.prod(shape))].reshape(shape)
    return out


def pack(parts):
    flat = np.zeros(N_PARAMS, dtype=np.float64)
    for name, (o, shape) in OFFSETS.items():
        flat[o:o + int(np.prod(shape))] = np.asarray(parts[name], dtype=np.float64).ravel()
    return flat


def init_params(seed):
    """He-normal matrices, zero biases; the fixed initialisation the trainer
    uses (numpy PCG64 from the seed)."""
    rng = np.random.default_rng(seed)
    parts = {
        "W1": rng.standard_normal((IN, HID)) * np.sqrt(2.0 / IN),
        "W2": rng.standard_normal((HID, HID)) * np.sqrt(2.0 / HID),
        "W3": rng.standard_normal((HID, HID)) * np.sqrt(2.0 / HID),
        "w4": rng.standard_normal(HID) * np.sqrt(1.0 / HID),
        "b1": np.zeros(HID), "b2": np.zeros(HID), "b3": np.zeros(HID), "b4": np.zeros(1),
    }
    return pack(parts).astype(np.float32)


def _r(x, emulate):
    return bf16_round(x) if emulate else x


def forward(params, X, emulate=True):
    """y [B] and the cached activations."""
    P = unpack(params)
    X = _r(np.asarray(X, dtype=np.float64), emulate)
    W = [_r(P["W1"], emulate), _r(P["W2"], emulate), _r(P["W3"], emulate)]
    b = [P["b1"], P["b2"], P["b3"]]
    H = [X]
    for i in range(3):
        z = H[-1] @ W[i] + b[i]
        H.append(_r(np.maximum(z, 0.0), emulate))
    y = H[3] @ P["w4"] + P["b4"][0]
    return y, {"H": H, "W": W}


def loss_and_grad(params, X, t, emulate=True):
    """(loss, flat gradient [N_PARAMS], per-row squared error)."""
    P = unpack(params)
    t = np.asarray(t, dtype=np.float64)
    B = t.shape[0]
    y, c = forward(params, X, emulate)
    H, W = c["H"], c["W"]
    d = y - t
    sq = d * d
    dy = 2.0 * d / B
    g = {"w4": H[3].T @ dy, "b4": np.array([dy.sum()])}
    G = _r(np.where(H[3] > 0, np.outer(dy, P["w4"]), 0.0), emulate)
    for i in (2, 1, 0):
        g[f"W{i + 1}"] = H[i].T @ G
        g[f"b{i + 1}"] = G.sum(axis=0)
        if i > 0:
            G = _r(np.where(H[i] > 0, G @ W[i].T, 0.0), emulate)
    return sq.mean(), pack(g), sq


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected, step counted from 1); returns new
    (params, m, v) in float64."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return params - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
