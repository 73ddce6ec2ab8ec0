"""ORACLE (test infrastructure only): binary64 restatement of the learned
proposal model's loss and gradient.

Only tests/ use this module, as the checker of the device kernels.  The
model is 48 -> 128 -> 128 -> 32 with ReLU; parameters are one vector in the
order W1[48][128], b1[128], W2[128][128], b2[128], W3[128][32], b3[32]
(row-major [in][out]).  The loss is the pos_weight-weighted binary cross
entropy of the 32 logits against labels Y over valid slots M, divided by
32 n.  With ``rounding=True`` the values the device stores in bfloat16 --
the inputs, the weights, the hidden activations and the output, second and
first layer gradients -- are rounded to bfloat16 (nearest even) at the same
points; everything else is binary64.  With ``rounding=False`` it is the
plain binary64 model, whose gradient the tests check against central finite
differences.
"""
import numpy as np

IN, HID, OUT = 48, 128, 32
SHAPES = (("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)),
          ("b2", (HID,)), ("W3", (HID, OUT)), ("b3", (OUT,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)  # 26912


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[n], np.float64).ravel()
                           for n, _ in SHAPES])


def _layers(params, X, rounding):
    """Weights as the device holds them and the forward pass: returns
    (W1, W2, W3, h1, h2, z3)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    P = unpack(params)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    b1 = P["b1"].astype(np.float32).astype(np.float64) if rounding else P["b1"]
    b2 = P["b2"].astype(np.float32).astype(np.float64) if rounding else P["b2"]
    b3 = P["b3"].astype(np.float32).astype(np.float64) if rounding else P["b3"]
    X = np.asarray(X, np.float64)
    h1 = r(np.maximum(X @ W1 + b1, 0.0))
    h2 = r(np.maximum(h1 @ W2 + b2, 0.0))
    z3 = h2 @ W3 + b3
    return W1, W2, W3, h1, h2, z3


def forward(params, X, rounding=True):
    """Logits (n, 32) of the model for inputs X (n, 48); with ``rounding``
    the weights and hidden activations pass through bfloat16 as on the
    device (the inference kernel's rounding points are the training
    kernel's)."""
    return _layers(params, X, rounding)[5]


def loss_and_grad(params, X, Y, M, pos_weight, rounding=True):
    """X: (n, 48) float64 inputs (already bf16 values when rounding)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3, h1, h2, z3 = _layers(params, X, rounding)
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    inv = 1.0 / (n * OUT)
    w = np.where(Y > 0.5, pos_weight, 1.0)
    p = 1.0 / (1.0 + np.exp(-z3))
    sp = np.maximum(z3, 0.0) + np.log1p(np.exp(-np.abs(z3)))
    loss = float(np.sum(M * w * (sp - Y * z3)) * inv)
    dz3 = r(M * w * (p - Y) * inv)
    dz2 = r(np.where(h2 > 0, dz3 @ W3.T, 0.0))
    dz1 = r(np.where(h1 > 0, dz2 @ W2.T, 0.0))
    g = {"W1": X.T @ dz1, "b1": dz1.sum(0), "W2": h1.T @ dz2,
         "b2": dz2.sum(0), "W3": h2.T @ dz3, "b3": dz3.sum(0)}
    return loss, pack(g)


def loss_only(params, X, Y, M, pos_weight):
    return loss_and_grad(params, X, Y, M, pos_weight, rounding=False)[0]


def xavier(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = rng.normal(0, 0.05, shape)
    return pack(d)
