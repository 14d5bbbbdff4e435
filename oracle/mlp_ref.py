"""FP64 numpy reference of the warm-start MLP (test infrastructure only).

Restates the network the GPU kernels implement (14-128-128-128-2, SiLU,
row-major [in][out] weights; see include/shapes_gpu.h) with the same BF16
roundings of inputs, weights and hidden activations, but FP64 arithmetic
everywhere else, so the GPU's forward pass and gradients can be checked
against it.
"""
import numpy as np

IN, HID = 14, 128
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (HID, 2)), ("b4", (2,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    out = {}
    k = 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[n], dtype=np.float64).ravel() for n, _ in SHAPES])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), shape)
        else:
            d[name] = np.zeros(shape)
    d["W4"] *= 0.1
    return pack(d).astype(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, X, emulate_bf16=True):
    p = unpack(params)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(X)
    zs, hs = [], [x]
    h = x
    for L in ("1", "2", "3"):
        z = h @ r(p["W" + L]) + p["b" + L]
        h = r(silu(z).astype(np.float32)) if emulate_bf16 else silu(z)
        zs.append(z)
        hs.append(h)
    y = h @ r(p["W4"]) + p["b4"]
    return y, zs, hs


def gradient(params, X, T, emulate_bf16=True):
    """Gradient of mean over rows of sum_k (y_k - t_k)^2."""
    p = unpack(params)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, zs, hs = forward(params, X, emulate_bf16)
    n = len(X)
    dy = 2.0 * (y - T) / n
    g = {"W4": hs[3].T @ dy, "b4": dy.sum(0)}
    dh = dy @ r(p["W4"]).T
    for L, idx in (("3", 2), ("2", 1), ("1", 0)):
        dz = dh * dsilu(zs[idx])
        g["W" + L] = hs[idx].T @ dz
        g["b" + L] = dz.sum(0)
        dh = dz @ r(p["W" + L]).T
    loss = float(((y - T) ** 2).sum() / n)
    return pack(g), loss
