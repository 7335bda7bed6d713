"""numpy reference of the action surrogate's network (fjmodels_amd/csrc/fj_mlp.hpp):
5 -> 128 -> 128 -> 128 -> 2, tanh hidden layers, mean-squared-error loss.

`emulate=True` rounds to bf16 exactly where the GPU path does (inputs,
weights, hidden activations fed to the next product, output and hidden
gradients fed to products); products are summed in float64 here and in fp32
on the GPU.  Test helper only.
"""
import numpy as np

IN, H, OUT = 5, 128, 2
SIZES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("W4", (H, OUT)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round float32 values to bf16 (nearest, ties to even); returns float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def _r(x, emulate):
    return bf16(x) if emulate else np.asarray(x, dtype=np.float64)


def forward(p, x, emulate=True):
    P = unpack(p)
    f32 = (lambda v: np.asarray(v, np.float32).astype(np.float64)) if emulate else (lambda v: v)
    a = _r(x, emulate)
    hs, h16 = [], [a]
    for l in range(3):
        pre = f32(h16[-1] @ _r(P[f"W{l + 1}"], emulate))
        h = f32(np.tanh(f32(pre + P[f"b{l + 1}"])))
        hs.append(h)
        h16.append(_r(h, emulate))
    y = f32(f32(h16[-1] @ _r(P["W4"], emulate)) + P["b4"])
    return y, hs, h16


def grad(p, x, t, emulate=True):
    """Gradient of mean((y - t)^2) with respect to the flat parameters."""
    P = unpack(p)
    y, hs, h16 = forward(p, x, emulate)
    M = len(x)
    dy = _r(2 * (y - t) / (M * OUT), emulate)
    g = {"W4": h16[3].T @ dy, "b4": dy.sum(0)}
    dh = dy @ _r(P["W4"], emulate).T
    for l in (2, 1, 0):
        dp = dh * (1 - hs[l] ** 2)
        g[f"b{l + 1}"] = dp.sum(0)
        dp16 = _r(dp, emulate)
        g[f"W{l + 1}"] = h16[l].T @ dp16
        if l > 0:
            dh = dp16 @ _r(P[f"W{l + 1}"], emulate).T
    loss = float(np.mean((y - t) ** 2))
    return pack(g), loss


def glorot(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SIZES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shp[0] + shp[1]))
            d[name] = rng.uniform(-a, a, shp)
        else:
            d[name] = rng.normal(0, 0.1, shp)
    return pack(d)
