"""numpy reference of the boundary MLP for tests (forward, loss, backward).

Parameter layout matches the engine: per layer W[in][out] then bias[out].
With ``bf16=True`` the operands of every product are rounded to BF16 the way
the engine's MFMA GEMMs round them (activations, weights, output gradients);
accumulation is float64.  With ``bf16=False`` the function is smooth, which
is what the finite-difference check of the backward pass needs.
"""
import numpy as np

SIZES = [(64, 128), (128, 128), (128, 128), (128, 16)]
C_GELU = np.sqrt(2.0 / np.pi)


def offsets():
    out, o = [], 0
    for i, j in SIZES:
        out.append((o, o + i * j))
        o += i * j + j
    return out, o


def bf16_round(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * x * x)


def unpack(params, bf16):
    offs, _ = offsets()
    Ws, bs = [], []
    for (i, j), (ow, ob) in zip(SIZES, offs):
        W = np.asarray(params[ow:ow + i * j], np.float64).reshape(i, j)
        b = np.asarray(params[ob:ob + j], np.float64)
        if bf16:
            W = bf16_round(W)
            b = bf16_round(b)
        Ws.append(W)
        bs.append(b)
    return Ws, bs


def forward(params, x, bf16=False):
    Ws, bs = unpack(params, bf16)
    r = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    a = np.asarray(x, np.float64)
    acts, zs = [a], []
    for l in range(4):
        z = r(a.astype(np.float32) if bf16 else a) @ Ws[l] + bs[l]
        zs.append(z)
        a = gelu(z) if l < 3 else z
        acts.append(a)
    return zs, acts


def loss_and_grad(params, x, y, bf16=False):
    """Mean squared error over all outputs and its gradient (flat, layout of
    the parameters)."""
    Ws, bs = unpack(params, bf16)
    r = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    zs, acts = forward(params, x, bf16)
    yh = zs[3]
    n = yh.size
    d = yh - y
    loss = float(np.sum(d * d) / n)
    dz = 2.0 * d / n
    offs, total = offsets()
    g = np.zeros(total)
    for l in range(3, -1, -1):
        xin = acts[l]
        (ow, ob), (i, j) = offs[l], SIZES[l]
        g[ow:ow + i * j] = (r(xin.astype(np.float32) if bf16 else xin).T @
                            r(dz.astype(np.float32) if bf16 else dz)).ravel()
        g[ob:ob + j] = dz.sum(axis=0)
        if l > 0:
            da = r(dz.astype(np.float32) if bf16 else dz) @ Ws[l].T
            dz = da * gelu_grad(zs[l - 1])
    return loss, g


def random_params(rng, scale=1.0):
    offs, total = offsets()
    p = np.zeros(total, np.float32)
    for (i, j), (ow, ob) in zip(SIZES, offs):
        a = scale * np.sqrt(3.0 / i)
        p[ow:ow + i * j] = rng.uniform(-a, a, i * j)
        p[ob:ob + j] = rng.uniform(-0.1, 0.1, j)
    return p
