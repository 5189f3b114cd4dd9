"""CPU oracle of the learned tuner's MLP — TEST INFRASTRUCTURE ONLY.

Restates the GPU kernels' arithmetic in numpy: the matrix products take
BF16-rounded operands (round to nearest even) and accumulate in FP32 (here in
FP64, then rounded to FP32, which bounds the GPU's FP32 accumulation error);
hidden layers ReLU, linear output, mean squared error over the real output
columns, gradients as the GPU forms them (dZ rounded to BF16 before the
products).  With round=False every rounding is dropped and the gradients are
the exact derivatives of the loss, which tests/test_oracle_mlp.py pins against
central finite differences.

Parameter layout (as faz_mlp_get_params): per layer W [K x N] row-major (padded
shapes) then b [N].
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    nan = np.isnan(x)
    out = r.view(np.float32).copy()
    out[nan] = x[nan]
    return out


def split(params, K, N):
    ws, bs, at = [], [], 0
    for k, n in zip(K, N):
        ws.append(params[at:at + k * n].reshape(k, n))
        at += k * n
        bs.append(params[at:at + n])
        at += n
    return ws, bs


def forward(params, K, N, x, round=True):
    """Outputs (padded width) and the per-layer (input, pre-activation)."""
    rd = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if round else (lambda a: a)
    ws, bs = split(params, K, N)
    B = x.shape[0]
    a = np.zeros((B, K[0]), np.float64)
    a[:, :x.shape[1]] = x
    a = rd(a)
    cache = []
    for l, (w, b) in enumerate(zip(ws, bs)):
        z = a @ rd(w)
        if round:
            z = z.astype(np.float32).astype(np.float64)
        z = z + b
        if round:
            z = z.astype(np.float32).astype(np.float64)
        cache.append((a, z))
        if l + 1 < len(ws):
            a = rd(np.maximum(z, 0.0))
    return cache[-1][1], cache


def gradients(params, K, N, x, t, round=True):
    """(loss, flat gradient in the parameter layout)."""
    rd = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if round else (lambda a: a)
    ws, _ = split(params, K, N)
    y, cache = forward(params, K, N, x, round)
    B, dout = t.shape
    d = y[:, :dout] - t
    loss = float((d * d).sum() / (B * dout))
    dz = np.zeros_like(y)
    dz[:, :dout] = 2.0 * d / (B * dout)
    dz = rd(dz)
    grads = [None] * len(ws)
    for l in range(len(ws) - 1, -1, -1):
        a, z = cache[l]
        gw = a.T @ dz
        gb = dz.sum(axis=0)
        grads[l] = (gw, gb)
        if l:
            da = dz @ rd(ws[l]).T
            if round:
                da = da.astype(np.float32).astype(np.float64)
            _, zp = cache[l - 1]
            dz = rd(np.where(zp > 0, da, 0.0))
    flat = np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])
    return loss, flat
