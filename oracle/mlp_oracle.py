"""CPU oracle of the surrogate MLP -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.

numpy restatement of the forward pass, mean-squared loss and backward pass of the library's
Fourier-feature MLP (3 -> 2F sin/cos features -> depth x width SiLU -> 1), with the device's
numerics emulated: every matrix-product operand rounded to bfloat16 (round to nearest even), FP32
accumulation, FP32 bias and activation, and silu'(z) of the hidden layers stored in bfloat16.  ``emulate=False`` runs the same algebra in float64 with no
rounding, which the finite-difference test pins.  The parameter vector is packed as the library
packs it: W_l (K_l x width, row-major) for every layer, then every b_l, then w_out, then b_out.
"""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u.astype(np.uint64) + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(nan, (u | 0x00400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32)


def unpack(p, F, width, depth):
    p = np.asarray(p)
    Ws, bs, off = [], [], 0
    Ks = [2 * F] + [width] * (depth - 1)
    for K in Ks:
        Ws.append(p[off:off + K * width].reshape(K, width))
        off += K * width
    for _ in range(depth):
        bs.append(p[off:off + width])
        off += width
    wout = p[off:off + width]
    off += width
    return Ws, bs, wout, p[off]


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, B, x, F, width, depth, emulate=True):
    """Returns (phi, cache) for points x (n x 3)."""
    dt = np.float32 if emulate else np.float64
    r = (lambda v: bf16(v)) if emulate else (lambda v: v.astype(np.float64))
    Ws, bs, wout, bout = unpack(p, F, width, depth)
    x = np.asarray(x, dtype=dt)
    B = np.asarray(B, dtype=dt)
    # the device forms B . x in the working precision and takes sin and cos of 2 pi times it (its
    # hardware sine works in revolutions): emulated as the float64 sine of the rounded dot product
    rev = (B[:, 0][None, :] * x[:, 0:1] + B[:, 1][None, :] * x[:, 1:2]) + B[:, 2][None, :] * x[:, 2:3]
    arg = 2 * np.pi * rev.astype(np.float64)
    a = r(np.concatenate([np.sin(arg), np.cos(arg)], axis=1).astype(dt))
    acts, zs = [a], []
    for l in range(depth):
        z = (a.astype(dt) @ r(Ws[l]).astype(dt)).astype(dt) + bs[l].astype(dt)
        zs.append(z)
        a = r(z * _sig(z))
        acts.append(a)
    phi = (a.astype(dt) @ r(wout).astype(dt)).astype(dt) + dt(bout)
    return phi, (acts, zs)


def loss_grad(p, B, x, y, F, width, depth, emulate=True):
    """Mean squared error and its gradient w.r.t. the packed parameters."""
    dt = np.float32 if emulate else np.float64
    r = (lambda v: bf16(v)) if emulate else (lambda v: v.astype(np.float64))
    Ws, bs, wout, bout = unpack(p, F, width, depth)
    phi, (acts, zs) = forward(p, B, x, F, width, depth, emulate)
    n = x.shape[0]
    res = phi - np.asarray(y, dtype=dt)
    loss = float(np.sum(res * res) / n)
    dphi = dt(2.0) * res / dt(n)
    dsil = lambda z: _sig(z) * (1 + z * (1 - _sig(z)))
    gW, gb = [None] * depth, [None] * depth
    g_wout = acts[-1].astype(dt).T @ dphi
    g_bout = np.sum(dphi)
    dz = r(dphi[:, None] * r(wout)[None, :] * dsil(zs[-1]))
    for l in range(depth - 1, -1, -1):
        gW[l] = acts[l].astype(dt).T @ dz.astype(dt)
        gb[l] = np.sum(dz.astype(dt), axis=0)
        if l > 0:
            # the device keeps silu'(z) of the hidden layers in bfloat16 between its passes
            dz = r((dz.astype(dt) @ r(Ws[l]).astype(dt).T) * r(dsil(zs[l - 1])))
    g = np.concatenate([w.reshape(-1) for w in gW] + list(gb) + [g_wout, [g_bout]])
    return loss, g
