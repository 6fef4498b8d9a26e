"""Oracle (test infrastructure only) of the learned denoiser: the same U-Net in
numpy float64 with BF16 rounding (round to nearest even) at the points where
the device stores BF16, so the device's FP32-accumulated matrix-core results
can be compared with it to accumulation-order tolerance.  The denoiser has no
counterpart in the reference; its layer maths is pinned by the tests against
torch.nn.functional.conv2d (CPU, float64) as a library routine.

Layouts follow the device: images NHWC, conv weights W[tap][ci][co] with
tap = 3 * ky + kx, biases b[co].
"""
from __future__ import annotations

import numpy as np

NET = [(10, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 128, 1), (128, 128, 1), (192, 64, 1), (64, 64, 1),
       (96, 32, 1), (32, 3, 0)]


ROUND = True  # False: plain float64 everywhere (to check the maths against autograd)


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    if not ROUND:
        return np.asarray(a, np.float64)
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None, relu=False):
    """y[b,h,w,co] = sum_{ky,kx,ci} x[b, h+ky-1, w+kx-1, ci] w[3ky+kx, ci, co] (+ b), zero padded."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, W, w.shape[2]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        y += np.einsum("bhwc,cd->bhwd", xp[:, ky:ky + H, kx:kx + W, :], w[t], optimize=True)
    if b is not None:
        y += b
    return np.maximum(y, 0.0) if relu else y


def conv3x3_data_grad(g, w):
    """dx[b,h,w,ci] = sum_{tap,co} g[b, h-ky+1, w-kx+1, co] w[tap, ci, co]."""
    B, H, W, _ = g.shape
    gp = np.pad(g, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dx = np.zeros((B, H, W, w.shape[1]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dx += np.einsum("bhwd,cd->bhwc", gp[:, 2 - ky:2 - ky + H, 2 - kx:2 - kx + W, :], w[t], optimize=True)
    return dx


def conv3x3_weight_grad(x, g):
    """dW[tap, ci, co] = sum_p x[p + shift(tap), ci] g[p, co]; db = sum_p g[p]."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((9, x.shape[3], g.shape[3]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dw[t] = np.einsum("bhwc,bhwd->cd", xp[:, ky:ky + H, kx:kx + W, :], g, optimize=True)
    return dw, g.sum(axis=(0, 1, 2))


def pool2(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def split_params(p):
    """Flat parameter vector -> [(W[9][cin][cout], b[cout])] per layer."""
    out, off = [], 0
    for cin, cout, _ in NET:
        w = p[off:off + 9 * cin * cout].reshape(9, cin, cout)
        off += 9 * cin * cout
        b = p[off:off + cout]
        off += cout
        out.append((np.asarray(w, np.float64), np.asarray(b, np.float64)))
    assert off == len(p)
    return out


def features(rad, aux):
    """Network input: log1p(radiance), albedo, normal / 2 + 1/2, log1p(depth)."""
    f = np.concatenate([np.log1p(np.maximum(rad, 0)), aux[..., 0:3], aux[..., 3:6] * 0.5 + 0.5,
                        np.log1p(np.maximum(aux[..., 6:7], 0))], axis=-1)
    return bf16(f.astype(np.float32) if ROUND else f)


def forward(params, rad, aux):
    """U-Net forward; returns (residual y [B,H,W,3], activations for backward)."""
    L = split_params(params)
    Wb = [(bf16(w.astype(np.float32)) if ROUND else w, b) for w, b in L]
    A = {}
    A["x0"] = features(rad, aux)

    def layer(i, x):
        w, b = Wb[i]
        y = conv3x3(x, w, b, NET[i][2])
        return bf16(y.astype(np.float32)) if ROUND else y

    A["a0"] = layer(0, A["x0"])
    A["s0"] = layer(1, A["a0"])
    A["p1"] = bf16(pool2(A["s0"]).astype(np.float32)) if ROUND else pool2(A["s0"])
    A["a1"] = layer(2, A["p1"])
    A["s1"] = layer(3, A["a1"])
    A["p2"] = bf16(pool2(A["s1"]).astype(np.float32)) if ROUND else pool2(A["s1"])
    A["a2"] = layer(4, A["p2"])
    A["b2"] = layer(5, A["a2"])
    A["cat1"] = np.concatenate([up2(A["b2"]), A["s1"]], axis=-1)
    A["d1a"] = layer(6, A["cat1"])
    A["d1b"] = layer(7, A["d1a"])
    A["cat0"] = np.concatenate([up2(A["d1b"]), A["s0"]], axis=-1)
    A["d0a"] = layer(8, A["cat0"])
    w, b = Wb[9]
    y = conv3x3(A["d0a"], w, b, False)
    return y, A, Wb


def denoise(params, rad, aux):
    y, _, _ = forward(params, rad, aux)
    return np.expm1(np.maximum(np.log1p(np.maximum(rad, 0)) + y, 0))


def loss_grad(params, rad, aux, target):
    """L1 loss of log1p(rad) + y against log1p(target), mean over pixel-channels,
    and its gradient with respect to the flat parameter vector."""
    y, A, Wb = forward(params, rad, aux)
    n = y.size
    d = np.log1p(np.maximum(rad, 0)) + y - np.log1p(np.maximum(target, 0))
    loss = np.abs(d).sum() / n
    dy = np.sign(d) / n
    grads = [None] * len(NET)

    def back(i, x, y_out, dy_out):
        w, _ = Wb[i]
        g = dy_out * (y_out > 0) if NET[i][2] else dy_out
        g = bf16(g.astype(np.float32)) if ROUND else g
        grads[i] = conv3x3_weight_grad(x, g)
        return conv3x3_data_grad(g, w)

    d_d0a = back(9, A["d0a"], None, dy)
    d_cat0 = back(8, A["cat0"], A["d0a"], d_d0a)
    d_up = d_cat0[..., :64]
    B, H, W, _ = d_up.shape
    d_d1b = d_up.reshape(B, H // 2, 2, W // 2, 2, 64).sum(axis=(2, 4))
    d_d1a = back(7, A["d1a"], A["d1b"], d_d1b)
    d_cat1 = back(6, A["cat1"], A["d1a"], d_d1a)
    d_up = d_cat1[..., :128]
    B1, H1, W1, _ = d_up.shape
    d_b2 = d_up.reshape(B1, H1 // 2, 2, W1 // 2, 2, 128).sum(axis=(2, 4))
    d_a2 = back(5, A["a2"], A["b2"], d_b2)
    d_p2 = back(4, A["p2"], A["a2"], d_a2)
    d_s1 = d_cat1[..., 128:] + up2(d_p2) * 0.25
    d_a1 = back(3, A["a1"], A["s1"], d_s1)
    d_p1 = back(2, A["p1"], A["a1"], d_a1)
    d_s0 = d_cat0[..., 64:] + up2(d_p1) * 0.25
    d_a0 = back(1, A["a0"], A["s0"], d_s0)
    back(0, A["x0"], A["a0"], d_a0)
    flat = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in grads])
    return loss, flat
