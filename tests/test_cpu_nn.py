"""The denoiser's oracle (oracle/nn_oracle.py) against torch's CPU float64
convolution and autograd, used as library routines: the 3x3 convolution, its
data and weight gradients, and the whole U-Net's loss gradient with the BF16
rounding switched off.  CPU only."""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch
import torch.nn.functional as F

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import nn_oracle as NN  # noqa: E402


def to_torch_w(w):  # W[tap][ci][co] -> [co][ci][3][3]
    return torch.from_numpy(w.reshape(3, 3, w.shape[1], w.shape[2]).transpose(3, 2, 0, 1).copy())


@pytest.mark.parametrize("cin,cout", [(10, 32), (32, 16), (64, 3)])
def test_conv_forward_and_gradients_vs_torch(cin, cout):
    rng = np.random.default_rng(cin + cout)
    x = rng.normal(size=(2, 6, 7, cin))
    w = rng.normal(size=(9, cin, cout))
    b = rng.normal(size=cout)
    g = rng.normal(size=(2, 6, 7, cout))
    xt = torch.from_numpy(x.transpose(0, 3, 1, 2).copy()).requires_grad_()
    wt = to_torch_w(w).requires_grad_()
    bt = torch.from_numpy(b).requires_grad_()
    yt = F.conv2d(xt, wt, bt, padding=1)
    y = NN.conv3x3(x, w, b)
    assert np.allclose(y, yt.detach().numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-10)
    yt.backward(torch.from_numpy(g.transpose(0, 3, 1, 2).copy()))
    dx = NN.conv3x3_data_grad(g, w)
    dw, db = NN.conv3x3_weight_grad(x, g)
    assert np.allclose(dx, xt.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-10)
    assert np.allclose(to_torch_w(dw).numpy(), wt.grad.numpy(), rtol=1e-12, atol=1e-10)
    assert np.allclose(db, bt.grad.numpy())


def torch_unet(params, feat):
    """The same network in torch (float64), from the flat parameter vector."""
    L = NN.split_params(params)
    W = [(to_torch_w(w).requires_grad_(), torch.from_numpy(b.copy()).requires_grad_()) for w, b in L]

    def c(i, x):
        w, b = W[i]
        y = F.conv2d(x, w, b, padding=1)
        return torch.relu(y) if NN.NET[i][2] else y

    x0 = torch.from_numpy(feat.transpose(0, 3, 1, 2).copy())
    s0 = c(1, c(0, x0))
    s1 = c(3, c(2, F.avg_pool2d(s0, 2)))
    b2 = c(5, c(4, F.avg_pool2d(s1, 2)))
    d1 = c(7, c(6, torch.cat([F.interpolate(b2, scale_factor=2, mode="nearest"), s1], 1)))
    y = c(9, c(8, torch.cat([F.interpolate(d1, scale_factor=2, mode="nearest"), s0], 1)))
    return y, W


def test_unet_loss_gradient_vs_torch_autograd():
    rng = np.random.default_rng(3)
    n = sum(9 * a * b + b for a, b, _ in NN.NET)
    params = rng.normal(size=n) * 0.05
    rad = rng.gamma(2.0, 0.5, size=(1, 8, 12, 3))
    aux = np.concatenate([rng.random((1, 8, 12, 3)), rng.normal(size=(1, 8, 12, 3)), rng.random((1, 8, 12, 1))], -1)
    target = rng.gamma(2.0, 0.5, size=(1, 8, 12, 3))
    NN.ROUND = False
    try:
        loss, g = NN.loss_grad(params, rad, aux, target)
        feat = NN.features(rad, aux)
    finally:
        NN.ROUND = True
    y, W = torch_unet(params, feat)
    pred = torch.log1p(torch.from_numpy(rad.transpose(0, 3, 1, 2).copy())) + y
    lt = (pred - torch.log1p(torch.from_numpy(target.transpose(0, 3, 1, 2).copy()))).abs().mean()
    lt.backward()
    gt = np.concatenate([np.concatenate([to_torch_w_inv(w.grad.numpy()).ravel(), b.grad.numpy()]) for w, b in W])
    assert abs(loss - lt.item()) < 1e-12
    assert np.allclose(g, gt, rtol=1e-9, atol=1e-12)


def to_torch_w_inv(wt):  # [co][ci][3][3] -> W[tap][ci][co]
    co, ci = wt.shape[:2]
    return wt.transpose(2, 3, 1, 0).reshape(9, ci, co)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -2.0e-3, 3.0e38, np.inf], np.float32)
    r = NN.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert np.isinf(r[5])
    u = r.astype(np.float32).view(np.uint32)
    assert np.all(u & 0xFFFF == 0)
