"""The denoiser's matrix-core kernels against the float64 oracle
(oracle/nn_oracle.py) on BF16-rounded inputs: a 3x3 convolution forward, its
data and weight gradients, and the whole U-Net's loss and parameter gradient.

Bars: single layers within 1e-4 of the output scale (FP32 accumulation of
exactly representable BF16 products, so only the summation order differs);
the U-Net loss within 1e-4 relative and the gradient within 2% relative L2
(BF16 rounding of intermediate activations can flip where FP32 sums differ in
the last bit); and training on rendered pairs lowers the loss.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import linrt

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import nn_oracle as NN  # noqa: E402

pytestmark = pytest.mark.gpu


def bfr(a):
    return NN.bf16(np.asarray(a, np.float32))


@pytest.mark.parametrize("cin,cout,relu", [(10, 32, 1), (32, 32, 1), (64, 128, 0), (192, 64, 1), (96, 32, 1), (32, 3, 0),
                                           (128, 128, 1)])
def test_conv_forward(cin, cout, relu):
    rng = np.random.default_rng(cin * 7 + cout)
    x = rng.normal(size=(2, 5, 70, cin)).astype(np.float32)
    w = (rng.normal(size=(9, cin, cout)) / np.sqrt(9 * cin)).astype(np.float32)
    b = rng.normal(size=cout).astype(np.float32)
    y = linrt.conv3x3(0, w, x=x, b=b, relu=relu)
    ref = NN.conv3x3(bfr(x), bfr(w), b, relu)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"conv {cin}->{cout}: max err / scale {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("cin,cout", [(32, 32), (10, 3)])
def test_conv_forward_many_tiles(cin, cout):
    """A frame of 256 x 1000 pixels (1,024 tiles of 64 x 4, more than the
    GPU holds blocks at once) through a single-chunk layer: every block
    walks several tiles, staging each next tile while computing the current
    one, and the last column of tiles is partial."""
    rng = np.random.default_rng(cin * 3 + cout)
    x = rng.normal(size=(1, 256, 1000, cin)).astype(np.float32)
    w = (rng.normal(size=(9, cin, cout)) / np.sqrt(9 * cin)).astype(np.float32)
    b = rng.normal(size=cout).astype(np.float32)
    y = linrt.conv3x3(0, w, x=x, b=b, relu=1)
    ref = NN.conv3x3(bfr(x), bfr(w), b, True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"conv {cin}->{cout} over 256 x 1000: max err / scale {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("cin,cout", [(10, 32), (32, 3), (64, 128), (192, 64), (96, 32)])
def test_conv_data_grad(cin, cout):
    rng = np.random.default_rng(cin * 11 + cout)
    g = rng.normal(size=(2, 6, 67, cout)).astype(np.float32)
    w = (rng.normal(size=(9, cin, cout)) / np.sqrt(9 * cout)).astype(np.float32)
    dx = linrt.conv3x3(1, w, g=g)
    ref = NN.conv3x3_data_grad(bfr(g), bfr(w))
    err = np.abs(dx - ref).max() / np.abs(ref).max()
    print(f"data grad {cin}->{cout}: max err / scale {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("cin,cout", [(10, 32), (32, 3), (64, 128), (192, 64), (32, 16)])
def test_conv_weight_grad(cin, cout):
    rng = np.random.default_rng(cin * 13 + cout)
    x = rng.normal(size=(2, 7, 45, cin)).astype(np.float32)
    g = rng.normal(size=(2, 7, 45, cout)).astype(np.float32)
    w = np.zeros((9, cin, cout), np.float32)
    dw, db = linrt.conv3x3(2, w, x=x, g=g)
    rdw, rdb = NN.conv3x3_weight_grad(bfr(x), bfr(g))
    err = np.abs(dw - rdw).max() / np.abs(rdw).max()
    errb = np.abs(db - rdb).max() / np.abs(rdb).max()
    print(f"weight grad {cin}->{cout}: max err / scale {err:.2e}, bias {errb:.2e}")
    assert err < 1e-4 and errb < 1e-4


def synthetic_batch(rng, B, H, W):
    """Smooth radiance with albedo / normal structure plus heavy-tailed noise."""
    yy, xx = np.mgrid[0:H, 0:W] / max(H, W)
    clean = np.stack([0.5 + 0.4 * np.sin(6 * xx + k) * np.cos(4 * yy) for k in range(3)], -1)
    clean = np.broadcast_to(clean, (B, H, W, 3)) * rng.uniform(0.5, 2.0, size=(B, 1, 1, 3))
    noisy = clean * rng.gamma(1.0, 1.0, size=(B, H, W, 3))
    aux = np.concatenate([np.clip(clean / 2, 0, 1), np.broadcast_to([0, 1, 0], (B, H, W, 3)),
                          np.full((B, H, W, 1), 2.0)], -1)
    return noisy.astype(np.float32), aux.astype(np.float32), clean.astype(np.float32)


def test_unet_loss_and_gradient_vs_oracle():
    rng = np.random.default_rng(5)
    den = linrt.Denoiser(seed=3)
    p = den.params
    # larger output weights so every layer carries gradient
    rad, aux, tgt = synthetic_batch(rng, 2, 16, 24)
    loss, g = den.loss_grad(rad, aux, tgt)
    rloss, rg = NN.loss_grad(p.astype(np.float64), rad.astype(np.float64), aux.astype(np.float64),
                             tgt.astype(np.float64))
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print(f"loss {loss:.6f} vs {rloss:.6f}; gradient relative L2 error {rel:.2e}")
    assert abs(loss - rloss) < 1e-4 * abs(rloss)
    assert rel < 2e-2
    off = 0
    for i, (cin, cout, _) in enumerate(NN.NET):
        n = 9 * cin * cout + cout
        gi, ri = g[off:off + n], rg[off:off + n]
        off += n
        li = np.linalg.norm(gi - ri) / max(np.linalg.norm(ri), 1e-30)
        print(f"  layer {i} ({cin}->{cout}): rel L2 {li:.2e}")
        assert li < 5e-2
    out = den.forward(rad, aux)
    ref = NN.denoise(p.astype(np.float64), rad.astype(np.float64), aux.astype(np.float64))
    assert np.abs(out - ref).max() < 1e-2 * np.abs(ref).max()


def test_forward_with_biases_vs_oracle():
    """The fused forward (features read while staging, pooled / upsampled /
    concatenated inputs, the image written by the last layer) with nonzero
    biases in every layer, on a frame whose width is not a multiple of the
    64-pixel tile, against the float64 oracle of the network."""
    rng = np.random.default_rng(11)
    den = linrt.Denoiser(seed=5)
    p = den.params.copy()
    off = 0
    for cin, cout, _ in NN.NET:
        off += 9 * cin * cout
        p[off:off + cout] = rng.normal(scale=0.2, size=cout)
        off += cout
    den.params = p
    rad, aux, _ = synthetic_batch(rng, 2, 20, 72)
    out = den.forward(rad, aux)
    ref = NN.denoise(p.astype(np.float64), rad.astype(np.float64), aux.astype(np.float64))
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"forward with biases: max err / scale {err:.2e}")
    assert err < 1e-2
    # the output layer's bias moves every channel: a zero-bias forward differs
    q = p.copy()
    q[off - 3:off] = 0.0
    den.params = q
    assert np.abs(den.forward(rad, aux) - out).max() > 10 * err * np.abs(ref).max()


def test_device_inputs_written_by_torch_just_before_the_call():
    """Inputs that torch writes on its default stream right before a
    device-pointer call are complete when the call's kernels read them (the
    library's streams are non-blocking, so the call orders itself after the
    null stream's work): a frame whose radiance is recomputed into a block
    that last held NaNs, each time just before the forward, denoises to the
    host-pointer forward's output every time."""
    import torch
    rng = np.random.default_rng(21)
    B, H, W = 1, 540, 960
    rad, aux, _ = synthetic_batch(rng, B, H, W)
    den = linrt.Denoiser(seed=3)
    ref = den.forward(rad, aux)
    r64 = torch.from_numpy(rad.astype(np.float64)).cuda()
    a = torch.from_numpy(aux).cuda()
    torch.cuda.synchronize()
    for k in range(4):
        junk = torch.full(rad.shape, float("nan"), dtype=torch.float32, device="cuda")
        del junk  # the caching allocator hands this block to r
        r = (r64 * 1.0).float()
        out = torch.full_like(r, float("nan"))
        den.forward_device(r.data_ptr(), a.data_ptr(), B, H, W, out.data_ptr())
        got = out.cpu().numpy()
        assert np.isfinite(got).all(), f"call {k}: non-finite output"
        assert np.array_equal(got, ref), f"call {k}: differs from the host-pointer forward"


def test_training_lowers_loss():
    rng = np.random.default_rng(9)
    den = linrt.Denoiser(seed=4)
    rad, aux, tgt = synthetic_batch(rng, 4, 32, 32)
    first = den.loss_grad(rad, aux, tgt, grads=False)[0]
    for _ in range(60):
        den.loss_grad(rad, aux, tgt, grads=False)
        den.adam_step(1e-3)
    last = den.loss_grad(rad, aux, tgt, grads=False)[0]
    print(f"loss {first:.4f} -> {last:.4f}")
    assert last < 0.7 * first


def test_trained_denoiser_gain():
    """The committed network on a held-out scene (seed and sky outside the
    training set), a 1920-wide band at 4 spp: its relMSE against a 512-spp
    render of the same band is far below the noisy input's, i.e. what the
    plain path tracer would need many times the samples to reach."""
    import torch
    from linrt import denoise as dn
    sky = linrt.load_sky_set()[5]
    gen = linrt.displaced_sphere_variant(seed=9001, sky=sky)
    sc = linrt.Scene(gen.desc_ptr)
    W, H, rows = 1920, 1080, (300, 428)
    rad = dn.radiance(sc, W, H, 4, 9001, 0, rows=rows)[rows[0]:rows[1]].contiguous()
    aux = dn.features(sc, W, H, 4, 9001, 0, rows=rows)[rows[0]:rows[1]].contiguous()
    tgt = dn.radiance(sc, W, H, 512, 9001, 1 << 20, rows=rows)[rows[0]:rows[1]].cpu().numpy()
    den = linrt.Denoiser()
    dn.load(den)
    out = torch.zeros_like(rad)
    den.forward_device(rad.data_ptr(), aux.data_ptr(), 1, rows[1] - rows[0], W, out.data_ptr())
    torch.cuda.synchronize()
    e_in = dn.relmse(rad.cpu().numpy(), tgt)
    e_out = dn.relmse(out.cpu().numpy(), tgt)
    print(f"held-out band: relMSE 4 spp {e_in:.4f} -> denoised {e_out:.5f} ({e_in / e_out:.1f}x)")
    assert e_out < 0.1 * e_in
    assert e_out < 0.01
