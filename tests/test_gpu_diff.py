"""Differentiable rendering on the inverse-rendering scene (the displaced
sphere textured all over, a rough copper ground, a rough glass ball): the
forward kernel against lrt_render (same samples, same radiance), the exact
BSDF derivatives against differences of the oracle's BSDF, the path-replay
adjoint against central finite differences with common random numbers, its
mean against differences of the expected image, the C++ autograd Function
against the adjoint it wraps, and inverse rendering lowering the error.

Bars: the forward image equal to lrt_render's to 1e-5 on >= 99.9% of
pixel-channels.  d(f cos)/d alpha and d/d IOR within 2e-3 relative of
Richardson-extrapolated differences of the oracle on >= 99% of the tuples.
Texel gradients of 32 seeded random texel parameters: the image is a
polynomial in each albedo with fixed samples (no roulette), so adjoint and FD
with common random numbers agree to 1e-2 relative.  Alpha and IOR: the adjoint
holds the sampling fixed (detached), so it is checked to 1e-2 against FD of
the held-sampling estimator it differentiates (lrt_diff_render_held), and its
mean over independent seeds lies within 3 sigma (adjoint and FD standard errors
combined) of central differences of the expected image.  Gradient sums are
order-independent fixed-point integers: two adjoint calls agree bit for bit,
and the inverse-rendering run is reproducible, its texture error falling by
at least 40%.
"""
import ctypes as C

import numpy as np
import pytest

import linrt
from linrt import diff

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    import torch
    torch.cuda.init()
    gen = linrt.inverse_sphere(seed=1, segments=128, rings=64, aspect=1.0)
    return gen, linrt.Scene(gen.desc_ptr)


def base_theta(gen, dp):
    d = gen.desc
    m0 = d.materials[dp.tex_material]
    tex = np.tile(np.array(list(m0.diffuse), np.float32), dp.tex_width * dp.tex_height)
    return np.concatenate([tex, [d.materials[dp.alpha_material[0]].alpha, d.materials[dp.alpha_material[1]].alpha,
                                 d.materials[dp.ior_material].int_ior]]).astype(np.float32)


def test_forward_matches_path_tracer(small):
    import torch
    gen, sc = small
    dp = diff.params()
    W = H = 96
    R = diff.Renderer(sc, dp, W, H, 4, max_depth=8, rr_depth=5, seed=3)
    img = R.render(torch.from_numpy(base_theta(gen, dp)).cuda()).cpu().numpy().astype(np.float64)
    s, q, st = sc.render(W, H, 4, 8, 5, seed=3)
    ref = s / 4
    close = np.isclose(img, ref, rtol=1e-5, atol=1e-7)
    print(f"diff kernel vs lrt_render: {close.mean():.5%} within 1e-5, max |d| {np.abs(img - ref).max():.2e}")
    assert close.mean() >= 0.999


def objective(R, theta, w):
    return float((R.render(theta).double() * w).sum())


def test_adjoint_vs_finite_differences(small):
    import torch
    gen, sc = small
    dp = diff.params()
    W = H = 128
    R = diff.Renderer(sc, dp, W, H, 16, max_depth=8, seed=5)
    rng = np.random.default_rng(2)
    theta = torch.from_numpy(base_theta(gen, dp) * np.float32(0.9)).cuda()
    theta[-1] = 1.5
    w = torch.from_numpy(rng.uniform(0.5, 1.5, size=(H, W, 3))).cuda()
    g = R.adjoint(theta, w.float()).cpu().numpy()
    n = R.n
    texels = np.sort(rng.choice(n - 3, 32, replace=False))
    print(f"texture parameters with a nonzero gradient: {np.count_nonzero(g[:n - 3])} of {n - 3}")
    eps = 1e-3
    rows = []
    for k in texels:
        tp, tm = theta.clone(), theta.clone()
        tp[k] += eps
        tm[k] -= eps
        fd = (objective(R, tp, w) - objective(R, tm, w)) / (2 * eps)
        rows.append((k, g[k], fd))
        print(f"  param {k}: adjoint {g[k]:+.5e}  fd {fd:+.5e}")
    tex = np.array([(a, f) for k, a, f in rows])
    assert np.count_nonzero(tex[:, 1]) >= 28
    assert np.all(np.abs(tex[:, 0] - tex[:, 1]) <= 1e-2 * np.abs(tex[:, 1]) + 1e-4 * np.abs(tex[:, 1]).max())
    # scalars (64^2): FD of the held-sampling estimator.  With the directions
    # held, the rough glass's transmission value drops to zero where the
    # refraction's half-vector crosses the horizon as the IOR moves, so a step
    # of 1e-3 now and then straddles such a jump; the IOR takes a step of 1e-4.
    Rs = diff.Renderer(sc, dp, 64, 64, 16, max_depth=8, seed=5)
    ws = w[::2, ::2].contiguous()
    gs = Rs.adjoint(theta, ws.float()).cpu().numpy()
    t0 = theta.cpu().numpy()
    wn = ws.cpu().numpy()
    for k, step in ((n - 3, 1e-3), (n - 2, 1e-3), (n - 1, 1e-4)):
        tp, tm = t0.copy(), t0.copy()
        tp[k] += step
        tm[k] -= step
        fd = ((Rs.render_held(tp, t0).astype(np.float64) * wn).sum() -
              (Rs.render_held(tm, t0).astype(np.float64) * wn).sum()) / (2 * step)
        print(f"  param {k}: adjoint {gs[k]:+.5e}  fd of the held estimator {fd:+.5e} (step {step:g})")
        assert abs(gs[k] - fd) <= 1e-2 * abs(fd) + 1e-3


def test_exact_bsdf_derivatives_vs_oracle_differences():
    """lrt_bsdf_param_grad, the derivative the adjoint takes at a vertex
    (forward-mode dual numbers through GGX D, Smith G1 and the dielectric
    Fresnel term), against Richardson-extrapolated central differences of the
    oracle's BSDF evaluation (h = 2e-3 and 1e-3, truncation O(h^4)), on seeded
    tuples of the rough copper (alpha 0.1) and the rough glass (alpha 0.2 and
    0.05; both faces, reflection and transmission), in both transport modes."""
    import oracle_lib as OL
    from common import bsdf_tuples
    m = linrt.material
    cu_eta, cu_k = (0.200438, 0.924033, 1.102212), (3.912948, 2.452847, 2.142188)
    mats = [m(linrt.BSDF_CONDUCTOR, alpha=0.1, cond_eta=cu_eta, cond_k=cu_k),
            m(linrt.BSDF_ROUGH_DIELECTRIC, int_ior=1.5, ext_ior=1.0, alpha=0.2),
            m(linrt.BSDF_ROUGH_DIELECTRIC, int_ior=1.5, ext_ior=1.0, alpha=0.05)]
    names = ["rough copper a0.1", "rough glass a0.2", "rough glass a0.05"]
    n = 1 << 16
    mat, wi, wo, front, _, _ = bsdf_tuples(n, seed=21, n_mats=len(mats))
    for mode in (linrt.MODE_RADIANCE, linrt.MODE_IMPORTANCE):
        for which, field in ((0, "alpha"), (1, "int_ior")):
            g = linrt.bsdf_param_grad(mats, None, mat, wi, wo, front, which, mode=mode).astype(np.float64)

            def f_at(dx):
                ms = [linrt.Material.from_buffer_copy(x) for x in mats]
                for x in ms:
                    setattr(x, field, getattr(x, field) + dx)
                f, _ = OL.oracle.bsdf_eval_pdf(ms, None, mat, wi, wo, front, measure=linrt.MEASURE_SOLID_ANGLE,
                                               mode=mode)
                return f.astype(np.float64)

            h = 2e-3
            d1 = (f_at(h) - f_at(-h)) / (2 * h)
            d2 = (f_at(h / 2) - f_at(-h / 2)) / h
            fd = (4 * d2 - d1) / 3
            f0 = np.abs(f_at(0.0))
            ok = np.all(np.abs(g - fd) <= 2e-3 * np.maximum(np.abs(g), np.abs(fd)) + 1e-3 * f0 + 1e-9, axis=1)
            for k in range(len(mats)):
                sel = (mat == k) & np.any(f0 > 0, axis=1)
                live = sel & np.any(np.abs(fd) > 1e-3 * f0, axis=1)
                frac = ok[sel].mean()
                print(f"  mode {mode} d/d {field:7s} {names[k]:18s}: {sel.sum()} tuples with f > 0, "
                      f"{live.sum()} with a derivative, {frac:.5f} within 2e-3")
                if which == 1 and k == 0:
                    assert np.all(g[mat == 0] == 0)  # the conductor has no IOR parameter
                else:
                    assert live.sum() > 1000 and frac >= 0.99


def test_scalar_fd_through_cpu_oracle(small):
    """The oracle's CPU path tracer (same Philox samples) gives the same
    finite differences of the image for alpha and IOR as the GPU kernel."""
    import torch
    import oracle_lib as OL
    gen, sc = small
    dp = diff.params()
    W = H = 24
    spp, depth = 4, 8
    R = diff.Renderer(sc, dp, W, H, spp, max_depth=depth, rr_depth=1 << 20, seed=9)
    theta = torch.from_numpy(base_theta(gen, dp)).cuda()
    d = gen.desc
    eps = 1e-3
    for slot, (mat, field) in enumerate([(dp.alpha_material[0], "alpha"), (dp.alpha_material[1], "alpha"),
                                         (dp.ior_material, "int_ior")]):
        fds = []
        gfd = []
        for sgn in (+1, -1):
            mats = (linrt.Material * d.n_materials)(*[d.materials[i] for i in range(d.n_materials)])
            setattr(mats[mat], field, getattr(mats[mat], field) + sgn * eps)
            desc = linrt.SceneDesc(d.n_objects, d.objects, d.n_materials, mats, d.n_lights, d.lights, d.n_tables,
                                   d.tables, d.camera)
            o = OL.oracle.scene(C.pointer(desc)).render(W, H, spp, depth, 1 << 20, seed=9, mode=0, threads=8)
            fds.append(o["sum"] / spp)
            t = theta.clone()
            t[-3 + slot] += sgn * eps
            gfd.append(R.render(t).cpu().numpy().astype(np.float64))
        d_cpu = (fds[0] - fds[1]) / (2 * eps)
        d_gpu = (gfd[0] - gfd[1]) / (2 * eps)
        err = np.abs(d_cpu - d_gpu).sum() / max(np.abs(d_cpu).sum(), 1e-12)
        print(f"  {field} of material {mat}: |FD gpu - FD cpu| / |FD cpu| = {err:.2e}, |FD| {np.abs(d_cpu).sum():.3e}")
        assert err < 0.05


def test_adjoint_is_unbiased():
    """The adjoint's mean over 24 independent seeds against central
    differences of the expected image, estimated by the wavefront path tracer
    (another implementation) at 16 x 1024 spp with common random numbers; the
    differences' standard error from the spread of the 16 batches."""
    import torch
    gen = linrt.inverse_sphere(seed=1, segments=128, rings=64, aspect=1.0)
    d = gen.desc
    W = H = 48

    def scene_with(a1, a2, ior):
        mats = (linrt.Material * d.n_materials)(*[d.materials[i] for i in range(d.n_materials)])
        mats[1].alpha, mats[2].alpha, mats[1].int_ior = a1, a2, ior
        desc = linrt.SceneDesc(d.n_objects, d.objects, d.n_materials, mats, d.n_lights, d.lights, d.n_tables,
                               d.tables, d.camera)
        return linrt.Scene(C.pointer(desc)), (mats, desc)

    base = [0.2, 0.1, 1.5]
    eps = 0.01
    batches, bspp = 16, 1024
    fd, fd_se = [], []
    for k in range(3):
        v = np.zeros((2, batches))
        for i, sgn in enumerate((1, -1)):
            p = list(base)
            p[k] += sgn * eps
            sc, keep = scene_with(*p)
            for b in range(batches):
                s, _, _ = sc.render(W, H, bspp, 8, 1 << 20, seed=77, spp_offset=b * bspp, sumsq=False)
                v[i, b] = s.sum() / bspp
            sc.close()
        per = (v[0] - v[1]) / (2 * eps)
        fd.append(per.mean())
        fd_se.append(per.std(ddof=1) / np.sqrt(batches))
    sc, keep = scene_with(*base)
    dp = diff.params()
    th = torch.from_numpy(base_theta(gen, dp)).cuda()
    th[-3:] = torch.tensor(base)
    w = torch.ones((H, W, 3), device="cuda:0")
    gs = np.array([diff.Renderer(sc, dp, W, H, 64, max_depth=8, seed=s).adjoint(th, w).cpu().numpy()[-3:]
                   for s in range(1, 25)])
    m, se = gs.mean(0), gs.std(0, ddof=1) / np.sqrt(len(gs))
    for k, name in enumerate(["alpha 1", "alpha 2", "ior"]):
        sigma = np.hypot(se[k], fd_se[k])
        print(f"  {name}: adjoint {m[k]:+.3f} +- {se[k]:.3f}, FD of the expectation {fd[k]:+.3f} +- {fd_se[k]:.3f}: "
              f"{abs(m[k] - fd[k]) / sigma:.2f} sigma")
        assert abs(m[k] - fd[k]) <= 3 * sigma


def test_adjoint_is_bit_reproducible(small):
    """Gradient sums are fixed-point integers, so two adjoint calls (both
    kernel families) and two short optimisation runs agree bit for bit."""
    import torch
    gen, sc = small
    dp = diff.params()
    W = H = 64
    theta = torch.from_numpy(base_theta(gen, dp) * np.float32(0.7)).cuda()
    w = torch.from_numpy(np.random.default_rng(6).uniform(-1, 1, size=(H, W, 3)).astype(np.float32)).cuda()
    for per_sample in (False, True):
        R = diff.Renderer(sc, dp, W, H, 8, max_depth=8, seed=4, per_sample=per_sample)
        a, b = R.adjoint(theta, w), R.adjoint(theta, w)
        print(f"per_sample={per_sample}: {int((a != 0).sum())} nonzero entries, equal: {bool(torch.equal(a, b))}")
        assert torch.equal(a, b)
    r1 = diff.optimize(iters=3, res=64, spp=4, lr=0.05, log=print)
    r2 = diff.optimize(iters=3, res=64, spp=4, lr=0.05, log=print)
    assert r1["loss"] == r2["loss"] and np.array_equal(r1["theta"], r2["theta"])


def test_cpp_autograd_function(small):
    """diff.render, the C++ torch::autograd::Function of linrt_torch: its
    forward is lrt_diff_render's image and its backward the adjoint of the
    upstream gradient replayed on other samples, bit for bit."""
    import torch
    gen, sc = small
    dp = diff.params()
    W = H = 64
    R = diff.Renderer(sc, dp, W, H, 8, max_depth=8, seed=3)
    theta = torch.from_numpy(base_theta(gen, dp) * np.float32(0.8)).cuda().requires_grad_()
    img = diff.render(theta, R)
    print(f"grad_fn: {img.grad_fn.name()}")
    assert "DiffRender" in img.grad_fn.name()
    assert torch.equal(img.detach(), R.render(theta))
    w = torch.from_numpy(np.random.default_rng(8).uniform(0, 1, size=(H, W, 3)).astype(np.float32)).cuda()
    (img * w).sum().backward()
    ref = R.adjoint(theta.detach(), w, decorrelate=True)
    print(f"|grad| {theta.grad.norm():.4e}, equal to the adjoint: {bool(torch.equal(theta.grad, ref))}")
    assert torch.equal(theta.grad, ref)
    with pytest.raises(RuntimeError, match="entries"):
        diff.render(theta[:-1], R)


def test_inverse_rendering_recovers_parameters():
    """Adam through the C++ autograd Function, seed pinned (the trajectory is
    reproducible): over 150 iterations at 128^2 x 16 spp the texture error
    over the texels the image sees falls by at least 40% (measured: 53%) and
    every scalar's error shrinks to a quarter of its initial one or less."""
    r = diff.optimize(iters=150, res=128, spp=16, lr=0.05, seed=7, log=print)
    th, tr, t0 = r["theta"], r["truth"], r["theta0"]
    print(f"scalars {t0[-3:]} -> {th[-3:]} (true {tr[-3:]}); texture RMSE over {r['texels_seen']} seen texels "
          f"{r['texture_rmse_seen0']:.4f} -> {r['texture_rmse_seen']:.4f}")
    assert r["texels_seen"] > 3500
    assert r["texture_rmse_seen"] < 0.60 * r["texture_rmse_seen0"]
    assert np.all(np.abs(th[-3:] - tr[-3:]) <= 0.25 * np.abs(t0[-3:] - tr[-3:]))


def test_wavefront_matches_per_sample_kernels(small):
    """The wavefront forward and adjoint against the per-sample kernels
    (LRT_DIFF_PER_SAMPLE), an independent implementation of the same
    estimator on the same samples: images equal to 1e-5 (summation order),
    gradients equal up to FP32 rounding (the per-sample kernels add a lane's
    samples in float before the fixed-point sum)."""
    import torch
    gen, sc = small
    dp = diff.params()
    W = H = 64
    theta = torch.from_numpy(base_theta(gen, dp) * np.float32(0.9)).cuda()
    theta[-1] = 1.45
    Rw = diff.Renderer(sc, dp, W, H, 8, max_depth=8, seed=9)
    Rp = diff.Renderer(sc, dp, W, H, 8, max_depth=8, seed=9, per_sample=True)
    iw, ip = Rw.render(theta).cpu().numpy(), Rp.render(theta).cpu().numpy()
    close = np.isclose(iw, ip, rtol=1e-5, atol=1e-7).mean()
    print(f"forward: {close:.5%} of pixel-channels within 1e-5")
    assert close >= 0.999
    rng = np.random.default_rng(4)
    w = torch.from_numpy(rng.uniform(0.5, 1.5, size=(H, W, 3)).astype(np.float32)).cuda()
    gw, gp = Rw.adjoint(theta, w).cpu().numpy(), Rp.adjoint(theta, w).cpu().numpy()
    n = Rw.n
    rel = np.linalg.norm(gw[:n - 3] - gp[:n - 3]) / np.linalg.norm(gp[:n - 3])
    print(f"adjoint: texture gradients relative L2 difference {rel:.2e}; scalars {gw[n - 3:]} vs {gp[n - 3:]}")
    assert rel < 1e-3
    assert np.all(np.abs(gw[n - 3:] - gp[n - 3:]) <= 1e-3 * np.abs(gp[n - 3:]) + 1e-5 * np.abs(gp[n - 3:]).max())
