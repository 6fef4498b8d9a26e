"""Differentiable rendering on the GPU (lrt_diff_*): a C++
torch::autograd::Function (csrc/torch/diff_function.cpp, module linrt_torch)
whose forward is the library's path tracer and whose backward is its
path-replay adjoint, and the inverse-rendering workload built on it: recover
a 64x64 RGB albedo texture covering the displaced sphere, two GGX roughnesses
and one IOR (12,291 scalars) from a target image, by Adam.

  python -m linrt.diff [--iters 200] [--res 512] [--spp 32]
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.machinery
import importlib.util
import shutil
import sys
import time
from pathlib import Path

import numpy as np

import linrt

TEX = 64
DIFF_PER_SAMPLE = 4  # lrt_render_params.flags: LRT_DIFF_PER_SAMPLE
EXT_NAME = "linrt_torch"
EXT_DIR = Path(__file__).resolve().parent / "_build"
EXT_SRC = linrt.PKG_ROOT / "csrc" / "torch" / "diff_function.cpp"
INCLUDE = linrt.PKG_ROOT.parent / "include"
_ext = None


def build_torch_ext(verbose=False):
    """Compiles the C++ autograd Function module into EXT_DIR (always anew)."""
    from torch.utils import cpp_extension
    shutil.rmtree(EXT_DIR, ignore_errors=True)
    EXT_DIR.mkdir(parents=True)
    cpp_extension.load(name=EXT_NAME, sources=[str(EXT_SRC)], extra_include_paths=[str(INCLUDE)],
                       extra_cflags=["-O2"], build_directory=str(EXT_DIR), verbose=verbose)


def torch_ext():
    """The built linrt_torch module, bound to the loaded liblinrt_hip.so; raises
    when it is not built."""
    global _ext
    if _ext is None:
        import torch  # noqa: F401  (the module links against torch's libraries)
        path = EXT_DIR / f"{EXT_NAME}.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built: run linrt.diff.build_torch_ext()")
        loader = importlib.machinery.ExtensionFileLoader(EXT_NAME, str(path))
        spec = importlib.util.spec_from_file_location(EXT_NAME, str(path), loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        mod.bind(str(linrt.LIB_PATH))
        _ext = mod
    return _ext


class DiffParams(C.Structure):
    _fields_ = [("tex_material", C.c_int32), ("tex_width", C.c_uint32), ("tex_height", C.c_uint32),
                ("alpha_material", C.c_int32 * 2), ("ior_material", C.c_int32)]


def params(tex_material=0, tex=TEX, alpha_materials=(1, 2), ior_material=1):
    """Materials of the inverse-rendering scene (linrt.inverse_sphere): 0 the
    sphere's diffuse (textured), 1 the rough glass ball (alpha and IOR), 2 the
    rough copper ground (alpha)."""
    a = (C.c_int32 * 2)(*alpha_materials)
    return DiffParams(tex_material, tex, tex, a, ior_material)


def count(dp):
    n = C.c_uint32()
    linrt._check(linrt.LIB.lrt_diff_param_count(C.byref(dp), C.byref(n)))
    return n.value


class Renderer:
    """Scene + parameter layout + render settings; forward/adjoint on torch
    device tensors (theta float32 [n], image float32 [H, W, 3])."""

    def __init__(self, scene, dp, width, height, spp, max_depth=8, rr_depth=1 << 20, seed=1, spp_offset=0,
                 per_sample=False):
        """per_sample: the per-sample kernels (LRT_DIFF_PER_SAMPLE) instead of the
        wavefront, an independent implementation of the same estimator."""
        self.scene, self.dp = scene, dp
        self.prm = linrt.RenderParams(width, height, spp, spp_offset, max_depth, rr_depth, seed, 0, 0,
                                      DIFF_PER_SAMPLE if per_sample else 0)
        self.shape = (height, width, 3)
        self.n = count(dp)

    def render(self, theta):
        import torch
        theta = theta.detach().contiguous()
        img = torch.empty(self.shape, dtype=torch.float32, device=theta.device)
        linrt._check(linrt.LIB.lrt_diff_render(self.scene._h, C.byref(self.dp), C.c_void_p(theta.data_ptr()),
                                               C.byref(self.prm), C.c_void_p(img.data_ptr()), 1))
        return img

    def render_held(self, theta, theta_sampling):
        """The estimator the adjoint differentiates: BSDF values at theta,
        sampling at theta_sampling (host arrays in, host image out)."""
        t = np.ascontiguousarray(theta, np.float32)
        ts = np.ascontiguousarray(theta_sampling, np.float32)
        img = np.zeros(self.shape, np.float32)
        linrt._check(linrt.LIB.lrt_diff_render_held(self.scene._h, C.byref(self.dp), linrt._p(t), linrt._p(ts),
                                                    C.byref(self.prm), linrt._p(img)))
        return img

    def adjoint(self, theta, d_image, decorrelate=False):
        """Gradient of sum(d_image * image).  decorrelate: replay other samples
        than the forward image's (a loss gradient d_image computed from that
        image would otherwise correlate with the replayed paths and bias the
        result, e.g. for an L2 loss)."""
        import torch
        theta = theta.detach().contiguous()
        d_image = d_image.detach().contiguous().float()
        g = torch.empty(self.n, dtype=torch.float32, device=theta.device)
        prm = self.prm
        if decorrelate:
            prm = linrt.RenderParams(prm.width, prm.height, prm.spp, prm.spp_offset + (1 << 30), prm.max_depth,
                                     prm.rr_depth, prm.seed, 0, 0, prm.flags)
        linrt._check(linrt.LIB.lrt_diff_adjoint(self.scene._h, C.byref(self.dp), C.c_void_p(theta.data_ptr()),
                                                C.byref(prm), C.c_void_p(d_image.data_ptr()),
                                                C.c_void_p(g.data_ptr()), 1))
        return g


def render(theta, R):
    """R's image of theta, differentiable in theta: the C++ autograd Function
    (forward lrt_diff_render, backward lrt_diff_adjoint replaying other samples
    than the forward's)."""
    dp = [R.dp.tex_material, R.dp.tex_width, R.dp.tex_height, R.dp.alpha_material[0], R.dp.alpha_material[1],
          R.dp.ior_material]
    p = R.prm
    prm = [p.width, p.height, p.spp, p.spp_offset, p.max_depth, p.rr_depth, p.seed, p.row_begin, p.row_end, p.flags]
    return torch_ext().render(theta, R.scene._h.value, dp, prm)


def scene(res=512, segments=512, rings=256):
    """The inverse-rendering scene (displaced sphere of 262,144 tris by
    default, copper ground, glass ball, Hosek sky), square camera."""
    gen = linrt.inverse_sphere(seed=1, segments=segments, rings=rings, aspect=1.0)
    return gen, linrt.Scene(gen.desc_ptr)


def hidden_theta(n, seed):
    """Seeded parameters: a smooth random texture, alphas and IOR in range."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:TEX, 0:TEX] / TEX
    tex = np.stack([0.45 + 0.35 * np.sin(2 * np.pi * (rng.uniform(1, 4) * xx + rng.uniform(0, 1)))
                    * np.cos(2 * np.pi * (rng.uniform(1, 4) * yy + rng.uniform(0, 1))) for _ in range(3)], -1)
    theta = np.concatenate([tex.ravel(), [rng.uniform(0.1, 0.35), rng.uniform(0.05, 0.25), rng.uniform(1.3, 1.7)]])
    assert theta.size == n
    return theta.astype(np.float32)


def clamp_(theta):
    """Keep parameters in their physical ranges (in place)."""
    t = theta.data
    t[:-3].clamp_(0.0, 1.0)
    t[-3:-1].clamp_(0.02, 0.8)
    t[-1:].clamp_(1.05, 2.5)


def optimize(iters=200, res=512, spp=32, max_depth=8, lr=0.02, seed=7, log=print):
    """Inverse rendering: random initial theta, L2 loss against the target
    rendered with hidden parameters, Adam; returns the history."""
    import torch
    gen, sc = scene(res)
    dp = params()
    R = Renderer(sc, dp, res, res, spp, max_depth)
    truth = torch.from_numpy(hidden_theta(R.n, seed)).cuda()
    target = Renderer(sc, dp, res, res, 4 * spp, max_depth, seed=99).render(truth)
    rng = np.random.default_rng(seed + 1)
    theta0 = np.concatenate([rng.uniform(0, 1, R.n - 3), [rng.uniform(0.05, 0.5), rng.uniform(0.05, 0.5),
                                                           rng.uniform(1.2, 1.8)]]).astype(np.float32)
    theta = torch.from_numpy(theta0).cuda().requires_grad_()
    opt = torch.optim.Adam([theta], lr=lr)
    hist = []
    t0 = time.perf_counter()
    for it in range(iters):
        R.prm.seed = 1000 + it  # fresh samples each iteration
        opt.zero_grad()
        img = render(theta, R)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        bad = int((~torch.isfinite(theta.grad)).sum())
        if bad:  # a non-finite gradient would poison Adam's moments
            theta.grad.nan_to_num_(nan=0.0, posinf=0.0, neginf=0.0)
            log(f"  iter {it}: {bad} non-finite gradient entries zeroed")
        opt.step()
        clamp_(theta)
        hist.append(loss.item())
        if not np.isfinite(hist[-1]):
            log(f"  iter {it}: non-finite loss, stopping")
            break
        if it % 10 == 0 or it == iters - 1:
            t = theta.detach()
            log(f"  iter {it}: loss {hist[-1]:.6f}, alpha {t[-3]:.3f}/{t[-2]:.3f} (true {truth[-3]:.3f}/{truth[-2]:.3f}), "
                f"ior {t[-1]:.3f} (true {truth[-1]:.3f}), {time.perf_counter() - t0:.1f}s")
    torch.cuda.synchronize()
    t = theta.detach()
    # texels the image sees (nonzero gradient of the loss at the end)
    cover = R.adjoint(t, torch.ones(R.shape, device=t.device)).abs()[:-3].cpu().numpy()
    seen = cover > 0.05 * cover.max()
    th, tr, t0_ = t.cpu().numpy(), truth.cpu().numpy(), theta0
    return {"loss": hist, "seconds": time.perf_counter() - t0, "theta": th, "truth": tr, "theta0": t0_,
            "iters": iters, "res": res, "spp": spp, "texels_seen": int(seen.sum()),
            "texture_rmse_seen": float(np.sqrt(np.mean((th[:-3][seen] - tr[:-3][seen]) ** 2))),
            "texture_rmse_seen0": float(np.sqrt(np.mean((t0_[:-3][seen] - tr[:-3][seen]) ** 2)))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--spp", type=int, default=32)
    args = ap.parse_args(argv)
    import json
    import torch
    torch.cuda.init()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    r = optimize(args.iters, args.res, args.spp, log=log)
    print(json.dumps({"iters": r["iters"], "seconds": round(r["seconds"], 2), "loss_first": r["loss"][0],
                      "loss_last": r["loss"][-1], "scalars": r["theta"][-3:].tolist(), "truth": r["truth"][-3:].tolist(),
                      "scalars0": r["theta0"][-3:].tolist(), "texels_seen": r["texels_seen"],
                      "texture_rmse_seen": r["texture_rmse_seen"], "texture_rmse_seen0": r["texture_rmse_seen0"]}))


if __name__ == "__main__":
    main()
