"""Training, evaluation and the time-to-quality measurement of the learned
denoiser (lrt_denoiser_*), all on one GPU.

Data: frames of seeded displaced-sphere scenes (random materials, mesh and
one of eight cooked skies) rendered by the library's own path tracer: a few
noisy low-spp frames with their first-hit features per scene and one
high-spp frame as the target.  Frames stay in device memory (torch tensors
are used only as allocations); every network step runs the library's BF16
matrix-core kernels through the C ABI.

  python -m linrt.denoise train [--scenes N] [--steps K] [--out path] [--evaluate]
  python -m linrt.denoise oracle --spp 65536 [--out path]   (the FP64 65,536-spp oracle of the whole frame)
  python -m linrt.denoise evaluate                            (time to relMSE <= 1e-2 over the whole frame)
"""
from __future__ import annotations

import argparse
import math
import sys
import time
from pathlib import Path

import numpy as np

import linrt

DATA = Path(__file__).resolve().parent.parent / "data"
WEIGHTS = DATA / "denoiser_weights.npz"
MAX_DEPTH, RR_DEPTH = 16, 5
EPS = 1e-2  # relMSE = mean((x - ref)^2 / (ref^2 + EPS))


def _torch():
    """torch (device allocations only), with its GPU runtime initialised: it
    must start before the library's first HIP call in the process."""
    import torch
    if not torch.cuda.is_initialized():
        torch.cuda.init()
    return torch


def radiance(scene, W, H, spp, seed, offset, torch_dev="cuda:0", rows=None):
    """Mean radiance of spp samples per pixel, float32 [H, W, 3] on the device
    (rows outside `rows` stay 0)."""
    torch = _torch()
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device=torch_dev)
    rb, re = rows if rows is not None else (0, 0)
    prm = linrt.RenderParams(W, H, spp, offset, MAX_DEPTH, RR_DEPTH, seed, rb, re, 0)
    st = linrt.RenderStats()
    linrt._check(linrt.LIB.lrt_render(scene._h, linrt.C.byref(prm), linrt.C.c_void_p(acc.data_ptr()), None, 1,
                                      linrt.C.byref(st)))
    scene.sync()
    return (acc / spp).float()


def radiance_features(scene, W, H, spp, seed, offset, torch_dev="cuda:0"):
    """radiance() and features() of the same samples in one render pass
    (lrt_render_features): the features come from the render's first bounce."""
    torch = _torch()
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device=torch_dev)
    aov = torch.zeros((H, W, 7), dtype=torch.float32, device=torch_dev)
    scene.render_features_device(W, H, spp, MAX_DEPTH, RR_DEPTH, acc.data_ptr(), aov.data_ptr(), seed=seed,
                                 spp_offset=offset)
    scene.sync()
    return (acc / spp).float(), aov


def features(scene, W, H, spp, seed, offset, torch_dev="cuda:0", rows=None):
    torch = _torch()
    aov = torch.zeros((H, W, 7), dtype=torch.float32, device=torch_dev)
    scene.render_aov_device(W, H, spp, aov.data_ptr(), seed=seed, spp_offset=offset, rows=rows)
    return aov


def relmse(x, ref, eps=EPS):
    """Relative MSE (per pixel-channel, against ref^2 + eps), float64, over
    the pixels whose reference is finite; a non-finite pixel of x counts as
    black, as the reference's image writer stores it (image_write.cpp:11-13)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    ok = np.isfinite(ref).all(axis=-1)
    x = np.where(np.isfinite(x), x, 0.0)
    return float(np.mean(((x - ref) ** 2 / (ref * ref + eps))[ok]))


def make_dataset(n_scenes=160, W=1920, H=1080, band=128, spp_in=(2, 4, 8), spp_tgt=512, seed0=100, log=print):
    """Per scene (seeded materials, mesh and sky): a band of `band` rows of the
    W x H frame at a seeded height, rendered once per entry of spp_in (noisy
    input + features) and once at spp_tgt (target), so the network sees the
    pixel scale of full-resolution frames.  Device tensors rad [N*len(spp_in),
    band, W, 3], aux [.., 7], tgt [N, band, W, 3] and each input's target index."""
    torch = _torch()
    skies = linrt.load_sky_set()
    rad, aux, tgt, idx = [], [], [], []
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed0)
    for s in range(n_scenes):
        seed = seed0 + s
        gen = linrt.displaced_sphere_variant(seed=seed, aspect=W / H, sky=skies[s % len(skies)])
        sc = linrt.Scene(gen.desc_ptr)
        r0 = int(rng.integers(0, (H - band) // 4 + 1)) * 4
        rows = (r0, r0 + band)
        off = 0
        for k, spp in enumerate(spp_in):
            rad.append(radiance(sc, W, H, spp, seed, off, rows=rows)[r0:r0 + band].clone())
            aux.append(features(sc, W, H, spp, seed, off, rows=rows)[r0:r0 + band].clone())
            idx.append(s)
            off += spp
        tgt.append(radiance(sc, W, H, spp_tgt, seed, 1 << 20, rows=rows)[r0:r0 + band].clone())
        sc.close()
        gen.close()
        if (s + 1) % 16 == 0:
            log(f"  dataset: {s + 1}/{n_scenes} scenes, {time.perf_counter() - t0:.1f}s")
    return torch.stack(rad), torch.stack(aux), torch.stack(tgt), torch.tensor(idx, device="cuda:0")


def random_sky(rng, turbidity=None):
    """A sky of a seeded sun (elevation 10..60 degrees, any azimuth) cooked by
    the library with the 1080p workload's turbidity 3 (or a seeded one in
    2..6 when turbidity is "random"), ground albedo 0.3 and radiance scale."""
    el = math.radians(rng.uniform(10.0, 60.0))
    az = rng.uniform(0.0, 2.0 * math.pi)
    d = [math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)]
    t = float(rng.uniform(2.0, 6.0)) if turbidity == "random" else (3.0 if turbidity is None else float(turbidity))
    return linrt.cook_sky(d, turbidity=t, ground_albedo=0.3, scale=0.05)


def crop_camera(desc, rng, crop=256, W=1920, H=1080):
    """A copy of the generated scene description whose camera looks at a
    seeded crop x crop window of the W x H frame: aimed at the window's centre
    ray, square, with the frame's angular pixel size (vertical field of view
    scaled by crop / H), so the crop sees the scene at full-frame resolution."""
    d = linrt.SceneDesc.from_buffer_copy(desc.contents)
    cam = d.camera
    o = np.array(cam.origin[:], np.float64)
    la = np.array(cam.look_at[:], np.float64)
    up = np.array(cam.up[:], np.float64)
    h = math.tan(math.radians(cam.fov_deg) / 2)
    vh, vw = 2 * h, 2 * h * (W / H)
    w = (o - la) / np.linalg.norm(o - la)
    u = np.cross(up, w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    llc = o - vw / 2 * u - vh / 2 * v - w
    cx = rng.uniform(crop / 2, W - crop / 2)
    cy = rng.uniform(crop / 2, H - crop / 2)
    ray = llc + cx / (W - 1) * vw * u + cy / (H - 1) * vh * v - o
    ray /= np.linalg.norm(ray)
    for k in range(3):
        cam.look_at[k] = float(o[k] + ray[k])
    cam.fov_deg = float(math.degrees(2 * math.atan(h * crop / H)))
    cam.aspect = 1.0
    return d


def make_crop_dataset(n_scenes=2000, crop=256, spp_in=(1, 2, 4), spp_tgt=1024, segments=1024, rings=512, seed0=10000,
                      builders=4, log=print):
    """The training set: n_scenes seeded scenes of the 1080p workload's
    generator (the displaced sphere at segments x rings, materials and mesh
    from the seed, a sky of a seeded sun angle), each seen through a crop x
    crop window of the 1080p frame at a seeded position; noisy frames (with
    their first-hit features) at each spp of spp_in and a target at spp_tgt.
    Scenes are generated and their kd trees built on `builders` host threads
    ahead of the renders.  Device tensors rad [N * len(spp_in), crop, crop,
    3], aux [.., 7], tgt [N, crop, crop, 3] and each input's target index."""
    import collections
    from concurrent.futures import ThreadPoolExecutor
    torch = _torch()
    n_in = len(spp_in)
    rad = torch.empty((n_scenes * n_in, crop, crop, 3), dtype=torch.float32, device="cuda:0")
    aux = torch.empty((n_scenes * n_in, crop, crop, 7), dtype=torch.float32, device="cuda:0")
    tgt = torch.empty((n_scenes, crop, crop, 3), dtype=torch.float32, device="cuda:0")
    idx = torch.arange(n_scenes, device="cuda:0").repeat_interleave(n_in)

    def build(s):
        seed = seed0 + s
        rng = np.random.default_rng(seed)
        gen = linrt.displaced_sphere_variant(seed=seed, segments=segments, rings=rings, aspect=1920 / 1080,
                                             sky=random_sky(rng))
        desc = crop_camera(gen.desc_ptr, rng, crop)
        return seed, gen, desc, linrt.Scene(linrt.C.byref(desc))

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=builders) as ex:
        pending = collections.deque(ex.submit(build, s) for s in range(min(n_scenes, builders + 2)))
        nxt = len(pending)
        for s in range(n_scenes):
            seed, gen, desc, sc = pending.popleft().result()
            if nxt < n_scenes:
                pending.append(ex.submit(build, nxt))
                nxt += 1
            off = 0
            for k, spp in enumerate(spp_in):
                rad[s * n_in + k] = radiance(sc, crop, crop, spp, seed, off)
                aux[s * n_in + k] = features(sc, crop, crop, spp, seed, off)
                off += spp
            tgt[s] = radiance(sc, crop, crop, spp_tgt, seed, 1 << 20)
            sc.close()
            gen.close()
            if (s + 1) % 50 == 0:
                log(f"  dataset: {s + 1}/{n_scenes} scenes, {time.perf_counter() - t0:.1f}s")
    return rad, aux, tgt, idx


def train(den, data, steps=1500, batch=8, crop=128, lr=2e-3, seed=0, log=print):
    """Adam on random crops (with flips) of the dataset; cosine learning rate."""
    torch = _torch()
    rad, aux, tgt, idx = data
    N, H, W, _ = rad.shape
    g = torch.Generator(device="cpu").manual_seed(seed)
    t0 = time.perf_counter()
    hist = []
    skipped = 0
    for it in range(steps):
        f = torch.randint(0, N, (batch,), generator=g)
        ys = torch.randint(0, H - crop + 1, (batch,), generator=g)
        xs = torch.randint(0, W - crop + 1, (batch,), generator=g)
        r = torch.stack([rad[f[i], ys[i]:ys[i] + crop, xs[i]:xs[i] + crop] for i in range(batch)])
        a = torch.stack([aux[f[i], ys[i]:ys[i] + crop, xs[i]:xs[i] + crop] for i in range(batch)])
        t = torch.stack([tgt[idx[f[i]], ys[i]:ys[i] + crop, xs[i]:xs[i] + crop] for i in range(batch)])
        if it % 2:
            r, a, t = r.flip(2), a.flip(2), t.flip(2)
            a[..., 3] = -a[..., 3]  # mirrored normals
        r, a, t = r.contiguous(), a.contiguous(), t.contiguous()
        loss = den.loss_grad_device(r.data_ptr(), a.data_ptr(), t.data_ptr(), batch, crop, crop)
        # a step whose loss is non-finite or far above the recent mean (a
        # sample of extreme fireflies) is not applied: one such update can
        # push the ReLUs dead and end the training
        recent = np.mean(hist[-200:]) if len(hist) >= 200 else None
        if math.isfinite(loss) and (recent is None or loss <= 4.0 * recent):
            den.adam_step(lr * 0.5 * (1 + math.cos(math.pi * it / steps)))
        else:
            skipped += 1
        hist.append(loss if math.isfinite(loss) else (recent or 0.0))
        if (it + 1) % 250 == 0:
            log(f"  step {it + 1}/{steps}: loss {np.mean(hist[-250:]):.4f} ({time.perf_counter() - t0:.1f}s, "
                f"{skipped} steps skipped)")
    torch.cuda.synchronize()
    return hist


def save(den, path=WEIGHTS):
    np.savez_compressed(path, params=den.params.astype(np.float16), layers=np.array(linrt_layers()))


def load(den, path=WEIGHTS):
    p = np.load(path)["params"].astype(np.float32)
    den.params = p


def linrt_layers():
    return [10, 32, 32, 32, 32, 64, 64, 64, 64, 128, 128, 128, 192, 64, 64, 64, 96, 32, 32, 3]


BF16_PEAK = 2.5e15  # dense bf16 MFMA flop/s of one MI355X
HBM_PEAK = 8.0e12   # HBM bytes/s


def layer_work(H, W):
    """Algorithmic work of the forward's 10 convolutions at an H x W frame:
    (name, flops, bytes) per layer.  flops = 2 * 9 * cin * cout per output
    pixel (real channel counts); bytes = what the layer must read and write
    once: its input sources (the pooled and upsampled ones at their own
    resolution, the features as 10 floats, the first layer's radiance and
    the last layer's radiance in and image out as 3 floats) and its bf16
    output."""
    P0, P1, P2 = H * W, (H // 2) * (W // 2), (H // 4) * (W // 4)
    bf = 2
    L = [  # name, cin, cout, output pixels, bytes read, bytes written
        ("e0a features->32", 10, 32, P0, P0 * 40, P0 * 32 * bf),
        ("e0b 32->32", 32, 32, P0, P0 * 32 * bf, P0 * 32 * bf),
        ("e1a pool->64", 32, 64, P1, P0 * 32 * bf, P1 * 64 * bf),
        ("e1b 64->64", 64, 64, P1, P1 * 64 * bf, P1 * 64 * bf),
        ("e2a pool->128", 64, 128, P2, P1 * 64 * bf, P2 * 128 * bf),
        ("e2b 128->128", 128, 128, P2, P2 * 128 * bf, P2 * 128 * bf),
        ("d1a up+skip 192->64", 192, 64, P1, P2 * 128 * bf + P1 * 64 * bf, P1 * 64 * bf),
        ("d1b 64->64", 64, 64, P1, P1 * 64 * bf, P1 * 64 * bf),
        ("d0a up+skip 96->32", 96, 32, P0, P1 * 64 * bf + P0 * 32 * bf, P0 * 32 * bf),
        ("out 32->3 image", 32, 3, P0, P0 * 32 * bf + P0 * 12, P0 * 12),
    ]
    return [(n, 2 * 9 * ci * co * p, rd + wr) for n, ci, co, p, rd, wr in L]


def layer_roofline(H, W, layer_ms):
    """Per layer: achieved flop/s, its bound min(bf16 MFMA peak, HBM peak x
    flops/bytes) and the fraction reached, from measured milliseconds."""
    out = []
    for (name, fl, by), ms in zip(layer_work(H, W), layer_ms):
        bound = min(BF16_PEAK, HBM_PEAK * fl / by)
        ach = fl / (ms * 1e-3) if ms > 0 else 0.0
        out.append({"layer": name, "gflop": round(fl / 1e9, 3), "mbytes": round(by / 1e6, 2),
                    "intensity": round(fl / by, 1), "ms": round(ms, 4), "tflops": round(ach / 1e12, 1),
                    "bound_tflops": round(bound / 1e12, 1), "bound": "mfma" if bound >= BF16_PEAK else "hbm",
                    "frac_of_bound": round(ach / bound, 4)})
    return out


def c1_scene():
    gen = linrt.displaced_sphere(seed=1)
    return gen, linrt.Scene(gen.desc_ptr)


ORACLE = DATA / "oracle_1080p_65536spp.npz"
ORACLE_SEED = (1 << 32) + 7  # a Philox key no timed render uses: the oracle's errors are independent of theirs


def make_oracle(spp=65536, out=ORACLE, chunk=512, accum=None, log=print):
    """The error judge of the 1080p frame: spp samples per pixel of the whole
    frame added into float64 sums and sums of squares on the device (the
    reference accumulates in float32, src/main.cpp:120-136).  Saves the
    float32 mean and the oracle's own relMSE, mean(var(mean) / (mean^2 +
    eps)) from the sums of squares; `accum`: also the float64 accumulators
    (lrt_accum_write, resumable)."""
    torch = _torch()
    gen, sc = c1_scene()
    W, H = 1920, 1080
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda:0")
    acc2 = torch.zeros_like(acc)
    t0 = time.perf_counter()
    last = t0
    for off in range(0, spp, chunk):
        n = min(chunk, spp - off)
        prm = linrt.RenderParams(W, H, n, off, MAX_DEPTH, RR_DEPTH, ORACLE_SEED, 0, 0, 0)
        st = linrt.RenderStats()
        linrt._check(linrt.LIB.lrt_render(sc._h, linrt.C.byref(prm), linrt.C.c_void_p(acc.data_ptr()),
                                          linrt.C.c_void_p(acc2.data_ptr()), 1, linrt.C.byref(st)))
        sc.sync()
        if time.perf_counter() - last > 30 or off + n == spp:
            last = time.perf_counter()
            log(f"  oracle: {off + n}/{spp} spp, {last - t0:.1f}s")
    s, q = acc.cpu().numpy(), acc2.cpu().numpy()
    if accum is not None:
        linrt.write_accum(accum, s, q, spp_done=spp, next_sample=spp, seed=ORACLE_SEED, max_depth=MAX_DEPTH,
                          rr_depth=RR_DEPTH)
    mean = s / spp
    var_mean = np.maximum(q / spp - mean * mean, 0.0) / (spp - 1)
    fin = np.isfinite(mean)
    self_relmse = float(np.mean((var_mean / (mean * mean + EPS))[fin]))
    np.savez_compressed(out, mean=mean.astype(np.float32), spp=spp, seed=ORACLE_SEED, max_depth=MAX_DEPTH,
                        rr_depth=RR_DEPTH, eps=EPS, self_relmse=self_relmse, nan_pixels=int((~fin).any(-1).sum()))
    log(f"oracle: {spp} spp in {time.perf_counter() - t0:.1f}s, own relMSE {self_relmse:.3e}, "
        f"non-finite pixels {int((~fin).any(-1).sum())}")
    sc.close()
    gen.close()
    return mean


def time_to_quality(den=None, target=0.01, log=print, spp_list=(1, 2, 4), weights=WEIGHTS, oracle=ORACLE):
    """Wall time for the plain path tracer and for denoised frames to reach
    relMSE <= target over the whole 1080p frame, judged by the FP64-accumulated
    65,536-spp oracle.  Each timed render includes its device accumulation
    (and for the denoised frames the features and the network)."""
    torch = _torch()
    d = np.load(oracle)
    ref = torch.from_numpy(d["mean"]).to("cuda:0")
    if den is None:
        den = linrt.Denoiser()
        load(den, weights)
    gen, sc = c1_scene()
    W, H = 1920, 1080
    out = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    fin = torch.isfinite(ref)

    def err(x):
        x = torch.nan_to_num(x, nan=0.0, posinf=0.0, neginf=0.0)
        e = (x - ref) ** 2 / (ref * ref + EPS)
        return float(e[fin].mean())

    r, a = radiance_features(sc, W, H, 1, 1, 0)
    den.forward_device(r.data_ptr(), a.data_ptr(), 1, H, W, out.data_ptr())
    torch.cuda.synchronize()
    res = {"oracle": {"file": oracle.name, "spp": int(d["spp"]), "accumulation": "float64",
                      "self_relmse": float(d["self_relmse"]), "pixels": W * H},
           "eps": EPS, "target": target, "denoised": [], "plain": []}
    best = None
    for spp in spp_list:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r, a = radiance_features(sc, W, H, spp, 1, 0)
        den.forward_device(r.data_ptr(), a.data_ptr(), 1, H, W, out.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        e, e0 = err(out), err(r)
        res["denoised"].append({"spp": spp, "seconds": dt, "relmse": e, "relmse_input": e0})
        log(f"  denoised {spp} spp: {dt * 1e3:.1f} ms, relMSE {e:.5f} (input {e0:.4f})")
        if e <= target and best is None:
            best = (spp, dt, e)
    res["denoised_best"] = None if best is None else {"spp": best[0], "seconds": best[1], "relmse": best[2]}
    # the network alone at 1080p: per-layer times (HIP events around each
    # convolution) against each layer's roofline
    den.timing(True)
    n_fw = 20
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_fw):
        den.forward_device(r.data_ptr(), a.data_ptr(), 1, H, W, out.data_ptr())
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / n_fw
    ms, calls = den.timing(False)
    ms = [m / max(calls, 1) for m in ms]
    layers = layer_roofline(H, W, ms)
    fl = sum(f for _, f, _ in layer_work(H, W))
    res["network"] = {"forward_ms": wall * 1e3, "conv_ms": sum(ms), "tflop": fl / 1e12,
                      "tflops": fl / (sum(ms) * 1e-3) / 1e12, "layers": layers}
    log(f"  network: {wall * 1e3:.3f} ms per 1080p forward, {fl / (sum(ms) * 1e-3) / 1e12:.1f} TFLOP/s over the convs")
    # plain: relMSE ~ c / spp; measure at 64 spp, render the predicted spp for
    # the target and step up by 10% until it is reached
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = radiance(sc, W, H, 64, 1, 0)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    e64 = err(r)
    res["plain"].append({"spp": 64, "seconds": t64, "relmse": e64})
    need = max(2, int(math.ceil(64 * e64 / target / 2)) * 2)
    for _ in range(6):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = radiance(sc, W, H, need, 1, 0)
        torch.cuda.synchronize()
        tn = time.perf_counter() - t0
        en = err(r)
        res["plain"].append({"spp": need, "seconds": tn, "relmse": en})
        log(f"  plain {need} spp: {tn:.2f} s, relMSE {en:.5f}")
        if en <= target:
            break
        need = int(math.ceil(need * 1.1 / 2)) * 2
    res["plain_best"] = {"spp": need, "seconds": tn, "relmse": en}
    res["speedup"] = tn / best[1] if (best is not None and en <= target) else None
    four = [x for x in res["denoised"] if x["spp"] == 4]
    res["speedup_4spp"] = tn / four[0]["seconds"] if (four and four[0]["relmse"] <= target and en <= target) else None
    sc.close()
    gen.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "oracle", "evaluate"])
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--spp-target", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=16384)
    ap.add_argument("--out", type=str, default=None)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--weights", type=str, default=None)
    ap.add_argument("--dataset", choices=["crops", "bands"], default="crops",
                    help="crops: 256x256 windows of 2,000 seeded scenes with 1,024-spp targets; "
                         "bands: 128-row bands of 1080p frames")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--evaluate", action="store_true", help="after training, time to relMSE with the new weights")
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    _torch()
    if args.cmd == "oracle":
        make_oracle(args.spp, Path(args.out) if args.out else ORACLE, log=log)
        return
    if args.cmd == "evaluate":
        import json
        print(json.dumps(time_to_quality(log=log, weights=Path(args.weights) if args.weights else WEIGHTS)))
        return
    den = linrt.Denoiser(seed=args.seed)
    t0 = time.perf_counter()
    if args.dataset == "crops":
        data = make_crop_dataset(args.scenes, spp_tgt=args.spp_target, log=log)
    else:
        data = make_dataset(args.scenes, spp_tgt=args.spp_target, log=log)
    log(f"dataset {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    train(den, data, steps=args.steps, batch=args.batch, lr=args.lr, log=log)
    log(f"training {time.perf_counter() - t0:.1f}s")
    out = Path(args.out) if args.out else WEIGHTS
    save(den, out)
    if args.evaluate:
        import json
        del data
        r = time_to_quality(den=den, log=log)
        print(json.dumps({k: v for k, v in r.items() if k != "network"}))


if __name__ == "__main__":
    main()
