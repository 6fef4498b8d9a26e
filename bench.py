#!/usr/bin/env python3
"""Headline benchmark: path-traced samples per second on one MI355X.

Workload (a "step" = one full frame): the seeded displaced sphere of
1,048,576 triangles (7 material regions: diffuse, rough dielectric, rough
conductor, smooth plastic, rough plastic, smooth coating, rough coating) on a
ground quad under a Hosek sky, 1920x1080 at 64 samples per pixel, max depth
16, Russian roulette from depth 5.  Scene data is resident in HBM before the
timed region; every step renders the same frame (same seed and sample
indices) so each step does identical work.

Prints ONE JSON line on stdout; progress goes to stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "ray-tracing_amd"))

METRIC = "Msamples/s (1080p, 1M tris); time to relMSE≤1e-2: denoised vs plain"
W, H, SPP, MAX_DEPTH, RR_DEPTH, SEED = 1920, 1080, 64, 16, 5, 1
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(gen, budget_s=15.0):
    """The oracle's CPU restatement (tests/oracle_lib.py) on this host's cores:
    a band of full-width rows of the same frame at 64 spp, sized to about
    budget_s seconds of render-loop time (scene build timed apart)."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib  # noqa: E402
    if oracle_lib.oracle is None:
        return None
    threads = len(os.sched_getaffinity(0)) or 1
    t0 = time.perf_counter()
    sc = oracle_lib.oracle.scene(gen.desc_ptr)
    build_s = time.perf_counter() - t0
    mid = H // 2
    t0 = time.perf_counter()
    sc.render(W, H, 4, MAX_DEPTH, RR_DEPTH, seed=SEED, threads=threads, rows=(mid, mid + 1))
    probe = time.perf_counter() - t0
    rows = max(1, min(H // 2, int(budget_s / max(probe * 16, 1e-3))))
    t0 = time.perf_counter()
    r = sc.render(W, H, SPP, MAX_DEPTH, RR_DEPTH, seed=SEED, threads=threads, rows=(mid, mid + rows))
    dt = time.perf_counter() - t0
    samples = W * rows * SPP
    value = samples / dt / 1e6
    model = "unknown"
    try:
        model = next(ln.split(":", 1)[1].strip() for ln in Path("/proc/cpuinfo").read_text().splitlines()
                     if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    out = {"value": value, "unit": "Msamples/s", "cores": threads, "kind": "port",
           "sample": f"rows {mid}..{mid + rows - 1} of the 1920x1080 frame at 64 spp ({samples} samples, "
                     f"{dt:.1f} s render loop; kd build {build_s:.1f} s not timed)",
           "cpu_model": model, "logical_cpus": os.cpu_count(), "threads": threads,
           "extension_rays": int(r["ext"]), "shadow_rays": int(r["shadow"])}
    # the restatement's speed relative to the reference's own code, measured
    # on one host with the same scenes and sample streams (oracle/calibrate_cpu.py)
    cal = REPO / "oracle" / "cpu_calibration.json"
    if cal.exists():
        c = json.loads(cal.read_text())
        out["calibration"] = {"ratio_port_over_reference": c["ratio"], "measured_on": c["host"],
                              "cases": [{k: v[k] for k in ("workload", "ratio")} for v in c["cases"]],
                              "reference_equivalent_value": round(value / c["ratio"], 4)}
    return out


def ray_bench(args, torch, linrt):
    """Incoherent-ray traversal throughput (closest hit) on the same 1M-tri
    mesh: area-uniform origins on the displaced sphere offset 1e-4 along the
    normal, cosine-distributed directions (secondary rays)."""
    sys.path.insert(0, str(REPO / "tests"))
    from common import surface_rays
    gen = linrt.displaced_sphere(seed=SEED)
    scene = linrt.Scene(gen.desc_ptr, device=0)
    pos, idx = gen.mesh(0)
    n = args.rays
    rays = surface_rays(pos, idx, n, seed=77)
    d_rays = torch.from_numpy(rays).to("cuda:0")
    d_op = torch.empty((n, 2), dtype=torch.int32, device="cuda:0")
    d_t = torch.empty(n, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    a, b, c = scene.trace_count(rays)
    bytes_per_ray = (32 + 16) + (8 * a.astype(np.int64) + 4 * b + 36 * c.astype(np.int64)).mean()
    for _ in range(args.warmup):
        scene.trace_closest_device(d_rays.data_ptr(), n, d_op.data_ptr(), d_t.data_ptr())
    scene.sync()
    scene.kernel_timing(linrt.KERNEL_CLOSEST, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scene.trace_closest_device(d_rays.data_ptr(), n, d_op.data_ptr(), d_t.data_ptr())
    scene.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms, nl, _ = scene.kernel_timing(linrt.KERNEL_CLOSEST)
    avg = ms / max(nl, 1)
    achieved = bytes_per_ray * n / (avg * 1e-3) / 1e9
    hit = (d_op[:, 0] >= 0).float().mean().item()
    out = {"metric": "M incoherent rays/s (closest hit, 1M tris)", "value": round(n * args.steps / el / 1e6, 2),
           "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(el / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded secondary rays)",
           "config": {"workload": f"{n} secondary rays on the displaced sphere (1,048,576 tris)"},
           "hit_fraction": round(hit, 4),
           "per_ray": {"nodes": float(a.mean()), "refs": float(b.mean()), "tris": float(c.mean()),
                       "bytes": float(bytes_per_ray)},
           "roofline": {"bound": "hbm", "kernel": "k_trace_closest", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "avg_launch_ms": round(avg, 4)}}
    print(json.dumps(out), flush=True)


def network_bench(args, torch, linrt):
    """The denoiser's U-Net alone: forwards of the committed network on one
    1920x1080 frame (radiance + first-hit features of the 1080p workload at
    1 spp, resident on the device), per-layer times from HIP events on the
    denoiser's stream against each layer's roofline."""
    from linrt import denoise as dn
    gen = linrt.displaced_sphere(seed=SEED)
    scene = linrt.Scene(gen.desc_ptr, device=0)
    rad = dn.radiance(scene, W, H, 1, SEED, 0)
    aux = dn.features(scene, W, H, 1, SEED, 0)
    out = torch.zeros_like(rad)
    den = linrt.Denoiser()
    dn.load(den)
    for _ in range(max(args.warmup, 1)):
        den.forward_device(rad.data_ptr(), aux.data_ptr(), 1, H, W, out.data_ptr())
    torch.cuda.synchronize()
    den.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        den.forward_device(rad.data_ptr(), aux.data_ptr(), 1, H, W, out.data_ptr())
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    ms, calls = den.timing(False)
    ms = [m / max(calls, 1) for m in ms]
    fl = sum(f for _, f, _ in dn.layer_work(H, W))
    ach = fl / (sum(ms) * 1e-3)
    out_line = {"metric": "denoiser U-Net forward at 1920x1080 (TFLOP/s over its convolutions)",
                "value": round(ach / 1e12, 1), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": round(el * 1e3, 4), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
                "data": "1-spp radiance and features of the 1080p workload",
                "config": {"workload": "U-Net 10->32/64/128->3, 3x3 convolutions, 1920x1080, batch 1"},
                "roofline": {"bound": "mfma", "kernel": "k_conv_fwd", "achieved": round(ach / 1e12, 1),
                             "peak": dn.BF16_PEAK / 1e12, "unit": "TFLOP/s", "frac": round(ach / dn.BF16_PEAK, 4),
                             "traffic": None, "flop_per_forward": fl, "conv_ms": round(sum(ms), 4)},
                "per_layer": dn.layer_roofline(H, W, ms)}
    print(json.dumps(out_line), flush=True)


def photon_bench(args, torch, linrt):
    """Photon-mapper workload: 4,194,304 photons traced (max depth 8, RR from
    5) from the 0.5 x 0.5 ceiling light through the 262,144-tri displaced
    sphere inside the box room, the kd photon map built over them, then the
    photon-integrator eye pass at 1024x1024 (k = 64, radius 0.05, eye depth
    limit 128).  One step = photon pass + map build + eye pass."""
    gen = linrt.photon_box(seed=SEED)
    scene = linrt.Scene(gen.desc_ptr, device=0)
    n_ph, res, spp = 4194304, 1024, args.spp if args.spp != SPP else 16
    scene.photon_map_load(np.zeros((0, 12), np.float32), np.zeros((0, 2), np.uint32), 1.0, 0.05, 64,
                          linrt.PHOTON_CAUSTIC)

    def step():
        st = scene.photon_build(n_ph, 8, 5, seed=SEED, radius=0.05, k=64)
        s, q, rs = scene.render_photon(res, res, spp, 128, seed=SEED, sumsq=False)
        return st, rs

    for _ in range(args.warmup):
        step()
    # untimed 1-spp counting pass: the gathers' algorithmic work per query
    if args.warmup == 0:
        step()
    _, _, cnt = scene.render_photon(res, res, 1, 128, seed=SEED, sumsq=False, count=True)
    torch.cuda.synchronize()
    scene.kernel_timing(linrt.KERNEL_GATHER, reset=True)
    t0 = time.perf_counter()
    rec = []
    for k in range(args.steps):
        rec.append(step())
        log(f"step {k}: photons {rec[-1][0].seconds:.3f}s (trace {rec[-1][0].trace_ms:.1f} ms, map "
            f"{rec[-1][0].sort_ms:.1f} ms), eye {rec[-1][1].seconds:.3f}s")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ph_s = sum(r[0].seconds for r in rec)
    eye_s = sum(r[1].seconds for r in rec)
    eye = rec[-1][1]
    samples = res * res * spp
    # gather roofline: B_query = 56 (query) + 16 n_node + 16 n_tested + 52 n_summed + 24 (sum out) bytes
    g_ms, g_n, _ = scene.kernel_timing(linrt.KERNEL_GATHER)
    q_c = max(cnt.segments, 1)
    per_query = 56 + 24 + (16 * cnt.nodes_visited + 16 * cnt.refs_read + 52 * cnt.tris_tested) / q_c
    queries_per_launch = eye.segments * args.steps / max(g_n, 1)
    g_avg_ms = g_ms / max(g_n, 1)
    g_achieved = per_query * queries_per_launch / (g_avg_ms * 1e-3) / 1e9
    out = {"metric": "photon-mapped frame: eye Msamples/s (1024x1024, 4M photons, 262k tris)",
           "value": round(samples * args.steps / eye_s / 1e6, 3), "unit": "Msamples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded scene generator)",
           "config": {"workload": f"photon box: {n_ph} photons (depth 8), {res}x{res} at {spp} spp, k 64, r 0.05",
                      "photons": n_ph, "width": res, "height": res, "spp": spp},
           "photon_pass_s": round(ph_s / args.steps, 4),
           "mphotons_per_s": round(n_ph * args.steps / ph_s / 1e6, 2),
           "eye_pass_s": round(eye_s / args.steps, 4),
           "eye_rays_per_sample": round((eye.extension_rays + eye.shadow_rays) / samples, 2),
           "eye_mrays_per_s": round((eye.extension_rays + eye.shadow_rays) / eye.seconds / 1e6, 1),
           "eye_gathers_per_sample": round(eye.segments / samples, 2),
           "eye_kernel_ms": {"extend": round(eye.trace_ms, 1), "shadow": round(eye.occluded_ms, 1),
                             "shade": round(eye.shade_ms, 1), "gather": round(g_ms / args.steps, 1)},
           "gather_work_per_query": {"nodes": round(cnt.nodes_visited / q_c, 2),
                                     "photons_tested": round(cnt.refs_read / q_c, 2),
                                     "photons_summed": round(cnt.tris_tested / q_c, 2)},
           "roofline": {"bound": "hbm", "kernel": "k_eye_gather (k-nearest photon gathers)",
                        "achieved": round(g_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(g_achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "bytes_per_launch": int(per_query * queries_per_launch), "avg_launch_ms": round(g_avg_ms, 4),
                        "launches": int(g_n)}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--rays", type=int, default=0, help="ray-traversal benchmark with this many rays instead")
    ap.add_argument("--photon", action="store_true", help="photon-mapper benchmark instead")
    ap.add_argument("--network", action="store_true", help="the denoiser's network alone at 1080p instead")
    ap.add_argument("--no-denoise", action="store_true", help="skip the time-to-relMSE leg")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only copy this many MiB device-to-device 5 times (HBM counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--gather-calibration", type=int, default=0,
                    help="only gather this many MiB of rows at random indices 5 times (counter calibration for "
                         "scattered reads, row size --gather-row bytes)")
    ap.add_argument("--gather-row", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by pixel/sample ranges "
              f"across GPUs but no multi-GPU run is part of this build", file=sys.stderr)
        sys.exit(2)

    import torch
    import linrt

    if args.copy_calibration:
        n = args.copy_calibration * 2**20 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_calibration, "launches": 5, "bytes_read_per_launch": n * 4,
                          "bytes_written_per_launch": n * 4}))
        return
    if args.gather_calibration:
        row = args.gather_row // 4
        n = args.gather_calibration * 2**20 // args.gather_row
        src = torch.ones((n, row), dtype=torch.float32, device="cuda:0")
        idx = torch.randperm(n, generator=torch.Generator().manual_seed(1)).to("cuda:0")
        out = torch.empty_like(src)
        for _ in range(5):
            torch.index_select(src, 0, idx, out=out)
        torch.cuda.synchronize()
        print(json.dumps({"gather_mib": args.gather_calibration, "row_bytes": args.gather_row, "rows": n,
                          "launches": 5, "bytes_read_per_launch": n * args.gather_row + n * 8,
                          "bytes_written_per_launch": n * args.gather_row}))
        return
    if args.rays:
        return ray_bench(args, torch, linrt)
    if args.photon:
        return photon_bench(args, torch, linrt)
    if args.network:
        return network_bench(args, torch, linrt)

    if linrt.device_count() < 1 or not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    gen = linrt.displaced_sphere(seed=SEED)
    gen_s = time.perf_counter() - t0
    scene = linrt.Scene(gen.desc_ptr, device=0)
    info = scene.info
    log(f"scene: {info.n_triangles} tris, {info.n_nodes} nodes, {info.n_refs} refs, depth {info.max_tree_depth}, "
        f"{info.device_bytes / 2**20:.1f} MiB, gen {gen_s:.2f}s build {info.build_seconds:.2f}s")
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    # untimed counting pass over the same frame: algorithmic work per ray
    st_c = scene.render_device(W, H, args.spp, MAX_DEPTH, RR_DEPTH, acc.data_ptr(), seed=SEED, count=True)
    scene.sync()
    log(f"count pass: ext {st_c.extension_rays} shadow {st_c.shadow_rays} nodes {st_c.nodes_visited} "
        f"refs {st_c.refs_read} tris {st_c.tris_tested} shadow nodes {st_c.shadow_nodes_visited} "
        f"refs {st_c.shadow_refs_read} tris {st_c.shadow_tris_tested}")
    for _ in range(args.warmup):
        scene.render_device(W, H, args.spp, MAX_DEPTH, RR_DEPTH, acc.data_ptr(), seed=SEED)
    scene.sync()
    torch.cuda.synchronize()
    scene.kernel_timing(linrt.KERNEL_CLOSEST, reset=True)
    scene.kernel_timing(linrt.KERNEL_OCCLUDED, reset=True)
    scene.kernel_timing(linrt.KERNEL_SHADE, reset=True)
    scene.kernel_timing(linrt.KERNEL_TAIL, reset=True)

    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(scene.render_device(W, H, args.spp, MAX_DEPTH, RR_DEPTH, acc.data_ptr(), seed=SEED))
        log(f"step {k}: {stats[-1].seconds:.3f}s")
    scene.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    samples = sum(s.samples for s in stats)
    ext = sum(s.extension_rays for s in stats)
    shadow = sum(s.shadow_rays for s in stats)
    ms_c, n_c, _ = scene.kernel_timing(linrt.KERNEL_CLOSEST)
    ms_o, n_o, _ = scene.kernel_timing(linrt.KERNEL_OCCLUDED)
    ms_s, n_s, _ = scene.kernel_timing(linrt.KERNEL_SHADE)
    ms_t, n_t, _ = scene.kernel_timing(linrt.KERNEL_TAIL)
    value = samples / elapsed / 1e6

    # roofline of the dominant kernels, the traversal launches (k_extend for
    # camera rays, then per bounce k_trace_queues over the shadow rays and the
    # next extension rays) and the tail launches (k_tail: once few paths are
    # left it finishes them, tracing their remaining rays and shading them, so
    # its whole time is priced as traversal time): B_ray = 32 (ray) + 8 n_node + 4 n_ref + 36 n_tri
    # + 16 (result) bytes, n_* from the counting pass.  `frac` prices the
    # reference's traversal work (full kd window, closest-hit shadow rays);
    # `own_work` prices what the timed kernels do (windows clipped to the mesh
    # box, shadow rays ended by their first hit), counted by the same rules,
    # with B = 32 + 8 n_node + 36 n_tri + 16: the kernels read each leaf's
    # triangle records in reference order and no reference index per test.
    rays = st_c.extension_rays + st_c.shadow_rays
    per_frame_bytes = (32 + 16) * rays + \
        8 * (st_c.nodes_visited + st_c.shadow_nodes_visited) + 4 * (st_c.refs_read + st_c.shadow_refs_read) + \
        36 * (st_c.tris_tested + st_c.shadow_tris_tested)
    own_frame_bytes = (32 + 16) * rays + \
        8 * (st_c.own_nodes_visited + st_c.own_shadow_nodes_visited) + \
        36 * (st_c.own_tris_tested + st_c.own_shadow_tris_tested)
    launches_per_frame = (n_c + n_t) / max(args.steps, 1)
    bytes_per_launch = per_frame_bytes / launches_per_frame
    own_bytes_per_launch = own_frame_bytes / launches_per_frame
    avg_ms = (ms_c + ms_t) / max(n_c + n_t, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    own_achieved = own_bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tf = REPO / "profiles" / "traffic_trace_displaced_sphere.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        traffic = tj.get("hbm_bytes_per_launch")
        traffic_src = (f"{tf.relative_to(REPO)}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, measured "
                       f"apart on build {tj.get('build', '?')} and not re-measured by this run (L2-to-fabric "
                       f"bytes; the 256 MB Infinity Cache sits between them and HBM)")

    cpu = None if args.no_cpu else cpu_baseline(gen)
    ttq = None
    if not args.no_denoise:
        # the second half of the metric: wall time to relMSE <= 1e-2 over the
        # whole frame, judged by the FP64-accumulated 65,536-spp oracle of it;
        # denoised low-spp frames (render + features + network) vs the plain
        # path tracer at the spp that reaches the target
        from linrt import denoise as dn
        r = dn.time_to_quality(log=log, spp_list=(1, 2, 4))
        rnd = lambda d: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in d.items()}  # noqa: E731
        ttq = {"target_relmse": r["target"], "eps": r["eps"],
               "oracle": {**r["oracle"], "self_relmse": float(f"{r['oracle']['self_relmse']:.3e}"),
                          "what": "the whole 1920x1080 frame at 65,536 spp, float64 sums (data/oracle_1080p_65536spp.npz)"},
               "denoised": [rnd(d) for d in r["denoised"]],
               "plain_best": rnd(r["plain_best"]),
               "denoised_best": None if r["denoised_best"] is None else rnd(r["denoised_best"]),
               "speedup": None if r["speedup"] is None else round(r["speedup"], 2),
               "speedup_4spp": None if r["speedup_4spp"] is None else round(r["speedup_4spp"], 2)}
        # the denoiser's own roofline: the U-Net's 10 fused convolutions at
        # 1080p, flops 2*9*cin*cout per output pixel over the summed layer
        # times (HIP events on the denoiser's stream), against the dense bf16
        # MFMA peak and against each layer's own bound
        nw = r["network"]
        ttq["network_roofline"] = {
            "bound": "mfma", "kernel": "k_conv_fwd (U-Net forward, 10 layers, 1920x1080)",
            "achieved": round(nw["tflops"], 1), "peak": dn.BF16_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": round(nw["tflops"] * 1e12 / dn.BF16_PEAK, 4), "traffic": None,
            "flop_per_forward": round(nw["tflop"] * 1e12), "conv_ms": round(nw["conv_ms"], 4),
            "forward_ms": round(nw["forward_ms"], 4),
            "per_layer": nw["layers"]}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": "displaced sphere 1,048,576 tris + ground, 7 BSDF regions, Hosek sky; "
                               f"{W}x{H}, {args.spp} spp, max depth {MAX_DEPTH}, RR from depth {RR_DEPTH}",
                   "width": W, "height": H, "spp": args.spp, "max_depth": MAX_DEPTH, "rr_depth": RR_DEPTH,
                   "triangles": int(info.n_triangles), "parallelism": "single GPU"},
        "mrays_per_s": round((ext + shadow) / elapsed / 1e6, 2),
        "rays_per_sample": round((ext + shadow) / max(samples, 1), 3),
        "kernel_ms_per_step": {"trace": round(ms_c / args.steps, 2), "tail": round(ms_t / args.steps, 2),
                               "shade": round(ms_s / args.steps, 2)},
        "roofline": {"bound": "hbm",
                     "kernel": "k_extend + k_trace_queues + k_tail (kd traversal: extension rays closest hit, "
                               "shadow rays any hit; k_tail also shades the last paths)",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(avg_ms, 4),
                     "launches": int(n_c + n_t), "tail_launches": int(n_t),
                     "frac_basis": "the reference's traversal work: full kd window, closest-hit shadow rays",
                     "own_work": {"what": "the timed kernels' own traversal: box-clipped windows, any-hit shadow rays, "
                                          "leaf triangles read in reference order",
                                  "bytes_per_launch": int(own_bytes_per_launch), "achieved": round(own_achieved, 1),
                                  "frac": round(own_achieved / HBM_PEAK_GBS, 4),
                                  "per_ray": {"nodes": round((st_c.own_nodes_visited + st_c.own_shadow_nodes_visited)
                                                             / max(rays, 1), 2),
                                              "tris": round((st_c.own_tris_tested + st_c.own_shadow_tris_tested)
                                                            / max(rays, 1), 2)}},
                     "reference_work_per_ray": {"nodes": round((st_c.nodes_visited + st_c.shadow_nodes_visited)
                                                               / max(rays, 1), 2),
                                                "tris": round((st_c.tris_tested + st_c.shadow_tris_tested)
                                                              / max(rays, 1), 2)}},
        "cpu_baseline": cpu,
        "time_to_relmse": ttq,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
