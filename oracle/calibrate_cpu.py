"""CPU calibration of the oracle against the reference: the restatement
(oracle/liborc.so) and the reference's own path integrator compiled from its
sources (oracle/_ref/libref.so) render the same rows of the same scenes with
the same sample streams (the reference's sampler), on the same host and the
same number of threads; the ratio of their speeds lets the CPU baseline that
bench.py times on another host with the restatement be read as reference
speed.  Test infrastructure: it runs only where the reference tree and
libref.so are present, and writes oracle/cpu_calibration.json.

Both sides run the reference's constants (max depth 50, Russian roulette from
depth 5, path_integrator.h:173,19), which the reference hard-codes.

  python oracle/calibrate_cpu.py [--threads N]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
sys.path.insert(0, str(REPO / "ray-tracing_amd"))
sys.path.insert(0, str(REPO / "tests"))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def timed(fn, repeats=2):
    """Best of `repeats` wall times of fn()."""
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    return best


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    args = ap.parse_args(argv)
    import linrt
    import oracle_lib as OL
    if OL.REF is None:
        raise SystemExit("oracle/_ref/libref.so or the reference tree is missing: nothing to calibrate against")
    th = args.threads
    cases = []
    # C0's scene (box room + icosphere L5, 20,480 tris) at 256x256, 16 spp: the whole frame
    g0 = linrt.box_icosphere(seed=1, subdivisions=5)
    # a crop of a 1,048,576-triangle frame: the displaced sphere of the 1080p
    # workload in one material (the reference takes one material per mesh),
    # copper ground, glass ball, Hosek sky; 8 full-width rows at 16 spp
    g1 = linrt.inverse_sphere(seed=1, segments=1024, rings=512, aspect=16.0 / 9.0)
    for name, gen, W, H, spp, rows in (("C0 box + icosphere L5, 256x256, 16 spp (whole frame)", g0, 256, 256, 16, None),
                                       ("1,048,576-tri displaced sphere, 1920x1080 rows 540..547, 16 spp", g1, 1920, 1080,
                                        16, (540, 548))):
        orc = OL.oracle.scene(gen.desc_ptr)
        ref = OL.reference.scene(gen.desc_ptr)
        n = W * ((rows[1] - rows[0]) if rows else H) * spp
        t_orc = timed(lambda: orc.render(W, H, spp, 50, 5, mode=1, threads=th, rows=rows, fsum=True))
        t_ref = timed(lambda: ref.render(W, H, spp, 50, 5, threads=th, rows=rows, fsum=True))
        cases.append({"workload": name, "samples": n, "restatement_msamples_s": round(n / t_orc / 1e6, 4),
                      "reference_msamples_s": round(n / t_ref / 1e6, 4), "ratio": round(t_ref / t_orc, 4)})
        print(json.dumps(cases[-1]), flush=True)
    out = {"host": {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count(), "threads": th},
           "depth": {"max_depth": 50, "rr_depth": 5, "sampler": "the reference's (sobol.h)"},
           "what": "render-loop wall time (best of 2), scene build excluded; ratio = restatement speed / reference "
                   "speed",
           "cases": cases,
           "ratio": round(sum(c["ratio"] for c in cases) / len(cases), 4)}
    (HERE / "cpu_calibration.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
