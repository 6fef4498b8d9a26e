"""Hit parity on the 1080p workload's own mesh: displaced_sphere(seed=1), the
1,048,576-triangle displaced sphere (kd depth limit 34, the deepest trees and
traversal stacks of any workload) with its ground plane and sky.

Rays: (a) 2^22 synthetic incoherent surface rays (area-uniform origins offset
1e-4 along the normal, cosine directions); (b) the rays a 1-spp 1920x1080
frame (depth 16, RR from 5) traces after its camera rays, captured from the
GPU path tracer itself (lrt_render_capture_rays): extension rays of bounces
>= 1 and next-event shadow rays.  Both sides trace the identical float32 rays.

Bar (closest hit): same (object, primitive) and t within 1e-6 relative on
>= 99.999% of rays, every mismatch listed.  Shadow rays are also checked as
the occlusion query the path tracer makes (any surface other than the ignored
emitter's in [0.001, tmax]) against the oracle's.
"""
import os

import numpy as np
import pytest

import linrt
from common import hit_parity, surface_rays
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 8)


def _report(name, frac, bad, op_g, t_g, op_o, t_o, rays):
    print(f"{name}: parity {frac:.7f} over {len(t_g)} rays, {len(bad)} mismatches", flush=True)
    for i in bad[:500]:
        print(f"  ray {i}: o={rays[0:3, i]} d={rays[3:6, i]} window=[{rays[6, i]}, {rays[7, i]}] "
              f"gpu={op_g[i]} t={t_g[i]!r} oracle={op_o[i]} t={t_o[i]!r}")
    if len(bad) > 500:
        print(f"  ... {len(bad) - 500} more")


@pytest.fixture(scope="module")
def c1():
    g = linrt.displaced_sphere(seed=1)
    pos, idx = g.mesh(0)
    assert idx.shape[0] == 1_048_576
    return g, linrt.Scene(g.desc_ptr), oracle.scene(g.desc_ptr)


def _closest(c1, rays, name):
    g, gs, os_ = c1
    op_g, t_g, _ = gs.trace_closest(rays)
    op_o, t_o = os_.trace(rays, threads=THREADS)
    frac, bad = hit_parity(op_g, t_g, op_o, t_o)
    _report(name, frac, bad, op_g, t_g, op_o, t_o, rays)
    return frac, op_o


def test_c1_synthetic_surface_rays(c1):
    g = c1[0]
    pos, idx = g.mesh(0)
    rays = surface_rays(pos, idx, 1 << 22, seed=101)
    frac, op_o = _closest(c1, rays, "1M-tri synthetic surface rays")
    print(f"  hit fraction {(op_o[:, 0] >= 0).mean():.4f}")
    assert frac >= 0.99999


@pytest.fixture(scope="module")
def frame_rays(c1):
    _, gs, _ = c1
    ext, sh, ign = gs.capture_rays(1920, 1080, 1, 16, 5, seed=7, ext_cap=1 << 23, shadow_cap=1 << 23)
    print(f"captured {ext.shape[1]} extension rays (bounces >= 1), {sh.shape[1]} shadow rays", flush=True)
    return ext, sh, ign


def test_c1_frame_extension_rays(c1, frame_rays):
    ext, _, _ = frame_rays
    assert ext.shape[1] >= 1 << 21 and ext.shape[1] < 1 << 23  # nothing dropped
    assert np.all(ext[6] == np.float32(0.001)) and np.all(np.isinf(ext[7]))
    frac, _ = _closest(c1, ext, "1080p frame extension rays")
    assert frac >= 0.99999


def test_c1_frame_shadow_rays(c1, frame_rays):
    _, sh, ign = frame_rays
    assert sh.shape[1] >= 1 << 20 and sh.shape[1] < 1 << 23
    # the reference's shadow test is a closest-hit query over [0.001, tmax]
    frac, _ = _closest(c1, sh, "1080p frame shadow rays (closest hit over the window)")
    assert frac >= 0.99999
    _, gs, os_ = c1
    a = gs.trace_occluded(sh, ign)
    b = os_.occluded(sh, ign, threads=THREADS)
    bad = np.nonzero(a != b)[0]
    print(f"1080p frame shadow rays (occlusion): {len(bad)} of {len(a)} differ, occluded fraction {b.mean():.4f}")
    for i in bad[:200]:
        print(f"  ray {i}: o={sh[0:3, i]} d={sh[3:6, i]} tmax={sh[7, i]!r} ignore={ign[i]} gpu={a[i]} oracle={b[i]}")
    assert len(bad) <= len(a) * 1e-5
