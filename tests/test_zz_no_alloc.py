"""lrt_trace_closest_noalloc (the entry point that allocates nothing) through
the C ABI against lrt_trace_closest on the same device buffers: bit-identical
hits on two calls in a row, the second with fewer rays (the shared scratch
counter must be cleared by each call), and the size limit rejected with a
clear error before anything is launched."""
import ctypes as C

import numpy as np
import pytest
import torch

import linrt
from common import random_rays, surface_rays

pytestmark = pytest.mark.gpu


def _run(sc, rays, noalloc):
    n = rays.shape[1]
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).to("cuda:0")
    op = torch.full((n, 2), -7, dtype=torch.int32, device="cuda:0")
    t = torch.full((n,), -7.0, dtype=torch.float32, device="cuda:0")
    uv = torch.full((n, 2), -7.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    f = sc.trace_closest_noalloc_device if noalloc else sc.trace_closest_device
    f(d_rays.data_ptr(), n, op.data_ptr(), t.data_ptr(), uv.data_ptr())
    sc.sync()
    return op.cpu().numpy(), t.cpu().numpy(), uv.cpu().numpy()


def test_noalloc_matches_trace_closest():
    torch.cuda.init()
    g = linrt.displaced_sphere(seed=3, segments=256, rings=128)
    sc = linrt.Scene(g.desc_ptr)
    pos, idx = g.mesh(0)
    big = np.concatenate([surface_rays(pos, idx, 300_000, seed=5),
                          random_rays(np.array([-3.0, -1.0, -3.0]), np.array([3.0, 3.0, 3.0]), 20_000, seed=6)], axis=1)
    small = surface_rays(pos, idx, 70_001, seed=7)
    for rays in (big, small):
        want = _run(sc, rays, noalloc=False)
        got = _run(sc, rays, noalloc=True)
        n = rays.shape[1]
        print(f"{n} rays: hits {(want[0][:, 0] >= 0).mean():.4f}")
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
    # a zero-length call is a no-op (nothing is launched or dereferenced)
    sc.trace_closest_noalloc_device(16, 0, 16, 16)
    # beyond the maximum: LRT_ERR_ARG with a message, nothing launched
    with pytest.raises(linrt.LinrtError) as e:
        linrt._check(linrt.LIB.lrt_trace_closest_noalloc(sc._h, C.c_void_p(16), C.c_uint32((1 << 31) + 1),
                                                          C.c_void_p(16), C.c_void_p(16), None))
    assert e.value.code == linrt.ERR_ARG and "maximum" in str(e.value)
    sc.sync()
