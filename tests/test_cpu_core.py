"""CPU tests: the C-ABI library loads and exports every declared symbol; the
oracle is pinned (Philox known-answer vectors, the reference's own kd-tree and
per-sample path radiance through oracle/_ref); host-side product code (kd
build, precompute) matches the oracle/reference exactly.  No GPU call."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import linrt
import oracle_lib as OL

REPO = Path(__file__).resolve().parent.parent
needs_ref = pytest.mark.skipif(OL.reference is None, reason="oracle/_ref not built (reference tree absent)")


def test_library_exports_every_declared_symbol():
    decl = set(re.findall(r"\b(lrt_[a-z0-9_]+)\s*\(", (REPO / "include" / "linrt.h").read_text()))
    decl = {d for d in decl if not d.endswith("_t")}
    missing = [d for d in sorted(decl) if not hasattr(linrt.LIB, d)]
    assert not missing, missing
    assert set(linrt.EXPORTED) <= decl


def test_philox_known_answers():
    """Random123 philox4x32-10 known-answer vectors (kat_vectors)."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert OL.philox(ctr, key) == want
        assert linrt.philox(ctr, key) == want


def _boxes(pos, idx):
    tri = pos[idx]
    return np.concatenate([tri.min(1), tri.max(1)], 1).astype(np.float32)


def _same_tree(a, b):
    for k in ("leaf", "axis", "split", "left", "right", "first", "count", "refs"):
        assert np.array_equal(a[k], b[k]), k
    assert a["depth_limit"] == b["depth_limit"]


@pytest.mark.parametrize("kind", ["icosphere", "room", "sphere_small", "random"])
def test_kdtree_product_equals_oracle(kind):
    if kind == "random":
        rng = np.random.default_rng(5)
        lo = rng.random((3000, 3)).astype(np.float32)
        hi = lo + rng.random((3000, 3)).astype(np.float32) * 0.05
        hi[::7, 1] = lo[::7, 1]  # planar boxes
        boxes = np.concatenate([lo, hi], 1)
    elif kind == "sphere_small":
        g = linrt.displaced_sphere(seed=2, segments=128, rings=64)
        boxes = _boxes(*g.mesh(0))
    else:
        g = linrt.box_icosphere(seed=1, subdivisions=4)
        boxes = _boxes(*g.mesh(1 if kind == "icosphere" else 0))
    _same_tree(linrt.kdtree_build(boxes), OL.oracle.kd_build(boxes))


@needs_ref
@pytest.mark.parametrize("kind", ["icosphere", "sphere_small", "random"])
def test_kdtree_oracle_equals_reference(kind):
    if kind == "random":
        rng = np.random.default_rng(6)
        lo = rng.random((2000, 3)).astype(np.float32)
        boxes = np.concatenate([lo, lo + 0.03 * rng.random((2000, 3)).astype(np.float32)], 1)
    elif kind == "sphere_small":
        boxes = _boxes(*linrt.displaced_sphere(seed=2, segments=128, rings=64).mesh(0))
    else:
        boxes = _boxes(*linrt.box_icosphere(seed=1, subdivisions=4).mesh(1))
    _same_tree(OL.oracle.kd_build(boxes), OL.reference.kd_build(boxes))


def test_kdtree_empty_and_tiny():
    t = linrt.kdtree_build(np.zeros((0, 6), np.float32))
    assert len(t["leaf"]) == 0
    one = np.array([[0, 0, 0, 1, 1, 1]], np.float32)
    t = linrt.kdtree_build(one)
    assert list(t["leaf"]) == [1] and list(t["refs"]) == [0]


@needs_ref
def test_oracle_render_is_the_reference_per_sample():
    """Reference-sampler mode of the oracle reproduces the reference's own
    path_integrator (depth 50, RR from 5) bit for bit on every sample."""
    g = linrt.box_icosphere(seed=1, subdivisions=3)
    a = OL.oracle.scene(g.desc_ptr).render(20, 20, 4, 50, 5, mode=1, per_sample=True, fsum=True)
    b = OL.reference.scene(g.desc_ptr).render(20, 20, 4, 50, 5, per_sample=True, fsum=True)
    assert np.array_equal(a["per_sample"], b["per_sample"])
    assert np.array_equal(a["fsum"], b["fsum"])


@needs_ref
def test_oracle_hits_are_the_reference():
    from common import random_rays, surface_rays
    g = linrt.box_icosphere(seed=1, subdivisions=4)
    pos, idx = g.mesh(1)
    rays = np.concatenate([surface_rays(pos, idx, 20000, 1), random_rays(np.full(3, -0.9), np.full(3, 0.9), 20000, 2)], 1)
    oa, ta = OL.oracle.scene(g.desc_ptr).trace(rays)
    ob, tb = OL.reference.scene(g.desc_ptr).trace(rays)
    assert np.array_equal(oa, ob) and np.array_equal(ta, tb)


def test_oracle_kd_traversal_matches_brute_force():
    from common import hit_parity, surface_rays
    g = linrt.displaced_sphere(seed=4, segments=96, rings=48)
    sc = OL.oracle.scene(g.desc_ptr)
    pos, idx = g.mesh(0)
    rays = surface_rays(pos, idx, 5000, 3)
    oa, ta = sc.trace(rays)
    ob, tb = sc.trace_brute(rays)
    frac, bad = hit_parity(oa, ta, ob, tb)
    assert frac >= 0.999, bad[:20]


@needs_ref
def test_sky_cook_is_the_reference():
    """lrt_sky_cook / lrt_sky_elevation against the reference's own
    arhosek_rgb_skymodelstate_alloc_init and env_emitter elevation (compiled
    from its sources), bit for bit, over turbidity, albedo and sun directions."""
    rng = np.random.default_rng(5)
    n = 0
    for t in (1.0, 1.5, 2.0, 3.0, 4.7, 6.25, 9.0, 9.99, 10.0):
        for a in (0.0, 0.1, 0.3, 0.77, 1.0):
            for _ in range(4):
                sun = rng.normal(size=3).astype(np.float32)
                sun[1] = abs(sun[1]) + 0.01
                el_ref = C.c_double()
                OL.REF.ref_sky_elevation(OL._p(sun), C.byref(el_ref))
                el = linrt.sky_elevation(sun)
                assert el == el_ref.value
                tf, af = float(np.float32(t)), float(np.float32(a))
                cfg, r3 = linrt.sky_cook(tf, af, el)
                rc = np.zeros(27, np.float64)
                rr = np.zeros(3, np.float64)
                OL.REF.ref_hosek_cook(C.c_float(t), C.c_float(a), C.c_double(el), OL._p(rc, C.c_double),
                                      OL._p(rr, C.c_double))
                assert np.array_equal(cfg.ravel(), rc) and np.array_equal(r3, rr), (t, a, el)
                n += 1
    print(f"{n} cooks bit-identical")


def test_device_code_defines_registered_scratch():
    """The library's gfx950 code object defines the static device array that
    its host code registers for lrt_trace_closest_noalloc (a library whose
    host and device passes saw different sources would abort when HIP loads
    it).  Reads the library built here, with the ROCm tools."""
    import shutil
    import subprocess
    import tempfile
    bundler = Path("/opt/rocm/llvm/bin/clang-offload-bundler")
    readelf = Path("/opt/rocm/llvm/bin/llvm-readelf")
    if not (bundler.exists() and readelf.exists() and shutil.which("objcopy")):
        pytest.skip("ROCm LLVM tools or objcopy not found")
    with tempfile.TemporaryDirectory() as d:
        fat, co = Path(d) / "fat.bin", Path(d) / "gfx950.co"
        subprocess.run(["objcopy", "--dump-section", f".hip_fatbin={fat}", str(linrt.LIB_PATH), str(Path(d) / "x.so")],
                       check=True, capture_output=True)
        subprocess.run([str(bundler), "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={fat}",
                        f"--output={co}", "--unbundle"], check=True, capture_output=True)
        syms = subprocess.run([str(readelf), "-s", str(co)], check=True, capture_output=True, text=True).stdout
    assert "_ZL7scratch" in syms
    assert "k_trace_closest" in syms


@needs_ref
def test_mesh_and_sphere_lights_oracle_is_the_reference():
    """Area lights on a non-convex mesh and on a sphere (area_emitter.h over
    triangle_mesh::sample_position and sphere::sample_position, the pdf's ray
    cast of the emitter's own shape, the closest-hit shadow test): the oracle's
    per-sample radiance (reference sampler, depth 50, RR 5) and its traced
    particles equal the reference's own, bit for bit."""
    from common import EmitterRoom
    er = EmitterRoom()
    ps_o = OL.oracle.scene(er.desc_ptr).render(16, 16, 4, 50, 5, mode=1, per_sample=True)["per_sample"]
    ps_r = OL.reference.scene(er.desc_ptr).render(16, 16, 4, 50, 5, per_sample=True)["per_sample"]
    assert np.array_equal(ps_o, ps_r)
    assert (ps_r > 0).mean() > 0.5
    po, do = OL.oracle.scene(er.desc_ptr).trace_particles(0, 800, 8, 5, mode=1)
    pr, dr = OL.reference.scene(er.desc_ptr).trace_particles(0, 800, 8, 5)
    assert len(po) > 200 and np.array_equal(po, pr) and np.array_equal(do, dr)


def test_torch_function_module_loads():
    """The C++ autograd Function module (linrt_torch) loads, binds the
    library's differentiable-rendering entry points and refuses a CPU tensor
    before any GPU call."""
    import torch
    from linrt import diff
    ext = diff.torch_ext()
    with pytest.raises(RuntimeError, match="GPU tensor"):
        ext.render(torch.zeros(12291), 1, [0, 64, 64, 1, 2, 1], [8, 8, 1, 0, 8, 1 << 20, 1, 0, 0, 0])
    with pytest.raises(RuntimeError, match="6 integers"):
        ext.render(torch.zeros(12291), 1, [0, 64], [8, 8, 1, 0, 8, 1 << 20, 1, 0, 0, 0])
