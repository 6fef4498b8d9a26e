"""OBJ loading (lrt_load_obj, obj_loader.h:10-141) and the README scene of
src/main.cpp:55-87 (lrt_generate_obj_scene), on the CPU.

The library's loader runs on the host, so it is compared here with the
reference's own load_obj (oracle/_ref, built from the reference's sources)
on synthetic files that exercise every record form and rule, and on the
reference's bunny.obj where the reference tree is present: positions,
indices and corner normals must be bit-identical.  Where the reference build
is absent the synthetic cases are checked against hand-computed values.
The README bunny scene itself is pinned by rendering it with the oracle in
the reference-sampler mode and comparing every sample with the reference's
own path_integrator on the same scene.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import linrt
import oracle_lib as OL

BUNNY = Path("/root/reference/assets/bunny.obj")

# every rule of load_obj: comments and blank lines, CRLF endings, vertex and
# normal records, the corner forms p, p/t, p/t/n and p//n, negative (relative)
# indices, polygons fanned around their first corner, a zero-area triangle
# that is dropped, and vertices not referenced by any face (moved and scaled
# with the rest, but not part of the bounding box)
SYNTH = """# comment line
v 0 0 0
v 2 0 0\r
v 2 1 0
v 0 1 0

vn 0 0 1
vn 0.6 0 0.8
v 5 5 5
f 1 2 3
f 1/1 3/2 4/3
f -5//-2 -4//-1 -3//-2 -2//-1
v 1 0 0
v 3 0 0
f 1 6 7
v 0.5 0.5 2.0
f 1/7/1 2/7/2 3/7/1 4/7/2 8/7/1
   # indented comment
"""


def product(path):
    return linrt.load_obj(path)


def reference(path):
    nv, nc = C.c_uint32(), C.c_uint32()
    rc = OL.REF.ref_load_obj(str(path).encode(), C.byref(nv), C.byref(nc), None, None, None)
    if rc != 0:
        return None
    pos = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros(nc.value, np.uint32)
    cn = np.zeros((nc.value, 3), np.float32)
    assert OL.REF.ref_load_obj(str(path).encode(), C.byref(nv), C.byref(nc), OL._p(pos), OL._p(idx, C.c_uint32),
                               OL._p(cn)) == 0
    return pos, idx.reshape(-1, 3), cn.reshape(-1, 3, 3)


def test_synthetic_obj_rules(tmp_path):
    f = tmp_path / "s.obj"
    f.write_bytes(SYNTH.encode())
    pos, idx, cn = product(f)
    # 1 + 1 + 2 (quad) + 0 (zero-area) + 3 (pentagon) triangles
    assert idx.shape == (7, 3) and pos.shape == (8, 3)
    # the fans and the relative indices
    assert idx.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 2], [0, 2, 3], [0, 1, 2], [0, 2, 3], [0, 3, 7]]
    # the unreferenced vertex 5 5 5 is outside the box: referenced vertices
    # span x 0..2, y 0..1, z 0..2 -> centre (1, 0.5, 1), scale 1/2
    np.testing.assert_array_equal(pos[0], np.float32([-0.5, -0.25, -0.5]))
    np.testing.assert_array_equal(pos[4], np.float32([2.0, 2.25, 2.0]))
    # face normals where a corner has none, the file's where it has one
    np.testing.assert_array_equal(cn[0], np.float32([[0, 0, 1]] * 3))
    np.testing.assert_array_equal(cn[2], np.float32([[0, 0, 1], [0.6, 0, 0.8], [0, 0, 1]]))
    if OL.REF is not None:
        ref = reference(f)
        for a, b in zip((pos, idx, cn), ref):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("text", [
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",        # index past the vertices
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n",          # fewer than 3 corners
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",        # index 0
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//2 2 3\n",     # normal index past the normals
    "v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n",        # only a zero-area triangle
    "# nothing\n",
])
def test_obj_errors_like_the_reference(tmp_path, text):
    f = tmp_path / "e.obj"
    f.write_text(text)
    with pytest.raises(linrt.LinrtError) as e:
        product(f)
    assert e.value.code == linrt.ERR_ARG
    if OL.REF is not None:
        assert reference(f) is None


def test_missing_obj_file(tmp_path):
    with pytest.raises(linrt.LinrtError):
        product(tmp_path / "none.obj")


@pytest.mark.skipif(OL.REF is None or not BUNNY.exists(), reason="needs the reference tree and its build")
def test_bunny_is_the_reference():
    pos, idx, cn = product(BUNNY)
    ref = reference(BUNNY)
    print(f"bunny: {pos.shape[0]} vertices, {idx.shape[0]} triangles")
    for a, b in zip((pos, idx, cn), ref):
        assert np.array_equal(a, b)


@pytest.mark.skipif(OL.REF is None or not BUNNY.exists(), reason="needs the reference tree and its build")
@pytest.mark.parametrize("name", linrt.README_MATERIALS)
def test_readme_bunny_scene_oracle_is_the_reference(name):
    """The README scenes (the bunny in each of the seven materials, ground
    sphere, rectangle light) from lrt_generate_obj_scene: the oracle's
    per-sample radiance with the reference sampler equals the reference's own
    path_integrator on them."""
    mats, tables = linrt.readme_materials(name)
    g = linrt.ObjScene(BUNNY, mats, tables)
    W = H = 12
    o = OL.oracle.scene(g.desc_ptr).render(W, H, 2, 50, 5, mode=1, per_sample=True)
    r = OL.reference.scene(g.desc_ptr).render(W, H, 2, 50, 5, per_sample=True)
    assert np.array_equal(o["per_sample"], r["per_sample"])
    assert o["per_sample"].max() > 0.0


def test_readme_materials_build():
    for name in linrt.README_MATERIALS:
        mats, tables = linrt.readme_materials(name)
        assert len(mats) >= 1 and tables.shape == (3, 101)
    with pytest.raises(ValueError):
        linrt.readme_materials("chrome")
