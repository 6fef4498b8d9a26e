"""Python host layer of liblinrt_hip.so (the MI355X render hot path of LinRT).

A thin ctypes mirror of ``include/linrt.h``: the structs keep the C layout and
every call goes through the C ABI, so what runs is the hand-written HIP code.
The roles follow the reference's interfaces:

* accelerator  -- ``kdtree_build``            (sah_kdtree.h:35-57)
* scene / hit  -- ``Scene.trace_closest``      (scene.h:36-38, hittable_list.h:19-34)
* materials    -- ``bsdf_eval_pdf`` / ``bsdf_sample`` (bsdf.h:78-88)
* emitters     -- ``sky_eval``                 (env_emitter.h:71-88)
* integrator   -- ``Scene.render``             (path_integrator.h:18-170 driven
                  like the frame loop of src/main.cpp:103-147)

The library must be built (``__graft_entry__.build()``); importing this
module without it raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
PKG_ROOT = _HERE.parent
# LINRT_LIB selects another build of the same C ABI (e.g. a tuning variant)
LIB_PATH = Path(os.environ.get("LINRT_LIB", str(PKG_ROOT / "liblinrt_hip.so")))
DATA_DIR = PKG_ROOT / "data"

OK = 0
ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_UNSUPPORTED, ERR_INTERNAL = -1, -2, -3, -4, -5

# BSDF flags (bsdf.h:18-28)
DIFFUSE, SPECULAR, REFLECTIVE, TRANSPARENT, GLOSSY, METALLIC, DIELECTRIC, REFRACTIVE = (1 << i for i in range(8))
MEASURE_SOLID_ANGLE, MEASURE_DISCRETE = 1, 5
MODE_RADIANCE, MODE_IMPORTANCE = 0, 1
(BSDF_NONE, BSDF_DIFFUSE, BSDF_CONDUCTOR, BSDF_DIELECTRIC, BSDF_ROUGH_DIELECTRIC, BSDF_PLASTIC,
 BSDF_ROUGH_PLASTIC, BSDF_COATING, BSDF_ROUGH_COATING) = range(9)
SHAPE_MESH, SHAPE_RECTANGLE, SHAPE_SPHERE = 0, 1, 2
LIGHT_AREA, LIGHT_SKY = 0, 1
GEN_BOX_ICOSPHERE, GEN_DISPLACED_SPHERE, GEN_PHOTON_BOX, GEN_DISPLACED_SPHERE_VARIANT, GEN_INVERSE_SPHERE = range(5)
PHOTON_SURFACE, PHOTON_CAUSTIC = 0, 1
KERNEL_CLOSEST, KERNEL_OCCLUDED, KERNEL_SHADE, KERNEL_GATHER, KERNEL_TAIL = 0, 1, 2, 3, 4

F3 = C.c_float * 3


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("nested", C.c_int32), ("int_ior", C.c_float), ("ext_ior", C.c_float),
                ("alpha", C.c_float), ("nonlinear", C.c_int32), ("diffuse", F3), ("specular_reflectance", F3),
                ("specular_transmittance", F3), ("cond_eta", F3), ("cond_k", F3), ("sigma_a", F3),
                ("thickness", C.c_float), ("fdr_int", C.c_float), ("table_ext", C.c_int32), ("table_int", C.c_int32)]


class ObjectDesc(C.Structure):
    _fields_ = [("shape", C.c_int32), ("material", C.c_int32), ("emitter", C.c_int32),
                ("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32),
                ("positions", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint32)),
                ("corner_normals", C.POINTER(C.c_float)), ("triangle_material", C.POINTER(C.c_uint16)),
                ("corner", F3), ("edge_u", F3), ("edge_v", F3), ("center", F3), ("radius", C.c_float)]


class LightDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("object", C.c_int32), ("radiance", F3), ("sun_dir", F3),
                ("turbidity", C.c_float), ("ground_albedo", C.c_float), ("scale", C.c_float),
                ("config", (C.c_double * 9) * 3), ("channel_radiance", C.c_double * 3)]


class CameraDesc(C.Structure):
    _fields_ = [("origin", F3), ("look_at", F3), ("up", F3), ("fov_deg", C.c_float), ("aspect", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("n_objects", C.c_uint32), ("objects", C.POINTER(ObjectDesc)),
                ("n_materials", C.c_uint32), ("materials", C.POINTER(Material)),
                ("n_lights", C.c_uint32), ("lights", C.POINTER(LightDesc)),
                ("n_tables", C.c_uint32), ("tables", C.POINTER(C.c_float)), ("camera", CameraDesc)]


class SceneInfo(C.Structure):
    _fields_ = [("n_triangles", C.c_uint64), ("n_nodes", C.c_uint64), ("n_refs", C.c_uint64),
                ("device_bytes", C.c_uint64), ("max_tree_depth", C.c_int32),
                ("build_seconds", C.c_double), ("upload_seconds", C.c_double),
                ("bounding_radius", C.c_float), ("has_delta_bsdf", C.c_int32)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("spp_offset", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("seed", C.c_uint64),
                ("row_begin", C.c_uint32), ("row_end", C.c_uint32), ("flags", C.c_uint32)]

RENDER_COUNT = 1
RENDER_REFERENCE_SAMPLER = 2
RENDER_WAVEFRONT_ONLY = 8


class RenderStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("segments", C.c_uint64), ("seconds", C.c_double), ("trace_ms", C.c_double),
                ("occluded_ms", C.c_double), ("shade_ms", C.c_double),
                ("trace_launches", C.c_uint64), ("occluded_launches", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("refs_read", C.c_uint64), ("tris_tested", C.c_uint64),
                ("shadow_nodes_visited", C.c_uint64), ("shadow_refs_read", C.c_uint64),
                ("shadow_tris_tested", C.c_uint64),
                ("own_nodes_visited", C.c_uint64), ("own_refs_read", C.c_uint64), ("own_tris_tested", C.c_uint64),
                ("own_shadow_nodes_visited", C.c_uint64), ("own_shadow_refs_read", C.c_uint64),
                ("own_shadow_tris_tested", C.c_uint64)]


class PhotonParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("requested", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("seed", C.c_uint64), ("radius", C.c_float), ("k", C.c_uint32), ("tree", C.c_uint32)]

PHOTON_TREE_DEVICE, PHOTON_TREE_REFERENCE = 0, 1


class PhotonStats(C.Structure):
    _fields_ = [("photons", C.c_uint64), ("shot", C.c_uint64), ("particles_traced", C.c_uint64),
                ("returned", C.c_uint64), ("full", C.c_int32), ("trace_ms", C.c_double), ("sort_ms", C.c_double),
                ("seconds", C.c_double)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("segments", C.c_uint32), ("rings", C.c_uint32),
                ("subdivisions", C.c_uint32), ("aspect", C.c_float)]


class LinrtError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"linrt error {code}: {message}")
        self.code = code


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    # torch ships its own copy of the HIP runtime, found by a different file
    # name than the system copy this library links against: loading torch
    # first makes the library bind to that same copy, so both share one
    # runtime in a process that uses torch for device buffers
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    lib.lrt_last_error.restype = C.c_char_p
    lib.lrt_generated_desc.restype = C.POINTER(SceneDesc)
    lib.lrt_generated_desc.argtypes = [C.c_void_p]
    return lib


LIB = _load()
EXPORTED = [
    "lrt_last_error", "lrt_version", "lrt_device_count", "lrt_kdtree_build_export", "lrt_scene_create",
    "lrt_scene_release", "lrt_scene_info_get", "lrt_trace_closest", "lrt_trace_occluded", "lrt_trace_count",
    "lrt_scene_sync", "lrt_scene_kernel_timing", "lrt_bsdf_eval_pdf", "lrt_bsdf_sample", "lrt_bsdf_flags",
    "lrt_fresnel_diffuse_reflectance", "lrt_rough_transmittance_slice", "lrt_sky_eval", "lrt_render",
    "lrt_photon_build", "lrt_photon_export", "lrt_photon_map_load", "lrt_photon_estimate", "lrt_render_photon",
    "lrt_generate_scene", "lrt_generated_desc", "lrt_generated_free", "lrt_philox4x32_10", "lrt_render_aov",
    "lrt_denoiser_create", "lrt_denoiser_release", "lrt_denoiser_param_count", "lrt_denoiser_params",
    "lrt_denoiser_forward", "lrt_denoiser_loss_grad", "lrt_denoiser_adam_step", "lrt_conv3x3", "lrt_denoiser_timing", "lrt_render_features",
    "lrt_diff_param_count", "lrt_diff_render", "lrt_diff_adjoint", "lrt_diff_render_held",
    "lrt_load_obj", "lrt_mesh_view", "lrt_mesh_free", "lrt_generate_obj_scene", "lrt_image_write",
    "lrt_image_to_bytes", "lrt_render_capture_rays", "lrt_trace_closest_noalloc", "lrt_sky_cook",
    "lrt_sky_elevation", "lrt_accum_write", "lrt_accum_read", "lrt_photon_map_tree", "lrt_bsdf_param_grad",
]


def _check(rc):
    if rc != OK:
        raise LinrtError(rc, LIB.lrt_last_error().decode())


def _p(a, ctype=C.c_float):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def device_count():
    return int(LIB.lrt_device_count())


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    _check(LIB.lrt_philox4x32_10(c, k, o))
    return list(o)


def kdtree_build(boxes):
    """SAH kd-tree in the reference's node numbering (sah_kdtree.h:35-57)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = boxes.shape[0]
    nn, nr, dl = C.c_uint32(), C.c_uint32(), C.c_int32()
    z = None
    _check(LIB.lrt_kdtree_build_export(_p(boxes), C.c_uint32(n), C.byref(nn), C.byref(nr), C.byref(dl),
                                       z, z, z, z, z, z, z, z))
    out = dict(leaf=np.zeros(nn.value, np.uint8), axis=np.zeros(nn.value, np.int32),
               split=np.zeros(nn.value, np.float32), left=np.zeros(nn.value, np.int32),
               right=np.zeros(nn.value, np.int32), first=np.zeros(nn.value, np.int32),
               count=np.zeros(nn.value, np.int32), refs=np.zeros(max(nr.value, 1), np.uint32))
    _check(LIB.lrt_kdtree_build_export(_p(boxes), C.c_uint32(n), C.byref(nn), C.byref(nr), C.byref(dl),
                                       _p(out["leaf"], C.c_uint8), _p(out["axis"], C.c_int32),
                                       _p(out["split"]), _p(out["left"], C.c_int32), _p(out["right"], C.c_int32),
                                       _p(out["first"], C.c_int32), _p(out["count"], C.c_int32),
                                       _p(out["refs"], C.c_uint32)))
    out["refs"] = out["refs"][: nr.value]
    out["depth_limit"] = dl.value
    return out


def fresnel_diffuse_reflectance(eta):
    o = C.c_float()
    _check(LIB.lrt_fresnel_diffuse_reflectance(C.c_float(eta), C.byref(o)))
    return o.value


def rough_transmittance_slice(table_bytes, eta, alpha):
    out = np.zeros(101, np.float32)
    buf = C.create_string_buffer(table_bytes, len(table_bytes))
    _check(LIB.lrt_rough_transmittance_slice(buf, C.c_size_t(len(table_bytes)), C.c_float(eta),
                                             C.c_float(alpha), _p(out)))
    return out


# ---- committed data: rough-transmittance slices and cooked sky --------------
def load_tables(name="tables_displaced_sphere.json"):
    d = json.loads((DATA_DIR / name).read_text())
    return np.asarray(d["tables"], np.float32).reshape(-1, 101), d


def load_sky(name="sky_t3_a0.3_e30.json"):
    d = json.loads((DATA_DIR / name).read_text())
    L = LightDesc()
    L.kind = LIGHT_SKY
    L.object = -1
    for i in range(3):
        L.sun_dir[i] = d["sun_dir"][i]
    L.turbidity = d["turbidity"]
    L.ground_albedo = d["ground_albedo"]
    L.scale = d["scale"]
    for c in range(3):
        for k in range(9):
            L.config[c][k] = d["config"][c][k]
        L.channel_radiance[c] = d["channel_radiance"][c]
    return L


_HOSEK = None


def hosek_dataset():
    """The Hosek-Wilkie RGB datasets (coefficients [3, 1080], radiances [3, 120])
    from data/hosek_rgb_dataset.json (BSD-3, notice in the file)."""
    global _HOSEK
    if _HOSEK is None:
        d = json.loads((DATA_DIR / "hosek_rgb_dataset.json").read_text())
        _HOSEK = (np.ascontiguousarray(d["coefficients"], np.float64), np.ascontiguousarray(d["radiances"], np.float64))
    return _HOSEK


def sky_elevation(sun_dir):
    """Solar elevation the sky emitter cooks its model at (lrt_sky_elevation)."""
    e = C.c_double()
    _check(LIB.lrt_sky_elevation(_p(np.ascontiguousarray(sun_dir, np.float32)), C.byref(e)))
    return e.value


def sky_cook(turbidity, albedo, elevation):
    """Hosek-Wilkie RGB state (lrt_sky_cook): config [3, 9], radiance [3]."""
    coef, rad = hosek_dataset()
    cfg = np.zeros((3, 9), np.float64)
    r3 = np.zeros(3, np.float64)
    _check(LIB.lrt_sky_cook(_p(coef, C.c_double), _p(rad, C.c_double), C.c_double(turbidity), C.c_double(albedo),
                            C.c_double(elevation), _p(cfg, C.c_double), _p(r3, C.c_double)))
    return cfg, r3


def cook_sky(sun_dir, turbidity=3.0, ground_albedo=0.3, scale=0.05):
    """A sky light (env_emitter, env_emitter.h:13-30) cooked by the library:
    the sun direction, turbidity and albedo are float, as the emitter keeps
    them, and the model is cooked at the emitter's elevation."""
    L = LightDesc()
    L.kind = LIGHT_SKY
    L.object = -1
    for i in range(3):
        L.sun_dir[i] = float(sun_dir[i])
    L.turbidity = float(turbidity)
    L.ground_albedo = float(ground_albedo)
    L.scale = float(scale)
    # the emitter passes its float turbidity and albedo to the double cook
    cfg, r3 = sky_cook(L.turbidity, L.ground_albedo, sky_elevation(np.array(L.sun_dir[:], np.float32)))
    for c in range(3):
        for k in range(9):
            L.config[c][k] = cfg[c, k]
        L.channel_radiance[c] = r3[c]
    return L


class Generated:
    """A seeded scene owned by the library; ``desc`` stays valid while alive."""

    def __init__(self, kind, seed=1, segments=1024, rings=512, subdivisions=5, aspect=0.0, tables=None, sky=None):
        p = GenParams(kind, seed, segments, rings, subdivisions, aspect)
        self._tables = None if tables is None else np.ascontiguousarray(tables, np.float32)
        self._sky = sky
        h = C.c_void_p()
        nt = 0 if self._tables is None else self._tables.size // 101
        _check(LIB.lrt_generate_scene(C.byref(p), _p(self._tables), C.c_uint32(nt),
                                      C.byref(sky) if sky is not None else None, C.byref(h)))
        self._h = h
        self.desc_ptr = LIB.lrt_generated_desc(h)
        self.desc = self.desc_ptr.contents

    def close(self):
        if self._h and LIB is not None:
            LIB.lrt_generated_free(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def mesh(self, obj):
        o = self.desc.objects[obj]
        pos = np.ctypeslib.as_array(o.positions, (o.n_vertices * 3,)).reshape(-1, 3).copy()
        idx = np.ctypeslib.as_array(o.indices, (o.n_triangles * 3,)).reshape(-1, 3).copy()
        return pos, idx


class ObjScene(Generated):
    """The scene of src/main.cpp:55-87 around the mesh of an OBJ file
    (lrt_generate_obj_scene): materials[0] on the mesh, the others nested by
    it; the ground sphere takes the next material index."""

    def __init__(self, path, materials, tables=None, aspect=1.0):
        self._tables = None if tables is None else np.ascontiguousarray(tables, np.float32)
        self._mats = (Material * len(materials))(*materials)
        h = C.c_void_p()
        nt = 0 if self._tables is None else self._tables.size // 101
        self._h = None
        _check(LIB.lrt_generate_obj_scene(str(path).encode(), self._mats, C.c_uint32(len(materials)),
                                          _p(self._tables), C.c_uint32(nt), C.c_float(aspect), C.byref(h)))
        self._h = h
        self.desc_ptr = LIB.lrt_generated_desc(h)
        self.desc = self.desc_ptr.contents


def load_obj(path):
    """load_obj (obj_loader.h:44-141) through lrt_load_obj: positions [V, 3],
    indices [T, 3], corner normals [T, 3, 3]."""
    h = C.c_void_p()
    _check(LIB.lrt_load_obj(str(path).encode(), C.byref(h)))
    try:
        nv, nt = C.c_uint32(), C.c_uint32()
        pp, ip, npp = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)(), C.POINTER(C.c_float)()
        _check(LIB.lrt_mesh_view(h, C.byref(nv), C.byref(nt), C.byref(pp), C.byref(ip), C.byref(npp)))
        pos = np.ctypeslib.as_array(pp, (nv.value * 3,)).reshape(-1, 3).copy()
        idx = np.ctypeslib.as_array(ip, (nt.value * 3,)).reshape(-1, 3).copy()
        cn = np.ctypeslib.as_array(npp, (nt.value * 9,)).reshape(-1, 3, 3).copy()
    finally:
        LIB.lrt_mesh_free(h)
    return pos, idx, cn


IMAGE_FORMATS = {"png": 0, "jpeg": 1, "ppm": 2, "hdr": 3, "pfm": 4, "raw": 5, "exr": 6}


class AccumInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp_done", C.c_uint64),
                ("next_sample", C.c_uint64), ("seed", C.c_uint64), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("has_sumsq", C.c_int32)]


def write_accum(path, sum_rgb, sumsq_rgb=None, spp_done=0, next_sample=None, seed=1, max_depth=16, rr_depth=5):
    """lrt_accum_write: float64 sums [H, W, 3] (and sums of squares) of a
    progressive render with their JSON sidecar (path + ".json")."""
    s = np.ascontiguousarray(sum_rgb, np.float64)
    q = None if sumsq_rgb is None else np.ascontiguousarray(sumsq_rgb, np.float64)
    H, W = s.shape[:2]
    info = AccumInfo(W, H, spp_done, spp_done if next_sample is None else next_sample, seed, max_depth, rr_depth, 0)
    _check(LIB.lrt_accum_write(str(path).encode(), _p(s, C.c_double), _p(q, C.c_double), C.byref(info)))


def read_accum(path, info_only=False):
    """lrt_accum_read: (sum [H, W, 3], sumsq [H, W, 3] or None, info dict)."""
    info = AccumInfo()
    _check(LIB.lrt_accum_read(str(path).encode(), None, None, C.byref(info)))
    meta = {k: getattr(info, k) for k, _ in AccumInfo._fields_}
    if info_only:
        return None, None, meta
    s = np.zeros((info.height, info.width, 3), np.float64)
    q = np.zeros_like(s) if info.has_sumsq else None
    _check(LIB.lrt_accum_read(str(path).encode(), _p(s, C.c_double), _p(q, C.c_double), C.byref(info)))
    return s, q, meta


def write_image(path, rgb, fmt=None, flip=False):
    """lrt_image_write: rgb float [H, W, 3] (row 0 first); fmt from the file
    suffix by default (png, ppm, hdr, pfm, raw, exr).  The reference's main.cpp
    saves its bottom-up frame as PNG with flip=True."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    fmt = fmt or str(path).rsplit(".", 1)[-1].lower()
    _check(LIB.lrt_image_write(str(path).encode(), _p(rgb), C.c_uint32(rgb.shape[1]), C.c_uint32(rgb.shape[0]),
                               C.c_int(IMAGE_FORMATS[fmt]), C.c_int(1 if flip else 0)))


def image_to_bytes(rgb, flip=False):
    """The reference's 8-bit mapping of a float frame (image_write.cpp:10-20)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(rgb.shape, np.uint8)
    _check(LIB.lrt_image_to_bytes(_p(rgb), C.c_uint32(rgb.shape[1]), C.c_uint32(rgb.shape[0]),
                                  C.c_int(1 if flip else 0), _p(out, C.c_uint8)))
    return out


# the bunny materials of src/main.cpp:38-47, by the names of the README renders
README_MATERIALS = ("diffuse", "rough-conductor", "plastic", "rough-plastic", "coating", "rough-coating",
                    "rough-dielectric")


def readme_materials(name):
    """(materials, tables) of one README render: the mesh material first, the
    materials it nests after it; tables from data/tables_readme.json."""
    tables, _ = load_tables("tables_readme.json")
    f32 = np.float32
    eta_p = float(f32(1.49) / f32(1.000277))
    plastic = material(BSDF_PLASTIC, int_ior=1.49, ext_ior=1.000277, diffuse=(0.8, 0.8, 0.5),
                       fdr_int=fresnel_diffuse_reflectance(float(f32(1.0) / f32(eta_p))))
    rough_glass = material(BSDF_ROUGH_DIELECTRIC, int_ior=1.5, ext_ior=1.0, alpha=0.05)
    if name == "diffuse":
        mats = [material(BSDF_DIFFUSE, diffuse=(0.3, 0.6, 0.6))]
    elif name == "rough-conductor":
        mats = [material(BSDF_CONDUCTOR, alpha=0.3, cond_eta=(0.200438, 0.924033, 1.102212),
                         cond_k=(3.912948, 2.452847, 2.142188))]
    elif name == "plastic":
        mats = [plastic]
    elif name == "rough-plastic":
        mats = [material(BSDF_ROUGH_PLASTIC, int_ior=1.49, ext_ior=1.000277, alpha=0.1, diffuse=(0.8, 0.1, 0.05),
                         table_ext=0, table_int=1)]
    elif name == "coating":
        mats = [material(BSDF_COATING, int_ior=1.7, ext_ior=1.000277, nested=1, specular_reflectance=(0.3, 0.6, 0.6)),
                plastic]
    elif name == "rough-coating":
        mats = [material(BSDF_ROUGH_COATING, int_ior=1.7, ext_ior=1.000277, alpha=0.3, nested=1, table_ext=2),
                rough_glass]
    elif name == "rough-dielectric":
        mats = [rough_glass]
    else:
        raise ValueError(f"unknown material {name!r}: one of {README_MATERIALS}")
    return mats, tables


def box_icosphere(seed=1, subdivisions=5, aspect=1.0):
    """Closed box room + icosphere + ceiling light (the 256x256 16-spp workload)."""
    return Generated(GEN_BOX_ICOSPHERE, seed=seed, subdivisions=subdivisions, aspect=aspect)


def displaced_sphere(seed=1, segments=1024, rings=512, aspect=16.0 / 9.0):
    """Displaced sphere with 7 material regions, ground and Hosek sky (1080p workload)."""
    tables, _ = load_tables()
    return Generated(GEN_DISPLACED_SPHERE, seed=seed, segments=segments, rings=rings, aspect=aspect,
                     tables=tables, sky=load_sky())


def displaced_sphere_variant(seed=1, segments=1024, rings=512, aspect=16.0 / 9.0, sky=None):
    """The displaced-sphere scene with seeded materials (and mesh) and a given
    sky (default: the one of the 1080p workload): denoiser training scenes."""
    tables, _ = load_tables()
    return Generated(GEN_DISPLACED_SPHERE_VARIANT, seed=seed, segments=segments, rings=rings, aspect=aspect,
                     tables=tables, sky=sky if sky is not None else load_sky())


def load_sky_set(name="skies_train.json"):
    """Cooked Hosek skies of several sun positions and turbidities."""
    out = []
    for d in json.loads((DATA_DIR / name).read_text())["skies"]:
        L = LightDesc()
        L.kind = LIGHT_SKY
        L.object = -1
        for i in range(3):
            L.sun_dir[i] = d["sun_dir"][i]
        L.turbidity, L.ground_albedo, L.scale = d["turbidity"], d["ground_albedo"], d["scale"]
        for c in range(3):
            for k in range(9):
                L.config[c][k] = d["config"][c][k]
            L.channel_radiance[c] = d["channel_radiance"][c]
        out.append(L)
    return out


def photon_box(seed=1, segments=512, rings=256, aspect=1.0):
    """Box room around a 262,144-triangle displaced sphere (7 material
    regions, radius 0.5) under a 0.5 x 0.5 ceiling light (photon workload)."""
    tables, _ = load_tables()
    return Generated(GEN_PHOTON_BOX, seed=seed, segments=segments, rings=rings, aspect=aspect, tables=tables)


def inverse_sphere(seed=1, segments=512, rings=256, aspect=1.0):
    """The differentiable renderer's scene: the displaced sphere (262,144
    triangles by default) in one diffuse material (0, the one a texture covers),
    a rough copper ground (2) and a rough glass ball (1), under the Hosek sky."""
    return Generated(GEN_INVERSE_SPHERE, seed=seed, segments=segments, rings=rings, aspect=aspect, sky=load_sky())


class Scene:
    """Scene uploaded to one GPU (lrt_scene_create)."""

    def __init__(self, desc_ptr, device=0):
        self._h = None
        h = C.c_void_p()
        _check(LIB.lrt_scene_create(C.c_int(device), desc_ptr, C.byref(h)))
        self._h = h

    def close(self):
        if self._h and LIB is not None:
            LIB.lrt_scene_release(self._h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def info(self):
        i = SceneInfo()
        _check(LIB.lrt_scene_info_get(self._h, C.byref(i)))
        return i

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(8, -1)
        n = rays.shape[1]
        op = np.zeros((n, 2), np.int32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        _check(LIB.lrt_trace_closest(self._h, _p(rays), C.c_uint32(n), _p(op, C.c_int32), _p(t), _p(uv), 0))
        return op, t, uv

    def trace_occluded(self, rays, ignore=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(8, -1)
        n = rays.shape[1]
        out = np.zeros(n, np.uint8)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        _check(LIB.lrt_trace_occluded(self._h, _p(rays), C.c_uint32(n), _p(ig, C.c_int32), _p(out, C.c_uint8), 0))
        return out.astype(bool)

    def trace_count(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(8, -1)
        n = rays.shape[1]
        a, b, c = (np.zeros(n, np.uint32) for _ in range(3))
        _check(LIB.lrt_trace_count(self._h, _p(rays), C.c_uint32(n), _p(a, C.c_uint32), _p(b, C.c_uint32),
                                   _p(c, C.c_uint32), 0))
        return a, b, c

    def capture_rays(self, width, height, spp, max_depth, rr_depth, seed=1, ext_cap=1 << 22, shadow_cap=1 << 22):
        """The rays a frame traces after its camera rays (lrt_render_capture_rays):
        (extension rays [8, n], shadow rays [8, m], shadow ignore [m])."""
        ext = np.zeros((8, ext_cap), np.float32)
        sh = np.zeros((8, shadow_cap), np.float32)
        ign = np.zeros(shadow_cap, np.int32)
        ne, ns = C.c_uint32(), C.c_uint32()
        prm = RenderParams(width, height, spp, 0, max_depth, rr_depth, seed, 0, 0, 0)
        _check(LIB.lrt_render_capture_rays(self._h, C.byref(prm), _p(ext), C.c_uint32(ext_cap), _p(sh),
                                           _p(ign, C.c_int32), C.c_uint32(shadow_cap), C.byref(ne), C.byref(ns)))
        e = ext.reshape(-1)[: 8 * ne.value].reshape(8, ne.value).copy()
        s = sh.reshape(-1)[: 8 * ns.value].reshape(8, ns.value).copy()
        return e, s, ign[: ns.value].copy()

    def trace_closest_device(self, rays_ptr, n, obj_prim_ptr, t_ptr, uv_ptr=None):
        _check(LIB.lrt_trace_closest(self._h, C.c_void_p(rays_ptr), C.c_uint32(n), C.c_void_p(obj_prim_ptr),
                                     C.c_void_p(t_ptr), C.c_void_p(uv_ptr) if uv_ptr else None, 1))

    def trace_closest_noalloc_device(self, rays_ptr, n, obj_prim_ptr, t_ptr, uv_ptr=None):
        """lrt_trace_closest_noalloc: device pointers, asynchronous, no allocation."""
        _check(LIB.lrt_trace_closest_noalloc(self._h, C.c_void_p(rays_ptr), C.c_uint32(n), C.c_void_p(obj_prim_ptr),
                                             C.c_void_p(t_ptr), C.c_void_p(uv_ptr) if uv_ptr else None))

    def sync(self):
        _check(LIB.lrt_scene_sync(self._h))

    def kernel_timing(self, kernel=KERNEL_CLOSEST, reset=False):
        ms, nl, it = C.c_double(), C.c_uint64(), C.c_uint64()
        _check(LIB.lrt_scene_kernel_timing(self._h, C.c_int(kernel), C.c_int(int(reset)), C.byref(ms),
                                           C.byref(nl), C.byref(it)))
        return ms.value, nl.value, it.value

    def render(self, width, height, spp, max_depth, rr_depth, seed=1, spp_offset=0, rows=None, sumsq=True,
               sum_rgb=None, sumsq_rgb=None, reference_sampler=False, wavefront_only=False):
        """Adds spp samples per pixel into (sum, sumsq) float64 buffers [H, W, 3]
        (row 0 = bottom, like src/main.cpp:122) and returns them with stats.
        reference_sampler: the reference's Sobol sampler instead of Philox.
        wavefront_only: no tail kernel (every bounce a wavefront launch)."""
        if sum_rgb is None:
            sum_rgb = np.zeros((height, width, 3), np.float64)
        if sumsq and sumsq_rgb is None:
            sumsq_rgb = np.zeros((height, width, 3), np.float64)
        rb, re = rows if rows is not None else (0, 0)
        prm = RenderParams(width, height, spp, spp_offset, max_depth, rr_depth, seed, rb, re,
                           (RENDER_REFERENCE_SAMPLER if reference_sampler else 0) |
                           (RENDER_WAVEFRONT_ONLY if wavefront_only else 0))
        st = RenderStats()
        _check(LIB.lrt_render(self._h, C.byref(prm), _p(sum_rgb, C.c_double),
                              _p(sumsq_rgb, C.c_double) if sumsq else None, 0, C.byref(st)))
        return sum_rgb, sumsq_rgb, st

    def render_features_device(self, width, height, spp, max_depth, rr_depth, sum_ptr, aov_ptr, sumsq_ptr=0, seed=1,
                               spp_offset=0, rows=None):
        """lrt_render_features on device buffers: radiance sums [H, W, 3] float64
        and the first-hit features [H, W, 7] float32 of the same camera rays."""
        rb, re = rows if rows is not None else (0, 0)
        prm = RenderParams(width, height, spp, spp_offset, max_depth, rr_depth, seed, rb, re, 0)
        st = RenderStats()
        _check(LIB.lrt_render_features(self._h, C.byref(prm), C.c_void_p(sum_ptr),
                                       C.c_void_p(sumsq_ptr) if sumsq_ptr else None, C.c_void_p(aov_ptr), 1,
                                       C.byref(st)))
        return st

    def render_device(self, width, height, spp, max_depth, rr_depth, sum_ptr, sumsq_ptr=0, seed=1, spp_offset=0,
                      count=False, wavefront_only=False):
        prm = RenderParams(width, height, spp, spp_offset, max_depth, rr_depth, seed, 0, 0,
                           (RENDER_COUNT if count else 0) | (RENDER_WAVEFRONT_ONLY if wavefront_only else 0))
        st = RenderStats()
        _check(LIB.lrt_render(self._h, C.byref(prm), C.c_void_p(sum_ptr), C.c_void_p(sumsq_ptr) if sumsq_ptr else None,
                              1, C.byref(st)))
        return st

    def render_aov(self, width, height, spp, seed=1, spp_offset=0, rows=None):
        """First-hit (albedo, normal, depth) averaged over the frame's camera
        rays: [H, W, 7] float32 (rows bottom-up like render)."""
        aov = np.zeros((height, width, 7), np.float32)
        rb, re = rows if rows is not None else (0, 0)
        prm = RenderParams(width, height, spp, spp_offset, 1, 0, seed, rb, re, 0)
        _check(LIB.lrt_render_aov(self._h, C.byref(prm), _p(aov), 0))
        return aov

    def render_aov_device(self, width, height, spp, aov_ptr, seed=1, spp_offset=0, rows=None):
        rb, re = rows if rows is not None else (0, 0)
        prm = RenderParams(width, height, spp, spp_offset, 1, 0, seed, rb, re, 0)
        _check(LIB.lrt_render_aov(self._h, C.byref(prm), C.c_void_p(aov_ptr), 1))

    # ---- photon mapping (photon_integrator.h:15-123) ----
    def photon_build(self, requested, max_depth, rr_depth, seed=1, radius=0.05, k=120, kind=PHOTON_SURFACE,
                     tree=PHOTON_TREE_DEVICE):
        """Traces particles until `requested` photons are stored and builds map
        `kind` (photon_gather.h:15-101, photon_map.h:26-114); tree: the device
        median tree (default) or the reference's nth_element tree (host)."""
        prm = PhotonParams(kind, requested, max_depth, rr_depth, seed, radius, k, tree)
        st = PhotonStats()
        _check(LIB.lrt_photon_build(self._h, C.byref(prm), C.byref(st)))
        return st

    def photon_export(self, kind=PHOTON_SURFACE):
        """Map `kind`'s photons in particle order: (photons [n, 12] float32 =
        position, direction, normal, power; depth_particle [n, 2] uint32; shot)."""
        n, shot = C.c_uint64(), C.c_uint64()
        _check(LIB.lrt_photon_export(self._h, C.c_int32(kind), None, None, C.c_uint64(0), C.byref(n), C.byref(shot)))
        ph = np.zeros((n.value, 12), np.float32)
        dp = np.zeros((n.value, 2), np.uint32)
        _check(LIB.lrt_photon_export(self._h, C.c_int32(kind), _p(ph), _p(dp, C.c_uint32), C.c_uint64(n.value),
                                     C.byref(n), C.byref(shot)))
        return ph, dp, shot.value

    def photon_map_tree(self, kind=PHOTON_SURFACE):
        """Map `kind`'s kd tree (lrt_photon_map_tree): nodes [m, 4] uint32
        (interior {split bits, axis, left, right}, leaf {0, 3, first, count},
        pre-order) and order [n] (particle-order index of each tree position)."""
        m, n = C.c_uint64(), C.c_uint64()
        _check(LIB.lrt_photon_map_tree(self._h, C.c_int32(kind), None, C.c_uint64(0), None, C.c_uint64(0),
                                       C.byref(m), C.byref(n)))
        nodes = np.zeros((m.value, 4), np.uint32)
        order = np.zeros(n.value, np.uint32)
        _check(LIB.lrt_photon_map_tree(self._h, C.c_int32(kind), _p(nodes, C.c_uint32), C.c_uint64(m.value),
                                       _p(order, C.c_uint32), C.c_uint64(n.value), C.byref(m), C.byref(n)))
        return nodes, order

    def photon_map_load(self, photons, depth_particle, scale, radius, k=120, kind=PHOTON_SURFACE):
        ph = np.ascontiguousarray(photons, np.float32).reshape(-1, 12)
        dp = np.ascontiguousarray(depth_particle, np.uint32).reshape(-1, 2)
        _check(LIB.lrt_photon_map_load(self._h, C.c_int32(kind), _p(ph), _p(dp, C.c_uint32), C.c_uint64(ph.shape[0]),
                                       C.c_float(scale), C.c_float(radius), C.c_uint32(k)))

    def photon_estimate(self, queries, max_depth, kind=PHOTON_SURFACE):
        """photon_map::estimate_irradiance of map `kind` at (point, normal) rows [n, 6]."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 6)
        out = np.zeros((q.shape[0], 3), np.float32)
        _check(LIB.lrt_photon_estimate(self._h, C.c_int32(kind), _p(q), C.c_uint32(q.shape[0]), C.c_int(max_depth),
                                       _p(out)))
        return out

    def render_photon(self, width, height, spp, max_depth=128, seed=1, spp_offset=0, rows=None, sumsq=True,
                      count=False):
        """photon_integrator::radiance per (pixel, sample) with the scene's maps
        (photon_integrator.h:39-123); returns (sum, sumsq, stats) like render.
        count: stats also count the gathers' work (nodes, photons tested,
        photons summed)."""
        sum_rgb = np.zeros((height, width, 3), np.float64)
        sumsq_rgb = np.zeros((height, width, 3), np.float64) if sumsq else None
        rb, re = rows if rows is not None else (0, 0)
        prm = RenderParams(width, height, spp, spp_offset, max_depth, 0, seed, rb, re, RENDER_COUNT if count else 0)
        st = RenderStats()
        _check(LIB.lrt_render_photon(self._h, C.byref(prm), _p(sum_rgb, C.c_double),
                                     _p(sumsq_rgb, C.c_double) if sumsq else None, 0, C.byref(st)))
        return sum_rgb, sumsq_rgb, st


class PhotonIntegrator:
    """photon_integrator (photon_integrator.h:13-123) on the GPU: preprocess()
    builds the global and caustic maps with the reference's constants, and
    render() runs the eye pass."""
    MAX_DEPTH, RR_DEPTH, LOOKUP = 128, 5, 120
    GLOBAL_PHOTONS = CAUSTIC_PHOTONS = 250000

    def __init__(self, global_photons=GLOBAL_PHOTONS, caustic_photons=CAUSTIC_PHOTONS, lookup=LOOKUP,
                 max_depth=MAX_DEPTH, seed=1):
        self.global_photons, self.caustic_photons = global_photons, caustic_photons
        self.lookup, self.max_depth, self.seed = lookup, max_depth, seed

    def preprocess(self, scene):
        info = scene.info
        r = info.bounding_radius
        caustic = self.caustic_photons if info.has_delta_bsdf else 0
        self.stats = [scene.photon_build(self.global_photons, self.max_depth - 1, self.RR_DEPTH, self.seed,
                                         0.05 * r, self.lookup, PHOTON_SURFACE)]
        if caustic:
            self.stats.append(scene.photon_build(caustic, self.max_depth - 1, self.RR_DEPTH, self.seed,
                                                 0.0125 * r, self.lookup, PHOTON_CAUSTIC))
        else:
            scene.photon_map_load(np.zeros((0, 12), np.float32), np.zeros((0, 2), np.uint32), 1.0,
                                  0.0125 * r, self.lookup, PHOTON_CAUSTIC)
        return self.stats

    def render(self, scene, width, height, spp, spp_offset=0):
        return scene.render_photon(width, height, spp, self.max_depth, self.seed, spp_offset)


class Denoiser:
    """The BF16 U-Net denoiser (lrt_denoiser_*): forward on host or device
    arrays, loss + gradient, Adam steps, parameter get/set."""

    def __init__(self, device=0, seed=1):
        self._h = None
        h = C.c_void_p()
        _check(LIB.lrt_denoiser_create(C.c_int(device), C.c_uint64(seed), C.byref(h)))
        self._h = h
        n = C.c_uint64()
        _check(LIB.lrt_denoiser_param_count(self._h, C.byref(n)))
        self.n_params = n.value

    def close(self):
        if self._h and LIB is not None:
            LIB.lrt_denoiser_release(self._h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(LIB.lrt_denoiser_params(self._h, _p(p), None))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(LIB.lrt_denoiser_params(self._h, None, _p(p)))

    def forward(self, rad, aux):
        """rad [B,H,W,3], aux [B,H,W,7] float32 -> denoised [B,H,W,3]."""
        rad = np.ascontiguousarray(rad, np.float32)
        aux = np.ascontiguousarray(aux, np.float32)
        if rad.ndim == 3:
            rad, aux = rad[None], aux[None]
        B, H, W, _ = rad.shape
        out = np.zeros((B, H, W, 3), np.float32)
        _check(LIB.lrt_denoiser_forward(self._h, _p(rad), _p(aux), C.c_uint32(B), C.c_uint32(H), C.c_uint32(W), _p(out), 0))
        return out

    def timing(self, enable=True):
        """Per-layer milliseconds summed over the forwards since the previous
        call (10 convolution layers) and the number of those forwards; timing
        stays on for the following forwards if `enable`."""
        ms = (C.c_double * 10)()
        n = C.c_uint64()
        _check(LIB.lrt_denoiser_timing(self._h, C.c_int(int(enable)), ms, C.byref(n)))
        return list(ms), n.value

    def forward_device(self, rad_ptr, aux_ptr, B, H, W, out_ptr):
        _check(LIB.lrt_denoiser_forward(self._h, C.c_void_p(rad_ptr), C.c_void_p(aux_ptr), C.c_uint32(B), C.c_uint32(H),
                                        C.c_uint32(W), C.c_void_p(out_ptr), 1))

    def loss_grad(self, rad, aux, target, grads=True):
        rad = np.ascontiguousarray(rad, np.float32)
        aux = np.ascontiguousarray(aux, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        B, H, W, _ = rad.shape
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32) if grads else None
        _check(LIB.lrt_denoiser_loss_grad(self._h, _p(rad), _p(aux), _p(target), C.c_uint32(B), C.c_uint32(H),
                                          C.c_uint32(W), C.byref(loss), _p(g), 0))
        return loss.value, g

    def loss_grad_device(self, rad_ptr, aux_ptr, tgt_ptr, B, H, W):
        loss = C.c_float()
        _check(LIB.lrt_denoiser_loss_grad(self._h, C.c_void_p(rad_ptr), C.c_void_p(aux_ptr), C.c_void_p(tgt_ptr),
                                          C.c_uint32(B), C.c_uint32(H), C.c_uint32(W), C.byref(loss), None, 1))
        return loss.value

    def adam_step(self, lr):
        _check(LIB.lrt_denoiser_adam_step(self._h, C.c_float(lr)))


def conv3x3(op, w, x=None, b=None, g=None, relu=False, device=0):
    """One 3x3 convolution layer through lrt_conv3x3 (op 0 forward, 1 data
    gradient, 2 weight + bias gradient); arrays NHWC float32, w [9][cin][cout]."""
    w = np.ascontiguousarray(w, np.float32)
    cin, cout = w.shape[1], w.shape[2]
    ref = x if x is not None else g
    B, H, W = ref.shape[:3]
    x = None if x is None else np.ascontiguousarray(x, np.float32)
    g = None if g is None else np.ascontiguousarray(g, np.float32)
    b = None if b is None else np.ascontiguousarray(b, np.float32)
    if op == 0:
        out = np.zeros((B, H, W, cout), np.float32)
    elif op == 1:
        out = np.zeros((B, H, W, cin), np.float32)
    else:
        out = np.zeros(9 * cin * cout + cout, np.float32)
    _check(LIB.lrt_conv3x3(C.c_int(device), C.c_int(op), _p(x), _p(w), _p(b), _p(g), C.c_uint32(B), C.c_uint32(H),
                           C.c_uint32(W), C.c_uint32(cin), C.c_uint32(cout), C.c_int(int(relu)), _p(out)))
    if op == 2:
        return out[:9 * cin * cout].reshape(9, cin, cout), out[9 * cin * cout:]
    return out


def _mats_arg(materials):
    arr = (Material * len(materials))(*materials)
    return arr, len(materials)


def bsdf_eval_pdf(materials, tables, mat, wi, wo, front, measure=MEASURE_SOLID_ANGLE, mode=MODE_RADIANCE,
                  component=-1, type_mask=0xFFFFFFFF, device=0):
    marr, nm = _mats_arg(materials)
    tb = np.ascontiguousarray(tables if tables is not None else np.zeros((0, 101)), np.float32)
    mat = np.ascontiguousarray(mat, np.int32)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    front = np.ascontiguousarray(front, np.uint8)
    n = mat.size
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    _check(LIB.lrt_bsdf_eval_pdf(C.c_int(device), marr, C.c_uint32(nm), _p(tb), C.c_uint32(tb.size // 101),
                                 _p(mat, C.c_int32), _p(wi), _p(wo), _p(front, C.c_uint8), C.c_uint32(n),
                                 C.c_int(measure), C.c_int(mode), C.c_int(component), C.c_uint32(type_mask),
                                 _p(f), _p(pdf)))
    return f, pdf


def bsdf_param_grad(materials, tables, mat, wi, wo, front, which, mode=MODE_RADIANCE, device=0):
    """lrt_bsdf_param_grad: d (f*cos) / d alpha (which 0) or d interior IOR
    (which 1) of BSDF tuples, [n, 3]."""
    marr, nm = _mats_arg(materials)
    tb = np.ascontiguousarray(tables if tables is not None else np.zeros((0, 101)), np.float32)
    mat = np.ascontiguousarray(mat, np.int32)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    front = np.ascontiguousarray(front, np.uint8)
    n = mat.size
    df = np.zeros((n, 3), np.float32)
    _check(LIB.lrt_bsdf_param_grad(C.c_int(device), marr, C.c_uint32(nm), _p(tb), C.c_uint32(tb.size // 101),
                                   _p(mat, C.c_int32), _p(wi), _p(wo), _p(front, C.c_uint8), C.c_uint32(n),
                                   C.c_int(mode), C.c_int(which), _p(df)))
    return df


def bsdf_sample(materials, tables, mat, wi, front, u, lobe_u, mode=MODE_RADIANCE, component=-1,
                type_mask=0xFFFFFFFF, device=0):
    marr, nm = _mats_arg(materials)
    tb = np.ascontiguousarray(tables if tables is not None else np.zeros((0, 101)), np.float32)
    mat = np.ascontiguousarray(mat, np.int32)
    wi = np.ascontiguousarray(wi, np.float32)
    front = np.ascontiguousarray(front, np.uint8)
    u = np.ascontiguousarray(u, np.float32)
    lobe_u = np.ascontiguousarray(lobe_u, np.float32)
    n = mat.size
    w = np.zeros((n, 3), np.float32)
    wo = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    eta = np.zeros(n, np.float32)
    st = np.zeros(n, np.uint32)
    sc = np.zeros(n, np.int32)
    _check(LIB.lrt_bsdf_sample(C.c_int(device), marr, C.c_uint32(nm), _p(tb), C.c_uint32(tb.size // 101),
                               _p(mat, C.c_int32), _p(wi), _p(front, C.c_uint8), _p(u), _p(lobe_u), C.c_uint32(n),
                               C.c_int(mode), C.c_int(component), C.c_uint32(type_mask), _p(w), _p(wo), _p(pdf),
                               _p(eta), _p(st, C.c_uint32), _p(sc, C.c_int32)))
    return dict(weight=w, wo=wo, pdf=pdf, eta=eta, sampled_type=st, sampled_component=sc)


def bsdf_flags(materials, index):
    marr, nm = _mats_arg(materials)
    f, c = C.c_uint32(), C.c_int32()
    _check(LIB.lrt_bsdf_flags(marr, C.c_uint32(nm), C.c_int32(index), C.byref(f), C.byref(c)))
    return f.value, c.value


def sky_eval(sky, dirs, device=0):
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = dirs.shape[0]
    rgb = np.zeros((n, 3), np.float32)
    _check(LIB.lrt_sky_eval(C.c_int(device), C.byref(sky), _p(dirs), C.c_uint32(n), _p(rgb)))
    return rgb


def material(type, **kw):
    """lrt_material with the reference constructors' defaults."""
    m = Material()
    m.type = type
    m.nested = -1
    m.table_ext = m.table_int = -1
    m.int_ior, m.ext_ior = 1.5, 1.0
    m.thickness = 1.0
    for i in range(3):
        m.specular_reflectance[i] = 1.0
        m.specular_transmittance[i] = 1.0
        m.diffuse[i] = 0.5
    for k, v in kw.items():
        if isinstance(v, (list, tuple, np.ndarray)):
            arr = getattr(m, k)
            for i in range(3):
                arr[i] = float(v[i])
        else:
            setattr(m, k, v)
    return m
