"""ctypes access to the oracle (oracle/liborc.so, the CPU restatement) and, where
it was built in this container, to oracle/_ref/libref.so (the reference's own
sources compiled through oracle/ref_harness.cpp).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORC_PATH = REPO / "oracle" / "liborc.so"
REF_PATH = REPO / "oracle" / "_ref" / "libref.so"
REF_ROOT = Path("/root/reference")


def _p(a, ctype=C.c_float):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def _load(path):
    lib = C.CDLL(str(path))
    for name in ("orc_scene_create", "ref_scene_create"):
        if hasattr(lib, name):
            getattr(lib, name).restype = C.c_void_p
    for name in ("orc_trace_particles", "orc_gather_photons", "ref_trace_particles"):
        if hasattr(lib, name):
            getattr(lib, name).restype = C.c_int64
    return lib


ORC = _load(ORC_PATH) if ORC_PATH.exists() else None
REF = _load(REF_PATH) if REF_PATH.exists() and REF_ROOT.exists() else None


class _Impl:
    """Common calls of the oracle (prefix orc_) and the reference harness (ref_)."""

    def __init__(self, lib, prefix):
        self.lib, self.pre = lib, prefix

    def f(self, name):
        return getattr(self.lib, self.pre + name)

    def kd_build(self, boxes):
        boxes = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        n = boxes.shape[0]
        nn, nr, dl = C.c_uint32(), C.c_uint32(), C.c_int32()
        z = None
        self.f("kd_build_export")(_p(boxes), C.c_uint32(n), C.byref(nn), C.byref(nr), C.byref(dl), z, z, z, z, z, z, z, z)
        o = dict(leaf=np.zeros(nn.value, np.uint8), axis=np.zeros(nn.value, np.int32), split=np.zeros(nn.value, np.float32),
                 left=np.zeros(nn.value, np.int32), right=np.zeros(nn.value, np.int32),
                 first=np.zeros(nn.value, np.int32), count=np.zeros(nn.value, np.int32),
                 refs=np.zeros(max(nr.value, 1), np.uint32))
        self.f("kd_build_export")(_p(boxes), C.c_uint32(n), C.byref(nn), C.byref(nr), C.byref(dl),
                                  _p(o["leaf"], C.c_uint8), _p(o["axis"], C.c_int32), _p(o["split"]),
                                  _p(o["left"], C.c_int32), _p(o["right"], C.c_int32), _p(o["first"], C.c_int32),
                                  _p(o["count"], C.c_int32), _p(o["refs"], C.c_uint32))
        o["refs"] = o["refs"][: nr.value]
        o["depth_limit"] = dl.value
        return o

    def scene(self, desc_ptr):
        h = self.f("scene_create")(desc_ptr)
        if not h:
            raise RuntimeError(f"{self.pre}scene_create failed")
        return _SceneHandle(self, h)

    def bsdf_eval_pdf(self, materials, tables, mat, wi, wo, front, measure=1, mode=0, component=-1,
                      type_mask=0xFFFFFFFF):
        from linrt import Material
        marr = (Material * len(materials))(*materials)
        tb = np.ascontiguousarray(tables if tables is not None else np.zeros((1, 101)), np.float32)
        mat = np.ascontiguousarray(mat, np.int32)
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        front = np.ascontiguousarray(front, np.uint8)
        n = mat.size
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        rc = self.f("bsdf_eval_pdf")(marr, C.c_uint32(len(materials)), _p(tb), C.c_uint32(tb.size // 101),
                                     _p(mat, C.c_int32), _p(wi), _p(wo), _p(front, C.c_uint8), C.c_uint32(n),
                                     C.c_int(measure), C.c_int(mode), C.c_int(component), C.c_uint32(type_mask),
                                     _p(f), _p(pdf))
        assert rc == 0
        return f, pdf

    def bsdf_sample(self, materials, tables, mat, wi, front, u, lobe_u, mode=0, component=-1, type_mask=0xFFFFFFFF):
        from linrt import Material
        marr = (Material * len(materials))(*materials)
        tb = np.ascontiguousarray(tables if tables is not None else np.zeros((1, 101)), np.float32)
        mat = np.ascontiguousarray(mat, np.int32)
        wi = np.ascontiguousarray(wi, np.float32)
        front = np.ascontiguousarray(front, np.uint8)
        u = np.ascontiguousarray(u, np.float32)
        lobe_u = np.ascontiguousarray(lobe_u, np.float32)
        n = mat.size
        w, wo = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        pdf, eta = np.zeros(n, np.float32), np.zeros(n, np.float32)
        st, sc = np.zeros(n, np.uint32), np.zeros(n, np.int32)
        rc = self.f("bsdf_sample")(marr, C.c_uint32(len(materials)), _p(tb), C.c_uint32(tb.size // 101),
                                   _p(mat, C.c_int32), _p(wi), _p(front, C.c_uint8), _p(u), _p(lobe_u), C.c_uint32(n),
                                   C.c_int(mode), C.c_int(component), C.c_uint32(type_mask), _p(w), _p(wo), _p(pdf),
                                   _p(eta), _p(st, C.c_uint32), _p(sc, C.c_int32))
        assert rc == 0
        return dict(weight=w, wo=wo, pdf=pdf, eta=eta, sampled_type=st, sampled_component=sc)

    def bsdf_flags(self, materials, index):
        from linrt import Material
        marr = (Material * len(materials))(*materials)
        f, c = C.c_uint32(), C.c_int32()
        assert self.f("bsdf_flags")(marr, C.c_uint32(len(materials)), C.c_int32(index), C.byref(f), C.byref(c)) == 0
        return f.value, c.value


class _SceneHandle:
    def __init__(self, impl, h):
        self.impl, self.h = impl, C.c_void_p(h)

    def close(self):
        if self.h:
            self.impl.f("scene_free")(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def trace(self, rays, threads=8, counts=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(8, -1)
        n = rays.shape[1]
        op = np.zeros((n, 2), np.int32)
        t = np.zeros(n, np.float32)
        if self.impl.pre == "ref_":
            self.impl.f("scene_trace")(self.h, _p(rays), C.c_uint32(n), _p(op, C.c_int32), _p(t), None, C.c_int(threads))
            return op, t
        uv = np.zeros((n, 2), np.float32)
        cnt = np.zeros((n, 3), np.uint32) if counts else None
        self.impl.f("scene_trace")(self.h, _p(rays), C.c_uint32(n), _p(op, C.c_int32), _p(t), _p(uv),
                                   _p(cnt, C.c_uint32), C.c_int(threads))
        return (op, t, uv, cnt) if counts else (op, t)

    def trace_brute(self, rays, threads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(8, -1)
        n = rays.shape[1]
        op = np.zeros((n, 2), np.int32)
        t = np.zeros(n, np.float32)
        self.impl.f("scene_trace_brute")(self.h, _p(rays), C.c_uint32(n), _p(op, C.c_int32), _p(t), C.c_int(threads))
        return op, t

    def occluded(self, rays, ignore=None, threads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(8, -1)
        n = rays.shape[1]
        out = np.zeros(n, np.uint8)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        self.impl.f("scene_occluded")(self.h, _p(rays), C.c_uint32(n), _p(ig, C.c_int32), _p(out, C.c_uint8),
                                      C.c_int(threads))
        return out.astype(bool)

    def render(self, W, H, spp, max_depth, rr_depth, seed=1, mode=0, threads=0, rows=None, spp_offset=0,
               per_sample=False, fsum=False):
        """Oracle frame loop; returns dict(sum, sumsq, per_sample, fsum, ext, shadow)."""
        rb, re = rows if rows is not None else (0, 0)
        s = np.zeros((H, W, 3), np.float64)
        q = np.zeros((H, W, 3), np.float64)
        ps = np.zeros((H * W * spp * 3,), np.float32) if per_sample else None
        fs = np.zeros((H, W, 3), np.float32) if fsum else None
        st = (C.c_uint64 * 2)()
        if self.impl.pre == "ref_":
            self.impl.f("render")(self.h, C.c_int(W), C.c_int(H), C.c_int(spp), C.c_int(rb), C.c_int(re), _p(ps),
                                  _p(fs), C.c_int(threads or 8))
            return dict(per_sample=None if ps is None else ps.reshape(H, W, spp, 3), fsum=fs)
        self.impl.f("render")(self.h, C.c_int(W), C.c_int(H), C.c_int(spp), C.c_int(spp_offset), C.c_int(max_depth),
                              C.c_int(rr_depth), C.c_uint64(seed), C.c_int(mode), C.c_int(threads), C.c_int(rb),
                              C.c_int(re), _p(s, C.c_double), _p(q, C.c_double), _p(ps), _p(fs), st)
        return dict(sum=s, sumsq=q, per_sample=None if ps is None else ps.reshape(H, W, spp, 3), fsum=fs,
                    ext=st[0], shadow=st[1])


    # ---- photon mapping ----
    def trace_particles(self, first, count, max_depth, rr_depth, kind=0, seed=1, mode=0):
        """Particles [first, first+count) traced one after another
        (photon_tracer.h:17-88): (photons [n, 12], depth_particle [n, 2])."""
        if self.impl.pre == "ref_":
            args = (self.h, C.c_int(kind), C.c_uint32(first), C.c_uint32(count), C.c_int(max_depth), C.c_int(rr_depth))
        else:
            args = (self.h, C.c_int(kind), C.c_uint32(first), C.c_uint32(count), C.c_int(max_depth), C.c_int(rr_depth),
                    C.c_uint64(seed), C.c_int(mode))
        f = self.impl.f("trace_particles")
        n = f(*args, None, None, C.c_int64(0))
        ph = np.zeros((max(n, 1), 12), np.float32)
        dp = np.zeros((max(n, 1), 2), np.uint32)
        f(*args, _p(ph), _p(dp, C.c_uint32), C.c_int64(n))
        return ph[:n], dp[:n]

    def gather_photons(self, requested, max_depth, rr_depth, kind=0, seed=1):
        """Oracle quota gather in particle order: (photons, depth_particle, shot)."""
        f = self.impl.f("gather_photons")
        shot = C.c_uint64()
        n = f(self.h, C.c_int(kind), C.c_uint32(requested), C.c_int(max_depth), C.c_int(rr_depth), C.c_uint64(seed),
              None, None, C.byref(shot))
        ph = np.zeros((max(n, 1), 12), np.float32)
        dp = np.zeros((max(n, 1), 2), np.uint32)
        f(self.h, C.c_int(kind), C.c_uint32(requested), C.c_int(max_depth), C.c_int(rr_depth), C.c_uint64(seed),
          _p(ph), _p(dp, C.c_uint32), C.byref(shot))
        return ph[:n], dp[:n], shot.value

    def render_photon(self, W, H, spp, maps, max_depth=128, k=120, seed=1, mode=0, threads=8, rows=None, tree=0):
        """photon_integrator::radiance per (pixel, sample) over maps built from
        maps = [(photons, depth_particle, scale, radius) or None] * 2 (global,
        caustic); tree (oracle only): 0 the reference's tree, 1 the library's
        device tree.  Returns per-sample radiance [H, W, spp, 3] (float32)."""
        rb, re = rows if rows is not None else (0, 0)
        ps = np.zeros(H * W * spp * 3, np.float32)
        args = []
        keep = []
        for m in maps:
            if m is None:
                args += [None, None, C.c_int64(0), C.c_float(1.0), C.c_float(1.0)]
            else:
                ph = np.ascontiguousarray(m[0], np.float32).reshape(-1, 12)
                dp = np.ascontiguousarray(m[1], np.uint32).reshape(-1, 2)
                keep += [ph, dp]
                args += [_p(ph), _p(dp, C.c_uint32), C.c_int64(ph.shape[0]), C.c_float(m[2]), C.c_float(m[3])]
        if self.impl.pre == "ref_":
            assert max_depth == 128 and k == 120, "the reference's photon integrator fixes depth 128 and k 120"
            self.impl.f("render_photon")(self.h, C.c_int(W), C.c_int(H), C.c_int(spp), *args, C.c_int(rb), C.c_int(re),
                                         _p(ps), C.c_int(threads))
        else:
            self.impl.f("render_photon")(self.h, C.c_int(W), C.c_int(H), C.c_int(spp), C.c_int(max_depth),
                                         C.c_uint64(seed), C.c_int(mode), *args, C.c_int(k), C.c_int(threads),
                                         C.c_int(rb), C.c_int(re), None, None, _p(ps), C.c_int(tree))
        return ps.reshape(H, W, spp, 3)

    def bounds(self):
        """(bounding sphere radius, has a delta BSDF) of the reference scene."""
        r, d = C.c_float(), C.c_int()
        self.impl.f("scene_bounds")(self.h, C.byref(r), C.byref(d))
        return r.value, bool(d.value)


def photon_estimate(impl, photons, depth_particle, scale, queries, radius, max_depth, k, threads=8, tree=0):
    """photon_map build + estimate_irradiance through the oracle or the reference;
    tree (oracle only): 0 the reference's tree, 1 the library's device tree."""
    ph = np.ascontiguousarray(photons, np.float32).reshape(-1, 12)
    dp = np.ascontiguousarray(depth_particle, np.uint32).reshape(-1, 2)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 6)
    out = np.zeros((q.shape[0], 3), np.float32)
    if impl.pre == "ref_":
        rc = impl.f("photon_estimate")(_p(ph), _p(dp, C.c_uint32), C.c_int64(ph.shape[0]), C.c_float(scale), _p(q),
                                       C.c_uint32(q.shape[0]), C.c_float(radius), C.c_int(max_depth), C.c_int(k),
                                       _p(out))
    else:
        rc = impl.f("photon_estimate")(_p(ph), _p(dp, C.c_uint32), C.c_int64(ph.shape[0]), C.c_float(scale), _p(q),
                                       C.c_uint32(q.shape[0]), C.c_float(radius), C.c_int(max_depth), C.c_int(k),
                                       _p(out), C.c_int(threads), C.c_int(tree))
    assert rc == 0
    return out


def photon_tree(xyz, rule="device"):
    """The oracle's photon kd tree over positions [n, 3] (input order): rule
    "device" (the library's median tree) or "reference" (photon_map's):
    (nodes [m, 4] uint32, order [n] uint32)."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    f = ORC.orc_photon_tree if rule == "device" else ORC.orc_photon_tree_reference
    f.restype = C.c_int64
    m = f(_p(xyz), C.c_int64(xyz.shape[0]), None, None)
    nodes = np.zeros((m, 4), np.uint32)
    order = np.zeros(xyz.shape[0], np.uint32)
    f(_p(xyz), C.c_int64(xyz.shape[0]), _p(nodes, C.c_uint32), _p(order, C.c_uint32))
    return nodes, order


oracle = _Impl(ORC, "orc_") if ORC is not None else None
reference = _Impl(REF, "ref_") if REF is not None else None


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    ORC.orc_philox(c, k, o)
    return list(o)


def sky_eval(sky, dirs):
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    rgb = np.zeros_like(dirs)
    ORC.orc_sky_eval(C.byref(sky), _p(dirs), C.c_uint32(dirs.shape[0]), _p(rgb))
    return rgb
