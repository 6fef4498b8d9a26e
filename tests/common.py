"""Seeded inputs shared by the tests: incoherent rays, BSDF tuples, materials."""
from __future__ import annotations

import math

import numpy as np

import linrt


def tri_normals(pos, idx):
    p0, p1, p2 = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    g = np.cross(p1 - p0, p2 - p0)
    a = np.linalg.norm(g, axis=1)
    return g, a


def surface_rays(pos, idx, n, seed, tmin=0.001, offset=1e-4):
    """Area-uniform origins on the mesh offset along the geometric normal,
    cosine-distributed directions around it (incoherent secondary rays)."""
    rng = np.random.default_rng(seed)
    g, a = tri_normals(pos.astype(np.float64), idx)
    ok = a > 0
    cdf = np.cumsum(np.where(ok, a, 0.0))
    t = np.searchsorted(cdf, rng.random(n) * cdf[-1], side="right").clip(0, len(a) - 1)
    u, v = rng.random(n), rng.random(n)
    su = np.sqrt(u)
    b0, b1 = 1 - su, v * su
    p = (b0[:, None] * pos[idx[t, 0]] + b1[:, None] * pos[idx[t, 1]] + (1 - b0 - b1)[:, None] * pos[idx[t, 2]])
    nrm = g[t] / a[t][:, None]
    o = p + offset * nrm
    # cosine hemisphere around nrm
    r1, r2 = rng.random(n), rng.random(n)
    ct, st, ph = np.sqrt(r1), np.sqrt(1 - r1), 2 * math.pi * r2
    helper = np.where(np.abs(nrm[:, :1]) > 0.9, [[0, 1, 0]], [[1, 0, 0]])
    tt = np.cross(nrm, helper)
    tt /= np.linalg.norm(tt, axis=1, keepdims=True)
    ss = np.cross(nrm, tt)
    d = ss * (np.cos(ph) * st)[:, None] + tt * (np.sin(ph) * st)[:, None] + nrm * ct[:, None]
    return pack_rays(o, d, tmin, np.inf)


def random_rays(lo, hi, n, seed, tmin=0.0):
    rng = np.random.default_rng(seed)
    o = lo + (hi - lo) * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return pack_rays(o, d, tmin, np.inf)


def pack_rays(o, d, tmin, tmax):
    n = o.shape[0]
    r = np.empty((8, n), np.float32)
    r[0:3] = o.T
    r[3:6] = d.T
    r[6] = tmin
    r[7] = tmax
    return r


def hit_parity(op_a, t_a, op_b, t_b, rel=1e-6):
    """Fraction of rays with the same (object, primitive) and t within rel; the
    list of mismatching ray indices."""
    same = np.all(op_a == op_b, axis=1)
    hit = op_a[:, 0] >= 0
    tol = rel * np.maximum(np.abs(t_b), 1e-30)
    t_ok = np.where(hit, np.abs(t_a - t_b) <= tol, True)
    good = same & t_ok
    return good.mean(), np.nonzero(~good)[0]


def material_zoo(tables):
    """One material of every reference BSDF class (plus the smooth conductor and
    two coatings), with the parameters the displaced-sphere scene uses."""
    m = linrt.material
    cu_eta, cu_k = (0.200438, 0.924033, 1.102212), (3.912948, 2.452847, 2.142188)
    mats = [
        m(linrt.BSDF_DIFFUSE, diffuse=(0.3, 0.6, 0.6)),
        m(linrt.BSDF_CONDUCTOR, alpha=0.1, cond_eta=cu_eta, cond_k=cu_k),
        m(linrt.BSDF_CONDUCTOR, alpha=0.0, cond_eta=cu_eta, cond_k=cu_k),
        m(linrt.BSDF_DIELECTRIC, int_ior=1.5, ext_ior=1.0),
        m(linrt.BSDF_ROUGH_DIELECTRIC, int_ior=1.5, ext_ior=1.0, alpha=0.2),
        m(linrt.BSDF_PLASTIC, int_ior=1.49, ext_ior=1.000277, diffuse=(0.8, 0.8, 0.5)),
        m(linrt.BSDF_ROUGH_PLASTIC, int_ior=1.49, ext_ior=1.000277, alpha=0.3, diffuse=(0.8, 0.1, 0.05),
          table_ext=0, table_int=1),
        m(linrt.BSDF_COATING, int_ior=1.7, ext_ior=1.000277, nested=5, specular_reflectance=(0.3, 0.6, 0.6)),
        m(linrt.BSDF_ROUGH_COATING, int_ior=1.7, ext_ior=1.000277, alpha=0.25, nested=9, table_ext=2),
        m(linrt.BSDF_ROUGH_DIELECTRIC, int_ior=1.5, ext_ior=1.0, alpha=0.05),
    ]
    mats[5].fdr_int = linrt.fresnel_diffuse_reflectance(np.float32(1.0) / (np.float32(1.49) / np.float32(1.000277)))
    return mats


NAMES = ["diffuse", "rough_conductor", "conductor", "dielectric", "rough_dielectric", "plastic", "rough_plastic",
         "coating", "rough_coating", "rough_dielectric_a005"]


def bsdf_tuples(n, seed, n_mats):
    """Seeded (material, wi, wo, front_face, u, lobe_u) tuples: directions
    uniform on the sphere with wi mostly in the upper hemisphere, plus exact
    mirror and refraction partners so the delta lobes are exercised."""
    rng = np.random.default_rng(seed)
    mat = rng.integers(0, n_mats, n).astype(np.int32)

    def sphere(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    wi = sphere(n)
    flip = rng.random(n) < 0.9
    wi[:, 2] = np.where(flip, np.abs(wi[:, 2]), wi[:, 2])
    wo = sphere(n)
    mirror = rng.random(n) < 0.1
    wo[mirror] = wi[mirror] * np.array([-1, -1, 1])
    front = (rng.random(n) < 0.8).astype(np.uint8)
    u = rng.random((n, 2))
    lobe = rng.random(n)
    return (mat, wi.astype(np.float32), wo.astype(np.float32), front, u.astype(np.float32), lobe.astype(np.float32))


class SphereRoom:
    """The box room of the 256x256 workload with four analytic spheres (smooth
    dielectric, rough conductor, smooth plastic, smooth coating over diffuse)
    under its ceiling light: every photon-integrator branch (specular lobes,
    glossy recursion, pure-diffuse gathers) with one material per shape, as the
    reference's scene graph requires."""

    def __init__(self, seed=1):
        self.gen = linrt.box_icosphere(seed=seed, subdivisions=1)
        d = self.gen.desc
        m = linrt.material
        cu_eta, cu_k = (0.200438, 0.924033, 1.102212), (3.912948, 2.452847, 2.142188)
        mats = [m(linrt.BSDF_DIFFUSE, diffuse=(0.62, 0.55, 0.45)),
                m(linrt.BSDF_DIELECTRIC, int_ior=1.5, ext_ior=1.0),
                m(linrt.BSDF_CONDUCTOR, alpha=0.15, cond_eta=cu_eta, cond_k=cu_k),
                m(linrt.BSDF_PLASTIC, int_ior=1.49, ext_ior=1.000277, diffuse=(0.2, 0.5, 0.8)),
                m(linrt.BSDF_COATING, int_ior=1.7, ext_ior=1.000277, nested=5, specular_reflectance=(0.9, 0.9, 0.9)),
                m(linrt.BSDF_DIFFUSE, diffuse=(0.7, 0.2, 0.2))]
        mats[3].fdr_int = linrt.fresnel_diffuse_reflectance(np.float32(1.0) / (np.float32(1.49) / np.float32(1.000277)))
        self.materials = (linrt.Material * len(mats))(*mats)
        room = linrt.ObjectDesc()
        C_ = __import__("ctypes")
        C_.pointer(room)[0] = d.objects[0]
        room.material = 0
        spheres = [((-0.45, -0.6, -0.3), 0.3, 1), ((0.45, -0.65, -0.35), 0.3, 2), ((-0.3, -0.75, 0.35), 0.22, 3),
                   ((0.35, -0.72, 0.3), 0.25, 4)]
        objs = [room]
        for c, r, mi in spheres:
            o = linrt.ObjectDesc()
            o.shape, o.material, o.emitter, o.radius = linrt.SHAPE_SPHERE, mi, -1, r
            for i in range(3):
                o.center[i] = c[i]
            objs.append(o)
        light_obj = None
        for i in range(d.n_objects):
            if d.objects[i].shape == linrt.SHAPE_RECTANGLE:
                light_obj = d.objects[i]
        lo = linrt.ObjectDesc()
        C_.pointer(lo)[0] = light_obj
        lo.emitter = 0
        objs.append(lo)
        self.objects = (linrt.ObjectDesc * len(objs))(*objs)
        ld = linrt.LightDesc()
        C_.pointer(ld)[0] = d.lights[0]
        ld.object = len(objs) - 1
        self.lights = (linrt.LightDesc * 1)(ld)
        self.desc = linrt.SceneDesc(len(objs), self.objects, len(mats), self.materials, 1, self.lights, 0, None,
                                    d.camera)
        self.desc_ptr = C_.pointer(self.desc)


class EmitterRoom:
    """The box room of the 256x256 workload lit by three area lights of
    different shapes (area_emitter.h over triangle_mesh / sphere / rectangle):
    light 0 a non-convex emissive mesh (an icosphere with radial bumps, its
    corner normals interpolated, so a ray toward one part of it can meet
    another part first), light 1 an emissive analytic sphere, light 2 the
    ceiling rectangle; a diffuse sphere and a plastic sphere receive the
    light.  Every emitter also carries a diffuse BSDF, as a reference shape
    may."""

    def __init__(self, seed=1, bumps=0.25):
        import ctypes as C_
        self.gen = linrt.box_icosphere(seed=seed, subdivisions=1)
        d = self.gen.desc
        m = linrt.material
        mats = [m(linrt.BSDF_DIFFUSE, diffuse=(0.62, 0.55, 0.45)),
                m(linrt.BSDF_DIFFUSE, diffuse=(0.3, 0.6, 0.3)),
                m(linrt.BSDF_PLASTIC, int_ior=1.49, ext_ior=1.000277, diffuse=(0.2, 0.5, 0.8))]
        mats[2].fdr_int = linrt.fresnel_diffuse_reflectance(np.float32(1.0) / (np.float32(1.49) / np.float32(1.000277)))
        self.materials = (linrt.Material * len(mats))(*mats)
        room = linrt.ObjectDesc()
        C_.pointer(room)[0] = d.objects[0]
        room.material = 0
        # the emissive mesh: icosphere (level 2) with seeded radial bumps
        pos, idx = linrt.box_icosphere(seed=seed, subdivisions=2).mesh(1)
        pos = pos.astype(np.float64)
        rng = np.random.default_rng(seed)
        rad = np.linalg.norm(pos, axis=1, keepdims=True)
        unit = pos / rad
        r = 0.16 * (1.0 + bumps * rng.uniform(-1.0, 1.0, (len(pos), 1)))
        self.mesh_pos = np.ascontiguousarray((unit * r + np.array([0.35, 0.35, -0.1])).astype(np.float32))
        self.mesh_idx = np.ascontiguousarray(idx.astype(np.uint32))
        self.mesh_cn = np.ascontiguousarray(unit[self.mesh_idx.reshape(-1)].astype(np.float32).reshape(-1, 9))
        mesh = linrt.ObjectDesc()
        mesh.shape, mesh.material, mesh.emitter = linrt.SHAPE_MESH, 0, 0
        mesh.n_vertices, mesh.n_triangles = len(self.mesh_pos), len(self.mesh_idx)
        mesh.positions = self.mesh_pos.ctypes.data_as(C_.POINTER(C_.c_float))
        mesh.indices = self.mesh_idx.ctypes.data_as(C_.POINTER(C_.c_uint32))
        mesh.corner_normals = self.mesh_cn.ctypes.data_as(C_.POINTER(C_.c_float))
        objs = [room, mesh]
        for c, rr, mi, em in (((-0.4, 0.45, 0.2), 0.12, 0, 1), ((-0.3, -0.65, -0.2), 0.3, 1, -1),
                              ((0.35, -0.7, 0.25), 0.28, 2, -1)):
            o = linrt.ObjectDesc()
            o.shape, o.material, o.emitter, o.radius = linrt.SHAPE_SPHERE, mi, em, rr
            for i in range(3):
                o.center[i] = c[i]
            objs.append(o)
        rect = linrt.ObjectDesc()
        for i in range(d.n_objects):
            if d.objects[i].shape == linrt.SHAPE_RECTANGLE:
                C_.pointer(rect)[0] = d.objects[i]
        rect.emitter = 2
        objs.append(rect)
        self.objects = (linrt.ObjectDesc * len(objs))(*objs)
        lights = []
        for obj, rgb in ((1, (6.0, 5.0, 4.0)), (2, (2.0, 3.0, 5.0)), (len(objs) - 1, (5.0, 5.0, 5.0))):
            ld = linrt.LightDesc()
            ld.kind, ld.object = linrt.LIGHT_AREA, obj
            for i in range(3):
                ld.radiance[i] = rgb[i]
            lights.append(ld)
        self.lights = (linrt.LightDesc * len(lights))(*lights)
        self.desc = linrt.SceneDesc(len(objs), self.objects, len(mats), self.materials, len(lights), self.lights, 0,
                                    None, d.camera)
        self.desc_ptr = C_.pointer(self.desc)
