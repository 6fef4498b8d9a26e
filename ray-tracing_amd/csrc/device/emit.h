// Device emitters: one-sided rectangle area lights (area_emitter.h:17-86) and
// the Hosek-Wilkie RGB sky with sun-cone next-event sampling
// (env_emitter.h:13-88), plus the scene-level light selection and MIS pdf
// (scene.h:62-108).
#pragma once

#include "trace.h"

namespace lrt {

struct DevLight {
    int32_t kind, object, sun_active, pad;
    float radiance[3];
    float sun[3], fs[3], ft[3], fn[3];  // sun direction and its frame
    float scale, cos_cutoff, cone_pdf, pad2;
    double cfg[3][9];
    double rad[3];
    uint32_t cdf_base, n_tris;  // mesh light: its triangles' running areas and vertices in DevLights
    float pad3[2];
};

constexpr int kMaxLights = 4;

struct DevLights {
    int32_t n;
    int32_t pad;
    const float* cdf;   // running triangle areas of every mesh light (triangle_mesh.h:16-25)
    const float* tri9;  // their triangles' vertices p0, p1, p2 (9 floats each)
    DevLight l[kMaxLights];
};

// ArHosekSkyModel.cc:291-304 and :637-651 in FP64 (published model formula),
// one channel; cg = cos(gamma), ct = cos(theta), zenith = sqrt(ct) are the
// same for the three channels and computed once.  The Mie term's x^1.5 is
// x * sqrt(x) (one correctly rounded square root and one rounded product,
// within an ulp of pow like the device pow itself; the channel's value is then
// rounded to float)
LRT_D double hosek(const double* c, double rad, double gamma, double cg, double ct, double zenith) {
    const double expM = exp(c[4] * gamma);
    const double rayM = cg * cg;
    const double x = 1.0 + c[8] * c[8] - 2.0 * c[8] * cg;
    const double mieM = (1.0 + cg * cg) / (x * sqrt(x));
    return (1.0 + c[0] * exp(c[1] / (ct + 0.01))) * (c[2] + c[3] * expM + c[5] * rayM + c[6] * mieM + c[7] * zenith) * rad;
}

LRT_D f3 sky_eval(const DevLight& L, f3 direction) {
    const f3 dir = normalize(direction);
    if (dir.y <= 0.0f) return splat(1.0f);  // below the horizon: 1, unscaled
    // the angles as env_emitter.h:79-80 computes them: the float acos of a
    // float, correctly rounded (here from double precision; the device acosf
    // can be an ulp off).  For 0 < y < 1.6e-8 the rounded theta is
    // float(pi/2), which lies above pi/2: cos(theta) < 0 and the zenith term
    // sqrt(cos(theta)) (ArHosekSkyModel.cc:300) is NaN, here as in the
    // reference, and the NaN reaches the sample's radiance
    const float theta = (float)acos((double)fminf(fmaxf(dir.y, 0.0f), 1.0f));
    const float gamma = (float)acos((double)fminf(fmaxf(dot(dir, ld3(L.sun)), -1.0f), 1.0f));
    const double g = gamma, cg = cos(g), ct = cos((double)theta), zenith = sqrt(ct);
    return f3{(float)hosek(L.cfg[0], L.rad[0], g, cg, ct, zenith), (float)hosek(L.cfg[1], L.rad[1], g, cg, ct, zenith),
              (float)hosek(L.cfg[2], L.rad[2], g, cg, ct, zenith)} * L.scale;
}

LRT_D f3 env_eval(const DevLights& Ls, f3 d) {
    f3 e = splat(0.0f);
    for (int i = 0; i < Ls.n; ++i)
        if (Ls.l[i].kind == LRT_LIGHT_SKY) e = e + sky_eval(Ls.l[i], d);
    return e;
}

struct DirectSample {
    f3 dir, point;
    float pdf, dist;
    f3 radiance;
    int emitter;
};

// hittable::sample_position of an area light's shape: a point, the shape's
// unit normal there and the area density (rectangle.h:67-71, sphere.h:59-67,
// triangle_mesh.h:109-138: a triangle chosen by area with upper_bound over the
// running areas, the remainder of x turned into the first barycentric
// coordinate, the geometric normal of the triangle)
LRT_D bool shape_sample(const DevScene& S, const DevLights& Ls, const DevLight& L, float sx, float sy, f3& q, f3& n,
                        float& pdf) {
    const DevObject& ob = S.obj[L.object];
    if (ob.shape == kShapeRect) {
        q = ld3(ob.corner) + sx * ld3(ob.eu) + sy * ld3(ob.ev);
        n = ld3(ob.rn);
    } else if (ob.shape == kShapeSphere) {
        const float z = 1.0f - 2.0f * sy;
        const float ring = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float phi = 2.0f * kPiF * sx;
        float sp, cp;
        sincos_rn(phi, &sp, &cp);
        n = f3{ring * cp, ring * sp, z};
        q = ld3(ob.center) + ob.radius * n;
    } else {
        const uint32_t m = L.n_tris;
        if (m == 0) return false;
        const float* cdf = Ls.cdf + L.cdf_base;
        const float target = fminf(sx, 0.999999f) * ob.area;
        uint32_t lo = 0, hi = m;  // first running area > target
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (target < cdf[mid]) hi = mid;
            else lo = mid + 1;
        }
        const uint32_t tri = lo >= m ? m - 1 : lo;
        const float prev = tri == 0 ? 0.0f : cdf[tri - 1];
        const float span = cdf[tri] - prev;
        const float u0 = span > 0.0f ? (target - prev) / span : 0.0f;
        const float su = sqrtf(u0);
        const float b0 = 1.0f - su;
        const float b1 = sy * su;
        const float b2 = 1.0f - b0 - b1;
        const float* v = Ls.tri9 + 9 * ((size_t)L.cdf_base + tri);
        const f3 a0 = ld3(v), a1 = ld3(v + 3), a2 = ld3(v + 6);
        q = b0 * a0 + b1 * a1 + b2 * a2;
        const f3 g = cross(a1 - a0, a2 - a0);
        n = g / len(g);
    }
    pdf = 1.0f / ob.area;
    return true;
}

LRT_D bool light_sample(const DevScene& S, const DevLights& Ls, const DevLight& L, f3 p, float sx, float sy,
                        DirectSample& out) {
    if (L.kind == LRT_LIGHT_SKY) {
        if (!L.sun_active) return false;
        const float ct = (1.0f - sx) + sx * L.cos_cutoff;
        const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const float phi = 2.0f * kPiF * sy;
        float sp, cp;
        sincos_rn(phi, &sp, &cp);
        const f3 local{cp * st, sp * st, ct};
        out.dir = local.x * ld3(L.fs) + local.y * ld3(L.ft) + local.z * ld3(L.fn);
        out.pdf = L.cone_pdf;
        out.dist = kInfF;
        out.radiance = sky_eval(L, out.dir);
        return true;
    }
    const DevObject& ob = S.obj[L.object];
    if (ob.area <= 0.0f) return false;
    // area_emitter::sample (area_emitter.h:17-45)
    f3 q, n;
    float pdf_area;
    if (!shape_sample(S, Ls, L, sx, sy, q, n, pdf_area)) return false;
    const f3 to = q - p;
    const float d2 = len2(to);
    if (d2 < 1e-8f) return false;
    const float d = sqrtf(d2);
    const f3 dir = to / d;
    const float cl = dot(n, -dir);
    if (cl <= 0.0f) return false;
    out.point = q;
    out.dir = dir;
    out.dist = d;
    out.pdf = pdf_area * d2 / cl;
    out.radiance = ld3(L.radiance);
    return true;
}

// area_emitter::pdf (area_emitter.h:47-71): the emitter's own shape ray-cast
// over [0.001, inf), front face only.  MESH: compiled with the mesh lights'
// kd traversal (private stack), for kernels launched on scenes that have one;
// without it a mesh light's pdf is never asked for.
template <bool MESH>
LRT_D float light_pdf(const DevScene& S, const DevLight& L, f3 p, f3 direction) {
    if (L.kind == LRT_LIGHT_SKY) {
        if (!L.sun_active) return 0.0f;
        const f3 dir = normalize(direction);
        if (dot(dir, ld3(L.sun)) + 1e-5f < L.cos_cutoff) return 0.0f;
        return L.cone_pdf;
    }
    const DevObject& ob = S.obj[L.object];
    if (ob.area <= 0.0f) return 0.0f;
    const f3 dir = normalize(direction);
    const Ray r{p, dir};
    float t;
    f3 nrm;
    if (ob.shape == kShapeRect) {
        float a, b;
        if (!rect_test(ob, r, 0.001f, kInfF, t, a, b)) return 0.0f;
        nrm = ld3(ob.rn);
        if (!(dot(r.d, nrm) < 0)) return 0.0f;  // back face
    } else if (ob.shape == kShapeSphere) {
        if (!sphere_test(ob, r, 0.001f, kInfF, t)) return 0.0f;
        nrm = (r.o + t * r.d - ld3(ob.center)) / ob.radius;
        if (!(dot(r.d, nrm) < 0)) return 0.0f;
    } else {
        if (!MESH) return 0.0f;
        float inv[3] = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
        float t0 = 0.001f, c = kInfF, u = 0.0f, v = 0.0f;
        const int o = L.object;
        if (!mesh_box_clip(S, o, r, inv, t0, c)) return 0.0f;
        int32_t prim = -1;
        Counters cnt{0, 0, 0};
        if (!kd_traverse<false, false, false>(S, ob, r, inv, t0, c, prim, u, v, nullptr, cnt)) return 0.0f;
        t = c;
        bool front;
        mesh_hit_normal(S, ob, prim, u, v, r.d, nrm, front);
        if (!front) return 0.0f;
    }
    const float cl = dot(nrm, -dir);
    if (cl <= 0.0f) return 0.0f;
    return t * t / (ob.area * cl);
}

LRT_D bool sample_emitter_direction(const DevScene& S, const DevLights& Ls, f3 p, float sx, float sy, DirectSample& out) {
    const int count = Ls.n;
    if (count <= 0) return false;
    const float scaled = sx * (float)count;
    int index = (int)scaled;
    if (index >= count) index = count - 1;
    if (!light_sample(S, Ls, Ls.l[index], p, scaled - (float)index, sy, out)) return false;
    out.emitter = index;
    out.pdf *= 1.0f / (float)count;
    return out.pdf > 0.0f;
}

template <bool MESH = false>
LRT_D float pdf_emitter_direction(const DevScene& S, const DevLights& Ls, f3 p, f3 d) {
    const int count = Ls.n;
    if (count <= 0) return 0.0f;
    const float sel = 1.0f / (float)count;
    float sum = 0.0f;
    for (int i = 0; i < count; ++i) sum += sel * light_pdf<MESH>(S, Ls.l[i], p, d);
    return sum;
}

LRT_D float power_heuristic(float a_pdf, float b_pdf) {  // path_integrator.h:177-182
    const float a = a_pdf * a_pdf, b = b_pdf * b_pdf;
    const float s = a + b;
    return s > 0.0f ? a / s : 0.0f;
}

}  // namespace lrt
