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
};

constexpr int kMaxLights = 4;

struct DevLights {
    int32_t n;
    DevLight l[kMaxLights];
};

// ArHosekSkyModel.cc:291-304 and :637-651 in FP64 (published model formula)
LRT_D double hosek(const double* c, double rad, double theta, double gamma) {
    const double cg = cos(gamma);
    const double expM = exp(c[4] * gamma);
    const double rayM = cg * cg;
    const double mieM = (1.0 + cg * cg) / pow((1.0 + c[8] * c[8] - 2.0 * c[8] * cg), 1.5);
    const double ct = cos(theta);
    const double zenith = sqrt(ct);
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
    return f3{(float)hosek(L.cfg[0], L.rad[0], theta, gamma), (float)hosek(L.cfg[1], L.rad[1], theta, gamma),
              (float)hosek(L.cfg[2], L.rad[2], theta, gamma)} * L.scale;
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

LRT_D bool light_sample(const DevScene& S, const DevLight& L, f3 p, float sx, float sy, DirectSample& out) {
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
    // rectangle::sample_position (rectangle.h:67-71)
    const f3 q = ld3(ob.corner) + sx * ld3(ob.eu) + sy * ld3(ob.ev);
    const f3 n = ld3(ob.rn);
    const float pdf_area = 1.0f / ob.area;
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
    float t, a, b;
    if (!rect_test(ob, r, 0.001f, kInfF, t, a, b)) return 0.0f;
    const f3 rn = ld3(ob.rn);
    if (!(dot(r.d, rn) < 0)) return 0.0f;  // back face
    const float cl = dot(rn, -dir);
    if (cl <= 0.0f) return 0.0f;
    return t * t / (ob.area * cl);
}

LRT_D bool sample_emitter_direction(const DevScene& S, const DevLights& Ls, f3 p, float sx, float sy, DirectSample& out) {
    const int count = Ls.n;
    if (count <= 0) return false;
    const float scaled = sx * (float)count;
    int index = (int)scaled;
    if (index >= count) index = count - 1;
    if (!light_sample(S, Ls.l[index], p, scaled - (float)index, sy, out)) return false;
    out.emitter = index;
    out.pdf *= 1.0f / (float)count;
    return out.pdf > 0.0f;
}

LRT_D float pdf_emitter_direction(const DevScene& S, const DevLights& Ls, f3 p, f3 d) {
    const int count = Ls.n;
    if (count <= 0) return 0.0f;
    const float sel = 1.0f / (float)count;
    float sum = 0.0f;
    for (int i = 0; i < count; ++i) sum += sel * light_pdf(S, Ls.l[i], p, d);
    return sum;
}

LRT_D float power_heuristic(float a_pdf, float b_pdf) {  // path_integrator.h:177-182
    const float a = a_pdf * a_pdf, b = b_pdf * b_pdf;
    const float s = a + b;
    return s > 0.0f ? a / s : 0.0f;
}

}  // namespace lrt
