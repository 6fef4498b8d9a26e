// Device BSDFs.  One function family per reference class, evaluated in the
// local +Z shading frame; eval and pdf are computed together (the reference
// recomputes the shared terms in separate virtual calls; the values are the
// same because every term keeps its operation order).  Coatings nest exactly
// one base BSDF (the host rejects a coating over a coating).
#pragma once

#include "../../../include/linrt.h"
#include "common.h"

namespace lrt {

constexpr uint32_t F_DIFFUSE = LRT_FLAG_DIFFUSE, F_SPECULAR = LRT_FLAG_SPECULAR, F_REFLECTIVE = LRT_FLAG_REFLECTIVE,
                   F_GLOSSY = LRT_FLAG_GLOSSY, F_METALLIC = LRT_FLAG_METALLIC, F_DIELECTRIC = LRT_FLAG_DIELECTRIC,
                   F_REFRACTIVE = LRT_FLAG_REFRACTIVE;
constexpr int M_SOLID = LRT_MEASURE_SOLID_ANGLE, M_DISCRETE = LRT_MEASURE_DISCRETE;

struct BsdfCtx {
    int mode = LRT_MODE_RADIANCE;
    uint32_t mask = 0xffffffffu;
    int component = -1;
    float lobe = 0.0f;
};

struct SRec {
    f3 wi, wo;
    float eta;
    float pdf;
    uint32_t type;
    int comp;
    bool front;
};

struct Mats {
    const lrt_material* m;
    const float* tables;  // 101 floats per slice
};

struct EP {
    f3 f;
    float pdf;
};
LRT_D EP ep0() { return EP{splat(0.0f), 0.0f}; }

LRT_D f3 col(const float* c) { return f3{c[0], c[1], c[2]}; }
LRT_D float lum(const float* c) { return 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]; }
LRT_D f3 reflect_n(f3 w, f3 n) { return 2.0f * dot(w, n) * n - w; }  // frame.h:14-16
LRT_D f3 refract_n(f3 wi, f3 n, float ct, float ei, float et) {     // frame.h:20-24
    const float s = ei / et;
    return n * (s * dot(wi, n) - ct) - wi * s;
}
LRT_D bool aligned(f3 a, f3 b) { return fabsf(dot(a, b) - 1.0f) < 1e-3f; }
LRT_D float sgn(float v) { return (float)((v > 0.0f) - (v < 0.0f)); }
LRT_D f3 zplus() { return f3{0.0f, 0.0f, 1.0f}; }

// ---- GGX (microfacet.h:10-80) ----------------------------------------------
struct Ggx {
    float a;
};
LRT_D Ggx ggx(float alpha) { return Ggx{fmaxf(alpha, 0.0001f)}; }
LRT_D float ggx_D(Ggx g, f3 m) {
    if (m.z <= 0.0f) return 0.0f;
    const float a2 = g.a * g.a;
    const float t = (m.x * m.x + m.y * m.y) / a2 + m.z * m.z;
    return 1.0f / (kPiF * a2 * t * t);
}
LRT_D float ggx_G1(Ggx g, f3 v, f3 m) {
    const float cv = fabsf(v.z);
    if (cv < 1e-6f || dot(v, m) * v.z <= 0.0f) return 0.0f;
    const float t = (g.a * g.a) * (v.x * v.x + v.y * v.y) / (v.z * v.z);
    const float lambda = 0.5f * (-1.0f + sqrtf(1.0f + t));
    return 1.0f / (1.0f + lambda);
}
LRT_D float ggx_G(Ggx g, f3 wi, f3 wo, f3 m) { return ggx_G1(g, wi, m) * ggx_G1(g, wo, m); }
LRT_D float ggx_pdf(Ggx g, f3 wi, f3 m) {
    if (wi.z <= 1e-6f) return 0.0f;
    return ggx_G1(g, wi, m) * fmaxf(0.0f, dot(wi, m)) * ggx_D(g, m) / wi.z;
}
// visible-normal sampling (Heitz 2018)
LRT_D f3 ggx_sample(Ggx g, f3 wi, float ux, float uy) {
    const f3 vh = normalize(f3{g.a * wi.x, g.a * wi.y, wi.z});
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const f3 t1 = lensq > 0.0f ? f3{-vh.y, vh.x, 0.0f} / sqrtf(lensq) : f3{1.0f, 0.0f, 0.0f};
    const f3 t2 = cross(vh, t1);
    const float r = sqrtf(ux);
    const float phi = 2.0f * kPiF * uy;
    float sp, cp;
    sincos_rn(phi, &sp, &cp);
    const float dx = r * cp;
    float dy = r * sp;
    const float s = 0.5f * (1.0f + vh.z);
    dy = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - dx * dx)) + s * dy;
    const f3 nh = dx * t1 + dy * t2 + sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy)) * vh;
    return normalize(f3{g.a * nh.x, g.a * nh.y, fmaxf(0.0f, nh.z)});
}

// ---- Fresnel (fresnel.h) ----------------------------------------------------
// Conductor Fresnel with complex IOR in FP32 (fresnel.h:43-56): complex
// products written out, quotients by Smith's method, the square root by the
// principal-branch half-angle form.
struct cpx {
    float re, im;
};
LRT_D cpx cmul(cpx a, cpx b) { return cpx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
LRT_D cpx cdiv(cpx a, cpx b) {
    if (fabsf(b.re) < fabsf(b.im)) {
        const float r = b.re / b.im, d = (b.re * r) + b.im;
        return cpx{((a.re * r) + a.im) / d, ((a.im * r) - a.re) / d};
    }
    const float r = b.im / b.re, d = (b.im * r) + b.re;
    return cpx{((a.im * r) + a.re) / d, (a.im - (a.re * r)) / d};
}
LRT_D cpx csqrt_(cpx z) {
    if (z.im == 0.0f) {
        if (z.re < 0.0f) return cpx{0.0f, copysignf(sqrtf(-z.re), z.im)};
        return cpx{fabsf(sqrtf(z.re)), copysignf(0.0f, z.im)};
    }
    const float d = hypotf(z.re, z.im);
    float r, s;
    if (z.re > 0.0f) {
        r = sqrtf(0.5f * (d + z.re));
        s = 0.5f * (z.im / r);
    } else {
        s = sqrtf(0.5f * (d - z.re));
        r = fabsf(0.5f * (z.im / s));
    }
    return cpx{r, copysignf(s, z.im)};
}
LRT_D float fresnel_conductor1(float ci, float eta, float k) {
    ci = fminf(fmaxf(ci, 0.0f), 1.0f);
    const cpx ior{eta, k};
    const float s2i = 1.0f - ci * ci;
    const cpx s2t = cdiv(cpx{s2i, 0.0f}, cmul(ior, ior));
    const cpx ct = csqrt_(cpx{1.0f - s2t.re, 0.0f - s2t.im});
    const cpx ic{ior.re * ci, ior.im * ci};
    const cpx rpa = cdiv(cpx{ic.re - ct.re, ic.im - ct.im}, cpx{ic.re + ct.re, ic.im + ct.im});
    const cpx ict = cmul(ior, ct);
    const cpx rpe = cdiv(cpx{ci - ict.re, 0.0f - ict.im}, cpx{ci + ict.re, 0.0f + ict.im});
    return 0.5f * ((rpa.re * rpa.re + rpa.im * rpa.im) + (rpe.re * rpe.re + rpe.im * rpe.im));
}
LRT_D f3 fresnel_conductor(float ci, const float* eta, const float* k) {
    return f3{fresnel_conductor1(ci, eta[0], k[0]), fresnel_conductor1(ci, eta[1], k[1]), fresnel_conductor1(ci, eta[2], k[2])};
}
struct Fres {
    float R, ct;
};
LRT_D Fres fresnel_dielectric(float ci, float ei, float et) {  // fresnel.h:72-95
    Fres r{1.0f, 0.0f};
    if (ei <= 0.0f || et <= 0.0f) return r;
    ci = fminf(fmaxf(ci, 0.0f), 1.0f);
    const float s2i = 1.0f - ci * ci;
    const float s = ei / et;
    const float s2t = s * s * s2i;
    if (s2t > 1.0f) return r;
    r.ct = sqrtf(1.0f - s2t);
    const float rs = (ei * ci - et * r.ct) / (ei * ci + et * r.ct);
    const float rp = (et * ci - ei * r.ct) / (et * ci + ei * r.ct);
    r.R = 0.5f * (rs * rs + rp * rp);
    return r;
}
LRT_D float fresnel_ext(float ci, float& ct_out, float eta) {  // fresnel.h:99-118
    if (eta == 1.0f) { ct_out = -ci; return 0.0f; }
    const float s = ci > 0.0f ? (1.0f / eta) : eta;
    const float ct2 = 1.0f - (1.0f - ci * ci) * (s * s);
    if (ct2 <= 0.0f) { ct_out = 0.0f; return 1.0f; }
    const float ai = fabsf(ci), at = sqrtf(ct2);
    const float rs = (ai - eta * at) / (ai + eta * at);
    const float rp = (eta * ai - at) / (eta * ai + at);
    ct_out = ci > 0.0f ? -at : at;
    return 0.5f * (rs * rs + rp * rp);
}
LRT_D float fresnel_ext(float ci, float eta) {
    float t;
    return fresnel_ext(ci, t, eta);
}

LRT_D f3 cos_hemi(float ux, float uy, float& pdf) {
    const float ct = sqrtf(ux), st = sqrtf(1.0f - ux);
    const float phi = 2.0f * kPiF * uy;
    float sp, cp;
    sincos_rn(phi, &sp, &cp);
    pdf = ct / kPiF;
    return f3{cp * st, sp * st, ct};
}

// rough-transmittance slice lookup (rough_transmittance.h:17-29)
LRT_D float trans_eval(const float* tb, float ct) {
    if (!(ct >= 0.0f)) return 0.0f;
    const float x = qrt_rn(ct);
    if (!(x >= 0.0f && x <= 1.0f)) return 0.0f;
    float t = ((x - 0.0f) * 99.0f) / (1.0f - 0.0f);
    int k = (int)t;
    if (k > 98) k = 98;
    const float f0 = tb[k], f1 = tb[k + 1];
    const float d0 = k > 0 ? 0.5f * (tb[k + 1] - tb[k - 1]) : (tb[k + 1] - tb[k]);
    const float d1 = (k + 2 < 100) ? 0.5f * (tb[k + 2] - tb[k]) : (tb[k + 1] - tb[k]);
    t -= (float)k;
    const float t2 = t * t, t3 = t2 * t;
    const float r = (2.0f * t3 - 3.0f * t2 + 1.0f) * f0 + (-2.0f * t3 + 3.0f * t2) * f1 + (t3 - 2.0f * t2 + t) * d0 +
                    (t3 - t2) * d1;
    return fminf(fmaxf(r, 0.0f), 1.0f);
}
LRT_D float trans_diffuse(const float* tb) { return fminf(fmaxf(tb[100], 0.0f), 1.0f); }

// ---- flags / components (bsdf.h:58-62 per class) --------------------------
LRT_D uint32_t base_flags(const lrt_material& M) {
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return F_DIFFUSE | F_REFLECTIVE;
        case LRT_BSDF_CONDUCTOR: return M.alpha <= 0.0f ? (F_SPECULAR | F_REFLECTIVE | F_METALLIC) : (F_GLOSSY | F_REFLECTIVE | F_METALLIC);
        case LRT_BSDF_DIELECTRIC: return F_SPECULAR | F_REFLECTIVE | F_REFRACTIVE | F_DIELECTRIC;
        case LRT_BSDF_ROUGH_DIELECTRIC: return F_GLOSSY | F_REFLECTIVE | F_REFRACTIVE | F_DIELECTRIC;
        case LRT_BSDF_PLASTIC: return F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC | F_DIFFUSE;
        case LRT_BSDF_ROUGH_PLASTIC: return F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC | F_DIFFUSE;
        default: return 0u;
    }
}
LRT_D int base_count(const lrt_material& M) {
    return (M.type == LRT_BSDF_DIELECTRIC || M.type == LRT_BSDF_PLASTIC || M.type == LRT_BSDF_ROUGH_PLASTIC) ? 2 : 1;
}
LRT_D uint32_t bsdf_flags(const Mats& S, int m) {
    const lrt_material& M = S.m[m];
    if (M.type == LRT_BSDF_COATING) return base_flags(S.m[M.nested]) | (F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC);
    if (M.type == LRT_BSDF_ROUGH_COATING) return base_flags(S.m[M.nested]) | (F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC);
    return base_flags(M);
}
// component_count / component_flags (bsdf.h:60-62 and the overrides in
// dielectric_bsdf.h:23-29, plastic_bsdf.h:33-39, rough_plastic_bsdf.h:39-43,
// coating_bsdf.h:33-39, rough_coating_bsdf.h:39-45)
LRT_D int bsdf_component_count(const Mats& S, int m) {
    const lrt_material& M = S.m[m];
    if (M.type == LRT_BSDF_COATING || M.type == LRT_BSDF_ROUGH_COATING) return base_count(S.m[M.nested]) + 1;
    return base_count(M);
}
LRT_D uint32_t base_component_flags(const lrt_material& M, int c) {
    switch (M.type) {
        case LRT_BSDF_DIELECTRIC:
            return c == 0 ? (F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC) : (F_SPECULAR | F_REFRACTIVE | F_DIELECTRIC);
        case LRT_BSDF_PLASTIC: return c == 0 ? (F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC) : (F_DIFFUSE | F_REFLECTIVE);
        case LRT_BSDF_ROUGH_PLASTIC: return c == 0 ? (F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC) : (F_DIFFUSE | F_REFLECTIVE);
        default: return base_flags(M);
    }
}
LRT_D uint32_t bsdf_component_flags(const Mats& S, int m, int c) {
    const lrt_material& M = S.m[m];
    if (M.type == LRT_BSDF_COATING || M.type == LRT_BSDF_ROUGH_COATING) {
        const lrt_material& N = S.m[M.nested];
        if (c == base_count(N))
            return M.type == LRT_BSDF_COATING ? (F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC)
                                              : (F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC);
        return base_component_flags(N, c);
    }
    return base_component_flags(M, c);
}

LRT_D void clear(SRec& r) {
    r.pdf = 0.0f;
    r.eta = 1.0f;
    r.type = 0;
    r.comp = -1;
}

// ---- diffuse (diffuse_bsdf.h:18-72) --------------------------------------------
LRT_D EP diffuse_ep(const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    if (meas != M_SOLID || !(c.mask & (F_DIFFUSE | F_REFLECTIVE)) || (c.component != -1 && c.component != 0) ||
        r.wi.z <= 0.0f || r.wo.z <= 0.0f)
        return ep0();
    return EP{col(M.diffuse) * (r.wo.z / kPiF), r.wo.z / kPiF};
}
LRT_D f3 diffuse_sample(const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    clear(r);
    if (!(c.mask & (F_DIFFUSE | F_REFLECTIVE)) || (c.component != -1 && c.component != 0) || r.wi.z <= 0.0f) return splat(0.0f);
    float pdf;
    r.wo = cos_hemi(ux, uy, pdf);
    r.pdf = pdf;
    r.type = F_DIFFUSE | F_REFLECTIVE;
    r.comp = 0;
    return col(M.diffuse);
}

// ---- conductor (rough_conductor_bsdf.h:28-128) ---------------------------------
LRT_D EP conductor_lobe_ep(const lrt_material& M, f3 wi, f3 wo) {
    if (wi.z <= 1e-6f || wo.z <= 1e-6f) return ep0();
    const f3 h = wi + wo;
    if (len2(h) <= 1e-12f) return ep0();
    const f3 m = normalize(h);
    const float wim = dot(wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return ep0();
    const Ggx g = ggx(M.alpha);
    const float d = ggx_D(g, m), gg = ggx_G(g, wi, wo, m);
    const f3 F = fresnel_conductor(wim, M.cond_eta, M.cond_k) * col(M.specular_reflectance);
    return EP{F * (d * gg / (4.0f * wi.z)), ggx_pdf(g, wi, m) / (4.0f * wim)};
}
LRT_D uint32_t conductor_lobe(const lrt_material& M) {
    return M.alpha <= 0.0f ? (F_SPECULAR | F_REFLECTIVE | F_METALLIC) : (F_GLOSSY | F_REFLECTIVE | F_METALLIC);
}
LRT_D EP conductor_ep(const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    if ((c.mask & conductor_lobe(M)) == 0 || (c.component != -1 && c.component != 0)) return ep0();
    if (M.alpha <= 0.0f) {
        if (meas != M_DISCRETE || r.wo.z <= 0.0f || !aligned(f3{-r.wi.x, -r.wi.y, r.wi.z}, r.wo)) return ep0();
        return EP{fresnel_conductor(r.wi.z, M.cond_eta, M.cond_k) * col(M.specular_reflectance), 1.0f};
    }
    if (meas != M_SOLID) return ep0();
    return conductor_lobe_ep(M, r.wi, r.wo);
}
LRT_D f3 conductor_sample(const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    const uint32_t lobe = conductor_lobe(M);
    clear(r);
    if ((c.mask & lobe) == 0 || (c.component != -1 && c.component != 0) || r.wi.z < 0.0f) return splat(0.0f);
    if (M.alpha <= 0.0f) {
        r.wo = f3{-r.wi.x, -r.wi.y, r.wi.z};
        r.type = lobe;
        r.comp = 0;
        r.pdf = 1.0f;
        return fresnel_conductor(r.wi.z, M.cond_eta, M.cond_k) * col(M.specular_reflectance);
    }
    const f3 m = ggx_sample(ggx(M.alpha), r.wi, ux, uy);
    const float wim = dot(r.wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return splat(0.0f);
    r.wo = reflect_n(r.wi, m);
    if (r.wo.z <= 0.0f) return splat(0.0f);
    const EP e = conductor_lobe_ep(M, r.wi, r.wo);
    if (e.pdf <= 0.0f) return splat(0.0f);
    r.pdf = e.pdf;
    r.type = lobe;
    r.comp = 0;
    return e.f / e.pdf;
}

// ---- smooth dielectric (dielectric_bsdf.h:31-157) ------------------------------
constexpr uint32_t kDielR = F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC, kDielT = F_SPECULAR | F_REFRACTIVE | F_DIELECTRIC;
LRT_D EP dielectric_ep(const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    if (meas != M_DISCRETE || r.wi.z <= 0.0f) return ep0();
    const bool ar = (c.mask & kDielR) != 0 && (c.component == -1 || c.component == 0);
    const bool at = (c.mask & kDielT) != 0 && (c.component == -1 || c.component == 1);
    const float ei = r.front ? M.ext_ior : M.int_ior, et = r.front ? M.int_ior : M.ext_ior;
    const Fres F = fresnel_dielectric(r.wi.z, ei, et);
    if (r.wi.z * r.wo.z >= 0.0f) {
        if (!ar || !aligned(reflect_n(r.wi, zplus()), r.wo)) return ep0();
        return EP{col(M.specular_reflectance) * F.R, at ? F.R : 1.0f};
    }
    if (!at || F.R >= 1.0f) return ep0();
    if (!aligned(refract_n(r.wi, zplus(), F.ct, ei, et), r.wo)) return ep0();
    f3 v = col(M.specular_transmittance) * (1.0f - F.R);
    if (c.mode == LRT_MODE_RADIANCE) {
        const float e = ei / et;
        v = v * (e * e);
    }
    return EP{v, ar ? (1.0f - F.R) : 1.0f};
}
LRT_D f3 dielectric_sample(const lrt_material& M, const BsdfCtx& c, SRec& r, float ux) {
    clear(r);
    if (r.wi.z <= 0.0f) return splat(0.0f);
    const bool ar = (c.mask & kDielR) != 0 && (c.component == -1 || c.component == 0);
    const bool at = (c.mask & kDielT) != 0 && (c.component == -1 || c.component == 1);
    if (!ar && !at) return splat(0.0f);
    const float ei = r.front ? M.ext_ior : M.int_ior, et = r.front ? M.int_ior : M.ext_ior;
    const Fres F = fresnel_dielectric(r.wi.z, ei, et);
    if (ar && (!at || ux < F.R)) {
        r.wo = reflect_n(r.wi, zplus());
        r.type = kDielR;
        r.comp = 0;
        r.pdf = at ? F.R : 1.0f;
        return at ? col(M.specular_reflectance) : col(M.specular_reflectance) * F.R;
    }
    if (F.R >= 1.0f) return splat(0.0f);
    r.wo = refract_n(r.wi, zplus(), F.ct, ei, et);
    r.eta = ei / et;
    r.type = kDielT;
    r.comp = 1;
    r.pdf = ar ? (1.0f - F.R) : 1.0f;
    f3 w = col(M.specular_transmittance);
    if (!ar) w = w * (1.0f - F.R);
    if (c.mode == LRT_MODE_RADIANCE) w = w * (r.eta * r.eta);
    return w;
}

// ---- rough dielectric (rough_dielectric_bsdf.h:28-197) -------------------------
LRT_D bool rd_r_ok(const BsdfCtx& c) { return (c.mask & (F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC)) != 0 && (c.component == -1 || c.component == 0); }
LRT_D bool rd_t_ok(const BsdfCtx& c) { return (c.mask & (F_GLOSSY | F_REFRACTIVE | F_DIELECTRIC)) != 0 && (c.component == -1 || c.component == 1); }
LRT_D EP rd_reflection(const lrt_material& M, const BsdfCtx& c, f3 wi, f3 wo, float ei, float et, bool both) {
    if (!rd_r_ok(c) || wi.z <= 1e-6f || wo.z <= 1e-6f) return ep0();
    f3 m = wi + wo;
    if (len2(m) <= 1e-12f) return ep0();
    m = normalize(m);
    if (m.z < 0.0f) m = -m;
    const float wim = dot(wi, m);
    if (wim <= 0.0f) return ep0();
    const Ggx g = ggx(M.alpha);
    const float F = fresnel_dielectric(wim, ei, et).R;
    const float d = ggx_D(g, m), gg = ggx_G(g, wi, wo, m), pm = ggx_pdf(g, wi, m);
    const float lobe = both ? F : 1.0f;
    return EP{splat(F * d * gg / (4.0f * wi.z)), pm * lobe / (4.0f * wim)};
}
LRT_D EP rd_transmission(const lrt_material& M, const BsdfCtx& c, f3 wi, f3 wo, float ei, float et, bool both) {
    if (!rd_t_ok(c) || wi.z <= 1e-6f || wo.z >= -1e-6f || et <= 0.0f) return ep0();
    f3 m = -(ei * wi + et * wo);
    if (len2(m) <= 1e-12f) return ep0();
    m = normalize(m);
    if (m.z < 0.0f) m = -m;
    const float wim = dot(wi, m), wom = dot(wo, m);
    if (wim <= 0.0f) return ep0();
    const float F = fresnel_dielectric(wim, ei, et).R;
    const float denom = ei * wim + et * wom;
    if (F >= 1.0f || denom * denom <= 1e-12f) return ep0();
    const Ggx g = ggx(M.alpha);
    const float d = ggx_D(g, m), gg = ggx_G(g, wi, wo, m), pm = ggx_pdf(g, wi, m);
    const float awom = fabsf(wom);
    const float es = c.mode == LRT_MODE_RADIANCE ? ei : et;
    const float lobe = both ? (1.0f - F) : 1.0f;
    const float value = (1.0f - F) * d * gg * wim * awom * es * es / (wi.z * denom * denom);
    return EP{splat(value), pm * lobe * (et * et) * awom / (denom * denom)};
}
LRT_D EP rd_ep(const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    if (meas != M_SOLID) return ep0();
    if (r.wi.z <= 1e-6f || fabsf(r.wo.z) <= 1e-6f) return ep0();
    const float ei = r.front ? M.ext_ior : M.int_ior, et = r.front ? M.int_ior : M.ext_ior;
    const bool both = rd_r_ok(c) && rd_t_ok(c);
    if (r.wi.z * r.wo.z > 0.0f) return rd_reflection(M, c, r.wi, r.wo, ei, et, both);
    return rd_transmission(M, c, r.wi, r.wo, ei, et, both);
}
LRT_D f3 rd_sample(const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    clear(r);
    if (r.wi.z <= 0.0f) return splat(0.0f);
    const bool ar = rd_r_ok(c), at = rd_t_ok(c);
    if (!ar && !at) return splat(0.0f);
    const float ei = r.front ? M.ext_ior : M.int_ior, et = r.front ? M.int_ior : M.ext_ior;
    const f3 m = ggx_sample(ggx(M.alpha), r.wi, ux, uy);
    const float wim = dot(r.wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return splat(0.0f);
    const Fres F = fresnel_dielectric(wim, ei, et);
    const bool both = ar && at;
    if (ar && (!at || c.lobe < F.R)) {
        r.wo = reflect_n(r.wi, m);
        if (r.wo.z <= 0.0f) return splat(0.0f);
        const EP l = rd_reflection(M, c, r.wi, r.wo, ei, et, both);
        if (l.pdf <= 0.0f) return splat(0.0f);
        r.pdf = l.pdf;
        r.type = F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC;
        r.comp = 0;
        return col(M.specular_reflectance) * (l.f / l.pdf);
    }
    if (F.ct <= 0.0f) return splat(0.0f);
    r.wo = refract_n(r.wi, m, F.ct, ei, et);
    if (r.wo.z >= 0.0f) return splat(0.0f);
    r.eta = ei / et;
    const EP l = rd_transmission(M, c, r.wi, r.wo, ei, et, both);
    if (l.pdf <= 0.0f) return splat(0.0f);
    r.pdf = l.pdf;
    r.type = F_GLOSSY | F_REFRACTIVE | F_DIELECTRIC;
    r.comp = 1;
    return col(M.specular_transmittance) * (l.f / l.pdf);
}

// ---- smooth plastic (plastic_bsdf.h:41-191) ------------------------------------
constexpr uint32_t kPlS = F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC, kPlD = F_DIFFUSE | F_REFLECTIVE;
struct Plastic {
    float eta, inv_eta2, w;
};
LRT_D Plastic plastic_of(const lrt_material& M) {
    Plastic p;
    p.eta = M.int_ior / M.ext_ior;
    p.inv_eta2 = 1.0f / (p.eta * p.eta);
    p.w = lum(M.specular_reflectance) / (lum(M.diffuse) + lum(M.specular_reflectance));
    return p;
}
LRT_D float plastic_prob(const Plastic& p, float Fi) {
    const float s = Fi * p.w, d = (1.0f - Fi) * (1.0f - p.w);
    return s / (s + d);
}
LRT_D f3 scaled_diffuse(const lrt_material& M, float fdr) {
    const f3 kd = col(M.diffuse);
    if (M.nonlinear) return f3{kd.x / (1.0f - kd.x * fdr), kd.y / (1.0f - kd.y * fdr), kd.z / (1.0f - kd.z * fdr)};
    return kd / (1.0f - fdr);
}
LRT_D f3 plastic_dw(const lrt_material& M, const Plastic& p, float Fi, float Fo) {
    return scaled_diffuse(M, M.fdr_int) * (p.inv_eta2 * (1.0f - Fi) * (1.0f - Fo));
}
LRT_D EP plastic_ep(const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    const bool as = (c.mask & kPlS) != 0 && (c.component == -1 || c.component == 0);
    const bool ad = (c.mask & kPlD) != 0 && (c.component == -1 || c.component == 1);
    if (r.wi.z <= 0.0f) return ep0();
    const Plastic p = plastic_of(M);
    if (meas == M_DISCRETE) {
        if (!as || !aligned(reflect_n(r.wi, zplus()), r.wo)) return ep0();
        const float Fi = fresnel_ext(r.wi.z, p.eta);
        return EP{col(M.specular_reflectance) * Fi, ad ? plastic_prob(p, Fi) : 1.0f};
    }
    if (meas != M_SOLID || !ad || r.wo.z <= 0.0f) return ep0();
    const float Fi = fresnel_ext(r.wi.z, p.eta), Fo = fresnel_ext(r.wo.z, p.eta);
    const float ps = as ? plastic_prob(p, Fi) : 0.0f;
    return EP{plastic_dw(M, p, Fi, Fo) * (r.wo.z / kPiF), (r.wo.z / kPiF) * (1.0f - ps)};
}
LRT_D f3 plastic_sample(const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    clear(r);
    const bool as = (c.mask & kPlS) != 0 && (c.component == -1 || c.component == 0);
    const bool ad = (c.mask & kPlD) != 0 && (c.component == -1 || c.component == 1);
    if ((!as && !ad) || r.wi.z <= 0.0f) return splat(0.0f);
    const Plastic p = plastic_of(M);
    const float Fi = fresnel_ext(r.wi.z, p.eta);
    if (as && ad) {
        const float ps = plastic_prob(p, Fi);
        if (ux < ps) {
            r.wo = reflect_n(r.wi, zplus());
            r.type = kPlS; r.comp = 0; r.pdf = ps;
            return col(M.specular_reflectance) * (Fi / ps);
        }
        float dp;
        r.wo = cos_hemi((ux - ps) / (1.0f - ps), uy, dp);
        r.pdf = (1.0f - ps) * dp;
        r.type = kPlD; r.comp = 1;
        const float Fo = fresnel_ext(r.wo.z, p.eta);
        return plastic_dw(M, p, Fi, Fo) / (1.0f - ps);
    }
    if (as) {
        r.wo = reflect_n(r.wi, zplus());
        r.type = kPlS; r.comp = 0; r.pdf = 1.0f;
        return col(M.specular_reflectance) * Fi;
    }
    float dp;
    r.wo = cos_hemi(ux, uy, dp);
    r.pdf = dp;
    r.type = kPlD; r.comp = 1;
    return plastic_dw(M, p, Fi, fresnel_ext(r.wo.z, p.eta));
}

// ---- rough plastic (rough_plastic_bsdf.h:45-185) -------------------------------
constexpr uint32_t kRpG = F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC;
LRT_D float rp_glossy_prob(const lrt_material& M, const float* ext, float w, float ci) {
    const float R = 1.0f - trans_eval(ext, ci);
    const float gl = R * w, di = (1.0f - R) * (1.0f - w);
    return gl / (gl + di);
}
LRT_D EP rough_plastic_ep(const Mats& S, const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    if (meas != M_SOLID) return ep0();
    const bool ag = (c.mask & kRpG) != 0 && (c.component == -1 || c.component == 0);
    const bool ad = (c.mask & kPlD) != 0 && (c.component == -1 || c.component == 1);
    if (r.wi.z <= 0.0f || r.wo.z <= 0.0f || (!ag && !ad)) return ep0();
    const float* ext = S.tables + 101 * (size_t)M.table_ext;
    const float eta = M.int_ior / M.ext_ior;
    const float inv_eta2 = 1.0f / (eta * eta);
    const float w = lum(M.specular_reflectance) / (lum(M.diffuse) + lum(M.specular_reflectance));
    const Ggx g = ggx(M.alpha);
    f3 f = splat(0.0f);
    float pg = 1.0f, pd = 1.0f;
    if (ag && ad) {
        pg = rp_glossy_prob(M, ext, w, r.wi.z);
        pd = 1.0f - pg;
    }
    float dens = 0.0f;
    if (ag) {
        const f3 h = normalize(r.wi + r.wo);
        const float d = ggx_D(g, h);
        const float F = fresnel_ext(dot(r.wi, h), eta);
        const float gg = ggx_G(g, r.wi, r.wo, h);
        f = f + col(M.specular_reflectance) * (F * d * gg / (4.0f * r.wi.z));
        const float j = 1.0f / (4.0f * dot(r.wo, h));
        dens += ggx_pdf(g, r.wi, h) * j * pg;
    }
    if (ad) {
        const float* in = S.tables + 101 * (size_t)M.table_int;
        const float t12 = trans_eval(ext, r.wi.z), t21 = trans_eval(ext, r.wo.z);
        const float fdr = 1.0f - trans_diffuse(in);
        f = f + scaled_diffuse(M, fdr) * ((r.wo.z / kPiF) * t12 * t21 * inv_eta2);
        dens += pd * (r.wo.z / kPiF);
    }
    return EP{f, dens};
}
LRT_D f3 rough_plastic_sample(const Mats& S, const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    clear(r);
    const bool ag = (c.mask & kRpG) != 0 && (c.component == -1 || c.component == 0);
    const bool ad = (c.mask & kPlD) != 0 && (c.component == -1 || c.component == 1);
    if ((!ag && !ad) || r.wi.z <= 0.0f) return splat(0.0f);
    float lx = ux, ly = uy;
    bool glossy = ag;
    if (ag && ad) {
        const float w = lum(M.specular_reflectance) / (lum(M.diffuse) + lum(M.specular_reflectance));
        const float pg = rp_glossy_prob(M, S.tables + 101 * (size_t)M.table_ext, w, r.wi.z);
        if (ly < pg) ly /= pg;
        else { ly = (ly - pg) / (1.0f - pg); glossy = false; }
    }
    if (glossy) {
        const f3 m = ggx_sample(ggx(M.alpha), r.wi, lx, ly);
        r.wo = reflect_n(r.wi, m);
        r.type = kRpG; r.comp = 0;
        if (r.wo.z <= 0.0f) return splat(0.0f);
    } else {
        float cp;
        r.wo = cos_hemi(lx, ly, cp);
        r.type = kPlD; r.comp = 1;
    }
    const EP e = rough_plastic_ep(S, M, c, r, M_SOLID);
    if (e.pdf == 0.0f) return splat(0.0f);
    r.pdf = e.pdf;
    return e.f / e.pdf;
}

// ---- base dispatch ----------------------------------------------------------
LRT_D EP base_ep(const Mats& S, int m, const BsdfCtx& c, const SRec& r, int meas) {
    const lrt_material& M = S.m[m];
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return diffuse_ep(M, c, r, meas);
        case LRT_BSDF_CONDUCTOR: return conductor_ep(M, c, r, meas);
        case LRT_BSDF_DIELECTRIC: return dielectric_ep(M, c, r, meas);
        case LRT_BSDF_ROUGH_DIELECTRIC: return rd_ep(M, c, r, meas);
        case LRT_BSDF_PLASTIC: return plastic_ep(M, c, r, meas);
        case LRT_BSDF_ROUGH_PLASTIC: return rough_plastic_ep(S, M, c, r, meas);
        default: return ep0();
    }
}
LRT_D f3 base_sample(const Mats& S, int m, const BsdfCtx& c, SRec& r, float ux, float uy) {
    const lrt_material& M = S.m[m];
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return diffuse_sample(M, c, r, ux, uy);
        case LRT_BSDF_CONDUCTOR: return conductor_sample(M, c, r, ux, uy);
        case LRT_BSDF_DIELECTRIC: return dielectric_sample(M, c, r, ux);
        case LRT_BSDF_ROUGH_DIELECTRIC: return rd_sample(M, c, r, ux, uy);
        case LRT_BSDF_PLASTIC: return plastic_sample(M, c, r, ux, uy);
        case LRT_BSDF_ROUGH_PLASTIC: return rough_plastic_sample(S, M, c, r, ux, uy);
        default: clear(r); return splat(0.0f);
    }
}

// ---- coating helpers (coating_layer.h) -----------------------------------------
struct Refr {
    f3 d;
    float R;
};
LRT_D Refr refract_in(f3 wi, float eta, float inv_eta) {
    float ct;
    const float R = fresnel_ext(fabsf(wi.z), ct, eta);
    return Refr{f3{inv_eta * wi.x, inv_eta * wi.y, -sgn(wi.z) * ct}, R};
}
LRT_D Refr refract_out(f3 wi, float eta, float inv_eta) {
    float ct;
    const float R = fresnel_ext(fabsf(wi.z), ct, inv_eta);
    return Refr{f3{eta * wi.x, eta * wi.y, -sgn(wi.z) * ct}, R};
}
LRT_D f3 refract_to(bool interior, f3 wi, float eta, float inv_eta) {
    const float s = interior ? inv_eta : eta;
    const float st2 = s * s * (wi.x * wi.x + wi.y * wi.y);
    if (st2 >= 1.0f) return splat(0.0f);
    const float ct = sqrtf(1.0f - st2);
    return f3{s * wi.x, s * wi.y, wi.z > 0.0f ? ct : -ct};
}
LRT_D float sa_factor(float inv_eta, f3 wo, f3 woi) { return inv_eta * inv_eta * wo.z / woi.z; }
LRT_D float coat_w(const lrt_material& M) {
    const f3 o = col(M.sigma_a) * (-2.0f * M.thickness);
    const float avg = (exp_rn(o.x) + exp_rn(o.y) + exp_rn(o.z)) / 3.0f;
    return 1.0f / (avg + 1.0f);
}
LRT_D f3 absorb(const lrt_material& M, f3 wi, f3 wo) {
    if (is_zero(col(M.sigma_a))) return splat(1.0f);
    const float l = (1.0f / fabsf(wi.z)) + (1.0f / fabsf(wo.z));
    const f3 e = col(M.sigma_a) * (-M.thickness * l);
    return f3{exp_rn(e.x), exp_rn(e.y), exp_rn(e.z)};
}
LRT_D float coat_prob(float R, float w) {
    const float s = R * w, n = (1.0f - R) * (1.0f - w);
    return s / (s + n);
}

// ---- smooth coating (coating_bsdf.h:45-177) ------------------------------------
constexpr uint32_t kCoS = F_SPECULAR | F_REFLECTIVE | F_DIELECTRIC;
LRT_D EP coating_ep(const Mats& S, const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    const int sc = base_count(S.m[M.nested]);
    const bool as = (c.mask & kCoS) != 0 && (c.component == -1 || c.component == sc);
    const bool an = (c.mask & base_flags(S.m[M.nested])) != 0 && (c.component == -1 || c.component < sc);
    const float eta = M.int_ior / M.ext_ior, inv_eta = M.ext_ior / M.int_ior;
    const Refr in = refract_in(r.wi, eta, inv_eta);
    if (meas == M_DISCRETE && as && aligned(f3{-r.wi.x, -r.wi.y, r.wi.z}, r.wo))
        return EP{col(M.specular_reflectance) * fresnel_ext(fabsf(r.wi.z), eta), an ? coat_prob(in.R, coat_w(M)) : 1.0f};
    if (!an) return ep0();
    const Refr out = refract_in(r.wo, eta, inv_eta);
    if (in.R == 1.0f || out.R == 1.0f) return ep0();
    SRec inner = r;
    inner.wi = in.d;
    inner.wo = out.d;
    const EP b = base_ep(S, M.nested, c, inner, meas);
    f3 f = b.f * ((1.0f - in.R) * (1.0f - out.R));
    f = f * absorb(M, in.d, out.d);
    float pdf = b.pdf;
    if (meas == M_SOLID) {
        const float k = sa_factor(inv_eta, r.wo, out.d);
        f = f * k;
        pdf *= k;
    }
    if (as) pdf *= 1.0f - coat_prob(in.R, coat_w(M));
    return EP{f, pdf};
}
LRT_D f3 coating_sample(const Mats& S, const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    clear(r);
    const int sc = base_count(S.m[M.nested]);
    const bool as = (c.mask & kCoS) != 0 && (c.component == -1 || c.component == sc);
    const bool an = (c.mask & base_flags(S.m[M.nested])) != 0 && (c.component == -1 || c.component < sc);
    if (!as && !an) return splat(0.0f);
    const float eta = M.int_ior / M.ext_ior, inv_eta = M.ext_ior / M.int_ior;
    const Refr in = refract_in(r.wi, eta, inv_eta);
    const float ps = coat_prob(in.R, coat_w(M));
    float lx = ux, ly = uy;
    bool spec = as;
    if (as && an) {
        if (lx < ps) lx /= ps;
        else { lx = (lx - ps) / (1.0f - ps); spec = false; }
    }
    if (spec) {
        r.wo = f3{-r.wi.x, -r.wi.y, r.wi.z};
        r.type = kCoS;
        r.comp = sc;
        r.pdf = an ? ps : 1.0f;
        return col(M.specular_reflectance) * (in.R / r.pdf);
    }
    if (in.R == 1.0f) return splat(0.0f);
    const f3 wi_outer = r.wi;
    r.wi = in.d;
    f3 w = base_sample(S, M.nested, c, r, lx, ly);
    const f3 wo_inner = r.wo;
    r.wi = wi_outer;
    if (is_zero(w)) return splat(0.0f);
    const Refr out = refract_out(wo_inner, eta, inv_eta);
    r.wo = out.d;
    if (out.R == 1.0f) return splat(0.0f);
    w = w * absorb(M, in.d, wo_inner);
    w = w * ((1.0f - in.R) * (1.0f - out.R));
    if (as) {
        r.pdf *= 1.0f - ps;
        w = w / (1.0f - ps);
    }
    if ((r.type & F_SPECULAR) == 0) r.pdf *= sa_factor(inv_eta, r.wo, wo_inner);
    return w;
}

// ---- rough coating (rough_coating_bsdf.h:51-217) -------------------------------
constexpr uint32_t kRcS = F_GLOSSY | F_REFLECTIVE | F_DIELECTRIC;
LRT_D EP rough_coating_ep(const Mats& S, const lrt_material& M, const BsdfCtx& c, const SRec& r, int meas) {
    const int sc = base_count(S.m[M.nested]);
    const bool an = (c.mask & base_flags(S.m[M.nested])) != 0 && (c.component == -1 || c.component < sc);
    const bool as = (c.mask & kRcS) != 0 && (c.component == -1 || c.component == sc) && meas == M_SOLID;
    const float* tr = S.tables + 101 * (size_t)M.table_ext;
    const float eta = M.int_ior / M.ext_ior, inv_eta = M.ext_ior / M.int_ior;
    const Ggx g = ggx(M.alpha);
    float ps = 1.0f, pn = 1.0f;
    if (as && an) {
        ps = coat_prob(1.0f - trans_eval(tr, fabsf(r.wi.z)), coat_w(M));
        pn = 1.0f - ps;
    }
    f3 f = splat(0.0f);
    float dens = 0.0f;
    if (as && r.wo.z * r.wi.z > 0.0f) {
        f3 h = r.wi + r.wo;
        if (len2(h) > 1e-12f) {
            h = normalize(h) * sgn(r.wo.z);
            const float d = ggx_D(g, h);
            const float F = fresnel_ext(fabsf(dot(r.wi, h)), eta);
            const float gg = ggx_G(g, r.wi, r.wo, h);
            f = f + col(M.specular_reflectance) * (F * d * gg / (4.0f * fabsf(r.wi.z)));
            const float j = 1.0f / (4.0f * fabsf(dot(r.wo, h)));
            dens += ggx_pdf(g, r.wi, h) * j * ps;
        }
    }
    if (an) {
        SRec inner = r;
        inner.wi = refract_to(true, r.wi, eta, inv_eta);
        inner.wo = refract_to(true, r.wo, eta, inv_eta);
        if (!is_zero(inner.wi) && !is_zero(inner.wo)) {
            const EP b = base_ep(S, M.nested, c, inner, meas);
            f3 nf = b.f * trans_eval(tr, fabsf(r.wi.z)) * trans_eval(tr, fabsf(r.wo.z));
            nf = nf * absorb(M, inner.wi, inner.wo);
            float np = b.pdf;
            if (meas == M_SOLID) {
                const float k = sa_factor(inv_eta, r.wo, inner.wo);
                nf = nf * k;
                np *= k;
            }
            f = f + nf;
            dens += np * pn;
        }
    }
    return EP{f, dens};
}
LRT_D f3 rough_coating_sample(const Mats& S, const lrt_material& M, const BsdfCtx& c, SRec& r, float ux, float uy) {
    clear(r);
    const int sc = base_count(S.m[M.nested]);
    const bool as = (c.mask & kRcS) != 0 && (c.component == -1 || c.component == sc);
    const bool an = (c.mask & base_flags(S.m[M.nested])) != 0 && (c.component == -1 || c.component < sc);
    if (!as && !an) return splat(0.0f);
    const float eta = M.int_ior / M.ext_ior, inv_eta = M.ext_ior / M.int_ior;
    float lx = ux, ly = uy;
    bool spec = as;
    if (as && an) {
        const float ps = coat_prob(1.0f - trans_eval(S.tables + 101 * (size_t)M.table_ext, fabsf(r.wi.z)), coat_w(M));
        if (ly < ps) ly /= ps;
        else { ly = (ly - ps) / (1.0f - ps); spec = false; }
    }
    if (spec) {
        const f3 m = ggx_sample(ggx(M.alpha), r.wi, lx, ly);
        r.wo = reflect_n(r.wi, m);
        r.type = kRcS; r.comp = sc; r.eta = 1.0f;
        if (r.wo.z * r.wi.z <= 0.0f) return splat(0.0f);
    } else {
        const f3 wi_outer = r.wi;
        r.wi = refract_to(true, wi_outer, eta, inv_eta);
        const f3 w = base_sample(S, M.nested, c, r, lx, ly);
        r.wi = wi_outer;
        if (is_zero(w)) return splat(0.0f);
        r.wo = refract_to(false, r.wo, eta, inv_eta);
        if (is_zero(r.wo)) return splat(0.0f);
    }
    const int meas = (r.type & F_SPECULAR) ? M_DISCRETE : M_SOLID;
    const EP e = rough_coating_ep(S, M, c, r, meas);
    if (e.pdf == 0.0f) return splat(0.0f);
    r.pdf = e.pdf;
    return e.f / e.pdf;
}

// ---- top-level dispatch -------------------------------------------------------
LRT_D EP bsdf_ep(const Mats& S, int m, const BsdfCtx& c, const SRec& r, int meas) {
    const lrt_material& M = S.m[m];
    if (M.type == LRT_BSDF_COATING) return coating_ep(S, M, c, r, meas);
    if (M.type == LRT_BSDF_ROUGH_COATING) return rough_coating_ep(S, M, c, r, meas);
    return base_ep(S, m, c, r, meas);
}
LRT_D f3 bsdf_sample(const Mats& S, int m, const BsdfCtx& c, SRec& r, float ux, float uy) {
    const lrt_material& M = S.m[m];
    if (M.type == LRT_BSDF_COATING) return coating_sample(S, M, c, r, ux, uy);
    if (M.type == LRT_BSDF_ROUGH_COATING) return rough_coating_sample(S, M, c, r, ux, uy);
    return base_sample(S, m, c, r, ux, uy);
}

}  // namespace lrt
