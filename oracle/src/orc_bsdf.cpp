// Oracle (test infrastructure only): the reference's BSDFs restated with the
// same floating-point operation order.  Cited per function.
#include "orc.h"

namespace orc {
namespace {

constexpr uint32_t DIFFUSE = LRT_FLAG_DIFFUSE, SPECULAR = LRT_FLAG_SPECULAR, REFLECTIVE = LRT_FLAG_REFLECTIVE,
                   GLOSSY = LRT_FLAG_GLOSSY, METALLIC = LRT_FLAG_METALLIC, DIELECTRIC = LRT_FLAG_DIELECTRIC,
                   REFRACTIVE = LRT_FLAG_REFRACTIVE;
constexpr int SOLID = LRT_MEASURE_SOLID_ANGLE, DISCRETE = LRT_MEASURE_DISCRETE;

inline int measure_of(uint32_t t) { return (t & SPECULAR) ? DISCRETE : SOLID; }  // bsdf.h:68-70
inline v3 reflect(const v3& w, const v3& n) { return 2.0f * dot(w, n) * n - w; }  // frame.h:14-16
inline v3 refract(const v3& wi, const v3& n, float ct, float ei, float et) {      // frame.h:20-24
    const float s = ei / et;
    return n * (s * dot(wi, n) - ct) - wi * s;
}
inline bool delta_aligned(const v3& a, const v3& b) { return std::abs(dot(a, b) - 1.0f) < kDeltaEps; }
inline float luminance(const float* c) { return 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]; }
inline v3 col(const float* c) { return v3(c[0], c[1], c[2]); }
inline float signum(float v) { return (float)((v > 0.0f) - (v < 0.0f)); }

// ---- GGX (microfacet.h:10-80) ----------------------------------------------
struct ggx {
    float a;
    explicit ggx(float alpha) : a(std::max(alpha, 0.0001f)) {}
    float D(const v3& m) const {
        if (m.z <= 0.0f) return 0.0f;
        const float a2 = a * a;
        const float t = (m.x * m.x + m.y * m.y) / a2 + m.z * m.z;
        return 1.0f / (kPi * a2 * t * t);
    }
    float G1(const v3& v, const v3& m) const {
        const float cv = std::abs(v.z);
        if (cv < 1e-6f || dot(v, m) * v.z <= 0.0f) return 0.0f;
        const float t = (a * a) * (v.x * v.x + v.y * v.y) / (v.z * v.z);
        const float lambda = 0.5f * (-1.0f + std::sqrt(1.0f + t));
        return 1.0f / (1.0f + lambda);
    }
    float G(const v3& wi, const v3& wo, const v3& m) const { return G1(wi, m) * G1(wo, m); }
    float pdf(const v3& wi, const v3& m) const {
        if (wi.z <= 1e-6f) return 0.0f;
        return G1(wi, m) * std::max(0.0f, dot(wi, m)) * D(m) / wi.z;
    }
    v3 sample(const v3& wi, float ux, float uy) const {
        const v3 vh = normalize(v3(a * wi.x, a * wi.y, wi.z));
        const float lensq = vh.x * vh.x + vh.y * vh.y;
        const v3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) / std::sqrt(lensq) : v3(1.0f, 0.0f, 0.0f);
        const v3 t2 = cross(vh, t1);
        const float r = std::sqrt(ux);
        const float phi = 2.0f * kPi * uy;
        const float dx = r * std::cos(phi);
        float dy = r * std::sin(phi);
        const float s = 0.5f * (1.0f + vh.z);
        dy = (1.0f - s) * std::sqrt(std::max(0.0f, 1.0f - dx * dx)) + s * dy;
        const v3 nh = dx * t1 + dy * t2 + std::sqrt(std::max(0.0f, 1.0f - dx * dx - dy * dy)) * vh;
        return normalize(v3(a * nh.x, a * nh.y, std::max(0.0f, nh.z)));
    }
};

// ---- Fresnel (fresnel.h) ---------------------------------------------------
float fresnel_conductor1(float ci, float eta, float k) {  // fresnel.h:43-56
    ci = std::clamp(ci, 0.0f, 1.0f);
    const std::complex<float> ior(eta, k);
    const float s2i = 1.0f - ci * ci;
    const auto s2t = s2i / (ior * ior);
    const auto ct = std::sqrt(std::complex<float>(1.0f) - s2t);
    const auto rpa = (ior * ci - ct) / (ior * ci + ct);
    const auto rpe = (ci - ior * ct) / (ci + ior * ct);
    return 0.5f * (std::norm(rpa) + std::norm(rpe));
}
v3 fresnel_conductor(float ci, const float* eta, const float* k) {  // fresnel.h:58-64
    return v3(fresnel_conductor1(ci, eta[0], k[0]), fresnel_conductor1(ci, eta[1], k[1]),
              fresnel_conductor1(ci, eta[2], k[2]));
}
struct fres { float R = 1.0f, ct = 0.0f; };
fres fresnel_dielectric(float ci, float ei, float et) {  // fresnel.h:72-95
    fres r;
    if (ei <= 0.0f || et <= 0.0f) return r;
    ci = std::clamp(ci, 0.0f, 1.0f);
    const float s2i = 1.0f - ci * ci;
    const float s = ei / et;
    const float s2t = s * s * s2i;
    if (s2t > 1.0f) return r;
    r.ct = std::sqrt(1.0f - s2t);
    const float rs = (ei * ci - et * r.ct) / (ei * ci + et * r.ct);
    const float rp = (et * ci - ei * r.ct) / (et * ci + ei * r.ct);
    r.R = 0.5f * (rs * rs + rp * rp);
    return r;
}
float fresnel_ext(float ci, float& ct_out, float eta) {  // fresnel.h:99-118
    if (eta == 1.0f) { ct_out = -ci; return 0.0f; }
    const float s = ci > 0.0f ? (1.0f / eta) : eta;
    const float ct2 = 1.0f - (1.0f - ci * ci) * (s * s);
    if (ct2 <= 0.0f) { ct_out = 0.0f; return 1.0f; }
    const float ai = std::abs(ci);
    const float at = std::sqrt(ct2);
    const float rs = (ai - eta * at) / (ai + eta * at);
    const float rp = (eta * ai - at) / (eta * ai + at);
    ct_out = ci > 0.0f ? -at : at;
    return 0.5f * (rs * rs + rp * rp);
}
float fresnel_ext(float ci, float eta) { float t; return fresnel_ext(ci, t, eta); }

v3 cosine_hemisphere(float ux, float uy, float& pdf) {  // plastic_bsdf.h:175-181
    const float ct = std::sqrt(ux);
    const float st = std::sqrt(1.0f - ux);
    const float phi = 2.0f * kPi * uy;
    pdf = ct / kPi;
    return v3(std::cos(phi) * st, std::sin(phi) * st, ct);
}

// ---- Catmull-Rom 1D and the rough-transmittance slice lookup -------------
// catmull_rom.h:10-34, rough_transmittance.h:17-29 (eval / eval_diffuse)
float cubic1(float x, const float* v, size_t n, float lo, float hi) {
    if (!(x >= lo && x <= hi)) return 0.0f;
    float t = ((x - lo) * (float)(n - 1)) / (hi - lo);
    size_t k = (size_t)t;
    if (k > n - 2) k = n - 2;
    const float f0 = v[k], f1 = v[k + 1];
    const float d0 = k > 0 ? 0.5f * (v[k + 1] - v[k - 1]) : (v[k + 1] - v[k]);
    const float d1 = (k + 2 < n) ? 0.5f * (v[k + 2] - v[k]) : (v[k + 1] - v[k]);
    t -= (float)k;
    const float t2 = t * t, t3 = t2 * t;
    return (2.0f * t3 - 3.0f * t2 + 1.0f) * f0 + (-2.0f * t3 + 3.0f * t2) * f1 + (t3 - 2.0f * t2 + t) * d0 +
           (t3 - t2) * d1;
}
float trans_eval(const float* table, float ct) {
    if (!(ct >= 0.0f)) return 0.0f;
    const float w = std::pow(ct, 0.25f);
    return std::clamp(cubic1(w, table, 100, 0.0f, 1.0f), 0.0f, 1.0f);
}
float trans_diffuse(const float* table) { return std::clamp(table[100], 0.0f, 1.0f); }

// ---- coating helpers (coating_layer.h) -----------------------------------
struct refr { v3 d; float R = 1.0f; };
refr refract_in(const v3& wi, float eta, float inv_eta) {  // :16-22
    refr r;
    float ct = 0.0f;
    r.R = fresnel_ext(std::abs(wi.z), ct, eta);
    r.d = v3(inv_eta * wi.x, inv_eta * wi.y, -signum(wi.z) * ct);
    return r;
}
refr refract_out(const v3& wi, float eta, float inv_eta) {  // :25-31
    refr r;
    float ct = 0.0f;
    r.R = fresnel_ext(std::abs(wi.z), ct, inv_eta);
    r.d = v3(eta * wi.x, eta * wi.y, -signum(wi.z) * ct);
    return r;
}
v3 refract_to(bool interior, const v3& wi, float eta, float inv_eta) {  // :41-50
    const float s = interior ? inv_eta : eta;
    const float st2 = s * s * (wi.x * wi.x + wi.y * wi.y);
    if (st2 >= 1.0f) return v3(0.0f);
    const float ct = std::sqrt(1.0f - st2);
    return v3(s * wi.x, s * wi.y, wi.z > 0.0f ? ct : -ct);
}
float solid_angle_factor(float inv_eta, const v3& wo, const v3& woi) { return inv_eta * inv_eta * wo.z / woi.z; }
float coat_weight(const float* sa, float th) {  // :57-62
    const v3 o = col(sa) * (-2.0f * th);
    const float avg = (std::exp(o.x) + std::exp(o.y) + std::exp(o.z)) / 3.0f;
    return 1.0f / (avg + 1.0f);
}
v3 absorption(const float* sa, float th, const v3& wi, const v3& wo) {  // :65-73
    if (col(sa).is_zero()) return v3(1.0f);
    const float len = (1.0f / std::abs(wi.z)) + (1.0f / std::abs(wo.z));
    const v3 e = col(sa) * (-th * len);
    return v3(std::exp(e.x), std::exp(e.y), std::exp(e.z));
}
float coat_prob(float R, float w) {  // :76-81
    const float s = R * w;
    const float n = (1.0f - R) * (1.0f - w);
    return s / (s + n);
}

struct env {
    const std::vector<lrt_material>& mats;
    const float* tables;
    const float* table(int i) const { return tables + (size_t)i * 101; }
};

uint32_t flags_of(const env& E, int m) {
    const lrt_material& M = E.mats[(size_t)m];
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return DIFFUSE | REFLECTIVE;
        case LRT_BSDF_CONDUCTOR: return M.alpha <= 0.0f ? (SPECULAR | REFLECTIVE | METALLIC) : (GLOSSY | REFLECTIVE | METALLIC);
        case LRT_BSDF_DIELECTRIC: return (SPECULAR | REFLECTIVE | DIELECTRIC) | (SPECULAR | REFRACTIVE | DIELECTRIC);
        case LRT_BSDF_ROUGH_DIELECTRIC: return GLOSSY | REFLECTIVE | REFRACTIVE | DIELECTRIC;
        case LRT_BSDF_PLASTIC: return (SPECULAR | REFLECTIVE | DIELECTRIC) | (DIFFUSE | REFLECTIVE);
        case LRT_BSDF_ROUGH_PLASTIC: return (GLOSSY | REFLECTIVE | DIELECTRIC) | (DIFFUSE | REFLECTIVE);
        case LRT_BSDF_COATING: return flags_of(E, M.nested) | (SPECULAR | REFLECTIVE | DIELECTRIC);
        case LRT_BSDF_ROUGH_COATING: return flags_of(E, M.nested) | (GLOSSY | REFLECTIVE | DIELECTRIC);
        default: return 0;
    }
}
int count_of(const env& E, int m) {
    const lrt_material& M = E.mats[(size_t)m];
    switch (M.type) {
        case LRT_BSDF_DIELECTRIC: case LRT_BSDF_PLASTIC: case LRT_BSDF_ROUGH_PLASTIC: return 2;
        case LRT_BSDF_COATING: case LRT_BSDF_ROUGH_COATING: return count_of(E, M.nested) + 1;
        default: return 1;
    }
}

v3 eval_(const env& E, int m, const bsdf_ctx& ctx, const bsdf_rec& rec, int meas);
float pdf_(const env& E, int m, const bsdf_ctx& ctx, const bsdf_rec& rec, int meas);
v3 sample_(const env& E, int m, const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy);

// ---- diffuse (diffuse_bsdf.h:18-72) ---------------------------------------
v3 diffuse_sample(const lrt_material& M, const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) {
    const uint32_t lobe = DIFFUSE | REFLECTIVE;
    rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
    if (!(ctx.type_mask & lobe)) return v3(0.0f);
    if (ctx.component != -1 && ctx.component != 0) return v3(0.0f);
    if (rec.wi.z <= 0.0f) return v3(0.0f);
    const float ct = std::sqrt(ux);
    const float st = std::sqrt(1.0f - ux);
    const float phi = 2.0f * kPi * uy;
    rec.wo = v3(std::cos(phi) * st, std::sin(phi) * st, ct);
    rec.pdf = ct / kPi;
    rec.sampled_type = lobe;
    rec.sampled_component = 0;
    return col(M.diffuse);
}
bool diffuse_ok(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) {
    if (meas != SOLID) return false;
    if (!(ctx.type_mask & (DIFFUSE | REFLECTIVE))) return false;
    if (ctx.component != -1 && ctx.component != 0) return false;
    return !(rec.wi.z <= 0.0f || rec.wo.z <= 0.0f);
}

// ---- rough conductor (rough_conductor_bsdf.h:28-128) -----------------------
struct ep { v3 f; float pdf; };
ep conductor_eval_pdf(const lrt_material& M, const v3& wi, const v3& wo) {
    if (wi.z <= 1e-6f || wo.z <= 1e-6f) return {v3(0.0f), 0.0f};
    const v3 h = wi + wo;
    if (h.length_squared() <= 1e-12f) return {v3(0.0f), 0.0f};
    const v3 m = normalize(h);
    const float wim = dot(wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return {v3(0.0f), 0.0f};
    const ggx g(M.alpha);
    const float d = g.D(m);
    const float gg = g.G(wi, wo, m);
    const v3 F = fresnel_conductor(wim, M.cond_eta, M.cond_k) * col(M.specular_reflectance);
    const v3 value = F * (d * gg / (4.0f * wi.z));
    const float pm = g.pdf(wi, m);
    return {value, pm / (4.0f * wim)};
}
uint32_t conductor_lobe(const lrt_material& M) {
    return M.alpha <= 0.0f ? (SPECULAR | REFLECTIVE | METALLIC) : (GLOSSY | REFLECTIVE | METALLIC);
}
v3 conductor_sample(const lrt_material& M, const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) {
    const uint32_t lobe = conductor_lobe(M);
    rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
    if ((ctx.type_mask & lobe) == 0) return v3(0.0f);
    if (ctx.component != -1 && ctx.component != 0) return v3(0.0f);
    if (rec.wi.z < 0.0f) return v3(0.0f);
    if (M.alpha <= 0.0f) {
        rec.wo = v3(-rec.wi.x, -rec.wi.y, rec.wi.z);
        rec.sampled_type = lobe;
        rec.sampled_component = 0;
        rec.pdf = 1.0f;
        return fresnel_conductor(rec.wi.z, M.cond_eta, M.cond_k) * col(M.specular_reflectance);
    }
    const ggx g(M.alpha);
    const v3 m = g.sample(rec.wi, ux, uy);
    const float wim = dot(rec.wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return v3(0.0f);
    rec.wo = reflect(rec.wi, m);
    if (rec.wo.z <= 0.0f) return v3(0.0f);
    const ep e = conductor_eval_pdf(M, rec.wi, rec.wo);
    if (e.pdf <= 0.0f) return v3(0.0f);
    rec.pdf = e.pdf;
    rec.sampled_type = lobe;
    rec.sampled_component = 0;
    return e.f / e.pdf;
}

// ---- smooth dielectric (dielectric_bsdf.h:31-157) --------------------------
constexpr uint32_t kDR = SPECULAR | REFLECTIVE | DIELECTRIC, kDT = SPECULAR | REFRACTIVE | DIELECTRIC;
v3 dielectric_sample(const lrt_material& M, const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float) {
    rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
    if (rec.wi.z <= 0.0f) return v3(0.0f);
    const bool ar = (ctx.type_mask & kDR) != 0 && (ctx.component == -1 || ctx.component == 0);
    const bool at = (ctx.type_mask & kDT) != 0 && (ctx.component == -1 || ctx.component == 1);
    if (!ar && !at) return v3(0.0f);
    const float ei = rec.front_face ? M.ext_ior : M.int_ior;
    const float et = rec.front_face ? M.int_ior : M.ext_ior;
    const fres F = fresnel_dielectric(rec.wi.z, ei, et);
    if (ar && (!at || ux < F.R)) {
        rec.wo = reflect(rec.wi, v3(0.0f, 0.0f, 1.0f));
        rec.sampled_type = kDR;
        rec.sampled_component = 0;
        rec.pdf = at ? F.R : 1.0f;
        return at ? col(M.specular_reflectance) : col(M.specular_reflectance) * F.R;
    }
    if (F.R >= 1.0f) return v3(0.0f);
    rec.wo = refract(rec.wi, v3(0.0f, 0.0f, 1.0f), F.ct, ei, et);
    rec.eta = ei / et;
    rec.sampled_type = kDT;
    rec.sampled_component = 1;
    rec.pdf = ar ? (1.0f - F.R) : 1.0f;
    v3 w = col(M.specular_transmittance);
    if (!ar) w = w * (1.0f - F.R);
    if (ctx.mode == LRT_MODE_RADIANCE) w = w * (rec.eta * rec.eta);
    return w;
}
// shared body of eval (:63-98) and pdf (:100-131)
void dielectric_eval_pdf(const lrt_material& M, const bsdf_ctx& ctx, const bsdf_rec& rec, int meas, v3& f, float& pdf) {
    f = v3(0.0f); pdf = 0.0f;
    if (meas != DISCRETE || rec.wi.z <= 0.0f) return;
    const bool ar = (ctx.type_mask & kDR) != 0 && (ctx.component == -1 || ctx.component == 0);
    const bool at = (ctx.type_mask & kDT) != 0 && (ctx.component == -1 || ctx.component == 1);
    const float ei = rec.front_face ? M.ext_ior : M.int_ior;
    const float et = rec.front_face ? M.int_ior : M.ext_ior;
    const fres F = fresnel_dielectric(rec.wi.z, ei, et);
    if (rec.wi.z * rec.wo.z >= 0.0f) {
        if (!ar || !delta_aligned(reflect(rec.wi, v3(0.0f, 0.0f, 1.0f)), rec.wo)) return;
        f = col(M.specular_reflectance) * F.R;
        pdf = at ? F.R : 1.0f;
        return;
    }
    if (!at || F.R >= 1.0f) return;
    const v3 tr = refract(rec.wi, v3(0.0f, 0.0f, 1.0f), F.ct, ei, et);
    if (!delta_aligned(tr, rec.wo)) return;
    v3 value = col(M.specular_transmittance) * (1.0f - F.R);
    if (ctx.mode == LRT_MODE_RADIANCE) {
        const float eta = ei / et;
        value = value * (eta * eta);
    }
    f = value;
    pdf = ar ? (1.0f - F.R) : 1.0f;
}

// ---- rough dielectric (rough_dielectric_bsdf.h:28-197) ---------------------
bool rd_refl_ok(const bsdf_ctx& c) {
    return (c.type_mask & (GLOSSY | REFLECTIVE | DIELECTRIC)) != 0 && (c.component == -1 || c.component == 0);
}
bool rd_trans_ok(const bsdf_ctx& c) {
    return (c.type_mask & (GLOSSY | REFRACTIVE | DIELECTRIC)) != 0 && (c.component == -1 || c.component == 1);
}
ep rd_reflection(const lrt_material& M, const bsdf_ctx& ctx, const v3& wi, const v3& wo, float ei, float et, bool both) {
    if (!rd_refl_ok(ctx) || wi.z <= 1e-6f || wo.z <= 1e-6f) return {v3(0.0f), 0.0f};
    v3 m = wi + wo;
    if (m.length_squared() <= 1e-12f) return {v3(0.0f), 0.0f};
    m = normalize(m);
    if (m.z < 0.0f) m = -m;
    const float wim = dot(wi, m);
    if (wim <= 0.0f) return {v3(0.0f), 0.0f};
    const ggx g(M.alpha);
    const float F = fresnel_dielectric(wim, ei, et).R;
    const float d = g.D(m);
    const float gg = g.G(wi, wo, m);
    const float pm = g.pdf(wi, m);
    const float lobe = both ? F : 1.0f;
    const float value = F * d * gg / (4.0f * wi.z);
    return {v3(value), pm * lobe / (4.0f * wim)};
}
ep rd_transmission(const lrt_material& M, const bsdf_ctx& ctx, const v3& wi, const v3& wo, float ei, float et, bool both) {
    if (!rd_trans_ok(ctx) || wi.z <= 1e-6f || wo.z >= -1e-6f || et <= 0.0f) return {v3(0.0f), 0.0f};
    v3 m = -(ei * wi + et * wo);
    if (m.length_squared() <= 1e-12f) return {v3(0.0f), 0.0f};
    m = normalize(m);
    if (m.z < 0.0f) m = -m;
    const float wim = dot(wi, m);
    const float wom = dot(wo, m);
    if (wim <= 0.0f) return {v3(0.0f), 0.0f};
    const float F = fresnel_dielectric(wim, ei, et).R;
    const float denom = ei * wim + et * wom;
    if (F >= 1.0f || denom * denom <= 1e-12f) return {v3(0.0f), 0.0f};
    const ggx g(M.alpha);
    const float d = g.D(m);
    const float gg = g.G(wi, wo, m);
    const float pm = g.pdf(wi, m);
    const float awom = std::abs(wom);
    const float es = ctx.mode == LRT_MODE_RADIANCE ? ei : et;
    const float lobe = both ? (1.0f - F) : 1.0f;
    const float value = (1.0f - F) * d * gg * wim * awom * es * es / (wi.z * denom * denom);
    const float pdf = pm * lobe * (et * et) * awom / (denom * denom);
    return {v3(value), pdf};
}
ep rd_eval_pdf(const lrt_material& M, const bsdf_ctx& ctx, const bsdf_rec& rec) {
    if (rec.wi.z <= 1e-6f || std::abs(rec.wo.z) <= 1e-6f) return {v3(0.0f), 0.0f};
    const float ei = rec.front_face ? M.ext_ior : M.int_ior;
    const float et = rec.front_face ? M.int_ior : M.ext_ior;
    const bool both = rd_refl_ok(ctx) && rd_trans_ok(ctx);
    if (rec.wi.z * rec.wo.z > 0.0f) return rd_reflection(M, ctx, rec.wi, rec.wo, ei, et, both);
    return rd_transmission(M, ctx, rec.wi, rec.wo, ei, et, both);
}
v3 rd_sample(const lrt_material& M, const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) {
    rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
    if (rec.wi.z <= 0.0f) return v3(0.0f);
    const bool ar = rd_refl_ok(ctx), at = rd_trans_ok(ctx);
    if (!ar && !at) return v3(0.0f);
    const float ei = rec.front_face ? M.ext_ior : M.int_ior;
    const float et = rec.front_face ? M.int_ior : M.ext_ior;
    const ggx g(M.alpha);
    const v3 m = g.sample(rec.wi, ux, uy);
    const float wim = dot(rec.wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return v3(0.0f);
    const fres F = fresnel_dielectric(wim, ei, et);
    const bool both = ar && at;
    if (ar && (!at || ctx.lobe_sample < F.R)) {
        rec.wo = reflect(rec.wi, m);
        if (rec.wo.z <= 0.0f) return v3(0.0f);
        const ep l = rd_reflection(M, ctx, rec.wi, rec.wo, ei, et, both);
        if (l.pdf <= 0.0f) return v3(0.0f);
        rec.pdf = l.pdf;
        rec.sampled_type = GLOSSY | REFLECTIVE | DIELECTRIC;
        rec.sampled_component = 0;
        return col(M.specular_reflectance) * (l.f / l.pdf);
    }
    if (F.ct <= 0.0f) return v3(0.0f);
    rec.wo = refract(rec.wi, m, F.ct, ei, et);
    if (rec.wo.z >= 0.0f) return v3(0.0f);
    rec.eta = ei / et;
    const ep l = rd_transmission(M, ctx, rec.wi, rec.wo, ei, et, both);
    if (l.pdf <= 0.0f) return v3(0.0f);
    rec.pdf = l.pdf;
    rec.sampled_type = GLOSSY | REFRACTIVE | DIELECTRIC;
    rec.sampled_component = 1;
    return col(M.specular_transmittance) * (l.f / l.pdf);
}

// ---- smooth plastic (plastic_bsdf.h:41-191) --------------------------------
constexpr uint32_t kPS = SPECULAR | REFLECTIVE | DIELECTRIC, kPD = DIFFUSE | REFLECTIVE;
struct plastic {
    const lrt_material& M;
    float eta, inv_eta2, w_spec;
    explicit plastic(const lrt_material& m) : M(m) {
        eta = m.int_ior / m.ext_ior;
        inv_eta2 = 1.0f / (eta * eta);
        w_spec = luminance(m.specular_reflectance) / (luminance(m.diffuse) + luminance(m.specular_reflectance));
    }
    bool spec_ok(const bsdf_ctx& c) const { return (c.type_mask & kPS) != 0 && (c.component == -1 || c.component == 0); }
    bool diff_ok(const bsdf_ctx& c) const { return (c.type_mask & kPD) != 0 && (c.component == -1 || c.component == 1); }
    float lobe_prob(float Fi) const {
        const float s = Fi * w_spec;
        const float d = (1.0f - Fi) * (1.0f - w_spec);
        return s / (s + d);
    }
    v3 scaled() const {
        const v3 kd = col(M.diffuse);
        if (M.nonlinear)
            return v3(kd.x / (1.0f - kd.x * M.fdr_int), kd.y / (1.0f - kd.y * M.fdr_int), kd.z / (1.0f - kd.z * M.fdr_int));
        return kd / (1.0f - M.fdr_int);
    }
    v3 dweight(float Fi, float Fo) const { return scaled() * (inv_eta2 * (1.0f - Fi) * (1.0f - Fo)); }
    v3 sample(const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) const {
        rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
        const bool as = spec_ok(ctx), ad = diff_ok(ctx);
        if ((!as && !ad) || rec.wi.z <= 0.0f) return v3(0.0f);
        const float Fi = fresnel_ext(rec.wi.z, eta);
        if (as && ad) {
            const float ps = lobe_prob(Fi);
            if (ux < ps) {
                rec.wo = reflect(rec.wi, v3(0.0f, 0.0f, 1.0f));
                rec.sampled_type = kPS; rec.sampled_component = 0; rec.pdf = ps;
                return col(M.specular_reflectance) * (Fi / ps);
            }
            float dp = 0.0f;
            rec.wo = cosine_hemisphere((ux - ps) / (1.0f - ps), uy, dp);
            rec.pdf = (1.0f - ps) * dp;
            rec.sampled_type = kPD; rec.sampled_component = 1;
            const float Fo = fresnel_ext(rec.wo.z, eta);
            return dweight(Fi, Fo) / (1.0f - ps);
        }
        if (as) {
            rec.wo = reflect(rec.wi, v3(0.0f, 0.0f, 1.0f));
            rec.sampled_type = kPS; rec.sampled_component = 0; rec.pdf = 1.0f;
            return col(M.specular_reflectance) * Fi;
        }
        float dp = 0.0f;
        rec.wo = cosine_hemisphere(ux, uy, dp);
        rec.pdf = dp;
        rec.sampled_type = kPD; rec.sampled_component = 1;
        const float Fo = fresnel_ext(rec.wo.z, eta);
        return dweight(Fi, Fo);
    }
    v3 eval(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        const bool as = spec_ok(ctx), ad = diff_ok(ctx);
        if (rec.wi.z <= 0.0f) return v3(0.0f);
        if (meas == DISCRETE) {
            if (!as || !delta_aligned(reflect(rec.wi, v3(0.0f, 0.0f, 1.0f)), rec.wo)) return v3(0.0f);
            return col(M.specular_reflectance) * fresnel_ext(rec.wi.z, eta);
        }
        if (meas != SOLID || !ad || rec.wo.z <= 0.0f) return v3(0.0f);
        const float Fi = fresnel_ext(rec.wi.z, eta);
        const float Fo = fresnel_ext(rec.wo.z, eta);
        return dweight(Fi, Fo) * (rec.wo.z / kPi);
    }
    float pdf(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        const bool as = spec_ok(ctx), ad = diff_ok(ctx);
        if (rec.wi.z <= 0.0f) return 0.0f;
        if (meas == DISCRETE) {
            if (!as || !delta_aligned(reflect(rec.wi, v3(0.0f, 0.0f, 1.0f)), rec.wo)) return 0.0f;
            return ad ? lobe_prob(fresnel_ext(rec.wi.z, eta)) : 1.0f;
        }
        if (meas != SOLID || !ad || rec.wo.z <= 0.0f) return 0.0f;
        float ps = 0.0f;
        if (as) ps = lobe_prob(fresnel_ext(rec.wi.z, eta));
        return (rec.wo.z / kPi) * (1.0f - ps);
    }
};

// ---- rough plastic (rough_plastic_bsdf.h:45-185) ---------------------------
constexpr uint32_t kRPG = GLOSSY | REFLECTIVE | DIELECTRIC;
struct rough_plastic {
    const lrt_material& M;
    const float* ext;
    const float* in;
    float eta, inv_eta2, w_spec;
    ggx g;
    rough_plastic(const env& E, const lrt_material& m)
        : M(m), ext(E.table(m.table_ext)), in(E.table(m.table_int)), g(m.alpha) {
        eta = m.int_ior / m.ext_ior;
        inv_eta2 = 1.0f / (eta * eta);
        w_spec = luminance(m.specular_reflectance) / (luminance(m.diffuse) + luminance(m.specular_reflectance));
    }
    bool gl_ok(const bsdf_ctx& c) const { return (c.type_mask & kRPG) != 0 && (c.component == -1 || c.component == 0); }
    bool di_ok(const bsdf_ctx& c) const { return (c.type_mask & kPD) != 0 && (c.component == -1 || c.component == 1); }
    float glossy_prob(float ci) const {
        const float R = 1.0f - trans_eval(ext, ci);
        const float gl = R * w_spec;
        const float di = (1.0f - R) * (1.0f - w_spec);
        return gl / (gl + di);
    }
    v3 scaled(float fdr) const {
        const v3 kd = col(M.diffuse);
        if (M.nonlinear) return v3(kd.x / (1.0f - kd.x * fdr), kd.y / (1.0f - kd.y * fdr), kd.z / (1.0f - kd.z * fdr));
        return kd / (1.0f - fdr);
    }
    v3 eval(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        if (meas != SOLID) return v3(0.0f);
        const bool ag = gl_ok(ctx), ad = di_ok(ctx);
        if (rec.wi.z <= 0.0f || rec.wo.z <= 0.0f || (!ag && !ad)) return v3(0.0f);
        v3 r(0.0f);
        if (ag) {
            const v3 h = normalize(rec.wi + rec.wo);
            const float d = g.D(h);
            const float F = fresnel_ext(dot(rec.wi, h), eta);
            const float gg = g.G(rec.wi, rec.wo, h);
            r = r + col(M.specular_reflectance) * (F * d * gg / (4.0f * rec.wi.z));
        }
        if (ad) {
            const float t12 = trans_eval(ext, rec.wi.z);
            const float t21 = trans_eval(ext, rec.wo.z);
            const float fdr = 1.0f - trans_diffuse(in);
            r = r + scaled(fdr) * ((rec.wo.z / kPi) * t12 * t21 * inv_eta2);
        }
        return r;
    }
    float pdf(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        if (meas != SOLID) return 0.0f;
        const bool ag = gl_ok(ctx), ad = di_ok(ctx);
        if (rec.wi.z <= 0.0f || rec.wo.z <= 0.0f || (!ag && !ad)) return 0.0f;
        float pg = 1.0f, pd = 1.0f;
        if (ag && ad) { pg = glossy_prob(rec.wi.z); pd = 1.0f - pg; }
        float dens = 0.0f;
        if (ag) {
            const v3 h = normalize(rec.wi + rec.wo);
            const float j = 1.0f / (4.0f * dot(rec.wo, h));
            dens += g.pdf(rec.wi, h) * j * pg;
        }
        if (ad) dens += pd * (rec.wo.z / kPi);
        return dens;
    }
    v3 sample(const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) const {
        rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
        const bool ag = gl_ok(ctx), ad = di_ok(ctx);
        if ((!ag && !ad) || rec.wi.z <= 0.0f) return v3(0.0f);
        float lx = ux, ly = uy;
        bool glossy = ag;
        if (ag && ad) {
            const float pg = glossy_prob(rec.wi.z);
            if (ly < pg) ly /= pg;
            else { ly = (ly - pg) / (1.0f - pg); glossy = false; }
        }
        if (glossy) {
            const v3 m = g.sample(rec.wi, lx, ly);
            rec.wo = reflect(rec.wi, m);
            rec.sampled_type = kRPG; rec.sampled_component = 0;
            if (rec.wo.z <= 0.0f) return v3(0.0f);
        } else {
            float cp = 0.0f;
            rec.wo = cosine_hemisphere(lx, ly, cp);
            rec.sampled_type = kPD; rec.sampled_component = 1;
        }
        const float dens = pdf(ctx, rec, SOLID);
        if (dens == 0.0f) return v3(0.0f);
        rec.pdf = dens;
        return eval(ctx, rec, SOLID) / dens;
    }
};

// ---- smooth coating (coating_bsdf.h:45-177) --------------------------------
constexpr uint32_t kCS = SPECULAR | REFLECTIVE | DIELECTRIC;
struct coating {
    const env& E;
    const lrt_material& M;
    float eta, inv_eta, w_spec;
    int spec_comp;
    coating(const env& e, const lrt_material& m) : E(e), M(m) {
        eta = m.int_ior / m.ext_ior;
        inv_eta = m.ext_ior / m.int_ior;
        w_spec = coat_weight(m.sigma_a, m.thickness);
        spec_comp = count_of(e, m.nested);
    }
    bool spec_ok(const bsdf_ctx& c) const { return (c.type_mask & kCS) != 0 && (c.component == -1 || c.component == spec_comp); }
    bool nest_ok(const bsdf_ctx& c) const {
        return (c.type_mask & flags_of(E, M.nested)) != 0 && (c.component == -1 || c.component < spec_comp);
    }
    v3 sample(const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) const {
        rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
        const bool as = spec_ok(ctx), an = nest_ok(ctx);
        if (!as && !an) return v3(0.0f);
        const refr in = refract_in(rec.wi, eta, inv_eta);
        const float ps = coat_prob(in.R, w_spec);
        float lx = ux, ly = uy;
        bool spec = as;
        if (as && an) {
            if (lx < ps) lx /= ps;
            else { lx = (lx - ps) / (1.0f - ps); spec = false; }
        }
        if (spec) {
            rec.wo = v3(-rec.wi.x, -rec.wi.y, rec.wi.z);
            rec.sampled_type = kCS; rec.sampled_component = spec_comp;
            rec.pdf = an ? ps : 1.0f;
            return col(M.specular_reflectance) * (in.R / rec.pdf);
        }
        if (in.R == 1.0f) return v3(0.0f);
        const v3 wi_outer = rec.wi;
        rec.wi = in.d;
        v3 w = sample_(E, M.nested, ctx, rec, lx, ly);
        const v3 wo_inner = rec.wo;
        rec.wi = wi_outer;
        if (w.is_zero()) return v3(0.0f);
        const refr out = refract_out(wo_inner, eta, inv_eta);
        rec.wo = out.d;
        if (out.R == 1.0f) return v3(0.0f);
        w = w * absorption(M.sigma_a, M.thickness, in.d, wo_inner);
        w = w * ((1.0f - in.R) * (1.0f - out.R));
        if (as) {
            rec.pdf *= 1.0f - ps;
            w = w / (1.0f - ps);
        }
        if (measure_of(rec.sampled_type) == SOLID) rec.pdf *= solid_angle_factor(inv_eta, rec.wo, wo_inner);
        return w;
    }
    v3 eval(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        const bool as = spec_ok(ctx), an = nest_ok(ctx);
        if (meas == DISCRETE && as && delta_aligned(v3(-rec.wi.x, -rec.wi.y, rec.wi.z), rec.wo))
            return col(M.specular_reflectance) * fresnel_ext(std::abs(rec.wi.z), eta);
        if (!an) return v3(0.0f);
        const refr in = refract_in(rec.wi, eta, inv_eta);
        const refr out = refract_in(rec.wo, eta, inv_eta);
        if (in.R == 1.0f || out.R == 1.0f) return v3(0.0f);
        bsdf_rec inner = rec;
        inner.wi = in.d;
        inner.wo = out.d;
        v3 r = eval_(E, M.nested, ctx, inner, meas) * ((1.0f - in.R) * (1.0f - out.R));
        r = r * absorption(M.sigma_a, M.thickness, in.d, out.d);
        if (meas == SOLID) r = r * solid_angle_factor(inv_eta, rec.wo, out.d);
        return r;
    }
    float pdf(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        const bool as = spec_ok(ctx), an = nest_ok(ctx);
        const refr in = refract_in(rec.wi, eta, inv_eta);
        if (meas == DISCRETE && as && delta_aligned(v3(-rec.wi.x, -rec.wi.y, rec.wi.z), rec.wo))
            return an ? coat_prob(in.R, w_spec) : 1.0f;
        if (!an) return 0.0f;
        const refr out = refract_in(rec.wo, eta, inv_eta);
        if (in.R == 1.0f || out.R == 1.0f) return 0.0f;
        bsdf_rec inner = rec;
        inner.wi = in.d;
        inner.wo = out.d;
        float d = pdf_(E, M.nested, ctx, inner, meas);
        if (meas == SOLID) d *= solid_angle_factor(inv_eta, rec.wo, out.d);
        if (as) d *= 1.0f - coat_prob(in.R, w_spec);
        return d;
    }
};

// ---- rough coating (rough_coating_bsdf.h:51-217) ---------------------------
constexpr uint32_t kRCS = GLOSSY | REFLECTIVE | DIELECTRIC;
struct rough_coating {
    const env& E;
    const lrt_material& M;
    const float* tr;
    float eta, inv_eta, w_spec;
    int spec_comp;
    ggx g;
    rough_coating(const env& e, const lrt_material& m) : E(e), M(m), tr(e.table(m.table_ext)), g(m.alpha) {
        eta = m.int_ior / m.ext_ior;
        inv_eta = m.ext_ior / m.int_ior;
        w_spec = coat_weight(m.sigma_a, m.thickness);
        spec_comp = count_of(e, m.nested);
    }
    bool comp_ok(const bsdf_ctx& c) const { return (c.type_mask & kRCS) != 0 && (c.component == -1 || c.component == spec_comp); }
    bool spec_ok(const bsdf_ctx& c, int meas) const { return comp_ok(c) && meas == SOLID; }
    bool nest_ok(const bsdf_ctx& c) const {
        return (c.type_mask & flags_of(E, M.nested)) != 0 && (c.component == -1 || c.component < spec_comp);
    }
    float spec_prob(float ci) const { return coat_prob(1.0f - trans_eval(tr, std::abs(ci)), w_spec); }
    v3 nested_eval(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        bsdf_rec inner = rec;
        inner.wi = refract_to(true, rec.wi, eta, inv_eta);
        inner.wo = refract_to(true, rec.wo, eta, inv_eta);
        if (inner.wi.is_zero() || inner.wo.is_zero()) return v3(0.0f);
        v3 r = eval_(E, M.nested, ctx, inner, meas) * trans_eval(tr, std::abs(rec.wi.z)) * trans_eval(tr, std::abs(rec.wo.z));
        r = r * absorption(M.sigma_a, M.thickness, inner.wi, inner.wo);
        if (meas == SOLID) r = r * solid_angle_factor(inv_eta, rec.wo, inner.wo);
        return r;
    }
    float nested_pdf(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        bsdf_rec inner = rec;
        inner.wi = refract_to(true, rec.wi, eta, inv_eta);
        inner.wo = refract_to(true, rec.wo, eta, inv_eta);
        if (inner.wi.is_zero() || inner.wo.is_zero()) return 0.0f;
        float d = pdf_(E, M.nested, ctx, inner, meas);
        if (meas == SOLID) d *= solid_angle_factor(inv_eta, rec.wo, inner.wo);
        return d;
    }
    v3 eval(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        const bool an = nest_ok(ctx), as = spec_ok(ctx, meas);
        v3 r(0.0f);
        if (as && rec.wo.z * rec.wi.z > 0.0f) {
            v3 h = rec.wi + rec.wo;
            if (h.length_squared() > 1e-12f) {
                h = normalize(h) * signum(rec.wo.z);
                const float d = g.D(h);
                const float F = fresnel_ext(std::abs(dot(rec.wi, h)), eta);
                const float gg = g.G(rec.wi, rec.wo, h);
                r = r + col(M.specular_reflectance) * (F * d * gg / (4.0f * std::abs(rec.wi.z)));
            }
        }
        if (an) r = r + nested_eval(ctx, rec, meas);
        return r;
    }
    float pdf(const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) const {
        const bool an = nest_ok(ctx), as = spec_ok(ctx, meas);
        float ps = 1.0f, pn = 1.0f;
        if (as && an) { ps = spec_prob(rec.wi.z); pn = 1.0f - ps; }
        float dens = 0.0f;
        if (as && rec.wo.z * rec.wi.z > 0.0f) {
            v3 h = rec.wi + rec.wo;
            if (h.length_squared() > 1e-12f) {
                h = normalize(h) * signum(rec.wo.z);
                const float j = 1.0f / (4.0f * std::abs(dot(rec.wo, h)));
                dens += g.pdf(rec.wi, h) * j * ps;
            }
        }
        if (an) dens += nested_pdf(ctx, rec, meas) * pn;
        return dens;
    }
    v3 sample(const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) const {
        rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
        const bool as = comp_ok(ctx), an = nest_ok(ctx);
        if (!as && !an) return v3(0.0f);
        float lx = ux, ly = uy;
        bool spec = as;
        if (as && an) {
            const float ps = spec_prob(rec.wi.z);
            if (ly < ps) ly /= ps;
            else { ly = (ly - ps) / (1.0f - ps); spec = false; }
        }
        if (spec) {
            const v3 m = g.sample(rec.wi, lx, ly);
            rec.wo = reflect(rec.wi, m);
            rec.sampled_type = kRCS; rec.sampled_component = spec_comp; rec.eta = 1.0f;
            if (rec.wo.z * rec.wi.z <= 0.0f) return v3(0.0f);
        } else {
            const v3 wi_outer = rec.wi;
            rec.wi = refract_to(true, wi_outer, eta, inv_eta);
            const v3 w = sample_(E, M.nested, ctx, rec, lx, ly);
            rec.wi = wi_outer;
            if (w.is_zero()) return v3(0.0f);
            rec.wo = refract_to(false, rec.wo, eta, inv_eta);
            if (rec.wo.is_zero()) return v3(0.0f);
        }
        const int meas = measure_of(rec.sampled_type);
        const float dens = pdf(ctx, rec, meas);
        if (dens == 0.0f) return v3(0.0f);
        rec.pdf = dens;
        return eval(ctx, rec, meas) / dens;
    }
};

v3 eval_(const env& E, int m, const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) {
    const lrt_material& M = E.mats[(size_t)m];
    switch (M.type) {
        case LRT_BSDF_DIFFUSE:
            return diffuse_ok(ctx, rec, meas) ? col(M.diffuse) * (rec.wo.z / kPi) : v3(0.0f);
        case LRT_BSDF_CONDUCTOR: {
            const uint32_t lobe = conductor_lobe(M);
            if ((ctx.type_mask & lobe) == 0) return v3(0.0f);
            if (ctx.component != -1 && ctx.component != 0) return v3(0.0f);
            if (M.alpha <= 0.0f) {
                if (meas != DISCRETE || rec.wo.z <= 0.0f || !delta_aligned(v3(-rec.wi.x, -rec.wi.y, rec.wi.z), rec.wo))
                    return v3(0.0f);
                return fresnel_conductor(rec.wi.z, M.cond_eta, M.cond_k) * col(M.specular_reflectance);
            }
            if (meas != SOLID) return v3(0.0f);
            return conductor_eval_pdf(M, rec.wi, rec.wo).f;
        }
        case LRT_BSDF_DIELECTRIC: { v3 f; float p; dielectric_eval_pdf(M, ctx, rec, meas, f, p); return f; }
        case LRT_BSDF_ROUGH_DIELECTRIC: return meas != SOLID ? v3(0.0f) : rd_eval_pdf(M, ctx, rec).f;
        case LRT_BSDF_PLASTIC: return plastic(M).eval(ctx, rec, meas);
        case LRT_BSDF_ROUGH_PLASTIC: return rough_plastic(E, M).eval(ctx, rec, meas);
        case LRT_BSDF_COATING: return coating(E, M).eval(ctx, rec, meas);
        case LRT_BSDF_ROUGH_COATING: return rough_coating(E, M).eval(ctx, rec, meas);
        default: return v3(0.0f);
    }
}
float pdf_(const env& E, int m, const bsdf_ctx& ctx, const bsdf_rec& rec, int meas) {
    const lrt_material& M = E.mats[(size_t)m];
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return diffuse_ok(ctx, rec, meas) ? rec.wo.z / kPi : 0.0f;
        case LRT_BSDF_CONDUCTOR: {
            const uint32_t lobe = conductor_lobe(M);
            if ((ctx.type_mask & lobe) == 0) return 0.0f;
            if (ctx.component != -1 && ctx.component != 0) return 0.0f;
            if (M.alpha <= 0.0f) {
                if (meas != DISCRETE || rec.wo.z <= 0.0f || !delta_aligned(v3(-rec.wi.x, -rec.wi.y, rec.wi.z), rec.wo))
                    return 0.0f;
                return 1.0f;
            }
            if (meas != SOLID) return 0.0f;
            return conductor_eval_pdf(M, rec.wi, rec.wo).pdf;
        }
        case LRT_BSDF_DIELECTRIC: { v3 f; float p; dielectric_eval_pdf(M, ctx, rec, meas, f, p); return p; }
        case LRT_BSDF_ROUGH_DIELECTRIC: return meas != SOLID ? 0.0f : rd_eval_pdf(M, ctx, rec).pdf;
        case LRT_BSDF_PLASTIC: return plastic(M).pdf(ctx, rec, meas);
        case LRT_BSDF_ROUGH_PLASTIC: return rough_plastic(E, M).pdf(ctx, rec, meas);
        case LRT_BSDF_COATING: return coating(E, M).pdf(ctx, rec, meas);
        case LRT_BSDF_ROUGH_COATING: return rough_coating(E, M).pdf(ctx, rec, meas);
        default: return 0.0f;
    }
}
v3 sample_(const env& E, int m, const bsdf_ctx& ctx, bsdf_rec& rec, float ux, float uy) {
    const lrt_material& M = E.mats[(size_t)m];
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return diffuse_sample(M, ctx, rec, ux, uy);
        case LRT_BSDF_CONDUCTOR: return conductor_sample(M, ctx, rec, ux, uy);
        case LRT_BSDF_DIELECTRIC: return dielectric_sample(M, ctx, rec, ux, uy);
        case LRT_BSDF_ROUGH_DIELECTRIC: return rd_sample(M, ctx, rec, ux, uy);
        case LRT_BSDF_PLASTIC: return plastic(M).sample(ctx, rec, ux, uy);
        case LRT_BSDF_ROUGH_PLASTIC: return rough_plastic(E, M).sample(ctx, rec, ux, uy);
        case LRT_BSDF_COATING: return coating(E, M).sample(ctx, rec, ux, uy);
        case LRT_BSDF_ROUGH_COATING: return rough_coating(E, M).sample(ctx, rec, ux, uy);
        default:
            rec.pdf = 0.0f; rec.eta = 1.0f; rec.sampled_type = 0; rec.sampled_component = -1;
            return v3(0.0f);
    }
}

}  // namespace

uint32_t bsdf_flags(const std::vector<lrt_material>& mats, int m) { return flags_of(env{mats, nullptr}, m); }
int bsdf_component_count(const std::vector<lrt_material>& mats, int m) { return count_of(env{mats, nullptr}, m); }
v3 bsdf_eval(const std::vector<lrt_material>& mats, const float* tables, int m, const bsdf_ctx& ctx,
             const bsdf_rec& rec, int measure) {
    return eval_(env{mats, tables}, m, ctx, rec, measure);
}
float bsdf_pdf(const std::vector<lrt_material>& mats, const float* tables, int m, const bsdf_ctx& ctx,
               const bsdf_rec& rec, int measure) {
    return pdf_(env{mats, tables}, m, ctx, rec, measure);
}
v3 bsdf_sample(const std::vector<lrt_material>& mats, const float* tables, int m, const bsdf_ctx& ctx,
               bsdf_rec& rec, float ux, float uy) {
    return sample_(env{mats, tables}, m, ctx, rec, ux, uy);
}

}  // namespace orc
