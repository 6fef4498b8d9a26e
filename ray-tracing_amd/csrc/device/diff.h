// Differentiable rendering: the path tracer of k_shade as one function per
// sample (same Philox dimensions, NEE with MIS, Russian roulette, so the same
// sample gives the same radiance), with an albedo texture on one material
// and the GGX alpha / interior IOR of given materials as parameters, and its
// path-replay adjoint (Vicini et al. 2021): the sample is traced once for its
// radiance L, then replayed; at each vertex the radiance still to come,
// L - (radiance gathered so far), times d log f / d theta of the BSDF value
// there gives that vertex's gradient, and every next-event term adds its own.
// Sampling densities, MIS weights and roulette probabilities are detached
// (held constant), which keeps the gradient estimate unbiased.  No tape is
// stored: a replay recomputes the path from its Philox seed.
//
// d f / d alpha and d f / d eta are exact (forward-mode dual numbers through
// the GGX lobes, diff_df), and so is the texture's derivative (a diffuse value
// is linear in its albedo).  A sample whose radiance is
// not finite (a few per 10^9) counts as black and adds no gradient, so one
// cannot poison a loss.
#pragma once

namespace lrt {

struct DiffParams {
    const float* tex;     // albedo texture, tex_h rows of tex_w RGB texels
    const lrt_material* eval_mats;  // non-null: BSDF values from these materials, sampling from M
    int32_t tex_w, tex_h, tex_mat;
    int32_t alpha_mat[2], ior_mat;
    uint32_t max_depth, rr_depth, W, H, spp, spp0;
    uint2 key;
};

constexpr float kDiffStep = 1e-3f;

// Texel of a point on the surface about the origin: its direction projected
// along z onto the unit disk (the hemisphere facing +z; a point behind takes
// the texel of its mirror image in z = 0), the disk mapped onto the square
// [-1, 1]^2 by the inverse of Shirley and Chiu's concentric map.  The
// concentric map preserves area and the projection along z is what a camera
// far out on +z sees, so from such a camera every texel covers about the same
// share of the surface's image and every texel is visible.  Row 0 is the top
// (+y).
LRT_D int diff_texel(const DiffParams& D, f3 p) {
    const f3 d = p * (1.0f / fmaxf(len(p), 1e-20f));
    const float x = d.x, y = d.y, r = sqrtf(fminf(x * x + y * y, 1.0f));
    float u, v;
    if (fabsf(x) >= fabsf(y)) {
        u = copysignf(r, x);
        v = x != 0.0f ? u * (4.0f / kPiF) * atanf(y / x) : 0.0f;
    } else {
        v = copysignf(r, y);
        u = v * (4.0f / kPiF) * atanf(x / y);
    }
    const int tx = min(max((int)((u + 1.0f) * 0.5f * (float)D.tex_w), 0), D.tex_w - 1);
    const int ty = min(max((int)((1.0f - v) * 0.5f * (float)D.tex_h), 0), D.tex_h - 1);
    return ty * D.tex_w + tx;
}

// the material seen at a hit: the textured one gets its texel as albedo
struct DiffMat {
    lrt_material m;
    Mats view;
    int index;
    int texel;  // -1 unless textured
};
LRT_D void diff_mat(const Mats& M, const DiffParams& D, int material, f3 p, DiffMat& out) {
    out.texel = -1;
    if (material == D.tex_mat && D.tex) {
        out.m = M.m[material];
        out.texel = diff_texel(D, p);
        const float* t = D.tex + 3 * (size_t)out.texel;
        out.m.diffuse[0] = t[0];
        out.m.diffuse[1] = t[1];
        out.m.diffuse[2] = t[2];
        out.view = Mats{&out.m, M.tables};
        out.index = 0;
    } else {
        out.view = M;
        out.index = material;
    }
}

// ---- d f / d alpha and d f / d eta of the GGX lobes, exact ----------------------
// Forward-mode dual numbers (value, derivative) through the rough conductor's
// and the rough dielectric's BSDF value (rough_conductor_bsdf.h:28-80,
// rough_dielectric_bsdf.h:28-140): the GGX distribution D and Smith's G1
// (microfacet.h:10-60), the dielectric Fresnel term (fresnel.h:72-95) and the
// transmission half vector, which moves with the interior IOR.  Every branch is
// decided on the value, as the float evaluation decides it; the conductor's
// Fresnel term does not depend on alpha.
struct Dual {
    float v, d;
};
LRT_D Dual K(float c) { return Dual{c, 0.0f}; }
LRT_D Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
LRT_D Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
LRT_D Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
LRT_D Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
LRT_D Dual operator/(Dual a, Dual b) { return Dual{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
LRT_D Dual dsqrt(Dual a) {
    const float s = sqrtf(a.v);
    return Dual{s, s > 0.0f ? a.d / (2.0f * s) : 0.0f};
}
LRT_D Dual dabs(Dual a) { return a.v < 0.0f ? -a : a; }
struct D3 {
    Dual x, y, z;
};
LRT_D D3 K3(f3 v) { return D3{K(v.x), K(v.y), K(v.z)}; }
LRT_D Dual ddot(D3 a, f3 b) { return a.x * K(b.x) + a.y * K(b.y) + a.z * K(b.z); }
LRT_D Dual ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

LRT_D Dual dggx_alpha(Dual a) { return a.v < 0.0001f ? K(0.0001f) : a; }  // microfacet.h:12-13
LRT_D Dual dggx_D(Dual a, D3 m) {
    if (m.z.v <= 0.0f) return K(0.0f);
    const Dual a2 = a * a;
    const Dual t = (m.x * m.x + m.y * m.y) / a2 + m.z * m.z;
    return K(1.0f) / (K(kPiF) * a2 * t * t);
}
LRT_D Dual dggx_G1(Dual a, f3 v, D3 m) {
    const float cv = fabsf(v.z);
    if (cv < 1e-6f || ddot(m, v).v * v.z <= 0.0f) return K(0.0f);
    const Dual t = (a * a) * K(v.x * v.x + v.y * v.y) / K(v.z * v.z);
    const Dual lambda = K(0.5f) * (K(-1.0f) + dsqrt(K(1.0f) + t));
    return K(1.0f) / (K(1.0f) + lambda);
}
LRT_D Dual dfresnel_dielectric(Dual ci, Dual ei, Dual et) {
    if (ei.v <= 0.0f || et.v <= 0.0f) return K(1.0f);
    if (ci.v < 0.0f) ci = K(0.0f);
    if (ci.v > 1.0f) ci = K(1.0f);
    const Dual s2i = K(1.0f) - ci * ci;
    const Dual s = ei / et;
    const Dual s2t = s * s * s2i;
    if (s2t.v > 1.0f) return K(1.0f);
    const Dual ct = dsqrt(K(1.0f) - s2t);
    const Dual rs = (ei * ci - et * ct) / (ei * ci + et * ct);
    const Dual rp = (et * ci - ei * ct) / (et * ci + ei * ct);
    return K(0.5f) * (rs * rs + rp * rp);
}

// d f / d alpha of the rough conductor (conductor_ep)
LRT_D f3 conductor_dalpha(const lrt_material& M, const BsdfCtx& c, const SRec& r) {
    if ((c.mask & conductor_lobe(M)) == 0 || (c.component != -1 && c.component != 0) || M.alpha <= 0.0f) return splat(0.0f);
    const f3 wi = r.wi, wo = r.wo;
    if (wi.z <= 1e-6f || wo.z <= 1e-6f) return splat(0.0f);
    const f3 h = wi + wo;
    if (len2(h) <= 1e-12f) return splat(0.0f);
    const f3 m = normalize(h);
    const float wim = dot(wi, m);
    if (m.z <= 0.0f || wim <= 0.0f) return splat(0.0f);
    const Dual a = dggx_alpha(Dual{M.alpha, 1.0f});
    const D3 dm = K3(m);
    const Dual dg = dggx_D(a, dm) * (dggx_G1(a, wi, dm) * dggx_G1(a, wo, dm));
    const f3 F = fresnel_conductor(wim, M.cond_eta, M.cond_k) * col(M.specular_reflectance);
    return F * (dg.d / (4.0f * wi.z));
}

// d f / d alpha (which 0) or d f / d interior IOR (which 1) of the rough
// dielectric (rd_ep: reflection or transmission by the sides of wi and wo)
LRT_D f3 rd_dparam(const lrt_material& M, const BsdfCtx& c, const SRec& r, int which) {
    const f3 wi = r.wi, wo = r.wo;
    if (wi.z <= 1e-6f || fabsf(wo.z) <= 1e-6f) return splat(0.0f);
    const Dual a = dggx_alpha(Dual{M.alpha, which == 0 ? 1.0f : 0.0f});
    const Dual eta_int{M.int_ior, which == 1 ? 1.0f : 0.0f};
    const Dual ei = r.front ? K(M.ext_ior) : eta_int, et = r.front ? eta_int : K(M.ext_ior);
    if (wi.z * wo.z > 0.0f) {  // rd_reflection
        if (!rd_r_ok(c) || wo.z <= 1e-6f) return splat(0.0f);
        f3 m = wi + wo;
        if (len2(m) <= 1e-12f) return splat(0.0f);
        m = normalize(m);
        if (m.z < 0.0f) m = -m;
        const float wim = dot(wi, m);
        if (wim <= 0.0f) return splat(0.0f);
        const D3 dm = K3(m);
        const Dual F = dfresnel_dielectric(K(wim), ei, et);
        const Dual v = F * dggx_D(a, dm) * (dggx_G1(a, wi, dm) * dggx_G1(a, wo, dm)) / K(4.0f * wi.z);
        return splat(v.d);
    }
    // rd_transmission
    if (!rd_t_ok(c) || wo.z >= -1e-6f || et.v <= 0.0f) return splat(0.0f);
    D3 m{-(ei * K(wi.x) + et * K(wo.x)), -(ei * K(wi.y) + et * K(wo.y)), -(ei * K(wi.z) + et * K(wo.z))};
    const Dual l2 = ddot(m, m);
    if (l2.v <= 1e-12f) return splat(0.0f);
    const Dual l = dsqrt(l2);
    m = D3{m.x / l, m.y / l, m.z / l};
    if (m.z.v < 0.0f) m = D3{-m.x, -m.y, -m.z};
    const Dual wim = ddot(m, wi), wom = ddot(m, wo);
    if (wim.v <= 0.0f) return splat(0.0f);
    const Dual F = dfresnel_dielectric(wim, ei, et);
    const Dual denom = ei * wim + et * wom;
    if (F.v >= 1.0f || denom.v * denom.v <= 1e-12f) return splat(0.0f);
    const Dual g = dggx_D(a, m) * (dggx_G1(a, wi, m) * dggx_G1(a, wo, m));
    const Dual es = c.mode == LRT_MODE_RADIANCE ? ei : et;
    const Dual v = (K(1.0f) - F) * g * wim * dabs(wom) * es * es / (K(wi.z) * denom * denom);
    return splat(v.d);
}

// d f / d theta of a base BSDF at the sampled directions: which 0 alpha, 1 the
// interior IOR.  Exact for the rough conductor and the rough dielectric (the
// parameters of the inverse-rendering workload); other BSDF types by central
// differences (step 1e-3).
LRT_D f3 diff_df(const Mats& M, int material, int which, const BsdfCtx& ctx, const SRec& sr) {
    const lrt_material& m0 = M.m[material];
    if (m0.type == LRT_BSDF_CONDUCTOR) return which == 0 ? conductor_dalpha(m0, ctx, sr) : splat(0.0f);
    if (m0.type == LRT_BSDF_ROUGH_DIELECTRIC) return rd_dparam(m0, ctx, sr, which);
    lrt_material a = m0, b = m0;
    if (which == 0) { a.alpha += kDiffStep; b.alpha -= kDiffStep; }
    else { a.int_ior += kDiffStep; b.int_ior -= kDiffStep; }
    const Mats va{&a, M.tables}, vb{&b, M.tables};
    const EP ea = base_ep(va, 0, ctx, sr, M_SOLID), eb = base_ep(vb, 0, ctx, sr, M_SOLID);
    return (ea.f - eb.f) * (0.5f / kDiffStep);
}

// Gradient sums: every contribution is rounded to a 64-bit fixed-point integer
// (value x scale, the scale a power of two picked per adjoint call from the
// largest |d_image| by k_grad_scale) and added with integer atomics, so a sum
// does not depend on the order its atomics land in and the adjoint is
// bit-reproducible run to run.  A per-sample pixel weight of the largest size
// is 2^kGradBits quanta; sums hold up to 2^63 quanta.
constexpr int kGradBits = 20;

LRT_D long long grad_quant(float v, float scale) {
    if (!isfinite(v)) return 0;
    return (long long)rintf(fminf(fmaxf(v * scale, -4.0e18f), 4.0e18f));
}

LRT_D void grad_add(unsigned long long* a, float v, float scale) {
    const long long q = grad_quant(v, scale);
    if (q != 0) atomicAdd(a, (unsigned long long)q);
}

// the wave's sum of a per-lane value, one atomic per wave (all lanes active)
LRT_D void grad_wave_add(unsigned long long* a, float v, float scale) {
    long long q = grad_quant(v, scale);
    for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off);
    if ((threadIdx.x & 63) == 0 && q != 0) atomicAdd(a, (unsigned long long)q);
}

// the fixed-point scale: 2^(kGradBits - floor(log2 m)), m the largest finite
// |d_img| / spp (1 when there is none); one block of 1024 threads
__global__ __launch_bounds__(1024) void k_grad_scale(const float* __restrict__ d_img, uint32_t n, float inv_spp,
                                                     float* __restrict__ scale) {
    __shared__ float red[16];
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
        const float a = fabsf(d_img[i]);
        if (a <= 3.4e38f) m = fmaxf(m, a);  // skips inf and NaN
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; ++w) m = fmaxf(m, red[w]);
        const float ms = m * inv_spp;
        const int e = ms > 0.0f ? min(max(kGradBits - ilogbf(ms), -100), 100) : 0;
        *scale = ldexpf(1.0f, e);
    }
}

// fixed-point sums -> float gradients
__global__ void k_grad_out(const unsigned long long* __restrict__ acc, uint32_t n, const float* __restrict__ scale,
                           float* __restrict__ out) {
    const double inv = 1.0 / (double)*scale;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (float)((double)(long long)acc[i] * inv);
}

struct DiffGrad {
    unsigned long long* tex;  // texture gradient (fixed-point sums)
    float scale;              // fixed-point scale
    float s[3];               // alpha 0, alpha 1, eta (per lane, reduced by the caller)
};

// one sample of path_integrator::radiance (as k_shade computes it); with ADJ
// the replay that adds delta-weighted gradients given the sample's radiance
template <bool ADJ>
LRT_D f3 diff_sample(const DevScene& S, const DevLights& Ls, const Mats& M, const DiffParams& D, const CameraDev& cam,
                     uint32_t x, uint32_t y, uint32_t smp, uint2* lds, f3 Ltot, f3 delta, DiffGrad& G) {
    const uint32_t pix = y * D.W + x;
    const uint4 v0 = philox(make_uint4(pix, smp, 0u, 0u), D.key);
    const float u = ((float)x + u24(v0.x)) / (float)(D.W - 1);
    const float vv = ((float)y + u24(v0.y)) / (float)(D.H - 1);
    Ray r{ld3(cam.origin), normalize(ld3(cam.llc) + u * ld3(cam.horizontal) + vv * ld3(cam.vertical) - ld3(cam.origin))};
    f3 T = splat(1.0f), L = splat(0.0f);
    float eta = 1.0f, sa_pdf = -1.0f;
    for (uint32_t depth = 0; depth < D.max_depth; ++depth) {
        const bool have = sa_pdf >= 0.0f;
        HitInfo h;
        Counters cnt{0, 0, 0};
        if (!trace_closest<false>(S, r, 0.001f, kInfF, h, lds, cnt)) {
            const f3 Le = env_eval(Ls, r.d);
            L = L + (have ? T * power_heuristic(sa_pdf, pdf_emitter_direction(S, Ls, splat(0.0f), r.d)) * Le : T * Le);
            break;
        }
        const SurfaceHit sh = surface(S, r, h.obj, h.prim, h.t, h.u, h.v);
        if (sh.emitter >= 0 && sh.front) {
            const f3 out = -normalize(r.d);
            const f3 Le = dot(sh.n, out) <= 0.0f ? splat(0.0f) : ld3(Ls.l[sh.emitter].radiance);
            if (!is_zero(Le))
                L = L + (have ? T * power_heuristic(sa_pdf, pdf_emitter_direction(S, Ls, r.o, r.d)) * Le : T * Le);
        }
        if (sh.material < 0) break;
        DiffMat dm;
        diff_mat(M, D, sh.material, sh.p, dm);
        const int pk = sh.material == D.alpha_mat[0] ? 0 : (sh.material == D.alpha_mat[1] ? 1 : -1);
        const bool pe = sh.material == D.ior_mat;
        const Frame fr = frame_from(sh.n);
        SRec sr;
        sr.wi = to_local(fr, -normalize(r.d));
        sr.front = sh.front;
        BsdfCtx ctx;
        const uint32_t d0 = 2u + 6u * depth;
        const uint4 A = philox(make_uint4(pix, smp, d0 >> 2, 0u), D.key);
        const uint4 B = philox(make_uint4(pix, smp, (d0 >> 2) + 1u, 0u), D.key);
        float u6[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const uint32_t d = d0 + (uint32_t)k;
            u6[k] = u24((d >> 2) == (d0 >> 2) ? pick(A, d & 3u) : pick(B, d & 3u));
        }
        ctx.lobe = u6[2];
        const uint32_t fl = bsdf_flags(dm.view, dm.index);
        if (fl & (F_DIFFUSE | F_GLOSSY)) {  // NEE (path_integrator.h:86-128)
            DirectSample ds;
            if (sample_emitter_direction(S, Ls, sh.p, u6[0], u6[1], ds)) {
                const f3 so = sh.p + sh.n * 1e-4f;
                f3 sd = ds.dir;
                float tmax = kInfF;
                bool target = true;
                if (!isinf(ds.dist)) {
                    const f3 to = ds.point - so;
                    const float dist = len(to);
                    if (dist < 1e-6f) target = false;
                    else { sd = to / dist; tmax = dist * (1.0f - 1e-4f); }
                }
                if (target) {
                    sr.wo = to_local(fr, sd);
                    const EP e = bsdf_ep(dm.view, dm.index, ctx, sr, M_SOLID);
                    if (!is_zero(e.f)) {
                        const float lp = pdf_emitter_direction(S, Ls, sh.p, sd);
                        if (lp > 0.0f && !trace_occluded(S, Ray{so, sd}, 0.001f, tmax, ds.emitter, lds)) {
                            const float w = power_heuristic(lp, e.pdf);
                            f3 fv = e.f;
                            if (!ADJ && D.eval_mats && dm.texel < 0)
                                fv = bsdf_ep(Mats{D.eval_mats, M.tables}, sh.material, ctx, sr, M_SOLID).f;
                            const f3 term = T * fv * ds.radiance * (w / lp);
                            L = L + term;
                            if (ADJ) {
                                if (dm.texel >= 0) {
                                    const float* rho = dm.m.diffuse;
                                    const float tv[3] = {term.x, term.y, term.z}, dv[3] = {delta.x, delta.y, delta.z};
                                    for (int c = 0; c < 3; ++c)
                                        if (rho[c] > 1e-20f && tv[c] != 0.0f)
                                            grad_add(G.tex + 3 * (size_t)dm.texel + c, dv[c] * tv[c] / rho[c], G.scale);
                                }
                                const f3 base = T * ds.radiance * (w / lp);
                                if (pk >= 0) {
                                    const float gv = dot(delta, base * diff_df(M, sh.material, 0, ctx, sr));
                                    if (isfinite(gv)) G.s[pk] += gv;
                                }
                                if (pe) {
                                    const float gv = dot(delta, base * diff_df(M, sh.material, 1, ctx, sr));
                                    if (isfinite(gv)) G.s[2] += gv;
                                }
                            }
                        }
                    }
                }
            }
        }
        if (depth + 1 >= D.max_depth) break;  // last bounce does NEE only
        f3 bw = bsdf_sample(dm.view, dm.index, ctx, sr, u6[3], u6[4]);
        if (is_zero(bw)) break;
        if (!ADJ && D.eval_mats && dm.texel < 0) {  // held sampling: weight f(eval) / p(sampling)
            const f3 f0 = bsdf_ep(dm.view, dm.index, ctx, sr, M_SOLID).f;
            const f3 f1 = bsdf_ep(Mats{D.eval_mats, M.tables}, sh.material, ctx, sr, M_SOLID).f;
            bw = f3{f0.x > 0.0f ? bw.x * (f1.x / f0.x) : 0.0f, f0.y > 0.0f ? bw.y * (f1.y / f0.y) : 0.0f,
                    f0.z > 0.0f ? bw.z * (f1.z / f0.z) : 0.0f};
        }
        if (ADJ) {
            // radiance still to come from this vertex on
            const f3 R = Ltot - L;
            if (dm.texel >= 0) {
                const float* rho = dm.m.diffuse;
                const float rv[3] = {R.x, R.y, R.z}, dv[3] = {delta.x, delta.y, delta.z};
                for (int c = 0; c < 3; ++c)
                    if (rho[c] > 1e-20f && rv[c] != 0.0f)
                        grad_add(G.tex + 3 * (size_t)dm.texel + c, dv[c] * rv[c] / rho[c], G.scale);
            }
            if (pk >= 0 || pe) {
                // d log f: channels whose value is too small to divide by carry no radiance
                const EP e = bsdf_ep(dm.view, dm.index, ctx, sr, M_SOLID);
                const f3 inv = f3{e.f.x > 1e-20f ? 1.0f / e.f.x : 0.0f, e.f.y > 1e-20f ? 1.0f / e.f.y : 0.0f,
                                  e.f.z > 1e-20f ? 1.0f / e.f.z : 0.0f};
                if (pk >= 0) {
                    const float gv = dot(delta, R * diff_df(M, sh.material, 0, ctx, sr) * inv);
                    if (isfinite(gv)) G.s[pk] += gv;
                }
                if (pe) {
                    const float gv = dot(delta, R * diff_df(M, sh.material, 1, ctx, sr) * inv);
                    if (isfinite(gv)) G.s[2] += gv;
                }
            }
        }
        T = T * bw;
        eta *= sr.eta;
        sa_pdf = ((sr.type & F_SPECULAR) == 0) && sr.pdf > 0.0f ? sr.pdf : -1.0f;
        r = Ray{sh.p, to_world(fr, sr.wo)};
        if (depth >= D.rr_depth) {  // Russian roulette (path_integrator.h:155-166), probability detached
            const float q = fminf(fmaxf(fmaxf(T.x, T.y), T.z) * eta * eta, 0.95f);
            if (u6[5] >= q) break;
            T = f3{T.x / q, T.y / q, T.z / q};
        }
    }
    return L;
}

// partial sums of the samples [j c, min(spp, (j + 1) c)) of every pixel:
// thread t takes pixel t % npix and chunk j = t / npix
__global__ __launch_bounds__(kBlock) LRT_SHADE_ATTR void k_diff_render(const DevScene S, const DevLights Ls, const Mats M,
                                                        const DiffParams D, const CameraDev cam, uint32_t chunk,
                                                        uint32_t nchunks, float* __restrict__ part) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    const uint32_t npix = D.W * D.H;
    const uint32_t n = npix * nchunks;
    DiffGrad G{nullptr, 0.0f, {0.0f, 0.0f, 0.0f}};
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t i = t % npix, j = t / npix;
        const uint32_t x = i % D.W, y = i / D.W;
        const uint32_t s0 = j * chunk, s1 = min(D.spp, s0 + chunk);
        f3 acc = splat(0.0f);
        for (uint32_t s = s0; s < s1; ++s) {
            const f3 L = diff_sample<false>(S, Ls, M, D, cam, x, y, D.spp0 + s, lds, splat(0.0f), splat(0.0f), G);
            if (isfinite(L.x + L.y + L.z)) acc = acc + L;  // a non-finite sample counts as black
        }
        float* o = part + 3 * (size_t)t;
        o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
    }
}

// image = (sum of the chunks in chunk order) / spp
__global__ __launch_bounds__(kBlock) void k_diff_resolve(const float* __restrict__ part, uint32_t npix, uint32_t nchunks,
                                                         uint32_t spp, float* __restrict__ img) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        f3 acc = splat(0.0f);
        for (uint32_t j = 0; j < nchunks; ++j) {
            const float* p = part + 3 * ((size_t)j * npix + i);
            acc = acc + f3{p[0], p[1], p[2]};
        }
        const float inv = 1.0f / (float)spp;
        img[3 * (size_t)i] = acc.x * inv;
        img[3 * (size_t)i + 1] = acc.y * inv;
        img[3 * (size_t)i + 2] = acc.z * inv;
    }
}

// gradient of sum(d_img * image) with respect to the parameters; threads over
// (pixel, sample chunk) like k_diff_render
__global__ __launch_bounds__(kBlock) LRT_SHADE_ATTR void k_diff_adjoint(const DevScene S, const DevLights Ls, const Mats M,
                                                         const DiffParams D, const CameraDev cam, uint32_t chunk,
                                                         uint32_t nchunks, const float* __restrict__ d_img,
                                                         unsigned long long* __restrict__ grad_tex,
                                                         unsigned long long* __restrict__ grad_s,
                                                         const float* __restrict__ gscale) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    const uint32_t npix = D.W * D.H;
    const uint32_t n = npix * nchunks;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n_round = (n + stride - 1) / stride * stride;
    DiffGrad G{grad_tex, *gscale, {0.0f, 0.0f, 0.0f}};
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_round; t += stride) {
        if (t < n) {
            const uint32_t i = t % npix, j = t / npix;
            const uint32_t x = i % D.W, y = i / D.W;
            const uint32_t s0 = j * chunk, s1 = min(D.spp, s0 + chunk);
            const float inv = 1.0f / (float)D.spp;
            const f3 delta = f3{d_img[3 * (size_t)i], d_img[3 * (size_t)i + 1], d_img[3 * (size_t)i + 2]} * inv;
            if (!is_zero(delta) && isfinite(delta.x + delta.y + delta.z))
                for (uint32_t s = s0; s < s1; ++s) {
                    const f3 L = diff_sample<false>(S, Ls, M, D, cam, x, y, D.spp0 + s, lds, splat(0.0f), splat(0.0f), G);
                    if (isfinite(L.x + L.y + L.z)) diff_sample<true>(S, Ls, M, D, cam, x, y, D.spp0 + s, lds, L, delta, G);
                }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) grad_wave_add(grad_s + k, G.s[k], G.scale);
}

// d f*cos / d theta of BSDF tuples (lrt_bsdf_param_grad): the derivative the
// adjoint uses at a vertex, in the solid-angle measure
__global__ void k_bsdf_param_grad(const Mats M, const int32_t* mat, const float* wi, const float* wo,
                                  const uint8_t* front, uint32_t n, int mode, int which, float* df) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        BsdfCtx c;
        c.mode = mode;
        SRec r;
        r.wi = ld3(wi + 3 * (size_t)i);
        r.wo = ld3(wo + 3 * (size_t)i);
        r.front = front[i] != 0;
        const f3 g = diff_df(M, mat[i], which, c, r);
        df[3 * (size_t)i] = g.x;
        df[3 * (size_t)i + 1] = g.y;
        df[3 * (size_t)i + 2] = g.z;
    }
}

}  // namespace lrt
