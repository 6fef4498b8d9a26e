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
// d f / d alpha and d f / d eta come from central differences of the BSDF
// value at the sampled directions (step 1e-3); the texture's derivative is
// exact (a diffuse value is linear in its albedo).  A sample whose radiance is
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

// spherical mapping about the origin: u from the azimuth, v from the polar angle
LRT_D int diff_texel(const DiffParams& D, f3 p) {
    const float r = len(p);
    const float u = 0.5f + atan2f(p.z, p.x) * (0.5f / kPiF);
    const float v = acosf(fminf(fmaxf(p.y / fmaxf(r, 1e-20f), -1.0f), 1.0f)) * (1.0f / kPiF);
    const int tx = min(max((int)(u * (float)D.tex_w), 0), D.tex_w - 1);
    const int ty = min(max((int)(v * (float)D.tex_h), 0), D.tex_h - 1);
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

// d f / d theta of a base BSDF by central differences: which 0 alpha, 1 eta
LRT_D f3 diff_df(const Mats& M, int material, int which, const BsdfCtx& ctx, const SRec& sr) {
    lrt_material a = M.m[material], b = M.m[material];
    if (which == 0) { a.alpha += kDiffStep; b.alpha -= kDiffStep; }
    else { a.int_ior += kDiffStep; b.int_ior -= kDiffStep; }
    const Mats va{&a, M.tables}, vb{&b, M.tables};
    const EP ea = base_ep(va, 0, ctx, sr, M_SOLID), eb = base_ep(vb, 0, ctx, sr, M_SOLID);
    return (ea.f - eb.f) * (0.5f / kDiffStep);
}

struct DiffGrad {
    float* tex;     // texture gradient (atomics)
    float s[3];     // alpha 0, alpha 1, eta (per lane, reduced by the caller)
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
                                        if (rho[c] > 1e-20f && tv[c] != 0.0f && isfinite(dv[c] * tv[c] / rho[c]))
                                            atomicAdd(G.tex + 3 * (size_t)dm.texel + c, dv[c] * tv[c] / rho[c]);
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
                    if (rho[c] > 1e-20f && rv[c] != 0.0f && isfinite(dv[c] * rv[c] / rho[c]))
                        atomicAdd(G.tex + 3 * (size_t)dm.texel + c, dv[c] * rv[c] / rho[c]);
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
    DiffGrad G{nullptr, {0.0f, 0.0f, 0.0f}};
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
                                                         float* __restrict__ grad_tex, float* __restrict__ grad_s) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    const uint32_t npix = D.W * D.H;
    const uint32_t n = npix * nchunks;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n_round = (n + stride - 1) / stride * stride;
    DiffGrad G{grad_tex, {0.0f, 0.0f, 0.0f}};
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
    for (int k = 0; k < 3; ++k) {
        float v = G.s[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0 && v != 0.0f) atomicAdd(grad_s + k, v);
    }
}

}  // namespace lrt
