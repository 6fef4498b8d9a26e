// Differentiable rendering as a wavefront: the samples of diff_sample
// (device/diff.h) traced by the path tracer's stages, so traversal runs in
// the persistent kernels (k_extend, k_shadow) at their occupancy and only
// shading carries the differentiable material logic.
//
//   forward:  k_raygen -> [k_extend -> k_shade_bins/order -> k_diff_shade<false> -> k_shadow]* -> k_diff_accum
//   adjoint:  the forward stages for the adjoint's samples, k_diff_keep (each path's total radiance), then the
//             replay: k_raygen -> [k_extend -> bins/order -> k_diff_shade<true> -> k_diff_shadow]*
//
// The replay recomputes every path from its Philox streams (no tape).  At a
// vertex the gradient of the BSDF-sampled continuation needs the radiance
// still to come, L_total - L, with L including this vertex's next-event term
// when that shadow ray is unoccluded; the shadow test runs after shading, so
// the vertex's factors (1 / rho for the texel, d f / d alpha / f and
// d f / d eta / f) are kept with the path and applied when the path is shaded
// next, where L holds exactly that sum.  A path that ends instead has no
// radiance still to come.  Next-event gradient terms travel with the shadow
// record and count only for unoccluded rays.  Sampling, MIS weights and
// roulette stay detached, as in diff_sample.
#pragma once

namespace lrt {

// per path, beside PathState
struct DiffPaths {
    float *ltr, *ltg, *ltb;        // total radiance of the path (replay)
    int32_t *ptex;                 // pending texel of the last BSDF sample, -1: none
    int32_t *pflag;                // pending factors: bit 0 alpha (slot in bit 1), bit 2 eta
    float *rr, *rg, *rb;           // rho of the pending texel
    float *far, *fag, *fab;        // d f / d alpha / f
    float *fer, *feg, *feb;        // d f / d eta / f
};

// per shadow record, beside ShadowBuf (replay): gradient terms of the
// next-event sample, counted when the ray is unoccluded
struct DiffShadow {
    int32_t* tex;                  // texel or -1
    float *gr, *gg, *gb;           // texture terms
    float *ga, *ge;                // alpha term (slot in `slot`), eta term
    int32_t* slot;                 // alpha slot 0 / 1, or -1
};

struct DiffWave {
    uint32_t depth, max_depth, rr_depth, sample;
    uint2 key;
    const float* d_img;            // replay: gradient image (W x H x 3); null: forward
    float inv_spp;
    const float* gscale;           // replay: fixed-point scale of the gradient sums
};

// the wave's sums of per-lane fixed-point values (lanes take paths in queue
// order, which varies run to run: integer sums keep the result independent of
// it), one atomic per wave and parameter; all lanes active
LRT_D void grad_reduce(const long long (&q)[3], unsigned long long* __restrict__ grad_s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        long long v = q[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0 && v != 0) atomicAdd(grad_s + k, (unsigned long long)v);
    }
}

// one bounce of diff_sample for every queued path (in material order)
template <bool ADJ>
__global__ __launch_bounds__(kBlock) LRT_SHADE_ATTR void k_diff_shade(const DevScene S, const DevLights Ls, const Mats M,
                                                                       const DiffParams D, PathState P,
                                                                       const uint32_t* __restrict__ queue,
                                                                       uint32_t* __restrict__ next_queue, uint32_t* counters,
                                                                       const HitBuf Hb, const uint32_t* __restrict__ order,
                                                                       ShadowBuf Sb, DiffShadow DS, DiffPaths DP, DiffWave prm,
                                                                       unsigned long long* __restrict__ grad_tex,
                                                                       unsigned long long* __restrict__ grad_s) {
    const uint32_t n = counters[0];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n_round = (n + stride - 1) / stride * stride;
    const float gsc = ADJ ? *prm.gscale : 0.0f;
    long long gs[3] = {0, 0, 0};
#if LRT_SHADE_BLOCK_APPEND
    __shared__ BlockAppend<kBlock> ba;
#endif
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_round; j += stride) {
        bool cont = false, shadow = false;
        uint32_t p = 0, octant = 0;
        f3 so, sd, contrib;
        float tmax = 0.0f;
        int ignore = -1;
        int stex = -1, sslot = -1;
        f3 sgt = splat(0.0f);
        float sga = 0.0f, sge = 0.0f;
        if (j < n) {
            const uint32_t i = order[j];
            p = queue[i];
            const uint32_t pix = P.pixel0 + p % P.npix, smp = prm.sample + p / P.npix;
            float4* pw = path_words(P, p);
            const float4 w0 = pw[0], w1 = pw[1], w2 = pw[2], w3 = pw[3];
            const Ray r{xyz(w0), xyz(w1)};
            f3 T = xyz(w2);
            f3 L = xyz(w3);
            float eta = w0.w;
            const float sa_pdf = w1.w;
            const bool have = sa_pdf >= 0.0f;
            f3 delta = splat(0.0f), Lt = splat(0.0f);
            bool live = false;
            if (ADJ) {
                delta = f3{prm.d_img[3 * (size_t)pix], prm.d_img[3 * (size_t)pix + 1], prm.d_img[3 * (size_t)pix + 2]} *
                        prm.inv_spp;
                Lt = f3{DP.ltr[p], DP.ltg[p], DP.ltb[p]};
                live = !is_zero(delta) && isfinite(delta.x + delta.y + delta.z) && isfinite(Lt.x + Lt.y + Lt.z);
                // the previous vertex's BSDF-sample gradient: radiance still to come from there
                const int pt = DP.ptex[p], pf = DP.pflag[p];
                if (live && (pt >= 0 || pf != 0)) {
                    const f3 R = Lt - L;
                    if (pt >= 0) {
                        const float rho[3] = {DP.rr[p], DP.rg[p], DP.rb[p]};
                        const float rv[3] = {R.x, R.y, R.z}, dv[3] = {delta.x, delta.y, delta.z};
                        for (int c = 0; c < 3; ++c)
                            if (rho[c] > 1e-20f && rv[c] != 0.0f)
                                grad_add(grad_tex + 3 * (size_t)pt + c, dv[c] * rv[c] / rho[c], gsc);
                    }
                    if (pf & 1) {
                        const float gv = dot(delta, R * f3{DP.far[p], DP.fag[p], DP.fab[p]});
                        gs[(pf >> 1) & 1] += grad_quant(gv, gsc);
                    }
                    if (pf & 4) {
                        const float gv = dot(delta, R * f3{DP.fer[p], DP.feg[p], DP.feb[p]});
                        gs[2] += grad_quant(gv, gsc);
                    }
                }
                if (pt >= 0 || pf != 0) {
                    DP.ptex[p] = -1;
                    DP.pflag[p] = 0;
                }
            }
            const HitRec h = hit_load(Hb, i);
            const int obj = h.obj;
            if (obj < 0) {  // escape: environment with MIS
                const f3 Le = env_eval(Ls, r.d);
                L = L + (have ? T * power_heuristic(sa_pdf, pdf_emitter_direction(S, Ls, splat(0.0f), r.d)) * Le : T * Le);
            } else {
                const SurfaceHit sh = surface(S, r, obj, h.prim, h.t, h.u, h.v);
                if (sh.emitter >= 0 && sh.front) {
                    const f3 out = -normalize(r.d);
                    const f3 Le = dot(sh.n, out) <= 0.0f ? splat(0.0f) : ld3(Ls.l[sh.emitter].radiance);
                    if (!is_zero(Le))
                        L = L + (have ? T * power_heuristic(sa_pdf, pdf_emitter_direction(S, Ls, r.o, r.d)) * Le : T * Le);
                }
                if (sh.material >= 0) {
                    DiffMat dm;
                    diff_mat(M, D, sh.material, sh.p, dm);
                    const int pk = sh.material == D.alpha_mat[0] ? 0 : (sh.material == D.alpha_mat[1] ? 1 : -1);
                    const bool pe = sh.material == D.ior_mat;
                    const Frame fr = frame_from(sh.n);
                    SRec sr;
                    sr.wi = to_local(fr, -normalize(r.d));
                    sr.front = sh.front;
                    BsdfCtx ctx;
                    const uint32_t d0 = 2u + 6u * prm.depth;
                    const uint4 A = philox(make_uint4(pix, smp, d0 >> 2, 0u), prm.key);
                    const uint4 B = philox(make_uint4(pix, smp, (d0 >> 2) + 1u, 0u), prm.key);
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
                            so = sh.p + sh.n * 1e-4f;
                            sd = ds.dir;
                            tmax = kInfF;
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
                                    if (lp > 0.0f) {
                                        const float w = power_heuristic(lp, e.pdf);
                                        contrib = T * e.f * ds.radiance * (w / lp);
                                        ignore = ds.emitter;
                                        shadow = true;
                                        if (ADJ && live) {
                                            if (dm.texel >= 0) {
                                                const float* rho = dm.m.diffuse;
                                                const float tv[3] = {contrib.x, contrib.y, contrib.z};
                                                const float dv[3] = {delta.x, delta.y, delta.z};
                                                float g3[3] = {0.0f, 0.0f, 0.0f};
                                                for (int c = 0; c < 3; ++c)
                                                    if (rho[c] > 1e-20f && tv[c] != 0.0f && isfinite(dv[c] * tv[c] / rho[c]))
                                                        g3[c] = dv[c] * tv[c] / rho[c];
                                                stex = dm.texel;
                                                sgt = f3{g3[0], g3[1], g3[2]};
                                            }
                                            const f3 base = T * ds.radiance * (w / lp);
                                            if (pk >= 0) {
                                                const float gv = dot(delta, base * diff_df(M, sh.material, 0, ctx, sr));
                                                if (isfinite(gv)) { sga = gv; sslot = pk; }
                                            }
                                            if (pe) {
                                                const float gv = dot(delta, base * diff_df(M, sh.material, 1, ctx, sr));
                                                if (isfinite(gv)) sge = gv;
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
                    if (prm.depth + 1 < prm.max_depth) {  // last bounce does NEE only
                        const f3 bw = bsdf_sample(dm.view, dm.index, ctx, sr, u6[3], u6[4]);
                        if (!is_zero(bw)) {
                            if (ADJ && live) {  // this vertex's factors, applied when the path is shaded next
                                if (dm.texel >= 0) {
                                    DP.ptex[p] = dm.texel;
                                    DP.rr[p] = dm.m.diffuse[0]; DP.rg[p] = dm.m.diffuse[1]; DP.rb[p] = dm.m.diffuse[2];
                                }
                                if (pk >= 0 || pe) {
                                    const EP e = bsdf_ep(dm.view, dm.index, ctx, sr, M_SOLID);
                                    const f3 inv = f3{e.f.x > 1e-20f ? 1.0f / e.f.x : 0.0f, e.f.y > 1e-20f ? 1.0f / e.f.y : 0.0f,
                                                      e.f.z > 1e-20f ? 1.0f / e.f.z : 0.0f};
                                    int flag = 0;
                                    if (pk >= 0) {
                                        const f3 fa = diff_df(M, sh.material, 0, ctx, sr) * inv;
                                        DP.far[p] = fa.x; DP.fag[p] = fa.y; DP.fab[p] = fa.z;
                                        flag |= 1 | (pk << 1);
                                    }
                                    if (pe) {
                                        const f3 fe = diff_df(M, sh.material, 1, ctx, sr) * inv;
                                        DP.fer[p] = fe.x; DP.feg[p] = fe.y; DP.feb[p] = fe.z;
                                        flag |= 4;
                                    }
                                    DP.pflag[p] = flag;
                                }
                            }
                            const f3 nd = to_world(fr, sr.wo);
                            T = T * bw;
                            eta *= sr.eta;
                            const bool hp = ((sr.type & F_SPECULAR) == 0) && sr.pdf > 0.0f;
                            cont = true;
                            if (prm.depth >= prm.rr_depth) {  // Russian roulette, probability detached
                                const float q = fminf(fmaxf(fmaxf(T.x, T.y), T.z) * eta * eta, 0.95f);
                                if (u6[5] >= q) cont = false;
                                else T = f3{T.x / q, T.y / q, T.z / q};
                            }
                            if (cont) {
                                octant = dir_octant(nd);
                                pw[0] = word(sh.p, eta);
                                pw[1] = word(nd, hp ? sr.pdf : -1.0f);
                                pw[2] = word(T, 0.0f);
                            }
                        }
                    }
                }
            }
            pw[3] = word(L, 0.0f);
        }
#if LRT_SHADE_BLOCK_APPEND
        // one append per block and queue, as k_shade
        const uint2 sl = block_append2<kBlock>(ba, cont, octant, shadow, 0u, &counters[1], &counters[2]);
        if (cont) next_queue[sl.x] = p;
        const uint32_t ss = sl.y;
#else
        const uint32_t slot = wave_append(&counters[1], cont);
        if (cont) next_queue[slot] = p;
        const uint32_t ss = wave_append(&counters[2], shadow);
#endif
        if (shadow) {
            Sb.ox[ss] = so.x; Sb.oy[ss] = so.y; Sb.oz[ss] = so.z;
            Sb.dx[ss] = sd.x; Sb.dy[ss] = sd.y; Sb.dz[ss] = sd.z;
            Sb.tmax[ss] = tmax;
            Sb.cr[ss] = contrib.x; Sb.cg[ss] = contrib.y; Sb.cb[ss] = contrib.z;
            Sb.ignore[ss] = ignore;
            Sb.path[ss] = p;
            if (ADJ) {
                DS.tex[ss] = stex;
                DS.gr[ss] = sgt.x; DS.gg[ss] = sgt.y; DS.gb[ss] = sgt.z;
                DS.ga[ss] = sga;
                DS.ge[ss] = sge;
                DS.slot[ss] = sslot;
            }
        }
    }
    if (ADJ) grad_reduce(gs, grad_s);
}

// replay shadow rays: unoccluded ones add their radiance and gradient terms
__global__ __launch_bounds__(kBlock) LRT_TRACE_ATTR void k_diff_shadow(const DevScene S, PathState P,
                                                                       uint32_t* __restrict__ counters, const ShadowBuf Sb,
                                                                       const DiffShadow DS,
                                                                       unsigned long long* __restrict__ grad_tex,
                                                                       unsigned long long* __restrict__ grad_s,
                                                                       const float* __restrict__ gscale) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    const uint32_t n = counters[2];
    const float gsc = *gscale;
    long long gs[3] = {0, 0, 0};
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 2;  // wave_refill: first fill pending
    for (;;) {
        const uint32_t got = wave_refill(!active, &counters[5], n, more);
        if (got != 0xffffffffu) {
            i = got;
            lane_shadow_start(S, T, Ray{f3{Sb.ox[i], Sb.oy[i], Sb.oz[i]}, f3{Sb.dx[i], Sb.dy[i], Sb.dz[i]}}, 0.001f,
                              Sb.tmax[i], Sb.ignore[i]);
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            if (!shadow_blocked(S, T, Sb.ignore[i])) {
                float4* w = path_words(P, Sb.path[i]);
                w[3] = word(xyz(w[3]) + f3{Sb.cr[i], Sb.cg[i], Sb.cb[i]}, 0.0f);
                const int t = DS.tex[i];
                if (t >= 0) {
                    const float g3[3] = {DS.gr[i], DS.gg[i], DS.gb[i]};
                    for (int c = 0; c < 3; ++c)
                        if (g3[c] != 0.0f) grad_add(grad_tex + 3 * (size_t)t + c, g3[c], gsc);
                }
                if (DS.slot[i] >= 0) gs[DS.slot[i]] += grad_quant(DS.ga[i], gsc);
                gs[2] += grad_quant(DS.ge[i], gsc);
            }
            active = false;
        }
    }
    grad_reduce(gs, grad_s);
}

// forward image: per pixel the pass's samples in slot order, a non-finite
// sample counting as black, added into acc (float, W x H x 3)
__global__ __launch_bounds__(kBlock) void k_diff_accum(const PathState P, uint32_t slots, float* __restrict__ acc) {
    const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= P.npix) return;
    const size_t o = 3 * ((size_t)P.pixel0 + ip);
    f3 a{acc[o], acc[o + 1], acc[o + 2]};
    for (uint32_t s = 0; s < slots; ++s) {
        const uint32_t p = s * P.npix + ip;
        const f3 L = xyz(path_words(P, p)[3]);
        if (isfinite(L.x + L.y + L.z)) a = a + L;
    }
    acc[o] = a.x; acc[o + 1] = a.y; acc[o + 2] = a.z;
}

__global__ __launch_bounds__(kBlock) void k_diff_scale(float* __restrict__ img, uint32_t n, float inv) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) img[i] = img[i] * inv;
}

// adjoint: keep each path's total radiance for the replay; clear its pending factors
__global__ __launch_bounds__(kBlock) void k_diff_keep(const PathState P, uint32_t n, DiffPaths DP) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const float4 w = path_words(P, p)[3];
        DP.ltr[p] = w.x; DP.ltg[p] = w.y; DP.ltb[p] = w.z;
        DP.ptex[p] = -1;
        DP.pflag[p] = 0;
    }
}

}  // namespace lrt
