// Photon-map eye pass: photon_integrator::li (photon_integrator.h:67-296) as a
// wavefront.  Each li() call of the reference's recursion is a work item:
//
//   k_eye_raygen  -> items of the camera rays (depth 1)
//   [k_eye_extend -> k_eye_shade -> k_eye_shadow, k_eye_gather]*  generation by generation
//   k_eye_resolve -> per-pixel sums
//
// k_eye_shade does at one hit what li() does there: the parent's BSDF-sampled
// direct light (bsdf_samples, :237-254), emission, the photon gathers at a
// pure-diffuse hit facing the ray (:95-106), one child per specular component
// while depth < 4 (specular_lobes, :126-155), the emitter samples as shadow
// rays (emitter_samples, :157-210) and one child per BSDF sample (:212-266).
// A child's radiance is weighted by the product of the factors on its way
// from the camera; every contribution goes to a fixed-point per-sample
// accumulator (2^-32 steps), so results do not depend on the order in which
// items run.
//
// Sampler streams: every item draws from Philox {pixel, sample, d/4, path}
// with its own dimension counter d; the camera item is path 0 and starts at
// d = 2 (dimensions 0 and 1 are the pixel jitter), and a child's path is
// eye_child(parent, code) with code = c for specular component c and 16 + i
// for BSDF sample i.

namespace lrt {

// Fixed point of the per-sample accumulators: a step of 2^-32 (the sample's
// radiance is rounded to float when it is resolved) and every contribution
// clamped to 2^24, so one add moves an accumulator by at most 2^56 and a
// sample takes 128 adds at the clamp before the int64 could wrap (radiance in
// these scenes stays below 10^3, four orders under the clamp).
constexpr double kEyeFix = 4294967296.0;  // 2^32
constexpr float kEyeClamp = 16777216.0f;  // 2^24

LRT_D uint32_t eye_child(uint32_t parent, uint32_t code) {
    uint32_t h = parent * 0x9e3779b9u + (code + 1u) * 0x632be5abu;
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h ? h : 1u;
}

struct EyeRng {
    uint32_t pixel, sample, path, d;
    uint2 key;
    uint32_t block;
    uint4 cache;
    LRT_D float next() {
        const uint32_t b = d >> 2;
        if (b != block) {
            cache = philox(make_uint4(pixel, sample, b, path), key);
            block = b;
        }
        return u24(pick(cache, (d++) & 3u));
    }
};

// non-finite contributions are dropped; magnitudes are clamped to kEyeClamp
LRT_D void eye_add(unsigned long long* acc, uint32_t sid, f3 c) {
    const float v[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        if (v[ch] == 0.0f || !isfinite(v[ch])) continue;
        const float x = fminf(fmaxf(v[ch], -kEyeClamp), kEyeClamp);
        atomicAdd(acc + 3 * (size_t)sid + ch, (unsigned long long)(long long)llrint((double)x * kEyeFix));
    }
}

__global__ __launch_bounds__(kBlock) void k_eye_raygen(const CameraDev cam, uint32_t W, uint32_t H, uint32_t n,
                                                       EyeParams prm, EyeItems It) {
    // item i takes the pixels in 8 x 8 tiles (whole-row chunks): a wave traces
    // one tile, and the children it queues start in one patch of the scene.
    // A sample's id (pixel, sample) keys its streams and its accumulator, so
    // the order changes no result.
    const bool tiled = LRT_CAMERA_TILES && prm.npix % W == 0u && prm.pixel0 % W == 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = i % prm.npix, ip = tiled ? tile_pixel(r, W, prm.npix / W) : r;
        const uint32_t sid = i - r + ip;
        const uint32_t pix = prm.pixel0 + ip;
        const uint32_t x = pix % W, y = pix / W;
        const uint4 v = philox(make_uint4(pix, prm.sample + i / prm.npix, 0u, 0u), prm.key);
        const float jx = u24(v.x), jy = u24(v.y);
        const float u = ((float)x + jx) / (float)(W - 1);
        const float vv = ((float)y + jy) / (float)(H - 1);
        const f3 d = normalize(ld3(cam.llc) + u * ld3(cam.horizontal) + vv * ld3(cam.vertical) - ld3(cam.origin));
        It.sid[i] = sid;
        It.path[i] = 0u;
        It.meta[i] = 1u | kEyeEmission | kEyeLi;
        It.ox[i] = cam.origin[0]; It.oy[i] = cam.origin[1]; It.oz[i] = cam.origin[2];
        It.dx[i] = d.x; It.dy[i] = d.y; It.dz[i] = d.z;
        It.tr[i] = 1.0f; It.tg[i] = 1.0f; It.tb[i] = 1.0f;
    }
}

// closest hits of the items; persistent (lanes refilled from *next)
__global__ __launch_bounds__(kBlock) LRT_TRACE_ATTR void k_eye_extend(const DevScene S, const EyeItems It, uint32_t n,
                                                                      HitBuf Hb, uint32_t* __restrict__ next) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 2;  // wave_refill: first fill pending
    for (;;) {
        const uint32_t got = wave_refill(!active, next, n, more);
        if (got != 0xffffffffu) {
            i = got;
            lane_trace_start(T, Ray{f3{It.ox[i], It.oy[i], It.oz[i]}, f3{It.dx[i], It.dy[i], It.dz[i]}}, 0.001f, kInfF, -1);
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            hit_store(Hb, i, T.hit ? T.obj() : -1, T.prim, T.closest, T.u, T.v);
            active = false;
        }
    }
}

LRT_D void eye_push(EyeItems& O, uint32_t s, uint32_t sid, uint32_t path, uint32_t meta, f3 o, f3 d, f3 T, f3 W,
                    float pdf) {
    O.sid[s] = sid;
    O.path[s] = path;
    O.meta[s] = meta;
    O.ox[s] = o.x; O.oy[s] = o.y; O.oz[s] = o.z;
    O.dx[s] = d.x; O.dy[s] = d.y; O.dz[s] = d.z;
    O.tr[s] = T.x; O.tg[s] = T.y; O.tb[s] = T.z;
    O.wr[s] = W.x; O.wg[s] = W.y; O.wb[s] = W.z;
    O.pdf[s] = pdf;
}

// counters: [0] children, [1] shadow rays, [2] photon queries
template <bool MESH_LIGHTS>  // as k_shade
__global__ __launch_bounds__(kBlock) LRT_SHADE_ATTR void k_eye_shade(const DevScene S, const DevLights Ls, const Mats M,
                                                      const EyeItems It, uint32_t n, const HitBuf Hb, EyeItems Out,
                                                      EyeShadow Sh, EyeQuery Q, uint32_t* counters, EyeParams prm,
                                                      unsigned long long* __restrict__ acc) {
    const uint32_t stride = gridDim.x * blockDim.x;
    // a grid-uniform trip count, and loops over samples that run while any
    // lane of the block has one left: every thread reaches the appends'
    // barriers (one append per block and queue, not per wave)
    const uint32_t n_round = (n + stride - 1) / stride * stride;
#if LRT_EYE_BLOCK_APPEND
    __shared__ BlockAppend1<kBlock> ba;
#define LRT_EYE_APPEND(ctr, want) block_append1<kBlock>(ba, (want), (ctr))
#define LRT_EYE_ANY(c) __syncthreads_or(c)
#else
#define LRT_EYE_APPEND(ctr, want) wave_append((ctr), (want))
#define LRT_EYE_ANY(c) __any(c)
#endif
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
        const bool valid = i < n;
        uint32_t sid = 0, meta = 0, path = 0;
        Ray r{splat(0.0f), splat(0.0f)};
        f3 T = splat(0.0f);
        SurfaceHit sh;
        sh.material = -1;
        sh.emitter = -1;
        bool hit = false;
        if (valid) {
            sid = It.sid[i];
            meta = It.meta[i];
            path = It.path[i];
            r = Ray{f3{It.ox[i], It.oy[i], It.oz[i]}, f3{It.dx[i], It.dy[i], It.dz[i]}};
            T = f3{It.tr[i], It.tg[i], It.tb[i]};
            const HitRec h = hit_load(Hb, i);
            hit = h.obj >= 0;
            if (hit) sh = surface(S, r, h.obj, h.prim, h.t, h.u, h.v);
        }
        const uint32_t depth = meta & kEyeDepthMask;
        // (A) the parent's BSDF-sampled direct light (photon_integrator.h:237-254)
        if (valid && (meta & kEyeDirect)) {
            f3 emission = splat(0.0f);
            float lp = 0.0f;
            if (hit && sh.emitter >= 0 && sh.front) {
                const f3 out = -normalize(r.d);
                emission = dot(sh.n, out) <= 0.0f ? splat(0.0f) : ld3(Ls.l[sh.emitter].radiance);
                lp = pdf_emitter_direction<MESH_LIGHTS>(S, Ls, r.o, r.d);
            } else if (!hit) {
                emission = env_eval(Ls, r.d);
                lp = pdf_emitter_direction<MESH_LIGHTS>(S, Ls, r.o, r.d);
            }
            if (!is_zero(emission)) {
                const bool first = depth == 2u;
                const float ce = first ? 16.0f : 1.0f;
                const int cb = first ? ((meta & kEyeParentDiffuse) ? 16 : 32) : 1;
                const float bw = first ? 1.0f / (float)cb : 1.0f;
                const float mis = power_heuristic(It.pdf[i] * (float)cb, lp * ce) * bw;
                eye_add(acc, sid, f3{It.wr[i], It.wg[i], It.wb[i]} * emission * mis);
            }
        }
        // (B) li() at this hit
        bool shade = false;
        uint32_t fl = 0;
        bool diffuse = false, specular = false, smooth = false, exhaustive = false;
        if (valid && (meta & kEyeLi)) {
            if (!hit) {
                if (meta & kEyeEmission) eye_add(acc, sid, T * env_eval(Ls, r.d));
            } else {
                if ((meta & kEyeEmission) && sh.emitter >= 0 && sh.front) {
                    const f3 out = -normalize(r.d);
                    if (dot(sh.n, out) > 0.0f) eye_add(acc, sid, T * ld3(Ls.l[sh.emitter].radiance));
                }
                if (depth < prm.max_depth && sh.material >= 0) {
                    shade = true;
                    fl = bsdf_flags(M, sh.material);
                    diffuse = (fl & (F_DIFFUSE | F_GLOSSY | F_SPECULAR)) == F_DIFFUSE;
                    specular = (fl & F_SPECULAR) != 0;
                    smooth = (fl & (F_DIFFUSE | F_GLOSSY)) != 0;
                    exhaustive = specular && depth < 4u;
                }
            }
        }
        // photon gathers (photon_integrator.h:95-106)
        const bool gather = shade && diffuse && dot(sh.n, r.d) < 0.0f;
        f3 refl = splat(0.0f);
        if (gather) refl = col(M.m[sh.material].diffuse) * (1.0f / kPiF);
        for (uint32_t m = 0; m < 2; ++m) {
            const bool q = gather && ((prm.maps >> m) & 1u);
            const uint32_t qs = LRT_EYE_APPEND(&counters[2], q);
            if (q && qs < prm.cap_query) {
                Q.sid[qs] = sid;
                Q.meta[qs] = m | ((prm.max_depth - depth) << 1);
                Q.px[qs] = sh.p.x; Q.py[qs] = sh.p.y; Q.pz[qs] = sh.p.z;
                Q.nx[qs] = sh.n.x; Q.ny[qs] = sh.n.y; Q.nz[qs] = sh.n.z;
                Q.tr[qs] = T.x; Q.tg[qs] = T.y; Q.tb[qs] = T.z;
                Q.wr[qs] = refl.x; Q.wg[qs] = refl.y; Q.wb[qs] = refl.z;
            }
        }
        Frame fr;
        f3 wi = splat(0.0f);
        if (shade) {
            fr = frame_from(sh.n);
            wi = to_local(fr, -normalize(r.d));
        }
        // specular components while depth < 4 (photon_integrator.h:126-155)
        const int ncomp = (shade && exhaustive) ? bsdf_component_count(M, sh.material) : 0;
        for (int c = 0; LRT_EYE_ANY(c < ncomp); ++c) {
            bool push = false;
            SRec sr;
            f3 w = splat(0.0f);
            if (c < ncomp && (bsdf_component_flags(M, sh.material, c) & F_SPECULAR) != 0) {
                BsdfCtx ctx;
                ctx.component = c;
                sr.wi = wi;
                sr.front = sh.front;
                w = bsdf_sample(M, sh.material, ctx, sr, 0.5f, 0.5f);
                push = !is_zero(w);
            }
            const uint32_t s = LRT_EYE_APPEND(&counters[0], push);
            if (push && s < prm.cap_items)
                eye_push(Out, s, sid, eye_child(path, (uint32_t)c), (depth + 1u) | kEyeEmission | kEyeLi, sh.p,
                         to_world(fr, sr.wo), T * w, splat(0.0f), 0.0f);
        }
        EyeRng rng{prm.pixel0 + sid % prm.npix, prm.sample + sid / prm.npix, path, path == 0u ? 2u : 0u, prm.key, 0xffffffffu,
                   make_uint4(0, 0, 0, 0)};
        const bool first = depth == 1u;
        // emitter samples (photon_integrator.h:157-210)
        const int ce = (shade && smooth) ? (first ? 16 : 1) : 0;
        const int cbn = first ? (diffuse ? 16 : 32) : 1;
        const float ew = first ? 1.0f / 16.0f : 1.0f;
        const uint32_t mask = (!specular || exhaustive) ? (F_DIFFUSE | F_GLOSSY) : 0xffffffffu;
        for (int k = 0; LRT_EYE_ANY(k < ce); ++k) {
            bool push = false;
            f3 so = splat(0.0f), sd = splat(0.0f), contrib = splat(0.0f);
            float tmax = kInfF;
            int ignore = -1;
            if (k < ce) {
                const float sx = rng.next(), sy = rng.next();
                DirectSample ds;
                if (sample_emitter_direction(S, Ls, sh.p, sx, sy, ds)) {
                    so = sh.p + sh.n * 1e-4f;
                    sd = ds.dir;
                    bool target = true;
                    if (!isinf(ds.dist)) {
                        const f3 to = ds.point - so;
                        const float dist = len(to);
                        if (dist < 1e-6f) target = false;
                        else { sd = to / dist; tmax = dist * (1.0f - 1e-4f); }
                    }
                    if (target) {
                        BsdfCtx ctx;
                        ctx.mask = mask;
                        SRec sr;
                        sr.wi = wi;
                        sr.front = sh.front;
                        sr.wo = to_local(fr, sd);
                        const EP e = bsdf_ep(M, sh.material, ctx, sr, M_SOLID);
                        if (!is_zero(e.f)) {
                            const float lp = pdf_emitter_direction<MESH_LIGHTS>(S, Ls, sh.p, sd);
                            if (lp > 0.0f) {
                                const float w = power_heuristic(lp * (float)ce, e.pdf * (float)cbn) * ew;
                                contrib = T * (e.f * ds.radiance * (w / lp));
                                ignore = ds.emitter;
                                push = true;
                            }
                        }
                    }
                }
            }
            const uint32_t s = LRT_EYE_APPEND(&counters[1], push);
            if (push && s < prm.cap_shadow) {
                Sh.sid[s] = sid;
                Sh.ox[s] = so.x; Sh.oy[s] = so.y; Sh.oz[s] = so.z;
                Sh.dx[s] = sd.x; Sh.dy[s] = sd.y; Sh.dz[s] = sd.z;
                Sh.tmax[s] = tmax;
                Sh.cr[s] = contrib.x; Sh.cg[s] = contrib.y; Sh.cb[s] = contrib.z;
                Sh.ignore[s] = ignore;
            }
        }
        // BSDF samples (photon_integrator.h:212-266)
        const bool sample_direct = shade && (smooth || (specular && !exhaustive));
        const bool sample_indirect = shade && !diffuse && (smooth || (specular && !exhaustive));
        const int cb = (sample_direct || sample_indirect) ? cbn : 0;
        const float bw = first ? 1.0f / (float)cbn : 1.0f;
        const uint32_t child_meta = (depth + 1u) | (sample_direct ? kEyeDirect : 0u) | (sample_indirect ? kEyeLi : 0u) |
                                    (diffuse ? kEyeParentDiffuse : 0u);
        for (int k = 0; LRT_EYE_ANY(k < cb); ++k) {
            bool push = false;
            SRec sr;
            f3 w = splat(0.0f);
            if (k < cb) {
                BsdfCtx ctx;
                ctx.lobe = rng.next();
                ctx.mask = mask;
                sr.wi = wi;
                sr.front = sh.front;
                const float ux = rng.next(), uy = rng.next();
                w = bsdf_sample(M, sh.material, ctx, sr, ux, uy);
                push = !is_zero(w) && sr.pdf > 0.0f;
            }
            const uint32_t s = LRT_EYE_APPEND(&counters[0], push);
            if (push && s < prm.cap_items)
                eye_push(Out, s, sid, eye_child(path, 16u + (uint32_t)k), child_meta, sh.p, to_world(fr, sr.wo),
                         (T * w) * bw, T * w, sr.pdf);
        }
    }
}

#undef LRT_EYE_APPEND
#undef LRT_EYE_ANY

// unoccluded shadow rays add their contributions; persistent (lanes
// refilled from *next)
__global__ __launch_bounds__(kBlock) LRT_TRACE_ATTR void k_eye_shadow(const DevScene S, const EyeShadow Sh, uint32_t n,
                                                                      unsigned long long* __restrict__ acc,
                                                                      uint32_t* __restrict__ next) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 2;  // wave_refill: first fill pending
    for (;;) {
        const uint32_t got = wave_refill(!active, next, n, more);
        if (got != 0xffffffffu) {
            i = got;
            lane_shadow_start(S, T, Ray{f3{Sh.ox[i], Sh.oy[i], Sh.oz[i]}, f3{Sh.dx[i], Sh.dy[i], Sh.dz[i]}}, 0.001f,
                              Sh.tmax[i], Sh.ignore[i]);
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            if (!shadow_blocked(S, T, Sh.ignore[i])) eye_add(acc, Sh.sid[i], f3{Sh.cr[i], Sh.cg[i], Sh.cb[i]});
            active = false;
        }
    }
}

// sort key of a gather query: its map in bit 27 above the 27-bit Morton code
// of its position in the scene box (9 bits per axis), so that the queries of
// one gather wave lie close together and walk the same part of one kd tree
LRT_D uint32_t spread3(uint32_t v) {  // bit i of v -> bit 3 i
    v &= 0x1ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(kBlock) void k_eye_query_keys(const EyeQuery Q, uint32_t n, f3 lo, f3 scale,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const f3 u = (f3{Q.px[i], Q.py[i], Q.pz[i]} - lo) * scale;
        const uint32_t x = (uint32_t)fminf(fmaxf(u.x, 0.0f), 511.0f);
        const uint32_t y = (uint32_t)fminf(fmaxf(u.y, 0.0f), 511.0f);
        const uint32_t z = (uint32_t)fminf(fmaxf(u.z, 0.0f), 511.0f);
        keys[i] = ((Q.meta[i] & 1u) << 27) | (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
        vals[i] = i;
    }
}

// gather queries in the given order (a permutation of 0..n-1), one query per
// group of 8 lanes; S = key slots per lane (k <= 8 S)
template <int S>
__global__ __launch_bounds__(kGatherBlock) LRT_GATHER_ATTR void k_eye_gather(const PhotonMapDev G0, const PhotonMapDev G1, const EyeQuery Q,
                                                             const uint32_t* __restrict__ order, uint32_t n,
                                                             unsigned long long* __restrict__ acc) {
    __shared__ uint32_t stacks[(kGatherBlock / kGroup) * kPkStack];
    uint32_t* stack = stacks + (threadIdx.x / kGroup) * kPkStack;
    const uint32_t groups = gridDim.x * (kGatherBlock / kGroup);
    for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; t < n; t += groups) {
        const uint32_t i = order[t];
        const uint32_t meta = Q.meta[i];
        Counters c{0, 0, 0};
        const f3 E = photon_estimate_group<S, false>((meta & 1u) ? G1 : G0, f3{Q.px[i], Q.py[i], Q.pz[i]},
                                                     f3{Q.nx[i], Q.ny[i], Q.nz[i]}, (int)(meta >> 1), stack, c);
        if ((threadIdx.x & (kGroup - 1)) == 0)
            eye_add(acc, Q.sid[i], f3{Q.tr[i], Q.tg[i], Q.tb[i]} * (E * f3{Q.wr[i], Q.wg[i], Q.wb[i]}));
    }
}

// the same gathers with work counters (nodes visited, photon positions
// tested, photons summed) instead of contributions
template <int S>
__global__ __launch_bounds__(kGatherBlock) void k_eye_gather_count(const PhotonMapDev G0, const PhotonMapDev G1,
                                                                   const EyeQuery Q, uint32_t n,
                                                                   unsigned long long* __restrict__ totals) {
    __shared__ uint32_t stacks[(kGatherBlock / kGroup) * kPkStack];
    uint32_t* stack = stacks + (threadIdx.x / kGroup) * kPkStack;
    const uint32_t groups = gridDim.x * (kGatherBlock / kGroup);
    uint64_t a = 0, b = 0, c3 = 0;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; i < n; i += groups) {
        const uint32_t meta = Q.meta[i];
        Counters c{0, 0, 0};
        photon_estimate_group<S, true>((meta & 1u) ? G1 : G0, f3{Q.px[i], Q.py[i], Q.pz[i]}, f3{Q.nx[i], Q.ny[i], Q.nz[i]},
                                       (int)(meta >> 1), stack, c);
        a += c.nodes; b += c.refs; c3 += c.tris;
    }
    atomicAdd(&totals[0], (unsigned long long)a);
    atomicAdd(&totals[1], (unsigned long long)b);
    atomicAdd(&totals[2], (unsigned long long)c3);
}

// per-sample radiance (rounded to FP32 like the reference's color) into the
// pixel sums, a pixel's samples in sample order (slot j holds sample j of the
// chunk's npix pixels)
__global__ __launch_bounds__(kBlock) void k_eye_resolve(const unsigned long long* __restrict__ acc, uint32_t npix,
                                                        uint32_t slots, uint32_t pixel0, double* __restrict__ sum,
                                                        double* __restrict__ sumsq) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        for (uint32_t j = 0; j < slots; ++j) {
            const size_t sid = (size_t)j * npix + i;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const double L = (double)(float)((double)(long long)acc[3 * sid + ch] / kEyeFix);
                const size_t o = 3 * ((size_t)pixel0 + i) + ch;
                sum[o] += L;
                if (sumsq) sumsq[o] += L * L;
            }
        }
    }
}

__global__ void k_fill_u64(unsigned long long* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

}  // namespace lrt
