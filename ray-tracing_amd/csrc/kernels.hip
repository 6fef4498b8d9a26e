// Kernels of liblinrt_hip.so (gfx950, wave64).
//
// Ray queries: one lane per ray, grid-stride over a device-resident count so
// the whole path-tracing frame is enqueued without host round trips.
//
// Path tracer (path_integrator.h:18-170) as a wavefront over SoA path state,
// one path per pixel per sample pass:
//   k_raygen -> [k_trace (extension) -> k_shade -> k_shadow]* -> k_accumulate
// k_shade appends surviving paths and next-event shadow rays to queues with a
// wave ballot + one atomic per wave.  Radiance is added to a path's L in the
// reference's order (emission, then NEE, then the next bounce), so queue order
// never changes results.
#include "kernels.h"

namespace lrt {

// ---------------------------------------------------------------------------
// wave-aggregated queue append: one atomic per wave, lanes get consecutive slots
LRT_D uint32_t wave_append(uint32_t* counter, bool want) {
    const uint64_t mask = __ballot(want);
    if (!want) return 0xffffffffu;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    return base + rank;
}

// the octant of a direction (sign bits of x, y, z)
LRT_D uint32_t dir_octant(const f3& d) { return (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u); }

// Block-aggregated append to one queue, for kernels whose every thread calls
// it the same number of times: per wave, a ballot counts the items; thread 0
// reserves the block's slots with one atomic.  (A wave's own append makes one
// atomic per 64 items on the same counter, which the memory-side atomic unit
// serialises.)  Returns the thread's slot, UINT32_MAX where it does not append.
template <uint32_t NT>
struct BlockAppend1 {
    uint32_t cnt[NT / 64], at[NT / 64];
};
template <uint32_t NT>
LRT_D uint32_t block_append1(BlockAppend1<NT>& sm, bool want, uint32_t* counter) {
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t m = __ballot(want);
    if (lane == 0) sm.cnt[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0u;
        for (uint32_t v = 0; v < NT / 64; ++v) {
            sm.at[v] = t;
            t += sm.cnt[v];
        }
        const uint32_t b = t ? atomicAdd(counter, t) : 0u;
        for (uint32_t v = 0; v < NT / 64; ++v) sm.at[v] += b;
    }
    __syncthreads();
    return want ? sm.at[wv] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) : 0xffffffffu;
}

// The same for two queues A and B with a key (0..7) per item: each queue's
// items are laid out key by key, then wave by wave and lane by lane, one
// atomic per queue and block.  Returns the thread's slots in A and B (where
// a / b).
template <uint32_t NT>
struct BlockAppend {
    uint32_t cnt[NT / 64][16], at[NT / 64][16];  // [0, 8): A's keys, [8, 16): B's
};
template <uint32_t NT>
LRT_D uint2 block_append2(BlockAppend<NT>& sm, bool a, uint32_t key_a, bool b, uint32_t key_b, uint32_t* ca,
                          uint32_t* cb) {
    constexpr uint32_t kW = NT / 64;
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t rank_a = 0u, rank_b = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        const bool in_a = a && key_a == k, in_b = b && key_b == k;
        const uint64_t ma = __ballot(in_a), mb = __ballot(in_b);
        if (lane == 0) {
            sm.cnt[wv][k] = (uint32_t)__popcll(ma);
            sm.cnt[wv][8 + k] = (uint32_t)__popcll(mb);
        }
        if (in_a) rank_a = (uint32_t)__popcll(ma & lt);
        if (in_b) rank_b = (uint32_t)__popcll(mb & lt);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t[2] = {0u, 0u};
        for (uint32_t k = 0; k < 16; ++k)
            for (uint32_t v = 0; v < kW; ++v) {
                sm.at[v][k] = t[k >> 3];
                t[k >> 3] += sm.cnt[v][k];
            }
        const uint32_t ba = t[0] ? atomicAdd(ca, t[0]) : 0u;
        const uint32_t bb = t[1] ? atomicAdd(cb, t[1]) : 0u;
        for (uint32_t v = 0; v < kW; ++v)
            for (uint32_t k = 0; k < 16; ++k) sm.at[v][k] += k < 8 ? ba : bb;
    }
    __syncthreads();
    return make_uint2(a ? sm.at[wv][key_a] + rank_a : 0u, b ? sm.at[wv][8 + key_b] + rank_b : 0u);
}

// ---------------------------------------------------------------------------
// ray queries
// rays -> closest hits (SoA rays: ox, oy, oz, dx, dy, dz, tmin, tmax);
// persistent like k_extend, rays fetched from *next (zero at launch)
__global__ __launch_bounds__(kBlock) LRT_TRACE_ATTR void k_trace_closest(const DevScene S, const float* __restrict__ rays,
                                                                         uint32_t n, int32_t* __restrict__ obj_prim,
                                                                         float* __restrict__ t_out, float* __restrict__ uv_out,
                                                                         uint32_t* __restrict__ next) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 1;
    for (;;) {
        const uint32_t got = wave_refill<false>(!active, next, n, more);
        if (got != 0xffffffffu) {
            i = got;
            lane_trace_start(T, Ray{f3{rays[i], rays[n + i], rays[2 * (size_t)n + i]},
                                    f3{rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]}},
                             rays[6 * (size_t)n + i], rays[7 * (size_t)n + i], -1);
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            obj_prim[2 * (size_t)i] = T.hit ? T.obj() : -1;
            obj_prim[2 * (size_t)i + 1] = T.hit ? T.prim : -1;
            t_out[i] = T.hit ? T.closest : 0.0f;
            if (uv_out) {
                uv_out[2 * (size_t)i] = T.hit ? T.u : 0.0f;
                uv_out[2 * (size_t)i + 1] = T.hit ? T.v : 0.0f;
            }
            active = false;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_trace_count(const DevScene S, const float* __restrict__ rays, uint32_t n,
                                                        uint32_t* __restrict__ nodes, uint32_t* __restrict__ refs,
                                                        uint32_t* __restrict__ tris) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Ray r{f3{rays[i], rays[n + i], rays[2 * (size_t)n + i]},
                    f3{rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]}};
        HitInfo h;
        Counters c{0, 0, 0};
        trace_closest<true>(S, r, rays[6 * (size_t)n + i], rays[7 * (size_t)n + i], h, lds, c);
        nodes[i] = c.nodes;
        refs[i] = c.refs;
        tris[i] = c.tris;
    }
}

// rays -> occluded flags (a surface of an emitter other than ignore[i] in
// [tmin, tmax]); persistent like k_shadow, rays fetched from *next
__global__ __launch_bounds__(kBlock) LRT_TRACE_ATTR void k_trace_occluded(const DevScene S, const float* __restrict__ rays,
                                                                          uint32_t n, const int32_t* __restrict__ ignore,
                                                                          uint8_t* __restrict__ out,
                                                                          uint32_t* __restrict__ next) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 1;
    for (;;) {
        const uint32_t got = wave_refill<false>(!active, next, n, more);
        if (got != 0xffffffffu) {
            i = got;
            lane_shadow_start(S, T, Ray{f3{rays[i], rays[n + i], rays[2 * (size_t)n + i]},
                                       f3{rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]}},
                              rays[6 * (size_t)n + i], rays[7 * (size_t)n + i], ignore ? ignore[i] : -1);
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            out[i] = shadow_blocked(S, T, ignore ? ignore[i] : -1) ? 1 : 0;
            active = false;
        }
    }
}

// ---------------------------------------------------------------------------
// BSDF / sky batches
__global__ void k_bsdf_eval(const Mats M, const int32_t* mat, const float* wi, const float* wo, const uint8_t* front,
                            uint32_t n, int meas, int mode, int comp, uint32_t mask, float* f_cos, float* pdf) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        BsdfCtx c;
        c.mode = mode; c.component = comp; c.mask = mask;
        SRec r;
        r.wi = ld3(wi + 3 * (size_t)i);
        r.wo = ld3(wo + 3 * (size_t)i);
        r.front = front[i] != 0;
        const EP e = bsdf_ep(M, mat[i], c, r, meas);
        f_cos[3 * (size_t)i] = e.f.x;
        f_cos[3 * (size_t)i + 1] = e.f.y;
        f_cos[3 * (size_t)i + 2] = e.f.z;
        pdf[i] = e.pdf;
    }
}

__global__ void k_bsdf_sample(const Mats M, const int32_t* mat, const float* wi, const uint8_t* front, const float* u,
                              const float* lobe_u, uint32_t n, int mode, int comp, uint32_t mask, float* weight,
                              float* wo, float* pdf, float* eta, uint32_t* stype, int32_t* scomp) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        BsdfCtx c;
        c.mode = mode; c.component = comp; c.mask = mask; c.lobe = lobe_u[i];
        SRec r;
        r.wi = ld3(wi + 3 * (size_t)i);
        r.wo = splat(0.0f);
        r.front = front[i] != 0;
        const f3 w = bsdf_sample(M, mat[i], c, r, u[2 * (size_t)i], u[2 * (size_t)i + 1]);
        weight[3 * (size_t)i] = w.x; weight[3 * (size_t)i + 1] = w.y; weight[3 * (size_t)i + 2] = w.z;
        wo[3 * (size_t)i] = r.wo.x; wo[3 * (size_t)i + 1] = r.wo.y; wo[3 * (size_t)i + 2] = r.wo.z;
        pdf[i] = r.pdf; eta[i] = r.eta; stype[i] = r.type; scomp[i] = r.comp;
    }
}

__global__ void k_sky_eval(const DevLight L, const float* dirs, uint32_t n, float* rgb) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const f3 c = sky_eval(L, ld3(dirs + 3 * (size_t)i));
        rgb[3 * (size_t)i] = c.x; rgb[3 * (size_t)i + 1] = c.y; rgb[3 * (size_t)i + 2] = c.z;
    }
}

// ---------------------------------------------------------------------------
// path tracer stages

// Philox dimension d of sample (pixel, sample): block d/4, lane d%4
LRT_D float dim_u(uint32_t pix, uint32_t smp, uint32_t d, uint2 key) {
    const uint4 v = philox(make_uint4(pix, smp, d >> 2, 0u), key);
    return u24(pick(v, d & 3u));
}

// the pixel at position r of a W x hp frame walked in 8 x 8 tiles, row-major
// tiles, the right and bottom tiles of a ragged frame narrower / lower (the
// inverse of the rank k_raygen computes)
LRT_D uint32_t tile_pixel(uint32_t r, uint32_t W, uint32_t hp) {
    const uint32_t ty = r / (8u * W), r2 = r - ty * 8u * W;
    const uint32_t th = min(8u, hp - ty * 8u);
    const uint32_t tx = r2 / (8u * th), r3 = r2 - tx * 8u * th;
    const uint32_t tw = min(8u, W - tx * 8u);
    return (ty * 8u + r3 / tw) * W + tx * 8u + r3 % tw;
}

__global__ __launch_bounds__(kBlock) void k_raygen(PathState P, const CameraDev cam, uint32_t W, uint32_t H, uint32_t row0,
                                                   uint32_t n, uint32_t smp, uint2 key, const uint32_t* __restrict__ sobol,
                                                   uint32_t* queue, uint32_t* counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        counters[0] = n;  // active queue length
        counters[1] = 0;  // next queue
        counters[2] = 0;  // shadow queue
        counters[3] = 0;
        counters[4] = 0;  // next extension ray to fetch (k_extend)
        counters[5] = 0;  // next shadow ray to fetch (k_shadow)
    }
    if (i < 1024) {  // shading bins: counts, cursors
        counters[kBinCount + i] = 0;
        counters[kBinCursor + i] = 0;
    }
    if (i < 2 * kXcds) counters[kXcdFetchExtend + i * kXcdStride] = 0;  // both sets of XCD fetch counters
    if (i >= n) return;
    const uint32_t ip = i % P.npix, slot = i / P.npix;
    const uint32_t x = ip % W, y = row0 + ip / W;
    const uint32_t pix = y * W + x;
    float jx, jy;
    if (sobol) {  // next_2d of the reference: (dimension 1, dimension 0), see k_shade
        const uint32_t ps = mix_bits(pix);
        jx = ref_sample(sobol, ps, smp + slot, 1u);
        jy = ref_sample(sobol, ps, smp + slot, 0u);
    } else {
        const uint4 v = philox(make_uint4(pix, smp + slot, 0u, 0u), key);
        jx = u24(v.x);
        jy = u24(v.y);
    }
    const float u = ((float)x + jx) / (float)(W - 1);
    const float vv = ((float)y + jy) / (float)(H - 1);
    const f3 d = normalize(ld3(cam.llc) + u * ld3(cam.horizontal) + vv * ld3(cam.vertical) - ld3(cam.origin));
    float4* w = path_words(P, i);
    w[0] = make_float4(cam.origin[0], cam.origin[1], cam.origin[2], 1.0f);  // eta 1
    w[1] = word(d, -1.0f);  // no solid-angle pdf yet (camera ray)
    w[2] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    w[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#if LRT_CAMERA_TILES
    // the camera queue in 8 x 8 pixel tiles (a wave traces one tile), row-major
    // tiles; the right and bottom tiles of a ragged frame are narrower / lower
    const uint32_t hp = P.npix / W, ly = ip / W;
    const uint32_t tw = min(8u, W - (x & ~7u)), th = min(8u, hp - (ly & ~7u));
    const uint32_t rank = (ly & ~7u) * W + (x & ~7u) * th + (ly & 7u) * tw + (x & 7u);
    queue[slot * P.npix + rank] = i;
#else
    queue[i] = i;
#endif
}

// extension rays of the active queue -> hit records by queue slot.
// Persistent: every lane takes a new ray as soon as its ray is finished (the
// wave fetches rays for its idle lanes from counters[4]).
__global__ __launch_bounds__(kBlock) LRT_PATH_TRACE_ATTR void k_extend(const DevScene S, const PathState P,
                                                                  const uint32_t* __restrict__ queue,
                                                                  uint32_t* __restrict__ counters, HitBuf Hb) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    const uint32_t n = counters[0];
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 2;  // wave_refill: first fill pending
#if LRT_XCD_FETCH_EXTEND
    const XcdSplit split{n, 0u};
#endif
    for (;;) {
#if LRT_XCD_FETCH_EXTEND
        const uint32_t got = wave_refill_xcd(!active, &counters[kXcdFetchExtend], split, more);
#else
        const uint32_t got = wave_refill(!active, &counters[4], n, more);
#endif
        if (got != 0xffffffffu) {
            i = got;
            const float4* w = path_words(P, queue[i]);
            lane_trace_start(T, Ray{xyz(w[0]), xyz(w[1])}, 0.001f, kInfF, -1);
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            uint32_t cell = 0u;
            if (T.hit && S.obj[T.obj()].shape == kShapeMesh)
                cell = grid_cell(S.obj_box + 6 * T.obj(), T.r.o + T.closest * T.r.d);
            hit_store(Hb, i, T.hit ? T.obj() : -1, T.prim, T.closest, T.u, T.v, cell);
            active = false;
        }
    }
}

// same rays with algorithmic-work counters (node visits, references, tests):
// the reference's full-window traversal into totals[0..2], the product's own
// box-clipped traversal into totals[6..8]; *count is the queue's length
__global__ __launch_bounds__(kBlock) void k_extend_count(const DevScene S, const PathState P, const uint32_t* __restrict__ queue,
                                                         const uint32_t* __restrict__ count, unsigned long long* totals) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    const uint32_t n = *count;
    uint64_t a = 0, b = 0, c3 = 0, oa = 0, ob = 0, oc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* w = path_words(P, queue[i]);
        const Ray r{xyz(w[0]), xyz(w[1])};
        HitInfo h;
        Counters c{0, 0, 0}, own{0, 0, 0};
        trace_closest<true, true>(S, r, 0.001f, kInfF, h, lds, c);
        trace_closest<true, false>(S, r, 0.001f, kInfF, h, lds, own);
        a += c.nodes; b += c.refs; c3 += c.tris;
        oa += own.nodes; ob += own.refs; oc += own.tris;
    }
    atomicAdd(&totals[0], (unsigned long long)a);
    atomicAdd(&totals[1], (unsigned long long)b);
    atomicAdd(&totals[2], (unsigned long long)c3);
    atomicAdd(&totals[6], (unsigned long long)oa);
    atomicAdd(&totals[7], (unsigned long long)ob);
    atomicAdd(&totals[8], (unsigned long long)oc);
}

// the shadow queue's algorithmic work into totals[3..5]: the reference's
// shadow test is a closest-hit query over [0.001, tmax] (path_integrator.h:113-116);
// the product's own work (box-clipped, ended by the first hit) into totals[9..11]
__global__ __launch_bounds__(kBlock) void k_shadow_count(const DevScene S, const ShadowBuf Sb,
                                                         const uint32_t* __restrict__ counters, unsigned long long* totals) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    const uint32_t n = counters[2];
    uint64_t a = 0, b = 0, c3 = 0, oa = 0, ob = 0, oc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Ray r{f3{Sb.ox[i], Sb.oy[i], Sb.oz[i]}, f3{Sb.dx[i], Sb.dy[i], Sb.dz[i]}};
        HitInfo h;
        Counters c{0, 0, 0}, own{0, 0, 0};
        trace_closest<true, true>(S, r, 0.001f, Sb.tmax[i], h, lds, c);
        trace_occluded<true>(S, r, 0.001f, Sb.tmax[i], Sb.ignore[i], lds, own);
        a += c.nodes; b += c.refs; c3 += c.tris;
        oa += own.nodes; ob += own.refs; oc += own.tris;
    }
    atomicAdd(&totals[3], (unsigned long long)a);
    atomicAdd(&totals[4], (unsigned long long)b);
    atomicAdd(&totals[5], (unsigned long long)c3);
    atomicAdd(&totals[9], (unsigned long long)oa);
    atomicAdd(&totals[10], (unsigned long long)ob);
    atomicAdd(&totals[11], (unsigned long long)oc);
}

struct SurfaceHit {
    f3 p, n;  // shading normal, face-forwarded (triangle_mesh.h:81-96)
    bool front;
    int material, emitter;
};

LRT_D SurfaceHit surface(const DevScene& S, const Ray& r, int obj, int prim, float t, float u, float v) {
    SurfaceHit s;
    const DevObject& ob = S.obj[obj];
    s.p = r.o + t * r.d;
    s.emitter = ob.emitter;
    if (ob.shape == kShapeMesh) {
        const uint32_t g = ob.tri_base + (uint32_t)prim;
        mesh_hit_normal(S, ob, prim, u, v, r.d, s.n, s.front);
        s.material = ob.per_tri_material ? (int)S.tri_mat[g] : ob.material;
    } else if (ob.shape == kShapeRect) {
        const f3 rn = ld3(ob.rn);
        s.front = dot(r.d, rn) < 0;
        s.n = s.front ? rn : -rn;
        s.material = -1;
    } else {
        const f3 outward = (s.p - ld3(ob.center)) / ob.radius;
        s.front = dot(r.d, outward) < 0;
        s.n = s.front ? outward : -outward;
        s.material = ob.material;
    }
    return s;
}

// Shading order: hit slots grouped by where the hit lies (escapes, emitter
// shapes, then cells of the hit mesh's box in Morton order).  The shading
// kernel appends the continued paths and their shadow rays in this order, so
// the next extension and shadow rays leave from nearby points, and the
// traversal of nearby rays shares the tree's nodes in the caches.  Hits of one
// region mostly share a material as well.  The results do not depend on the
// order: every path's state is its own and the pixel sums are taken per path
// afterwards.
LRT_D uint32_t shade_bin(const DevScene& S, const HitBuf& Hb, uint32_t i) {
    const uint4 h = Hb.rec[2 * (size_t)i];
    const int obj = (int)h.x;
    if (obj < 0) return 0u;
    if (S.obj[obj].shape == kShapeRect) return 1u;
    return 2u + reinterpret_cast<const uint32_t*>(Hb.rec + 2 * (size_t)i + 1)[1];
}

// counts per bin into counters[kBinCount + b]
__global__ __launch_bounds__(kBlock) void k_shade_bins(const DevScene S, const HitBuf Hb, uint32_t* __restrict__ counters) {
    __shared__ uint32_t cnt[kShadeBins];
    for (int b = threadIdx.x; b < kShadeBins; b += blockDim.x) cnt[b] = 0u;
    __syncthreads();
    const uint32_t n = counters[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&cnt[shade_bin(S, Hb, i)], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kShadeBins; b += blockDim.x)
        if (cnt[b]) atomicAdd(&counters[kBinCount + b], cnt[b]);
}

// bin starts: exclusive prefix sum of the counts (one block of 1024 threads)
__global__ __launch_bounds__(1024) void k_shade_scan(uint32_t* __restrict__ counters) {
    __shared__ uint32_t v[1024];
    const int t = threadIdx.x;
    const uint32_t mine = t < kShadeBins ? counters[kBinCount + t] : 0u;
    v[t] = mine;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint32_t add = t >= d ? v[t - d] : 0u;
        __syncthreads();
        v[t] += add;
        __syncthreads();
    }
    if (t < kShadeBins) counters[kBinStart + t] = v[t] - mine;
}

// order[]: slots bin by bin.  Each block takes a contiguous chunk of slots and
// reserves its range in every bin it uses with one atomic (cursors
// counters[kBinCursor + b]).
__global__ __launch_bounds__(kBlock) void k_shade_order(const DevScene S, const HitBuf Hb, uint32_t* __restrict__ counters,
                                                        uint32_t* __restrict__ order) {
    __shared__ uint32_t cnt[kShadeBins], base[kShadeBins];
    for (int b = threadIdx.x; b < kShadeBins; b += blockDim.x) cnt[b] = 0u;
    __syncthreads();
    const uint32_t n = counters[0];
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t b0 = min(n, blockIdx.x * per), b1 = min(n, b0 + per);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) atomicAdd(&cnt[shade_bin(S, Hb, i)], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kShadeBins; b += blockDim.x) {
        base[b] = cnt[b] ? counters[kBinStart + b] + atomicAdd(&counters[kBinCursor + b], cnt[b]) : 0u;
        cnt[b] = 0u;
    }
    __syncthreads();
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const uint32_t b = shade_bin(S, Hb, i);
        order[base[b] + atomicAdd(&cnt[b], 1u)] = i;
    }
}

// first-hit albedo for the denoiser's features: diffuse colour, the
// normal-incidence Fresnel reflectance of a conductor, 1 for dielectrics
LRT_D f3 albedo_of(const Mats& M, int m) {
    const lrt_material& A = M.m[m];
    switch (A.type) {
        case LRT_BSDF_DIFFUSE:
        case LRT_BSDF_PLASTIC:
        case LRT_BSDF_ROUGH_PLASTIC: return col(A.diffuse);
        case LRT_BSDF_CONDUCTOR: {
            f3 f;
            float* o = &f.x;
            for (int c = 0; c < 3; ++c) {
                const float e = A.cond_eta[c], k = A.cond_k[c];
                o[c] = ((e - 1.0f) * (e - 1.0f) + k * k) / ((e + 1.0f) * (e + 1.0f) + k * k) * A.specular_reflectance[c];
            }
            return f;
        }
        case LRT_BSDF_COATING:
        case LRT_BSDF_ROUGH_COATING: {
            const lrt_material& N = M.m[A.nested];
            if (N.type == LRT_BSDF_DIFFUSE || N.type == LRT_BSDF_PLASTIC || N.type == LRT_BSDF_ROUGH_PLASTIC) return col(N.diffuse);
            return splat(1.0f);
        }
        default: return splat(1.0f);
    }
}

// What one bounce of a path leaves for the traversal: whether the path goes
// on (its next ray is in its record) and its next-event shadow ray with the
// contribution the ray adds if unoccluded.
struct ShadeOut {
    bool cont, shadow;
    uint32_t octant;  // of the continuing ray's direction (sign bits x, y, z)
    f3 so, sd, contrib;
    float tmax;
    int ignore;
};

// One bounce of path_integrator::radiance (path_integrator.h:32-166) for
// path p at hit h: escape or emission with MIS, next-event estimation, BSDF
// sampling and Russian roulette; the path record is updated in place.
// MESH_LIGHTS: compiled with the mesh lights' pdf ray cast (launched only for
// scenes that have one, so the other scenes' shading keeps its registers)
template <bool MESH_LIGHTS>
LRT_D void shade_path(const DevScene& S, const DevLights& Ls, const Mats& M, const PathState& P, uint32_t p,
                      const HitRec& h, const ShadeParams& prm, ShadeOut& o) {
    o.cont = false;
    o.shadow = false;
    o.tmax = 0.0f;
    o.ignore = -1;
    float4* pw = path_words(P, p);
    const float4 w0 = pw[0], w1 = pw[1], w2 = pw[2], w3 = pw[3];
    const Ray r{xyz(w0), xyz(w1)};
    f3 T = xyz(w2);
    f3 L = xyz(w3);
    float eta = w0.w;
    const float sa_pdf = w1.w;
    const bool have = sa_pdf >= 0.0f;
    const int obj = h.obj;
    float* aov = prm.depth == 0 ? prm.aov_path : nullptr;  // first-hit features, as k_aov computes them
    if (obj < 0) {  // escape: environment with MIS (path_integrator.h:32-45)
        if (aov) {
            float* o = aov + 7 * (size_t)p;
            o[0] = o[1] = o[2] = 1.0f;
            o[3] = o[4] = o[5] = o[6] = 0.0f;
        }
        const f3 Le = env_eval(Ls, r.d);
        if (!have) {
            L = L + T * Le;
        } else {
            const float lp = pdf_emitter_direction<MESH_LIGHTS>(S, Ls, splat(0.0f), r.d);
            L = L + T * power_heuristic(sa_pdf, lp) * Le;
        }
    } else {
        const SurfaceHit sh = surface(S, r, obj, h.prim, h.t, h.u, h.v);
        if (aov) {
            const f3 al = sh.material >= 0 ? albedo_of(M, sh.material) : splat(1.0f);
            float* o = aov + 7 * (size_t)p;
            o[0] = al.x; o[1] = al.y; o[2] = al.z;
            o[3] = sh.n.x; o[4] = sh.n.y; o[5] = sh.n.z;
            o[6] = h.t;
        }
        if (sh.emitter >= 0 && sh.front) {  // path_integrator.h:47-59
            const f3 out = -normalize(r.d);
            const f3 Le = dot(sh.n, out) <= 0.0f ? splat(0.0f) : ld3(Ls.l[sh.emitter].radiance);
            if (!is_zero(Le)) {
                if (!have) {
                    L = L + T * Le;
                } else {
                    const float lp = pdf_emitter_direction<MESH_LIGHTS>(S, Ls, r.o, r.d);
                    L = L + T * power_heuristic(sa_pdf, lp) * Le;
                }
            }
        }
        if (sh.material >= 0) {
            const Frame fr = frame_from(sh.n);
            SRec sr;
            sr.wi = to_local(fr, -normalize(r.d));
            sr.front = sh.front;
            BsdfCtx ctx;
            // six sampler dimensions per bounce (path_integrator.h:80-83)
            const uint32_t pix = P.pixel0 + p % P.npix, smp = prm.sample + p / P.npix;
            const uint32_t d0 = 2u + 6u * prm.depth;
            float u6[6];  // light uv, lobe, BSDF uv, roulette
            if (prm.sobol) {
                // sampler::next_2d is vec2(next_1d(), next_1d()) (sampler.h:27-29),
                // whose arguments the compiled reference evaluates right to
                // left: a 2D sample is (dimension d + 1, dimension d)
                const uint32_t ps = mix_bits(pix);
                u6[0] = ref_sample(prm.sobol, ps, smp, d0 + 1u);
                u6[1] = ref_sample(prm.sobol, ps, smp, d0);
                u6[2] = ref_sample(prm.sobol, ps, smp, d0 + 2u);
                u6[3] = ref_sample(prm.sobol, ps, smp, d0 + 4u);
                u6[4] = ref_sample(prm.sobol, ps, smp, d0 + 3u);
                u6[5] = ref_sample(prm.sobol, ps, smp, d0 + 5u);
            } else {
                const uint4 A = philox(make_uint4(pix, smp, d0 >> 2, 0u), prm.key);
                const uint4 B = philox(make_uint4(pix, smp, (d0 >> 2) + 1u, 0u), prm.key);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const uint32_t d = d0 + (uint32_t)k;
                    u6[k] = u24((d >> 2) == (d0 >> 2) ? pick(A, d & 3u) : pick(B, d & 3u));
                }
            }
            ctx.lobe = u6[2];
            const uint32_t fl = bsdf_flags(M, sh.material);
            if (fl & (F_DIFFUSE | F_GLOSSY)) {  // NEE (path_integrator.h:86-128)
                DirectSample ds;
                if (sample_emitter_direction(S, Ls, sh.p, u6[0], u6[1], ds)) {
                    o.so = sh.p + sh.n * 1e-4f;
                    o.sd = ds.dir;
                    o.tmax = kInfF;
                    bool target = true;
                    if (!isinf(ds.dist)) {
                        const f3 to = ds.point - o.so;
                        const float dist = len(to);
                        if (dist < 1e-6f) target = false;
                        else { o.sd = to / dist; o.tmax = dist * (1.0f - 1e-4f); }
                    }
                    if (target) {
                        sr.wo = to_local(fr, o.sd);
                        const EP e = bsdf_ep(M, sh.material, ctx, sr, M_SOLID);
                        if (!is_zero(e.f)) {
                            const float lp = pdf_emitter_direction<MESH_LIGHTS>(S, Ls, sh.p, o.sd);
                            if (lp > 0.0f) {
                                const float w = power_heuristic(lp, e.pdf);
                                o.contrib = T * e.f * ds.radiance * (w / lp);
                                o.ignore = ds.emitter;
                                o.shadow = true;
                            }
                        }
                    }
                }
            }
            if (prm.depth + 1 < prm.max_depth) {  // last bounce does NEE only
                const f3 bw = bsdf_sample(M, sh.material, ctx, sr, u6[3], u6[4]);
                if (!is_zero(bw)) {
                    const f3 nd = to_world(fr, sr.wo);
                    T = T * bw;
                    eta *= sr.eta;
                    const bool hp = ((sr.type & F_SPECULAR) == 0) && sr.pdf > 0.0f;
                    o.cont = true;
                    if (prm.depth >= prm.rr_depth) {  // Russian roulette (path_integrator.h:155-166)
                        const float q = fminf(fmaxf(fmaxf(T.x, T.y), T.z) * eta * eta, 0.95f);
                        if (u6[5] >= q) o.cont = false;
                        else T = f3{T.x / q, T.y / q, T.z / q};
                    }
                    if (o.cont) {
                        o.octant = dir_octant(nd);
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

// one bounce of path_integrator::radiance for every queued path, in the
// shading order of the hit slots
template <bool MESH_LIGHTS>
__global__ __launch_bounds__(kShadeBlock) LRT_SHADE_ATTR void k_shade(const DevScene S, const DevLights Ls, const Mats M, PathState P,
                                                  const uint32_t* __restrict__ queue, uint32_t* __restrict__ next_queue,
                                                  uint32_t* counters, const HitBuf Hb, const uint32_t* __restrict__ order,
                                                  ShadowBuf Sb, ShadeParams prm) {
    const uint32_t n = counters[0];
    const uint32_t stride = gridDim.x * blockDim.x;
    // grid-stride with a grid-uniform trip count: the queue appends see full
    // waves and every thread of a block reaches its barriers
    const uint32_t n_round = (n + stride - 1) / stride * stride;
#if LRT_SHADE_BLOCK_APPEND
    __shared__ BlockAppend<kShadeBlock> ba;
#endif
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_round; j += stride) {
        ShadeOut o;
        o.cont = false;
        o.shadow = false;
        uint32_t p = 0;
        if (j < n) {
            const uint32_t i = order[j];  // hit slots grouped by material
            p = queue[i];
            shade_path<MESH_LIGHTS>(S, Ls, M, P, p, hit_load(Hb, i), prm, o);
        }
#if LRT_SHADE_BLOCK_APPEND
        // one append per block and queue, the continuing paths laid out by
        // direction octant: a traversal wave takes rays of one shading cell
        // and one octant
        const uint2 sl = block_append2<kShadeBlock>(ba, o.cont, o.cont ? o.octant : 0u, o.shadow,
                                                    o.shadow && LRT_SHADOW_OCTANTS ? dir_octant(o.sd) : 0u,
                                                    &counters[1], &counters[2]);
        if (o.cont) next_queue[sl.x] = p;
        const uint32_t ss = sl.y;
#else
        const uint32_t slot = wave_append(&counters[1], o.cont);
        if (o.cont) next_queue[slot] = p;
        const uint32_t ss = wave_append(&counters[2], o.shadow);
#endif
        if (o.shadow) {
            Sb.ox[ss] = o.so.x; Sb.oy[ss] = o.so.y; Sb.oz[ss] = o.so.z;
            Sb.dx[ss] = o.sd.x; Sb.dy[ss] = o.sd.y; Sb.dz[ss] = o.sd.z;
            Sb.tmax[ss] = o.tmax;
            Sb.cr[ss] = o.contrib.x; Sb.cg[ss] = o.contrib.y; Sb.cb[ss] = o.contrib.z;
            Sb.ignore[ss] = o.ignore;
            Sb.path[ss] = p;
        }
    }
}

// The last bounces of a pass once few paths are left (0 < counters[0] <=
// limit, else it returns at once): each lane takes a path at its hit and
// finishes it, bounce after bounce (shade_path, its shadow ray, its next ray),
// so the pass does not pay one launch tail (the slowest ray of a launch) per
// remaining bounce.  The arithmetic and the order of the radiance additions
// are the wavefront's (emission, then the shadow ray's contribution, per
// bounce).  Extension and shadow rays traced are added to stats[0] / stats[1];
// k_tail_done then empties the queue for the remaining wavefront launches.
template <bool MESH_LIGHTS>
__global__ __launch_bounds__(kBlock) LRT_TAIL_ATTR void k_tail(const DevScene S, const DevLights Ls, const Mats M, PathState P,
                                                 const uint32_t* __restrict__ queue, const uint32_t* __restrict__ counters,
                                                 const HitBuf Hb, ShadeParams prm, uint32_t limit,
                                                 unsigned long long* __restrict__ stats) {
    const uint32_t n = counters[0];
    if (n == 0 || n > limit) return;
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint32_t ext = 0, shd = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t p = queue[j];
        HitRec h = hit_load(Hb, j);
        ++ext;
        for (uint32_t d = prm.depth; d < prm.max_depth; ++d) {
            ShadeParams sp = prm;
            sp.depth = d;
            ShadeOut o;
            shade_path<MESH_LIGHTS>(S, Ls, M, P, p, h, sp, o);
            if (o.shadow) {
                ++shd;
                if (!trace_occluded(S, Ray{o.so, o.sd}, 0.001f, o.tmax, o.ignore, lds)) {
                    float4* w = path_words(P, p);
                    w[3] = word(xyz(w[3]) + o.contrib, 0.0f);
                }
            }
            if (!o.cont) break;
            const float4* w = path_words(P, p);
            HitInfo hi;
            Counters c{0, 0, 0};
            trace_closest<false>(S, Ray{xyz(w[0]), xyz(w[1])}, 0.001f, kInfF, hi, lds, c);
            h = HitRec{hi.obj, hi.prim, hi.t, hi.u, hi.v};
            ++ext;
        }
    }
    // one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        ext += __shfl_xor(ext, off);
        shd += __shfl_xor(shd, off);
    }
    if ((threadIdx.x & 63u) == 0u) {
        if (ext) atomicAdd(&stats[0], (unsigned long long)ext);
        if (shd) atomicAdd(&stats[1], (unsigned long long)shd);
    }
}

// after k_tail: a queue it finished is emptied
__global__ void k_tail_done(uint32_t* counters, uint32_t limit) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && counters[0] <= limit) counters[0] = 0;
}


// per-pixel average over the frame's camera rays of (albedo, normal, depth)
__global__ __launch_bounds__(kBlock) void k_aov(const DevScene S, const Mats M, const CameraDev cam, uint32_t W,
                                                uint32_t H, uint32_t row0, uint32_t n, uint32_t spp, uint32_t spp0,
                                                uint2 key, float* __restrict__ aov) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t x = i % W, y = row0 + i / W;
        const uint32_t pix = y * W + x;
        f3 al = splat(0.0f), nr = splat(0.0f);
        float dp = 0.0f;
        for (uint32_t s = 0; s < spp; ++s) {
            const uint4 v = philox(make_uint4(pix, spp0 + s, 0u, 0u), key);
            const float u = ((float)x + u24(v.x)) / (float)(W - 1);
            const float vv = ((float)y + u24(v.y)) / (float)(H - 1);
            const Ray r{ld3(cam.origin),
                        normalize(ld3(cam.llc) + u * ld3(cam.horizontal) + vv * ld3(cam.vertical) - ld3(cam.origin))};
            HitInfo h;
            Counters c{0, 0, 0};
            if (trace_closest<false>(S, r, 0.001f, kInfF, h, lds, c)) {
                const SurfaceHit sh = surface(S, r, h.obj, h.prim, h.t, h.u, h.v);
                al = al + (sh.material >= 0 ? albedo_of(M, sh.material) : splat(1.0f));
                nr = nr + sh.n;
                dp += h.t;
            } else {
                al = al + splat(1.0f);
            }
        }
        const float inv = 1.0f / (float)spp;
        float* o = aov + 7 * (size_t)pix;
        o[0] = al.x * inv; o[1] = al.y * inv; o[2] = al.z * inv;
        o[3] = nr.x * inv; o[4] = nr.y * inv; o[5] = nr.z * inv;
        o[6] = dp * inv;
    }
}


// per pixel of a pass, its samples' first-hit features (written by shading
// at depth 0) added in sample order to aov (from zero when `first`), and
// scaled by `scale` when it is not zero (the last pass: 1 / spp): the sums and
// the rounding of k_aov over the same camera rays
__global__ __launch_bounds__(kBlock) void k_aov_resolve(const PathState P, uint32_t slots,
                                                        const float* __restrict__ aov_path, float* __restrict__ aov,
                                                        int first, float scale) {
    const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= P.npix) return;
    float* o = aov + 7 * ((size_t)P.pixel0 + ip);
    float a[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) a[c] = first ? 0.0f : o[c];
    for (uint32_t s = 0; s < slots; ++s) {
        const float* v = aov_path + 7 * ((size_t)s * P.npix + ip);
#pragma unroll
        for (int c = 0; c < 7; ++c) a[c] += v[c];
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) o[c] = scale != 0.0f ? a[c] * scale : a[c];
}

// shadow rays: unoccluded ones add their contribution to the path radiance.
// Persistent like k_extend (rays fetched from counters[5]).
__global__ __launch_bounds__(kBlock) LRT_PATH_TRACE_ATTR void k_shadow(const DevScene S, PathState P,
                                                                  uint32_t* __restrict__ counters, const ShadowBuf Sb) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    const uint32_t n = counters[2];
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
            }
            active = false;
        }
    }
}

// One persistent launch over two queues: the shadow rays of this bounce
// (counters[2] records, any hit, unoccluded ones add their contribution to
// the path radiance) and the extension rays of the next bounce (counters[1]
// paths of next_queue, closest hit into the hit records).  Both only read the
// state shading wrote, so they can share a launch, and the launch ends with the
// slowest ray of the two queues instead of one tail per queue.  Items are
// fetched from counters[5]: shadow records first, then extension rays.
__global__ __launch_bounds__(kBlock) LRT_PATH_TRACE_ATTR void k_trace_queues(const DevScene S, PathState P,
                                                                        const uint32_t* __restrict__ next_queue,
                                                                        uint32_t* __restrict__ counters, HitBuf Hb,
                                                                        const ShadowBuf Sb) {
    __shared__ uint2 stack[kLdsStack * kBlock];
    uint2* lds = stack + threadIdx.x;
    uint2 spill[kMaxStack - kLdsStack];
    const uint32_t ns = counters[2], n = ns + counters[1];
    LaneTrace T;
    uint32_t i = 0;
    bool active = false;
    int more = 2;  // wave_refill: first fill pending
#if LRT_XCD_FETCH_QUEUES
    const XcdSplit split{ns, n - ns};
#endif
    for (;;) {
#if LRT_XCD_FETCH_QUEUES
        const uint32_t got = wave_refill_xcd(!active, &counters[kXcdFetchQueues], split, more);
#else
        const uint32_t got = wave_refill(!active, &counters[5], n, more);
#endif
        if (got != 0xffffffffu) {
            i = got;
            if (i < ns) {
                lane_shadow_start(S, T, Ray{f3{Sb.ox[i], Sb.oy[i], Sb.oz[i]}, f3{Sb.dx[i], Sb.dy[i], Sb.dz[i]}},
                                  0.001f, Sb.tmax[i], Sb.ignore[i]);
            } else {
                const float4* w = path_words(P, next_queue[i - ns]);
                lane_trace_start(T, Ray{xyz(w[0]), xyz(w[1])}, 0.001f, kInfF, -1);
            }
            active = true;
        }
        if (!__ballot(active)) {
            if (!more) break;
            continue;
        }
        if (active && lane_trace_round<false>(S, T, lds, spill)) {
            if (i < ns) {
                if (!shadow_blocked(S, T, Sb.ignore[i])) {
                    float4* w = path_words(P, Sb.path[i]);
                    w[3] = word(xyz(w[3]) + f3{Sb.cr[i], Sb.cg[i], Sb.cb[i]}, 0.0f);
                }
            } else {
                uint32_t cell = 0u;
                if (T.hit && S.obj[T.obj()].shape == kShapeMesh)
                    cell = grid_cell(S.obj_box + 6 * T.obj(), T.r.o + T.closest * T.r.d);
                hit_store(Hb, i - ns, T.hit ? T.obj() : -1, T.prim, T.closest, T.u, T.v, cell);
            }
            active = false;
        }
    }
}

// Copies the rays a bounce is about to trace into a caller's buffers, in the
// layout of the batch ray queries (SoA: ox oy oz dx dy dz tmin tmax with a
// stride of the buffer's capacity): the next bounce's extension rays (window
// [0.001, inf), path_integrator.h:140) and this bounce's shadow rays (window
// [0.001, tmax] and the emitter they ignore, path_integrator.h:104-116).
// fill[0] / fill[1] count the rays offered; those beyond a capacity are dropped.
__global__ void k_capture_rays(const PathState P, const uint32_t* __restrict__ next_queue,
                               const uint32_t* __restrict__ counters, const ShadowBuf Sb, float* __restrict__ ext,
                               uint32_t ext_cap, float* __restrict__ sh, int32_t* __restrict__ sh_ignore,
                               uint32_t sh_cap, uint32_t* __restrict__ fill) {
    const uint32_t ne = counters[1], ns = counters[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ne + ns; i += gridDim.x * blockDim.x) {
        float r[8];
        float* out;
        uint32_t cap, slot;
        if (i < ne) {
            const float4* w = path_words(P, next_queue[i]);
            const float4 a = w[0], b = w[1];
            r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = b.x; r[4] = b.y; r[5] = b.z; r[6] = 0.001f; r[7] = kInfF;
            out = ext;
            cap = ext_cap;
            slot = atomicAdd(&fill[0], 1u);
        } else {
            const uint32_t j = i - ne;
            r[0] = Sb.ox[j]; r[1] = Sb.oy[j]; r[2] = Sb.oz[j]; r[3] = Sb.dx[j]; r[4] = Sb.dy[j]; r[5] = Sb.dz[j];
            r[6] = 0.001f; r[7] = Sb.tmax[j];
            out = sh;
            cap = sh_cap;
            slot = atomicAdd(&fill[1], 1u);
            if (slot < cap) sh_ignore[slot] = Sb.ignore[j];
        }
        if (slot < cap)
#pragma unroll
            for (int k = 0; k < 8; ++k) out[(size_t)k * cap + slot] = r[k];
    }
}

// swap: next queue becomes active, counters reset (no memset nodes, graph safe)
__global__ void k_advance(uint32_t* counters, unsigned long long* stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        stats[0] += counters[0];  // extension rays traced at this depth
        stats[1] += counters[2];  // shadow rays traced at this depth
        counters[0] = counters[1];
        counters[1] = 0;
        counters[2] = 0;
        counters[3] = 0;
        counters[4] = 0;
        counters[5] = 0;
    }
    for (uint32_t b = threadIdx.x; b < 1024; b += blockDim.x) {  // shading bins: counts, cursors
        counters[kBinCount + b] = 0;
        counters[kBinCursor + b] = 0;
    }
    if (threadIdx.x < 2 * kXcds) counters[kXcdFetchExtend + threadIdx.x * kXcdStride] = 0;
}

// per pixel, the pass's samples in sample order
__global__ __launch_bounds__(kBlock) void k_accumulate(const PathState P, uint32_t slots, double* __restrict__ sum,
                                                       double* __restrict__ sumsq) {
    const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= P.npix) return;
    const size_t pix = (size_t)P.pixel0 + ip;
    double s[3] = {sum[3 * pix], sum[3 * pix + 1], sum[3 * pix + 2]};
    double q[3] = {0.0, 0.0, 0.0};
    if (sumsq) { q[0] = sumsq[3 * pix]; q[1] = sumsq[3 * pix + 1]; q[2] = sumsq[3 * pix + 2]; }
    for (uint32_t j = 0; j < slots; ++j) {
        const size_t i = (size_t)j * P.npix + ip;
        const float4 w = path_words(P, (uint32_t)i)[3];
        const float l[3] = {w.x, w.y, w.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s[c] += (double)l[c];
            q[c] += (double)l[c] * (double)l[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        sum[3 * pix + c] = s[c];
        if (sumsq) sumsq[3 * pix + c] = q[c];
    }
}

}  // namespace lrt

// photon tracing, sorting and the kd photon map (uses surface() and wave_append)
#include "device/photon.h"

// the photon kd tree built on the device (uses the radix sort of photon.h)
#include "device/photon_tree.h"

// differentiable rendering (uses surface())
#include "device/diff.h"
