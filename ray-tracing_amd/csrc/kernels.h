// Shared structures of the render kernels and their host launcher.
#pragma once

#include "device/bsdf.h"
#include "device/emit.h"
#include "device/trace.h"

namespace lrt {

constexpr int kBlock = 256;

struct CameraDev {
    float origin[3], llc[3], horizontal[3], vertical[3];
};

// State of the paths of one sample pass (one path per pixel), 64 bytes per
// path read and written as four 16-byte words, so that a kernel visiting the
// paths in any order (shading in bin order, shadow results) touches one
// cache line per path rather than one per field:
//   word 0: origin, eta            word 1: direction, solid-angle pdf (< 0: none)
//   word 2: throughput, unused     word 3: radiance, unused
struct PathState {
    float4* rec;        // 4 words per path
    uint32_t pixel0;    // pixel index of path 0
    uint32_t npix;      // pixels per sample slice: path p is pixel pixel0 + p % npix of sample slot p / npix
};

LRT_D float4* path_words(const PathState& P, uint32_t p) { return P.rec + 4 * (size_t)p; }
LRT_D f3 xyz(float4 w) { return f3{w.x, w.y, w.z}; }
LRT_D float4 word(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// hit records of the active queue, by queue slot: 32 bytes per slot,
// {obj, prim, t bits, u bits} then {v bits, spatial cell of the hit point,
// unused}, so that shading in any order reads one cache line per hit
struct HitBuf {
    uint4* rec;  // 2 words per slot
};

struct HitRec {
    int32_t obj, prim;  // obj < 0: no hit
    float t, u, v;
};

LRT_D void hit_store(const HitBuf& H, uint32_t i, int32_t obj, int32_t prim, float t, float u, float v,
                     uint32_t cell = 0u) {
    uint4* w = H.rec + 2 * (size_t)i;
    w[0] = make_uint4((uint32_t)obj, (uint32_t)prim, __float_as_uint(t), __float_as_uint(u));
    *reinterpret_cast<uint2*>(w + 1) = make_uint2(__float_as_uint(v), cell);
}

// Shading order bins: 0 escapes, 1 emitter shapes, then 2 + the hit point's
// cell in an 8 x 8 x 8 grid over the hit mesh's box, in Morton order
constexpr int kCellBits = 3;
constexpr int kShadeBins = 2 + (1 << (3 * kCellBits));
// counters[] beyond the queue counters: per-bin counts, cursors and starts
constexpr int kBinCount = 64, kBinCursor = kBinCount + 1024, kBinStart = kBinCursor + 1024;
// then the per-XCD fetch counters of k_extend and of k_trace_queues (8 each, one
// 128-byte line apart), zeroed by k_raygen and k_advance
constexpr int kXcdFetchExtend = kBinStart + 1024, kXcdFetchQueues = kXcdFetchExtend + 256;
constexpr size_t kCounterWords = kXcdFetchQueues + 256;
static_assert(kShadeBins <= 1024, "shading bins");

// Morton cell of point q in the box lo..hi (6 floats), 3 bits per axis
LRT_D uint32_t grid_cell(const float* box, f3 q) {
    uint32_t c = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float x = (comp(q, a) - box[a]) / (box[3 + a] - box[a]) * (float)(1 << kCellBits);
        const uint32_t k = (uint32_t)fminf(fmaxf(x, 0.0f), (float)((1 << kCellBits) - 1));
#pragma unroll
        for (int b = 0; b < kCellBits; ++b) c |= ((k >> b) & 1u) << (3 * b + a);
    }
    return c;
}

LRT_D HitRec hit_load(const HitBuf& H, uint32_t i) {
    const uint4* w = H.rec + 2 * (size_t)i;
    const uint4 a = w[0];
    const uint32_t v = reinterpret_cast<const uint32_t*>(w + 1)[0];
    return HitRec{(int32_t)a.x, (int32_t)a.y, __uint_as_float(a.z), __uint_as_float(a.w), __uint_as_float(v)};
}

// pending next-event rays with their contribution if unoccluded
struct ShadowBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmax, *cr, *cg, *cb;
    int32_t* ignore;
    uint32_t* path;
};

// ---- photon-map eye pass (photon_integrator::li as a wavefront) ----
// One work item per li() call: its ray, the factor its radiance is weighted
// with, and for a BSDF-sample child also the parent's direct-light part
// (emission found by the sampled ray, MIS-weighted against the emitter
// samples).  Contributions go to fixed-point per-sample accumulators, so the
// order in which items run does not change the result.
enum : uint32_t {
    kEyeEmission = 1u << 16,   // add emission at the hit (query.emission)
    kEyeLi = 1u << 17,         // evaluate li() at the hit
    kEyeDirect = 1u << 18,     // parent's BSDF-sampled direct light
    kEyeParentDiffuse = 1u << 19,
    kEyeDepthMask = 0xffffu,
};

struct EyeItems {
    uint32_t *sid, *path, *meta;        // sample slot, Philox path, depth | flags
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tr, *tg, *tb;                // weight of li(item)
    float *wr, *wg, *wb;                // parent throughput x BSDF weight (direct part)
    float* pdf;                         // BSDF sample pdf (direct part)
};

struct EyeShadow {
    uint32_t* sid;
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmax, *cr, *cg, *cb;
    int32_t* ignore;
};

// photon queries at pure-diffuse hits: point, normal, weight T, albedo / pi,
// meta = map | photon depth << 1
struct EyeQuery {
    uint32_t *sid, *meta;
    float *px, *py, *pz, *nx, *ny, *nz, *tr, *tg, *tb, *wr, *wg, *wb;
};

struct EyeParams {
    uint32_t max_depth, sample, pixel0;
    uint2 key;
    uint32_t maps;   // bit m: map m (0 global, 1 caustic) is not empty
    uint32_t cap_items, cap_shadow, cap_query;  // appends beyond these are dropped
    uint32_t npix;   // pixels per sample slice: item sid is pixel pixel0 + sid % npix of sample sample + sid / npix
};

struct ShadeParams {
    uint32_t depth, max_depth, rr_depth, sample;
    uint2 key;
    const uint32_t* sobol;  // non-null: the reference's Sobol sampler (direction numbers), else Philox
    float* aov_path;        // non-null: at depth 0 each path's first-hit features, 7 floats (k_aov_resolve)
};

}  // namespace lrt
