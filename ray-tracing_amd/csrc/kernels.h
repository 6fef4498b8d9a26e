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

// SoA state of the paths of one sample pass (one path per pixel)
struct PathState {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tr, *tg, *tb, *lr, *lg, *lb;
    float *eta, *pdf;   // pdf < 0: no solid-angle pdf (camera ray or delta lobe)
    uint32_t pixel0;    // pixel index of path 0
    uint32_t npix;      // pixels per sample slice: path p is pixel pixel0 + p % npix of sample slot p / npix
};

// hit records of the active queue, by queue slot
struct HitBuf {
    int32_t *obj, *prim;
    float *t, *u, *v;
};

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
};

struct ShadeParams {
    uint32_t depth, max_depth, rr_depth, sample;
    uint2 key;
    const uint32_t* sobol;  // non-null: the reference's Sobol sampler (direction numbers), else Philox
};

}  // namespace lrt
