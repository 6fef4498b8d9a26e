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

struct ShadeParams {
    uint32_t depth, max_depth, rr_depth, sample;
    uint2 key;
};

}  // namespace lrt
