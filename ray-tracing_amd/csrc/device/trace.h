// kd-tree traversal and primitive tests on the device.
//
// Traversal restates sah_kdtree::intersect (sah_kdtree.h:61-121) over the
// 8-byte flattened nodes: front-to-back with a (node, mint) stack, the far
// child pushed only when the split plane lies inside the current [mint, maxt].
// A stack entry's maxt is not stored: it always equals the mint of the entry
// below it (or the initial maxt for the bottom entry), because an entry is
// pushed with the current maxt and the current maxt is always the mint of the
// top entry.  Entries live in LDS (kLdsStack per lane) and spill to a private
// array beyond that.
#pragma once

#include "common.h"

namespace lrt {

enum : int { kShapeMesh = 0, kShapeRect = 1, kShapeSphere = 2 };

struct DevObject {
    int32_t shape, material, emitter, per_tri_material;
    uint32_t node_base, tri_base, n_tris, pad;
    float corner[3], eu[3], ev[3], rn[3];  // rectangle
    float center[3], radius;               // sphere
    float area;
};

constexpr int kMaxObjects = 16;

struct DevScene {
    const uint2* nodes;        // flattened kd nodes of every mesh object
    const uint32_t* refs;      // global triangle index per leaf reference
    const float* tris;         // 9 floats per triangle: p0, p1-p0, p2-p0 (degenerate: edges zeroed)
    const float* leaf_tris;    // the same 9 floats per leaf reference, in reference order: a leaf's
                               // triangles lie together, so its tests read a few whole cache lines
    const float* cnormals;     // 9 floats per triangle: corner normals
    const uint16_t* tri_mat;   // per triangle material (or null)
    const float* obj_box;      // per object 6 floats: enlarged AABB of a mesh (min xyz, max xyz)
    int32_t n_objects;
    int32_t stack_depth;       // deepest root-to-leaf path over all trees
    uint32_t nonconvex_emitters;  // bit e: emitter e's shape is a mesh (its own surface can hide a blocker)
    uint32_t pad;
    DevObject obj[kMaxObjects];
};

// An emitter whose surface may lie between a shading point and the point
// sampled on it: the shadow test must be the reference's closest-hit form
// (blocked iff the nearest surface in the window is not the emitter's,
// path_integrator.h:113-116) rather than an any-hit that skips the emitter.
LRT_D bool emitter_nonconvex(const DevScene& S, int e) {
    return e >= 0 && e < 32 && ((S.nonconvex_emitters >> e) & 1u) != 0u;
}

#ifndef LRT_LDS_STACK
#define LRT_LDS_STACK 8
#endif
#ifndef LRT_LEAF_BATCH
#define LRT_LEAF_BATCH 2
#endif
// stack entries kept in LDS per lane (the rest spill to private memory), and
// triangles whose reference and vertex loads are issued together in a leaf
// waves per SIMD the traversal kernels are compiled to fit (their registers
// are capped accordingly; 0 = the compiler's choice), and the same for k_shade
#ifndef LRT_TRACE_WAVES
#define LRT_TRACE_WAVES 7
#endif
#if LRT_TRACE_WAVES > 0
#define LRT_TRACE_ATTR __attribute__((amdgpu_waves_per_eu(LRT_TRACE_WAVES)))
#else
#define LRT_TRACE_ATTR
#endif
// the path tracer's traversal launches (k_extend, k_trace_queues, k_shadow):
// measured faster at 8 waves/SIMD on the 1080p frame, while the batch ray
// queries are faster at 7
#ifndef LRT_PATH_TRACE_WAVES
#define LRT_PATH_TRACE_WAVES 8
#endif
#if LRT_PATH_TRACE_WAVES > 0
#define LRT_PATH_TRACE_ATTR __attribute__((amdgpu_waves_per_eu(LRT_PATH_TRACE_WAVES)))
#else
#define LRT_PATH_TRACE_ATTR
#endif
#ifndef LRT_SHADE_WAVES
#define LRT_SHADE_WAVES 2
#endif
#if LRT_SHADE_WAVES > 0
#define LRT_SHADE_ATTR __attribute__((amdgpu_waves_per_eu(LRT_SHADE_WAVES)))
#else
#define LRT_SHADE_ATTR
#endif
// k_tail (shading and both traversals in one lane per path): more waves than
// k_shade's 2 hide more of its dependent loads, at the price of spills
// (denoised 1-spp 1080p frame: 2 -> 14.9, 3 -> 14.3, 4 -> 14.3, 5 -> 14.0,
// 6 -> 14.1 ms; 2 spp 23.4 -> 21.8 ms at 5)
#ifndef LRT_TAIL_WAVES
#define LRT_TAIL_WAVES 5
#endif
#if LRT_TAIL_WAVES > 0
#define LRT_TAIL_ATTR __attribute__((amdgpu_waves_per_eu(LRT_TAIL_WAVES)))
#else
#define LRT_TAIL_ATTR
#endif
constexpr int kLdsStack = LRT_LDS_STACK;

constexpr int kLeafBatch = LRT_LEAF_BATCH;
constexpr int kMaxStack = 48;

struct Ray {
    f3 o, d;
};

struct HitInfo {
    float t;
    int32_t obj, prim;
    float u, v;
};

struct TriRec {
    f3 p0, e1, e2;
};
LRT_D TriRec load_tri(const float* __restrict__ tri) { return TriRec{ld3(tri), ld3(tri + 3), ld3(tri + 6)}; }

// Moller-Trumbore restated from triangle_mesh.h:43-79 with the voxel window
LRT_D bool tri_test(const TriRec& T, const Ray& r, float t0, float t1, float& t, float& u, float& v) {
    const f3 p0 = T.p0, e1 = T.e1, e2 = T.e2;
    const f3 pv = cross(r.d, e2);
    const float det = dot(e1, pv);
    if (fabsf(det) < 1e-8f) return false;
    const float inv = 1.0f / det;
    const f3 tv = r.o - p0;
    u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const f3 qv = cross(tv, e1);
    v = dot(r.d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qv) * inv;
    return !(t < t0 || t > t1);
}

// rectangle.h:18-48
LRT_D bool rect_test(const DevObject& ob, const Ray& r, float tmin, float tmax, float& t, float& a, float& b) {
    const f3 rn = ld3(ob.rn), c = ld3(ob.corner), eu = ld3(ob.eu), ev = ld3(ob.ev);
    const float denom = dot(rn, r.d);
    if (fabsf(denom) < 1e-8f) return false;
    t = dot(c - r.o, rn) / denom;
    if (t < tmin || t > tmax) return false;
    const f3 p = r.o + t * r.d;
    const f3 d = p - c;
    const float uu = dot(eu, eu), vv = dot(ev, ev), uv = dot(eu, ev);
    const float du = dot(d, eu), dv = dot(d, ev);
    const float det = uu * vv - uv * uv;
    if (fabsf(det) < 1e-12f) return false;
    a = (vv * du - uv * dv) / det;
    b = (uu * dv - uv * du) / det;
    return !(a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f);
}

// sphere.h:21-44
LRT_D bool sphere_test(const DevObject& ob, const Ray& r, float tmin, float tmax, float& t) {
    const f3 oc = r.o - ld3(ob.center);
    const float a = len2(r.d);
    const float hb = dot(oc, r.d);
    const float c = len2(oc) - ob.radius * ob.radius;
    const float disc = hb * hb - a * c;
    if (disc < 0) return false;
    const float sq = sqrtf(disc);
    float root = (-hb - sq) / a;
    if (root < tmin || root > tmax) {
        root = (-hb + sq) / a;
        if (root < tmin || root > tmax) return false;
    }
    t = root;
    return true;
}

struct Counters {
    uint32_t nodes, refs, tris;
};

#ifndef LRT_BOX_CLIP
#define LRT_BOX_CLIP 1
#endif

// Narrows [t0, t1] of the ray to the (enlarged) box of mesh object o and
// returns false when nothing is left.  The box holds every triangle with a
// margin of 1e-3 of its size (plus 2^-16 of |t| for the slab distances'
// rounding), so no triangle hit lies outside the narrowed window.  Traversing
// the mesh's tree over the narrowed window visits the reference's leaves in
// the reference's order minus those whose window lies outside the box, and
// accepts the same hits: the result, ties included, is unchanged.
LRT_D bool mesh_box_clip(const DevScene& S, int o, const Ray& r, const float inv[3], float& t0, float& t1) {
    const float* b = S.obj_box + 6 * o;
    float lo = t0, hi = t1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float org = comp(r.o, a), d = comp(r.d, a);
        if (d == 0.0f) {
            if (org < b[a] || org > b[3 + a]) return false;
            continue;
        }
        float ta = (b[a] - org) * inv[a], tb = (b[3 + a] - org) * inv[a];
        if (ta > tb) { const float x = ta; ta = tb; tb = x; }
        lo = fmaxf(lo, ta - fabsf(ta) * 0x1p-16f);
        hi = fminf(hi, tb + fabsf(tb) * 0x1p-16f);
    }
    if (!(lo <= hi)) return false;
    if (LRT_BOX_CLIP) {
        t0 = lo;
        t1 = hi;
    }
    return true;
}

#ifndef LRT_WHILE_WHILE
#define LRT_WHILE_WHILE 1
#endif


// One kd-tree, closest hit.  `closest` enters as the object's t_max and
// leaves as the nearest accepted t.  ANY: stop at the first accepted hit.
//
// The loop descends through interior nodes until the lane reaches a leaf and
// only then tests the leaf ("while-while"): lanes of a wave that reach their
// leaves at different depths wait in the cheap descent instead of running the
// leaf code once per node step.  The visit order is the reference's: the
// descent neither changes mint nor closest, so the reference's per-node test
// closest < mint can only change after a pop.
// USE_LDS false: the whole stack in private memory (lds unused), for callers
// without an LDS stack (the emitter pdf's ray cast of a mesh light).
template <bool ANY, bool COUNT, bool USE_LDS = true>
LRT_D bool kd_traverse(const DevScene& S, const DevObject& ob, const Ray& r, float inv[3], float tmin,
                       float& closest, int32_t& prim, float& hu, float& hv, uint2* lds, Counters& cnt) {
    const float maxt0 = closest;
    constexpr int kLds = USE_LDS ? kLdsStack : 0;
    uint2 spill[kMaxStack - kLds];
    int top = 0;
    float mint = tmin, maxt = closest;
    uint32_t node = ob.node_base;
    const uint32_t stride = blockDim.x;
    bool found = false;
    for (;;) {
        if (closest < mint) break;
        uint2 nd = S.nodes[node];
        if (COUNT) cnt.nodes++;
        uint32_t tag = nd.y & 3u;
        while (tag != 3u) {
            const int axis = (int)tag;
            const float split = __uint_as_float(nd.x);
            const float oa = comp(r.o, axis), da = comp(r.d, axis);
            const float ia = axis == 0 ? inv[0] : (axis == 1 ? inv[1] : inv[2]);
            const float t_plane = (split - oa) * ia;
            const bool left_first = oa < split || (oa == split && da <= 0.0f);
            const uint32_t left = node + 1, right = ob.node_base + (nd.y >> 2);
            const uint32_t first = left_first ? left : right;
            const uint32_t second = left_first ? right : left;
            if (t_plane > maxt || t_plane <= 0.0f) {
                node = first;
            } else if (t_plane < mint) {
                node = second;
            } else {
                const uint2 e = make_uint2(second, __float_as_uint(t_plane));
                if constexpr (USE_LDS) {
                    if (top < kLds) lds[top * stride] = e;
                    else spill[top - kLds] = e;
                } else {
                    spill[top] = e;
                }
                ++top;
                node = first;
                maxt = t_plane;
            }
#if LRT_WHILE_WHILE
            nd = S.nodes[node];
            if (COUNT) cnt.nodes++;
            tag = nd.y & 3u;
#else
            break;
#endif
        }
#if !LRT_WHILE_WHILE
        if (tag != 3u) continue;
#endif
        const uint32_t count = nd.y >> 2;
        const uint32_t first = nd.x;
        float limit = fminf(maxt, closest);
        for (uint32_t k = 0; k < count; k += kLeafBatch) {
            // issue the reference and vertex loads of a batch together,
            // then test in reference order (ties go to the later triangle)
            TriRec rec[kLeafBatch];
            // every load issued, past the leaf's end clamped to its last
            // triangle (a record of the same leaf, never tested), so no branch
            // separates them and the batch's loads are in flight together
#pragma unroll
            for (int j = 0; j < kLeafBatch; ++j)
                rec[j] = load_tri(S.leaf_tris + 9 * (size_t)(first + min(k + (uint32_t)j, count - 1u)));
#pragma unroll
            for (int j = 0; j < kLeafBatch; ++j) {
                if (k + j < count) {
                    if (COUNT) { cnt.refs++; cnt.tris++; }
                    float t, u, v;
                    if (tri_test(rec[j], r, mint, limit, t, u, v)) {
                        closest = t;
                        prim = (int32_t)(S.refs[first + k + j] - ob.tri_base);
                        hu = u;
                        hv = v;
                        found = true;
                        if (ANY) return true;
                    }
                    limit = fminf(maxt, closest);
                }
            }
        }
        if (top == 0) break;
        --top;
        uint2 e;
        if constexpr (USE_LDS) e = top < kLds ? lds[top * stride] : spill[top - kLds];
        else e = spill[top];
        node = e.x;
        mint = __uint_as_float(e.y);
        if (top == 0) maxt = maxt0;
        else {
            uint2 below;
            if constexpr (USE_LDS) below = (top - 1) < kLds ? lds[(top - 1) * stride] : spill[top - 1 - kLds];
            else below = spill[top - 1];
            maxt = __uint_as_float(below.y);
        }
    }
    return found;
}

// The normal triangle_mesh::hit records for a hit on triangle prim of mesh ob
// at barycentrics (u, v) by a ray of direction d (triangle_mesh.h:81-96):
// front face from the geometric normal; the normal is the interpolated corner
// normal, normalised and turned toward the ray's origin, or without corner
// normals the face-forwarded geometric normal.
LRT_D void mesh_hit_normal(const DevScene& S, const DevObject& ob, int prim, float u, float v, f3 d, f3& n, bool& front) {
    const uint32_t g = ob.tri_base + (uint32_t)prim;
    const float* tri = S.tris + 9 * (size_t)g;
    const f3 gn = cross(ld3(tri + 3), ld3(tri + 6));
    const f3 outward = gn / len(gn);
    front = dot(d, outward) < 0;
    n = front ? outward : -outward;
    const float* cn = S.cnormals + 9 * (size_t)g;
    const float b0 = 1.0f - u - v;
    f3 sh = b0 * ld3(cn) + u * ld3(cn + 3) + v * ld3(cn + 6);
    if (len2(sh) > 0.0f) {
        sh = sh / len(sh);
        if (dot(d, sh) > 0.0f) sh = -sh;
        n = sh;
    }
}

// hittable_list::hit over the object list (hittable_list.h:19-34); a later
// object wins an equal t.  `lds` points at this lane's first stack slot.
// FULL: traverse each mesh over the reference's full window instead of the
// window clipped to the mesh's box (the counting pass counts both).
template <bool COUNT, bool FULL = COUNT>
LRT_D bool trace_closest(const DevScene& S, const Ray& r, float tmin, float tmax, HitInfo& h, uint2* lds,
                         Counters& cnt) {
    float inv[3] = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    float closest = tmax;
    bool any = false;
    h.obj = -1;
    h.prim = -1;
    for (int o = 0; o < S.n_objects; ++o) {
        const DevObject& ob = S.obj[o];
        if (ob.shape == kShapeMesh) {
            float t0 = tmin, c = closest, u = 0.0f, v = 0.0f;
            if (!FULL && !mesh_box_clip(S, o, r, inv, t0, c)) continue;
            int32_t p = -1;
            if (kd_traverse<false, COUNT>(S, ob, r, inv, t0, c, p, u, v, lds, cnt)) {
                closest = c; h.obj = o; h.prim = p; h.u = u; h.v = v; any = true;
            }
        } else if (ob.shape == kShapeRect) {
            float t, a, b;
            if (rect_test(ob, r, tmin, closest, t, a, b)) {
                closest = t; h.obj = o; h.prim = 0; h.u = a; h.v = b; any = true;
            }
        } else {
            float t;
            if (sphere_test(ob, r, tmin, closest, t)) {
                closest = t; h.obj = o; h.prim = 0; h.u = 0.0f; h.v = 0.0f; any = true;
            }
        }
    }
    h.t = closest;
    return any;
}

// Shadow test of path_integrator.h:113-116: blocked iff a surface that does
// not belong to emitter `ignore` is hit in [tmin, tmax].  Exact when the
// ignored emitter is convex (rectangle, sphere): its own surface cannot be
// nearer than the sampled point on it; a mesh emitter takes the reference's
// closest-hit form.
// COUNT: also count the traversal work into cnt (box-clipped windows, the
// first hit ends the query: the work the persistent kernels do).
template <bool COUNT>
LRT_D bool trace_occluded(const DevScene& S, const Ray& r, float tmin, float tmax, int ignore, uint2* lds,
                          Counters& cnt) {
    float inv[3] = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    if (emitter_nonconvex(S, ignore)) {  // the reference's closest-hit form
        HitInfo h;
        return trace_closest<COUNT, false>(S, r, tmin, tmax, h, lds, cnt) && S.obj[h.obj].emitter != ignore;
    }
    for (int o = 0; o < S.n_objects; ++o) {
        const DevObject& ob = S.obj[o];
        if (ob.emitter >= 0 && ob.emitter == ignore) continue;
        if (ob.shape == kShapeMesh) {
            float t0 = tmin, c = tmax, u, v;
            if (!mesh_box_clip(S, o, r, inv, t0, c)) continue;
            int32_t p;
            if (kd_traverse<true, COUNT>(S, ob, r, inv, t0, c, p, u, v, lds, cnt)) return true;
        } else if (ob.shape == kShapeRect) {
            float t, a, b;
            if (rect_test(ob, r, tmin, tmax, t, a, b)) return true;
        } else {
            float t;
            if (sphere_test(ob, r, tmin, tmax, t)) return true;
        }
    }
    return false;
}

LRT_D bool trace_occluded(const DevScene& S, const Ray& r, float tmin, float tmax, int ignore, uint2* lds) {
    Counters cnt{0, 0, 0};
    return trace_occluded<false>(S, r, tmin, tmax, ignore, lds, cnt);
}

// ---- resumable traversal for persistent kernels ----------------------------
// A lane's ray is traced in rounds (one descent to a leaf, the leaf's tests
// and a pop per round) so that a kernel can hand the lane a new ray as soon
// as its ray is finished instead of waiting for the slowest lane of the wave.
// The visit order, windows and acceptance rules are those of kd_traverse and
// trace_closest / trace_occluded.

// which of the path tracer's traversal launches take their rays per XCD
// (wave_refill_xcd): the camera rays (k_extend, a pixel-ordered queue whose
// in-flight rays span a band of the image) and the bounce queues
// (k_trace_queues, whose in-flight rays already start in a few shading cells)
#ifndef LRT_XCD_FETCH_EXTEND
#define LRT_XCD_FETCH_EXTEND 1
#endif
#ifndef LRT_XCD_FETCH_QUEUES
#define LRT_XCD_FETCH_QUEUES 0
#endif
// k_shade appends to its two queues once per block (continuing paths grouped
// by direction octant) instead of once per wave
#ifndef LRT_SHADE_BLOCK_APPEND
#define LRT_SHADE_BLOCK_APPEND 1
#endif
// k_shade groups the block's shadow records by direction octant as well
// (no traversal gain on the 1080p frame, shading +0.6 ms: off)
#ifndef LRT_SHADOW_OCTANTS
#define LRT_SHADOW_OCTANTS 0
#endif
// the photon eye pass's shading kernel appends once per block as well
#ifndef LRT_EYE_BLOCK_APPEND
#define LRT_EYE_BLOCK_APPEND 1
#endif
// k_shade's block size (threads)
#ifndef LRT_SHADE_BLOCK
#define LRT_SHADE_BLOCK 256
#endif
constexpr uint32_t kShadeBlock = LRT_SHADE_BLOCK;
// the camera rays queued in 8 x 8 pixel tiles rather than in rows
#ifndef LRT_CAMERA_TILES
#define LRT_CAMERA_TILES 1
#endif
#ifndef LRT_REFILL
#define LRT_REFILL 48
#endif
constexpr int kRefill = LRT_REFILL;  // idle lanes that make a wave take new rays

// A lane's state between rounds.  The small integers share one register
// (ids: object being traversed, hit object + 1, ignored emitter + 1, stack
// top) and the window's upper end is not kept (it is the mint of the entry
// below the top, or maxt0), so the persistent kernels spill fewer registers.
struct LaneTrace {
    Ray r;
    float tmin, closest;   // closest: nearest accepted t (ANY: the ray's tmax)
    int32_t prim;          // hit so far (closest hit)
    float u, v;
    uint32_t ids;          // bits 0-4 object o, 5-9 hit object + 1, 10-13 ignored emitter + 1, 16-21 top
    uint32_t node;         // kd state of object o
    float mint, maxt0;
    bool in_mesh, hit;
    bool any;              // any hit ends the ray (shadow ray), also when ANY is false
    LRT_D int o() const { return (int)(ids & 31u); }
    LRT_D int obj() const { return (int)((ids >> 5) & 31u) - 1; }
    LRT_D int ignore() const { return (int)((ids >> 10) & 15u) - 1; }
    LRT_D int top() const { return (int)((ids >> 16) & 63u); }
    LRT_D void set_o(int v) { ids = (ids & ~31u) | (uint32_t)v; }
    LRT_D void set_obj(int v) { ids = (ids & ~(31u << 5)) | ((uint32_t)(v + 1) << 5); }
    LRT_D void set_top(int v) { ids = (ids & ~(63u << 16)) | ((uint32_t)v << 16); }
};
static_assert(kMaxObjects < 31 && kMaxStack < 64, "LaneTrace::ids fields");

LRT_D void lane_trace_start(LaneTrace& T, const Ray& r, float tmin, float tmax, int ignore) {
    T.r = r;
    T.tmin = tmin;
    T.closest = tmax;
    T.prim = -1;
    T.u = 0.0f;
    T.v = 0.0f;
    // an ignored emitter outside 0..14 matches no emitter (a scene has at most 4 lights)
    const int ig = ignore >= 0 && ignore < 15 ? ignore : -1;
    T.ids = (uint32_t)(ig + 1) << 10;  // o 0, no hit object, top 0
    T.in_mesh = false;
    T.hit = false;
    T.any = false;
}

// Shadow rays (path_integrator.h:104-116): an any-hit that skips the sampled
// emitter's surfaces, or for a non-convex emitter the reference's closest-hit
// form; shadow_blocked decides when the ray is finished.
LRT_D void lane_shadow_start(const DevScene& S, LaneTrace& T, const Ray& r, float tmin, float tmax, int ignore) {
    lane_trace_start(T, r, tmin, tmax, ignore);
    T.any = !emitter_nonconvex(S, ignore);
}
LRT_D bool shadow_blocked(const DevScene& S, const LaneTrace& T, int ignore) {
    return T.hit && (T.any || S.obj[T.obj()].emitter != ignore);
}

// One round; returns true when the ray is finished (T.hit / T.obj(), T.prim,
// T.closest, T.u, T.v hold the result).
// one interior node of kd_traverse: the child to visit next; the far child
// is pushed with its entry distance when the split plane lies in [mint, maxt]
LRT_D uint32_t kd_step(const Ray& r, uint2 nd, uint32_t node, uint32_t node_base, float mint,
                       float& maxt, int& top, uint2* lds, uint2* spill) {
    const int axis = (int)(nd.y & 3u);
    const float split = __uint_as_float(nd.x);
    const float oa = comp(r.o, axis), da = comp(r.d, axis);
    const float ia = 1.0f / da;  // recomputed per step: the same value, three fewer live registers
    const float t_plane = (split - oa) * ia;
    const bool left_first = oa < split || (oa == split && da <= 0.0f);
    const uint32_t left = node + 1, right = node_base + (nd.y >> 2);
    const uint32_t first = left_first ? left : right;
    const uint32_t second = left_first ? right : left;
    if (t_plane > maxt || t_plane <= 0.0f) return first;
    if (t_plane < mint) return second;
    const uint2 e = make_uint2(second, __float_as_uint(t_plane));
    if (top < kLdsStack) lds[top * (int)blockDim.x] = e;
    else spill[top - kLdsStack] = e;
    ++top;
    maxt = t_plane;
    return first;
}

template <bool ANY>
LRT_D bool lane_trace_round(const DevScene& S, LaneTrace& T, uint2* lds, uint2* spill) {
    const uint32_t stride = blockDim.x;
    const bool any = ANY || T.any;
    if (!T.in_mesh) {
        // the objects in order: analytic shapes at once, a mesh starts a kd traversal
        for (int o = T.o(); o < S.n_objects; ++o, T.set_o(o)) {
            const DevObject& ob = S.obj[o];
            if (any && ob.emitter >= 0 && ob.emitter == T.ignore()) continue;
            if (ob.shape == kShapeMesh) {
                float t0 = T.tmin, t1 = T.closest;
                const float inv[3] = {1.0f / T.r.d.x, 1.0f / T.r.d.y, 1.0f / T.r.d.z};
                if (!mesh_box_clip(S, o, T.r, inv, t0, t1)) continue;
                T.node = ob.node_base;
                T.mint = t0;
                T.maxt0 = t1;
                T.set_top(0);
                T.in_mesh = true;
                break;
            } else if (ob.shape == kShapeRect) {
                float t, a, b;
                if (rect_test(ob, T.r, T.tmin, T.closest, t, a, b)) {
                    T.hit = true;
                    if (any) return true;
                    T.closest = t; T.set_obj(o); T.prim = 0; T.u = a; T.v = b;
                }
            } else {
                float t;
                if (sphere_test(ob, T.r, T.tmin, T.closest, t)) {
                    T.hit = true;
                    if (any) return true;
                    T.closest = t; T.set_obj(o); T.prim = 0; T.u = 0.0f; T.v = 0.0f;
                }
            }
        }
        if (!T.in_mesh) return true;
    }
    const int o = T.o();
    const DevObject& ob = S.obj[o];
    const Ray& r = T.r;
    bool done = T.closest < T.mint;
    if (!done) {
        uint32_t node = T.node;
        const float mint = T.mint;
        int top = T.top();
        float maxt = T.maxt0;  // the window's end: the mint of the entry below the top
        if (top > 0) maxt = __uint_as_float((top - 1 < kLdsStack ? lds[(top - 1) * stride] : spill[top - 1 - kLdsStack]).y);
        uint2 nd = S.nodes[node];
        while ((nd.y & 3u) != 3u) {
            node = kd_step(r, nd, node, ob.node_base, mint, maxt, top, lds, spill);
            nd = S.nodes[node];
        }
        const uint32_t count = nd.y >> 2;
        const uint32_t firstref = nd.x;
        float limit = fminf(maxt, T.closest);
        for (uint32_t k = 0; k < count; k += kLeafBatch) {
            TriRec rec[kLeafBatch];
            // as in kd_traverse: the batch's loads issued together, clamped to the leaf
#pragma unroll
            for (int j = 0; j < kLeafBatch; ++j)
                rec[j] = load_tri(S.leaf_tris + 9 * (size_t)(firstref + min(k + (uint32_t)j, count - 1u)));
#pragma unroll
            for (int j = 0; j < kLeafBatch; ++j) {
                if (k + j < count) {
                    float t, u, v;
                    if (tri_test(rec[j], r, mint, limit, t, u, v)) {
                        T.hit = true;
                        if (any) return true;
                        T.closest = t;
                        T.set_obj(o);
                        T.prim = (int32_t)(S.refs[firstref + k + j] - ob.tri_base);
                        T.u = u;
                        T.v = v;
                    }
                    limit = fminf(maxt, T.closest);
                }
            }
        }
        if (top == 0) {
            done = true;
        } else {
            --top;
            const uint2 e = top < kLdsStack ? lds[top * stride] : spill[top - kLdsStack];
            T.node = e.x;
            T.mint = __uint_as_float(e.y);
            T.set_top(top);
        }
    }
    if (done) {  // object o finished: the next round goes on with the following objects
        T.in_mesh = false;
        T.set_o(o + 1);
    }
    return false;
}

// Wave-level work distribution for persistent kernels: when at least
// kRefill lanes are idle (or all are), the wave takes that many consecutive
// items from *next with one atomic.  Returns the lane's new item or
// UINT32_MAX; `more` turns 0 once the items are exhausted (uniform over the
// wave).  STATIC_FIRST (the path tracer's and eye pass's queues, often short
// or empty): the first call (more == 2, every lane idle) hands lane g of the
// grid item g without an atomic and later items are the grid's lane count +
// the counter's value, so a launch over an empty or short queue costs no
// atomic per wave (20k waves on one address took ~0.2 ms per launch).  The
// batch ray queries (the host knows their ray count) keep the plain form,
// which needs fewer registers.
template <bool STATIC_FIRST = true>
LRT_D uint32_t wave_refill(bool idle_lane, uint32_t* next, uint32_t n, int& more) {
    const uint32_t none = 0xffffffffu;
    if (!more) return none;
    const uint32_t lanes = STATIC_FIRST ? gridDim.x * blockDim.x : 0u;
    if (STATIC_FIRST && more == 2) {
        more = lanes < n ? 1 : 0;
        const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
        return idle_lane && g < n ? g : none;
    }
    const uint64_t idle = __ballot(idle_lane);
    const int nidle = __popcll(idle);
    if (nidle < kRefill && nidle < (int)__popcll(__ballot(true))) return none;
    const int lane = (int)(threadIdx.x & 63u);
    const int leader = __ffsll((long long)idle) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(next, (uint32_t)nidle);
    base = lanes + (uint32_t)__shfl((int)base, leader);  // n < 2^31: no wrap before more turns 0
    if (base + (uint32_t)nidle >= n) more = 0;
    if (!idle_lane) return none;
    const uint32_t i = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
    return i < n ? i : none;
}

// XCD-aware work distribution for the path tracer's traversal launches.
// MI355X's 8 XCDs each have their own L2, and the dispatcher places
// workgroups on them round-robin (block b on XCD b % 8).  The items are two
// queues, A = [0, na) and B = [na, na + nb), each written in shading-bin
// (Morton cell) order, so a contiguous stretch of either holds rays that start
// in one region of the scene.  XCD x owns the x-th eighth of each queue: its
// L2 then serves the nodes and triangles of that region instead of the whole
// scene's.  A wave's first call hands its lanes the XCD's first items without
// an atomic (as STATIC_FIRST above); later items come from the XCD's counter,
// and once its range is drained the wave takes items from the other XCDs'
// ranges, so the launch still ends with its slowest ray, not its slowest XCD.
// ctr: kXcds counters kXcdStride words apart (one cache line each), zero at
// launch.  `more` (uniform over the wave): 2 before the first call, 0 once
// every range is drained, else 1 + 4 x the mask of the ranges found drained.
constexpr uint32_t kXcds = 8, kXcdStride = 32;

struct XcdSplit {
    uint32_t na, nb;
    LRT_D uint32_t a0(uint32_t x) const { return (uint32_t)(((uint64_t)na * x) >> 3); }
    LRT_D uint32_t b0(uint32_t x) const { return (uint32_t)(((uint64_t)nb * x) >> 3); }
    LRT_D uint32_t len(uint32_t x) const { return a0(x + 1) - a0(x) + b0(x + 1) - b0(x); }
    // item j of XCD x's range: its share of A, then its share of B
    LRT_D uint32_t item(uint32_t x, uint32_t j) const {
        const uint32_t la = a0(x + 1) - a0(x);
        return j < la ? a0(x) + j : na + b0(x) + (j - la);
    }
};

LRT_D uint32_t wave_refill_xcd(bool idle_lane, uint32_t* ctr, const XcdSplit& q, int& more) {
    const uint32_t none = 0xffffffffu;
    if (!more) return none;
    const uint32_t home = blockIdx.x & (kXcds - 1u);
    if (more == 2) {
        more = 1;
        const uint32_t j = (blockIdx.x / kXcds) * blockDim.x + threadIdx.x;
        return idle_lane && j < q.len(home) ? q.item(home, j) : none;
    }
    const uint64_t idle = __ballot(idle_lane);
    const int nidle = __popcll(idle);
    if (nidle < kRefill && nidle < (int)__popcll(__ballot(true))) return none;
    const int lane = (int)(threadIdx.x & 63u);
    const int leader = __ffsll((long long)idle) - 1;
    uint32_t tried = (uint32_t)more >> 2, from = kXcds, start = 0u;
    if (lane == leader) {
        for (uint32_t t = 0; t < kXcds; ++t) {
            const uint32_t x = (home + t) & (kXcds - 1u);
            if ((tried >> x) & 1u) continue;
            const uint32_t len = q.len(x);
            // lanes of XCD x that took its first items in their first call
            const uint32_t first = ((gridDim.x + kXcds - 1u - x) / kXcds) * blockDim.x;
            if (len > first) {
                // len < 2^31 and at most one overshoot of 64 per wave and range: no wrap
                const uint32_t b = first + atomicAdd(&ctr[x * kXcdStride], (uint32_t)nidle);
                if (b < len) {
                    from = x;
                    start = b;
                    if (b + (uint32_t)nidle >= len) tried |= 1u << x;  // drained by this fetch
                    break;
                }
            }
            tried |= 1u << x;
        }
    }
    tried = (uint32_t)__builtin_amdgcn_readlane((int)tried, leader);
    from = (uint32_t)__builtin_amdgcn_readlane((int)from, leader);
    start = (uint32_t)__builtin_amdgcn_readlane((int)start, leader);
    more = from == kXcds || tried == (1u << kXcds) - 1u ? 0 : (int)(1u + 4u * tried);
    if (from == kXcds || !idle_lane) return none;
    const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
    return start + rank < q.len(from) ? q.item(from, start + rank) : none;
}

}  // namespace lrt
