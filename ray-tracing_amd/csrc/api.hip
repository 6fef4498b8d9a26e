// liblinrt_hip.so: the C ABI (include/linrt.h) over the HIP kernels.
// Host work here: validating descriptions, kd-tree build and flattening,
// device uploads, launch sequencing and HIP-event timing.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "host/host.hpp"
#include "kernels.hip"
#include "eye.hip"
#include "diff_wave.hip"
#include "device/nn.h"

// Scratch of lrt_trace_closest_noalloc: the persistent traversal's ray-fetch
// counter lives here instead of in memory the library allocates, so that
// callers which cannot allocate (code under graph capture, serving loops) can
// query rays.  256 bytes per device, held until the process ends; every call
// on a device shares its copy (calls that overlap on one device race).  Host
// code never passes `scratch` itself to a kernel (in host code the name is a
// stand-in, not a device address): lrt_scene_create resolves the device
// address once per device with hipGetSymbolAddress, and the entry point uses
// that address.
constexpr uint32_t kNoAllocScratchWords = 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];
// rays per call of lrt_trace_closest_noalloc: the fetch counter runs up to
// n plus 64 per wave of the grid, which must not wrap 32 bits
constexpr uint64_t kNoAllocMaxRays = 1ull << 31;

// paths per render pass (2^27: ~21 GiB of path records, hit records, shadow
// and path queues, one pass for a 64-spp 1080p frame); the environment
// variable LINRT_PASS_PATHS overrides it
#ifndef LRT_PASS_PATHS
#define LRT_PASS_PATHS (1u << 27)
#endif
// camera samples per eye-pass chunk (2^22: the work items of the first
// generations take up to ~25 GiB at 32 BSDF samples per first hit; bigger
// chunks mean fewer launches and host round trips per frame);
// LINRT_EYE_CHUNK overrides
#ifndef LRT_EYE_CHUNK
#define LRT_EYE_CHUNK (1u << 22)
#endif
static uint32_t eye_chunk() {
    static const uint32_t v = [] {
        const char* e = std::getenv("LINRT_EYE_CHUNK");
        const long long x = e ? std::atoll(e) : 0;
        return x >= 1024 && x <= (1ll << 24) ? (uint32_t)x : (uint32_t)LRT_EYE_CHUNK;
    }();
    return v;
}

static uint32_t pass_paths() {
    static const uint32_t v = [] {
        const char* e = std::getenv("LINRT_PASS_PATHS");
        const long long x = e ? std::atoll(e) : 0;
        return x >= 1024 && x < (1ll << 31) ? (uint32_t)x : (uint32_t)LRT_PASS_PATHS;
    }();
    return v;
}

// k_shade's grid, in blocks per CU (at most; 2 are resident): with more
// blocks each walks fewer strided items and the queues it writes follow the
// shading order more closely (64-spp frame: 4 -> 430, 8 -> 432, 32 -> 437,
// 128 -> 442 Msamples/s; one item per thread 426, the blocks' dispatch and
// barriers then dominate)
#ifndef LRT_SHADE_BLOCKS_PER_CU
#define LRT_SHADE_BLOCKS_PER_CU 128
#endif
// the photon eye pass's k_eye_shade (its generations' queues gain nothing
// from a larger grid: 128 blocks per CU made the eye pass 0.7% slower)
#ifndef LRT_EYE_SHADE_BLOCKS_PER_CU
#define LRT_EYE_SHADE_BLOCKS_PER_CU 16
#endif
// live paths at or below which a pass's remaining bounces run in k_tail
// (one lane finishes a path) instead of one wavefront launch per bounce, from
// the Russian-roulette depth on.  A large pass is served better by the
// wavefront down to few paths (its late launches still hold 10^5-10^6 paths),
// a small one by an early tail (each wavefront launch lasts at least as long
// as its slowest ray): LRT_TAIL_SMALL for passes of at most LRT_SMALL_PASS
// paths, else LRT_TAIL_LARGE (1080p, 64 spp = 2^27 paths: 2^17 -> 442.8,
// 2^18 -> 440.4, 2^19 -> 433.4 Msamples/s; 1 spp = 2^21 paths: 2^17 -> 17.3,
// 2^18 -> 15.0, 2^19 -> 14.6 ms per denoised frame, 2 and 4 spp likewise
// faster at 2^19).  LINRT_TAIL_PATHS sets a fixed threshold (0: never).
#ifndef LRT_TAIL_SMALL
#define LRT_TAIL_SMALL (1u << 19)
#endif
#ifndef LRT_TAIL_LARGE
#define LRT_TAIL_LARGE (1u << 17)
#endif
#ifndef LRT_SMALL_PASS
#define LRT_SMALL_PASS (1u << 23)
#endif
static uint32_t tail_paths(uint32_t pass_paths) {
    static const long long fixed = [] {
        const char* e = std::getenv("LINRT_TAIL_PATHS");
        if (!e) return -1ll;
        const long long x = std::atoll(e);
        return x >= 0 && x < (1ll << 31) ? x : -1ll;
    }();
    if (fixed >= 0) return (uint32_t)fixed;
    return pass_paths <= LRT_SMALL_PASS ? LRT_TAIL_SMALL : LRT_TAIL_LARGE;
}

namespace lrt {

struct Generated;
Generated* generate(const lrt_gen_params& p, const float* tables, uint32_t n_tables, const lrt_light_desc* sky);
void free_generated(Generated* g);
Generated* generate_obj_scene(const char* path, const lrt_material* materials, uint32_t n_materials,
                              const float* tables, uint32_t n_tables, float aspect);
const lrt_scene_desc* generated_desc(const Generated* g);

namespace {
thread_local std::string g_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                     \
    do {                                                                                                \
        const hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw Error(LRT_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The library's streams are non-blocking: they do not wait for the null
// stream, where a caller (torch's default stream, for one) may still be
// writing the inputs it passes by device pointer, or reading the outputs of
// an earlier call.  A call on device pointers therefore first orders its
// stream after the work queued on the null stream so far (an event recorded
// there and waited for on the device; no host synchronisation).
void after_caller(hipStream_t s, hipEvent_t ev) {
    HIPCHECK(hipEventRecord(ev, nullptr));
    HIPCHECK(hipStreamWaitEvent(s, ev, 0));
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return LRT_OK;
    } catch (const Error& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return LRT_ERR_ALLOC;
    } catch (const std::exception& e) {
        g_error = e.what();
        return LRT_ERR_ARG;
    } catch (...) {
        g_error = "unknown error";
        return LRT_ERR_INTERNAL;
    }
}
void require(bool c, const std::string& msg, int code = LRT_ERR_ARG) {
    if (!c) throw Error(code, msg);
}

// device buffer owned by RAII
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        if (b <= bytes && p) return;
        release();
        if (b == 0) return;
        const hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) throw Error(LRT_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        bytes = b;
    }
    template <typename T>
    T* as() const { return (T*)p; }
};

void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
    b.alloc(bytes);
    if (bytes) HIPCHECK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
}

int cu_count(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
}

void check_materials(const lrt_material* m, uint32_t nm, uint32_t n_tables) {
    for (uint32_t i = 0; i < nm; ++i) {
        const lrt_material& M = m[i];
        require(M.type >= LRT_BSDF_NONE && M.type <= LRT_BSDF_ROUGH_COATING, "material " + std::to_string(i) + ": unknown type");
        if (M.type == LRT_BSDF_COATING || M.type == LRT_BSDF_ROUGH_COATING) {
            require(M.nested >= 0 && (uint32_t)M.nested < nm, "material " + std::to_string(i) + ": nested index out of range");
            const int nt = m[M.nested].type;
            require(nt != LRT_BSDF_COATING && nt != LRT_BSDF_ROUGH_COATING && nt != LRT_BSDF_NONE,
                    "material " + std::to_string(i) + ": a coating must nest a base BSDF", LRT_ERR_UNSUPPORTED);
        }
        if (M.type == LRT_BSDF_ROUGH_PLASTIC)
            require(M.table_ext >= 0 && (uint32_t)M.table_ext < n_tables && M.table_int >= 0 && (uint32_t)M.table_int < n_tables,
                    "material " + std::to_string(i) + ": rough plastic needs table_ext and table_int");
        if (M.type == LRT_BSDF_ROUGH_COATING)
            require(M.table_ext >= 0 && (uint32_t)M.table_ext < n_tables,
                    "material " + std::to_string(i) + ": rough coating needs table_ext");
    }
}

// bsdf::flags of a material on the host (the device's bsdf_flags)
uint32_t host_bsdf_flags(const std::vector<lrt_material>& mats, int m) {
    const lrt_material& M = mats[(size_t)m];
    const uint32_t S = LRT_FLAG_SPECULAR, G = LRT_FLAG_GLOSSY, D = LRT_FLAG_DIFFUSE, R = LRT_FLAG_REFLECTIVE,
                   T = LRT_FLAG_REFRACTIVE, DI = LRT_FLAG_DIELECTRIC, ME = LRT_FLAG_METALLIC;
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: return D | R;
        case LRT_BSDF_CONDUCTOR: return (M.alpha <= 0.0f ? S : G) | R | ME;
        case LRT_BSDF_DIELECTRIC: return S | R | T | DI;
        case LRT_BSDF_ROUGH_DIELECTRIC: return G | R | T | DI;
        case LRT_BSDF_PLASTIC: return S | R | DI | D;
        case LRT_BSDF_ROUGH_PLASTIC: return G | R | DI | D;
        case LRT_BSDF_COATING:
        case LRT_BSDF_ROUGH_COATING:
            if (M.nested < 0 || (size_t)M.nested >= mats.size() || M.nested == m) return 0u;
            return host_bsdf_flags(mats, M.nested) | (M.type == LRT_BSDF_COATING ? S : G) | R | DI;
        default: return 0u;
    }
}

// the largest number of specular components of any material
int host_max_specular_components(const std::vector<lrt_material>& mats) {
    int best = 0;
    for (size_t m = 0; m < mats.size(); ++m) {
        const lrt_material& M = mats[m];
        auto base = [](const lrt_material& B) {
            if (B.type == LRT_BSDF_DIELECTRIC) return 2;
            if (B.type == LRT_BSDF_PLASTIC) return 1;
            if (B.type == LRT_BSDF_CONDUCTOR && B.alpha <= 0.0f) return 1;
            return 0;
        };
        int c = base(M);
        if ((M.type == LRT_BSDF_COATING || M.type == LRT_BSDF_ROUGH_COATING) && M.nested >= 0 &&
            (size_t)M.nested < mats.size())
            c = base(mats[(size_t)M.nested]) + (M.type == LRT_BSDF_COATING ? 1 : 0);
        best = std::max(best, c);
    }
    return best;
}

DevLight make_light(const lrt_light_desc& d, const std::vector<lrt_object_desc>& objs) {
    DevLight L;
    std::memset(&L, 0, sizeof(L));
    L.kind = d.kind;
    L.object = d.object;
    for (int c = 0; c < 3; ++c) L.radiance[c] = d.radiance[c];
    if (d.kind == LRT_LIGHT_SKY) {
        // env_emitter.h:13-30
        f3 s = ld3(d.sun_dir);
        const float l = std::sqrt(s.x * s.x + s.y * s.y + s.z * s.z);
        s = f3{s.x / l, s.y / l, s.z / l};
        L.sun[0] = s.x; L.sun[1] = s.y; L.sun[2] = s.z;
        const float l2 = std::sqrt(s.x * s.x + s.y * s.y + s.z * s.z);
        const f3 n{s.x / l2, s.y / l2, s.z / l2};
        const f3 a = (std::fabs(n.x) > 0.9f) ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
        f3 t = cross(n, a);
        const float lt = std::sqrt(t.x * t.x + t.y * t.y + t.z * t.z);
        t = f3{t.x / lt, t.y / lt, t.z / lt};
        const f3 sx = cross(n, t);
        L.fs[0] = sx.x; L.fs[1] = sx.y; L.fs[2] = sx.z;
        L.ft[0] = t.x; L.ft[1] = t.y; L.ft[2] = t.z;
        L.fn[0] = n.x; L.fn[1] = n.y; L.fn[2] = n.z;
        L.sun_active = s.y > 0.0f;
        const double radius = 0.255 * 3.14159265358979323846 / 180.0;
        const double cosine = std::cos(radius);
        L.cos_cutoff = (float)cosine;
        L.cone_pdf = (float)(1.0 / (2.0 * 3.14159265358979323846 * (1.0 - cosine)));
        L.scale = d.scale;
        std::memcpy(L.cfg, d.config, sizeof(L.cfg));
        std::memcpy(L.rad, d.channel_radiance, sizeof(L.rad));
    } else {
        require(d.object >= 0 && (size_t)d.object < objs.size(), "area light object out of range");
    }
    return L;
}

// HIP-event timing of one kernel kind
struct Timer {
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    double total_ms = 0.0;
    uint64_t launches = 0, items = 0;
    ~Timer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    std::pair<hipEvent_t, hipEvent_t> next() {
        while (pool.size() < used + 2) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        used += 2;
        return {pool[used - 2], pool[used - 1]};
    }
    void collect() {  // after a stream sync
        for (size_t i = 0; i + 1 < used; i += 2) {
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, pool[i], pool[i + 1]));
            total_ms += ms;
            launches++;
        }
        used = 0;
    }
};

}  // namespace
}  // namespace lrt

using namespace lrt;

struct PhotonState {
    PhotonBuf buf;
    DevBuf raw, per, sort, hist, perm, xyz, nodes, order, pos, pay, tree_work;
    PhotonMapDev map;
    uint64_t n = 0, shot = 0, n_nodes = 0;
    bool valid = false;
};

struct lrt_scene {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    DevScene S;
    DevLights Ls;
    Mats M;
    CameraDev cam;
    DevBuf nodes, refs, tris, leaf_tris, cnormals, trimat, mats, tables, obj_box;
    DevBuf light_cdf, light_tris;  // mesh lights: running triangle areas, triangle vertices
    bool mesh_lights = false;      // some area light is a mesh: shading kernels with its pdf ray cast
    lrt_scene_info info;
    std::vector<lrt_object_desc> objects;  // shapes only (pointers not kept)
    // path-tracer work buffers
    size_t cap = 0;
    DevBuf pstate, hits, shadow, queues, counters, stats, accum, accum2, totals;
    DevBuf aov_path, aov_dev;  // lrt_render_features: per-path first-hit features, a device copy of the output
    // ray-query staging
    DevBuf q_rays, q_out0, q_out1, q_out2, q_aux;
    // photon mapping
    PhotonState pm[2];   // global, caustic
    DevBuf eye_items[2], eye_hits, eye_shadow, eye_query, eye_acc, eye_sort, eye_hist;
    DevBuf sobol;  // the reference sampler's direction numbers (uploaded on first use)
    size_t eye_cap[2] = {0, 0}, eye_hit_cap = 0, eye_shadow_cap = 0, eye_query_cap = 0;
    EmitCdf emit;
    float bounding_radius = 0.0f;
    float box_lo[3] = {0.0f, 0.0f, 0.0f}, box_hi[3] = {0.0f, 0.0f, 0.0f};  // scene bounds
    bool has_delta = false;
    DevBuf diff_mats, diff_tex, diff_theta, diff_img, diff_dimg, diff_grad, diff_fix, diff_part, diff_paths, diff_shadow;
    std::vector<lrt_material> host_mats;
    int max_spec_components = 0;
    Timer timers[LRT_KERNEL_COUNT_];
    hipEvent_t caller = nullptr;  // after_caller's event
    ~lrt_scene() {
        if (caller) (void)hipEventDestroy(caller);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void build_scene(lrt_scene& sc, const lrt_scene_desc& d) {
    require(d.n_objects <= (uint32_t)kMaxObjects, "too many objects (max " + std::to_string(kMaxObjects) + ")", LRT_ERR_UNSUPPORTED);
    require(d.n_lights <= (uint32_t)kMaxLights, "too many lights (max " + std::to_string(kMaxLights) + ")", LRT_ERR_UNSUPPORTED);
    require(d.n_objects == 0 || d.objects, "objects is null");
    require(d.n_materials == 0 || d.materials, "materials is null");
    require(d.n_tables == 0 || d.tables, "tables is null");
    check_materials(d.materials, d.n_materials, d.n_tables);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> words, refs;
    std::vector<float> tris, cn;
    std::vector<uint16_t> tmat;
    // mesh emitters: running triangle areas and vertices (triangle_mesh.h:16-25)
    std::vector<float> lcdf, ltris;
    std::vector<uint32_t> obj_cdf_base(kMaxObjects, 0), obj_cdf_n(kMaxObjects, 0);
    bool any_tmat = false;
    int max_depth = 0;
    // per object: min xyz, max xyz of a mesh's triangles (empty box otherwise)
    std::vector<float> obj_box(6 * (size_t)kMaxObjects);
    for (size_t i = 0; i < obj_box.size(); ++i) obj_box[i] = i % 6 < 3 ? INFINITY : -INFINITY;
    std::memset(&sc.S, 0, sizeof(sc.S));
    sc.objects.assign(d.objects, d.objects + d.n_objects);
    for (uint32_t o = 0; o < d.n_objects; ++o) {
        const lrt_object_desc& od = d.objects[o];
        DevObject& ob = sc.S.obj[o];
        std::memset(&ob, 0, sizeof(ob));
        ob.shape = od.shape;
        ob.material = od.shape == LRT_SHAPE_RECTANGLE ? -1 : od.material;
        ob.emitter = od.emitter;
        require(ob.material < (int)d.n_materials, "object " + std::to_string(o) + ": material out of range");
        require(od.emitter < (int)d.n_lights, "object " + std::to_string(o) + ": emitter out of range");
        sc.objects[o].positions = nullptr;
        sc.objects[o].indices = nullptr;
        sc.objects[o].corner_normals = nullptr;
        sc.objects[o].triangle_material = nullptr;
        if (od.shape == LRT_SHAPE_MESH) {
            require(od.n_triangles == 0 || (od.positions && od.indices), "object " + std::to_string(o) + ": mesh arrays are null");
            const uint32_t nt = od.n_triangles;
            ob.tri_base = (uint32_t)(tris.size() / 9);
            ob.n_tris = nt;
            ob.per_tri_material = od.triangle_material ? 1 : 0;
            std::vector<float> boxes(6 * (size_t)nt);
            const bool emissive = od.emitter >= 0;
            if (emissive) obj_cdf_base[o] = (uint32_t)lcdf.size();
            float area = 0.0f;
            for (uint32_t t = 0; t < nt; ++t) {
                const uint32_t i0 = od.indices[3 * t], i1 = od.indices[3 * t + 1], i2 = od.indices[3 * t + 2];
                require(i0 < od.n_vertices && i1 < od.n_vertices && i2 < od.n_vertices,
                        "object " + std::to_string(o) + ": vertex index out of range");
                const f3 p0 = ld3(od.positions + 3 * (size_t)i0), p1 = ld3(od.positions + 3 * (size_t)i1),
                         p2 = ld3(od.positions + 3 * (size_t)i2);
                f3 e1 = p1 - p0, e2 = p2 - p0;
                {  // triangle_mesh.h:19-25: area += 0.5 * |cross(p1 - p0, p2 - p0)|
                    const f3 g = cross(e1, e2);
                    area += 0.5f * std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
                    if (emissive) {
                        lcdf.push_back(area);
                        const float v9[9] = {p0.x, p0.y, p0.z, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z};
                        ltris.insert(ltris.end(), v9, v9 + 9);
                    }
                }
                if (len2(cross(e1, e2)) <= 0.0f) e1 = e2 = splat(0.0f);  // triangle_mesh.h:52-54
                const float rec[9] = {p0.x, p0.y, p0.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z};
                tris.insert(tris.end(), rec, rec + 9);
                float* bx = obj_box.data() + 6 * (size_t)o;
                for (const f3& q : {p0, p1, p2})
                    for (int a = 0; a < 3; ++a) {
                        const float v = comp(q, a);
                        bx[a] = std::min(bx[a], v);
                        bx[3 + a] = std::max(bx[3 + a], v);
                    }
                for (int k = 0; k < 9; ++k) cn.push_back(od.corner_normals ? od.corner_normals[9 * (size_t)t + k] : 0.0f);
                float* b = &boxes[6 * (size_t)t];
                for (int a = 0; a < 3; ++a) {
                    const float v0 = comp(p0, a), v1 = comp(p1, a), v2 = comp(p2, a);
                    b[a] = std::min(std::min(v0, v1), v2);
                    b[3 + a] = std::max(std::max(v0, v1), v2);
                }
                if (od.triangle_material) {
                    require(od.triangle_material[t] < d.n_materials, "object " + std::to_string(o) + ": triangle material out of range");
                    tmat.push_back(od.triangle_material[t]);
                    any_tmat = true;
                } else {
                    tmat.push_back((uint16_t)std::max(0, ob.material));
                }
            }
            ob.area = area;
            if (emissive) obj_cdf_n[o] = nt;
            KdTree kt;
            kd_build(boxes.data(), nt, kt);
            FlatTree ft;
            kd_flatten(kt, ft);
            require(ft.max_depth <= kMaxStack, "kd-tree deeper than the traversal stack");
            max_depth = std::max(max_depth, ft.max_depth);
            ob.node_base = (uint32_t)(words.size() / 2);
            const uint32_t ref_base = (uint32_t)refs.size();
            for (size_t k = 0; k < ft.words.size(); k += 2) {
                uint32_t w0 = ft.words[k];
                if ((ft.words[k + 1] & 3u) == 3u) w0 += ref_base;
                words.push_back(w0);
                words.push_back(ft.words[k + 1]);
            }
            for (uint32_t r : ft.refs) refs.push_back(ob.tri_base + r);
            if (nt == 0) {  // empty mesh: one empty leaf
                ob.node_base = (uint32_t)(words.size() / 2);
                words.push_back(0);
                words.push_back(3u);
            }
        } else if (od.shape == LRT_SHAPE_RECTANGLE) {
            const f3 eu = ld3(od.edge_u), ev = ld3(od.edge_v);
            const f3 n = cross(eu, ev);
            const float area = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
            for (int c = 0; c < 3; ++c) {
                ob.corner[c] = od.corner[c];
                ob.eu[c] = od.edge_u[c];
                ob.ev[c] = od.edge_v[c];
                ob.rn[c] = area > 0.0f ? comp(n, c) / area : 0.0f;
            }
            ob.area = area;
        } else if (od.shape == LRT_SHAPE_SPHERE) {
            for (int c = 0; c < 3; ++c) ob.center[c] = od.center[c];
            ob.radius = od.radius;
            ob.area = 4.0f * kPiF * od.radius * od.radius;
        } else {
            throw Error(LRT_ERR_ARG, "object " + std::to_string(o) + ": unknown shape");
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (words.empty()) { words.push_back(0); words.push_back(3u); }
    const size_t n_nodes = words.size() / 2;
    if (refs.empty()) refs.push_back(0);
    if (tris.empty()) tris.assign(9, 0.0f);
    if (cn.empty()) cn.assign(9, 0.0f);
    if (tmat.empty()) tmat.push_back(0);
    upload(sc.nodes, words.data(), words.size() * 4, sc.stream);
    upload(sc.refs, refs.data(), refs.size() * 4, sc.stream);
    upload(sc.tris, tris.data(), tris.size() * 4, sc.stream);
    {
        // the triangle records again in leaf-reference order
        std::vector<float> leaf(refs.size() * 9);
        for (size_t i = 0; i < refs.size(); ++i)
            std::memcpy(&leaf[9 * i], &tris[9 * (size_t)refs[i]], 9 * sizeof(float));
        upload(sc.leaf_tris, leaf.data(), leaf.size() * 4, sc.stream);
        HIPCHECK(hipStreamSynchronize(sc.stream));
    }
    upload(sc.cnormals, cn.data(), cn.size() * 4, sc.stream);
    upload(sc.trimat, tmat.data(), tmat.size() * 2, sc.stream);
    std::vector<lrt_material> mats(d.materials, d.materials + d.n_materials);
    if (mats.empty()) { lrt_material z; std::memset(&z, 0, sizeof(z)); mats.push_back(z); }
    upload(sc.mats, mats.data(), mats.size() * sizeof(lrt_material), sc.stream);
    std::vector<float> tb(d.tables ? d.tables : nullptr, d.tables ? d.tables + (size_t)d.n_tables * 101 : nullptr);
    if (tb.empty()) tb.assign(101, 0.0f);
    // the mesh boxes, enlarged so that rounding in the slab test cannot reject
    // a ray whose hit lies on the box (a triangle hit lies inside its mesh's box)
    for (uint32_t o = 0; o < d.n_objects; ++o) {
        float* bx = obj_box.data() + 6 * (size_t)o;
        if (!(bx[0] <= bx[3])) continue;
        float ext = 0.0f, mag = 0.0f;
        for (int a = 0; a < 3; ++a) {
            ext = std::max(ext, bx[3 + a] - bx[a]);
            mag = std::max(mag, std::max(std::fabs(bx[a]), std::fabs(bx[3 + a])));
        }
        const float m = 1e-3f * std::max(ext, mag) + 1e-4f;
        for (int a = 0; a < 3; ++a) { bx[a] -= m; bx[3 + a] += m; }
    }
    upload(sc.obj_box, obj_box.data(), obj_box.size() * 4, sc.stream);
    upload(sc.tables, tb.data(), tb.size() * 4, sc.stream);
    HIPCHECK(hipStreamSynchronize(sc.stream));
    (void)any_tmat;
    sc.S.obj_box = sc.obj_box.as<float>();
    sc.S.nodes = sc.nodes.as<uint2>();
    sc.S.refs = sc.refs.as<uint32_t>();
    sc.S.tris = sc.tris.as<float>();
    sc.S.leaf_tris = sc.leaf_tris.as<float>();
    sc.S.cnormals = sc.cnormals.as<float>();
    sc.S.tri_mat = sc.trimat.as<uint16_t>();
    sc.S.n_objects = (int32_t)d.n_objects;
    sc.S.stack_depth = max_depth;
    sc.M.m = sc.mats.as<lrt_material>();
    sc.M.tables = sc.tables.as<float>();
    std::memset(&sc.Ls, 0, sizeof(sc.Ls));
    sc.Ls.n = (int32_t)d.n_lights;
    sc.S.nonconvex_emitters = 0u;
    sc.mesh_lights = false;
    for (uint32_t i = 0; i < d.n_lights; ++i) {
        sc.Ls.l[i] = make_light(d.lights[i], sc.objects);
        if (d.lights[i].kind == LRT_LIGHT_AREA && d.objects[d.lights[i].object].shape == LRT_SHAPE_MESH) {
            const int o = d.lights[i].object;
            sc.Ls.l[i].cdf_base = obj_cdf_base[(size_t)o];
            sc.Ls.l[i].n_tris = obj_cdf_n[(size_t)o];
            sc.S.nonconvex_emitters |= 1u << i;
            sc.mesh_lights = true;
        }
    }
    if (lcdf.empty()) lcdf.push_back(0.0f);
    if (ltris.empty()) ltris.assign(9, 0.0f);
    upload(sc.light_cdf, lcdf.data(), lcdf.size() * 4, sc.stream);
    upload(sc.light_tris, ltris.data(), ltris.size() * 4, sc.stream);
    sc.Ls.cdf = sc.light_cdf.as<float>();
    sc.Ls.tri9 = sc.light_tris.as<float>();
    for (uint32_t o = 0; o < d.n_objects; ++o)
        if (d.objects[o].emitter >= 0) {
            const lrt_light_desc& L = d.lights[d.objects[o].emitter];
            require(L.kind == LRT_LIGHT_AREA && L.object == (int)o, "object " + std::to_string(o) + ": emitter does not name this object");
        }
    // emission CDF by luminance(power) over area lights (scene.h:21-34)
    std::memset(&sc.emit, 0, sizeof(sc.emit));
    float esum = 0.0f;
    for (uint32_t i = 0; i < d.n_lights; ++i) {
        const lrt_light_desc& L = d.lights[i];
        if (L.kind != LRT_LIGHT_AREA) continue;
        const float area = sc.S.obj[L.object].area;
        const float pa = area <= 0.0f ? 0.0f : kPiF * area;
        const float w = 0.212671f * (L.radiance[0] * pa) + 0.715160f * (L.radiance[1] * pa) + 0.072169f * (L.radiance[2] * pa);
        if (w <= 0.0f) continue;
        esum += w;
        sc.emit.index[sc.emit.n] = (int32_t)i;
        sc.emit.cdf[sc.emit.n] = esum;
        sc.emit.n++;
    }
    sc.emit.total = esum;
    // scene::bounding_sphere_radius and has_delta_bsdf (scene.h:41-57)
    {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        auto grow = [&](const float* q) {
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], q[a]); hi[a] = std::max(hi[a], q[a]); }
        };
        std::vector<char> delta_mat(mats.size(), 0);
        for (size_t m = 0; m < mats.size() && d.n_materials; ++m)
            delta_mat[m] = (host_bsdf_flags(mats, (int)m) & LRT_FLAG_SPECULAR) != 0;
        bool delta = false;
        for (uint32_t o = 0; o < d.n_objects; ++o) {
            const lrt_object_desc& od = d.objects[o];
            float b0[3] = {INFINITY, INFINITY, INFINITY}, b1[3] = {-INFINITY, -INFINITY, -INFINITY};
            auto bgrow = [&](const float* q) {
                for (int a = 0; a < 3; ++a) { b0[a] = std::min(b0[a], q[a]); b1[a] = std::max(b1[a], q[a]); }
            };
            if (od.shape == LRT_SHAPE_MESH) {
                for (uint32_t v = 0; v < od.n_vertices; ++v) bgrow(od.positions + 3 * (size_t)v);
                if (od.triangle_material) {
                    for (uint32_t t = 0; t < od.n_triangles; ++t)
                        if (od.triangle_material[t] < delta_mat.size() && delta_mat[od.triangle_material[t]]) delta = true;
                } else if (od.material >= 0 && (size_t)od.material < delta_mat.size() && delta_mat[(size_t)od.material]) {
                    delta = true;
                }
            } else if (od.shape == LRT_SHAPE_RECTANGLE) {
                float q[3];
                bgrow(od.corner);
                for (int a = 0; a < 3; ++a) q[a] = od.corner[a] + od.edge_u[a];
                bgrow(q);
                for (int a = 0; a < 3; ++a) q[a] = od.corner[a] + od.edge_v[a];
                bgrow(q);
                for (int a = 0; a < 3; ++a) q[a] = (od.corner[a] + od.edge_u[a]) + od.edge_v[a];
                bgrow(q);
            } else {
                float q[3];
                for (int a = 0; a < 3; ++a) q[a] = od.center[a] - od.radius;
                bgrow(q);
                for (int a = 0; a < 3; ++a) q[a] = od.center[a] + od.radius;
                bgrow(q);
                if (od.material >= 0 && (size_t)od.material < delta_mat.size() && delta_mat[(size_t)od.material]) delta = true;
            }
            grow(b0);
            grow(b1);
        }
        float h[3];
        for (int a = 0; a < 3; ++a) h[a] = (hi[a] - lo[a]) * 0.5f;
        if (d.n_objects)
            for (int a = 0; a < 3; ++a) { sc.box_lo[a] = lo[a]; sc.box_hi[a] = hi[a]; }
        sc.bounding_radius = d.n_objects ? std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) : 0.0f;
        sc.has_delta = delta;
        sc.max_spec_components = d.n_materials ? host_max_specular_components(mats) : 0;
        sc.host_mats = mats;
    }
    const CameraBasis cb = camera_basis(d.camera);
    std::memcpy(&sc.cam, &cb, sizeof(sc.cam));
    const auto t2 = std::chrono::steady_clock::now();
    std::memset(&sc.info, 0, sizeof(sc.info));
    sc.info.n_triangles = tris.size() / 9;
    sc.info.n_nodes = n_nodes;
    sc.info.n_refs = refs.size();
    sc.info.device_bytes = sc.nodes.bytes + sc.refs.bytes + sc.tris.bytes + sc.leaf_tris.bytes + sc.cnormals.bytes + sc.trimat.bytes;
    sc.info.max_tree_depth = max_depth;
    sc.info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    sc.info.upload_seconds = std::chrono::duration<double>(t2 - t1).count();
    sc.info.bounding_radius = sc.bounding_radius;
    sc.info.has_delta_bsdf = sc.has_delta ? 1 : 0;
}

uint32_t grid_for(const lrt_scene& sc, uint64_t n, int per_cu) {
    const uint64_t need = (n + kBlock - 1) / kBlock;
    const uint64_t cap = (uint64_t)sc.cus * (uint64_t)per_cu;
    return (uint32_t)std::max<uint64_t>(1, std::min(need, cap));
}

template <typename T>
T* alloc_carve(char*& p, size_t n) {
    T* r = (T*)p;
    p += ((n * sizeof(T) + 255) / 256) * 256;
    return r;
}

void stage_rays(lrt_scene& sc, const float* rays, uint32_t n, int device_ptrs, const float*& d_rays) {
    if (device_ptrs) {
        after_caller(sc.stream, sc.caller);
        d_rays = rays;
        return;
    }
    upload(sc.q_rays, rays, (size_t)n * 8 * 4, sc.stream);
    d_rays = sc.q_rays.as<float>();
}

// ---- BSDF and sky batches (host arrays) ---------------------------------------
struct BatchDev {
    hipStream_t s = nullptr;
    std::vector<void*> bufs;
    explicit BatchDev(int dev) {
        HIPCHECK(hipSetDevice(dev));
        HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    }
    ~BatchDev() {
        if (s) (void)hipStreamSynchronize(s);
        for (void* p : bufs) (void)hipFree(p);
        if (s) (void)hipStreamDestroy(s);
    }
    template <typename T>
    T* put(const T* src, size_t count) {
        void* p = nullptr;
        const size_t bytes = std::max<size_t>(count * sizeof(T), 4);
        if (hipMalloc(&p, bytes) != hipSuccess) throw Error(LRT_ERR_ALLOC, "hipMalloc failed");
        bufs.push_back(p);
        if (src && count) HIPCHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
        return (T*)p;
    }
    template <typename T>
    void get(T* dst, const T* src, size_t count) {
        HIPCHECK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

}  // namespace

extern "C" {

const char* lrt_last_error(void) { return g_error.c_str(); }
int lrt_version(void) { return 1; }
int lrt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lrt_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    return guarded([&] {
        require(ctr && key && out, "null pointer");
        philox(ctr, key, out);
    });
}

int lrt_kdtree_build_export(const float* boxes, uint32_t n, uint32_t* n_nodes, uint32_t* n_refs, int32_t* depth_limit,
                            uint8_t* leaf, int32_t* axis, float* split, int32_t* left, int32_t* right, int32_t* first,
                            int32_t* count, uint32_t* refs) {
    return guarded([&] {
        require(n == 0 || boxes, "boxes is null");
        KdTree t;
        kd_build(boxes, n, t);
        if (n_nodes) *n_nodes = (uint32_t)t.nodes.size();
        if (n_refs) *n_refs = (uint32_t)t.refs.size();
        if (depth_limit) *depth_limit = t.depth_limit;
        if (leaf) {
            require(axis && split && left && right && first && count && refs, "output arrays are null");
            for (size_t i = 0; i < t.nodes.size(); ++i) {
                const KdNode& nd = t.nodes[i];
                leaf[i] = nd.leaf; axis[i] = nd.axis; split[i] = nd.split; left[i] = nd.left;
                right[i] = nd.right; first[i] = nd.first; count[i] = nd.count;
            }
            if (!t.refs.empty()) std::memcpy(refs, t.refs.data(), t.refs.size() * 4);
        }
    });
}

int lrt_fresnel_diffuse_reflectance(float eta, float* out) {
    return guarded([&] {
        require(out != nullptr, "out is null");
        *out = fresnel_diffuse_reflectance(eta);
    });
}

int lrt_rough_transmittance_slice(const void* table_file, size_t table_bytes, float eta, float alpha, float* out101) {
    return guarded([&] {
        require(table_file && out101, "null pointer");
        rough_transmittance_slice(table_file, table_bytes, eta, alpha, out101);
    });
}

// device address of `scratch` per device, resolved once per device (with the
// device current) by lrt_scene_create
constexpr int kMaxDevices = 64;
static uint32_t* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static void scratch_setup(int dev) {
    require(dev < kMaxDevices, "device index beyond " + std::to_string(kMaxDevices), LRT_ERR_UNSUPPORTED);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return;
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "hipGetSymbolAddress(scratch) returned null", LRT_ERR_HIP);
    g_scratch_ptr[dev] = (uint32_t*)p;
}

int lrt_scene_create(int hip_device, const lrt_scene_desc* desc, lrt_scene** out) {
    return guarded([&] {
        require(desc && out, "null pointer");
        *out = nullptr;
        int ndev = 0;
        require(hipGetDeviceCount(&ndev) == hipSuccess && hip_device >= 0 && hip_device < ndev,
                "no HIP device " + std::to_string(hip_device), LRT_ERR_HIP);
        HIPCHECK(hipSetDevice(hip_device));
        scratch_setup(hip_device);
        auto sc = std::make_unique<lrt_scene>();
        sc->device = hip_device;
        sc->cus = cu_count(hip_device);
        HIPCHECK(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&sc->caller, hipEventDisableTiming));
        build_scene(*sc, *desc);
        *out = sc.release();
    });
}

int lrt_scene_release(lrt_scene* scene) {
    return guarded([&] {
        if (!scene) return;
        (void)hipSetDevice(scene->device);
        (void)hipStreamSynchronize(scene->stream);
        delete scene;
    });
}

int lrt_scene_info_get(const lrt_scene* scene, lrt_scene_info* info) {
    return guarded([&] {
        require(scene && info, "null pointer");
        *info = scene->info;
    });
}

int lrt_scene_sync(lrt_scene* scene) {
    return guarded([&] {
        require(scene, "null scene");
        HIPCHECK(hipSetDevice(scene->device));
        HIPCHECK(hipStreamSynchronize(scene->stream));
        for (auto& t : scene->timers) t.collect();
    });
}

int lrt_scene_kernel_timing(lrt_scene* scene, int kernel, int reset, double* total_ms, uint64_t* launches, uint64_t* items) {
    return guarded([&] {
        require(scene && kernel >= 0 && kernel < LRT_KERNEL_COUNT_, "bad arguments");
        HIPCHECK(hipSetDevice(scene->device));
        HIPCHECK(hipStreamSynchronize(scene->stream));
        Timer& t = scene->timers[kernel];
        t.collect();
        if (total_ms) *total_ms = t.total_ms;
        if (launches) *launches = t.launches;
        if (items) *items = t.items;
        if (reset) { t.total_ms = 0.0; t.launches = 0; t.items = 0; }
    });
}

// ---- ray queries -------------------------------------------------------------
int lrt_trace_closest(lrt_scene* sc, const float* rays, uint32_t n, int32_t* hit_obj_prim, float* hit_t, float* hit_uv,
                      int device_ptrs) {
    return guarded([&] {
        require(sc && rays && hit_obj_prim && hit_t, "null pointer");
        if (n == 0) return;
        HIPCHECK(hipSetDevice(sc->device));
        const float* d_rays;
        stage_rays(*sc, rays, n, device_ptrs, d_rays);
        int32_t* d_op = hit_obj_prim;
        float* d_t = hit_t;
        float* d_uv = hit_uv;
        if (!device_ptrs) {
            sc->q_out0.alloc((size_t)n * 8);
            sc->q_out1.alloc((size_t)n * 4);
            d_op = sc->q_out0.as<int32_t>();
            d_t = sc->q_out1.as<float>();
            if (hit_uv) { sc->q_out2.alloc((size_t)n * 8); d_uv = sc->q_out2.as<float>(); }
        }
        Timer& tm = sc->timers[LRT_KERNEL_CLOSEST];
        auto ev = tm.next();
        HIPCHECK(hipEventRecord(ev.first, sc->stream));
        sc->counters.alloc(256);
        uint32_t* next = sc->counters.as<uint32_t>() + 8;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sc->stream, next, 1u, 0u);
        hipLaunchKernelGGL(k_trace_closest, dim3(grid_for(*sc, n, 8)), dim3(kBlock), 0, sc->stream, sc->S, d_rays, n, d_op, d_t,
                           d_uv, next);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ev.second, sc->stream));
        tm.items += n;
        if (!device_ptrs) {
            HIPCHECK(hipMemcpyAsync(hit_obj_prim, d_op, (size_t)n * 8, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipMemcpyAsync(hit_t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost, sc->stream));
            if (hit_uv) HIPCHECK(hipMemcpyAsync(hit_uv, d_uv, (size_t)n * 8, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipStreamSynchronize(sc->stream));
        }
    });
}

int lrt_trace_closest_noalloc(lrt_scene* sc, const float* rays, uint32_t n, int32_t* hit_obj_prim, float* hit_t,
                              float* hit_uv) {
    // no allocation on any path that succeeds: no strings are built unless a
    // check fails, no timers, no staging; the fetch counter is `scratch` word 0
    // (cleared by a kernel, not a memset node)
    return guarded([&] {
        if (!sc || !rays || !hit_obj_prim || !hit_t) throw Error(LRT_ERR_ARG, "null pointer");
        if ((uint64_t)n > kNoAllocMaxRays)
            throw Error(LRT_ERR_ARG, "lrt_trace_closest_noalloc: n = " + std::to_string(n) +
                                         " rays exceeds the maximum of 2^31 rays per call");
        if (n == 0) return;
        uint32_t* next = (sc->device >= 0 && sc->device < kMaxDevices) ? g_scratch_ptr[sc->device] : nullptr;
        if (!next) throw Error(LRT_ERR_INTERNAL, "scratch of the device was not set up");
        HIPCHECK(hipSetDevice(sc->device));
        after_caller(sc->stream, sc->caller);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sc->stream, next, 1u, 0u);
        hipLaunchKernelGGL(k_trace_closest, dim3(grid_for(*sc, n, 8)), dim3(kBlock), 0, sc->stream, sc->S, rays, n,
                           hit_obj_prim, hit_t, hit_uv, next);
        HIPCHECK(hipGetLastError());
    });
}

int lrt_trace_occluded(lrt_scene* sc, const float* rays, uint32_t n, const int32_t* ignore_emitter, uint8_t* occluded,
                       int device_ptrs) {
    return guarded([&] {
        require(sc && rays && occluded, "null pointer");
        if (n == 0) return;
        HIPCHECK(hipSetDevice(sc->device));
        const float* d_rays;
        stage_rays(*sc, rays, n, device_ptrs, d_rays);
        const int32_t* d_ign = ignore_emitter;
        uint8_t* d_out = occluded;
        if (!device_ptrs) {
            if (ignore_emitter) { upload(sc->q_aux, ignore_emitter, (size_t)n * 4, sc->stream); d_ign = sc->q_aux.as<int32_t>(); }
            sc->q_out0.alloc(n);
            d_out = sc->q_out0.as<uint8_t>();
        }
        Timer& tm = sc->timers[LRT_KERNEL_OCCLUDED];
        auto ev = tm.next();
        HIPCHECK(hipEventRecord(ev.first, sc->stream));
        sc->counters.alloc(256);
        uint32_t* next = sc->counters.as<uint32_t>() + 8;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sc->stream, next, 1u, 0u);
        hipLaunchKernelGGL(k_trace_occluded, dim3(grid_for(*sc, n, 8)), dim3(kBlock), 0, sc->stream, sc->S, d_rays, n, d_ign,
                           d_out, next);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ev.second, sc->stream));
        tm.items += n;
        if (!device_ptrs) {
            HIPCHECK(hipMemcpyAsync(occluded, d_out, n, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipStreamSynchronize(sc->stream));
        }
    });
}

int lrt_trace_count(lrt_scene* sc, const float* rays, uint32_t n, uint32_t* nodes, uint32_t* refs, uint32_t* tris,
                    int device_ptrs) {
    return guarded([&] {
        require(sc && rays && nodes && refs && tris, "null pointer");
        if (n == 0) return;
        HIPCHECK(hipSetDevice(sc->device));
        const float* d_rays;
        stage_rays(*sc, rays, n, device_ptrs, d_rays);
        uint32_t *a = nodes, *b = refs, *c = tris;
        if (!device_ptrs) {
            sc->q_out0.alloc((size_t)n * 4);
            sc->q_out1.alloc((size_t)n * 4);
            sc->q_out2.alloc((size_t)n * 4);
            a = sc->q_out0.as<uint32_t>(); b = sc->q_out1.as<uint32_t>(); c = sc->q_out2.as<uint32_t>();
        }
        hipLaunchKernelGGL(k_trace_count, dim3(grid_for(*sc, n, 20)), dim3(kBlock), 0, sc->stream, sc->S, d_rays, n, a, b, c);
        HIPCHECK(hipGetLastError());
        if (!device_ptrs) {
            HIPCHECK(hipMemcpyAsync(nodes, a, (size_t)n * 4, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipMemcpyAsync(refs, b, (size_t)n * 4, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipMemcpyAsync(tris, c, (size_t)n * 4, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipStreamSynchronize(sc->stream));
        }
    });
}

int lrt_bsdf_eval_pdf(int hip_device, const lrt_material* materials, uint32_t n_materials, const float* tables,
                      uint32_t n_tables, const int32_t* mat, const float* wi, const float* wo, const uint8_t* front_face,
                      uint32_t n, int measure, int mode, int component, uint32_t type_mask, float* f_cos, float* pdf) {
    return guarded([&] {
        require(materials && mat && wi && wo && front_face && f_cos && pdf, "null pointer");
        require(n_tables == 0 || tables, "tables is null");
        check_materials(materials, n_materials, n_tables);
        for (uint32_t i = 0; i < n; ++i) require(mat[i] >= 0 && (uint32_t)mat[i] < n_materials, "material index out of range");
        if (n == 0) return;
        BatchDev B(hip_device);
        Mats M{B.put(materials, n_materials), B.put(tables, (size_t)n_tables * 101)};
        const int32_t* dm = B.put(mat, n);
        const float* dwi = B.put(wi, 3 * (size_t)n);
        const float* dwo = B.put(wo, 3 * (size_t)n);
        const uint8_t* df = B.put(front_face, n);
        float* dfc = B.put<float>(nullptr, 3 * (size_t)n);
        float* dp = B.put<float>(nullptr, n);
        hipLaunchKernelGGL(k_bsdf_eval, dim3((n + 255) / 256), dim3(256), 0, B.s, M, dm, dwi, dwo, df, n, measure, mode,
                           component, type_mask, dfc, dp);
        HIPCHECK(hipGetLastError());
        B.get(f_cos, dfc, 3 * (size_t)n);
        B.get(pdf, dp, n);
        HIPCHECK(hipStreamSynchronize(B.s));
    });
}

int lrt_bsdf_param_grad(int hip_device, const lrt_material* materials, uint32_t n_materials, const float* tables,
                        uint32_t n_tables, const int32_t* mat, const float* wi, const float* wo,
                        const uint8_t* front_face, uint32_t n, int mode, int which, float* df) {
    return guarded([&] {
        require(materials && mat && wi && wo && front_face && df, "null pointer");
        require(n_tables == 0 || tables, "tables is null");
        require(which == 0 || which == 1, "which must be 0 (alpha) or 1 (interior IOR)");
        check_materials(materials, n_materials, n_tables);
        for (uint32_t i = 0; i < n; ++i) require(mat[i] >= 0 && (uint32_t)mat[i] < n_materials, "material index out of range");
        if (n == 0) return;
        BatchDev B(hip_device);
        Mats M{B.put(materials, n_materials), B.put(tables, (size_t)n_tables * 101)};
        const int32_t* dm = B.put(mat, n);
        const float* dwi = B.put(wi, 3 * (size_t)n);
        const float* dwo = B.put(wo, 3 * (size_t)n);
        const uint8_t* dfr = B.put(front_face, n);
        float* dd = B.put<float>(nullptr, 3 * (size_t)n);
        hipLaunchKernelGGL(k_bsdf_param_grad, dim3((n + 255) / 256), dim3(256), 0, B.s, M, dm, dwi, dwo, dfr, n, mode,
                           which, dd);
        HIPCHECK(hipGetLastError());
        B.get(df, dd, 3 * (size_t)n);
        HIPCHECK(hipStreamSynchronize(B.s));
    });
}

int lrt_bsdf_sample(int hip_device, const lrt_material* materials, uint32_t n_materials, const float* tables,
                    uint32_t n_tables, const int32_t* mat, const float* wi, const uint8_t* front_face, const float* u,
                    const float* lobe_u, uint32_t n, int mode, int component, uint32_t type_mask, float* weight, float* wo,
                    float* pdf, float* eta, uint32_t* sampled_type, int32_t* sampled_component) {
    return guarded([&] {
        require(materials && mat && wi && front_face && u && lobe_u && weight && wo && pdf && eta && sampled_type &&
                    sampled_component, "null pointer");
        require(n_tables == 0 || tables, "tables is null");
        check_materials(materials, n_materials, n_tables);
        for (uint32_t i = 0; i < n; ++i) require(mat[i] >= 0 && (uint32_t)mat[i] < n_materials, "material index out of range");
        if (n == 0) return;
        BatchDev B(hip_device);
        Mats M{B.put(materials, n_materials), B.put(tables, (size_t)n_tables * 101)};
        const int32_t* dm = B.put(mat, n);
        const float* dwi = B.put(wi, 3 * (size_t)n);
        const uint8_t* df = B.put(front_face, n);
        const float* du = B.put(u, 2 * (size_t)n);
        const float* dl = B.put(lobe_u, n);
        float* dw = B.put<float>(nullptr, 3 * (size_t)n);
        float* dwo = B.put<float>(nullptr, 3 * (size_t)n);
        float* dp = B.put<float>(nullptr, n);
        float* de = B.put<float>(nullptr, n);
        uint32_t* dt = B.put<uint32_t>(nullptr, n);
        int32_t* dc = B.put<int32_t>(nullptr, n);
        hipLaunchKernelGGL(k_bsdf_sample, dim3((n + 255) / 256), dim3(256), 0, B.s, M, dm, dwi, df, du, dl, n, mode,
                           component, type_mask, dw, dwo, dp, de, dt, dc);
        HIPCHECK(hipGetLastError());
        B.get(weight, dw, 3 * (size_t)n);
        B.get(wo, dwo, 3 * (size_t)n);
        B.get(pdf, dp, n);
        B.get(eta, de, n);
        B.get(sampled_type, dt, n);
        B.get(sampled_component, dc, n);
        HIPCHECK(hipStreamSynchronize(B.s));
    });
}

int lrt_bsdf_flags(const lrt_material* materials, uint32_t n_materials, int32_t index, uint32_t* flags, int32_t* component_count) {
    return guarded([&] {
        require(materials && flags && component_count && index >= 0 && (uint32_t)index < n_materials, "bad arguments");
        check_materials(materials, n_materials, 0xffffffffu);
        auto base = [](const lrt_material& M) -> uint32_t {
            switch (M.type) {
                case LRT_BSDF_DIFFUSE: return LRT_FLAG_DIFFUSE | LRT_FLAG_REFLECTIVE;
                case LRT_BSDF_CONDUCTOR: return M.alpha <= 0.0f ? (LRT_FLAG_SPECULAR | LRT_FLAG_REFLECTIVE | LRT_FLAG_METALLIC)
                                                                : (LRT_FLAG_GLOSSY | LRT_FLAG_REFLECTIVE | LRT_FLAG_METALLIC);
                case LRT_BSDF_DIELECTRIC: return LRT_FLAG_SPECULAR | LRT_FLAG_REFLECTIVE | LRT_FLAG_REFRACTIVE | LRT_FLAG_DIELECTRIC;
                case LRT_BSDF_ROUGH_DIELECTRIC: return LRT_FLAG_GLOSSY | LRT_FLAG_REFLECTIVE | LRT_FLAG_REFRACTIVE | LRT_FLAG_DIELECTRIC;
                case LRT_BSDF_PLASTIC: return LRT_FLAG_SPECULAR | LRT_FLAG_REFLECTIVE | LRT_FLAG_DIELECTRIC | LRT_FLAG_DIFFUSE;
                case LRT_BSDF_ROUGH_PLASTIC: return LRT_FLAG_GLOSSY | LRT_FLAG_REFLECTIVE | LRT_FLAG_DIELECTRIC | LRT_FLAG_DIFFUSE;
                default: return 0u;
            }
        };
        auto count = [](const lrt_material& M) {
            return (M.type == LRT_BSDF_DIELECTRIC || M.type == LRT_BSDF_PLASTIC || M.type == LRT_BSDF_ROUGH_PLASTIC) ? 2 : 1;
        };
        const lrt_material& M = materials[index];
        if (M.type == LRT_BSDF_COATING || M.type == LRT_BSDF_ROUGH_COATING) {
            const lrt_material& N = materials[M.nested];
            *flags = base(N) | (M.type == LRT_BSDF_COATING ? LRT_FLAG_SPECULAR : LRT_FLAG_GLOSSY) | LRT_FLAG_REFLECTIVE |
                     LRT_FLAG_DIELECTRIC;
            *component_count = count(N) + 1;
        } else {
            *flags = base(M);
            *component_count = count(M);
        }
    });
}

int lrt_sky_cook(const double* coefficients, const double* radiances, double turbidity, double albedo,
                 double elevation, double* config27, double* radiance3) {
    return guarded([&] {
        require(coefficients && radiances && config27 && radiance3, "null pointer");
        require(turbidity >= 1.0 && turbidity <= 10.0, "turbidity outside [1, 10]");
        require(albedo >= 0.0 && albedo <= 1.0, "ground albedo outside [0, 1]");
        sky_cook(coefficients, radiances, turbidity, albedo, elevation, config27, radiance3);
    });
}

int lrt_sky_elevation(const float* sun_dir, double* elevation) {
    return guarded([&] {
        require(sun_dir && elevation, "null pointer");
        // vec3 normalize (each component over the float length), then
        // std::asin of the clamped float y
        const float l = std::sqrt(sun_dir[0] * sun_dir[0] + sun_dir[1] * sun_dir[1] + sun_dir[2] * sun_dir[2]);
        require(l > 0.0f, "zero sun direction");
        const float y = sun_dir[1] / l;
        *elevation = (double)std::asin(std::min(std::max(y, -1.0f), 1.0f));
    });
}

int lrt_sky_eval(int hip_device, const lrt_light_desc* sky, const float* dirs, uint32_t n, float* rgb) {
    return guarded([&] {
        require(sky && dirs && rgb, "null pointer");
        require(sky->kind == LRT_LIGHT_SKY, "not a sky light");
        if (n == 0) return;
        const DevLight L = make_light(*sky, {});
        BatchDev B(hip_device);
        const float* dd = B.put(dirs, 3 * (size_t)n);
        float* dc = B.put<float>(nullptr, 3 * (size_t)n);
        hipLaunchKernelGGL(k_sky_eval, dim3((n + 255) / 256), dim3(256), 0, B.s, L, dd, n, dc);
        HIPCHECK(hipGetLastError());
        B.get(rgb, dc, 3 * (size_t)n);
        HIPCHECK(hipStreamSynchronize(B.s));
    });
}

// ---- path tracer ----------------------------------------------------------------
// device buffers that receive a frame's rays (lrt_render_capture_rays)
struct RayCapture {
    float* ext;
    uint32_t ext_cap;
    float* sh;
    int32_t* sh_ignore;
    uint32_t sh_cap;
    uint32_t* fill;  // 2 counters
};

// aov_out (lrt_render_features): also the first-hit features of the camera
// rays, 7 floats per pixel of the frame (host or device memory as the sums)
static void render_impl(lrt_scene* sc, const lrt_render_params* prm, double* sum_rgb, double* sumsq_rgb,
                        int device_ptrs, lrt_render_stats* stats, const RayCapture* capture = nullptr,
                        float* aov_out = nullptr) {
    {
        require(sc && prm && sum_rgb, "null pointer");
        const uint32_t W = prm->width, H = prm->height;
        require(W >= 2 && H >= 2, "image must be at least 2x2");
        require(prm->max_depth >= 1, "max_depth must be >= 1");
        const uint32_t rb = prm->row_end > prm->row_begin ? prm->row_begin : 0;
        const uint32_t re = prm->row_end > prm->row_begin ? prm->row_end : H;
        require(re <= H, "row range outside the image");
        const uint64_t npix64 = (uint64_t)W * (re - rb);
        require(npix64 < (1ull << 31), "too many pixels");
        const uint32_t npix = (uint32_t)npix64;
        // samples per pass: several samples of every pixel run side by side
        // (up to LRT_PASS_PATHS paths), so the late bounces, whose launches
        // last as long as their slowest ray whatever their size, serve many
        // samples at once
        const uint32_t slots = std::max<uint32_t>(1, std::min<uint32_t>(prm->spp, pass_paths() / npix));
        const uint32_t n = npix * slots;
        HIPCHECK(hipSetDevice(sc->device));
        if (device_ptrs) after_caller(sc->stream, sc->caller);
        const auto t0 = std::chrono::steady_clock::now();
        // work buffers: path state (64 bytes), hits (32 bytes), shadow queue (12), two queues
        if (sc->cap < n) {
            sc->pstate.alloc((size_t)n * 64 + 256);
            sc->hits.alloc((size_t)n * 32 + 256);
            sc->shadow.alloc((size_t)n * 12 * 4 + 12 * 256);
            sc->queues.alloc((size_t)n * 3 * 4 + 3 * 256);
            sc->cap = n;
        }
        sc->counters.alloc(kCounterWords * 4);
        sc->stats.alloc(256);
        PathState P;
        char* c = sc->pstate.as<char>();
        P.rec = alloc_carve<float4>(c, 4 * (size_t)n);
        P.pixel0 = rb * W;
        P.npix = npix;
        HitBuf Hb;
        c = sc->hits.as<char>();
        Hb.rec = alloc_carve<uint4>(c, 2 * (size_t)n);
        ShadowBuf Sb;
        c = sc->shadow.as<char>();
        float** sf[10] = {&Sb.ox, &Sb.oy, &Sb.oz, &Sb.dx, &Sb.dy, &Sb.dz, &Sb.tmax, &Sb.cr, &Sb.cg, &Sb.cb};
        for (float** f : sf) *f = alloc_carve<float>(c, n);
        Sb.ignore = alloc_carve<int32_t>(c, n);
        Sb.path = alloc_carve<uint32_t>(c, n);
        c = sc->queues.as<char>();
        uint32_t* qa = alloc_carve<uint32_t>(c, n);
        uint32_t* qb = alloc_carve<uint32_t>(c, n);
        uint32_t* order = alloc_carve<uint32_t>(c, n);
        uint32_t* counters = sc->counters.as<uint32_t>();
        unsigned long long* dstats = sc->stats.as<unsigned long long>();
        // accumulation target
        double* dsum = sum_rgb;
        double* dsq = sumsq_rgb;
        const size_t img = (size_t)W * H * 3;
        if (!device_ptrs) {
            sc->accum.alloc(img * 8);
            HIPCHECK(hipMemcpyAsync(sc->accum.p, sum_rgb, img * 8, hipMemcpyHostToDevice, sc->stream));
            dsum = sc->accum.as<double>();
            if (sumsq_rgb) {
                sc->accum2.alloc(img * 8);
                HIPCHECK(hipMemcpyAsync(sc->accum2.p, sumsq_rgb, img * 8, hipMemcpyHostToDevice, sc->stream));
                dsq = sc->accum2.as<double>();
            }
        }
        const unsigned long long zero[4] = {0, 0, 0, 0};
        HIPCHECK(hipMemcpyAsync(dstats, zero, sizeof(zero), hipMemcpyHostToDevice, sc->stream));
        const bool counting = (prm->flags & LRT_RENDER_COUNT) != 0;
        float* aov_path = nullptr;
        float* daov = aov_out;
        const size_t aov_bytes = (size_t)W * H * 7 * 4;
        if (aov_out) {
            sc->aov_path.alloc((size_t)n * 7 * 4);
            aov_path = sc->aov_path.as<float>();
            if (!device_ptrs) {
                sc->aov_dev.alloc(aov_bytes);
                daov = sc->aov_dev.as<float>();
                HIPCHECK(hipMemcpyAsync(daov, aov_out, aov_bytes, hipMemcpyHostToDevice, sc->stream));
            }
        }
        sc->totals.alloc(256);
        // traversal work: reference extension, reference shadow, own extension, own shadow
        const unsigned long long zero12[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        HIPCHECK(hipMemcpyAsync(sc->totals.p, zero12, sizeof(zero12), hipMemcpyHostToDevice, sc->stream));
        const uint2 key = make_uint2((uint32_t)prm->seed, (uint32_t)(prm->seed >> 32));
        const uint32_t* sobol = nullptr;
        if (prm->flags & LRT_RENDER_REFERENCE_SAMPLER) {
            if (!sc->sobol.p) {
                uint32_t dirs[16 * 32];
                sobol_directions(dirs);
                upload(sc->sobol, dirs, sizeof(dirs), sc->stream);
            }
            sobol = sc->sobol.as<uint32_t>();
        }
        const uint32_t g_all = (n + kBlock - 1) / kBlock;
        const uint32_t g_trace = (uint32_t)std::min<uint64_t>(g_all, (uint64_t)sc->cus * 20);
        const uint32_t g_shade = (uint32_t)std::min<uint64_t>(((uint64_t)n + kShadeBlock - 1) / kShadeBlock,
                                                              (uint64_t)sc->cus * LRT_SHADE_BLOCKS_PER_CU * 256 / kShadeBlock);
        const uint32_t g_bins = (uint32_t)std::min<uint64_t>(g_all, (uint64_t)sc->cus * 4);
        // the tail kernel: not with the counting kernels (they count per bounce), ray capture or
        // LRT_RENDER_WAVEFRONT_ONLY
        const bool tail_on = !(counting || capture || (prm->flags & LRT_RENDER_WAVEFRONT_ONLY));
        Timer& tc = sc->timers[LRT_KERNEL_CLOSEST];
        Timer& to = sc->timers[LRT_KERNEL_OCCLUDED];
        Timer& ts = sc->timers[LRT_KERNEL_SHADE];
        Timer& tt = sc->timers[LRT_KERNEL_TAIL];
        for (uint32_t s = 0; s < prm->spp; s += slots) {
            const uint32_t smp = prm->spp_offset + s;
            const uint32_t ns = std::min(slots, prm->spp - s);  // samples in this pass
            const uint32_t np = npix * ns;
            const uint32_t tail = tail_on ? tail_paths(np) : 0u;
            const uint32_t g_tail = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(((uint64_t)tail + kBlock - 1) / kBlock, (uint64_t)sc->cus * 8));
            hipLaunchKernelGGL(k_raygen, dim3((std::max(np, 1024u) + kBlock - 1) / kBlock), dim3(kBlock), 0, sc->stream, P, sc->cam, W, H, rb,
                               np, smp, key, sobol, qa, counters);
            uint32_t* q = qa;
            uint32_t* qn = qb;
            unsigned long long* totals = sc->totals.as<unsigned long long>();
            // camera rays; every later bounce traces its extension rays in the
            // previous bounce's k_trace_queues launch
            if (counting)
                hipLaunchKernelGGL(k_extend_count, dim3(g_trace), dim3(kBlock), 0, sc->stream, sc->S, P, q, counters, totals);
            auto e0 = tc.next();
            HIPCHECK(hipEventRecord(e0.first, sc->stream));
            hipLaunchKernelGGL(k_extend, dim3(g_trace), dim3(kBlock), 0, sc->stream, sc->S, P, q, counters, Hb);
            HIPCHECK(hipEventRecord(e0.second, sc->stream));
            for (uint32_t depth = 0; depth < prm->max_depth; ++depth) {
                ShadeParams sp{depth, prm->max_depth, prm->rr_depth, smp, key, sobol, aov_path};
                if (tail && depth >= prm->rr_depth) {  // few paths left: finish them in one launch
                    auto et = tt.next();
                    HIPCHECK(hipEventRecord(et.first, sc->stream));
                    hipLaunchKernelGGL(sc->mesh_lights ? k_tail<true> : k_tail<false>, dim3(g_tail), dim3(kBlock), 0,
                                       sc->stream, sc->S, sc->Ls, sc->M, P, q, counters, Hb, sp, tail, dstats);
                    HIPCHECK(hipEventRecord(et.second, sc->stream));
                    hipLaunchKernelGGL(k_tail_done, dim3(1), dim3(64), 0, sc->stream, counters, tail);
                    // once the tail has finished the pass, the remaining bounces' launches would find
                    // nothing: read the queue length back (one 4-byte copy) and end the pass
                    uint32_t left = 0;
                    HIPCHECK(hipMemcpyAsync(&left, counters, 4, hipMemcpyDeviceToHost, sc->stream));
                    HIPCHECK(hipStreamSynchronize(sc->stream));
                    if (left == 0) break;
                }
                auto e2 = ts.next();
                HIPCHECK(hipEventRecord(e2.first, sc->stream));
                hipLaunchKernelGGL(k_shade_bins, dim3(g_bins), dim3(kBlock), 0, sc->stream, sc->S, Hb, counters);
                hipLaunchKernelGGL(k_shade_scan, dim3(1), dim3(1024), 0, sc->stream, counters);
                hipLaunchKernelGGL(k_shade_order, dim3(g_bins), dim3(kBlock), 0, sc->stream, sc->S, Hb, counters, order);
                hipLaunchKernelGGL(sc->mesh_lights ? k_shade<true> : k_shade<false>, dim3(g_shade), dim3(kShadeBlock), 0,
                                   sc->stream, sc->S, sc->Ls, sc->M, P, q, qn,
                                   counters, Hb, order, Sb, sp);
                HIPCHECK(hipEventRecord(e2.second, sc->stream));
                if (capture)
                    hipLaunchKernelGGL(k_capture_rays, dim3(g_bins), dim3(kBlock), 0, sc->stream, P, qn, counters, Sb,
                                       capture->ext, capture->ext_cap, capture->sh, capture->sh_ignore,
                                       capture->sh_cap, capture->fill);
                if (counting) {
                    hipLaunchKernelGGL(k_shadow_count, dim3(g_trace), dim3(kBlock), 0, sc->stream, sc->S, Sb, counters,
                                       totals);
                    hipLaunchKernelGGL(k_extend_count, dim3(g_trace), dim3(kBlock), 0, sc->stream, sc->S, P, qn,
                                       counters + 1, totals);
                }
                auto e3 = tc.next();
                HIPCHECK(hipEventRecord(e3.first, sc->stream));
                hipLaunchKernelGGL(k_trace_queues, dim3(g_trace), dim3(kBlock), 0, sc->stream, sc->S, P, qn, counters, Hb,
                                   Sb);
                HIPCHECK(hipEventRecord(e3.second, sc->stream));
                hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, sc->stream, counters, dstats);
                std::swap(q, qn);
            }
            hipLaunchKernelGGL(k_accumulate, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, sc->stream, P, ns, dsum, dsq);
            if (aov_path)
                hipLaunchKernelGGL(k_aov_resolve, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, sc->stream, P, ns,
                                   aov_path, daov, s == 0 ? 1 : 0, s + ns >= prm->spp ? 1.0f / (float)prm->spp : 0.0f);
            HIPCHECK(hipGetLastError());
        }
        if (aov_out && !device_ptrs) HIPCHECK(hipMemcpyAsync(aov_out, daov, aov_bytes, hipMemcpyDeviceToHost, sc->stream));
        unsigned long long hst[4] = {0, 0, 0, 0}, work[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        HIPCHECK(hipMemcpyAsync(hst, dstats, sizeof(hst), hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipMemcpyAsync(work, sc->totals.p, sizeof(work), hipMemcpyDeviceToHost, sc->stream));
        if (!device_ptrs) {
            HIPCHECK(hipMemcpyAsync(sum_rgb, dsum, img * 8, hipMemcpyDeviceToHost, sc->stream));
            if (sumsq_rgb) HIPCHECK(hipMemcpyAsync(sumsq_rgb, dsq, img * 8, hipMemcpyDeviceToHost, sc->stream));
        }
        HIPCHECK(hipStreamSynchronize(sc->stream));
        const auto t1 = std::chrono::steady_clock::now();
        const double before_c = tc.total_ms, before_o = to.total_ms, before_s = ts.total_ms;
        const uint64_t lc0 = tc.launches, lo0 = to.launches;
        tc.collect();
        to.collect();
        ts.collect();
        tt.collect();
        tc.items += hst[0];
        to.items += hst[1];
        if (stats) {
            stats->samples = (uint64_t)npix * prm->spp;
            stats->extension_rays = hst[0];
            stats->shadow_rays = hst[1];
            stats->segments = hst[0];
            stats->seconds = std::chrono::duration<double>(t1 - t0).count();
            stats->trace_ms = tc.total_ms - before_c;
            stats->occluded_ms = to.total_ms - before_o;
            stats->shade_ms = ts.total_ms - before_s;
            stats->trace_launches = tc.launches - lc0;
            stats->occluded_launches = to.launches - lo0;
            stats->nodes_visited = work[0];
            stats->refs_read = work[1];
            stats->tris_tested = work[2];
            stats->shadow_nodes_visited = work[3];
            stats->shadow_refs_read = work[4];
            stats->shadow_tris_tested = work[5];
            stats->own_nodes_visited = work[6];
            stats->own_refs_read = work[7];
            stats->own_tris_tested = work[8];
            stats->own_shadow_nodes_visited = work[9];
            stats->own_shadow_refs_read = work[10];
            stats->own_shadow_tris_tested = work[11];
        }
    }
}

int lrt_render(lrt_scene* sc, const lrt_render_params* prm, double* sum_rgb, double* sumsq_rgb, int device_ptrs,
               lrt_render_stats* stats) {
    return guarded([&] { render_impl(sc, prm, sum_rgb, sumsq_rgb, device_ptrs, stats); });
}

int lrt_render_features(lrt_scene* sc, const lrt_render_params* prm, double* sum_rgb, double* sumsq_rgb, float* aov,
                        int device_ptrs, lrt_render_stats* stats) {
    return guarded([&] {
        require(aov != nullptr, "null pointer");
        render_impl(sc, prm, sum_rgb, sumsq_rgb, device_ptrs, stats, nullptr, aov);
    });
}

int lrt_render_capture_rays(lrt_scene* sc, const lrt_render_params* prm, float* ext_rays, uint32_t ext_cap,
                            float* shadow_rays, int32_t* shadow_ignore, uint32_t shadow_cap, uint32_t* n_ext,
                            uint32_t* n_shadow) {
    return guarded([&] {
        require(sc && prm && n_ext && n_shadow, "null pointer");
        require((ext_cap == 0 || ext_rays) && (shadow_cap == 0 || (shadow_rays && shadow_ignore)), "null pointer");
        HIPCHECK(hipSetDevice(sc->device));
        DevBuf img, e, s, si, fill;
        const size_t npix = (size_t)prm->width * prm->height;
        img.alloc(npix * 3 * 8);
        HIPCHECK(hipMemsetAsync(img.p, 0, npix * 3 * 8, sc->stream));
        e.alloc((size_t)ext_cap * 32);
        s.alloc((size_t)shadow_cap * 32);
        si.alloc((size_t)shadow_cap * 4);
        fill.alloc(8);
        HIPCHECK(hipMemsetAsync(fill.p, 0, 8, sc->stream));
        const RayCapture cap{e.as<float>(), ext_cap, s.as<float>(), si.as<int32_t>(), shadow_cap, fill.as<uint32_t>()};
        render_impl(sc, prm, img.as<double>(), nullptr, 1, nullptr, &cap);
        uint32_t f[2];
        HIPCHECK(hipMemcpyAsync(f, fill.p, 8, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
        *n_ext = std::min(f[0], ext_cap);
        *n_shadow = std::min(f[1], shadow_cap);
        // compact the SoA columns to the stored count
        for (int k = 0; k < 8; ++k) {
            if (*n_ext)
                HIPCHECK(hipMemcpyAsync(ext_rays + (size_t)k * *n_ext, e.as<float>() + (size_t)k * ext_cap,
                                        (size_t)*n_ext * 4, hipMemcpyDeviceToHost, sc->stream));
            if (*n_shadow)
                HIPCHECK(hipMemcpyAsync(shadow_rays + (size_t)k * *n_shadow, s.as<float>() + (size_t)k * shadow_cap,
                                        (size_t)*n_shadow * 4, hipMemcpyDeviceToHost, sc->stream));
        }
        if (*n_shadow)
            HIPCHECK(hipMemcpyAsync(shadow_ignore, si.p, (size_t)*n_shadow * 4, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

int lrt_render_aov(lrt_scene* sc, const lrt_render_params* prm, float* aov, int device_ptrs) {
    return guarded([&] {
        require(sc && prm && aov, "null pointer");
        require(prm->width >= 2 && prm->height >= 2 && prm->spp >= 1, "bad render parameters");
        const uint32_t W = prm->width, H = prm->height;
        const uint32_t rb = prm->row_end > prm->row_begin ? prm->row_begin : 0;
        const uint32_t re = prm->row_end > prm->row_begin ? prm->row_end : H;
        require(re <= H, "row range outside the image");
        HIPCHECK(hipSetDevice(sc->device));
        if (device_ptrs) after_caller(sc->stream, sc->caller);
        const uint32_t n = (re - rb) * W;
        const size_t bytes = (size_t)W * H * 7 * 4;
        float* d = aov;
        if (!device_ptrs) {
            sc->q_out0.alloc(bytes);
            d = sc->q_out0.as<float>();
            HIPCHECK(hipMemcpyAsync(d, aov, bytes, hipMemcpyHostToDevice, sc->stream));
        }
        const uint2 key = make_uint2((uint32_t)prm->seed, (uint32_t)(prm->seed >> 32));
        hipLaunchKernelGGL(k_aov, dim3(grid_for(*sc, n, 20)), dim3(kBlock), 0, sc->stream, sc->S, sc->M, sc->cam, W, H, rb, n,
                           prm->spp, prm->spp_offset, key, d);
        HIPCHECK(hipGetLastError());
        if (!device_ptrs) HIPCHECK(hipMemcpyAsync(aov, d, bytes, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

// ---- differentiable rendering --------------------------------------------------------
static uint32_t diff_count(const lrt_diff_params& dp) { return dp.tex_width * dp.tex_height * 3 + 3; }

// samples per thread and chunk count so a launch has ~2M threads
static void diff_chunks(const lrt_render_params& prm, uint32_t& chunk, uint32_t& nchunks) {
    const uint32_t npix = prm.width * prm.height;
    const uint32_t want = std::max<uint32_t>(1, std::min<uint32_t>(prm.spp, (1u << 21) / std::max<uint32_t>(npix, 1)));
    chunk = (prm.spp + want - 1) / want;
    nchunks = (prm.spp + chunk - 1) / chunk;
}

static void diff_launch_render(lrt_scene& sc, const Mats& M, const DiffParams& D, const lrt_render_params& prm,
                               float* img) {
    uint32_t chunk, nchunks;
    diff_chunks(prm, chunk, nchunks);
    const uint32_t npix = prm.width * prm.height;
    sc.diff_part.alloc((size_t)npix * nchunks * 12);
    hipLaunchKernelGGL(k_diff_render, dim3(grid_for(sc, (uint64_t)npix * nchunks, 16)), dim3(kBlock), 0, sc.stream, sc.S,
                       sc.Ls, M, D, sc.cam, chunk, nchunks, sc.diff_part.as<float>());
    hipLaunchKernelGGL(k_diff_resolve, dim3(grid_for(sc, npix, 8)), dim3(kBlock), 0, sc.stream, sc.diff_part.as<float>(),
                       npix, nchunks, prm.spp, img);
    HIPCHECK(hipGetLastError());
}

// The wavefront forward (img != null: mean radiance per pixel, W x H x 3,
// device) or adjoint (dimg != null: gradient of sum(dimg * image) added into
// the fixed-point sums gfix at scale *gscale, device) of the differentiable
// renderer, over passes of up to LRT_PASS_PATHS paths like lrt_render.
static void diff_wave_run(lrt_scene& sc, const Mats& M, const DiffParams& D, const lrt_render_params& prm, float* img,
                          const float* dimg, unsigned long long* gfix, const float* gscale, uint32_t n_params) {
    const uint32_t W = prm.width, H = prm.height, npix = W * H;
    const uint32_t slots = std::max<uint32_t>(1, std::min<uint32_t>(prm.spp, pass_paths() / npix));
    const uint32_t n = npix * slots;
    if (sc.cap < n) {
        sc.pstate.alloc((size_t)n * 64 + 256);
        sc.hits.alloc((size_t)n * 32 + 256);
        sc.shadow.alloc((size_t)n * 12 * 4 + 12 * 256);
        sc.queues.alloc((size_t)n * 3 * 4 + 3 * 256);
        sc.cap = n;
    }
    sc.counters.alloc(kCounterWords * 4);
    sc.stats.alloc(256);
    PathState P;
    char* c = sc.pstate.as<char>();
    P.rec = alloc_carve<float4>(c, 4 * (size_t)n);
    P.pixel0 = 0;
    P.npix = npix;
    HitBuf Hb;
    c = sc.hits.as<char>();
    Hb.rec = alloc_carve<uint4>(c, 2 * (size_t)n);
    ShadowBuf Sb;
    c = sc.shadow.as<char>();
    float** sf[10] = {&Sb.ox, &Sb.oy, &Sb.oz, &Sb.dx, &Sb.dy, &Sb.dz, &Sb.tmax, &Sb.cr, &Sb.cg, &Sb.cb};
    for (float** f : sf) *f = alloc_carve<float>(c, n);
    Sb.ignore = alloc_carve<int32_t>(c, n);
    Sb.path = alloc_carve<uint32_t>(c, n);
    c = sc.queues.as<char>();
    uint32_t* qa = alloc_carve<uint32_t>(c, n);
    uint32_t* qb = alloc_carve<uint32_t>(c, n);
    uint32_t* order = alloc_carve<uint32_t>(c, n);
    DiffPaths DP{};
    DiffShadow DS{};
    const bool adj = dimg != nullptr;
    if (adj) {
        sc.diff_paths.alloc((size_t)n * 14 * 4 + 14 * 256);
        c = sc.diff_paths.as<char>();
        float** df[12] = {&DP.ltr, &DP.ltg, &DP.ltb, &DP.rr, &DP.rg, &DP.rb, &DP.far, &DP.fag, &DP.fab, &DP.fer, &DP.feg, &DP.feb};
        for (float** f : df) *f = alloc_carve<float>(c, n);
        DP.ptex = alloc_carve<int32_t>(c, n);
        DP.pflag = alloc_carve<int32_t>(c, n);
        sc.diff_shadow.alloc((size_t)n * 7 * 4 + 7 * 256);
        c = sc.diff_shadow.as<char>();
        DS.tex = alloc_carve<int32_t>(c, n);
        float** sg[5] = {&DS.gr, &DS.gg, &DS.gb, &DS.ga, &DS.ge};
        for (float** f : sg) *f = alloc_carve<float>(c, n);
        DS.slot = alloc_carve<int32_t>(c, n);
    }
    uint32_t* counters = sc.counters.as<uint32_t>();
    unsigned long long* dstats = sc.stats.as<unsigned long long>();
    const uint32_t g_all = (n + kBlock - 1) / kBlock;
    const uint32_t g_trace = (uint32_t)std::min<uint64_t>(g_all, (uint64_t)sc.cus * 20);
    const uint32_t g_shade = (uint32_t)std::min<uint64_t>(g_all, (uint64_t)sc.cus * 8);  // 128 per CU: 4% slower
    const uint32_t g_bins = (uint32_t)std::min<uint64_t>(g_all, (uint64_t)sc.cus * 4);
    if (img) hipLaunchKernelGGL(k_zero_f32, dim3(grid_for(sc, (uint64_t)npix * 3, 4)), dim3(kBlock), 0, sc.stream, img,
                                (long long)npix * 3);
    unsigned long long* grad_s = adj ? gfix + (n_params - 3) : nullptr;
    const float inv_spp = 1.0f / (float)prm.spp;
    auto pass = [&](uint32_t smp, uint32_t np, bool replay) {
        hipLaunchKernelGGL(k_raygen, dim3((std::max(np, 1024u) + kBlock - 1) / kBlock), dim3(kBlock), 0, sc.stream, P, sc.cam, W, H, 0u, np,
                           smp, D.key, (const uint32_t*)nullptr, qa, counters);
        uint32_t* q = qa;
        uint32_t* qn = qb;
        for (uint32_t depth = 0; depth < prm.max_depth; ++depth) {
            hipLaunchKernelGGL(k_extend, dim3(g_trace), dim3(kBlock), 0, sc.stream, sc.S, P, q, counters, Hb);
            hipLaunchKernelGGL(k_shade_bins, dim3(g_bins), dim3(kBlock), 0, sc.stream, sc.S, Hb, counters);
            hipLaunchKernelGGL(k_shade_scan, dim3(1), dim3(1024), 0, sc.stream, counters);
            hipLaunchKernelGGL(k_shade_order, dim3(g_bins), dim3(kBlock), 0, sc.stream, sc.S, Hb, counters, order);
            const DiffWave dw{depth, prm.max_depth, prm.rr_depth, smp, D.key, replay ? dimg : nullptr, inv_spp, gscale};
            if (replay) {
                hipLaunchKernelGGL(k_diff_shade<true>, dim3(g_shade), dim3(kBlock), 0, sc.stream, sc.S, sc.Ls, M, D, P, q, qn,
                                   counters, Hb, order, Sb, DS, DP, dw, gfix, grad_s);
                hipLaunchKernelGGL(k_diff_shadow, dim3(g_trace), dim3(kBlock), 0, sc.stream, sc.S, P, counters, Sb, DS, gfix,
                                   grad_s, gscale);
            } else {
                hipLaunchKernelGGL(k_diff_shade<false>, dim3(g_shade), dim3(kBlock), 0, sc.stream, sc.S, sc.Ls, M, D, P, q,
                                   qn, counters, Hb, order, Sb, DS, DP, dw, gfix, grad_s);
                hipLaunchKernelGGL(k_shadow, dim3(g_trace), dim3(kBlock), 0, sc.stream, sc.S, P, counters, Sb);
            }
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, sc.stream, counters, dstats);
            std::swap(q, qn);
        }
        HIPCHECK(hipGetLastError());
    };
    for (uint32_t s = 0; s < prm.spp; s += slots) {
        const uint32_t smp = prm.spp_offset + s;
        const uint32_t ns = std::min(slots, prm.spp - s);
        const uint32_t np = npix * ns;
        pass(smp, np, false);
        if (img) {
            hipLaunchKernelGGL(k_diff_accum, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, sc.stream, P, ns, img);
        } else {
            hipLaunchKernelGGL(k_diff_keep, dim3(grid_for(sc, np, 8)), dim3(kBlock), 0, sc.stream, P, np, DP);
            pass(smp, np, true);
        }
    }
    if (img) hipLaunchKernelGGL(k_diff_scale, dim3(grid_for(sc, (uint64_t)npix * 3, 4)), dim3(kBlock), 0, sc.stream, img,
                                npix * 3, inv_spp);
    HIPCHECK(hipGetLastError());
}

// validates dp, applies theta's scalars to a copy of the materials and
// uploads it; returns the kernel parameters (texture pointer set by caller)
static DiffParams diff_setup(lrt_scene& sc, const lrt_diff_params& dp, const float* theta_host_scalars,
                             const lrt_render_params& prm) {
    const int nm = (int)sc.host_mats.size();
    require(prm.width >= 2 && prm.height >= 2 && prm.spp >= 1 && prm.max_depth >= 1, "bad render parameters");
    require(dp.tex_material < 0 || (dp.tex_material < nm && sc.host_mats[dp.tex_material].type == LRT_BSDF_DIFFUSE),
            "tex_material must be a diffuse material");
    require(dp.tex_material < 0 || (dp.tex_width >= 1 && dp.tex_height >= 1), "bad texture size");
    for (int k = 0; k < 2; ++k) {
        const int m = dp.alpha_material[k];
        require(m < 0 || (m < nm && (sc.host_mats[m].type == LRT_BSDF_CONDUCTOR || sc.host_mats[m].type == LRT_BSDF_ROUGH_DIELECTRIC)),
                "alpha_material must be a conductor or a rough dielectric");
    }
    require(dp.ior_material < 0 || (dp.ior_material < nm && (sc.host_mats[dp.ior_material].type == LRT_BSDF_DIELECTRIC ||
                                                             sc.host_mats[dp.ior_material].type == LRT_BSDF_ROUGH_DIELECTRIC)),
            "ior_material must be a (rough) dielectric");
    std::vector<lrt_material> mats = sc.host_mats;
    for (int k = 0; k < 2; ++k)
        if (dp.alpha_material[k] >= 0) mats[dp.alpha_material[k]].alpha = theta_host_scalars[k];
    if (dp.ior_material >= 0) mats[dp.ior_material].int_ior = theta_host_scalars[2];
    upload(sc.diff_mats, mats.data(), mats.size() * sizeof(lrt_material), sc.stream);
    DiffParams D;
    D.tex = nullptr;
    D.eval_mats = nullptr;
    D.tex_w = (int)dp.tex_width;
    D.tex_h = (int)dp.tex_height;
    D.tex_mat = dp.tex_material;
    D.alpha_mat[0] = dp.alpha_material[0];
    D.alpha_mat[1] = dp.alpha_material[1];
    D.ior_mat = dp.ior_material;
    D.max_depth = prm.max_depth;
    D.rr_depth = prm.rr_depth;
    D.W = prm.width;
    D.H = prm.height;
    D.spp = prm.spp;
    D.spp0 = prm.spp_offset;
    D.key = make_uint2((uint32_t)prm.seed, (uint32_t)(prm.seed >> 32));
    return D;
}

// theta (host or device) -> device copy and host scalars
static const float* diff_theta(lrt_scene& sc, const lrt_diff_params& dp, const float* theta, int device_ptrs, float s[3]) {
    const uint32_t n = diff_count(dp);
    if (device_ptrs) {
        after_caller(sc.stream, sc.caller);
        HIPCHECK(hipMemcpyAsync(s, theta + (n - 3), 12, hipMemcpyDeviceToHost, sc.stream));
        HIPCHECK(hipStreamSynchronize(sc.stream));
        return theta;
    }
    s[0] = theta[n - 3];
    s[1] = theta[n - 2];
    s[2] = theta[n - 1];
    upload(sc.diff_theta, theta, (size_t)n * 4, sc.stream);
    return sc.diff_theta.as<float>();
}

int lrt_diff_param_count(const lrt_diff_params* dp, uint32_t* n) {
    return guarded([&] {
        require(dp && n, "null pointer");
        *n = diff_count(*dp);
    });
}

int lrt_diff_render(lrt_scene* sc, const lrt_diff_params* dp, const float* theta, const lrt_render_params* prm,
                    float* image, int device_ptrs) {
    return guarded([&] {
        require(sc && dp && theta && prm && image, "null pointer");
        require(!sc->mesh_lights, "the differentiable renderer takes rectangle, sphere and sky lights (no mesh lights)",
                LRT_ERR_UNSUPPORTED);
        HIPCHECK(hipSetDevice(sc->device));
        float sv[3];
        const float* dth = diff_theta(*sc, *dp, theta, device_ptrs, sv);
        DiffParams D = diff_setup(*sc, *dp, sv, *prm);
        D.tex = dp->tex_material >= 0 ? dth : nullptr;
        const Mats M{sc->diff_mats.as<lrt_material>(), sc->M.tables};
        const size_t bytes = (size_t)prm->width * prm->height * 12;
        float* img = image;
        if (!device_ptrs) {
            sc->diff_img.alloc(bytes);
            img = sc->diff_img.as<float>();
        }
        if (prm->flags & LRT_DIFF_PER_SAMPLE) diff_launch_render(*sc, M, D, *prm, img);
        else diff_wave_run(*sc, M, D, *prm, img, nullptr, nullptr, nullptr, diff_count(*dp));
        if (!device_ptrs) HIPCHECK(hipMemcpyAsync(image, img, bytes, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

int lrt_diff_render_held(lrt_scene* sc, const lrt_diff_params* dp, const float* theta, const float* theta_sampling,
                         const lrt_render_params* prm, float* image) {
    return guarded([&] {
        require(sc && dp && theta && theta_sampling && prm && image, "null pointer");
        require(!sc->mesh_lights, "the differentiable renderer takes rectangle, sphere and sky lights (no mesh lights)",
                LRT_ERR_UNSUPPORTED);
        HIPCHECK(hipSetDevice(sc->device));
        const uint32_t np = diff_count(*dp);
        // BSDF values at theta (texture and scalars), sampling at theta_sampling's scalars
        float se[3] = {theta[np - 3], theta[np - 2], theta[np - 1]};
        diff_setup(*sc, *dp, se, *prm);  // materials at theta -> copied aside for the BSDF values
        DevBuf eval;
        eval.alloc(sc->host_mats.size() * sizeof(lrt_material));
        HIPCHECK(hipMemcpyAsync(eval.p, sc->diff_mats.p, sc->host_mats.size() * sizeof(lrt_material), hipMemcpyDeviceToDevice,
                                sc->stream));
        float ss[3] = {theta_sampling[np - 3], theta_sampling[np - 2], theta_sampling[np - 1]};
        DiffParams D = diff_setup(*sc, *dp, ss, *prm);
        upload(sc->diff_theta, theta, (size_t)np * 4, sc->stream);
        D.tex = dp->tex_material >= 0 ? sc->diff_theta.as<float>() : nullptr;
        D.eval_mats = eval.as<lrt_material>();
        const Mats M{sc->diff_mats.as<lrt_material>(), sc->M.tables};
        const size_t bytes = (size_t)prm->width * prm->height * 12;
        sc->diff_img.alloc(bytes);
        diff_launch_render(*sc, M, D, *prm, sc->diff_img.as<float>());
        HIPCHECK(hipMemcpyAsync(image, sc->diff_img.p, bytes, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

int lrt_diff_adjoint(lrt_scene* sc, const lrt_diff_params* dp, const float* theta, const lrt_render_params* prm,
                     const float* d_image, float* grad, int device_ptrs) {
    return guarded([&] {
        require(sc && dp && theta && prm && d_image && grad, "null pointer");
        require(!sc->mesh_lights, "the differentiable renderer takes rectangle, sphere and sky lights (no mesh lights)",
                LRT_ERR_UNSUPPORTED);
        HIPCHECK(hipSetDevice(sc->device));
        float sv[3];
        const float* dth = diff_theta(*sc, *dp, theta, device_ptrs, sv);
        DiffParams D = diff_setup(*sc, *dp, sv, *prm);
        D.tex = dp->tex_material >= 0 ? dth : nullptr;
        const Mats M{sc->diff_mats.as<lrt_material>(), sc->M.tables};
        const uint32_t np = diff_count(*dp);
        const size_t bytes = (size_t)prm->width * prm->height * 12;
        const float* dimg = d_image;
        float* g = grad;
        if (!device_ptrs) {
            upload(sc->diff_dimg, d_image, bytes, sc->stream);
            dimg = sc->diff_dimg.as<float>();
            sc->diff_grad.alloc((size_t)np * 4);
            g = sc->diff_grad.as<float>();
        }
        // fixed-point sums (np words) and their scale
        sc->diff_fix.alloc((size_t)np * 8 + 256);
        unsigned long long* gfix = sc->diff_fix.as<unsigned long long>();
        float* gscale = reinterpret_cast<float*>(gfix + np);
        hipLaunchKernelGGL(k_zero_f32, dim3(grid_for(*sc, 2 * (uint64_t)np, 4)), dim3(kBlock), 0, sc->stream,
                           reinterpret_cast<float*>(gfix), 2 * (long long)np);
        hipLaunchKernelGGL(k_grad_scale, dim3(1), dim3(1024), 0, sc->stream, dimg, prm->width * prm->height * 3,
                           1.0f / (float)prm->spp, gscale);
        if (prm->flags & LRT_DIFF_PER_SAMPLE) {
            uint32_t chunk, nchunks;
            diff_chunks(*prm, chunk, nchunks);
            const uint64_t n = (uint64_t)prm->width * prm->height * nchunks;
            hipLaunchKernelGGL(k_diff_adjoint, dim3(grid_for(*sc, n, 16)), dim3(kBlock), 0, sc->stream, sc->S, sc->Ls, M, D,
                               sc->cam, chunk, nchunks, dimg, gfix, gfix + (np - 3), gscale);
        } else {
            diff_wave_run(*sc, M, D, *prm, nullptr, dimg, gfix, gscale, np);
        }
        hipLaunchKernelGGL(k_grad_out, dim3(grid_for(*sc, np, 4)), dim3(kBlock), 0, sc->stream, gfix, np, gscale, g);
        HIPCHECK(hipGetLastError());
        if (!device_ptrs) HIPCHECK(hipMemcpyAsync(grad, g, (size_t)np * 4, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

// ---- photon mapping -----------------------------------------------------------------
static void radix_sort(lrt_scene& sc, DevBuf& hbuf, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n,
                       int bits, uint32_t*& out_k, uint32_t*& out_v) {
    const uint32_t nb = std::max<uint32_t>(1, (n + kRsTile - 1) / kRsTile);
    hbuf.alloc((size_t)nb * 16 * 4);
    uint32_t* hist = hbuf.as<uint32_t>();
    for (int shift = 0; shift < bits; shift += 4) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kRsBlock), 0, sc.stream, k0, n, shift, hist, nb);
        hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, sc.stream, hist, nb * 16);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kRsBlock), 0, sc.stream, k0, v0, k1, v1, n, shift, hist, nb);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    HIPCHECK(hipGetLastError());
    out_k = k0;
    out_v = v0;
}

// nodes of a median subtree over c photons (leaves of at most 8)
static uint64_t pk_subtree_nodes(uint64_t c, std::map<uint64_t, uint64_t>& memo) {
    if (c <= 8) return 1;
    auto it = memo.find(c);
    if (it != memo.end()) return it->second;
    const uint64_t v = 1 + pk_subtree_nodes(c / 2, memo) + pk_subtree_nodes(c - c / 2, memo);
    memo[c] = v;
    return v;
}

// The device median tree over the n photon positions xyz (particle order):
// S.nodes (pre-order), S.order (tree position -> particle-order index).
// Level l holds 2^l segments of floor(n / 2^l) or that + 1 photons; levels
// whose smallest segment has 9 or more photons split every segment, the first
// level below that is finished per segment (leaves, or a 9 into 4 + 5).
// Returns the depth (nodes on the longest root-to-leaf path).
static int build_photon_tree_device(lrt_scene& sc, PhotonState& S, uint32_t n) {
    std::map<uint64_t, uint64_t> memo;
    const uint64_t n_nodes = pk_subtree_nodes(n, memo);
    int L = 0;  // levels that split every segment
    while ((n >> L) >= 9u) ++L;
    const uint32_t m_max = 1u << L;
    const size_t tiles = (3 * (size_t)n + kPkTile - 1) / kPkTile;
    const size_t words = 3 * (size_t)n * 3 + 2 * (size_t)n + 2 * 4 * (size_t)m_max + tiles + (size_t)n / 4 + 64;
    S.tree_work.alloc(words * 4 + 16 * 256);
    char* c = S.tree_work.as<char>();
    uint32_t* lists = alloc_carve<uint32_t>(c, 3 * (size_t)n);
    uint32_t* lists2 = alloc_carve<uint32_t>(c, 3 * (size_t)n);
    uint32_t* prefix = alloc_carve<uint32_t>(c, 3 * (size_t)n);
    uint32_t* seg = alloc_carve<uint32_t>(c, n);
    uint32_t* seg2 = alloc_carve<uint32_t>(c, n);
    PkLevel lv[2];
    for (PkLevel& l : lv) {
        l.start = alloc_carve<uint32_t>(c, m_max);
        l.count = alloc_carve<uint32_t>(c, m_max);
        l.pre = alloc_carve<uint32_t>(c, m_max);
        l.axis = alloc_carve<uint32_t>(c, m_max);
    }
    uint32_t* tile_sum = alloc_carve<uint32_t>(c, tiles);
    uint8_t* side = alloc_carve<uint8_t>(c, n);
    S.nodes.alloc(n_nodes * 16);
    S.order.alloc((size_t)n * 4);
    const float* xyz = S.xyz.as<float>();
    const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, (uint32_t)sc.cus * 16));
    // lists sorted by (coordinate key, index) per axis: stable radix sort of the keys
    uint32_t* kbuf = prefix;  // scratch for the sorts (3n words: keys, and the sort's second buffers)
    uint32_t* kb2 = lists2;
    for (int a = 0; a < 3; ++a) {
        uint32_t* keys = kbuf;
        uint32_t* ids = kbuf + n;
        hipLaunchKernelGGL(k_pk_keys, dim3(g), dim3(256), 0, sc.stream, xyz, n, a, keys, ids);
        uint32_t *ok, *ov;
        radix_sort(sc, S.hist, keys, ids, kb2, kb2 + n, n, 32, ok, ov);
        HIPCHECK(hipMemcpyAsync(lists + (size_t)a * n, ov, (size_t)n * 4, hipMemcpyDeviceToDevice, sc.stream));
    }
    // level 0: the root segment
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, sc.stream, seg, n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sc.stream, lv[0].start, 1u, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sc.stream, lv[0].count, 1u, n);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sc.stream, lv[0].pre, 1u, 0u);
    uint4* nodes = S.nodes.as<uint4>();
    const uint32_t g3 = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)((3 * (size_t)n + 255) / 256), (uint32_t)sc.cus * 16));
    for (int l = 0; l < L; ++l) {
        const uint32_t m = 1u << l;
        const uint32_t a_next = n >> (l + 1);
        const uint32_t s_lo = (uint32_t)pk_subtree_nodes(a_next, memo), s_hi = (uint32_t)pk_subtree_nodes(a_next + 1, memo);
        PkLevel& cur = lv[l & 1];
        PkLevel& nxt = lv[(l + 1) & 1];
        hipLaunchKernelGGL(k_pk_choose, dim3((m + 255) / 256), dim3(256), 0, sc.stream, xyz, lists, n, cur, m, nxt,
                           a_next, s_lo, s_hi, nodes);
        hipLaunchKernelGGL(k_pk_side, dim3(g), dim3(256), 0, sc.stream, lists, n, seg, cur, side);
        hipLaunchKernelGGL(k_pk_scan_tiles, dim3((uint32_t)tiles), dim3(kPkScanBlock), 0, sc.stream, lists, side,
                           3 * (size_t)n, tile_sum);
        hipLaunchKernelGGL(k_pk_scan_sums, dim3(1), dim3(1024), 0, sc.stream, tile_sum, (uint32_t)tiles);
        hipLaunchKernelGGL(k_pk_scan_apply, dim3((uint32_t)tiles), dim3(kPkScanBlock), 0, sc.stream, lists, side,
                           3 * (size_t)n, tile_sum, prefix);
        hipLaunchKernelGGL(k_pk_scatter, dim3(g3), dim3(256), 0, sc.stream, lists, n, seg, cur, side, prefix, lists2,
                           seg2);
        std::swap(lists, lists2);
        std::swap(seg, seg2);
    }
    const uint32_t m = 1u << L;
    hipLaunchKernelGGL(k_pk_finish, dim3((m + 255) / 256), dim3(256), 0, sc.stream, xyz, lists, n, lv[L & 1], m, nodes,
                       S.order.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    S.n_nodes = n_nodes;
    const uint64_t deepest = ((uint64_t)n + m - 1) / m;  // the right-most segment of the last level
    return L + 1 + (deepest > 8 ? 1 : 0);
}

// raw photons (S.buf, count n_raw) -> kept in particle order -> kd tree (on
// the device, or the reference's tree built on the host) -> photons packed in
// tree order
static void build_photon_map(lrt_scene& sc, PhotonState& S, uint32_t n_raw, uint32_t n_keep, float scale, float radius,
                             uint32_t k, uint32_t tree) {
    const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((n_raw + 255) / 256, (uint32_t)sc.cus * 16));
    S.sort.alloc((size_t)std::max<uint32_t>(n_raw, 1) * 4 * 4 + 4 * 256);
    char* c = S.sort.as<char>();
    uint32_t* ka = alloc_carve<uint32_t>(c, std::max<uint32_t>(n_raw, 1));
    uint32_t* va = alloc_carve<uint32_t>(c, std::max<uint32_t>(n_raw, 1));
    uint32_t* kb = alloc_carve<uint32_t>(c, std::max<uint32_t>(n_raw, 1));
    uint32_t* vb = alloc_carve<uint32_t>(c, std::max<uint32_t>(n_raw, 1));
    S.valid = false;
    // 1. particle order: sort by particle << 8 | depth, keep the first n_keep
    uint32_t *sk = ka, *sv = va;
    if (n_raw) {
        HIPCHECK(hipMemcpyAsync(ka, S.buf.key, (size_t)n_raw * 4, hipMemcpyDeviceToDevice, sc.stream));
        hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, sc.stream, va, n_raw);
        radix_sort(sc, S.hist, ka, va, kb, vb, n_raw, 32, sk, sv);
    }
    S.perm.alloc((size_t)std::max<uint32_t>(n_keep, 1) * 4);
    if (n_keep) HIPCHECK(hipMemcpyAsync(S.perm.p, sv, (size_t)n_keep * 4, hipMemcpyDeviceToDevice, sc.stream));
    // 2. the kd tree over the kept positions, in particle order
    S.xyz.alloc((size_t)std::max<uint32_t>(n_keep, 1) * 12);
    if (n_keep)
        hipLaunchKernelGGL(k_photon_positions, dim3(g), dim3(256), 0, sc.stream, S.buf, S.perm.as<uint32_t>(), n_keep,
                           S.xyz.as<float>());
    HIPCHECK(hipGetLastError());
    if (tree == LRT_PHOTON_TREE_DEVICE) {
        if (n_keep) require(build_photon_tree_device(sc, S, n_keep) < kPkStack, "photon kd tree deeper than the search stack");
    } else {
        std::vector<float> xyz((size_t)n_keep * 3);
        if (n_keep) HIPCHECK(hipMemcpyAsync(xyz.data(), S.xyz.p, (size_t)n_keep * 12, hipMemcpyDeviceToHost, sc.stream));
        HIPCHECK(hipStreamSynchronize(sc.stream));
        PhotonKd kd;
        photon_kd_build(xyz.data(), n_keep, kd);
        require(kd.depth < kPkStack, "photon kd tree deeper than the search stack");
        for (size_t i = 0; i < kd.nodes.size(); i += 4)  // the device search loads a leaf's photons 8 at a time
            require(kd.nodes[i + 1] != 3u || kd.nodes[i + 3] <= 8u, "photon kd leaf with more than 8 photons");
        S.nodes.alloc(std::max<size_t>(kd.nodes.size(), 4) * 4);
        S.order.alloc((size_t)std::max<uint32_t>(n_keep, 1) * 4);
        if (n_keep) {
            HIPCHECK(hipMemcpyAsync(S.nodes.p, kd.nodes.data(), kd.nodes.size() * 4, hipMemcpyHostToDevice, sc.stream));
            HIPCHECK(hipMemcpyAsync(S.order.p, kd.order.data(), (size_t)n_keep * 4, hipMemcpyHostToDevice, sc.stream));
        }
        S.n_nodes = kd.nodes.size() / 4;
        HIPCHECK(hipStreamSynchronize(sc.stream));  // the host vectors are released on return
    }
    // 3. photons in tree order
    S.pos.alloc((size_t)std::max<uint32_t>(n_keep, 1) * 16);
    S.pay.alloc((size_t)std::max<uint32_t>(n_keep, 1) * 36);
    if (n_keep) {
        hipLaunchKernelGGL(k_photon_pack, dim3(g), dim3(256), 0, sc.stream, S.buf, S.perm.as<uint32_t>(),
                           S.order.as<uint32_t>(), n_keep, S.pos.as<float4>(), S.pay.as<float>());
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(sc.stream));
    S.map.pos = S.pos.as<float4>();
    S.map.pay = S.pay.as<float>();
    S.map.nodes = S.nodes.as<uint4>();
    S.map.n = n_keep;
    S.map.scale = scale;
    S.map.radius = radius;
    S.map.k = (int32_t)k;
    S.n = n_keep;
    S.valid = true;
}

static void alloc_photons(PhotonState& S, uint32_t capacity) {
    S.raw.alloc((size_t)capacity * 13 * 4 + 13 * 256 + 256);
    char* c = S.raw.as<char>();
    float** f[12] = {&S.buf.px, &S.buf.py, &S.buf.pz, &S.buf.dx, &S.buf.dy, &S.buf.dz,
                     &S.buf.nx, &S.buf.ny, &S.buf.nz, &S.buf.wr, &S.buf.wg, &S.buf.wb};
    for (float** p : f) *p = alloc_carve<float>(c, capacity);
    S.buf.key = alloc_carve<uint32_t>(c, capacity);
    S.buf.count = alloc_carve<uint32_t>(c, 64);
    S.buf.capacity = capacity;
}

static PhotonState& photon_state(lrt_scene* sc, int32_t kind) {
    require(kind == 0 || kind == 1, "kind must be 0 (global map) or 1 (caustic map)");
    return sc->pm[kind];
}

int lrt_photon_build(lrt_scene* sc, const lrt_photon_params* prm, lrt_photon_stats* stats) {
    return guarded([&] {
        require(sc && prm, "null pointer");
        require(prm->k >= 1 && prm->k <= (uint32_t)kMaxK, "k must be in [1, 128]");
        require(prm->radius > 0.0f && prm->max_depth >= 1, "bad photon parameters");
        require(prm->max_depth < 256, "photon depth is stored in 8 bits: max_depth must be < 256");
        require(prm->tree == LRT_PHOTON_TREE_DEVICE || prm->tree == LRT_PHOTON_TREE_REFERENCE, "unknown photon tree");
        PhotonState& S = photon_state(sc, prm->kind);
        HIPCHECK(hipSetDevice(sc->device));
        const auto t0 = std::chrono::steady_clock::now();
        S.valid = false;
        S.n = 0;
        S.shot = 0;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (prm->requested == 0) return;
        uint32_t capacity = prm->requested + prm->requested / 2 + 65536;
        alloc_photons(S, capacity);
        std::vector<uint32_t> per;  // photons per particle, particle order
        uint64_t returned = 0;
        uint32_t traced = 0;
        double trace_ms = 0.0;
        bool failed = false;
        const uint2 key = make_uint2((uint32_t)prm->seed, (uint32_t)(prm->seed >> 32));
        HIPCHECK(hipMemsetAsync(S.buf.count, 0, 4, sc->stream));
        uint32_t batch = std::max<uint32_t>(65536, prm->requested / 4);
        for (int iter = 0; iter < 256 && returned < prm->requested; ++iter) {
            // photon_gather.h:99-101: stop when far too few photons are stored
            if (traced > 100000 && (returned == 0 || returned < traced / 1024)) { failed = true; break; }
            require((uint64_t)traced + batch < (1ull << 24), "photon tracing needs more than 2^24 particles");
            S.per.alloc((size_t)batch * 4);
            hipEvent_t a, b;
            HIPCHECK(hipEventCreate(&a));
            HIPCHECK(hipEventCreate(&b));
            HIPCHECK(hipEventRecord(a, sc->stream));
            PhotonTraceParams tp{traced, batch, prm->max_depth, prm->rr_depth, (uint32_t)prm->kind, key};
            const uint32_t g = std::min<uint32_t>((batch + 255) / 256, (uint32_t)sc->cus * 8);
            hipLaunchKernelGGL(k_photon_trace, dim3(g), dim3(256), 0, sc->stream, sc->S, sc->Ls, sc->M, sc->emit, S.buf,
                               S.per.as<uint32_t>(), tp);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipEventRecord(b, sc->stream));
            const size_t off = per.size();
            per.resize(off + batch);
            HIPCHECK(hipMemcpyAsync(per.data() + off, S.per.p, (size_t)batch * 4, hipMemcpyDeviceToHost, sc->stream));
            uint32_t total = 0;
            HIPCHECK(hipMemcpyAsync(&total, S.buf.count, 4, hipMemcpyDeviceToHost, sc->stream));
            HIPCHECK(hipStreamSynchronize(sc->stream));
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, a, b));
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            trace_ms += ms;
            traced += batch;
            returned = total;
            if (total > capacity) {  // grow and trace everything again
                capacity = total + total / 4 + 65536;
                alloc_photons(S, capacity);
                HIPCHECK(hipMemsetAsync(S.buf.count, 0, 4, sc->stream));
                per.clear();
                traced = 0;
                returned = 0;
                continue;
            }
            if (returned < prm->requested) {
                const double rate = returned > 0 ? (double)returned / traced : 0.01;
                batch = (uint32_t)std::min<double>(1u << 22, std::max(4096.0, (prm->requested - returned) / rate * 1.1 + 1024));
            }
        }
        // shot: particles up to the one whose photons first reach `requested`
        uint64_t cum = 0, shot = traced;
        for (uint32_t i = 0; i < traced; ++i) {
            cum += per[i];
            if (cum >= prm->requested) { shot = (uint64_t)i + 1; break; }
        }
        const bool full = !failed && returned >= prm->requested;
        const auto t1 = std::chrono::steady_clock::now();
        if (full) {
            build_photon_map(*sc, S, (uint32_t)returned, prm->requested, 1.0f / (float)shot, prm->radius, prm->k,
                             prm->tree);
            S.shot = shot;
        }
        const auto t2 = std::chrono::steady_clock::now();
        if (stats) {
            stats->photons = full ? prm->requested : std::min<uint64_t>(returned, prm->requested);
            stats->shot = shot;
            stats->particles_traced = traced;
            stats->returned = returned;
            stats->full = full ? 1 : 0;
            stats->trace_ms = trace_ms;
            stats->sort_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
            stats->seconds = std::chrono::duration<double>(t2 - t0).count();
        }
    });
}

int lrt_photon_map_load(lrt_scene* sc, int32_t kind, const float* photons12, const uint32_t* depth_particle, uint64_t n,
                        float scale, float radius, uint32_t k) {
    return guarded([&] {
        require(sc && (n == 0 || (photons12 && depth_particle)), "null pointer");
        require(k >= 1 && k <= (uint32_t)kMaxK && radius > 0.0f, "bad map parameters");
        require(n < (1ull << 24), "photon map load: more than 2^24 photons");
        PhotonState& S = photon_state(sc, kind);
        HIPCHECK(hipSetDevice(sc->device));
        S.valid = false;
        S.n = 0;
        if (n == 0) return;
        alloc_photons(S, (uint32_t)n);
        std::vector<float> col((size_t)n);
        float* dst[12] = {S.buf.px, S.buf.py, S.buf.pz, S.buf.dx, S.buf.dy, S.buf.dz,
                          S.buf.nx, S.buf.ny, S.buf.nz, S.buf.wr, S.buf.wg, S.buf.wb};
        for (int j = 0; j < 12; ++j) {
            for (uint64_t i = 0; i < n; ++i) col[i] = photons12[12 * i + j];
            HIPCHECK(hipMemcpy(dst[j], col.data(), n * 4, hipMemcpyHostToDevice));
        }
        // keys keep the given order: position << 8 | depth
        std::vector<uint32_t> keys((size_t)n);
        for (uint64_t i = 0; i < n; ++i) {
            require(depth_particle[2 * i] < 256, "photon depth must be < 256");
            keys[i] = ((uint32_t)i << 8) | depth_particle[2 * i];
        }
        HIPCHECK(hipMemcpy(S.buf.key, keys.data(), n * 4, hipMemcpyHostToDevice));
        build_photon_map(*sc, S, (uint32_t)n, (uint32_t)n, scale, radius, k, LRT_PHOTON_TREE_REFERENCE);
        S.shot = (uint64_t)std::llround(1.0 / scale);
    });
}

int lrt_photon_map_tree(lrt_scene* sc, int32_t kind, uint32_t* nodes4, uint64_t node_capacity, uint32_t* order,
                        uint64_t order_capacity, uint64_t* n_nodes, uint64_t* n_photons) {
    return guarded([&] {
        require(sc && n_nodes && n_photons, "null pointer");
        PhotonState& S = photon_state(sc, kind);
        const bool have = S.valid && S.n > 0;
        *n_nodes = have ? S.n_nodes : 0;
        *n_photons = have ? S.n : 0;
        if (!have || (!nodes4 && !order)) return;
        require(nodes4 && order && node_capacity >= S.n_nodes && order_capacity >= S.n, "buffers too small");
        HIPCHECK(hipSetDevice(sc->device));
        HIPCHECK(hipMemcpyAsync(nodes4, S.nodes.p, S.n_nodes * 16, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipMemcpyAsync(order, S.order.p, S.n * 4, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

int lrt_photon_estimate(lrt_scene* sc, int32_t kind, const float* queries6, uint32_t n, int max_depth, float* out3) {
    return guarded([&] {
        require(sc && queries6 && out3, "null pointer");
        PhotonState& S = photon_state(sc, kind);
        if (n == 0) return;
        if (!S.valid || S.n == 0) {  // photon_map::estimate_irradiance on an empty map
            std::memset(out3, 0, (size_t)n * 12);
            return;
        }
        HIPCHECK(hipSetDevice(sc->device));
        upload(sc->q_rays, queries6, (size_t)n * 24, sc->stream);
        sc->q_out0.alloc((size_t)n * 12);
        const uint32_t per_block = kGatherBlock / kGroup;
        const uint32_t g = std::min<uint32_t>((n + per_block - 1) / per_block, (uint32_t)sc->cus * 32);
        hipLaunchKernelGGL(gather_slots(S.map.k) == 8 ? k_photon_estimate<8> : k_photon_estimate<16>, dim3(g),
                           dim3(kGatherBlock), 0, sc->stream, S.map, sc->q_rays.as<float>(), n,
                           max_depth, sc->q_out0.as<float>());
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(out3, sc->q_out0.p, (size_t)n * 12, hipMemcpyDeviceToHost, sc->stream));
        HIPCHECK(hipStreamSynchronize(sc->stream));
    });
}

int lrt_photon_export(lrt_scene* sc, int32_t kind, float* photons12, uint32_t* depth_particle, uint64_t capacity,
                      uint64_t* n, uint64_t* shot) {
    return guarded([&] {
        require(sc != nullptr, "null pointer");
        PhotonState& S = photon_state(sc, kind);
        const uint64_t count = S.valid ? S.n : 0;
        if (n) *n = count;
        if (shot) *shot = S.valid ? S.shot : 0;
        if (!photons12 || count == 0) return;
        require(depth_particle != nullptr, "null pointer");
        HIPCHECK(hipSetDevice(sc->device));
        const uint64_t m = std::min<uint64_t>(capacity, count);
        std::vector<uint32_t> perm((size_t)m);
        HIPCHECK(hipMemcpy(perm.data(), S.perm.p, m * 4, hipMemcpyDeviceToHost));
        const uint32_t cap = S.buf.capacity;
        std::vector<float> col(cap);
        float* src[12] = {S.buf.px, S.buf.py, S.buf.pz, S.buf.dx, S.buf.dy, S.buf.dz,
                          S.buf.nx, S.buf.ny, S.buf.nz, S.buf.wr, S.buf.wg, S.buf.wb};
        for (int j = 0; j < 12; ++j) {
            HIPCHECK(hipMemcpy(col.data(), src[j], (size_t)cap * 4, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < m; ++i) photons12[12 * i + j] = col[perm[i]];
        }
        std::vector<uint32_t> keys(cap);
        HIPCHECK(hipMemcpy(keys.data(), S.buf.key, (size_t)cap * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < m; ++i) {
            depth_particle[2 * i] = keys[perm[i]] & 0xffu;
            depth_particle[2 * i + 1] = keys[perm[i]] >> 8;
        }
    });
}

// ---- photon-map eye pass ----------------------------------------------------------------
static void carve_items(DevBuf& b, size_t cap, EyeItems& It) {
    b.alloc(cap * 16 * 4 + 16 * 256);
    char* c = b.as<char>();
    It.sid = alloc_carve<uint32_t>(c, cap);
    It.path = alloc_carve<uint32_t>(c, cap);
    It.meta = alloc_carve<uint32_t>(c, cap);
    float** f[13] = {&It.ox, &It.oy, &It.oz, &It.dx, &It.dy, &It.dz, &It.tr, &It.tg, &It.tb, &It.wr, &It.wg, &It.wb, &It.pdf};
    for (float** p : f) *p = alloc_carve<float>(c, cap);
}

static void carve_hits(DevBuf& b, size_t cap, HitBuf& Hb) {
    b.alloc(cap * 32 + 256);
    char* c = b.as<char>();
    Hb.rec = alloc_carve<uint4>(c, 2 * (size_t)cap);
}

static void carve_shadow(DevBuf& b, size_t cap, EyeShadow& Sh) {
    b.alloc(cap * 12 * 4 + 12 * 256);
    char* c = b.as<char>();
    Sh.sid = alloc_carve<uint32_t>(c, cap);
    float** f[10] = {&Sh.ox, &Sh.oy, &Sh.oz, &Sh.dx, &Sh.dy, &Sh.dz, &Sh.tmax, &Sh.cr, &Sh.cg, &Sh.cb};
    for (float** p : f) *p = alloc_carve<float>(c, cap);
    Sh.ignore = alloc_carve<int32_t>(c, cap);
}

static void carve_query(DevBuf& b, size_t cap, EyeQuery& Q) {
    b.alloc(cap * 14 * 4 + 14 * 256);
    char* c = b.as<char>();
    Q.sid = alloc_carve<uint32_t>(c, cap);
    Q.meta = alloc_carve<uint32_t>(c, cap);
    float** f[12] = {&Q.px, &Q.py, &Q.pz, &Q.nx, &Q.ny, &Q.nz, &Q.tr, &Q.tg, &Q.tb, &Q.wr, &Q.wg, &Q.wb};
    for (float** p : f) *p = alloc_carve<float>(c, cap);
}

// grows a buffer by half again when it must be reallocated
static size_t grow_cap(size_t have, size_t need) { return need <= have ? have : std::max(need, have + have / 2); }

static void render_photon_impl(lrt_scene* sc, const lrt_render_params* prm, double* sum_rgb, double* sumsq_rgb,
                               int device_ptrs, lrt_render_stats* stats) {
    require(sc && prm && sum_rgb, "null pointer");
    require(prm->width >= 2 && prm->height >= 2 && prm->max_depth >= 1 && prm->max_depth < 65535, "bad render parameters");
    const uint32_t W = prm->width, H = prm->height;
    const uint32_t rb = prm->row_end > prm->row_begin ? prm->row_begin : 0;
    const uint32_t re = prm->row_end > prm->row_begin ? prm->row_end : H;
    require(re <= H, "row range outside the image");
    require((uint64_t)W * H < (1ull << 31), "image too large");
    HIPCHECK(hipSetDevice(sc->device));
    if (device_ptrs) after_caller(sc->stream, sc->caller);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t img = (size_t)W * H * 3;
    double* dsum = sum_rgb;
    double* dsq = sumsq_rgb;
    if (!device_ptrs) {
        sc->accum.alloc(img * 8);
        HIPCHECK(hipMemcpyAsync(sc->accum.p, sum_rgb, img * 8, hipMemcpyHostToDevice, sc->stream));
        dsum = sc->accum.as<double>();
        if (sumsq_rgb) {
            sc->accum2.alloc(img * 8);
            HIPCHECK(hipMemcpyAsync(sc->accum2.p, sumsq_rgb, img * 8, hipMemcpyHostToDevice, sc->stream));
            dsq = sc->accum2.as<double>();
        }
    }
    const int max_spec = sc->max_spec_components;
    const uint32_t maps = (sc->pm[0].valid && sc->pm[0].n ? 1u : 0u) | (sc->pm[1].valid && sc->pm[1].n ? 2u : 0u);
    PhotonMapDev G0 = sc->pm[0].map, G1 = sc->pm[1].map;
    if (!(maps & 1u)) std::memset(&G0, 0, sizeof(G0));
    if (!(maps & 2u)) std::memset(&G1, 0, sizeof(G1));
    const bool slots8 = gather_slots(std::max(G0.k, G1.k)) == 8;
    const bool counting = (prm->flags & LRT_RENDER_COUNT) != 0;
    if (counting) {
        sc->totals.alloc(256);
        HIPCHECK(hipMemsetAsync(sc->totals.p, 0, 32, sc->stream));
    }
    const uint64_t pixels = (uint64_t)(re - rb) * W;
    // a chunk: up to eye_chunk() camera samples, several samples of every
    // pixel when the band has fewer pixels
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(pixels, eye_chunk());
    const uint32_t eslots = std::max<uint32_t>(1, std::min<uint32_t>(prm->spp, (uint32_t)(eye_chunk() / chunk)));
    sc->counters.alloc(256);
    uint32_t* counters = sc->counters.as<uint32_t>();
    sc->eye_acc.alloc((size_t)chunk * eslots * 3 * 8);
    unsigned long long* acc = sc->eye_acc.as<unsigned long long>();
    const uint2 key = make_uint2((uint32_t)prm->seed, (uint32_t)(prm->seed >> 32));
    const f3 box_lo{sc->box_lo[0], sc->box_lo[1], sc->box_lo[2]};
    auto cells = [&](int a) {  // Morton cells per unit length along axis a
        const float ext = sc->box_hi[a] - sc->box_lo[a];
        return ext > 0.0f ? 512.0f / ext : 0.0f;
    };
    const f3 box_scale{cells(0), cells(1), cells(2)};
    uint64_t n_ext = 0, n_shadow = 0, n_query = 0;
    Timer& tc = sc->timers[LRT_KERNEL_CLOSEST];
    Timer& to = sc->timers[LRT_KERNEL_OCCLUDED];
    Timer& ts = sc->timers[LRT_KERNEL_SHADE];
    Timer& tg = sc->timers[LRT_KERNEL_GATHER];
    HIPCHECK(hipStreamSynchronize(sc->stream));
    tc.collect();
    to.collect();
    ts.collect();
    tg.collect();
    const double before_c = tc.total_ms, before_o = to.total_ms, before_s = ts.total_ms;
    const uint64_t lc0 = tc.launches, lo0 = to.launches;
    for (uint32_t s = 0; s < prm->spp; s += eslots) {
        const uint32_t ns = std::min(eslots, prm->spp - s);
        for (uint64_t c0 = 0; c0 < pixels; c0 += chunk) {
            const uint32_t nc = (uint32_t)std::min<uint64_t>(chunk, pixels - c0);
            const uint32_t n0 = nc * ns;  // camera samples of this chunk
            EyeParams ep{prm->max_depth, prm->spp_offset + s, (uint32_t)((uint64_t)rb * W + c0), key, maps, 0, 0, 0, nc};
            int cur = 0;
            sc->eye_cap[cur] = grow_cap(sc->eye_cap[cur], n0);
            EyeItems A, B;
            carve_items(sc->eye_items[cur], sc->eye_cap[cur], A);
            hipLaunchKernelGGL(k_fill_u64, dim3(std::min<uint32_t>((3 * n0 + 255) / 256, 4096)), dim3(256), 0, sc->stream,
                               acc, (size_t)3 * n0);
            hipLaunchKernelGGL(k_eye_raygen, dim3((n0 + kBlock - 1) / kBlock), dim3(kBlock), 0, sc->stream, sc->cam, W, H,
                               n0, ep, A);
            HIPCHECK(hipGetLastError());
            uint32_t n = n0;
            for (uint32_t depth = 1; n > 0; ++depth) {
                // children bound: 32 BSDF samples at depth 1, else 1, plus the
                // specular components while depth < 4
                const size_t per = (depth == 1 ? 32 : 1) + (depth < 4 ? (size_t)max_spec : 0);
                const size_t need_children = (size_t)n * per;
                const size_t need_shadow = (size_t)n * (depth == 1 ? 16 : 1);
                const int nxt = cur ^ 1;
                // buffers only grow, and only after the kernels that read them are done
                if (need_children > sc->eye_cap[nxt] || n > sc->eye_hit_cap || need_shadow > sc->eye_shadow_cap ||
                    (size_t)n * 2 > sc->eye_query_cap)
                    HIPCHECK(hipStreamSynchronize(sc->stream));
                sc->eye_cap[nxt] = grow_cap(sc->eye_cap[nxt], need_children);
                carve_items(sc->eye_items[nxt], sc->eye_cap[nxt], B);
                carve_items(sc->eye_items[cur], sc->eye_cap[cur], A);
                sc->eye_hit_cap = grow_cap(sc->eye_hit_cap, n);
                HitBuf Hb;
                carve_hits(sc->eye_hits, sc->eye_hit_cap, Hb);
                sc->eye_shadow_cap = grow_cap(sc->eye_shadow_cap, need_shadow);
                EyeShadow Sh;
                carve_shadow(sc->eye_shadow, sc->eye_shadow_cap, Sh);
                sc->eye_query_cap = grow_cap(sc->eye_query_cap, (size_t)n * 2);
                EyeQuery Q;
                carve_query(sc->eye_query, sc->eye_query_cap, Q);
                ep.cap_items = (uint32_t)std::min<size_t>(sc->eye_cap[nxt], 0xffffffffu);
                ep.cap_shadow = (uint32_t)std::min<size_t>(sc->eye_shadow_cap, 0xffffffffu);
                ep.cap_query = (uint32_t)std::min<size_t>(sc->eye_query_cap, 0xffffffffu);
                // [0..2] appends of k_eye_shade, [4] / [5] ray fetches of k_eye_extend / k_eye_shadow
                const uint32_t zero8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                HIPCHECK(hipMemcpyAsync(counters, zero8, sizeof(zero8), hipMemcpyHostToDevice, sc->stream));
                const uint32_t g = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, (uint64_t)sc->cus * LRT_EYE_SHADE_BLOCKS_PER_CU);
                auto e1 = tc.next();
                HIPCHECK(hipEventRecord(e1.first, sc->stream));
                hipLaunchKernelGGL(k_eye_extend, dim3(g), dim3(kBlock), 0, sc->stream, sc->S, A, n, Hb, counters + 4);
                HIPCHECK(hipEventRecord(e1.second, sc->stream));
                auto e2 = ts.next();
                HIPCHECK(hipEventRecord(e2.first, sc->stream));
                hipLaunchKernelGGL(sc->mesh_lights ? k_eye_shade<true> : k_eye_shade<false>, dim3(g), dim3(kBlock), 0,
                                   sc->stream, sc->S, sc->Ls, sc->M, A, n, Hb, B, Sh,
                                   Q, counters, ep, acc);
                HIPCHECK(hipEventRecord(e2.second, sc->stream));
                HIPCHECK(hipGetLastError());
                uint32_t cnt[4] = {0, 0, 0, 0};
                HIPCHECK(hipMemcpyAsync(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost, sc->stream));
                HIPCHECK(hipStreamSynchronize(sc->stream));
                require(cnt[0] <= ep.cap_items && cnt[1] <= ep.cap_shadow && cnt[2] <= ep.cap_query,
                        "eye pass queue overflow");
                n_ext += n;
                n_shadow += cnt[1];
                n_query += cnt[2];
                if (cnt[1]) {
                    auto e3 = to.next();
                    HIPCHECK(hipEventRecord(e3.first, sc->stream));
                    hipLaunchKernelGGL(k_eye_shadow,
                                       dim3((uint32_t)std::min<uint64_t>((cnt[1] + kBlock - 1) / kBlock, (uint64_t)sc->cus * 16)),
                                       dim3(kBlock), 0, sc->stream, sc->S, Sh, cnt[1], acc, counters + 5);
                    HIPCHECK(hipEventRecord(e3.second, sc->stream));
                }
                if (cnt[2]) {
                    // queries in Morton order of their positions (the estimates
                    // and the fixed-point sums do not depend on the order)
                    const uint32_t nq = cnt[2];
                    sc->eye_sort.alloc((size_t)nq * 16 + 4 * 256);
                    char* c = sc->eye_sort.as<char>();
                    uint32_t* ka = alloc_carve<uint32_t>(c, nq);
                    uint32_t* va = alloc_carve<uint32_t>(c, nq);
                    uint32_t* kb = alloc_carve<uint32_t>(c, nq);
                    uint32_t* vb = alloc_carve<uint32_t>(c, nq);
                    hipLaunchKernelGGL(k_eye_query_keys,
                                       dim3((uint32_t)std::min<uint64_t>((nq + kBlock - 1) / kBlock, (uint64_t)sc->cus * 16)),
                                       dim3(kBlock), 0, sc->stream, Q, nq, box_lo, box_scale, ka, va);
                    uint32_t *sk = ka, *sv = va;
                    radix_sort(*sc, sc->eye_hist, ka, va, kb, vb, nq, 28, sk, sv);
                    const uint32_t gq = (uint32_t)std::min<uint64_t>(
                        (nq + kGatherBlock / kGroup - 1) / (kGatherBlock / kGroup), (uint64_t)sc->cus * 32);
                    if (counting)
                        hipLaunchKernelGGL(slots8 ? k_eye_gather_count<8> : k_eye_gather_count<16>, dim3(gq),
                                           dim3(kGatherBlock), 0, sc->stream, G0, G1, Q, nq,
                                           sc->totals.as<unsigned long long>());
                    auto e4 = tg.next();
                    HIPCHECK(hipEventRecord(e4.first, sc->stream));
                    hipLaunchKernelGGL(slots8 ? k_eye_gather<8> : k_eye_gather<16>, dim3(gq), dim3(kGatherBlock), 0,
                                       sc->stream, G0, G1, Q, sv, nq, acc);
                    HIPCHECK(hipEventRecord(e4.second, sc->stream));
                }
                HIPCHECK(hipGetLastError());
                cur = nxt;
                std::swap(A, B);
                n = cnt[0];
            }
            hipLaunchKernelGGL(k_eye_resolve, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, sc->stream, acc, nc, ns,
                               ep.pixel0, dsum, dsq);
            HIPCHECK(hipGetLastError());
        }
    }
    if (!device_ptrs) {
        HIPCHECK(hipMemcpyAsync(sum_rgb, dsum, img * 8, hipMemcpyDeviceToHost, sc->stream));
        if (sumsq_rgb) HIPCHECK(hipMemcpyAsync(sumsq_rgb, dsq, img * 8, hipMemcpyDeviceToHost, sc->stream));
    }
    HIPCHECK(hipStreamSynchronize(sc->stream));
    const auto t1 = std::chrono::steady_clock::now();
    tc.collect();
    to.collect();
    ts.collect();
    tg.collect();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->samples = pixels * prm->spp;
        stats->extension_rays = n_ext;
        stats->shadow_rays = n_shadow;
        stats->segments = n_query;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->trace_ms = tc.total_ms - before_c;
        stats->occluded_ms = to.total_ms - before_o;
        stats->shade_ms = ts.total_ms - before_s;
        stats->trace_launches = tc.launches - lc0;
        stats->occluded_launches = to.launches - lo0;
        if (counting) {
            unsigned long long t[3];
            HIPCHECK(hipMemcpy(t, sc->totals.p, sizeof(t), hipMemcpyDeviceToHost));
            stats->nodes_visited = t[0];
            stats->refs_read = t[1];
            stats->tris_tested = t[2];
        }
    }
}

int lrt_render_photon(lrt_scene* sc, const lrt_render_params* prm, double* sum_rgb, double* sumsq_rgb, int device_ptrs,
                      lrt_render_stats* stats) {
    return guarded([&] { render_photon_impl(sc, prm, sum_rgb, sumsq_rgb, device_ptrs, stats); });
}

// ---- generators ------------------------------------------------------------------
int lrt_generate_scene(const lrt_gen_params* params, const float* tables, uint32_t n_tables, const lrt_light_desc* sky,
                       lrt_generated** out) {
    return guarded([&] {
        require(params && out, "null pointer");
        *out = (lrt_generated*)generate(*params, tables, n_tables, sky);
    });
}
int lrt_generate_obj_scene(const char* obj_path, const lrt_material* materials, uint32_t n_materials,
                           const float* tables, uint32_t n_tables, float aspect, lrt_generated** out) {
    return guarded([&] {
        require(obj_path && materials && out, "null pointer");
        require(n_materials >= 1, "the mesh needs a material");
        *out = nullptr;
        try {
            *out = (lrt_generated*)generate_obj_scene(obj_path, materials, n_materials, tables, n_tables, aspect);
        } catch (const std::invalid_argument& e) {
            throw Error(LRT_ERR_ARG, e.what());
        }
    });
}
const lrt_scene_desc* lrt_generated_desc(const lrt_generated* g) { return g ? generated_desc((const Generated*)g) : nullptr; }
int lrt_generated_free(lrt_generated* g) {
    free_generated((Generated*)g);
    return LRT_OK;
}

// ---- image files ---------------------------------------------------------
int lrt_image_write(const char* path, const float* rgb, uint32_t width, uint32_t height, int format,
                    int flip_vertically) {
    return guarded([&] {
        require(path && rgb, "null pointer");
        require(width >= 1 && height >= 1, "empty image");
        if (format == LRT_IMAGE_JPEG) throw Error(LRT_ERR_UNSUPPORTED, "JPEG is not written (lossy)");
        require(format >= LRT_IMAGE_PNG && format <= LRT_IMAGE_EXR, "unknown image format");
        std::string err;
        if (!write_image(path, rgb, width, height, format, flip_vertically != 0, err)) throw Error(LRT_ERR_ARG, err);
    });
}

int lrt_accum_write(const char* path, const double* sum, const double* sumsq, const lrt_accum_info* info) {
    return guarded([&] {
        require(path && sum && info, "null pointer");
        require(info->width >= 1 && info->height >= 1 && info->width <= 65536 && info->height <= 65536, "bad image size");
        std::string err;
        if (!write_accum(path, sum, sumsq, *info, err)) throw Error(LRT_ERR_ARG, err);
    });
}

int lrt_accum_read(const char* path, double* sum, double* sumsq, lrt_accum_info* info) {
    return guarded([&] {
        require(path && info, "null pointer");
        std::string err;
        if (!read_accum(path, sum, sumsq, *info, err)) throw Error(LRT_ERR_ARG, err);
    });
}

int lrt_image_to_bytes(const float* rgb, uint32_t width, uint32_t height, int flip_vertically, uint8_t* out) {
    return guarded([&] {
        require(rgb && out, "null pointer");
        image_to_bytes(rgb, width, height, flip_vertically != 0, out);
    });
}

// ---- OBJ meshes ----------------------------------------------------------
struct lrt_mesh {
    std::vector<float> positions, corner_normals;
    std::vector<uint32_t> indices;
};

int lrt_load_obj(const char* path, lrt_mesh** out) {
    return guarded([&] {
        require(path && out, "null pointer");
        *out = nullptr;
        auto m = std::make_unique<lrt_mesh>();
        std::string err;
        if (!load_obj(path, m->positions, m->indices, m->corner_normals, err)) throw Error(LRT_ERR_ARG, "load_obj: " + err);
        *out = m.release();
    });
}

int lrt_mesh_view(const lrt_mesh* m, uint32_t* n_vertices, uint32_t* n_triangles, const float** positions,
                  const uint32_t** indices, const float** corner_normals) {
    return guarded([&] {
        require(m && n_vertices && n_triangles, "null pointer");
        *n_vertices = (uint32_t)(m->positions.size() / 3);
        *n_triangles = (uint32_t)(m->indices.size() / 3);
        if (positions) *positions = m->positions.data();
        if (indices) *indices = m->indices.data();
        if (corner_normals) *corner_normals = m->corner_normals.data();
    });
}

int lrt_mesh_free(lrt_mesh* m) {
    delete m;
    return LRT_OK;
}

}  // extern "C"

namespace lrt {
void set_error(const std::string& msg) { g_error = msg; }
}  // namespace lrt

#include "denoise.hip"
