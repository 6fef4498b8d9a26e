// Oracle (test infrastructure only): extern "C" entry points used by tests/
// (ctypes), __graft_entry__.smoke() and bench.py's cpu_baseline leg.
#include <cstring>
#include <thread>

#include "orc.h"

using namespace orc;


template <typename F>
static void parallel_for(uint32_t n, int threads, F&& f) {
    const int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    auto body = [&](int t) {
        const uint64_t b = (uint64_t)n * t / nt, e = (uint64_t)n * (t + 1) / nt;
        for (uint64_t i = b; i < e; ++i) f((uint32_t)i);
    };
    for (int t = 1; t < nt; ++t) pool.emplace_back(body, t);
    body(0);
    for (auto& th : pool) th.join();
}


extern "C" {

int orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox4x32_10(ctr, key, out);
    return 0;
}

int orc_kd_build_export(const float* boxes, uint32_t n, uint32_t* n_nodes, uint32_t* n_refs, int32_t* depth_limit,
                        uint8_t* leaf, int32_t* axis, float* split, int32_t* left, int32_t* right, int32_t* first,
                        int32_t* count, uint32_t* refs) {
    std::vector<box> bx(n);
    for (uint32_t i = 0; i < n; ++i) {
        bx[i].lo = load3(boxes + 6 * (size_t)i);
        bx[i].hi = load3(boxes + 6 * (size_t)i + 3);
    }
    kdtree kd;
    kd.build(bx);
    if (n_nodes) *n_nodes = (uint32_t)kd.nodes.size();
    if (n_refs) *n_refs = (uint32_t)kd.refs.size();
    if (depth_limit) *depth_limit = kd.depth_limit;
    if (leaf) {
        for (size_t i = 0; i < kd.nodes.size(); ++i) {
            const kd_node& nd = kd.nodes[i];
            leaf[i] = nd.leaf; axis[i] = nd.axis; split[i] = nd.split; left[i] = nd.left;
            right[i] = nd.right; first[i] = nd.first; count[i] = nd.count;
        }
        std::memcpy(refs, kd.refs.data(), kd.refs.size() * 4);
    }
    return 0;
}

void* orc_scene_create(const lrt_scene_desc* desc) {
    try {
        return make_scene(desc).release();
    } catch (...) {
        return nullptr;
    }
}
void orc_scene_free(void* h) { delete (scene*)h; }


static ray ray_at(const float* rays, uint32_t n, uint32_t i, float& tmin, float& tmax) {
    ray r{v3(rays[i], rays[n + i], rays[2 * (size_t)n + i]),
          v3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i])};
    tmin = rays[6 * (size_t)n + i];
    tmax = rays[7 * (size_t)n + i];
    return r;
}

// closest hit per ray; counts (3*n: nodes, refs, tris) optional
int orc_scene_trace(void* h, const float* rays, uint32_t n, int32_t* obj_prim, float* t, float* uv, uint32_t* counts,
                    int threads) {
    const scene& S = *(scene*)h;
    parallel_for(n, threads, [&](uint32_t i) {
        float tmin, tmax;
        const ray r = ray_at(rays, n, i, tmin, tmax);
        hit_rec rec;
        kd_counts c;
        const bool any = S.hit(r, tmin, tmax, rec, counts ? &c : nullptr);
        obj_prim[2 * (size_t)i] = any ? rec.obj : -1;
        obj_prim[2 * (size_t)i + 1] = any ? rec.prim : -1;
        t[i] = any ? rec.t : 0.0f;
        if (uv) { uv[2 * (size_t)i] = any ? rec.u : 0.0f; uv[2 * (size_t)i + 1] = any ? rec.v : 0.0f; }
        if (counts) { counts[3 * (size_t)i] = (uint32_t)c.nodes; counts[3 * (size_t)i + 1] = (uint32_t)c.refs; counts[3 * (size_t)i + 2] = (uint32_t)c.tris; }
    });
    return 0;
}

// brute force: every triangle of every mesh object in order, window [tmin, closest]
int orc_scene_trace_brute(void* h, const float* rays, uint32_t n, int32_t* obj_prim, float* t, int threads) {
    const scene& S = *(scene*)h;
    parallel_for(n, threads, [&](uint32_t i) {
        float tmin, tmax;
        const ray r = ray_at(rays, n, i, tmin, tmax);
        float closest = tmax;
        int bo = -1, bp = -1;
        for (size_t o = 0; o < S.objects.size(); ++o) {
            const object& ob = S.objects[o];
            hit_rec rec;
            if (ob.shape == LRT_SHAPE_MESH) {
                for (uint32_t tri = 0; tri < ob.idx.size() / 3; ++tri)
                    if (ob.hit_mesh_tri(tri, r, tmin, closest, closest, rec)) { bo = (int)o; bp = (int)tri; }
            } else if (ob.intersect(r, tmin, closest, rec)) {
                closest = rec.t; bo = (int)o; bp = 0;
            }
        }
        obj_prim[2 * (size_t)i] = bo;
        obj_prim[2 * (size_t)i + 1] = bp;
        t[i] = bo >= 0 ? closest : 0.0f;
    });
    return 0;
}

// blocked test of path_integrator.h:113-116: a closest hit in [tmin, tmax]
// whose emitter differs from ignore_emitter[i]
int orc_scene_occluded(void* h, const float* rays, uint32_t n, const int32_t* ignore, uint8_t* out, int threads) {
    const scene& S = *(scene*)h;
    parallel_for(n, threads, [&](uint32_t i) {
        float tmin, tmax;
        const ray r = ray_at(rays, n, i, tmin, tmax);
        hit_rec rec;
        const bool any = S.hit(r, tmin, tmax, rec);
        // the reference compares against a real emitter (never null); -1 here
        // means no emitter is exempt, so any hit blocks
        const int ig = ignore ? ignore[i] : -1;
        out[i] = (uint8_t)(any && (ig < 0 || rec.emitter != ig));
    });
    return 0;
}

int orc_bsdf_eval_pdf(const lrt_material* mats, uint32_t nm, const float* tables, uint32_t nt, const int32_t* mat,
                      const float* wi, const float* wo, const uint8_t* front, uint32_t n, int measure, int mode,
                      int component, uint32_t type_mask, float* f_cos, float* pdf) {
    (void)nt;
    const std::vector<lrt_material> M(mats, mats + nm);
    for (uint32_t i = 0; i < n; ++i) {
        bsdf_ctx ctx;
        ctx.mode = mode; ctx.component = component; ctx.type_mask = type_mask;
        bsdf_rec rec;
        rec.wi = load3(wi + 3 * (size_t)i);
        rec.wo = load3(wo + 3 * (size_t)i);
        rec.front_face = front[i] != 0;
        const v3 f = bsdf_eval(M, tables, mat[i], ctx, rec, measure);
        f_cos[3 * (size_t)i] = f.x; f_cos[3 * (size_t)i + 1] = f.y; f_cos[3 * (size_t)i + 2] = f.z;
        pdf[i] = bsdf_pdf(M, tables, mat[i], ctx, rec, measure);
    }
    return 0;
}

int orc_bsdf_sample(const lrt_material* mats, uint32_t nm, const float* tables, uint32_t nt, const int32_t* mat,
                    const float* wi, const uint8_t* front, const float* u, const float* lobe_u, uint32_t n, int mode,
                    int component, uint32_t type_mask, float* weight, float* wo, float* pdf, float* eta,
                    uint32_t* sampled_type, int32_t* sampled_component) {
    (void)nt;
    const std::vector<lrt_material> M(mats, mats + nm);
    for (uint32_t i = 0; i < n; ++i) {
        bsdf_ctx ctx;
        ctx.mode = mode; ctx.component = component; ctx.type_mask = type_mask;
        ctx.lobe_sample = lobe_u[i];
        bsdf_rec rec;
        rec.wi = load3(wi + 3 * (size_t)i);
        rec.front_face = front[i] != 0;
        const v3 w = bsdf_sample(M, tables, mat[i], ctx, rec, u[2 * (size_t)i], u[2 * (size_t)i + 1]);
        weight[3 * (size_t)i] = w.x; weight[3 * (size_t)i + 1] = w.y; weight[3 * (size_t)i + 2] = w.z;
        wo[3 * (size_t)i] = rec.wo.x; wo[3 * (size_t)i + 1] = rec.wo.y; wo[3 * (size_t)i + 2] = rec.wo.z;
        pdf[i] = rec.pdf; eta[i] = rec.eta; sampled_type[i] = rec.sampled_type; sampled_component[i] = rec.sampled_component;
    }
    return 0;
}

int orc_bsdf_flags(const lrt_material* mats, uint32_t nm, int32_t m, uint32_t* flags, int32_t* count) {
    const std::vector<lrt_material> M(mats, mats + nm);
    *flags = bsdf_flags(M, m);
    *count = bsdf_component_count(M, m);
    return 0;
}

int orc_sky_eval(const lrt_light_desc* sky, const float* dirs, uint32_t n, float* rgb) {
    lrt_scene_desc d{};
    d.n_lights = 1;
    d.lights = sky;
    d.camera.fov_deg = 40.0f; d.camera.aspect = 1.0f; d.camera.origin[2] = 1.0f; d.camera.up[1] = 1.0f;
    auto s = make_scene(&d);
    for (uint32_t i = 0; i < n; ++i) {
        const v3 c = sky_eval(s->lights[0], load3(dirs + 3 * (size_t)i));
        rgb[3 * (size_t)i] = c.x; rgb[3 * (size_t)i + 1] = c.y; rgb[3 * (size_t)i + 2] = c.z;
    }
    return 0;
}

// stats: [extension rays, shadow rays]
int orc_render(void* h, int W, int H, int spp, int spp_offset, int max_depth, int rr_depth, uint64_t seed, int mode,
               int threads, int row_begin, int row_end, double* sum, double* sumsq, float* per_sample, float* fsum,
               uint64_t* stats) {
    render_job J{};
    J.S = (scene*)h;
    J.W = W; J.H = H; J.spp = spp; J.spp_offset = spp_offset; J.max_depth = max_depth; J.rr_depth = rr_depth;
    J.mode = mode; J.threads = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    J.key[0] = (uint32_t)seed; J.key[1] = (uint32_t)(seed >> 32);
    J.row_begin = row_end > row_begin ? row_begin : 0;
    J.row_end = row_end > row_begin ? row_end : H;
    J.sum = sum; J.sumsq = sumsq; J.per_sample = per_sample; J.fsum = fsum;
    orc::render(J);
    if (stats) { stats[0] = J.ext; stats[1] = J.shadow; }
    return 0;
}

}  // extern "C"
