// Test harness around the reference's OWN sources (compiled from
// /root/reference by oracle/Makefile into oracle/_ref/libref.so).  It holds no
// copy of reference code: it includes the reference headers where they lie
// and exposes extern "C" entry points mirroring the oracle's (orc_*), so tests
// can pin the oracle restatement against the real reference.
//
// Private members of sah_kdtree (nodes, ordered), triangle_mesh (kdtree) and
// rough_transmittance (trans, diff_trans) are read through the access macro
// below; the photon integrator's maps and radii are set through it from given
// photon lists (its own preprocess gathers in a thread-dependent order).
#include "../include/linrt.h"

#define private public
#include "accelerator/sah_kdtree.h"
#include "core/scene.h"
#include "core/sampler.h"
#include "emitter/area_emitter.h"
#include "emitter/env_emitter.h"
#include "geometry/rectangle.h"
#include "geometry/sphere.h"
#include "geometry/triangle_mesh.h"
#include "integrator/path/path_integrator.h"
#include "integrator/photonmapper/photon_integrator.h"
#include "integrator/photonmapper/photon_map.h"
#include "integrator/photonmapper/photon_tracer.h"
#include "io/image_write.h"
#include "io/obj_loader.h"
#include "material/coating_bsdf.h"
#include "material/dielectric_bsdf.h"
#include "material/diffuse_bsdf.h"
#include "material/plastic_bsdf.h"
#include "material/rough_coating_bsdf.h"
#include "material/rough_conductor_bsdf.h"
#include "material/rough_dielectric_bsdf.h"
#include "material/rough_plastic_bsdf.h"
#undef private

#include <unistd.h>

using namespace LinRT;

namespace {

std::string g_root = "/root/reference";

struct cwd_guard {  // rough_transmittance opens third_party/data/ggx.dat relative to the CWD
    char old[4096];
    cwd_guard() {
        if (!getcwd(old, sizeof(old))) old[0] = 0;
        if (chdir(g_root.c_str()) != 0) throw std::runtime_error("chdir");
    }
    ~cwd_guard() {
        if (old[0] && chdir(old) != 0) {}
    }
};

vec3 V(const float* p) { return vec3(p[0], p[1], p[2]); }

std::shared_ptr<bsdf> make_bsdf(const lrt_material* mats, int m, std::vector<std::shared_ptr<bsdf>>& cache) {
    if (cache[(size_t)m]) return cache[(size_t)m];
    const lrt_material& M = mats[m];
    std::shared_ptr<bsdf> b;
    cwd_guard g;
    switch (M.type) {
        case LRT_BSDF_DIFFUSE: b = std::make_shared<diffuse_bsdf>(V(M.diffuse)); break;
        case LRT_BSDF_CONDUCTOR:
            b = std::make_shared<rough_conductor_bsdf>(conductor_ior{V(M.cond_eta), V(M.cond_k)}, M.alpha,
                                                       V(M.specular_reflectance));
            break;
        case LRT_BSDF_DIELECTRIC:
            b = std::make_shared<dielectric_bsdf>(M.int_ior, M.ext_ior, V(M.specular_reflectance),
                                                  V(M.specular_transmittance));
            break;
        case LRT_BSDF_ROUGH_DIELECTRIC:
            b = std::make_shared<rough_dielectric_bsdf>(M.int_ior, M.ext_ior, M.alpha, V(M.specular_reflectance),
                                                        V(M.specular_transmittance));
            break;
        case LRT_BSDF_PLASTIC:
            b = std::make_shared<plastic_bsdf>(M.int_ior, M.ext_ior, V(M.specular_reflectance), V(M.diffuse),
                                               M.nonlinear != 0);
            break;
        case LRT_BSDF_ROUGH_PLASTIC:
            b = std::make_shared<rough_plastic_bsdf>(M.int_ior, M.ext_ior, M.alpha, V(M.specular_reflectance),
                                                     V(M.diffuse), M.nonlinear != 0);
            break;
        case LRT_BSDF_COATING:
            b = std::make_shared<coating_bsdf>(M.int_ior, M.ext_ior, make_bsdf(mats, M.nested, cache),
                                               V(M.specular_reflectance), V(M.sigma_a), M.thickness);
            break;
        case LRT_BSDF_ROUGH_COATING:
            b = std::make_shared<rough_coating_bsdf>(M.int_ior, M.ext_ior, M.alpha, make_bsdf(mats, M.nested, cache),
                                                     V(M.specular_reflectance), V(M.sigma_a), M.thickness);
            break;
        default: return nullptr;
    }
    cache[(size_t)m] = b;
    return b;
}

struct ref_scene {
    std::vector<std::shared_ptr<hittable>> objects;
    std::vector<int> kind;
    std::unique_ptr<scene> world;
    std::vector<std::shared_ptr<emitter>> emitters;
};

ray ray_at(const float* rays, uint32_t n, uint32_t i, float& tmin, float& tmax) {
    ray r(vec3(rays[i], rays[n + i], rays[2 * (size_t)n + i]),
          vec3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]));
    tmin = rays[6 * (size_t)n + i];
    tmax = rays[7 * (size_t)n + i];
    return r;
}

template <typename F>
void parallel_for(uint32_t n, int threads, F&& f) {
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

}  // namespace

extern "C" {

// LinRT::load_obj (obj_loader.h:44-141): returns 0 and the counts (and the
// arrays where the pointers are not null), or 1 where it returns nullptr
int ref_load_obj(const char* path, uint32_t* n_vertices, uint32_t* n_corners, float* positions, uint32_t* indices,
                 float* corner_normals) {
    const auto m = LinRT::load_obj(path);
    if (!m) return 1;
    *n_vertices = (uint32_t)m->positions.size();
    *n_corners = (uint32_t)m->indices.size();
    for (size_t i = 0; positions && i < m->positions.size(); ++i) {
        positions[3 * i] = m->positions[i].x;
        positions[3 * i + 1] = m->positions[i].y;
        positions[3 * i + 2] = m->positions[i].z;
    }
    for (size_t i = 0; indices && i < m->indices.size(); ++i) indices[i] = m->indices[i];
    for (size_t i = 0; corner_normals && i < m->corner_normals.size(); ++i) {
        corner_normals[3 * i] = m->corner_normals[i].x;
        corner_normals[3 * i + 1] = m->corner_normals[i].y;
        corner_normals[3 * i + 2] = m->corner_normals[i].z;
    }
    return 0;
}

// ImageWriter::save (image_write.cpp:22-100); format: 0 PNG, 1 JPEG, 2 PPM, 3 HDR
int ref_image_save(const char* path, int w, int h, const float* rgb, int format, int flip) {
    std::vector<LinRT::color> px((size_t)w * h);
    for (size_t i = 0; i < px.size(); ++i) px[i] = LinRT::color(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    const LinRT::ImageFormat f[4] = {LinRT::ImageFormat::PNG, LinRT::ImageFormat::JPEG, LinRT::ImageFormat::PPM,
                                     LinRT::ImageFormat::HDR};
    return LinRT::ImageWriter::save(path, w, h, px, f[format], flip != 0) ? 0 : 1;
}

int ref_set_root(const char* root) {
    g_root = root;
    return 0;
}

int ref_kd_build_export(const float* boxes, uint32_t n, uint32_t* n_nodes, uint32_t* n_refs, int32_t* depth_limit,
                        uint8_t* leaf, int32_t* axis, float* split, int32_t* left, int32_t* right, int32_t* first,
                        int32_t* count, uint32_t* refs) {
    std::vector<aabb> bx(n);
    for (uint32_t i = 0; i < n; ++i) {
        bx[i].extend(V(boxes + 6 * (size_t)i));
        bx[i].extend(V(boxes + 6 * (size_t)i + 3));
    }
    sah_kdtree kd;
    kd.build(bx);
    if (n_nodes) *n_nodes = (uint32_t)kd.nodes.size();
    if (n_refs) *n_refs = (uint32_t)kd.ordered.size();
    if (depth_limit) *depth_limit = kd.max_depth;
    if (leaf) {
        for (size_t i = 0; i < kd.nodes.size(); ++i) {
            leaf[i] = kd.nodes[i].leaf; axis[i] = kd.nodes[i].axis; split[i] = kd.nodes[i].split;
            left[i] = kd.nodes[i].left; right[i] = kd.nodes[i].right; first[i] = kd.nodes[i].first;
            count[i] = kd.nodes[i].count;
        }
        std::memcpy(refs, kd.ordered.data(), kd.ordered.size() * 4);
    }
    return 0;
}

void* ref_scene_create(const lrt_scene_desc* d) {
    try {
        auto* s = new ref_scene();
        std::vector<std::shared_ptr<bsdf>> cache(d->n_materials);
        hittable_list list;
        for (uint32_t i = 0; i < d->n_objects; ++i) {
            const lrt_object_desc& o = d->objects[i];
            std::shared_ptr<hittable> h;
            if (o.shape == LRT_SHAPE_MESH) {
                if (o.triangle_material) { delete s; return nullptr; }  // one material per mesh in the reference
                std::vector<vec3> pos(o.n_vertices);
                for (uint32_t v = 0; v < o.n_vertices; ++v) pos[v] = V(o.positions + 3 * (size_t)v);
                std::vector<uint32_t> idx(o.indices, o.indices + 3 * (size_t)o.n_triangles);
                std::vector<vec3> cn(3 * (size_t)o.n_triangles);
                for (size_t k = 0; k < cn.size(); ++k) cn[k] = o.corner_normals ? V(o.corner_normals + 3 * k) : vec3(0.0f);
                h = std::make_shared<triangle_mesh>(std::move(pos), std::move(idx), std::move(cn));
            } else if (o.shape == LRT_SHAPE_RECTANGLE) {
                h = std::make_shared<rectangle>(V(o.corner), V(o.edge_u), V(o.edge_v));
            } else {
                h = std::make_shared<sphere>(V(o.center), o.radius);
            }
            if (o.material >= 0 && o.shape != LRT_SHAPE_RECTANGLE) h->set_material(make_bsdf(d->materials, o.material, cache));
            s->objects.push_back(h);
            s->kind.push_back(o.shape);
            list.add(h);
        }
        for (uint32_t i = 0; i < d->n_lights; ++i) {
            const lrt_light_desc& L = d->lights[i];
            if (L.kind == LRT_LIGHT_AREA) {
                s->emitters.push_back(std::make_shared<area_emitter>(s->objects[(size_t)L.object], V(L.radiance)));
            } else {
                s->emitters.push_back(std::make_shared<env_emitter>(V(L.sun_dir), L.turbidity, L.ground_albedo, L.scale));
            }
        }
        camera cam;
        cam.origin = V(d->camera.origin);
        cam.look_at = V(d->camera.look_at);
        cam.update(V(d->camera.up), d->camera.fov_deg, d->camera.aspect);
        s->world = std::make_unique<scene>(std::move(list), s->emitters, cam);
        return s;
    } catch (...) {
        return nullptr;
    }
}

void ref_scene_free(void* h) { delete (ref_scene*)h; }

// hittable_list::hit with (object, primitive) ids: meshes go through their own
// kd-tree and hit_triangle exactly as triangle_mesh::hit does.
int ref_scene_trace(void* h, const float* rays, uint32_t n, int32_t* obj_prim, float* t, float* uv, int threads) {
    (void)uv;
    ref_scene& S = *(ref_scene*)h;
    parallel_for(n, threads, [&](uint32_t i) {
        float tmin, tmax;
        const ray r = ray_at(rays, n, i, tmin, tmax);
        float closest_so_far = tmax;
        int bo = -1, bp = -1;
        float bt = 0.0f;
        for (size_t o = 0; o < S.objects.size(); ++o) {
            hit_record rec;
            if (S.kind[o] == LRT_SHAPE_MESH) {
                auto* m = (triangle_mesh*)S.objects[o].get();
                bool any = false;
                int prim = -1;
                float closest = closest_so_far;
                m->kdtree.intersect(r, tmin, closest, [&](std::uint32_t tri, float mint, float maxt) {
                    if (m->hit_triangle(tri, r, mint, maxt, closest, rec)) { any = true; prim = (int)tri; }
                });
                if (any) { closest_so_far = rec.t; bo = (int)o; bp = prim; bt = rec.t; }
            } else if (S.objects[o]->hit(r, tmin, closest_so_far, rec)) {
                closest_so_far = rec.t; bo = (int)o; bp = 0; bt = rec.t;
            }
        }
        obj_prim[2 * (size_t)i] = bo;
        obj_prim[2 * (size_t)i + 1] = bp;
        t[i] = bt;
    });
    return 0;
}

int ref_bsdf_eval_pdf(const lrt_material* mats, uint32_t nm, const float*, uint32_t, const int32_t* mat,
                      const float* wi, const float* wo, const uint8_t* front, uint32_t n, int meas, int mode,
                      int component, uint32_t type_mask, float* f_cos, float* pdf) {
    try {
        std::vector<std::shared_ptr<bsdf>> cache(nm);
        for (uint32_t i = 0; i < n; ++i) {
            auto b = make_bsdf(mats, mat[i], cache);
            bsdf_context ctx;
            ctx.mode = mode == LRT_MODE_RADIANCE ? bsdf_context::transport_mode::RADIANCE
                                                 : bsdf_context::transport_mode::IMPORTANCE;
            ctx.component = component;
            ctx.type_mask = type_mask;
            bsdf_sample_record rec;
            rec.wi = V(wi + 3 * (size_t)i);
            rec.wo = V(wo + 3 * (size_t)i);
            rec.front_face = front[i] != 0;
            const measure m = meas == LRT_MEASURE_DISCRETE ? measure::discrete : measure::solid_angle;
            const color f = b->eval(ctx, rec, m);
            f_cos[3 * (size_t)i] = f.x; f_cos[3 * (size_t)i + 1] = f.y; f_cos[3 * (size_t)i + 2] = f.z;
            pdf[i] = b->pdf(ctx, rec, m);
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_bsdf_sample(const lrt_material* mats, uint32_t nm, const float*, uint32_t, const int32_t* mat,
                    const float* wi, const uint8_t* front, const float* u, const float* lobe_u, uint32_t n, int mode,
                    int component, uint32_t type_mask, float* weight, float* wo, float* pdf, float* eta,
                    uint32_t* sampled_type, int32_t* sampled_component) {
    try {
        std::vector<std::shared_ptr<bsdf>> cache(nm);
        for (uint32_t i = 0; i < n; ++i) {
            auto b = make_bsdf(mats, mat[i], cache);
            bsdf_context ctx;
            ctx.mode = mode == LRT_MODE_RADIANCE ? bsdf_context::transport_mode::RADIANCE
                                                 : bsdf_context::transport_mode::IMPORTANCE;
            ctx.component = component;
            ctx.type_mask = type_mask;
            ctx.lobe_sample = lobe_u[i];
            bsdf_sample_record rec;
            rec.wi = V(wi + 3 * (size_t)i);
            rec.front_face = front[i] != 0;
            const color w = b->sample(ctx, rec, vec2(u[2 * (size_t)i], u[2 * (size_t)i + 1]));
            weight[3 * (size_t)i] = w.x; weight[3 * (size_t)i + 1] = w.y; weight[3 * (size_t)i + 2] = w.z;
            wo[3 * (size_t)i] = rec.wo.x; wo[3 * (size_t)i + 1] = rec.wo.y; wo[3 * (size_t)i + 2] = rec.wo.z;
            pdf[i] = rec.pdf; eta[i] = rec.eta; sampled_type[i] = rec.sampled_type;
            sampled_component[i] = rec.sampled_component;
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_bsdf_flags(const lrt_material* mats, uint32_t nm, int32_t m, uint32_t* flags, int32_t* count) {
    try {
        std::vector<std::shared_ptr<bsdf>> cache(nm);
        auto b = make_bsdf(mats, m, cache);
        *flags = b->flags();
        *count = b->component_count();
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_fdr(float eta, float* out) {
    *out = fresnel_diffuse_reflectance(eta);
    return 0;
}

int ref_trans_slice(float eta, float alpha, float* out101) {
    try {
        rough_transmittance t(eta, alpha, g_root + "/third_party/data/ggx.dat");
        for (int i = 0; i < 100; ++i) out101[i] = t.trans[(size_t)i];
        out101[100] = t.diff_trans[0];
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_hosek_cook(float turbidity, float albedo, double elevation, double* cfg27, double* rad3) {
    ArHosekSkyModelState* st = arhosek_rgb_skymodelstate_alloc_init(turbidity, albedo, elevation);
    for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < 9; ++k) cfg27[c * 9 + k] = st->configs[c][k];
        rad3[c] = st->radiances[c];
    }
    arhosekskymodelstate_free(st);
    return 0;
}

// env_emitter: elevation the constructor passes to the cook (env_emitter.h:26)
int ref_sky_elevation(const float* sun, double* elevation) {
    const vec3 s = normalize(V(sun));
    *elevation = std::asin(std::clamp(s.y, -1.0f, 1.0f));
    return 0;
}

int ref_sky_eval(const float* sun, float turbidity, float albedo, float scale, const float* dirs, uint32_t n, float* rgb) {
    env_emitter e(V(sun), turbidity, albedo, scale);
    for (uint32_t i = 0; i < n; ++i) {
        const color c = e.eval(V(dirs + 3 * (size_t)i));
        rgb[3 * (size_t)i] = c.x; rgb[3 * (size_t)i + 1] = c.y; rgb[3 * (size_t)i + 2] = c.z;
    }
    return 0;
}

// sampler sequence: n calls of next_2d after start(pixel, sample)
int ref_sampler_seq(uint32_t pixel, uint32_t sample, uint32_t n, float* out) {
    LinRT::sampler s;
    s.start(pixel, sample);
    for (uint32_t i = 0; i < n; ++i) {
        const vec2 v = s.next_2d();
        out[2 * i] = v.x;
        out[2 * i + 1] = v.y;
    }
    return 0;
}

// photon_tracer::trace_particle for particles [first, first+count), each with
// its sampler started at (0, index) as photon_gather does (photon_gather.h:45),
// one after another (the reference interleaves chunks over threads)
int64_t ref_trace_particles(void* h, int kind, uint32_t first, uint32_t count, int max_depth, int rr_depth,
                            float* out12, uint32_t* depth_particle, int64_t capacity) {
    ref_scene& S = *(ref_scene*)h;
    std::vector<photon> ph;
    std::vector<uint32_t> owner;
    for (uint32_t i = 0; i < count; ++i) {
        LinRT::sampler smp;
        smp.start(0, first + i);
        const size_t before = ph.size();
        photon_tracer::trace_particle(*S.world, kind == 1 ? photon_tracer::gather::caustic : photon_tracer::gather::surface,
                                      smp, max_depth, rr_depth, ph);
        owner.resize(ph.size(), first + i);
        (void)before;
    }
    if (out12) {
        const int64_t n = std::min<int64_t>((int64_t)ph.size(), capacity);
        for (int64_t i = 0; i < n; ++i) {
            const photon& p = ph[(size_t)i];
            float* o = out12 + 12 * i;
            o[0] = p.position.x; o[1] = p.position.y; o[2] = p.position.z;
            o[3] = p.direction.x; o[4] = p.direction.y; o[5] = p.direction.z;
            o[6] = p.normal.x; o[7] = p.normal.y; o[8] = p.normal.z;
            o[9] = p.power.x; o[10] = p.power.y; o[11] = p.power.z;
            depth_particle[2 * i] = p.depth;
            depth_particle[2 * i + 1] = owner[(size_t)i];
        }
    }
    return (int64_t)ph.size();
}

// photon_map::build + estimate_irradiance (photon_map.h:26-70)
int ref_photon_estimate(const float* photons12, const uint32_t* depth_particle, int64_t n_photons, float scale,
                        const float* queries6, uint32_t n_queries, float radius, int max_depth, int k, float* out3) {
    std::vector<photon> ph((size_t)n_photons);
    for (int64_t i = 0; i < n_photons; ++i) {
        const float* p = photons12 + 12 * i;
        photon& o = ph[(size_t)i];
        o.position = V(p); o.direction = V(p + 3); o.normal = V(p + 6); o.power = V(p + 9);
        o.depth = (std::uint16_t)depth_particle[2 * i];
    }
    photon_map m;
    m.build(std::move(ph), scale);
    for (uint32_t q = 0; q < n_queries; ++q) {
        const color e = m.estimate_irradiance(V(queries6 + 6 * (size_t)q), V(queries6 + 6 * (size_t)q + 3), radius,
                                              max_depth, (size_t)k);
        out3[3 * (size_t)q] = e.x; out3[3 * (size_t)q + 1] = e.y; out3[3 * (size_t)q + 2] = e.z;
    }
    return 0;
}

// src/main.cpp:118-136 loop with the reference path_integrator (max depth 50,
// RR from depth 5, path_integrator.h:173,19) and the reference sampler.
// per_sample: 3 floats per (pixel, sample); fsum: FP32 per-pixel sum.
int ref_render(void* h, int W, int H, int spp, int row_begin, int row_end, float* per_sample, float* fsum, int threads) {
    ref_scene& S = *(ref_scene*)h;
    const path_integrator integ;
    const int rb = row_end > row_begin ? row_begin : 0, re = row_end > row_begin ? row_end : H;
    const uint32_t rows = (uint32_t)(re - rb);
    parallel_for(rows, threads, [&](uint32_t ry) {
        const int y = rb + (int)ry;
        for (int x = 0; x < W; ++x) {
            color pc(0.0f);
            for (int s = 0; s < spp; ++s) {
                const auto pix = static_cast<uint32_t>(y * W + x);
                LinRT::sampler smp;
                smp.start(pix, static_cast<uint32_t>(s));
                const auto off = smp.next_2d();
                const auto u = (static_cast<float>(x) + off.x) / static_cast<float>(W - 1);
                const auto v = (static_cast<float>(y) + off.y) / static_cast<float>(H - 1);
                auto r = S.world->sensor().get_ray(u, v);
                const color L = integ.radiance(r, *S.world, smp);
                pc += L;
                if (per_sample) {
                    float* o = per_sample + ((size_t)pix * spp + s) * 3;
                    o[0] = L.x; o[1] = L.y; o[2] = L.z;
                }
            }
            if (fsum) {
                float* o = fsum + (size_t)(y * W + x) * 3;
                o[0] = pc.x; o[1] = pc.y; o[2] = pc.z;
            }
        }
    });
    return 0;
}

// photon_integrator::radiance (photon_integrator.h:39-43) per (pixel, sample)
// in the src/main.cpp:118-136 loop, with the global and caustic maps built
// from the given photons (n = 0: empty map) and the given radii.
int ref_render_photon(void* h, int W, int H, int spp, const float* g_ph, const uint32_t* g_dp, int64_t g_n,
                      float g_scale, float g_radius, const float* c_ph, const uint32_t* c_dp, int64_t c_n,
                      float c_scale, float c_radius, int row_begin, int row_end, float* per_sample, int threads) {
    ref_scene& S = *(ref_scene*)h;
    photon_integrator integ;
    auto load = [](const float* ph, const uint32_t* dp, int64_t n) {
        std::vector<photon> v((size_t)std::max<int64_t>(n, 0));
        for (int64_t i = 0; i < n; ++i) {
            const float* p = ph + 12 * i;
            photon& o = v[(size_t)i];
            o.position = V(p); o.direction = V(p + 3); o.normal = V(p + 6); o.power = V(p + 9);
            o.depth = (std::uint16_t)dp[2 * i];
        }
        return v;
    };
    if (g_n > 0) integ.global_map.build(load(g_ph, g_dp, g_n), g_scale);
    if (c_n > 0) integ.caustic_map.build(load(c_ph, c_dp, c_n), c_scale);
    integ.global_radius = g_radius;
    integ.caustic_radius = c_radius;
    const int rb = row_end > row_begin ? row_begin : 0, re = row_end > row_begin ? row_end : H;
    parallel_for((uint32_t)(re - rb), threads, [&](uint32_t ry) {
        const int y = rb + (int)ry;
        for (int x = 0; x < W; ++x)
            for (int s = 0; s < spp; ++s) {
                const auto pix = static_cast<uint32_t>(y * W + x);
                LinRT::sampler smp;
                smp.start(pix, static_cast<uint32_t>(s));
                const auto off = smp.next_2d();
                const auto u = (static_cast<float>(x) + off.x) / static_cast<float>(W - 1);
                const auto v = (static_cast<float>(y) + off.y) / static_cast<float>(H - 1);
                auto r = S.world->sensor().get_ray(u, v);
                const color L = integ.radiance(r, *S.world, smp);
                float* o = per_sample + ((size_t)pix * spp + s) * 3;
                o[0] = L.x; o[1] = L.y; o[2] = L.z;
            }
    });
    return 0;
}

// scene::bounding_sphere_radius and has_delta_bsdf (scene.h:41-57)
int ref_scene_bounds(void* h, float* radius, int* has_delta) {
    ref_scene& S = *(ref_scene*)h;
    *radius = S.world->bounding_sphere_radius();
    *has_delta = S.world->has_delta_bsdf() ? 1 : 0;
    return 0;
}

}  // extern "C"
