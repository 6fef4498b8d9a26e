// Seeded scene generators for the two path-tracer workloads:
//   box_icosphere:    closed box room [-1,1]^3 of 12 axis-aligned quads
//                     (24 tris), an icosphere of `subdivisions` levels
//                     (20*4^L tris, radius 0.3, centred), a 0.25 x 0.25
//                     ceiling rectangle light of radiance 17; Lambertian
//                     albedo ~ U[0.2,0.8]^3 per object; pinhole 40 deg.
//   displaced_sphere: UV sphere of segments x rings x 2 tris, radially
//                     displaced by 5% fBm value noise (8 octaves), split into
//                     7 seeded Voronoi regions with one material each,
//                     a ground quad and a Hosek sky.
// All randomness is Philox-4x32-10 keyed by the seed.
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>

#include "host.hpp"

namespace lrt {
namespace {

struct Rng {
    uint32_t key[2];
    uint32_t stream;
    uint32_t counter = 0;
    Rng(uint64_t seed, uint32_t s) : stream(s) { key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32); }
    float next() {
        const uint32_t c[4] = {counter++, stream, 0x5ce4e5u, 0u};
        uint32_t o[4];
        philox(c, key, o);
        return (float)(o[0] >> 8) * (1.0f / 16777216.0f);
    }
};

struct Mesh {
    std::vector<float> pos;      // 3 per vertex
    std::vector<uint32_t> idx;   // 3 per triangle
    std::vector<float> cn;       // 9 per triangle
    std::vector<uint16_t> tmat;  // per triangle (optional)
};

void add_quad(Mesh& m, const float* a, const float* b, const float* c, const float* d, const float* n) {
    const uint32_t base = (uint32_t)(m.pos.size() / 3);
    for (const float* p : {a, b, c, d}) m.pos.insert(m.pos.end(), p, p + 3);
    const uint32_t t[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
    for (int k = 0; k < 6; ++k) {
        m.idx.push_back(t[k]);
        m.cn.insert(m.cn.end(), n, n + 3);
    }
}

void icosphere(Mesh& m, int levels, float radius) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    std::vector<float> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                            0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2,
                               10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5,
                               2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto normalize_v = [&](uint32_t i) {
        float* p = &v[3 * (size_t)i];
        const float l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        p[0] /= l; p[1] /= l; p[2] /= l;
    };
    for (uint32_t i = 0; i < 12; ++i) normalize_v(i);
    for (int l = 0; l < levels; ++l) {
        std::vector<uint32_t> nf;
        // edge midpoints, keyed by the lower vertex id
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj(v.size() / 3);
        auto mid2 = [&](uint32_t a, uint32_t b) {
            const uint32_t lo = std::min(a, b), hi = std::max(a, b);
            for (auto& e : adj[lo])
                if (e.first == hi) return e.second;
            const uint32_t id = (uint32_t)(v.size() / 3);
            for (int c = 0; c < 3; ++c) v.push_back(0.5f * (v[3 * (size_t)a + c] + v[3 * (size_t)b + c]));
            normalize_v(id);
            adj[lo].emplace_back(hi, id);
            return id;
        };
        for (size_t i = 0; i < f.size(); i += 3) {
            const uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            const uint32_t ab = mid2(a, b), bc = mid2(b, c), ca = mid2(c, a);
            const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    const uint32_t base = (uint32_t)(m.pos.size() / 3);
    for (size_t i = 0; i < v.size(); ++i) m.pos.push_back(v[i] * radius);
    for (uint32_t k : f) {
        m.idx.push_back(base + k);
        m.cn.insert(m.cn.end(), &v[3 * (size_t)k], &v[3 * (size_t)k] + 3);  // unit position = smooth normal
    }
}

// 3D value noise on an integer lattice (hashed with Philox), quintic fade.
struct ValueNoise {
    uint32_t key[2];
    explicit ValueNoise(uint64_t seed) { key[0] = (uint32_t)(seed * 0x9E3779B97F4A7C15ull); key[1] = (uint32_t)(seed >> 32) ^ 0x6a09e667u; }
    float lattice(int x, int y, int z, int oct) const {
        const uint32_t c[4] = {(uint32_t)x, (uint32_t)y, (uint32_t)z, (uint32_t)oct};
        uint32_t o[4];
        philox(c, key, o);
        return (float)(o[0] >> 8) * (2.0f / 16777216.0f) - 1.0f;
    }
    static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
    float eval(float x, float y, float z, int oct) const {
        const float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        const float u = fade(x - fx), v = fade(y - fy), w = fade(z - fz);
        float r = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
            const float wt = (dx ? u : 1.0f - u) * (dy ? v : 1.0f - v) * (dz ? w : 1.0f - w);
            r += wt * lattice(ix + dx, iy + dy, iz + dz, oct);
        }
        return r;
    }
    float fbm(float x, float y, float z) const {  // 8 octaves, lacunarity 2, gain 0.5, in [-1, 1]
        float sum = 0.0f, amp = 1.0f, norm = 0.0f, f = 2.0f;
        for (int o = 0; o < 8; ++o) {
            sum += amp * eval(x * f, y * f, z * f, o);
            norm += amp;
            amp *= 0.5f;
            f *= 2.0f;
        }
        return sum / norm;
    }
};

void displaced_sphere(Mesh& m, uint32_t segs, uint32_t rings, uint64_t seed, const float (*sites)[3], int n_sites) {
    // vertices: two poles + (rings-1) latitude rows of `segs` vertices, the
    // seam shared; every quad gives 2 triangles, so pole quads carry one
    // zero-area triangle each (rejected by the triangle test)
    const uint32_t nv = 2 + (rings - 1) * segs;
    m.pos.assign(3 * (size_t)nv, 0.0f);
    const ValueNoise noise(seed);
    const double pi = 3.14159265358979323846;
    auto vid = [&](uint32_t ring, uint32_t s) -> uint32_t {
        if (ring == 0) return 0;
        if (ring == rings) return 1;
        return 2 + (ring - 1) * segs + (s % segs);
    };
    std::vector<float> unit(3 * (size_t)nv);
    auto set_unit = [&](uint32_t id, double th, double ph) {
        unit[3 * (size_t)id] = (float)(std::sin(th) * std::cos(ph));
        unit[3 * (size_t)id + 1] = (float)std::cos(th);
        unit[3 * (size_t)id + 2] = (float)(std::sin(th) * std::sin(ph));
    };
    set_unit(0, 0.0, 0.0);
    set_unit(1, pi, 0.0);
    for (uint32_t r = 1; r < rings; ++r)
        for (uint32_t s = 0; s < segs; ++s) set_unit(vid(r, s), pi * r / rings, 2.0 * pi * s / segs);
    const unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            for (uint32_t i = t; i < nv; i += nt) {
                const float* u = &unit[3 * (size_t)i];
                const float r = 1.0f + 0.05f * noise.fbm(u[0], u[1], u[2]);
                for (int c = 0; c < 3; ++c) m.pos[3 * (size_t)i + c] = u[c] * r;
            }
        });
    for (auto& th : pool) th.join();
    const size_t ntri = (size_t)segs * rings * 2;
    m.idx.reserve(3 * ntri);
    for (uint32_t r = 0; r < rings; ++r)
        for (uint32_t s = 0; s < segs; ++s) {
            const uint32_t a = vid(r, s), b = vid(r, s + 1), c = vid(r + 1, s), d = vid(r + 1, s + 1);
            // outward winding: (a, b, d) and (a, d, c)
            const uint32_t t[6] = {a, b, d, a, d, c};
            m.idx.insert(m.idx.end(), t, t + 6);
        }
    // area-weighted vertex normals
    std::vector<double> vn(3 * (size_t)nv, 0.0);
    for (size_t t = 0; t < ntri; ++t) {
        const float* p0 = &m.pos[3 * (size_t)m.idx[3 * t]];
        const float* p1 = &m.pos[3 * (size_t)m.idx[3 * t + 1]];
        const float* p2 = &m.pos[3 * (size_t)m.idx[3 * t + 2]];
        const double e1[3] = {(double)p1[0] - p0[0], (double)p1[1] - p0[1], (double)p1[2] - p0[2]};
        const double e2[3] = {(double)p2[0] - p0[0], (double)p2[1] - p0[1], (double)p2[2] - p0[2]};
        const double g[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) vn[3 * (size_t)m.idx[3 * t + k] + c] += g[c];
    }
    m.cn.resize(9 * ntri);
    m.tmat.resize(ntri);
    for (size_t t = 0; t < ntri; ++t) {
        double cen[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = m.idx[3 * t + k];
            const double l = std::sqrt(vn[3 * (size_t)v] * vn[3 * (size_t)v] + vn[3 * (size_t)v + 1] * vn[3 * (size_t)v + 1] +
                                       vn[3 * (size_t)v + 2] * vn[3 * (size_t)v + 2]);
            for (int c = 0; c < 3; ++c) {
                m.cn[9 * t + 3 * k + c] = l > 0 ? (float)(vn[3 * (size_t)v + c] / l) : 0.0f;
                cen[c] += unit[3 * (size_t)v + c];
            }
        }
        int best = 0;
        double bd = -1e30;
        for (int s = 0; s < n_sites; ++s) {
            const double d = cen[0] * sites[s][0] + cen[1] * sites[s][1] + cen[2] * sites[s][2];
            if (d > bd) { bd = d; best = s; }
        }
        m.tmat[t] = (uint16_t)best;
    }
}

lrt_material blank() {
    lrt_material m;
    std::memset(&m, 0, sizeof(m));
    m.nested = -1;
    m.table_ext = m.table_int = -1;
    for (int c = 0; c < 3; ++c) {
        m.specular_reflectance[c] = 1.0f;
        m.specular_transmittance[c] = 1.0f;
        m.diffuse[c] = 0.5f;
    }
    m.thickness = 1.0f;
    m.int_ior = 1.5f;
    m.ext_ior = 1.0f;
    return m;
}

}  // namespace

struct Generated {
    std::vector<Mesh> meshes;
    std::vector<lrt_object_desc> objects;
    std::vector<lrt_material> materials;
    std::vector<lrt_light_desc> lights;
    std::vector<float> tables;
    lrt_scene_desc desc;
};

Generated* generate(const lrt_gen_params& p, const float* tables, uint32_t n_tables, const lrt_light_desc* sky) {
    auto g = std::make_unique<Generated>();
    std::memset(&g->desc, 0, sizeof(g->desc));
    if (tables && n_tables) g->tables.assign(tables, tables + (size_t)n_tables * 101);
    auto obj_for_mesh = [](const Mesh& m, int material) {
        lrt_object_desc o;
        std::memset(&o, 0, sizeof(o));
        o.shape = LRT_SHAPE_MESH;
        o.material = material;
        o.emitter = -1;
        o.n_vertices = (uint32_t)(m.pos.size() / 3);
        o.n_triangles = (uint32_t)(m.idx.size() / 3);
        o.positions = m.pos.data();
        o.indices = m.idx.data();
        o.corner_normals = m.cn.data();
        o.triangle_material = m.tmat.empty() ? nullptr : m.tmat.data();
        return o;
    };
    lrt_camera_desc& cam = g->desc.camera;
    // each face of [-1,1]^3 as two quads, normals facing inward (12 quads)
    auto build_room = [](Mesh& room) {
        for (int axis = 0; axis < 3; ++axis)
            for (int side = -1; side <= 1; side += 2) {
                const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
                float n[3] = {0, 0, 0};
                n[axis] = (float)-side;
                for (int half = 0; half < 2; ++half) {
                    float q[4][3];
                    const float u0 = half == 0 ? -1.0f : 0.0f, u1 = half == 0 ? 0.0f : 1.0f;
                    const float uv[4][2] = {{u0, -1.0f}, {u1, -1.0f}, {u1, 1.0f}, {u0, 1.0f}};
                    for (int k = 0; k < 4; ++k) {
                        q[k][axis] = (float)side;
                        q[k][a1] = uv[k][0];
                        q[k][a2] = uv[k][1];
                    }
                    // inward winding: cross(e1, e2) must point along n
                    const float e1[3] = {q[1][0] - q[0][0], q[1][1] - q[0][1], q[1][2] - q[0][2]};
                    const float e2[3] = {q[2][0] - q[0][0], q[2][1] - q[0][1], q[2][2] - q[0][2]};
                    const float c = e1[(axis + 1) % 3] * e2[(axis + 2) % 3] - e1[(axis + 2) % 3] * e2[(axis + 1) % 3];
                    if (c * n[axis] > 0) add_quad(room, q[0], q[1], q[2], q[3], n);
                    else add_quad(room, q[0], q[3], q[2], q[1], n);
                }
            }
    };
    // ceiling rectangle light of edge `size` just below y = 1, facing down
    auto ceiling_light = [&](float size, float radiance, int object) {
        lrt_object_desc L;
        std::memset(&L, 0, sizeof(L));
        L.shape = LRT_SHAPE_RECTANGLE;
        L.material = -1;
        L.emitter = (int)g->lights.size();
        const float corner[3] = {-0.5f * size, 0.999f, -0.5f * size}, eu[3] = {size, 0.0f, 0.0f}, ev[3] = {0.0f, 0.0f, size};
        std::memcpy(L.corner, corner, 12);
        std::memcpy(L.edge_u, eu, 12);
        std::memcpy(L.edge_v, ev, 12);
        g->objects.push_back(L);
        lrt_light_desc ld;
        std::memset(&ld, 0, sizeof(ld));
        ld.kind = LRT_LIGHT_AREA;
        ld.object = object;
        ld.radiance[0] = ld.radiance[1] = ld.radiance[2] = radiance;
        g->lights.push_back(ld);
    };
    // 7 seeded Voronoi sites on the sphere and the region materials:
    // 0..6 = diffuse, rough dielectric, rough conductor, plastic, rough plastic,
    // coating, rough coating (README.md:34 order); 7 = the rough glass nested
    // in the rough coating; 8 = grey diffuse.  Tables 0/1/2 = rough plastic
    // external/internal and rough coating slices.
    auto region_setup = [&](Rng& rng, float sites[7][3]) {
        if (g->tables.size() < 3 * 101) throw std::runtime_error("this scene needs 3 rough-transmittance slices");
        for (int s = 0; s < 7; ++s) {
            const float z = 1.0f - 2.0f * rng.next(), ph = 6.28318530718f * rng.next();
            const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
            sites[s][0] = r * std::cos(ph); sites[s][1] = z; sites[s][2] = r * std::sin(ph);
        }
        const float cu_eta[3] = {0.200438f, 0.924033f, 1.102212f}, cu_k[3] = {3.912948f, 2.452847f, 2.142188f};
        lrt_material m0 = blank();
        m0.type = LRT_BSDF_DIFFUSE;
        for (int c = 0; c < 3; ++c) m0.diffuse[c] = 0.2f + 0.6f * rng.next();
        lrt_material m1 = blank();
        m1.type = LRT_BSDF_ROUGH_DIELECTRIC; m1.int_ior = 1.5f; m1.ext_ior = 1.0f; m1.alpha = 0.2f;
        lrt_material m2 = blank();
        m2.type = LRT_BSDF_CONDUCTOR; m2.alpha = 0.1f;
        std::memcpy(m2.cond_eta, cu_eta, 12); std::memcpy(m2.cond_k, cu_k, 12);
        lrt_material m3 = blank();  // src/main.cpp:42
        m3.type = LRT_BSDF_PLASTIC; m3.int_ior = 1.49f; m3.ext_ior = 1.000277f;
        m3.diffuse[0] = 0.8f; m3.diffuse[1] = 0.8f; m3.diffuse[2] = 0.5f;
        m3.fdr_int = fresnel_diffuse_reflectance(1.0f / (m3.int_ior / m3.ext_ior));
        lrt_material m4 = blank();  // src/main.cpp:43 colours
        m4.type = LRT_BSDF_ROUGH_PLASTIC; m4.int_ior = 1.49f; m4.ext_ior = 1.000277f; m4.alpha = 0.3f;
        m4.diffuse[0] = 0.8f; m4.diffuse[1] = 0.1f; m4.diffuse[2] = 0.05f;
        m4.table_ext = 0; m4.table_int = 1;
        lrt_material m5 = blank();  // src/main.cpp:44-45
        m5.type = LRT_BSDF_COATING; m5.int_ior = 1.7f; m5.ext_ior = 1.000277f; m5.nested = 3;
        m5.specular_reflectance[0] = 0.3f; m5.specular_reflectance[1] = 0.6f; m5.specular_reflectance[2] = 0.6f;
        lrt_material m6 = blank();  // src/main.cpp:41,46-47
        m6.type = LRT_BSDF_ROUGH_COATING; m6.int_ior = 1.7f; m6.ext_ior = 1.000277f; m6.alpha = 0.25f; m6.nested = 7;
        m6.table_ext = 2;
        lrt_material m7 = blank();
        m7.type = LRT_BSDF_ROUGH_DIELECTRIC; m7.int_ior = 1.5f; m7.ext_ior = 1.0f; m7.alpha = 0.05f;
        lrt_material m8 = blank();
        m8.type = LRT_BSDF_DIFFUSE;
        m8.diffuse[0] = m8.diffuse[1] = m8.diffuse[2] = 0.5f;
        if (p.kind == LRT_GEN_DISPLACED_SPHERE_VARIANT) {
            // colours, the conductor's and the rough glass's roughness and the
            // glass's index drawn from the seed (the rough plastic and rough
            // coating keep their table-bound alpha and eta)
            Rng vr(p.seed, 4);
            for (int c = 0; c < 3; ++c) {
                m0.diffuse[c] = 0.05f + 0.85f * vr.next();
                m3.diffuse[c] = 0.05f + 0.85f * vr.next();
                m4.diffuse[c] = 0.05f + 0.85f * vr.next();
                m5.specular_reflectance[c] = 0.2f + 0.8f * vr.next();
                m8.diffuse[c] = 0.2f + 0.6f * vr.next();
            }
            m1.alpha = 0.05f + 0.35f * vr.next();
            m1.int_ior = 1.3f + 0.5f * vr.next();
            m2.alpha = 0.03f + 0.3f * vr.next();
            m3.fdr_int = fresnel_diffuse_reflectance(1.0f / (m3.int_ior / m3.ext_ior));
        }
        for (const auto& m : {m0, m1, m2, m3, m4, m5, m6, m7, m8}) g->materials.push_back(m);
    };
    if (p.kind == LRT_GEN_BOX_ICOSPHERE) {
        Rng rng(p.seed, 1);
        g->meshes.resize(2);
        build_room(g->meshes[0]);
        icosphere(g->meshes[1], (int)p.subdivisions, 0.3f);
        for (int i = 0; i < 2; ++i) {
            lrt_material m = blank();
            m.type = LRT_BSDF_DIFFUSE;
            for (int c = 0; c < 3; ++c) m.diffuse[c] = 0.2f + 0.6f * rng.next();
            g->materials.push_back(m);
            g->objects.push_back(obj_for_mesh(g->meshes[(size_t)i], i));
        }
        ceiling_light(0.25f, 17.0f, 2);
        const float o[3] = {0.0f, 0.0f, 0.95f}, la[3] = {0.0f, -0.05f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
        std::memcpy(cam.origin, o, 12);
        std::memcpy(cam.look_at, la, 12);
        std::memcpy(cam.up, up, 12);
        cam.fov_deg = 40.0f;
        cam.aspect = p.aspect > 0 ? p.aspect : 1.0f;
    } else if (p.kind == LRT_GEN_DISPLACED_SPHERE || p.kind == LRT_GEN_DISPLACED_SPHERE_VARIANT) {
        if (!sky) throw std::runtime_error("displaced sphere scene needs a sky light");
        Rng rng(p.seed, 2);
        float sites[7][3];
        region_setup(rng, sites);
        g->meshes.resize(2);
        displaced_sphere(g->meshes[0], p.segments ? p.segments : 1024, p.rings ? p.rings : 512, p.seed, sites, 7);
        // ground quad (2 triangles) below the sphere
        Mesh& ground = g->meshes[1];
        const float y = -1.06f, e = 20.0f, n[3] = {0.0f, 1.0f, 0.0f};
        const float a[3] = {-e, y, -e}, b[3] = {-e, y, e}, c[3] = {e, y, e}, d[3] = {e, y, -e};
        add_quad(ground, a, b, c, d, n);
        g->objects.push_back(obj_for_mesh(g->meshes[0], 0));
        g->objects.push_back(obj_for_mesh(g->meshes[1], 8));
        g->lights.push_back(*sky);
        const float o[3] = {0.0f, 0.45f, 3.4f}, la[3] = {0.0f, -0.15f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
        std::memcpy(cam.origin, o, 12);
        std::memcpy(cam.look_at, la, 12);
        std::memcpy(cam.up, up, 12);
        cam.fov_deg = 40.0f;
        cam.aspect = p.aspect > 0 ? p.aspect : 16.0f / 9.0f;
    } else if (p.kind == LRT_GEN_INVERSE_SPHERE) {
        // the differentiable renderer's scene: the displaced sphere in one
        // diffuse material (0, the textured one), a rough copper ground (2)
        // that mirrors its underside, and a rough glass ball (1) in front and
        // to the right, under the Hosek sky
        if (!sky) throw std::runtime_error("inverse-rendering scene needs a sky light");
        const float cu_eta[3] = {0.200438f, 0.924033f, 1.102212f}, cu_k[3] = {3.912948f, 2.452847f, 2.142188f};
        lrt_material m0 = blank();
        m0.type = LRT_BSDF_DIFFUSE;
        m0.diffuse[0] = m0.diffuse[1] = m0.diffuse[2] = 0.5f;
        lrt_material m1 = blank();
        m1.type = LRT_BSDF_ROUGH_DIELECTRIC; m1.int_ior = 1.5f; m1.ext_ior = 1.0f; m1.alpha = 0.2f;
        lrt_material m2 = blank();
        m2.type = LRT_BSDF_CONDUCTOR; m2.alpha = 0.1f;
        std::memcpy(m2.cond_eta, cu_eta, 12); std::memcpy(m2.cond_k, cu_k, 12);
        for (const auto& m : {m0, m1, m2}) g->materials.push_back(m);
        float sites[1][3] = {{0.0f, 1.0f, 0.0f}};
        g->meshes.resize(2);
        displaced_sphere(g->meshes[0], p.segments ? p.segments : 512, p.rings ? p.rings : 256, p.seed, sites, 1);
        g->meshes[0].tmat.clear();
        Mesh& ground = g->meshes[1];
        const float y = -1.06f, e = 20.0f, n[3] = {0.0f, 1.0f, 0.0f};
        const float a[3] = {-e, y, -e}, b[3] = {-e, y, e}, c[3] = {e, y, e}, d[3] = {e, y, -e};
        add_quad(ground, a, b, c, d, n);
        g->objects.push_back(obj_for_mesh(g->meshes[0], 0));
        g->objects.push_back(obj_for_mesh(g->meshes[1], 2));
        lrt_object_desc ball;
        std::memset(&ball, 0, sizeof(ball));
        ball.shape = LRT_SHAPE_SPHERE;
        ball.material = 1;
        ball.emitter = -1;
        ball.center[0] = 1.45f; ball.center[1] = -0.70f; ball.center[2] = 0.9f;
        ball.radius = 0.35f;
        g->objects.push_back(ball);
        g->lights.push_back(*sky);
        const float o[3] = {0.3f, 0.55f, 4.9f}, la[3] = {0.3f, -0.3f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
        std::memcpy(cam.origin, o, 12);
        std::memcpy(cam.look_at, la, 12);
        std::memcpy(cam.up, up, 12);
        cam.fov_deg = 40.0f;
        cam.aspect = p.aspect > 0 ? p.aspect : 1.0f;
    } else if (p.kind == LRT_GEN_PHOTON_BOX) {
        // the box room around the displaced sphere (radius 0.5, 7 regions),
        // lit by a 0.5 x 0.5 ceiling light
        Rng rng(p.seed, 3);
        float sites[7][3];
        region_setup(rng, sites);
        g->meshes.resize(2);
        build_room(g->meshes[0]);
        displaced_sphere(g->meshes[1], p.segments ? p.segments : 512, p.rings ? p.rings : 256, p.seed, sites, 7);
        for (size_t i = 0; i < g->meshes[1].pos.size(); i += 3) {
            g->meshes[1].pos[i] *= 0.5f;
            g->meshes[1].pos[i + 1] = g->meshes[1].pos[i + 1] * 0.5f - 0.45f;
            g->meshes[1].pos[i + 2] *= 0.5f;
        }
        lrt_material wall = blank();
        wall.type = LRT_BSDF_DIFFUSE;
        for (int c = 0; c < 3; ++c) wall.diffuse[c] = 0.2f + 0.6f * rng.next();
        g->materials.push_back(wall);  // material 9
        g->objects.push_back(obj_for_mesh(g->meshes[0], 9));
        g->objects.push_back(obj_for_mesh(g->meshes[1], 0));
        ceiling_light(0.5f, 10.0f, 2);
        const float o[3] = {0.0f, 0.0f, 0.95f}, la[3] = {0.0f, -0.3f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
        std::memcpy(cam.origin, o, 12);
        std::memcpy(cam.look_at, la, 12);
        std::memcpy(cam.up, up, 12);
        cam.fov_deg = 60.0f;
        cam.aspect = p.aspect > 0 ? p.aspect : 1.0f;
    } else {
        throw std::runtime_error("unknown generator kind");
    }
    g->desc.n_objects = (uint32_t)g->objects.size();
    g->desc.objects = g->objects.data();
    g->desc.n_materials = (uint32_t)g->materials.size();
    g->desc.materials = g->materials.data();
    g->desc.n_lights = (uint32_t)g->lights.size();
    g->desc.lights = g->lights.data();
    g->desc.n_tables = (uint32_t)(g->tables.size() / 101);
    g->desc.tables = g->tables.empty() ? nullptr : g->tables.data();
    return g.release();
}

// The scene of src/main.cpp:55-87 around a mesh loaded from an OBJ file:
// the mesh with materials[0] (the other given materials are those it nests),
// a ground sphere of radius 100 at (0, -100.5, 0) with diffuse albedo
// (0.8, 0.8, 0) (the material after the given ones), a 1.2 x 1.2 rectangle
// light of radiance 12 at y = 1.5 facing down, and the pinhole camera at
// (0, 0.5, 3) looking at the origin, 60 degrees.
Generated* generate_obj_scene(const char* path, const lrt_material* materials, uint32_t n_materials,
                              const float* tables, uint32_t n_tables, float aspect) {
    auto g = std::make_unique<Generated>();
    std::memset(&g->desc, 0, sizeof(g->desc));
    if (tables && n_tables) g->tables.assign(tables, tables + (size_t)n_tables * 101);
    g->meshes.resize(1);
    Mesh& m = g->meshes[0];
    std::string err;
    if (!load_obj(path, m.pos, m.idx, m.cn, err)) throw std::invalid_argument("load_obj: " + err);
    g->materials.assign(materials, materials + n_materials);
    lrt_material ground = blank();
    ground.type = LRT_BSDF_DIFFUSE;
    ground.diffuse[0] = 0.8f;
    ground.diffuse[1] = 0.8f;
    ground.diffuse[2] = 0.0f;
    g->materials.push_back(ground);
    lrt_object_desc o;
    std::memset(&o, 0, sizeof(o));
    o.shape = LRT_SHAPE_MESH;
    o.material = 0;
    o.emitter = -1;
    o.n_vertices = (uint32_t)(m.pos.size() / 3);
    o.n_triangles = (uint32_t)(m.idx.size() / 3);
    o.positions = m.pos.data();
    o.indices = m.idx.data();
    o.corner_normals = m.cn.data();
    g->objects.push_back(o);
    std::memset(&o, 0, sizeof(o));
    o.shape = LRT_SHAPE_SPHERE;
    o.material = (int32_t)n_materials;
    o.emitter = -1;
    o.center[1] = -100.5f;
    o.radius = 100.0f;
    g->objects.push_back(o);
    std::memset(&o, 0, sizeof(o));
    o.shape = LRT_SHAPE_RECTANGLE;
    o.material = -1;
    o.emitter = 0;
    const float corner[3] = {-0.6f, 1.5f, -0.6f}, eu[3] = {1.2f, 0.0f, 0.0f}, ev[3] = {0.0f, 0.0f, 1.2f};
    std::memcpy(o.corner, corner, 12);
    std::memcpy(o.edge_u, eu, 12);
    std::memcpy(o.edge_v, ev, 12);
    g->objects.push_back(o);
    lrt_light_desc ld;
    std::memset(&ld, 0, sizeof(ld));
    ld.kind = LRT_LIGHT_AREA;
    ld.object = 2;
    ld.radiance[0] = ld.radiance[1] = ld.radiance[2] = 12.0f;
    g->lights.push_back(ld);
    lrt_camera_desc& cam = g->desc.camera;
    const float org[3] = {0.0f, 0.5f, 3.0f}, la[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
    std::memcpy(cam.origin, org, 12);
    std::memcpy(cam.look_at, la, 12);
    std::memcpy(cam.up, up, 12);
    cam.fov_deg = 60.0f;
    cam.aspect = aspect > 0.0f ? aspect : 1.0f;
    g->desc.n_objects = (uint32_t)g->objects.size();
    g->desc.objects = g->objects.data();
    g->desc.n_materials = (uint32_t)g->materials.size();
    g->desc.materials = g->materials.data();
    g->desc.n_lights = (uint32_t)g->lights.size();
    g->desc.lights = g->lights.data();
    g->desc.n_tables = (uint32_t)(g->tables.size() / 101);
    g->desc.tables = g->tables.empty() ? nullptr : g->tables.data();
    return g.release();
}

void free_generated(Generated* g) { delete g; }
const lrt_scene_desc* generated_desc(const Generated* g) { return &g->desc; }

}  // namespace lrt
