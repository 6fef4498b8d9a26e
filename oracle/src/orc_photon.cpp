// Oracle (test infrastructure only): photon mapping restated from the
// reference: emitter rays (scene.h:113-142, area_emitter.h:93-122), particle
// tracing (photon_tracer.h:23-96), the photon quota in particle order
// (photon_gather.h:15-101, single-threaded order), the median kd photon map
// and its kNN irradiance estimate (photon_map.h:26-165), and the eye pass
// photon_integrator::li (photon_integrator.h:67-296).
#include <algorithm>
#include <cstring>
#include <queue>
#include <thread>

#include "orc.h"

namespace orc {

uint32_t orc_sobol_bits(uint32_t index, uint32_t dim);

namespace {
uint32_t mix_bits(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
float luminance(const v3& c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
}  // namespace

// particle sampler: Philox stream {index, 0xF0000000 + kind, d/4, 0}, or the
// reference sampler started at (pixel 0, sample index) (photon_gather.h:45)
struct psampler {
    int mode = 0;
    uint32_t index = 0, tag = 0, dim = 0;
    uint32_t key[2] = {0, 0};
    uint32_t pixel_seed = 0;
    float next() {
        const uint32_t d = dim++;
        if (mode == 1) {
            const uint32_t bits = orc_sobol_bits(index, d) ^ mix_bits(pixel_seed + d * 0x9e3779b9u);
            return (float)bits * (1.0f / 4294967296.0f);
        }
        const uint32_t ctr[4] = {index, tag, d >> 2, 0u};
        uint32_t o[4];
        philox4x32_10(ctr, key, o);
        return u24(o[d & 3]);
    }
    // vec2(next_1d(), next_1d()) as g++ compiles it (arguments right to left)
    void next2(float& x, float& y) {
        if (mode == 1) { y = next(); x = next(); }
        else { x = next(); y = next(); }
    }
};

struct emitter_ray {
    ray r;
    v3 n, weight;
    float pdf_pos = 0, pdf_dir = 0, sel = 1;
    int emitter = -1;
};

// area_emitter::sample_ray (area_emitter.h:93-122)
static bool area_sample_ray(const scene& S, const light& L, float sx, float sy, float dx, float dy, emitter_ray& out) {
    const object& shp = S.objects[(size_t)L.object];
    if (shp.area <= 0.0f) return false;
    v3 p, n;
    float pdf_area = 0.0f;
    shp.sample_position(sx, sy, p, n, pdf_area);
    if (pdf_area <= 0.0f) return false;
    const float ct = std::sqrt(dx);
    const float st = std::sqrt(1.0f - dx);
    const float phi = 2.0f * kPi * dy;
    const v3 local(std::cos(phi) * st, std::sin(phi) * st, ct);
    onb f;
    f.build(n);
    out.r = ray{p, f.to_world(local)};
    out.n = n;
    out.pdf_pos = pdf_area;
    out.pdf_dir = ct * (1.0f / kPi);
    out.weight = L.radiance * (kPi * shp.area);  // power(), area_emitter.h:141-147
    out.sel = 1.0f;
    return true;
}

// scene::sample_emitter_ray (scene.h:113-142): light chosen by luminance(power)
static bool sample_emitter_ray(const scene& S, float sx, float sy, float dx, float dy, emitter_ray& out) {
    std::vector<size_t> index;
    std::vector<float> cdf;
    float sum = 0.0f;
    for (size_t i = 0; i < S.lights.size(); ++i) {
        const light& L = S.lights[i];
        if (L.kind != LRT_LIGHT_AREA) continue;
        const object& shp = S.objects[(size_t)L.object];
        const v3 power = shp.area <= 0.0f ? v3(0.0f) : L.radiance * (kPi * shp.area);
        const float w = luminance(power);
        if (w <= 0.0f) continue;
        sum += w;
        index.push_back(i);
        cdf.push_back(sum);
    }
    if (sum <= 0.0f || cdf.empty()) return false;
    const float target = sx * sum;
    size_t bin = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), target) - cdf.begin());
    if (bin >= cdf.size()) bin = cdf.size() - 1;
    const float prev = bin == 0 ? 0.0f : cdf[bin - 1];
    const float span = cdf[bin] - prev;
    if (span <= 0.0f) return false;
    const float sel = span / sum;
    const float u = (target - prev) / span;
    if (!area_sample_ray(S, S.lights[index[bin]], u, sy, dx, dy, out)) return false;
    out.weight = v3(out.weight.x / sel, out.weight.y / sel, out.weight.z / sel);
    out.sel = sel;
    out.emitter = (int)index[bin];
    return true;
}

struct photon {
    v3 pos, dir, normal, power;
    uint16_t depth;
    uint32_t particle;
};

// photon_tracer::trace_particle + store_if_accepted (photon_tracer.h:23-96);
// kind 0 = surface (global) map, 1 = caustic map
static void trace_particle(const scene& S, int kind, psampler& smp, int max_depth, int rr_depth, std::vector<photon>& out) {
    emitter_ray em;
    float sx, sy, dx, dy;
    if (smp.mode == 1) {  // g++ evaluates the two next_2d() arguments right to left
        smp.next2(dx, dy);
        smp.next2(sx, sy);
    } else {
        smp.next2(sx, sy);
        smp.next2(dx, dy);
    }
    if (!sample_emitter_ray(S, sx, sy, dx, dy, em)) return;
    ray cur = em.r;
    v3 T(1.0f);
    bool delta = false;
    const float* tables = S.tables.empty() ? nullptr : S.tables.data();
    for (int depth = 1; !T.is_zero() && depth <= max_depth;) {
        hit_rec rec;
        if (!S.hit(cur, 0.001f, kInf, rec) || rec.material < 0) break;
        const uint32_t flags = bsdf_flags(S.materials, rec.material);
        const bool smooth = (flags & LRT_FLAG_DIFFUSE) != 0 || (flags & LRT_FLAG_GLOSSY) != 0;
        if (smooth && depth > 1 && (kind == 1 ? delta : !delta)) {
            photon ph;
            ph.pos = rec.p;
            ph.dir = normalize(cur.d);
            ph.normal = rec.normal;
            ph.power = T * em.weight;
            ph.depth = (uint16_t)depth;
            ph.particle = smp.index;
            out.push_back(ph);
        }
        onb basis;
        basis.build(rec.normal);
        bsdf_ctx ctx;
        ctx.mode = LRT_MODE_IMPORTANCE;
        ctx.lobe_sample = smp.next();
        bsdf_rec srec;
        srec.wi = basis.to_local(-normalize(cur.d));
        srec.front_face = rec.front_face;
        float ux, uy;
        smp.next2(ux, uy);
        const v3 w = bsdf_sample(S.materials, tables, rec.material, ctx, srec, ux, uy);
        if (w.is_zero()) break;
        const v3 outgoing = basis.to_world(srec.wo);
        const v3 incoming = basis.to_world(srec.wi);
        if (dot(rec.normal, incoming) * srec.wi.z <= 0.0f || dot(rec.normal, outgoing) * srec.wo.z <= 0.0f) break;
        T = T * w;
        delta = (srec.sampled_type & LRT_FLAG_SPECULAR) != 0;
        cur = ray{rec.p, outgoing};
        if (depth++ >= rr_depth) {
            const float surv = std::min(std::max({T.x, T.y, T.z}), 0.95f);
            if (smp.next() >= surv) break;
            T = v3(T.x / surv, T.y / surv, T.z / surv);
        }
    }
}

// photon_map (photon_map.h:26-165)
struct pmap {
    struct node { int left = -1, right = -1, start = 0, count = 0, axis = 0; float split = 0.0f; bool leaf = false; };
    std::vector<photon> ph;
    std::vector<node> nodes;
    float scale = 1.0f;
    int build_range(int start, int count) {
        const int index = (int)nodes.size();
        nodes.push_back(node{});
        if (count <= 8) {
            nodes[(size_t)index].leaf = true;
            nodes[(size_t)index].start = start;
            nodes[(size_t)index].count = count;
            return index;
        }
        box b;
        for (int i = 0; i < count; ++i) b.grow(ph[(size_t)(start + i)].pos);
        const v3 e = b.hi - b.lo;
        int axis = 0;
        if (e.y > e[axis]) axis = 1;
        if (e.z > e[axis]) axis = 2;
        const int mid = start + count / 2;
        std::nth_element(ph.begin() + start, ph.begin() + mid, ph.begin() + start + count,
                         [axis](const photon& a, const photon& c) { return a.pos[axis] < c.pos[axis]; });
        const int l = build_range(start, mid - start);
        const int r = build_range(mid, start + count - mid);
        nodes[(size_t)index].axis = axis;
        nodes[(size_t)index].split = ph[(size_t)mid].pos[axis];
        nodes[(size_t)index].left = l;
        nodes[(size_t)index].right = r;
        return index;
    }
    void build(std::vector<photon> g, float s) {
        ph = std::move(g);
        scale = s;
        nodes.clear();
        if (!ph.empty()) build_range(0, (int)ph.size());
    }

    // The library's device tree (no counterpart in the reference): the shape
    // and axis rule of build_range, but a node's left child takes the
    // floor(c / 2) photons first in (coordinate, input index) order on the
    // axis, its split is the coordinate of the right child's first photon in
    // that order (the median), and a leaf lists its photons in (x, input index)
    // order.  Coordinates compare by their order-preserving key (-0 before +0).
    static uint32_t fkey(float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
    struct by_axis {
        const std::vector<photon>* g;
        int a;
        bool operator()(uint32_t i, uint32_t j) const {
            const uint32_t ki = fkey((*g)[i].pos[a]), kj = fkey((*g)[j].pos[a]);
            return ki != kj ? ki < kj : i < j;
        }
    };
    int median_range(const std::vector<photon>& g, const std::vector<uint32_t>& xs, std::vector<uint32_t>& order,
                     std::vector<uint8_t>& mark) {
        const int index = (int)nodes.size();
        nodes.push_back(node{});
        const int c = (int)xs.size();
        if (c <= 8) {
            nodes[(size_t)index].leaf = true;
            nodes[(size_t)index].start = (int)order.size();
            nodes[(size_t)index].count = c;
            order.insert(order.end(), xs.begin(), xs.end());
            return index;
        }
        box b;
        for (uint32_t i : xs) b.grow(g[i].pos);
        const v3 e = b.hi - b.lo;
        int axis = 0;
        if (e.y > e[axis]) axis = 1;
        if (e.z > e[axis]) axis = 2;
        std::vector<uint32_t> sorted = xs;
        std::sort(sorted.begin(), sorted.end(), by_axis{&g, axis});
        const int k = c / 2;
        const float split = g[sorted[(size_t)k]].pos[axis];
        for (int i = 0; i < k; ++i) mark[sorted[(size_t)i]] = 1;
        std::vector<uint32_t> lx, rx;
        lx.reserve((size_t)k);
        rx.reserve((size_t)(c - k));
        for (uint32_t i : xs) (mark[i] ? lx : rx).push_back(i);
        for (int i = 0; i < k; ++i) mark[sorted[(size_t)i]] = 0;
        std::vector<uint32_t>().swap(sorted);
        const int l = median_range(g, lx, order, mark);
        const int r = median_range(g, rx, order, mark);
        nodes[(size_t)index].axis = axis;
        nodes[(size_t)index].split = split;
        nodes[(size_t)index].left = l;
        nodes[(size_t)index].right = r;
        return index;
    }
    // tree order: order[i] = input index of the photon at tree position i
    std::vector<uint32_t> tree_order;
    void build_median(std::vector<photon> g, float s) {
        scale = s;
        nodes.clear();
        ph.clear();
        tree_order.clear();
        if (g.empty()) return;
        std::vector<uint32_t> xs(g.size());
        for (size_t i = 0; i < g.size(); ++i) xs[i] = (uint32_t)i;
        std::sort(xs.begin(), xs.end(), by_axis{&g, 0});
        std::vector<uint8_t> mark(g.size(), 0);
        tree_order.reserve(g.size());
        median_range(g, xs, tree_order, mark);
        ph.resize(g.size());
        for (size_t i = 0; i < g.size(); ++i) ph[i] = g[tree_order[i]];
    }
    using nb = std::pair<float, uint32_t>;
    void search(int index, const v3& p, float r2, size_t k, std::priority_queue<nb>& heap) const {
        const node& cur = nodes[(size_t)index];
        float limit = r2;
        if (heap.size() == k) limit = std::min(limit, heap.top().first);
        if (cur.leaf) {
            for (int i = 0; i < cur.count; ++i) {
                const uint32_t pi = (uint32_t)(cur.start + i);
                const float d = (ph[pi].pos - p).length_squared();
                if (d > limit) continue;
                if (heap.size() == k) heap.pop();
                heap.emplace(d, pi);
                if (heap.size() == k) limit = std::min(r2, heap.top().first);
            }
            return;
        }
        const float delta = p[cur.axis] - cur.split;
        const int near = delta <= 0.0f ? cur.left : cur.right;
        const int far = delta <= 0.0f ? cur.right : cur.left;
        search(near, p, r2, k, heap);
        if (delta * delta <= limit) search(far, p, r2, k, heap);
    }
    v3 estimate(const v3& p, const v3& n, float radius, int max_depth, size_t k) const {
        if (ph.empty() || radius <= 0.0f || k == 0) return v3(0.0f);
        const float r2 = radius * radius;
        std::priority_queue<nb> heap;
        search(0, p, r2, k, heap);
        const float inv_r2 = 1.0f / r2;
        v3 sum(0.0f);
        while (!heap.empty()) {
            const nb e = heap.top();
            heap.pop();
            const photon& c = ph[e.second];
            if ((int)c.depth > max_depth) continue;
            const v3 in = -c.dir;
            const float ip = dot(c.normal, in);
            const float is = dot(n, in);
            if (is <= 0.0f || dot(c.normal, n) <= 1e-1f || ip <= 1e-2f) continue;
            const float sq = 1.0f - e.first * inv_r2;
            const float w = (is / ip) * (sq * sq);
            sum = sum + c.power * std::abs(w);
        }
        return sum * (scale * 3.0f * (1.0f / kPi) * inv_r2);
    }
};

static v3 col3(const float* c) { return v3(c[0], c[1], c[2]); }

// photon_gather::run (photon_gather.h:15-101) in particle order with one
// particle per work unit: particles are traced until the photons returned
// reach `requested` (or the gather is unsuccessful, :99-101); the first
// `requested` photons are kept and every traced particle counts in `shot`.
static void gather_quota(const scene& S, int kind, uint32_t requested, int max_depth, int rr_depth, uint64_t seed,
                         std::vector<photon>& out, uint64_t& shot) {
    out.clear();
    shot = 0;
    uint64_t returned = 0;
    for (uint32_t i = 0;; ++i) {
        if (returned >= requested) break;
        if (shot > 100000 && out.size() < requested && (out.empty() || out.size() < shot / 1024)) break;
        psampler smp;
        smp.index = i;
        smp.tag = 0xF0000000u + (uint32_t)kind;
        smp.key[0] = (uint32_t)seed;
        smp.key[1] = (uint32_t)(seed >> 32);
        std::vector<photon> ph;
        trace_particle(S, kind, smp, max_depth, rr_depth, ph);
        shot = i + 1;
        returned += ph.size();
        for (const photon& p : ph)
            if (out.size() < requested) out.push_back(p);
    }
}

// ---- eye pass: photon_integrator::li (photon_integrator.h:67-296) -----------
// The reference's recursive estimator: emission; the photon gathers at
// pure-diffuse hits facing the ray; every specular component recursed while
// depth < 4; direct light by emitter samples and BSDF samples with MIS
// weighted by the sample counts (16 / 16 or 32 at depth 1, else 1); and the
// non-diffuse indirect recursion through the BSDF samples with emission off.
//
// Sampler streams.  mode 1: the reference sampler, one sequence shared by the
// whole recursion in its depth-first order.  mode 0 (the device's scheme):
// every li call draws from its own Philox stream {pixel, sample, d/4, path},
// with path 0 for the camera ray (dimensions from 2 on, after the two pixel
// jitter dimensions) and path = eye_child(parent, code) for a child: code = c
// for specular component c, 16 + i for BSDF sample i.
uint32_t eye_child(uint32_t parent, uint32_t code) {
    uint32_t h = parent * 0x9e3779b9u + (code + 1u) * 0x632be5abu;
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h ? h : 1u;
}

struct estream {
    int mode = 0;
    psampler* shared = nullptr;  // mode 1
    uint32_t pixel = 0, sample = 0, path = 0, dim = 0;
    uint32_t key[2] = {0, 0};
    float next() {
        if (mode == 1) return shared->next();
        const uint32_t d = dim++;
        const uint32_t ctr[4] = {pixel, sample, d >> 2, path};
        uint32_t o[4];
        philox4x32_10(ctr, key, o);
        return u24(o[d & 3]);
    }
    void next2(float& x, float& y) {
        if (mode == 1) { shared->next2(x, y); return; }
        x = next();
        y = next();
    }
    estream child(uint32_t code) const {
        estream c = *this;
        if (mode == 0) { c.path = eye_child(path, code); c.dim = 0; }
        return c;
    }
};

struct eye_job {
    const scene* S;
    const pmap* maps[2];     // global, caustic (null or empty: skipped)
    float radius[2];
    int max_depth, k;
};

struct eye_query {
    int depth = 1;
    bool evaluated = false, missed = false;
    hit_rec hit;
    bool emission = true;
};

constexpr uint32_t kSmooth = LRT_FLAG_DIFFUSE | LRT_FLAG_GLOSSY;

static uint32_t component_flags(const std::vector<lrt_material>& mats, int m, int c) {
    const lrt_material& M = mats[(size_t)m];
    const uint32_t S = LRT_FLAG_SPECULAR, G = LRT_FLAG_GLOSSY, D = LRT_FLAG_DIFFUSE, R = LRT_FLAG_REFLECTIVE,
                   DI = LRT_FLAG_DIELECTRIC;
    switch (M.type) {
        case LRT_BSDF_DIELECTRIC: return c == 0 ? (S | R | DI) : (S | LRT_FLAG_REFRACTIVE | DI);
        case LRT_BSDF_PLASTIC: return c == 0 ? (S | R | DI) : (D | R);
        case LRT_BSDF_ROUGH_PLASTIC: return c == 0 ? (G | R | DI) : (D | R);
        case LRT_BSDF_COATING:
            return c == bsdf_component_count(mats, M.nested) ? (S | R | DI) : component_flags(mats, M.nested, c);
        case LRT_BSDF_ROUGH_COATING:
            return c == bsdf_component_count(mats, M.nested) ? (G | R | DI) : component_flags(mats, M.nested, c);
        default: return bsdf_flags(mats, m);
    }
}

struct eye_counts {
    int emitter = 1, bsdf = 1;
    float emitter_weight = 1.0f, bsdf_weight = 1.0f;
};
static eye_counts sample_counts(int depth, bool diffuse) {  // photon_integrator.h:277-286
    eye_counts c;
    if (depth == 1) {
        c.emitter = 16;
        c.bsdf = diffuse ? 16 : 32;
        c.emitter_weight = 1.0f / (float)c.emitter;
        c.bsdf_weight = 1.0f / (float)c.bsdf;
    }
    return c;
}

static float mis2(float a_pdf, float b_pdf) {  // photon_integrator.h:306-311
    const float a = a_pdf * a_pdf;
    const float b = b_pdf * b_pdf;
    const float s = a + b;
    return s > 0.0f ? a / s : 0.0f;
}

static v3 emitted(const scene& S, const hit_rec& rec, const v3& w) {  // area_emitter::emitted
    return dot(rec.normal, w) <= 0.0f ? v3(0.0f) : S.lights[(size_t)rec.emitter].radiance;
}

static v3 li(const eye_job& J, const ray& r, eye_query q, estream& smp);

static void prepare(const ray& r, const hit_rec& rec, const onb& basis, bsdf_rec& srec) {
    srec.wi = basis.to_local(-normalize(r.d));
    srec.front_face = rec.front_face;
}

static v3 specular_lobes(const eye_job& J, const ray& r, const hit_rec& rec, estream& smp, int depth) {
    const scene& S = *J.S;
    onb basis;
    basis.build(rec.normal);
    v3 result(0.0f);
    const float* tables = S.tables.empty() ? nullptr : S.tables.data();
    const int n = bsdf_component_count(S.materials, rec.material);
    for (int c = 0; c < n; ++c) {
        if ((component_flags(S.materials, rec.material, c) & LRT_FLAG_SPECULAR) == 0) continue;
        bsdf_ctx ctx;
        ctx.component = c;
        bsdf_rec srec;
        prepare(r, rec, basis, srec);
        const v3 w = bsdf_sample(S.materials, tables, rec.material, ctx, srec, 0.5f, 0.5f);
        if (w.is_zero()) continue;
        eye_query child;
        child.depth = depth + 1;
        estream cs = smp.child((uint32_t)c);
        result = result + w * li(J, ray{rec.p, basis.to_world(srec.wo)}, child, cs);
    }
    return result;
}

static v3 emitter_samples(const eye_job& J, const ray& r, const hit_rec& rec, estream& smp, int depth, bool diffuse,
                          bool specular, bool exhaustive) {
    const scene& S = *J.S;
    const eye_counts cnt = sample_counts(depth, diffuse);
    onb basis;
    basis.build(rec.normal);
    v3 result(0.0f);
    const float* tables = S.tables.empty() ? nullptr : S.tables.data();
    for (int i = 0; i < cnt.emitter; ++i) {
        float sx, sy;
        smp.next2(sx, sy);
        direct_sample ls;
        if (!S.sample_emitter_direction(rec.p, sx, sy, ls)) continue;
        const v3 so = rec.p + rec.normal * 1e-4f;
        v3 sd = ls.direction;
        float tmax = kInf;
        if (!std::isinf(ls.distance)) {
            const v3 to = ls.point - so;
            const float dist = to.length();
            if (dist < 1e-6f) continue;
            sd = to / dist;
            tmax = dist * (1.0f - 1e-4f);
        }
        bsdf_rec srec;
        prepare(r, rec, basis, srec);
        srec.wo = basis.to_local(sd);
        bsdf_ctx ctx;
        if (!specular || exhaustive) ctx.type_mask = kSmooth;
        const v3 value = bsdf_eval(S.materials, tables, rec.material, ctx, srec, LRT_MEASURE_SOLID_ANGLE);
        if (value.is_zero()) continue;
        hit_rec occ;
        if (S.hit(ray{so, sd}, 0.001f, tmax, occ) && occ.emitter != ls.emitter) continue;
        const float lp = S.pdf_emitter_direction(rec.p, sd);
        const float bp = bsdf_pdf(S.materials, tables, rec.material, ctx, srec, LRT_MEASURE_SOLID_ANGLE);
        if (lp <= 0.0f) continue;
        const float w = mis2(lp * (float)cnt.emitter, bp * (float)cnt.bsdf) * cnt.emitter_weight;
        result = result + value * ls.radiance * (w / lp);
    }
    return result;
}

static v3 bsdf_samples(const eye_job& J, const ray& r, const hit_rec& rec, estream& smp, int depth, bool diffuse,
                       bool specular, bool exhaustive, bool sample_direct, bool sample_indirect) {
    const scene& S = *J.S;
    const eye_counts cnt = sample_counts(depth, diffuse);
    onb basis;
    basis.build(rec.normal);
    v3 result(0.0f);
    const float* tables = S.tables.empty() ? nullptr : S.tables.data();
    for (int i = 0; i < cnt.bsdf; ++i) {
        bsdf_ctx ctx;
        ctx.lobe_sample = smp.next();
        if (!specular || exhaustive) ctx.type_mask = kSmooth;
        bsdf_rec srec;
        prepare(r, rec, basis, srec);
        float ux, uy;
        smp.next2(ux, uy);
        const v3 w = bsdf_sample(S.materials, tables, rec.material, ctx, srec, ux, uy);
        if (w.is_zero() || srec.pdf <= 0.0f) continue;
        const ray sc{rec.p, basis.to_world(srec.wo)};
        hit_rec occ;
        const bool hit = S.hit(sc, 0.001f, kInf, occ);
        if (sample_direct) {
            v3 emission(0.0f);
            float lp = 0.0f;
            if (hit && occ.emitter >= 0 && occ.front_face) {
                emission = emitted(S, occ, -normalize(sc.d));
                lp = S.pdf_emitter_direction(rec.p, sc.d);
            } else if (!hit) {
                emission = S.eval_environment(sc.d);
                lp = S.pdf_emitter_direction(rec.p, sc.d);
            }
            if (!emission.is_zero()) {
                const float mis = mis2(srec.pdf * (float)cnt.bsdf, lp * (float)cnt.emitter) * cnt.bsdf_weight;
                result = result + emission * w * mis;
            }
        }
        if (sample_indirect) {
            eye_query child;
            child.depth = depth + 1;
            child.evaluated = true;
            child.missed = !hit;
            child.hit = occ;
            child.emission = false;
            estream cs = smp.child(16u + (uint32_t)i);
            result = result + w * li(J, sc, child, cs) * cnt.bsdf_weight;
        }
    }
    return result;
}

static v3 li(const eye_job& J, const ray& r, eye_query q, estream& smp) {
    const scene& S = *J.S;
    if (!q.evaluated) {
        q.missed = !S.hit(r, 0.001f, kInf, q.hit);
        q.evaluated = true;
    }
    if (q.missed) return q.emission ? S.eval_environment(r.d) : v3(0.0f);
    v3 result(0.0f);
    const hit_rec& rec = q.hit;
    if (q.emission && rec.emitter >= 0 && rec.front_face) result = result + emitted(S, rec, -normalize(r.d));
    if (q.depth >= J.max_depth || rec.material < 0) return result;
    const uint32_t flags = bsdf_flags(S.materials, rec.material);
    const bool diffuse = (flags & (LRT_FLAG_DIFFUSE | LRT_FLAG_GLOSSY | LRT_FLAG_SPECULAR)) == LRT_FLAG_DIFFUSE;
    const bool specular = (flags & LRT_FLAG_SPECULAR) != 0;
    const bool smooth = (flags & kSmooth) != 0;
    const bool exhaustive = specular && q.depth < 4;
    if (diffuse && dot(rec.normal, r.d) < 0.0f) {
        const int photon_depth = J.max_depth - q.depth;
        const v3 refl = col3(S.materials[(size_t)rec.material].diffuse) * (1.0f / kPi);
        for (int m = 0; m < 2; ++m)
            if (J.maps[m] && !J.maps[m]->ph.empty())
                result = result + J.maps[m]->estimate(rec.p, rec.normal, J.radius[m], photon_depth, (size_t)J.k) * refl;
    }
    if (exhaustive) result = result + specular_lobes(J, r, rec, smp, q.depth);
    if (smooth) result = result + emitter_samples(J, r, rec, smp, q.depth, diffuse, specular, exhaustive);
    const bool sample_direct = smooth || (specular && !exhaustive);
    const bool sample_indirect = !diffuse && (smooth || (specular && !exhaustive));
    if (sample_direct || sample_indirect)
        result = result + bsdf_samples(J, r, rec, smp, q.depth, diffuse, specular, exhaustive, sample_direct,
                                       sample_indirect);
    return result;
}

}  // namespace orc

using namespace orc;

extern "C" {

// Eye pass over photon maps built from given photons (12 floats + (depth,
// particle) each; n = 0 leaves a map empty).  mode 0: Philox streams as on
// the device; mode 1: the reference sampler.  per_sample (optional): float
// radiance per (pixel, sample).
// tree: 0 the reference's photon_map tree, 1 the library's device tree
// (pmap::build_median).
int orc_render_photon(void* h, int W, int H, int spp, int max_depth, uint64_t seed, int mode, const float* g_ph,
                      const uint32_t* g_dp, int64_t g_n, float g_scale, float g_radius, const float* c_ph,
                      const uint32_t* c_dp, int64_t c_n, float c_scale, float c_radius, int k, int threads,
                      int row_begin, int row_end, double* sum, double* sumsq, float* per_sample, int tree) {
    const scene& S = *(scene*)h;
    pmap maps[2];
    const float* ph[2] = {g_ph, c_ph};
    const uint32_t* dp[2] = {g_dp, c_dp};
    const int64_t np[2] = {g_n, c_n};
    const float sc[2] = {g_scale, c_scale};
    for (int m = 0; m < 2; ++m) {
        std::vector<photon> v((size_t)std::max<int64_t>(np[m], 0));
        for (int64_t i = 0; i < np[m]; ++i) {
            const float* p = ph[m] + 12 * i;
            v[(size_t)i] = photon{load3(p), load3(p + 3), load3(p + 6), load3(p + 9), (uint16_t)dp[m][2 * i],
                                  dp[m][2 * i + 1]};
        }
        if (tree) maps[m].build_median(std::move(v), sc[m]);
        else maps[m].build(std::move(v), sc[m]);
    }
    eye_job J{&S, {&maps[0], &maps[1]}, {g_radius, c_radius}, max_depth, k};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int rb = row_end > row_begin ? row_begin : 0, re = row_end > row_begin ? row_end : H;
    const int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    std::vector<std::thread> pool;
    auto body = [&](int t) {
        for (int y = rb + t; y < re; y += nt)
            for (int x = 0; x < W; ++x) {
                const uint32_t pix = (uint32_t)(y * W + x);
                double s3[3] = {0, 0, 0}, q3[3] = {0, 0, 0};
                for (int si = 0; si < spp; ++si) {
                    psampler base;  // the reference sampler started at (pixel, sample)
                    base.mode = 1;
                    base.index = (uint32_t)si;
                    base.pixel_seed = mix_bits(pix);
                    estream es;
                    es.mode = mode;
                    es.shared = &base;
                    es.pixel = pix;
                    es.sample = (uint32_t)si;
                    es.key[0] = key[0];
                    es.key[1] = key[1];
                    float jx, jy;
                    es.next2(jx, jy);  // pixel jitter: dimensions 0 and 1 of path 0
                    const float u = ((float)x + jx) / (float)(W - 1);
                    const float v = ((float)y + jy) / (float)(H - 1);
                    const camera& c = S.cam;
                    const ray r{c.origin, normalize(c.llc + u * c.horizontal + v * c.vertical - c.origin)};
                    const v3 L = li(J, r, eye_query{}, es);
                    if (per_sample) {
                        float* o = per_sample + ((size_t)pix * spp + si) * 3;
                        o[0] = L.x; o[1] = L.y; o[2] = L.z;
                    }
                    for (int ch = 0; ch < 3; ++ch) { s3[ch] += L[ch]; q3[ch] += (double)L[ch] * L[ch]; }
                }
                for (int ch = 0; ch < 3; ++ch) {
                    if (sum) sum[(size_t)pix * 3 + ch] += s3[ch];
                    if (sumsq) sumsq[(size_t)pix * 3 + ch] += q3[ch];
                }
            }
    };
    for (int t = 1; t < nt; ++t) pool.emplace_back(body, t);
    body(0);
    for (auto& th : pool) th.join();
    return 0;
}

// quota gather only: photons (12 floats) + (depth, particle), in particle order
int64_t orc_gather_photons(void* h, int kind, uint32_t requested, int max_depth, int rr_depth, uint64_t seed,
                           float* out12, uint32_t* depth_particle, uint64_t* shot_out) {
    const scene& S = *(scene*)h;
    std::vector<photon> ph;
    uint64_t shot = 0;
    gather_quota(S, kind, requested, max_depth, rr_depth, seed, ph, shot);
    for (size_t i = 0; out12 && i < ph.size(); ++i) {
        const photon& p = ph[i];
        float* o = out12 + 12 * i;
        o[0] = p.pos.x; o[1] = p.pos.y; o[2] = p.pos.z;
        o[3] = p.dir.x; o[4] = p.dir.y; o[5] = p.dir.z;
        o[6] = p.normal.x; o[7] = p.normal.y; o[8] = p.normal.z;
        o[9] = p.power.x; o[10] = p.power.y; o[11] = p.power.z;
        depth_particle[2 * i] = p.depth;
        depth_particle[2 * i + 1] = p.particle;
    }
    if (shot_out) *shot_out = shot;
    return (int64_t)ph.size();
}

// Traces particles [first, first+count) one after another; writes photons in
// particle order (13 floats: pos, dir, normal, power, then depth and particle
// as floats-encoded integers in two uint32 arrays).  Returns the photon count;
// with out == null only counts.
int64_t orc_trace_particles(void* h, int kind, uint32_t first, uint32_t count, int max_depth, int rr_depth, uint64_t seed,
                            int mode, float* out12, uint32_t* depth_particle, int64_t capacity) {
    const scene& S = *(scene*)h;
    std::vector<photon> ph;
    for (uint32_t i = 0; i < count; ++i) {
        psampler smp;
        smp.mode = mode;
        smp.index = first + i;
        smp.tag = 0xF0000000u + (uint32_t)kind;
        smp.key[0] = (uint32_t)seed;
        smp.key[1] = (uint32_t)(seed >> 32);
        smp.pixel_seed = mix_bits(0u);
        trace_particle(S, kind, smp, max_depth, rr_depth, ph);
    }
    if (out12) {
        const int64_t n = std::min<int64_t>((int64_t)ph.size(), capacity);
        for (int64_t i = 0; i < n; ++i) {
            const photon& p = ph[(size_t)i];
            float* o = out12 + 12 * i;
            o[0] = p.pos.x; o[1] = p.pos.y; o[2] = p.pos.z;
            o[3] = p.dir.x; o[4] = p.dir.y; o[5] = p.dir.z;
            o[6] = p.normal.x; o[7] = p.normal.y; o[8] = p.normal.z;
            o[9] = p.power.x; o[10] = p.power.y; o[11] = p.power.z;
            depth_particle[2 * i] = p.depth;
            depth_particle[2 * i + 1] = p.particle;
        }
    }
    return (int64_t)ph.size();
}

// photon_map::build + estimate_irradiance for n queries (point, normal)
int orc_photon_estimate(const float* photons12, const uint32_t* depth_particle, int64_t n_photons, float scale,
                        const float* queries6, uint32_t n_queries, float radius, int max_depth, int k, float* out3,
                        int threads, int tree) {
    pmap m;
    std::vector<photon> ph((size_t)n_photons);
    for (int64_t i = 0; i < n_photons; ++i) {
        const float* p = photons12 + 12 * i;
        ph[(size_t)i] = photon{load3(p), load3(p + 3), load3(p + 6), load3(p + 9), (uint16_t)depth_particle[2 * i],
                               depth_particle[2 * i + 1]};
    }
    if (tree) m.build_median(std::move(ph), scale);
    else m.build(std::move(ph), scale);
    const int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    auto body = [&](int t) {
        for (uint32_t q = (uint32_t)t; q < n_queries; q += (uint32_t)nt) {
            const v3 e = m.estimate(load3(queries6 + 6 * (size_t)q), load3(queries6 + 6 * (size_t)q + 3), radius, max_depth, (size_t)k);
            out3[3 * (size_t)q] = e.x; out3[3 * (size_t)q + 1] = e.y; out3[3 * (size_t)q + 2] = e.z;
        }
    };
    for (int t = 1; t < nt; ++t) pool.emplace_back(body, t);
    body(0);
    for (auto& th : pool) th.join();
    return 0;
}

// The library's device photon tree over n positions (3 floats each, input
// order): nodes (4 words each, pre-order: interior {split bits, axis, left,
// right}, leaf {0, 3, first, count}) and order (input index per tree
// position).  Returns the node count; with null buffers only counts.
int64_t orc_photon_tree(const float* xyz, int64_t n, uint32_t* nodes4, uint32_t* order) {
    std::vector<photon> g((size_t)n);
    for (int64_t i = 0; i < n; ++i) g[(size_t)i].pos = load3(xyz + 3 * i);
    pmap m;
    m.build_median(std::move(g), 1.0f);
    if (nodes4) {
        for (size_t i = 0; i < m.nodes.size(); ++i) {
            const pmap::node& d = m.nodes[i];
            uint32_t* o = nodes4 + 4 * i;
            if (d.leaf) {
                o[0] = 0u; o[1] = 3u; o[2] = (uint32_t)d.start; o[3] = (uint32_t)d.count;
            } else {
                std::memcpy(&o[0], &d.split, 4);
                o[1] = (uint32_t)d.axis; o[2] = (uint32_t)d.left; o[3] = (uint32_t)d.right;
            }
        }
    }
    if (order) std::memcpy(order, m.tree_order.data(), m.tree_order.size() * 4);
    return (int64_t)m.nodes.size();
}

// The reference's photon_map tree (pmap::build) over n positions, same
// layout (the split read after both subtrees, photon_map.h:107-110).
int64_t orc_photon_tree_reference(const float* xyz, int64_t n, uint32_t* nodes4, uint32_t* order) {
    std::vector<photon> g((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        g[(size_t)i].pos = load3(xyz + 3 * i);
        g[(size_t)i].particle = (uint32_t)i;
    }
    pmap m;
    m.build(std::move(g), 1.0f);
    if (nodes4) {
        for (size_t i = 0; i < m.nodes.size(); ++i) {
            const pmap::node& d = m.nodes[i];
            uint32_t* o = nodes4 + 4 * i;
            if (d.leaf) {
                o[0] = 0u; o[1] = 3u; o[2] = (uint32_t)d.start; o[3] = (uint32_t)d.count;
            } else {
                std::memcpy(&o[0], &d.split, 4);
                o[1] = (uint32_t)d.axis; o[2] = (uint32_t)d.left; o[3] = (uint32_t)d.right;
            }
        }
    }
    if (order)
        for (int64_t i = 0; i < n; ++i) order[i] = m.ph[(size_t)i].particle;
    return (int64_t)m.nodes.size();
}

}  // extern "C"
