// Oracle (test infrastructure only): kd-tree build and traversal, shapes,
// scene list, camera and emitters restated from the reference.
#include <cstring>
#include <future>
#include <stdexcept>

#include "orc.h"

namespace orc {

// ---------------------------------------------------------------------------
// SAH kd-tree build: sah_kdtree.h:35-57 (build), :147-166 (split_probability),
// :168-173 (make_leaf), :175-347 (build_node).  The per-node N-byte side vector
// of the reference (:290) is replaced by one array reused across nodes; every
// entry a node reads is reset for that node's primitives first, so the
// partition is identical.
// ---------------------------------------------------------------------------
namespace {
constexpr float kTraversal = 15.0f, kQuery = 20.0f, kEmptyBonus = 0.9f;
constexpr int kStopPrims = 6, kMaxBad = 3;

struct event {
    float pos;
    uint32_t index;
    uint8_t type;  // 0 end, 1 planar, 2 start
    uint8_t axis;
};

inline void split_prob(const box& b, int axis, float lw, float rw, float& pl, float& pr) {
    const v3 e = b.hi - b.lo;
    const float denom = e.x * e.y + e.y * e.z + e.x * e.z;
    const float inv = 1.0f / denom;
    float face, perim;
    if (axis == 0) { face = e.y * e.z; perim = e.y + e.z; }
    else if (axis == 1) { face = e.x * e.z; perim = e.x + e.z; }
    else { face = e.x * e.y; perim = e.x + e.y; }
    pl = (face + perim * lw) * inv;
    pr = (face + perim * rw) * inv;
}
}  // namespace

void kdtree::build(const std::vector<box>& prim_bounds) {
    nodes.clear();
    refs.clear();
    own_bounds_ = prim_bounds;
    bounds_ = &own_bounds_;
    const int n = (int)own_bounds_.size();
    if (n <= 0) return;
    box root;
    for (const box& b : own_bounds_) root.grow(b);
    std::vector<uint32_t> prims((size_t)n);
    for (int i = 0; i < n; ++i) prims[(size_t)i] = (uint32_t)i;
    const int lim = std::min(48, (int)(8.0f + 1.3f * std::log2((float)n)));
    depth_limit = std::max(lim, 1);
    side_.assign((size_t)n, 0);
    build_node(prims, root, 1, 0);
    side_.clear();
    side_.shrink_to_fit();
}

// Appends a subtree built on its own (numbered from 0) after the current
// nodes: the pre-order numbering of the reference is kept because the left
// subtree is spliced first, then the right one, exactly where the sequential
// recursion would have put them.
void kdtree::splice(const kdtree& t) {
    const int no = (int)nodes.size(), ro = (int)refs.size();
    for (kd_node nd : t.nodes) {
        if (nd.leaf) nd.first += ro;
        else { nd.left += no; nd.right += no; }
        nodes.push_back(nd);
    }
    refs.insert(refs.end(), t.refs.begin(), t.refs.end());
}

void kdtree::make_leaf(int index, const std::vector<uint32_t>& prims) {
    kd_node& nd = nodes[(size_t)index];
    nd.leaf = true;
    nd.first = (int)refs.size();
    nd.count = (int)prims.size();
    refs.insert(refs.end(), prims.begin(), prims.end());
}

float kdtree::build_node(const std::vector<uint32_t>& prims, const box& voxel, int depth, int bad) {
    const int index = (int)nodes.size();
    nodes.emplace_back();
    const int n = (int)prims.size();
    const float leaf_cost = (float)n * kQuery;
    if (n <= kStopPrims || depth >= depth_limit || voxel.area() <= 0.0f) {
        make_leaf(index, prims);
        return leaf_cost;
    }

    std::vector<event> ev;
    ev.reserve((size_t)n * 6);
    for (uint32_t p : prims) {
        const box& b = (*bounds_)[p];
        for (uint8_t a = 0; a < 3; ++a) {
            const float lo = b.lo[a], hi = b.hi[a];
            if (lo == hi) {
                ev.push_back(event{lo, p, 1, a});
            } else {
                ev.push_back(event{lo, p, 2, a});
                ev.push_back(event{hi, p, 0, a});
            }
        }
    }
    std::sort(ev.begin(), ev.end(), [](const event& a, const event& b) {
        if (a.axis != b.axis) return a.axis < b.axis;
        if (a.pos != b.pos) return a.pos < b.pos;
        if (a.type != b.type) return a.type < b.type;
        return a.index < b.index;
    });

    int nl[3] = {0, 0, 0}, nr[3] = {n, n, n};
    float best = kInf, best_pos = 0.0f;
    int best_axis = 0;
    bool planar_left = false;
    for (size_t c = 0; c < ev.size();) {
        const int axis = ev[c].axis;
        const float pos = ev[c].pos;
        int n_end = 0, n_planar = 0, n_start = 0;
        while (c < ev.size() && ev[c].axis == axis && ev[c].pos == pos && ev[c].type == 0) { ++n_end; ++c; }
        while (c < ev.size() && ev[c].axis == axis && ev[c].pos == pos && ev[c].type == 1) { ++n_planar; ++c; }
        while (c < ev.size() && ev[c].axis == axis && ev[c].pos == pos && ev[c].type == 2) { ++n_start; ++c; }
        nr[axis] -= n_planar + n_end;
        if (pos > voxel.lo[axis] && pos < voxel.hi[axis]) {
            float pl, pr;
            split_prob(voxel, axis, pos - voxel.lo[axis], voxel.hi[axis] - pos, pl, pr);
            const int a = nl[axis], b = nr[axis];
            if (n_planar == 0) {
                float cost = kTraversal + kQuery * (pl * (float)a + pr * (float)b);
                if (a == 0 || b == 0) cost *= kEmptyBonus;
                if (cost < best) { best = cost; best_axis = axis; best_pos = pos; }
            } else {
                float cl = kTraversal + kQuery * (pl * (float)(a + n_planar) + pr * (float)b);
                float cr = kTraversal + kQuery * (pl * (float)a + pr * (float)(b + n_planar));
                if (a + n_planar == 0 || b == 0) cl *= kEmptyBonus;
                if (a == 0 || b + n_planar == 0) cr *= kEmptyBonus;
                if (cl < best || cr < best) {
                    planar_left = cl < cr;
                    best = planar_left ? cl : cr;
                    best_axis = axis;
                    best_pos = pos;
                }
            }
        }
        nl[axis] += n_start + n_planar;
    }

    if (best >= leaf_cost) {
        if ((best > 4.0f * leaf_cost && n < 16) || bad >= kMaxBad || best == kInf) {
            make_leaf(index, prims);
            return leaf_cost;
        }
        ++bad;
    }

    for (uint32_t p : prims) side_[p] = 0;
    for (const event& e : ev) {
        if (e.axis != best_axis) continue;
        if (e.type == 0 && e.pos <= best_pos) side_[e.index] = 1;
        else if (e.type == 2 && e.pos >= best_pos) side_[e.index] = 2;
        else if (e.type == 1) side_[e.index] = (e.pos < best_pos || (e.pos == best_pos && planar_left)) ? 1 : 2;
    }
    std::vector<uint32_t> lp, rp;
    for (uint32_t p : prims) {
        if (side_[p] == 1) lp.push_back(p);
        else if (side_[p] == 2) rp.push_back(p);
        else { lp.push_back(p); rp.push_back(p); }
    }
    ev.clear();
    ev.shrink_to_fit();

    box lv = voxel, rv = voxel;
    lv.hi[best_axis] = best_pos;
    rv.lo[best_axis] = best_pos;
    int li, ri;
    float lc, rc;
    if (n >= 65536 && depth <= 4) {
        // large subtrees on two threads, spliced back in pre-order
        kdtree L, R;
        for (kdtree* t : {&L, &R}) {
            t->bounds_ = bounds_;
            t->depth_limit = depth_limit;
            t->side_.assign(bounds_->size(), 0);
        }
        auto fut = std::async(std::launch::async, [&]() { return L.build_node(lp, lv, depth + 1, bad); });
        rc = R.build_node(rp, rv, depth + 1, bad);
        lc = fut.get();
        li = (int)nodes.size();
        splice(L);
        ri = (int)nodes.size();
        splice(R);
    } else {
        li = (int)nodes.size();
        lc = build_node(lp, lv, depth + 1, bad);
        ri = (int)nodes.size();
        rc = build_node(rp, rv, depth + 1, bad);
    }
    float pl, pr;
    split_prob(voxel, best_axis, best_pos - voxel.lo[best_axis], voxel.hi[best_axis] - best_pos, pl, pr);
    const float final_cost = kTraversal + pl * lc + pr * rc;
    if (final_cost < leaf_cost) {
        kd_node& nd = nodes[(size_t)index];
        nd.leaf = false;
        nd.axis = best_axis;
        nd.split = best_pos;
        nd.left = li;
        nd.right = ri;
        return final_cost;
    }
    make_leaf(index, prims);
    return leaf_cost;
}

// ---------------------------------------------------------------------------
// Shapes
// ---------------------------------------------------------------------------

// triangle_mesh.h:43-99 (Moller-Trumbore, voxel window [t0, t1], shading
// normal = normalised barycentric blend flipped toward the ray).
bool object::hit_mesh_tri(uint32_t tri, const ray& r, float t0, float t1, float& closest, hit_rec& rec) const {
    const size_t b = (size_t)tri * 3;
    const v3& p0 = pos[idx[b]];
    const v3& p1 = pos[idx[b + 1]];
    const v3& p2 = pos[idx[b + 2]];
    const v3 e1 = p1 - p0;
    const v3 e2 = p2 - p0;
    const v3 g = cross(e1, e2);
    if (g.length_squared() <= 0.0f) return false;
    const v3 pv = cross(r.d, e2);
    const float det = dot(e1, pv);
    if (std::abs(det) < 1e-8f) return false;
    const float inv = 1.0f / det;
    const v3 tv = r.o - p0;
    const float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const v3 qv = cross(tv, e1);
    const float v = dot(r.d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = dot(e2, qv) * inv;
    if (t < t0 || t > t1) return false;

    closest = t;
    rec.t = t;
    rec.p = r.at(t);
    rec.material = tri_mat.empty() ? material : (int)tri_mat[tri];
    rec.emitter = emitter;
    const v3 outward = g / g.length();
    rec.front_face = dot(r.d, outward) < 0;
    rec.normal = rec.front_face ? outward : -outward;
    const float b0 = 1.0f - u - v;
    v3 s = b0 * cn[b] + u * cn[b + 1] + v * cn[b + 2];
    if (s.length_squared() > 0.0f) {
        s = s / s.length();
        if (dot(r.d, s) > 0.0f) s = -s;
        rec.normal = s;
    }
    rec.prim = (int)tri;
    rec.u = u;
    rec.v = v;
    return true;
}

bool object::intersect(const ray& r, float t_min, float t_max, hit_rec& rec, kd_counts* cnt) const {
    if (shape == LRT_SHAPE_MESH) {  // triangle_mesh.h:30-41
        bool any = false;
        float closest = t_max;
        kd.intersect(r, t_min, closest, [&](uint32_t tri, float mint, float maxt) {
            if (hit_mesh_tri(tri, r, mint, maxt, closest, rec)) any = true;
        }, cnt);
        return any;
    }
    if (shape == LRT_SHAPE_RECTANGLE) {  // rectangle.h:18-55
        const float denom = dot(rn, r.d);
        if (std::abs(denom) < 1e-8f) return false;
        const float t = dot(corner - r.o, rn) / denom;
        if (t < t_min || t > t_max) return false;
        const v3 p = r.at(t);
        const v3 d = p - corner;
        const float uu = dot(eu, eu), vv = dot(ev, ev), uv = dot(eu, ev);
        const float du = dot(d, eu), dv = dot(d, ev);
        const float det = uu * vv - uv * uv;
        if (std::abs(det) < 1e-12f) return false;
        const float a = (vv * du - uv * dv) / det;
        const float c = (uu * dv - uv * du) / det;
        if (a < 0.0f || a > 1.0f || c < 0.0f || c > 1.0f) return false;
        rec.t = t;
        rec.p = p;
        rec.material = -1;  // rectangle.h:50: never a BSDF
        rec.emitter = emitter;
        rec.front_face = dot(r.d, rn) < 0;
        rec.normal = rec.front_face ? rn : -rn;
        rec.prim = 0;
        rec.u = a;
        rec.v = c;
        return true;
    }
    // sphere.h:21-49
    const v3 oc = r.o - center;
    const float a = r.d.length_squared();
    const float half_b = dot(oc, r.d);
    const float c = oc.length_squared() - radius * radius;
    const float disc = half_b * half_b - a * c;
    if (disc < 0) return false;
    const float sq = std::sqrt(disc);
    float root = (-half_b - sq) / a;
    if (root < t_min || root > t_max) {
        root = (-half_b + sq) / a;
        if (root < t_min || root > t_max) return false;
    }
    rec.t = root;
    rec.p = r.at(rec.t);
    rec.material = material;
    rec.emitter = emitter;
    const v3 outward = (rec.p - center) / radius;
    rec.front_face = dot(r.d, outward) < 0;
    rec.normal = rec.front_face ? outward : -outward;
    rec.prim = 0;
    rec.u = rec.v = 0.0f;
    return true;
}

// triangle_mesh.h:109-138, rectangle.h:67-71, sphere.h:59-67
void object::sample_position(float sx, float sy, v3& p, v3& n, float& pdf) const {
    if (shape == LRT_SHAPE_MESH) {
        pdf = 0.0f;
        if (area <= 0.0f || area_cdf.empty()) return;
        const float target = std::min(sx, 0.999999f) * area;
        size_t tri = (size_t)(std::upper_bound(area_cdf.begin(), area_cdf.end(), target) - area_cdf.begin());
        if (tri >= area_cdf.size()) tri = area_cdf.size() - 1;
        const float prev = tri == 0 ? 0.0f : area_cdf[tri - 1];
        const float span = area_cdf[tri] - prev;
        const float u0 = span > 0.0f ? (target - prev) / span : 0.0f;
        const float su = std::sqrt(u0);
        const float b0 = 1.0f - su;
        const float b1 = sy * su;
        const float b2 = 1.0f - b0 - b1;
        const v3& a0 = pos[idx[3 * tri]];
        const v3& a1 = pos[idx[3 * tri + 1]];
        const v3& a2 = pos[idx[3 * tri + 2]];
        p = b0 * a0 + b1 * a1 + b2 * a2;
        const v3 g = cross(a1 - a0, a2 - a0);
        n = g / g.length();
        pdf = 1.0f / area;
    } else if (shape == LRT_SHAPE_RECTANGLE) {
        p = corner + sx * eu + sy * ev;
        n = rn;
        pdf = 1.0f / area;
    } else {
        const float z = 1.0f - 2.0f * sy;
        const float ring = std::sqrt(std::max(0.0f, 1.0f - z * z));
        const float phi = 2.0f * kPi * sx;
        n = v3(ring * std::cos(phi), ring * std::sin(phi), z);
        p = center + radius * n;
        pdf = 1.0f / area;
    }
}

// hittable_list.h:19-34: closest over the object list; a later object wins
// an equal t because the object's t_max is inclusive.
bool scene::hit(const ray& r, float t_min, float t_max, hit_rec& rec, kd_counts* cnt) const {
    hit_rec tmp;
    bool any = false;
    float closest = t_max;
    for (size_t i = 0; i < objects.size(); ++i) {
        if (objects[i].intersect(r, t_min, closest, tmp, cnt)) {
            any = true;
            closest = tmp.t;
            tmp.obj = (int)i;
            rec = tmp;
        }
    }
    return any;
}

// ---------------------------------------------------------------------------
// Emitters: area_emitter.h:17-72, env_emitter.h:13-88, scene.h:62-108
// ---------------------------------------------------------------------------
static bool area_sample(const scene& s, const light& L, const v3& p, float sx, float sy, direct_sample& out) {
    const object& shp = s.objects[(size_t)L.object];
    if (shp.area <= 0.0f) return false;
    v3 q, n;
    float pdf_area = 0.0f;
    shp.sample_position(sx, sy, q, n, pdf_area);
    const v3 to = q - p;
    const float d2 = to.length_squared();
    if (d2 < 1e-8f) return false;
    const float d = std::sqrt(d2);
    const v3 dir = to / d;
    const float cos_l = dot(n, -dir);
    if (cos_l <= 0.0f) return false;
    out.point = q;
    out.normal = n;
    out.direction = dir;
    out.distance = d;
    out.pdf = pdf_area * d2 / cos_l;
    out.radiance = L.radiance;
    return true;
}

static float area_pdf(const scene& s, const light& L, const v3& p, const v3& direction) {
    const object& shp = s.objects[(size_t)L.object];
    if (shp.area <= 0.0f) return 0.0f;
    const v3 dir = normalize(direction);
    hit_rec rec;
    if (!shp.intersect(ray{p, dir}, 0.001f, kInf, rec) || !rec.front_face) return 0.0f;
    const float cos_l = dot(rec.normal, -dir);
    if (cos_l <= 0.0f) return 0.0f;
    return rec.t * rec.t / (shp.area * cos_l);
}

// ArHosekSkyModel.cc:291-304 (9-parameter radiance) and :637-651 (scaled by
// the cooked per-channel radiance); published Hosek-Wilkie 2012 formula.
double hosek_channel(const double c[9], double rad, double theta, double gamma) {
    const double expM = std::exp(c[4] * gamma);
    const double rayM = std::cos(gamma) * std::cos(gamma);
    const double mieM = (1.0 + std::cos(gamma) * std::cos(gamma)) /
                        std::pow((1.0 + c[8] * c[8] - 2.0 * c[8] * std::cos(gamma)), 1.5);
    const double zenith = std::sqrt(std::cos(theta));
    return (1.0 + c[0] * std::exp(c[1] / (std::cos(theta) + 0.01))) *
           (c[2] + c[3] * expM + c[5] * rayM + c[6] * mieM + c[7] * zenith) * rad;
}

v3 sky_eval(const light& L, const v3& direction) {
    const v3 dir = normalize(direction);
    if (dir.y <= 0.0f) return v3(1.0f);  // env_emitter.h:75-77, not scaled
    const float theta = std::acos(std::clamp(dir.y, 0.0f, 1.0f));
    const float gamma = std::acos(std::clamp(dot(dir, L.sun), -1.0f, 1.0f));
    return v3((float)hosek_channel(L.cfg[0], L.rad[0], theta, gamma),
              (float)hosek_channel(L.cfg[1], L.rad[1], theta, gamma),
              (float)hosek_channel(L.cfg[2], L.rad[2], theta, gamma)) * L.scale;
}

static bool sky_sample(const light& L, float sx, float sy, direct_sample& out) {
    if (!L.sun_active) return false;
    const float ct = (1.0f - sx) + sx * L.cos_cutoff;
    const float st = std::sqrt(std::max(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * kPi * sy;
    const v3 local(std::cos(phi) * st, std::sin(phi) * st, ct);
    out.direction = L.frame.to_world(local);
    out.pdf = L.cone_pdf;
    out.radiance = sky_eval(L, out.direction);
    return true;
}

static float sky_pdf(const light& L, const v3& direction) {
    if (!L.sun_active) return 0.0f;
    const v3 dir = normalize(direction);
    if (dot(dir, L.sun) + 1e-5f < L.cos_cutoff) return 0.0f;
    return L.cone_pdf;
}

v3 scene::eval_environment(const v3& d) const {
    v3 e(0.0f);
    for (const light& L : lights)
        if (L.kind == LRT_LIGHT_SKY) e = e + sky_eval(L, d);
    return e;
}

bool scene::sample_emitter_direction(const v3& p, float sx, float sy, direct_sample& out) const {
    const int count = (int)lights.size();
    if (count <= 0) return false;
    const float scaled = sx * (float)count;
    int index = (int)scaled;
    if (index >= count) index = count - 1;
    const float rx = scaled - (float)index;
    const light& L = lights[(size_t)index];
    const bool ok = L.kind == LRT_LIGHT_SKY ? sky_sample(L, rx, sy, out) : area_sample(*this, L, p, rx, sy, out);
    if (!ok) return false;
    out.emitter = index;
    out.pdf *= 1.0f / (float)count;
    return out.pdf > 0.0f;
}

float scene::pdf_emitter_direction(const v3& p, const v3& d) const {
    const int count = (int)lights.size();
    if (count <= 0) return 0.0f;
    const float sel = 1.0f / (float)count;
    float sum = 0.0f;
    for (const light& L : lights)
        sum += sel * (L.kind == LRT_LIGHT_SKY ? sky_pdf(L, d) : area_pdf(*this, L, p, d));
    return sum;
}

// camera.h:20-44 with aperture 0 and focus distance 1
void camera_setup(camera& c, const lrt_camera_desc& d) {
    const float focus = 1.0f;
    const float h = std::tan(d.fov_deg * kPi / 180.0f / 2.0f);
    const float vh = 2.0f * h;
    const float vw = d.aspect * vh;
    c.origin = load3(d.origin);
    c.w = normalize(c.origin - load3(d.look_at));
    c.u = normalize(cross(load3(d.up), c.w));
    c.v = cross(c.w, c.u);
    c.horizontal = focus * vw * c.u;
    c.vertical = focus * vh * c.v;
    c.llc = c.origin - c.horizontal / 2.0f - c.vertical / 2.0f - focus * c.w;
}

std::unique_ptr<scene> make_scene(const lrt_scene_desc* desc) {
    auto s = std::make_unique<scene>();
    s->materials.assign(desc->materials, desc->materials + desc->n_materials);
    s->tables.assign(desc->tables, desc->tables + (size_t)desc->n_tables * 101);
    s->objects.resize(desc->n_objects);
    for (uint32_t i = 0; i < desc->n_objects; ++i) {
        const lrt_object_desc& od = desc->objects[i];
        object& o = s->objects[i];
        o.shape = od.shape;
        o.material = od.shape == LRT_SHAPE_RECTANGLE ? -1 : od.material;
        o.emitter = od.emitter;
        if (od.shape == LRT_SHAPE_MESH) {
            o.pos.resize(od.n_vertices);
            for (uint32_t v = 0; v < od.n_vertices; ++v) o.pos[v] = load3(od.positions + 3 * (size_t)v);
            o.idx.assign(od.indices, od.indices + 3 * (size_t)od.n_triangles);
            o.cn.resize(3 * (size_t)od.n_triangles);
            for (size_t k = 0; k < o.cn.size(); ++k)
                o.cn[k] = od.corner_normals ? load3(od.corner_normals + 3 * k) : v3(0.0f);
            if (od.triangle_material) o.tri_mat.assign(od.triangle_material, od.triangle_material + od.n_triangles);
            // triangle_mesh.h:16-27: running area sum and per-triangle boxes
            o.area = 0.0f;
            std::vector<box> boxes(od.n_triangles);
            for (uint32_t t = 0; t < od.n_triangles; ++t) {
                const v3& a = o.pos[o.idx[3 * t]];
                const v3& b = o.pos[o.idx[3 * t + 1]];
                const v3& c = o.pos[o.idx[3 * t + 2]];
                o.area += 0.5f * cross(b - a, c - a).length();
                o.area_cdf.push_back(o.area);
                boxes[t].grow(a); boxes[t].grow(b); boxes[t].grow(c);
            }
            o.kd.build(boxes);
        } else if (od.shape == LRT_SHAPE_RECTANGLE) {
            o.corner = load3(od.corner); o.eu = load3(od.edge_u); o.ev = load3(od.edge_v);
            const v3 n = cross(o.eu, o.ev);
            o.area = n.length();
            if (o.area > 0.0f) o.rn = n / o.area;
        } else {
            o.center = load3(od.center);
            o.radius = od.radius;
            o.area = 4.0f * kPi * od.radius * od.radius;
        }
    }
    for (uint32_t i = 0; i < desc->n_lights; ++i) {
        const lrt_light_desc& ld = desc->lights[i];
        light L;
        L.kind = ld.kind;
        L.object = ld.object;
        L.radiance = load3(ld.radiance);
        if (ld.kind == LRT_LIGHT_SKY) {
            L.sun = normalize(load3(ld.sun_dir));
            L.frame.build(L.sun);
            L.sun_active = L.sun.y > 0.0f;
            const double radius = 0.255 * 3.14159265358979323846 / 180.0;
            const double cosine = std::cos(radius);
            L.cos_cutoff = (float)cosine;
            L.cone_pdf = (float)(1.0 / (2.0 * 3.14159265358979323846 * (1.0 - cosine)));
            L.scale = ld.scale;
            std::memcpy(L.cfg, ld.config, sizeof(L.cfg));
            std::memcpy(L.rad, ld.channel_radiance, sizeof(L.rad));
        }
        s->lights.push_back(L);
    }
    camera_setup(s->cam, desc->camera);
    return s;
}

}  // namespace orc
