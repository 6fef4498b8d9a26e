// Oracle (test infrastructure only): samplers and the path-tracing integrator
// restated from the reference, plus a tiled multi-threaded frame loop like
// src/main.cpp:103-147.
#include <atomic>
#include <thread>

#include "orc.h"

namespace orc {

// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11; Random123): 10 rounds of
// two 32x32->64 multiplies with Weyl key bumps between rounds.
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ---- reference sampler: sampler.h:12-32 + sobol.h:12-106 (Gray-code Sobol,
// Joe-Kuo primitive polynomials for 16 dimensions, per-pixel XOR scramble) ----
namespace {
uint32_t mix_bits(uint32_t x) {  // sobol.h:34-41
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
struct sobol_dirs {
    uint32_t v[16][32];
    sobol_dirs() {
        // (degree, a, m_1..m_degree) for dimensions 2..16 (Joe & Kuo)
        static const uint32_t prim[15][8] = {
            {1, 0, 1}, {2, 1, 1, 3}, {3, 1, 1, 3, 1}, {3, 2, 1, 1, 1}, {4, 1, 1, 1, 3, 3},
            {4, 4, 1, 3, 5, 13}, {5, 2, 1, 1, 5, 5, 17}, {5, 4, 1, 1, 5, 5, 5}, {5, 7, 1, 1, 7, 11, 19},
            {5, 11, 1, 1, 5, 1, 1}, {5, 13, 1, 1, 1, 3, 11}, {5, 14, 1, 3, 5, 5, 13},
            {6, 1, 1, 3, 3, 9, 7, 49}, {6, 13, 1, 1, 1, 15, 21, 21}, {6, 16, 1, 3, 1, 13, 27, 49}};
        for (int b = 0; b < 32; ++b) v[0][b] = 1u << (31 - b);
        for (int d = 0; d < 15; ++d) {
            const uint32_t s = prim[d][0], a = prim[d][1];
            uint32_t m[33] = {};
            for (uint32_t i = 1; i <= s; ++i) m[i] = prim[d][1 + i];
            for (int i = (int)s + 1; i <= 32; ++i) {
                uint32_t x = (m[i - s] << s) ^ m[i - s];
                for (int k = 1; k < (int)s; ++k)
                    if ((a >> (s - 1 - k)) & 1u) x ^= m[i - k] << k;
                if (i < 32) x &= (1u << i) - 1u;
                m[i] = x;
            }
            for (int i = 1; i <= 32; ++i) v[d + 1][i - 1] = m[i] << (32 - i);
        }
    }
};
const sobol_dirs& dirs() { static const sobol_dirs d; return d; }
uint32_t sobol_bits(uint32_t index, uint32_t dim) {
    if (dim >= 16) return mix_bits(index + dim * 0x9e3779b9u);
    const uint32_t g = index ^ (index >> 1);
    uint32_t x = 0;
    for (int b = 0; b < 32; ++b)
        if ((g >> b) & 1u) x ^= dirs().v[dim][b];
    return x;
}
}  // namespace

uint32_t orc_sobol_bits(uint32_t index, uint32_t dim) { return sobol_bits(index, dim); }

struct sampler {
    int mode = 0;  // 0 Philox, 1 reference Sobol
    uint32_t pixel = 0, sample = 0, dim = 0, pixel_seed = 0;
    uint32_t key[2] = {0, 0};
    void start(uint32_t p, uint32_t s) {
        pixel = p; sample = s; dim = 0;
        pixel_seed = mix_bits(p);
    }
    float next() {
        const uint32_t d = dim++;
        if (mode == 1) {
            const uint32_t bits = sobol_bits(sample, d) ^ mix_bits(pixel_seed + d * 0x9e3779b9u);
            return (float)bits * (1.0f / 4294967296.0f);
        }
        const uint32_t ctr[4] = {pixel, sample, d >> 2, 0u};
        uint32_t o[4];
        philox4x32_10(ctr, key, o);
        return u24(o[d & 3]);
    }
    // sampler::next_2d (sampler.h:27-29) is vec2(next_1d(), next_1d()); g++
    // evaluates those arguments right to left, so the compiled reference
    // returns (dimension d+1, dimension d).  Philox mode keeps (d, d+1).
    void next2(float& x, float& y) {
        if (mode == 1) { y = next(); x = next(); }
        else { x = next(); y = next(); }
    }
};

static float power_heuristic(float a_pdf, float b_pdf) {  // path_integrator.h:177-182
    const float a = a_pdf * a_pdf;
    const float b = b_pdf * b_pdf;
    const float s = a + b;
    return s > 0.0f ? a / s : 0.0f;
}

// path_integrator.h:18-170 with max_depth / rr_depth as parameters
struct path_stats { uint64_t ext = 0, shadow = 0; };

v3 radiance(const scene& S, ray r, sampler& smp, int max_depth, int rr_depth, path_stats& st) {
    float eta = 1.0f;
    v3 T(1.0f), L(0.0f);
    bool have_pdf = false;
    float sa_pdf = 0.0f;
    const float* tables = S.tables.empty() ? nullptr : S.tables.data();
    for (int depth = 0; depth < max_depth; ++depth) {
        hit_rec rec;
        st.ext++;
        if (!S.hit(r, 0.001f, kInf, rec)) {
            const v3 Le = S.eval_environment(r.d);
            if (!have_pdf) {
                L = L + T * Le;
            } else {
                const float lp = S.pdf_emitter_direction(rec.p, r.d);
                const float w = power_heuristic(sa_pdf, lp);
                L = L + T * w * Le;
            }
            break;
        }
        if (rec.emitter >= 0 && rec.front_face) {
            const v3 out = -normalize(r.d);
            const light& E = S.lights[(size_t)rec.emitter];
            const v3 Le = dot(rec.normal, out) <= 0.0f ? v3(0.0f) : E.radiance;  // area_emitter.h:80-86
            if (!Le.is_zero()) {
                if (!have_pdf) {
                    L = L + T * Le;
                } else {
                    const float lp = S.pdf_emitter_direction(r.o, r.d);
                    const float w = power_heuristic(sa_pdf, lp);
                    L = L + T * w * Le;
                }
            }
        }
        if (rec.material < 0) break;
        onb basis;
        basis.build(rec.normal);
        bsdf_rec srec;
        srec.wi = basis.to_local(-normalize(r.d));
        srec.front_face = rec.front_face;
        bsdf_ctx ctx;
        ctx.mode = LRT_MODE_RADIANCE;
        float lu, lv, bu, bv;
        smp.next2(lu, lv);
        ctx.lobe_sample = smp.next();
        smp.next2(bu, bv);
        const float rr = smp.next();

        if ((bsdf_flags(S.materials, rec.material) & (LRT_FLAG_DIFFUSE | LRT_FLAG_GLOSSY)) != 0) {
            direct_sample ds;
            if (S.sample_emitter_direction(rec.p, lu, lv, ds)) {
                const v3 so = rec.p + rec.normal * 1e-4f;
                v3 sd = ds.direction;
                float tmax = kInf;
                bool target = true;
                if (!std::isinf(ds.distance)) {
                    const v3 to = ds.point - so;
                    const float dist = to.length();
                    if (dist < 1e-6f) target = false;
                    else { sd = to / dist; tmax = dist * (1.0f - 1e-4f); }
                }
                if (target) {
                    srec.wo = basis.to_local(sd);
                    const v3 fc = bsdf_eval(S.materials, tables, rec.material, ctx, srec, LRT_MEASURE_SOLID_ANGLE);
                    if (!fc.is_zero()) {
                        hit_rec occ;
                        st.shadow++;
                        const bool any = S.hit(ray{so, sd}, 0.001f, tmax, occ);
                        const bool blocked = any && occ.emitter != ds.emitter;
                        if (!blocked) {
                            const float bp = bsdf_pdf(S.materials, tables, rec.material, ctx, srec, LRT_MEASURE_SOLID_ANGLE);
                            const float lp = S.pdf_emitter_direction(rec.p, sd);
                            if (lp > 0.0f) {
                                const float w = power_heuristic(lp, bp);
                                L = L + T * fc * ds.radiance * (w / lp);
                            }
                        }
                    }
                }
            }
        }
        if (depth == max_depth - 1) break;
        const v3 bw = bsdf_sample(S.materials, tables, rec.material, ctx, srec, bu, bv);
        if (bw.is_zero()) break;
        r.o = rec.p;
        r.d = basis.to_world(srec.wo);
        T = T * bw;
        eta *= srec.eta;
        have_pdf = ((srec.sampled_type & LRT_FLAG_SPECULAR) == 0) && srec.pdf > 0.0f;
        sa_pdf = srec.pdf;
        if (depth >= rr_depth) {
            const float q = std::min(std::max({T.x, T.y, T.z}) * eta * eta, 0.95f);
            if (rr >= q) break;
            T = v3(T.x / q, T.y / q, T.z / q);
        }
    }
    return L;
}

// Frame loop (src/main.cpp:110-146): 32x32 tiles over a thread pool; pixel
// index y*W+x with row 0 at the bottom; u = (x+jx)/(W-1), v = (y+jy)/(H-1).
// Accumulates the per-pixel sum (FP64) and sum of squares, or writes every
// sample when per_sample is given (3 floats per (pixel, sample)).


void render(render_job& J) {
    const int tile = 32;
    std::vector<std::pair<int, int>> tiles;
    for (int y0 = J.row_begin; y0 < J.row_end; y0 += tile)
        for (int x0 = 0; x0 < J.W; x0 += tile) tiles.emplace_back(x0, y0);
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> ext{0}, shadow{0};
    auto worker = [&]() {
        path_stats st;
        for (;;) {
            const size_t t = next.fetch_add(1);
            if (t >= tiles.size()) break;
            const int x0 = tiles[t].first, y0 = tiles[t].second;
            const int x1 = std::min(x0 + tile, J.W), y1 = std::min(y0 + tile, J.row_end);
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    const uint32_t pix = (uint32_t)(y * J.W + x);
                    v3 facc(0.0f);
                    double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
                    for (int k = 0; k < J.spp; ++k) {
                        sampler smp;
                        smp.mode = J.mode;
                        smp.key[0] = J.key[0]; smp.key[1] = J.key[1];
                        smp.start(pix, (uint32_t)(J.spp_offset + k));
                        float jx, jy;
                        smp.next2(jx, jy);
                        const float u = ((float)x + jx) / (float)(J.W - 1);
                        const float v = ((float)y + jy) / (float)(J.H - 1);
                        const camera& c = J.S->cam;
                        ray r{c.origin, normalize(c.llc + u * c.horizontal + v * c.vertical - c.origin)};
                        const v3 L = radiance(*J.S, r, smp, J.max_depth, J.rr_depth, st);
                        facc = facc + L;
                        for (int ch = 0; ch < 3; ++ch) { s[ch] += L[ch]; q[ch] += (double)L[ch] * L[ch]; }
                        if (J.per_sample) {
                            float* o = J.per_sample + ((size_t)pix * J.spp + k) * 3;
                            o[0] = L.x; o[1] = L.y; o[2] = L.z;
                        }
                    }
                    for (int ch = 0; ch < 3; ++ch) {
                        if (J.sum) J.sum[(size_t)pix * 3 + ch] += s[ch];
                        if (J.sumsq) J.sumsq[(size_t)pix * 3 + ch] += q[ch];
                        if (J.fsum) J.fsum[(size_t)pix * 3 + ch] = facc[ch];
                    }
                }
        }
        ext += st.ext;
        shadow += st.shadow;
    };
    const int nt = std::max(1, J.threads);
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    J.ext = ext.load();
    J.shadow = shadow.load();
}

}  // namespace orc
