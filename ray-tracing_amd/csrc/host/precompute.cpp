// Constructor-time precompute of the reference materials and camera, on the
// host: the device kernels receive only the resulting scalars and tables.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "host.hpp"

namespace lrt {
namespace {

// fresnel.h:99-118 (Mitsuba fresnelDielectricExt), the integrand of fdr
float fresnel_dielectric_ext(float ci, float eta) {
    if (eta == 1.0f) return 0.0f;
    const float s = ci > 0.0f ? (1.0f / eta) : eta;
    const float ct2 = 1.0f - (1.0f - ci * ci) * (s * s);
    if (ct2 <= 0.0f) return 1.0f;
    const float ai = std::abs(ci);
    const float at = std::sqrt(ct2);
    const float rs = (ai - eta * at) / (ai + eta * at);
    const float rp = (eta * ai - at) / (eta * ai + at);
    return 0.5f * (rs * rs + rp * rp);
}

// Adaptive Gauss-Lobatto quadrature (Gander & Gautschi 2000, as in Mitsuba's
// quad.cpp / QuantLib): 4-point Lobatto vs 7-point Kronrod on each interval,
// initial tolerance from a 13-point estimate.  gauss_lobatto.h:23-121.
struct GaussLobatto {
    size_t max_evals;
    float abs_err, rel_err;
    const float alpha = (float)std::sqrt(2.0 / 3.0);
    const float beta = (float)(1.0 / std::sqrt(5.0));
    const float x1 = (float)0.94288241569547971906;
    const float x2 = (float)0.64185334234578130578;
    const float x3 = (float)0.23638319966214988028;

    template <typename F>
    float tolerance(const F& f, float a, float b, size_t& evals) const {
        const float m = (a + b) * 0.5f, h = (b - a) * 0.5f;
        const float y1 = f(a), y3 = f(m - alpha * h), y5 = f(m - beta * h), y7 = f(m);
        const float y9 = f(m + beta * h), y11 = f(m + alpha * h), y13 = f(b);
        const float acc = h * ((float)0.0158271919734801831 * (y1 + y13) +
                               (float)0.0942738402188500455 * (f(m - x1 * h) + f(m + x1 * h)) +
                               (float)0.1550719873365853963 * (y3 + y11) +
                               (float)0.1888215739601824544 * (f(m - x2 * h) + f(m + x2 * h)) +
                               (float)0.1997734052268585268 * (y5 + y9) +
                               (float)0.2249264653333395270 * (f(m - x3 * h) + f(m + x3 * h)) +
                               (float)0.2426110719014077338 * y7);
        evals += 13;
        float r = 1.0f;
        const float i2 = (h / 6.0f) * (y1 + y13 + 5.0f * (y5 + y9));
        const float i1 = (h / 1470.0f) * (77.0f * (y1 + y13) + 432.0f * (y3 + y11) + 625.0f * (y5 + y9) + 672.0f * y7);
        if (std::abs(i2 - acc) != 0.0f) r = std::abs(i1 - acc) / std::abs(i2 - acc);
        if (r == 0.0f || r > 1.0f) r = 1.0f;
        const float eps = std::numeric_limits<float>::epsilon();
        float res = std::numeric_limits<float>::infinity();
        if (rel_err != 0.0f && acc != 0.0f) res = acc * std::max(rel_err, eps) / (r * eps);
        if (abs_err != 0.0f) res = std::min(res, abs_err / (r * eps));
        return res;
    }
    template <typename F>
    float step(const F& f, float a, float b, float fa, float fb, float acc, size_t& evals) const {
        const float h = (b - a) * 0.5f, m = (a + b) * 0.5f;
        const float mll = m - alpha * h, ml = m - beta * h, mr = m + beta * h, mrr = m + alpha * h;
        const float fmll = f(mll), fml = f(ml), fm = f(m), fmr = f(mr), fmrr = f(mrr);
        const float i2 = (h / 6.0f) * (fa + fb + 5.0f * (fml + fmr));
        const float i1 = (h / 1470.0f) * (77.0f * (fa + fb) + 432.0f * (fmll + fmrr) + 625.0f * (fml + fmr) + 672.0f * fm);
        evals += 5;
        if (evals >= max_evals) return i1;
        const float dist = acc + (i1 - i2);
        if (dist == acc || mll <= a || b <= mrr) return i1;
        return step(f, a, mll, fa, fmll, acc, evals) + step(f, mll, ml, fmll, fml, acc, evals) +
               step(f, ml, m, fml, fm, acc, evals) + step(f, m, mr, fm, fmr, acc, evals) +
               step(f, mr, mrr, fmr, fmrr, acc, evals) + step(f, mrr, b, fmrr, fb, acc, evals);
    }
    template <typename F>
    float integrate(const F& f, float a, float b) const {
        float factor = 1.0f;
        size_t evals = 0;
        if (a == b) return 0.0f;
        if (b < a) { std::swap(a, b); factor = -1.0f; }
        const float tol = tolerance(f, a, b, evals);
        evals += 2;
        return factor * step(f, a, b, f(a), f(b), tol, evals);
    }
};

// catmull_rom.h:10-154: uniform-knot cubic Hermite splines with central
// difference derivatives (one-sided at the ends), in 1, 2 and 3 dimensions.
float cubic1(float x, const float* v, size_t n, float lo, float hi) {
    if (!(x >= lo && x <= hi)) return 0.0f;
    float t = ((x - lo) * (float)(n - 1)) / (hi - lo);
    size_t k = (size_t)t;
    if (k > n - 2) k = n - 2;
    const float f0 = v[k], f1 = v[k + 1];
    const float d0 = k > 0 ? 0.5f * (v[k + 1] - v[k - 1]) : (v[k + 1] - v[k]);
    const float d1 = (k + 2 < n) ? 0.5f * (v[k + 2] - v[k]) : (v[k + 1] - v[k]);
    t -= (float)k;
    const float t2 = t * t, t3 = t2 * t;
    return (2.0f * t3 - 3.0f * t2 + 1.0f) * f0 + (-2.0f * t3 + 3.0f * t2) * f1 + (t3 - 2.0f * t2 + t) * d0 + (t3 - t2) * d1;
}

template <int D>
bool knot_weights(const float* p, const size_t* size, float w[D][4], size_t* knot) {
    for (int d = 0; d < D; ++d) {
        float* ww = w[d];
        if (!(p[d] >= 0.0f && p[d] <= 1.0f)) return false;
        float t = ((p[d] - 0.0f) * (float)(size[d] - 1)) / (1.0f - 0.0f);
        knot[d] = std::min((size_t)t, size[d] - 2);
        t -= (float)knot[d];
        const float t2 = t * t, t3 = t2 * t;
        ww[0] = 0.0f;
        ww[1] = 2.0f * t3 - 3.0f * t2 + 1.0f;
        ww[2] = -2.0f * t3 + 3.0f * t2;
        ww[3] = 0.0f;
        const float d0 = t3 - 2.0f * t2 + t, d1 = t3 - t2;
        if (knot[d] > 0) { ww[2] += 0.5f * d0; ww[0] -= 0.5f * d0; }
        else { ww[2] += d0; ww[1] -= d0; }
        if (knot[d] + 2 < size[d]) { ww[3] += 0.5f * d1; ww[1] -= 0.5f * d1; }
        else { ww[2] += d1; ww[1] -= d1; }
    }
    return true;
}

float cubic2(const float* p, const float* v, const size_t* size) {
    float w[2][4];
    size_t k[2];
    if (!knot_weights<2>(p, size, w, k)) return 0.0f;
    float r = 0.0f;
    for (int y = -1; y <= 2; ++y) {
        const float wy = w[1][y + 1];
        for (int x = -1; x <= 2; ++x) {
            const float wxy = w[0][x + 1] * wy;
            if (wxy == 0.0f) continue;
            r += v[(k[1] + (size_t)y) * size[0] + k[0] + (size_t)x] * wxy;
        }
    }
    return r;
}

float cubic3(const float* p, const float* v, const size_t* size) {
    float w[3][4];
    size_t k[3];
    if (!knot_weights<3>(p, size, w, k)) return 0.0f;
    float r = 0.0f;
    for (int z = -1; z <= 2; ++z) {
        const float wz = w[2][z + 1];
        for (int y = -1; y <= 2; ++y) {
            const float wyz = w[1][y + 1] * wz;
            for (int x = -1; x <= 2; ++x) {
                const float wxyz = w[0][x + 1] * wyz;
                if (wxyz == 0.0f) continue;
                r += v[((k[2] + (size_t)z) * size[1] + (k[1] + (size_t)y)) * size[0] + k[0] + (size_t)x] * wxyz;
            }
        }
    }
    return r;
}

uint64_t rd64(const unsigned char* p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
float rdf(const unsigned char* p) {
    const uint32_t b = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

}  // namespace

float fresnel_diffuse_reflectance(float eta) {
    const GaussLobatto q{1024, 0.0f, 1e-5f};
    return q.integrate([eta](float xi) { return fresnel_dielectric_ext(std::sqrt(xi), eta); }, 0.0f, 1.0f);
}

// ggx.dat layout (rough_transmittance.h:67-112): "MTS_TRANSMITTANCE", then
// u64 eta/alpha/theta counts, f32 eta_min/max, alpha_min/max, then for each of
// 2*eta*alpha (eta >= 1 block, then eta < 1 block): theta floats + 1 diffuse.
// Slicing (:129-185): 3D -> 2D at the warped eta, then 2D -> 1D at the warped
// alpha, with Catmull-Rom splines.
void rough_transmittance_slice(const void* data, size_t bytes, float eta, float alpha, float out101[101]) {
    const auto* p = (const unsigned char*)data;
    if (bytes < 17 + 24 + 16 || std::memcmp(p, "MTS_TRANSMITTANCE", 17) != 0)
        throw std::runtime_error("rough_transmittance: invalid header");
    const size_t ne = (size_t)rd64(p + 17), na = (size_t)rd64(p + 25), nth = (size_t)rd64(p + 33);
    const float eta_min = rdf(p + 41), eta_max = rdf(p + 45), a_min = rdf(p + 49), a_max = rdf(p + 53);
    if (nth != 100) throw std::runtime_error("rough_transmittance: expected 100 theta samples");
    const size_t tsz = 2 * ne * na * nth, dsz = 2 * ne * na;
    if (bytes < 57 + (tsz + dsz) * 4) throw std::runtime_error("rough_transmittance: truncated table");
    std::vector<float> trans(tsz), diff(dsz);
    const unsigned char* q = p + 57;
    size_t it = 0, id = 0;
    for (size_t i = 0; i < 2 * ne; ++i)
        for (size_t j = 0; j < na; ++j) {
            for (size_t k = 0; k < nth; ++k, q += 4) trans[it++] = rdf(q);
            diff[id++] = rdf(q);
            q += 4;
        }
    {  // check_eta / check_alpha (:119-127)
        float e = eta;
        if (e < 1.0f) e = 1.0f / e;
        if (e < eta_min || e > eta_max) throw std::runtime_error("rough_transmittance: eta outside table range");
        if (alpha < a_min || alpha > a_max) throw std::runtime_error("rough_transmittance: alpha outside table range");
    }
    const float* tsrc = trans.data();
    const float* dsrc = diff.data();
    if (eta < 1.0f) {
        tsrc += ne * na * nth;
        dsrc += ne * na;
        eta = 1.0f / eta;
    }
    if (eta < eta_min) eta = eta_min;
    const float we = std::pow((eta - eta_min) / (eta_max - eta_min), 0.25f);
    const float da = 1.0f / (float)(na - 1), dt = 1.0f / (float)(nth - 1);
    const size_t s3[3] = {nth, na, ne}, s2[2] = {na, ne};
    std::vector<float> st(na * nth), sd(na);
    for (size_t i = 0; i < na; ++i) {
        for (size_t j = 0; j < nth; ++j) {
            const float pt[3] = {(float)j * dt, (float)i * da, we};
            st[i * nth + j] = cubic3(pt, tsrc, s3);
        }
        const float pq[2] = {(float)i * da, we};
        sd[i] = cubic2(pq, dsrc, s2);
    }
    const float wa = std::pow((alpha - a_min) / (a_max - a_min), 0.25f);
    const size_t s2b[2] = {nth, na};
    for (size_t i = 0; i < nth; ++i) {
        const float pt[2] = {(float)i * dt, wa};
        out101[i] = cubic2(pt, st.data(), s2b);
    }
    out101[100] = cubic1(wa, sd.data(), na, 0.0f, 1.0f);
}

CameraBasis camera_basis(const lrt_camera_desc& c) {
    auto sub = [](const float* a, const float* b, float* o) { for (int i = 0; i < 3; ++i) o[i] = a[i] - b[i]; };
    auto norm = [](float* v) {
        const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int i = 0; i < 3; ++i) v[i] = v[i] / l;
    };
    auto cross = [](const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    const float pi = 3.14159265358979323846f;
    const float h = std::tan(c.fov_deg * pi / 180.0f / 2.0f);
    const float vh = 2.0f * h, vw = c.aspect * vh;
    float w[3], u[3], v[3];
    sub(c.origin, c.look_at, w);
    norm(w);
    cross(c.up, w, u);
    norm(u);
    cross(w, u, v);
    CameraBasis b;
    const float fh = 1.0f * vw, fv = 1.0f * vh;
    for (int i = 0; i < 3; ++i) {
        b.origin[i] = c.origin[i];
        b.horizontal[i] = u[i] * fh;
        b.vertical[i] = v[i] * fv;
    }
    for (int i = 0; i < 3; ++i)
        b.llc[i] = c.origin[i] - b.horizontal[i] / 2.0f - b.vertical[i] / 2.0f - w[i] * 1.0f;
    return b;
}

void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xD2511F53u * c0, b = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0, n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)b;
        c3 = (uint32_t)a;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Dimension 1 is van der Corput; dimensions 2..16 follow the Joe-Kuo
// recurrence m_i = 2 a_1 m_{i-1} ^ 4 a_2 m_{i-2} ^ ... ^ 2^s m_{i-s} ^ m_{i-s}
// from their primitive polynomials (degree s, coefficients a, initial m),
// the parameters the reference's table is built from (sobol.h:43-106).
void sobol_directions(uint32_t out[16 * 32]) {
    static const uint32_t prim[15][8] = {
        {1, 0, 1}, {2, 1, 1, 3}, {3, 1, 1, 3, 1}, {3, 2, 1, 1, 1}, {4, 1, 1, 1, 3, 3},
        {4, 4, 1, 3, 5, 13}, {5, 2, 1, 1, 5, 5, 17}, {5, 4, 1, 1, 5, 5, 5}, {5, 7, 1, 1, 7, 11, 19},
        {5, 11, 1, 1, 5, 1, 1}, {5, 13, 1, 1, 1, 3, 11}, {5, 14, 1, 3, 5, 5, 13},
        {6, 1, 1, 3, 3, 9, 7, 49}, {6, 13, 1, 1, 1, 15, 21, 21}, {6, 16, 1, 3, 1, 13, 27, 49}};
    for (int b = 0; b < 32; ++b) out[b] = 1u << (31 - b);
    for (int d = 0; d < 15; ++d) {
        const uint32_t deg = prim[d][0], a = prim[d][1];
        uint32_t m[33] = {};
        for (uint32_t i = 1; i <= deg; ++i) m[i] = prim[d][1 + i];
        for (int i = (int)deg + 1; i <= 32; ++i) {
            uint32_t x = (m[i - deg] << deg) ^ m[i - deg];
            for (int k = 1; k < (int)deg; ++k)
                if ((a >> (deg - 1 - k)) & 1u) x ^= m[i - k] << k;
            if (i < 32) x &= (1u << i) - 1u;
            m[i] = x;
        }
        for (int i = 1; i <= 32; ++i) out[(d + 1) * 32 + (i - 1)] = m[i] << (32 - i);
    }
}

}  // namespace lrt
