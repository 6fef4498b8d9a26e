// Device math for the render kernels (gfx950).  The whole library is
// compiled with -ffp-contract=off and IEEE-correct division and square root,
// and every expression keeps the reference's operation order, so products
// like the Moller-Trumbore t round exactly as on the reference CPU build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define LRT_D __device__ __forceinline__
#define LRT_HD __host__ __device__ __forceinline__

namespace lrt {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInfF = __builtin_huge_valf();

struct f3 {
    float x, y, z;
};
LRT_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
LRT_HD f3 splat(float s) { return f3{s, s, s}; }
LRT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
LRT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
LRT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
LRT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
LRT_HD f3 operator*(float s, f3 a) { return f3{a.x * s, a.y * s, a.z * s}; }
LRT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
LRT_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
LRT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LRT_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LRT_HD float len2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
LRT_D float len(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
LRT_D f3 normalize(f3 a) { return a / len(a); }
LRT_HD bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
LRT_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
LRT_HD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

// sin, cos, x^(1/4) and exp of a float as the reference's std::sin(float) etc.
// compute them on its CPU: in double precision, rounded once to float.  That
// is the correctly rounded float except where the double result lies within a
// double ulp of a float rounding boundary.  The host C library's float
// versions are within 0.56 ulp (glibc: they equal the correctly rounded value
// on 98.7% of sin/cos and 99.9% of pow/exp arguments); the device library's
// float versions are not correctly rounded and differ from the host more
// often, which forks sampled paths.
//
// sincos_rn serves the samplers, whose arguments are 2 pi u with u in [0, 1):
// Cody-Waite reduction by pi/2 (pi/2 split in three parts, the first two with
// 30 significant bits, so k * part is exact for |k| < 2^23) and the Taylor
// series of sin and cos on [-pi/4, pi/4] through the r^21 and r^22 terms
// (truncation below 2e-25), evaluated with fused multiply-adds: within 2
// double ulps of the exact values for |x| < 2^22, and a fraction of the
// registers of a general double sincos, which also handles huge arguments.
LRT_D void sincos_rn(float xf, float* s, float* c) {
    constexpr double kP1 = 0x1.921fb548p+0, kP2 = -0x1.de973dc8p-31, kP3 = -0x1.9d9cceba3f91fp-62;
    const double x = (double)xf;
    const double k = rint(x * 0.6366197723675814);  // nearest multiple of pi/2
    const double r = ((x - k * kP1) - k * kP2) - k * kP3;
    const double r2 = r * r;
    double ps = 1.9572941063391263e-20;  // (-1)^n / (2n + 1)!, n = 10 .. 1
    ps = fma(ps, r2, -8.22063524662433e-18);
    ps = fma(ps, r2, 2.8114572543455206e-15);
    ps = fma(ps, r2, -7.647163731819816e-13);
    ps = fma(ps, r2, 1.6059043836821613e-10);
    ps = fma(ps, r2, -2.505210838544172e-08);
    ps = fma(ps, r2, 2.7557319223985893e-06);
    ps = fma(ps, r2, -0.0001984126984126984);
    ps = fma(ps, r2, 0.008333333333333333);
    ps = fma(ps, r2, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = -8.896791392450574e-22;  // (-1)^n / (2n)!, n = 11 .. 1
    pc = fma(pc, r2, 4.110317623312165e-19);
    pc = fma(pc, r2, -1.5619206968586225e-16);
    pc = fma(pc, r2, 4.779477332387385e-14);
    pc = fma(pc, r2, -1.1470745597729725e-11);
    pc = fma(pc, r2, 2.08767569878681e-09);
    pc = fma(pc, r2, -2.755731922398589e-07);
    pc = fma(pc, r2, 2.48015873015873e-05);
    pc = fma(pc, r2, -0.001388888888888889);
    pc = fma(pc, r2, 0.041666666666666664);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int q = (int)k & 3;  // sin(r + q pi/2), cos(r + q pi/2)
    double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    if (q & 2) sv = -sv;
    if ((q + 1) & 2) cv = -cv;
    *s = (float)sv;
    *c = (float)cv;
}
// x^(1/4) as two correctly rounded double square roots (within 0.75 double ulp)
LRT_D float qrt_rn(float x) { return (float)sqrt(sqrt((double)x)); }
LRT_D float exp_rn(float x) { return (float)exp((double)x); }

// orthonormal basis around a normal (onb.h:8-23)
struct Frame {
    f3 s, t, n;
};
LRT_D Frame frame_from(f3 nrm) {
    Frame f;
    f.n = normalize(nrm);
    const f3 a = (fabsf(f.n.x) > 0.9f) ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    f.t = normalize(cross(f.n, a));
    f.s = cross(f.n, f.t);
    return f;
}
LRT_D f3 to_world(const Frame& f, f3 a) { return a.x * f.s + a.y * f.t + a.z * f.n; }
LRT_D f3 to_local(const Frame& f, f3 a) { return f3{dot(a, f.s), dot(a, f.t), dot(a, f.n)}; }

// Philox-4x32-10 counter-based generator (Salmon et al., SC'11)
LRT_D uint4 philox(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}
LRT_D float u24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// The reference's sampler (sampler.h:12-32, sobol.h:12-41): dimension d of
// sample `index` is the Gray-code Sobol value of that dimension (a hash
// beyond the table's 16 dimensions) XOR a scramble of (pixel, d), scaled by
// 2^-32.  dirs: the 16 x 32 direction numbers; pixel_seed = mix_bits(pixel).
LRT_D uint32_t mix_bits(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
LRT_D float ref_sample(const uint32_t* __restrict__ dirs, uint32_t pixel_seed, uint32_t index, uint32_t d) {
    uint32_t bits = 0;
    if (d >= 16u) {
        bits = mix_bits(index + d * 0x9e3779b9u);
    } else {
        for (uint32_t g = index ^ (index >> 1); g; g &= g - 1u) bits ^= dirs[d * 32u + (uint32_t)__builtin_ctz(g)];
    }
    bits ^= mix_bits(pixel_seed + d * 0x9e3779b9u);
    return (float)bits * (1.0f / 4294967296.0f);
}
LRT_D uint32_t pick(uint4 v, uint32_t lane) { return lane == 0 ? v.x : (lane == 1 ? v.y : (lane == 2 ? v.z : v.w)); }

}  // namespace lrt
