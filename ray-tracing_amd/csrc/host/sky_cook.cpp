// Hosek-Wilkie RGB sky state cooked on the host in double precision: the
// nine-parameter configuration and the radiance scale of each channel
// (ArHosekSkyModel.cc:147-227 CookConfiguration, :229-289
// CookRadianceConfiguration, :602-635 arhosek_rgb_skymodelstate_alloc_init).
//
// Every quantity of the model is a quintic Bezier curve over the cube root of
// the solar elevation normalised by pi/2, through six control points per
// (ground albedo 0 or 1, integer turbidity 1..10).  The cooked value blends the
// curves of the two albedos and the two integer turbidities around the given
// one bilinearly.  Rounding follows the published model step for step (Bezier
// weights from pow(), terms summed in control-point order, the four corners
// accumulated albedo 0 / 1 at the lower turbidity, then albedo 0 / 1 at the
// upper one, the upper turbidity skipped at turbidity 10), so the state equals
// the model's own to the last bit.
#include <cmath>

#include "host.hpp"

namespace lrt {
namespace {

constexpr double kHalfPi = 3.141592653589793 / 2.0;  // MATH_PI / 2 of the model

struct Bezier5 {
    double w[6];
};

// weights of the six control points at curve parameter s
Bezier5 bezier5(double s) {
    const double r = 1.0 - s;
    Bezier5 b;
    b.w[0] = std::pow(r, 5.0);
    b.w[1] = 5.0 * std::pow(r, 4.0) * s;
    b.w[2] = 10.0 * std::pow(r, 3.0) * std::pow(s, 2.0);
    b.w[3] = 10.0 * std::pow(r, 2.0) * std::pow(s, 3.0);
    b.w[4] = 5.0 * r * std::pow(s, 4.0);
    b.w[5] = std::pow(s, 5.0);
    return b;
}

// the curve through control points p[0], p[stride], ..., p[5 * stride]
double curve(const Bezier5& b, const double* p, int stride) {
    double v = b.w[0] * p[0];
    for (int k = 1; k < 6; ++k) v = v + b.w[k] * p[k * stride];
    return v;
}

// blend of the curve over the albedo / turbidity corners; `block` is the size
// of one (albedo, turbidity) block (6 control points x `stride` values), the
// value is element `elem` of each control point
double blend(const double* data, int block, int stride, int elem, double turbidity, double albedo, const Bezier5& b) {
    const int t = (int)turbidity;
    const double rem = turbidity - (double)t;
    const double* alb0 = data;               // albedo 0: turbidities 1..10
    const double* alb1 = data + 10 * block;  // albedo 1
    double v = (1.0 - albedo) * (1.0 - rem) * curve(b, alb0 + block * (t - 1) + elem, stride);
    v += albedo * (1.0 - rem) * curve(b, alb1 + block * (t - 1) + elem, stride);
    if (t == 10) return v;
    v += (1.0 - albedo) * rem * curve(b, alb0 + block * t + elem, stride);
    v += albedo * rem * curve(b, alb1 + block * t + elem, stride);
    return v;
}

}  // namespace

void sky_cook(const double* coefficients, const double* radiances, double turbidity, double albedo, double elevation,
              double* config27, double* radiance3) {
    // curve parameter: cube root of the elevation over pi/2 (NaN below the
    // horizon, as in the model)
    const Bezier5 b = bezier5(std::pow(elevation / kHalfPi, 1.0 / 3.0));
    for (int c = 0; c < 3; ++c) {
        const double* coef = coefficients + (size_t)c * kHosekCoefficients;
        for (int i = 0; i < 9; ++i) config27[9 * c + i] = blend(coef, 9 * 6, 9, i, turbidity, albedo, b);
        radiance3[c] = blend(radiances + (size_t)c * kHosekRadiances, 6, 1, 0, turbidity, albedo, b);
    }
}

}  // namespace lrt
