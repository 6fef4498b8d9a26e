// Kernels of the BF16 U-Net denoiser: 3x3 convolutions as implicit GEMMs on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), their
// data and weight gradients, and the elementwise layers around them.
//
// Tensors are NHWC with the batch folded into the pixel index: pixel
// q = (b * H + y) * W + x of a tensor with channel stride cs and offset off
// holds its channels at [q * cs + off, q * cs + off + C).  Activations are
// BF16 (uint16 bit patterns, round to nearest even); gradients that are
// summed from several consumers are FP32.
//
// MFMA fragments (one wave, 16x16x32): lane l holds
//   A[m = l % 16][k = 8 * (l / 16) + j]   (j = 0..7, 8 bf16),
//   B[k = 8 * (l / 16) + j][n = l % 16],
//   D[m = 4 * (l / 16) + j][n = l % 16]   (j = 0..3, 4 floats).
// Convolution weights are kept FP32 as W[tap][ci][co] (tap = 3 * ky + kx) and
// packed into B fragments: frag (tap, kc, nt), lane l, element j =
// W[tap][32 kc + 8 (l / 16) + j][16 nt + l % 16], zero outside the real
// channel counts.
#pragma once

#include "common.h"

namespace lrt {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// float -> bf16 rounded to nearest even (v_cvt_pk_bf16_f32; NaN stays a NaN)
LRT_D uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
LRT_D float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kConvPx = 64;        // output pixels per block (one row segment)
constexpr int kConvStage = kConvPx + 2;

// y = act(conv3x3(x) + bias): grid (ceil(W / 64), H, B), 256 threads; wave w
// computes output pixels x0 + 16 w .. + 15 for all COUT channels.  Input rows
// y-1..y+1 (66 columns, cin channels, zero padded) are staged in LDS.
// cin: multiple of 32; xcs, xoff: multiples of 8.  Outputs: bf16 y (channel
// stride ycs, offset yoff) and/or fp32 yf (stride yfcs, offset yfoff,
// accumulate if yf_acc); co >= cout_real is not written.
template <int COUT>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t* __restrict__ x, int H, int W, int cin, int xcs, int xoff,
                                                 const bf16x8* __restrict__ wfrag, const float* __restrict__ bias,
                                                 int cout_real, int relu, uint16_t* __restrict__ y, int ycs, int yoff,
                                                 float* __restrict__ yf, int yfcs, int yfoff, int yf_acc) {
    extern __shared__ __attribute__((aligned(16))) uint16_t conv_tile[];
    const int b = blockIdx.z, row = blockIdx.y, x0 = blockIdx.x * kConvPx;
    const size_t img = (size_t)b * H * W;
    const int cpp = cin >> 3;  // 16-byte chunks per pixel
    const int total = 3 * kConvStage * cpp;
    for (int idx = threadIdx.x; idx < total; idx += 256) {
        const int r = idx / (kConvStage * cpp);
        const int rem = idx - r * (kConvStage * cpp);
        const int px = rem / cpp, ch = rem - px * cpp;
        const int yy = row + r - 1, xx = x0 + px - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            v = *(const uint4*)(x + (img + (size_t)yy * W + xx) * xcs + xoff + ch * 8);
        *(uint4*)(conv_tile + (size_t)(r * kConvStage + px) * cin + ch * 8) = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    constexpr int NT = COUT / 16;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int nkc = cin >> 5;
    for (int tap = 0; tap < 9; ++tap) {
        const int ky = tap / 3, kx = tap - 3 * ky;
        const uint16_t* arow = conv_tile + (size_t)(ky * kConvStage + wave * 16 + i + kx) * cin + kq * 8;
        const bf16x8* bf = wfrag + (size_t)tap * nkc * NT * 64 + lane;
        for (int kc = 0; kc < nkc; ++kc) {
            const bf16x8 a = *(const bf16x8*)(arow + kc * 32);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bf[(kc * NT + t) * 64], acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + i;
        if (co >= cout_real) continue;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int px = x0 + wave * 16 + kq * 4 + j;
            if (px >= W) continue;
            float v = acc[t][j] + bv;
            if (relu) v = fmaxf(v, 0.0f);
            const size_t q = img + (size_t)row * W + px;
            if (y) y[q * ycs + yoff + co] = f2bf(v);
            if (yf) {
                float* o = yf + q * yfcs + yfoff + co;
                *o = yf_acc ? *o + v : v;
            }
        }
    }
}

// ---- fused forward convolution -------------------------------------------
//
// The forward pass reads its input through up to two sources that together
// form the conv's input channels (source 0: channels [0, s0.c), source 1:
// [s0.c, s0.c + s1.c), zero beyond), each read in one of four ways while the
// tile is staged, so no concatenated, upsampled, pooled or feature tensor is
// ever written:
//   kSrcDirect   a tensor at the conv's resolution;
//   kSrcUp2      a tensor at half the resolution, nearest upsampling;
//   kSrcPool2    a tensor at twice the resolution, 2x2 average (the four
//                values summed in float in raster order, times 0.25, rounded
//                to bf16: the arithmetic of k_pool2);
//   kSrcFeatures the network input computed from radiance (3 floats / px) and
//                the first-hit features (7 floats / px): log1p radiance,
//                albedo, normal * 0.5 + 0.5, log1p depth (k_nn_input's
//                arithmetic), channels 10..31 zero.
// The output is bf16 activations, fp32 values, or (kOutImage) the denoised
// radiance expm1(max(log1p(max(rad, 0)) + v, 0)) of channels 0..2.
enum : int { kSrcDirect = 0, kSrcUp2 = 1, kSrcPool2 = 2, kSrcFeatures = 3 };
enum : int { kOutBf16 = 0, kOutF32 = 1, kOutImage = 2 };

struct ConvSrc {
    const uint16_t* p;  // bf16 NHWC tensor (unused for kSrcFeatures)
    int cs, off, c;     // channel stride, channel offset, channel count (multiples of 8)
    int mode;
};
struct ConvIn {
    ConvSrc s0, s1;
    const float* rad;  // kSrcFeatures inputs
    const float* aux;
};
struct ConvOut {
    uint16_t* y;
    int ycs, yoff;
    float* yf;
    int yfcs, yfoff, yf_acc;
    float* img;
    const float* rad;
    int mode;
};

constexpr int kFwdTW = 64;          // output pixels per row of a block tile (four 16-pixel tiles per wave)
constexpr int kFwdSW = kFwdTW + 2;  // staged columns
constexpr int kFwdPitch = 32;       // bf16 per staged pixel: one 32-channel chunk, four 16-byte units

// LDS offset (bf16 elements) of 16-byte unit c (channels 8c .. 8c + 7) of
// staged pixel p.  The units of pixels p and p + 8 are XOR-swizzled apart
// (c ^ 2 when bit 3 of p is set): a fragment read has 16 lanes on 16
// consecutive pixels and two units each per 32-lane group, and with the
// swizzle those 32 reads of 16 bytes fall on 32 different 16-byte bank
// groups.
LRT_D int fwd_lds(int p, int c) { return p * kFwdPitch + ((c ^ ((p >> 2) & 2)) << 3); }

// block shape of k_conv_fwd<NTW>: four waves, one output row each, 4 x 16
// pixels per wave, NTW 16-channel output tiles per block
template <int NTW>
struct FwdShape {
    static constexpr int TR = 4;                             // output rows per tile
    static constexpr int SR = TR + 2;                        // staged rows
    static constexpr int PIX = SR * kFwdSW;                  // staged pixels
    static constexpr int IN_UNITS = PIX * 4;                 // 16-byte input units per 32-channel chunk
    static constexpr int W_UNITS = 9 * NTW * 64;             // 16-byte weight units per chunk (9 taps x NTW fragments)
    static constexpr int PER_IN = (IN_UNITS + 255) / 256;    // per thread
    static constexpr int PER_PIX = (PIX + 255) / 256;
    static constexpr int PER_W = (W_UNITS + 255) / 256;
    static constexpr int IN_ELEMS = PIX * kFwdPitch;         // bf16 elements of the staged input
};

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
// two floats -> packed bf16 pair (a in the low half), one conversion instruction
LRT_D uint32_t bf2x(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// 8 channels [c, c + 8) of source s at pixel (b, yy, xx) of the conv's grid
template <int MODE>
LRT_D uint4 conv_src_unit(const ConvIn& in, const ConvSrc& s, int b, int H, int W, int yy, int xx, int c) {
    // element offsets in 32 bits (the host checks that every tensor has fewer than 2^31 elements)
    if (MODE == kSrcDirect)
        return *(const uint4*)(s.p + (uint32_t)(((b * H + yy) * W + xx) * s.cs + s.off + c));
    if (MODE == kSrcUp2) {
        const int h = H >> 1, w = W >> 1;
        return *(const uint4*)(s.p + (uint32_t)(((b * h + (yy >> 1)) * w + (xx >> 1)) * s.cs + s.off + c));
    }
    if (MODE == kSrcPool2) {
        const int W2 = 2 * W;
        const uint32_t base = (uint32_t)((b * 2 * H + 2 * yy) * W2 + 2 * xx);
        const uint32_t o0 = base * (uint32_t)s.cs + (uint32_t)(s.off + c), ocs = (uint32_t)s.cs;
        union U { uint4 v; uint16_t h[8]; };
        U a, bb, cc, d, o;
        a.v = *(const uint4*)(s.p + o0);
        bb.v = *(const uint4*)(s.p + (o0 + ocs));
        cc.v = *(const uint4*)(s.p + (o0 + (uint32_t)W2 * ocs));
        d.v = *(const uint4*)(s.p + (o0 + ((uint32_t)W2 + 1u) * ocs));
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float sum = bf2f(a.h[k]) + bf2f(bb.h[k]) + bf2f(cc.h[k]) + bf2f(d.h[k]);
            o.h[k] = f2bf(0.25f * sum);
        }
        return o.v;
    }
    return make_uint4(0u, 0u, 0u, 0u);
}

// 8 input channels [ch, ch + 8) of pixel (b, yy, xx) of the conv's virtual
// input; M0 / M1: the modes of sources 0 / 1 (fixed per kernel instance, so a
// layer's registers are sized for its own reads)
template <int M0, int M1>
LRT_D uint4 conv_in_unit(const ConvIn& in, int b, int H, int W, int yy, int xx, int ch) {
    uint4 z = make_uint4(0u, 0u, 0u, 0u);
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) return z;
    const bool second = ch >= in.s0.c;
    if (second && M1 != M0) {
        const int c = ch - in.s0.c;
        return c < in.s1.c ? conv_src_unit<M1>(in, in.s1, b, H, W, yy, xx, c) : z;
    }
    const ConvSrc s = second ? in.s1 : in.s0;
    const int c = second ? ch - in.s0.c : ch;
    return c < s.c ? conv_src_unit<M0>(in, s, b, H, W, yy, xx, c) : z;
}

// the two 16-byte units (channels 0..7, 8..15) of the network input at pixel
// (b, yy, xx), zero outside the image (k_nn_input's arithmetic)
LRT_D void feature_units(const ConvIn& in, int b, int H, int W, int yy, int xx, uint4& u0, uint4& u1) {
    u0 = u1 = make_uint4(0u, 0u, 0u, 0u);
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) return;
    const uint32_t q = (uint32_t)((b * H + yy) * W + xx);
    const float* r = in.rad + 3u * q;
    const float* a = in.aux + 7u * q;
    u0 = make_uint4(bf2x(log1pf(fmaxf(r[0], 0.0f)), log1pf(fmaxf(r[1], 0.0f))), bf2x(log1pf(fmaxf(r[2], 0.0f)), a[0]),
                    bf2x(a[1], a[2]), bf2x(a[3] * 0.5f + 0.5f, a[4] * 0.5f + 0.5f));
    u1 = make_uint4(bf2x(a[5] * 0.5f + 0.5f, log1pf(fmaxf(a[6], 0.0f))), 0u, 0u, 0u);
}

// y = act(conv3x3(x) + bias) over the virtual input `in` (nkc chunks of 32
// channels), output channels [16 NTW g, 16 NTW (g + 1)) of group g =
// blockIdx.y, output written as OUT (out.mode).  Persistent blocks (grid
// (n, groups), 256 threads) walk the tiles of 64 x 4 output pixels
// t = blockIdx.x, + gridDim.x, ...; per chunk
// a tile stages its input (6 rows x 66 columns x 32 channels) and the
// chunk's weight fragments of the group (9 taps x NTW; once per block when
// there is one chunk) in LDS, so the tap loop reads only LDS, and the loads
// of the next (tile, chunk) are in flight while the current one's MFMAs run.
// The MFMA's A operand is the weight fragment and B the staged pixels, so D
// is [channel][pixel]: a lane ends with four consecutive channels of a pixel
// and writes them with one 8-byte store.
template <int NTW, int M0, int M1, int OUT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NTW >= 4 ? 2 : 3))) void k_conv_fwd(ConvIn in, int B, int H, int W, int nkc, int groups,
                                                  const bf16x8* __restrict__ wfrag, const float* __restrict__ bias,
                                                  int cout_real, int relu, ConvOut out) {
    using S = FwdShape<NTW>;
    constexpr bool FEAT = M0 == kSrcFeatures;
    __shared__ __attribute__((aligned(16))) uint16_t tile[S::IN_ELEMS];
    __shared__ __attribute__((aligned(16))) bf16x8 wtile[S::W_UNITS];
    const int g = blockIdx.y;
    const int tiles_x = (W + kFwdTW - 1) / kFwdTW, tiles_y = (H + S::TR - 1) / S::TR;
    const int ntiles = B * tiles_x * tiles_y;
    const int NT = groups * NTW;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    int t = blockIdx.x;
    if (t >= ntiles) return;
    f32x4 acc[4][NTW];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int tt = 0; tt < NTW; ++tt) acc[m][tt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // staging registers of one (tile, chunk)
    uint4 rin[FEAT ? 2 * S::PER_PIX : S::PER_IN];
    bf16x8 rw[S::PER_W];
    auto origin = [&](int tl, int& b, int& y0, int& x0) {
        const int tx = tl % tiles_x, r = tl / tiles_x;
        x0 = tx * kFwdTW;
        y0 = (r % tiles_y) * S::TR;
        b = r / tiles_y;
    };
    int nb = 0, ny0 = 0, nx0 = 0;  // origin of the tile being loaded
    auto load = [&](int tl, int kc, bool weights) {
        int b, y0, x0;
        origin(tl, b, y0, x0);
        nb = b;
        ny0 = y0;
        nx0 = x0;
        if constexpr (FEAT) {
#pragma unroll
            for (int k = 0; k < S::PER_PIX; ++k) {
                const int p = threadIdx.x + 256 * k;
                if (p < S::PIX) {
                    const int r = p / kFwdSW, px = p - r * kFwdSW;
                    feature_units(in, b, H, W, y0 + r - 1, x0 + px - 1, rin[2 * k], rin[2 * k + 1]);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < S::PER_IN; ++k) {
                const int u = threadIdx.x + 256 * k;
                if (u < S::IN_UNITS) {
                    const int c8 = u & 3, pr = u >> 2;
                    const int r = pr / kFwdSW, px = pr - r * kFwdSW;
                    rin[k] = conv_in_unit<M0, M1>(in, b, H, W, y0 + r - 1, x0 + px - 1, kc * 32 + c8 * 8);
                }
            }
        }
        if (weights) {
#pragma unroll
            for (int k = 0; k < S::PER_W; ++k) {
                const int u = threadIdx.x + 256 * k;
                if (u < S::W_UNITS) {
                    const int f = u >> 6;  // tap * NTW + tt
                    const int tap = f / NTW, tt = f - tap * NTW;
                    rw[k] = wfrag[((size_t)(tap * nkc + kc) * NT + g * NTW + tt) * 64 + (u & 63)];
                }
            }
        }
    };
    auto store = [&](bool weights) {
        if constexpr (FEAT) {
#pragma unroll
            for (int k = 0; k < S::PER_PIX; ++k) {
                const int p = threadIdx.x + 256 * k;
                if (p < S::PIX) {
                    *(uint4*)(tile + fwd_lds(p, 0)) = rin[2 * k];
                    *(uint4*)(tile + fwd_lds(p, 1)) = rin[2 * k + 1];
                    *(uint4*)(tile + fwd_lds(p, 2)) = make_uint4(0u, 0u, 0u, 0u);
                    *(uint4*)(tile + fwd_lds(p, 3)) = make_uint4(0u, 0u, 0u, 0u);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < S::PER_IN; ++k) {
                const int u = threadIdx.x + 256 * k;
                if (u < S::IN_UNITS) *(uint4*)(tile + fwd_lds(u >> 2, u & 3)) = rin[k];
            }
        }
        if (weights) {
#pragma unroll
            for (int k = 0; k < S::PER_W; ++k) {
                const int u = threadIdx.x + 256 * k;
                if (u < S::W_UNITS) wtile[u] = rw[k];
            }
        }
    };
    const bool w_once = nkc == 1;
    // this lane's bias values (its four channels of each output tile)
    float bias_r[NTW][4];
#pragma unroll
    for (int tt = 0; tt < NTW; ++tt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int co = (g * NTW + tt) * 16 + 4 * kq + j;
            bias_r[tt][j] = (bias && co < cout_real) ? bias[co] : 0.0f;
        }
    int kc = 0;
    bool first = true;
    load(t, 0, true);
    int cb = nb, cy0 = ny0, cx0 = nx0;  // origin of the tile being computed
    for (;;) {
        if (!first) __syncthreads();
        store(first || !w_once);
        __syncthreads();
        first = false;
        int tn = t, kcn = kc + 1;
        if (kcn == nkc) {
            kcn = 0;
            tn = t + gridDim.x;
        }
        const bool more = tn < ntiles;
        if (more) load(tn, kcn, !w_once);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap - 3 * ky;
            bf16x8 bw[NTW];
#pragma unroll
            for (int tt = 0; tt < NTW; ++tt) bw[tt] = wtile[(tap * NTW + tt) * 64 + lane];
            // pixel tiles m are 16 pixels apart: same swizzle, offset 16 m pitches
            const uint16_t* arow = tile + fwd_lds((wave + ky) * kFwdSW + i + kx, kq);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const bf16x8 a = *(const bf16x8*)(arow + 16 * m * kFwdPitch);
#pragma unroll
                for (int tt = 0; tt < NTW; ++tt)
                    acc[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[tt], a, acc[m][tt], 0, 0, 0);
            }
        }
        if (kc == nkc - 1) {
            const int row = cy0 + wave, x0 = cx0;
            if constexpr (OUT == kOutImage) {
                // channels 0..2 of pixel i sit in lane (i, kq = 0) of tile 0;
                // lane (i, j), j < 3, takes channel j of that pixel, so 48
                // lanes share the output transform instead of 16
                if (row < H) {
                    const uint32_t rowq = (uint32_t)((cb * H + row) * W);
                    const float bk = (bias && kq < 3 && kq < cout_real) ? bias[kq] : 0.0f;  // channel kq's bias
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const float c0 = __shfl(acc[m][0][0], i), c1 = __shfl(acc[m][0][1], i),
                                    c2 = __shfl(acc[m][0][2], i);
                        const int px = x0 + 16 * m + i;
                        if (g == 0 && kq < 3 && kq < cout_real && px < W) {
                            float v = (kq == 0 ? c0 : (kq == 1 ? c1 : c2)) + bk;
                            if (relu) v = fmaxf(v, 0.0f);
                            const uint32_t q = rowq + (uint32_t)px;
                            const float l = log1pf(fmaxf(out.rad[3u * q + kq], 0.0f)) + v;
                            out.img[3u * q + kq] = expm1f(fmaxf(l, 0.0f));
                        }
                    }
                }
            } else if (row < H) {
                const uint32_t rowq = (uint32_t)((cb * H + row) * W);
#pragma unroll
                for (int tt = 0; tt < NTW; ++tt) {
                    const int co0 = (g * NTW + tt) * 16 + 4 * kq;  // this lane's four channels co0 .. co0 + 3
                    if (co0 >= cout_real) continue;
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const int px = x0 + 16 * m + i;
                        if (px >= W) continue;
                        float v[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            v[j] = acc[m][tt][j] + bias_r[tt][j];
                            if (relu) v[j] = fmaxf(v[j], 0.0f);
                        }
                        const uint32_t q = rowq + (uint32_t)px;
                        if constexpr (OUT == kOutBf16) {
                            uint16_t* o = out.y + (q * (uint32_t)out.ycs + (uint32_t)(out.yoff + co0));
                            if (co0 + 3 < cout_real) {
                                *(uint2*)o = make_uint2(bf2x(v[0], v[1]), bf2x(v[2], v[3]));
                            } else {
#pragma unroll
                                for (int j = 0; j < 4; ++j)
                                    if (co0 + j < cout_real) o[j] = f2bf(v[j]);
                            }
                        } else if constexpr (OUT == kOutF32) {
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (co0 + j < cout_real) {
                                    float* o = out.yf + q * out.yfcs + out.yfoff + co0 + j;
                                    *o = out.yf_acc ? *o + v[j] : v[j];
                                }
                        }
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int tt = 0; tt < NTW; ++tt) acc[m][tt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        if (!more) break;
        t = tn;
        kc = kcn;
        cb = nb;
        cy0 = ny0;
        cx0 = nx0;
    }
}

// dW[tap][ci][co] += sum_q X[q shifted by tap][ci] * G[q][co] and, from the
// blocks with tap 0 and ci-tile 0, db[co] += sum_q G[q][co].
// grid (chunks, 9, cin / 16), 256 threads; each wave takes 32-pixel k-chunks
// of the block's pixel range and all COUT / 16 output tiles; the four waves
// are summed in LDS and one atomic add per weight is issued per block.
template <int COUT>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const uint16_t* __restrict__ x, int B, int H, int W, int xcs,
                                                       int xoff, int cin_real, const uint16_t* __restrict__ g, int gcs,
                                                       int goff, int cout_real, int px_per_block,
                                                       float* __restrict__ dw, float* __restrict__ db) {
    __shared__ float red[4][COUT / 16][64][4];
    const int tap = blockIdx.y, ky = tap / 3, kx = tap - 3 * ky;
    const int ci0 = blockIdx.z * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    constexpr int NT = COUT / 16;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bsum[t] = 0.0f;
    const bool do_bias = db && tap == 0 && blockIdx.z == 0;
    const long long P = (long long)B * H * W;
    const long long q0 = (long long)blockIdx.x * px_per_block;
    const long long q1 = min(P, q0 + px_per_block);
    const int ci = ci0 + i;
    for (long long kb = q0 + wave * 32; kb < q1; kb += 128) {
        union { bf16x8 v; uint16_t s[8]; } a;
        union { bf16x8 v; uint16_t s[8]; } bm[NT];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long long q = kb + kq * 8 + j;
            uint16_t av = 0;
            if (q < q1 && ci < cin_real) {
                const int xq = (int)(q % W);
                const long long rq = q / W;
                const int yq = (int)(rq % H);
                const int yy = yq + ky - 1, xx = xq + kx - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) av = x[(q + (long long)(ky - 1) * W + (kx - 1)) * xcs + xoff + ci];
            }
            a.s[j] = av;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int co = t * 16 + i;
                uint16_t gv = 0;
                if (q < q1 && co < cout_real) gv = g[q * gcs + goff + co];
                bm[t].s[j] = gv;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bm[t].v, acc[t], 0, 0, 0);
            if (do_bias)
                for (int j = 0; j < 8; ++j) bsum[t] += bf2f(bm[t].s[j]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[wave][t][lane][j] = acc[t][j];
    __syncthreads();
    // D[m = ci][n = co]: lane l, element j -> ci0 + 4 (l / 16) + j, co = 16 t + l % 16
    if (wave == 0) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float s = red[0][t][lane][j] + red[1][t][lane][j] + red[2][t][lane][j] + red[3][t][lane][j];
                const int cim = ci0 + kq * 4 + j, co = t * 16 + i;
                if (cim < cin_real && co < cout_real && s != 0.0f) atomicAdd(dw + ((size_t)tap * cin_real + cim) * cout_real + co, s);
            }
    }
    if (do_bias) {
        __syncthreads();
        float* rb = &red[0][0][0][0];
#pragma unroll
        for (int t = 0; t < NT; ++t) rb[(t * 256) + threadIdx.x] = bsum[t];
        __syncthreads();
        if (threadIdx.x < COUT) {
            const int t = threadIdx.x / 16, n = threadIdx.x % 16;
            float s = 0.0f;
            for (int w = 0; w < 4; ++w)
                for (int q4 = 0; q4 < 4; ++q4) s += rb[t * 256 + w * 64 + q4 * 16 + n];
            if (threadIdx.x < cout_real && s != 0.0f) atomicAdd(db + threadIdx.x, s);
        }
    }
}

// fp32 W[tap][ci][co] (cin x cout real) -> B fragments for a conv with
// cin_pad input and cout_pad output channels; flip: the data-gradient
// weights W'[tap][co][ci] = W[8 - tap][ci][co] (input channels = cout).
__global__ void k_pack_weights(const float* __restrict__ w, int cin, int cout, int flip, int kin_pad, int nout_pad,
                               bf16x8* __restrict__ frag) {
    const int nkc = kin_pad / 32, nt = nout_pad / 16;
    const int total = 9 * nkc * nt * 64;
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < total; id += gridDim.x * blockDim.x) {
        const int lane = id & 63;
        int rest = id >> 6;
        const int t = rest % nt;
        rest /= nt;
        const int kc = rest % nkc;
        const int tap = rest / nkc;
        union { bf16x8 v; uint16_t s[8]; } o;
        for (int j = 0; j < 8; ++j) {
            const int kin = kc * 32 + 8 * (lane >> 4) + j;  // input channel of this conv
            const int nout = t * 16 + (lane & 15);          // output channel of this conv
            float v = 0.0f;
            if (!flip) {
                if (kin < cin && nout < cout) v = w[((size_t)tap * cin + kin) * cout + nout];
            } else {
                if (kin < cout && nout < cin) v = w[((size_t)(8 - tap) * cin + nout) * cout + kin];
            }
            o.s[j] = f2bf(v);
        }
        frag[id] = o.v;
    }
}

// G = bf16(dY * (Y > 0)) (relu) or bf16(dY); channels [0, c), zero in [c, gcs)
__global__ void k_relu_grad(const float* __restrict__ dy, int dycs, int dyoff, const uint16_t* __restrict__ y, int ycs,
                            int yoff, int relu, long long P, int c, uint16_t* __restrict__ g, int gcs) {
    const long long total = P * gcs;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total; id += (long long)gridDim.x * blockDim.x) {
        const long long q = id / gcs;
        const int ch = (int)(id - q * gcs);
        float v = 0.0f;
        if (ch < c) {
            v = dy[q * dycs + dyoff + ch];
            if (relu && !(bf2f(y[q * ycs + yoff + ch]) > 0.0f)) v = 0.0f;
        }
        g[id] = f2bf(v);
    }
}

// 2x2 average pool (H, W even): bf16 -> bf16
__global__ void k_pool2(const uint16_t* __restrict__ x, int B, int H, int W, int c, int xcs, int xoff,
                        uint16_t* __restrict__ y, int ycs, int yoff) {
    const int Ho = H / 2, Wo = W / 2;
    const long long total = (long long)B * Ho * Wo * c;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total; id += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(id % c);
        long long q = id / c;
        const int xo = (int)(q % Wo);
        q /= Wo;
        const int yo = (int)(q % Ho);
        const int b = (int)(q / Ho);
        const long long base = ((long long)b * H + 2 * yo) * W + 2 * xo;
        const float s = bf2f(x[base * xcs + xoff + ch]) + bf2f(x[(base + 1) * xcs + xoff + ch]) +
                        bf2f(x[(base + W) * xcs + xoff + ch]) + bf2f(x[(base + W + 1) * xcs + xoff + ch]);
        y[(((long long)b * Ho + yo) * Wo + xo) * ycs + yoff + ch] = f2bf(0.25f * s);
    }
}

// gradient of the 2x2 average pool: dx += dy / 4 (fp32)
__global__ void k_pool2_grad(const float* __restrict__ dy, int B, int H, int W, int c, int dycs, int dyoff,
                             float* __restrict__ dx, int dxcs, int dxoff) {
    const long long total = (long long)B * H * W * c;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total; id += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(id % c);
        long long q = id / c;
        const int xx = (int)(q % W);
        q /= W;
        const int yy = (int)(q % H);
        const int b = (int)(q / H);
        const long long qi = ((long long)b * H + yy) * W + xx;
        const long long qo = ((long long)b * (H / 2) + yy / 2) * (W / 2) + xx / 2;
        dx[qi * dxcs + dxoff + ch] += 0.25f * dy[qo * dycs + dyoff + ch];
    }
}

// nearest 2x upsample (output H, W): bf16 -> bf16
__global__ void k_up2(const uint16_t* __restrict__ x, int B, int H, int W, int c, int xcs, int xoff,
                      uint16_t* __restrict__ y, int ycs, int yoff) {
    const long long total = (long long)B * H * W * c;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total; id += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(id % c);
        long long q = id / c;
        const int xx = (int)(q % W);
        q /= W;
        const int yy = (int)(q % H);
        const int b = (int)(q / H);
        const long long qi = ((long long)b * (H / 2) + yy / 2) * (W / 2) + xx / 2;
        y[(((long long)b * H + yy) * W + xx) * ycs + yoff + ch] = x[qi * xcs + xoff + ch];
    }
}

// gradient of the nearest upsample: dx = sum over the 2x2 block (fp32)
__global__ void k_up2_grad(const float* __restrict__ dy, int B, int H, int W, int c, int dycs, int dyoff,
                           float* __restrict__ dx, int dxcs, int dxoff) {
    const int Hi = H / 2, Wi = W / 2;
    const long long total = (long long)B * Hi * Wi * c;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total; id += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(id % c);
        long long q = id / c;
        const int xi = (int)(q % Wi);
        q /= Wi;
        const int yi = (int)(q % Hi);
        const int b = (int)(q / Hi);
        const long long base = ((long long)b * H + 2 * yi) * W + 2 * xi;
        const float s = dy[base * dycs + dyoff + ch] + dy[(base + 1) * dycs + dyoff + ch] +
                        dy[(base + W) * dycs + dyoff + ch] + dy[(base + W + 1) * dycs + dyoff + ch];
        dx[(((long long)b * Hi + yi) * Wi + xi) * dxcs + dxoff + ch] = s;
    }
}

// network input: log1p(radiance), albedo, normal * 0.5 + 0.5, log1p(depth);
// channels 10..cs-1 zero.  rad: 3 floats per pixel, aux: 7 (albedo, normal, depth)
__global__ void k_nn_input(const float* __restrict__ rad, const float* __restrict__ aux, long long P,
                           uint16_t* __restrict__ x, int cs) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < P; q += (long long)gridDim.x * blockDim.x) {
        float v[10];
        for (int c = 0; c < 3; ++c) v[c] = log1pf(fmaxf(rad[3 * q + c], 0.0f));
        for (int c = 0; c < 3; ++c) v[3 + c] = aux[7 * q + c];
        for (int c = 0; c < 3; ++c) v[6 + c] = aux[7 * q + 3 + c] * 0.5f + 0.5f;
        v[9] = log1pf(fmaxf(aux[7 * q + 6], 0.0f));
        uint16_t* o = x + q * cs;
        for (int c = 0; c < cs; ++c) o[c] = c < 10 ? f2bf(v[c]) : (uint16_t)0;
    }
}

// output (residual in the log domain): out = expm1(max(log1p(rad) + y, 0)),
// y: fp32, channel stride ycs
__global__ void k_nn_output(const float* __restrict__ y, int ycs, const float* __restrict__ rad, long long P,
                            float* __restrict__ out) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < P; q += (long long)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            const float l = log1pf(fmaxf(rad[3 * q + c], 0.0f)) + y[q * ycs + c];
            out[3 * q + c] = expm1f(fmaxf(l, 0.0f));
        }
}

// L1 loss in the log domain between log1p(rad) + y and log1p(target);
// dy = sign / (3 P) into channels 0..2 of dy (stride dycs); loss sum atomically
__global__ void k_nn_loss(const float* __restrict__ y, int ycs, const float* __restrict__ rad,
                          const float* __restrict__ target, long long P, float* __restrict__ dy, int dycs,
                          float* __restrict__ loss) {
    float part = 0.0f;
    const float inv = 1.0f / (3.0f * (float)P);
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < P; q += (long long)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            const float pred = log1pf(fmaxf(rad[3 * q + c], 0.0f)) + y[q * ycs + c];
            const float d = pred - log1pf(fmaxf(target[3 * q + c], 0.0f));
            part += fabsf(d);
            dy[q * dycs + c] = d > 0.0f ? inv : (d < 0.0f ? -inv : 0.0f);
        }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, part * inv);
}

// Adam (bias-corrected) over n parameters
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void k_zero_f32(float* __restrict__ p, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0.0f;
}

__global__ void k_f32_to_bf16(const float* __restrict__ s, long long n, uint16_t* __restrict__ d) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) d[i] = f2bf(s[i]);
}

}  // namespace lrt
