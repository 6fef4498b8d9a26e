// The learned denoiser: a 3-level U-Net (channels 32 / 64 / 128, 3x3
// convolutions, 2x2 average pooling, nearest upsampling, concatenated skips)
// over 10 input channels (log radiance, albedo, normal, log depth) that
// predicts the log-domain residual of the clean radiance.  Every layer runs on
// the kernels of device/nn.h: BF16 activations and weights on the matrix
// cores with FP32 accumulation, FP32 master weights and Adam moments.
//
//   e0a 10->32, e0b 32->32 (skip 0) | pool | e1a 32->64, e1b 64->64 (skip 1) |
//   pool | e2a 64->128, e2b 128->128 | up, concat skip 1 | d1a 192->64,
//   d1b 64->64 | up, concat skip 0 | d0a 96->32 | out 32->3 (linear)
//
// Skips are written by their producer straight into the channel slice of the
// concat buffer the decoder reads.  The forward convolutions read pooled,
// upsampled and concatenated inputs and the network's input features through
// their staging (k_conv_fwd), so inference writes none of those tensors; the
// training step materialises the ones its weight gradients read.
#include <map>
#include <mutex>

#include "device/nn.h"

namespace lrt {

struct NnLayer {
    int cin, cout, cinp, coutp, relu;
    size_t woff, boff;            // offsets of W[9][cin][cout] and b[cout] in the parameter vector
    DevBuf fwd, bwd;              // packed B fragments (forward; data gradient)
};

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

template <int C>
static void conv_launch(hipStream_t s, const uint16_t* x, int B, int H, int W, int cin, int xcs, int xoff, const void* frag,
                        const float* bias, int cout_real, int relu, uint16_t* y, int ycs, int yoff, float* yf, int yfcs,
                        int yfoff, int yf_acc) {
    const size_t lds = (size_t)3 * kConvStage * cin * 2;
    static size_t attr = 0;  // largest dynamic LDS size enabled so far
    if (lds > attr) {
        HIPCHECK(hipFuncSetAttribute((const void*)k_conv3x3<C>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = lds;
    }
    hipLaunchKernelGGL(k_conv3x3<C>, dim3((W + kConvPx - 1) / kConvPx, H, B), dim3(256), lds, s, x, H, W, cin, xcs, xoff,
                       (const bf16x8*)frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc);
}

// conv with a runtime output-channel count (padded to 16)
static void conv(hipStream_t s, int coutp, const uint16_t* x, int B, int H, int W, int cin, int xcs, int xoff,
                 const void* frag, const float* bias, int cout_real, int relu, uint16_t* y, int ycs, int yoff, float* yf,
                 int yfcs, int yfoff, int yf_acc) {
    require(cin % 32 == 0 && xcs % 8 == 0 && xoff % 8 == 0, "conv: input channels must be padded to 32");
    switch (coutp) {
        case 16: conv_launch<16>(s, x, B, H, W, cin, xcs, xoff, frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc); break;
        case 32: conv_launch<32>(s, x, B, H, W, cin, xcs, xoff, frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc); break;
        case 64: conv_launch<64>(s, x, B, H, W, cin, xcs, xoff, frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc); break;
        case 96: conv_launch<96>(s, x, B, H, W, cin, xcs, xoff, frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc); break;
        case 128: conv_launch<128>(s, x, B, H, W, cin, xcs, xoff, frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc); break;
        case 192: conv_launch<192>(s, x, B, H, W, cin, xcs, xoff, frag, bias, cout_real, relu, y, ycs, yoff, yf, yfcs, yfoff, yf_acc); break;
        default: throw Error(LRT_ERR_UNSUPPORTED, "conv: unsupported output channel count " + std::to_string(coutp));
    }
    HIPCHECK(hipGetLastError());
}

// a direct bf16 source / a bf16 destination
static ConvSrc src(const uint16_t* p, int cs, int off, int c, int mode = kSrcDirect) { return ConvSrc{p, cs, off, c, mode}; }
static ConvIn in1(ConvSrc a) { return ConvIn{a, ConvSrc{nullptr, 0, 0, 0, kSrcDirect}, nullptr, nullptr}; }
static ConvIn in2(ConvSrc a, ConvSrc b) { return ConvIn{a, b, nullptr, nullptr}; }
static ConvOut out_bf16(uint16_t* y, int ycs, int yoff) {
    return ConvOut{y, ycs, yoff, nullptr, 0, 0, 0, nullptr, nullptr, kOutBf16};
}
static ConvOut out_f32(float* yf, int yfcs, int yfoff, int acc) {
    return ConvOut{nullptr, 0, 0, yf, yfcs, yfoff, acc, nullptr, nullptr, kOutF32};
}

// blocks of a 256-thread kernel the current device holds at once (all CUs),
// per kernel
static int conv_resident_blocks(const void* kern) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> known;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = known.find({kern, dev});
    if (it != known.end()) return it->second;
    int per_cu = 0, cus = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int n = std::max(per_cu, 1) * std::max(cus, 1);
    known[{kern, dev}] = n;
    return n;
}

template <int M0, int M1, int OUT>
static void conv_fwd_nt(hipStream_t s, int coutp, const ConvIn& in, int B, int H, int W, int cin, const void* frag,
                        const float* bias, int cout_real, int relu, const ConvOut& out) {
    const int nkc = cin / 32;
    const bf16x8* f = (const bf16x8*)frag;
    auto launch = [&](auto kern, int ntw) {
        const int groups = coutp / 16 / ntw;
        const long long tiles = (long long)B * ((W + kFwdTW - 1) / kFwdTW) * ((H + 3) / 4);
        // persistent blocks (as many as the CUs hold at once): each tile's
        // loads overlap the previous tile's MFMAs, and with one chunk the
        // weights are staged once per block (measured against one block per
        // tile: 0.987 -> 0.964 ms per 1080p forward, the pooled 64->128 layer
        // 0.087 -> 0.062 ms, the 128->128 layer 0.065 -> 0.069 ms)
        const long long nb =
            std::min<long long>(tiles, std::max<long long>(1, (long long)conv_resident_blocks((const void*)kern) / groups));
        hipLaunchKernelGGL(kern, dim3((uint32_t)nb, groups), dim3(256), 0, s, in, B, H, W, nkc, groups, f, bias, cout_real,
                           relu, out);
    };
    switch (coutp) {
        case 16: launch(k_conv_fwd<1, M0, M1, OUT>, 1); break;
        case 32: launch(k_conv_fwd<2, M0, M1, OUT>, 2); break;
        case 64: launch(k_conv_fwd<4, M0, M1, OUT>, 4); break;
        case 128: launch(k_conv_fwd<4, M0, M1, OUT>, 4); break;
        default: throw Error(LRT_ERR_UNSUPPORTED, "conv: unsupported output channel count " + std::to_string(coutp));
    }
}

// the fused forward convolution over a virtual input of cin channels (cin a
// multiple of 32, the sources' channel counts adding up to at most cin)
static void conv_fwd(hipStream_t s, int coutp, const ConvIn& in, int B, int H, int W, int cin, const void* frag,
                     const float* bias, int cout_real, int relu, const ConvOut& out) {
    require(cin % 32 == 0 && in.s0.c % 8 == 0 && in.s1.c % 8 == 0 && in.s0.c + in.s1.c <= cin,
            "conv: input channels must be padded to 32");
    for (const ConvSrc* c : {&in.s0, &in.s1})
        if (c->c && c->mode != kSrcFeatures)
            require(c->p != nullptr && c->cs % 8 == 0 && c->off % 8 == 0 && c->off + c->c <= c->cs, "conv: bad source");
    if (in.s0.mode == kSrcUp2 || in.s1.mode == kSrcUp2) require(H % 2 == 0 && W % 2 == 0, "conv: upsampled source needs even H, W");
    // the kernel's element offsets are 32-bit: every tensor it reads or writes holds fewer than 2^31 elements
    const long long px = (long long)B * H * W, lim = 1ll << 31;
    for (const ConvSrc* c : {&in.s0, &in.s1}) {
        if (!c->c) continue;
        const long long n = c->mode == kSrcFeatures ? px * 7 : c->mode == kSrcUp2 ? px / 4 * c->cs
                          : c->mode == kSrcPool2 ? px * 4 * c->cs : px * c->cs;
        require(n < lim, "conv: input tensor too large for 32-bit offsets");
    }
    require(px * std::max({out.ycs, out.yfcs, 3}) < lim, "conv: output tensor too large for 32-bit offsets");
    const int m0 = in.s0.mode, m1 = in.s1.c ? in.s1.mode : m0;
    // the combinations the network and lrt_conv3x3 use
    if (m0 == kSrcFeatures && out.mode == kOutBf16)
        conv_fwd_nt<kSrcFeatures, kSrcFeatures, kOutBf16>(s, coutp, in, B, H, W, cin, frag, bias, cout_real, relu, out);
    else if (m0 == kSrcDirect && m1 == kSrcDirect && out.mode == kOutBf16)
        conv_fwd_nt<kSrcDirect, kSrcDirect, kOutBf16>(s, coutp, in, B, H, W, cin, frag, bias, cout_real, relu, out);
    else if (m0 == kSrcDirect && m1 == kSrcDirect && out.mode == kOutF32)
        conv_fwd_nt<kSrcDirect, kSrcDirect, kOutF32>(s, coutp, in, B, H, W, cin, frag, bias, cout_real, relu, out);
    else if (m0 == kSrcDirect && m1 == kSrcDirect && out.mode == kOutImage)
        conv_fwd_nt<kSrcDirect, kSrcDirect, kOutImage>(s, coutp, in, B, H, W, cin, frag, bias, cout_real, relu, out);
    else if (m0 == kSrcPool2 && m1 == kSrcPool2 && out.mode == kOutBf16)
        conv_fwd_nt<kSrcPool2, kSrcPool2, kOutBf16>(s, coutp, in, B, H, W, cin, frag, bias, cout_real, relu, out);
    else if (m0 == kSrcUp2 && m1 == kSrcDirect && out.mode == kOutBf16)
        conv_fwd_nt<kSrcUp2, kSrcDirect, kOutBf16>(s, coutp, in, B, H, W, cin, frag, bias, cout_real, relu, out);
    else
        throw Error(LRT_ERR_UNSUPPORTED, "conv: unsupported combination of input sources and output");
    HIPCHECK(hipGetLastError());
}

template <int C>
static void wgrad_launch(hipStream_t s, int cus, const uint16_t* x, int B, int H, int W, int xcs, int xoff, int cin,
                         const uint16_t* g, int gcs, int goff, int cout, float* dw, float* db) {
    const long long P = (long long)B * H * W;
    const int tiles = (cin + 15) / 16;
    // enough blocks to fill the chip, each covering a multiple of 128 pixels
    long long chunks = std::max<long long>(1, std::min<long long>((P + 127) / 128, (long long)cus * 8 / (9 * tiles) + 1));
    const int per = (int)(((P + chunks - 1) / chunks + 127) / 128 * 128);
    chunks = (P + per - 1) / per;
    hipLaunchKernelGGL(k_conv3x3_wgrad<C>, dim3((uint32_t)chunks, 9, tiles), dim3(256), 0, s, x, B, H, W, xcs, xoff, cin, g,
                       gcs, goff, cout, per, dw, db);
    HIPCHECK(hipGetLastError());
}

static void wgrad(hipStream_t s, int cus, int coutp, const uint16_t* x, int B, int H, int W, int xcs, int xoff, int cin,
                  const uint16_t* g, int gcs, int goff, int cout, float* dw, float* db) {
    switch (coutp) {
        case 16: wgrad_launch<16>(s, cus, x, B, H, W, xcs, xoff, cin, g, gcs, goff, cout, dw, db); break;
        case 32: wgrad_launch<32>(s, cus, x, B, H, W, xcs, xoff, cin, g, gcs, goff, cout, dw, db); break;
        case 64: wgrad_launch<64>(s, cus, x, B, H, W, xcs, xoff, cin, g, gcs, goff, cout, dw, db); break;
        case 128: wgrad_launch<128>(s, cus, x, B, H, W, xcs, xoff, cin, g, gcs, goff, cout, dw, db); break;
        default: throw Error(LRT_ERR_UNSUPPORTED, "wgrad: unsupported output channel count " + std::to_string(coutp));
    }
}

static uint32_t grid1(long long n, int cus) {
    return (uint32_t)std::max<long long>(1, std::min<long long>((n + 255) / 256, (long long)cus * 16));
}

}  // namespace lrt

struct lrt_denoiser {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    NnLayer L[10];
    size_t n_params = 0;
    DevBuf params, grads, m, v, loss;
    int adam_t = 0;
    // activations and gradients for the current shape
    int B = 0, H = 0, W = 0;
    DevBuf x0, a0, cat0, p1, a1, cat1, p2, a2, b2, d1a, d1b, d0a, out;
    DevBuf g, d_out, d_d0a, d_cat0, d_d1b, d_d1a, d_cat1, d_b2, d_a2, d_p2, d_a1, d_p1, d_a0;
    DevBuf in_rad, in_aux, in_tgt, res;
    // per-layer timing of lrt_denoiser_forward (lrt_denoiser_timing): events
    // around each convolution, their sums since the last read
    bool timing = false;
    hipEvent_t ev[11] = {};
    double layer_ms[10] = {};
    uint64_t timed_calls = 0;
    hipEvent_t caller = nullptr;  // after_caller's event
    ~lrt_denoiser() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (caller) (void)hipEventDestroy(caller);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace lrt {

// the U-Net's layers: (cin, cout, relu)
static const int kNet[10][3] = {{10, 32, 1},  {32, 32, 1}, {32, 64, 1}, {64, 64, 1}, {64, 128, 1},
                                {128, 128, 1}, {192, 64, 1}, {64, 64, 1}, {96, 32, 1}, {32, 3, 0}};

static void pack_all(lrt_denoiser& D) {
    const float* p = D.params.as<float>();
    for (NnLayer& l : D.L) {
        const int nf = 9 * (l.cinp / 32) * (l.coutp / 16) * 64;
        hipLaunchKernelGGL(k_pack_weights, dim3(grid1(nf, D.cus)), dim3(256), 0, D.stream, p + l.woff, l.cin, l.cout, 0,
                           l.cinp, l.coutp, l.fwd.as<bf16x8>());
        const int gin = round_up(l.coutp, 32);  // data-gradient conv: input channels = padded cout
        const int nb = 9 * (gin / 32) * (l.cinp / 16) * 64;
        hipLaunchKernelGGL(k_pack_weights, dim3(grid1(nb, D.cus)), dim3(256), 0, D.stream, p + l.woff, l.cin, l.cout, 1, gin,
                           l.cinp, l.bwd.as<bf16x8>());
    }
    HIPCHECK(hipGetLastError());
}

static void shape(lrt_denoiser& D, int B, int H, int W) {
    require(B >= 1 && H >= 4 && W >= 4 && H % 4 == 0 && W % 4 == 0, "denoiser: H and W must be multiples of 4");
    if (D.B == B && D.H == H && D.W == W) return;
    const size_t P = (size_t)B * H * W, P1 = P / 4, P2 = P / 16;
    D.x0.alloc(P * 32 * 2);
    D.a0.alloc(P * 32 * 2);
    D.cat0.alloc(P * 96 * 2);
    D.p1.alloc(P1 * 32 * 2);
    D.a1.alloc(P1 * 64 * 2);
    D.cat1.alloc(P1 * 192 * 2);
    D.p2.alloc(P2 * 64 * 2);
    D.a2.alloc(P2 * 128 * 2);
    D.b2.alloc(P2 * 128 * 2);
    D.d1a.alloc(P1 * 64 * 2);
    D.d1b.alloc(P1 * 64 * 2);
    D.d0a.alloc(P * 32 * 2);
    D.out.alloc(P * 16 * 4);
    D.B = B;
    D.H = H;
    D.W = W;
}

static void shape_grad(lrt_denoiser& D) {
    const size_t P = (size_t)D.B * D.H * D.W, P1 = P / 4, P2 = P / 16;
    D.g.alloc(P * 32 * 2);
    D.d_out.alloc(P * 16 * 4);
    D.d_d0a.alloc(P * 32 * 4);
    D.d_cat0.alloc(P * 96 * 4);
    D.d_d1b.alloc(P1 * 64 * 4);
    D.d_d1a.alloc(P1 * 64 * 4);
    D.d_cat1.alloc(P1 * 192 * 4);
    D.d_b2.alloc(P2 * 128 * 4);
    D.d_a2.alloc(P2 * 128 * 4);
    D.d_p2.alloc(P2 * 64 * 4);
    D.d_a1.alloc(P1 * 64 * 4);
    D.d_p1.alloc(P1 * 32 * 4);
    D.d_a0.alloc(P * 32 * 4);
}

// forward pass from device radiance (3 / px) and aux (7 / px): the denoised
// radiance into img (3 floats / px), or with `train` the last layer's fp32
// output into D.out (16 / px) and the tensors the weight gradients read (x0,
// p1, p2 and the upsampled halves of cat0 / cat1), which the fused
// convolutions themselves never write
static void forward(lrt_denoiser& D, const float* rad, const float* aux, bool train, float* img) {
    hipStream_t s = D.stream;
    const int B = D.B, H = D.H, W = D.W, H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
    const long long P = (long long)B * H * W;
    const float* p = D.params.as<float>();
    auto wf = [&](int i) { return (const void*)D.L[i].fwd.p; };
    auto bias = [&](int i) { return p + D.L[i].boff; };
    auto mark = [&](int k) {
        if (D.timing) HIPCHECK(hipEventRecord(D.ev[k], s));
    };
    uint16_t *a0 = D.a0.as<uint16_t>(), *cat0 = D.cat0.as<uint16_t>(), *a1 = D.a1.as<uint16_t>(),
             *cat1 = D.cat1.as<uint16_t>(), *a2 = D.a2.as<uint16_t>(), *b2 = D.b2.as<uint16_t>(),
             *d1a = D.d1a.as<uint16_t>(), *d1b = D.d1b.as<uint16_t>(), *d0a = D.d0a.as<uint16_t>();
    ConvIn feat = in1(ConvSrc{nullptr, 0, 0, 32, kSrcFeatures});
    feat.rad = rad;
    feat.aux = aux;
    mark(0);
    conv_fwd(s, 32, feat, B, H, W, 32, wf(0), bias(0), 32, 1, out_bf16(a0, 32, 0));
    mark(1);
    conv_fwd(s, 32, in1(src(a0, 32, 0, 32)), B, H, W, 32, wf(1), bias(1), 32, 1, out_bf16(cat0, 96, 64));
    mark(2);
    conv_fwd(s, 64, in1(src(cat0, 96, 64, 32, kSrcPool2)), B, H1, W1, 32, wf(2), bias(2), 64, 1, out_bf16(a1, 64, 0));
    mark(3);
    conv_fwd(s, 64, in1(src(a1, 64, 0, 64)), B, H1, W1, 64, wf(3), bias(3), 64, 1, out_bf16(cat1, 192, 128));
    mark(4);
    conv_fwd(s, 128, in1(src(cat1, 192, 128, 64, kSrcPool2)), B, H2, W2, 64, wf(4), bias(4), 128, 1, out_bf16(a2, 128, 0));
    mark(5);
    conv_fwd(s, 128, in1(src(a2, 128, 0, 128)), B, H2, W2, 128, wf(5), bias(5), 128, 1, out_bf16(b2, 128, 0));
    mark(6);
    conv_fwd(s, 64, in2(src(b2, 128, 0, 128, kSrcUp2), src(cat1, 192, 128, 64)), B, H1, W1, 192, wf(6), bias(6), 64, 1,
             out_bf16(d1a, 64, 0));
    mark(7);
    conv_fwd(s, 64, in1(src(d1a, 64, 0, 64)), B, H1, W1, 64, wf(7), bias(7), 64, 1, out_bf16(d1b, 64, 0));
    mark(8);
    conv_fwd(s, 32, in2(src(d1b, 64, 0, 64, kSrcUp2), src(cat0, 96, 64, 32)), B, H, W, 96, wf(8), bias(8), 32, 1,
             out_bf16(d0a, 32, 0));
    mark(9);
    ConvOut o = out_f32(D.out.as<float>(), 16, 0, 0);
    if (!train) {
        o.mode = kOutImage;
        o.img = img;
        o.rad = rad;
    }
    conv_fwd(s, 16, in1(src(d0a, 32, 0, 32)), B, H, W, 32, wf(9), bias(9), 3, 0, o);
    mark(10);
    if (train) {
        hipLaunchKernelGGL(k_nn_input, dim3(grid1(P, D.cus)), dim3(256), 0, s, rad, aux, P, D.x0.as<uint16_t>(), 32);
        hipLaunchKernelGGL(k_pool2, dim3(grid1(P / 4 * 32, D.cus)), dim3(256), 0, s, cat0, B, H, W, 32, 96, 64,
                           D.p1.as<uint16_t>(), 32, 0);
        hipLaunchKernelGGL(k_pool2, dim3(grid1(P / 16 * 64, D.cus)), dim3(256), 0, s, cat1, B, H1, W1, 64, 192, 128,
                           D.p2.as<uint16_t>(), 64, 0);
        hipLaunchKernelGGL(k_up2, dim3(grid1(P / 4 * 128, D.cus)), dim3(256), 0, s, b2, B, H1, W1, 128, 128, 0, cat1, 192, 0);
        hipLaunchKernelGGL(k_up2, dim3(grid1(P * 64, D.cus)), dim3(256), 0, s, d1b, B, H, W, 64, 64, 0, cat0, 96, 0);
    }
    HIPCHECK(hipGetLastError());
}

// backward pass from D.d_out (fp32, 16 / px, channels 0..2) into D.grads
static void backward(lrt_denoiser& D) {
    hipStream_t s = D.stream;
    const int B = D.B, H = D.H, W = D.W, H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
    const long long P = (long long)B * H * W, P1 = P / 4, P2 = P / 16;
    float* gp = D.grads.as<float>();
    uint16_t* G = D.g.as<uint16_t>();
    auto bw = [&](int i) { return (const void*)D.L[i].bwd.p; };
    auto dw = [&](int i) { return gp + D.L[i].woff; };
    auto db = [&](int i) { return gp + D.L[i].boff; };
    auto relu_grad = [&](const float* dy, int dycs, int dyoff, const DevBuf& y, int ycs, int yoff, int relu, long long n,
                         int c, int gcs) {
        hipLaunchKernelGGL(k_relu_grad, dim3(grid1(n * gcs, D.cus)), dim3(256), 0, s, dy, dycs, dyoff, y.as<uint16_t>(), ycs,
                           yoff, relu, n, c, G, gcs);
    };
    // out: 32 -> 3 (linear)
    relu_grad(D.d_out.as<float>(), 16, 0, D.d0a, 32, 0, 0, P, 3, 32);
    wgrad(s, D.cus, 16, D.d0a.as<uint16_t>(), B, H, W, 32, 0, 32, G, 32, 0, 3, dw(9), db(9));
    conv(s, 32, G, B, H, W, 32, 32, 0, bw(9), nullptr, 32, 0, nullptr, 0, 0, D.d_d0a.as<float>(), 32, 0, 0);
    // d0a: 96 -> 32
    relu_grad(D.d_d0a.as<float>(), 32, 0, D.d0a, 32, 0, 1, P, 32, 32);
    wgrad(s, D.cus, 32, D.cat0.as<uint16_t>(), B, H, W, 96, 0, 96, G, 32, 0, 32, dw(8), db(8));
    conv(s, 96, G, B, H, W, 32, 32, 0, bw(8), nullptr, 96, 0, nullptr, 0, 0, D.d_cat0.as<float>(), 96, 0, 0);
    // up(d1b) -> cat0[0:64)
    hipLaunchKernelGGL(k_up2_grad, dim3(grid1(P1 * 64, D.cus)), dim3(256), 0, s, D.d_cat0.as<float>(), B, H, W, 64, 96, 0,
                       D.d_d1b.as<float>(), 64, 0);
    // d1b: 64 -> 64
    relu_grad(D.d_d1b.as<float>(), 64, 0, D.d1b, 64, 0, 1, P1, 64, 64);
    wgrad(s, D.cus, 64, D.d1a.as<uint16_t>(), B, H1, W1, 64, 0, 64, G, 64, 0, 64, dw(7), db(7));
    conv(s, 64, G, B, H1, W1, 64, 64, 0, bw(7), nullptr, 64, 0, nullptr, 0, 0, D.d_d1a.as<float>(), 64, 0, 0);
    // d1a: 192 -> 64
    relu_grad(D.d_d1a.as<float>(), 64, 0, D.d1a, 64, 0, 1, P1, 64, 64);
    wgrad(s, D.cus, 64, D.cat1.as<uint16_t>(), B, H1, W1, 192, 0, 192, G, 64, 0, 64, dw(6), db(6));
    conv(s, 192, G, B, H1, W1, 64, 64, 0, bw(6), nullptr, 192, 0, nullptr, 0, 0, D.d_cat1.as<float>(), 192, 0, 0);
    // up(b2) -> cat1[0:128)
    hipLaunchKernelGGL(k_up2_grad, dim3(grid1(P2 * 128, D.cus)), dim3(256), 0, s, D.d_cat1.as<float>(), B, H1, W1, 128, 192,
                       0, D.d_b2.as<float>(), 128, 0);
    // e2b: 128 -> 128
    relu_grad(D.d_b2.as<float>(), 128, 0, D.b2, 128, 0, 1, P2, 128, 128);
    wgrad(s, D.cus, 128, D.a2.as<uint16_t>(), B, H2, W2, 128, 0, 128, G, 128, 0, 128, dw(5), db(5));
    conv(s, 128, G, B, H2, W2, 128, 128, 0, bw(5), nullptr, 128, 0, nullptr, 0, 0, D.d_a2.as<float>(), 128, 0, 0);
    // e2a: 64 -> 128
    relu_grad(D.d_a2.as<float>(), 128, 0, D.a2, 128, 0, 1, P2, 128, 128);
    wgrad(s, D.cus, 128, D.p2.as<uint16_t>(), B, H2, W2, 64, 0, 64, G, 128, 0, 128, dw(4), db(4));
    conv(s, 64, G, B, H2, W2, 128, 128, 0, bw(4), nullptr, 64, 0, nullptr, 0, 0, D.d_p2.as<float>(), 64, 0, 0);
    // pool(skip 1): d cat1[128:192) += d p2 / 4
    hipLaunchKernelGGL(k_pool2_grad, dim3(grid1(P1 * 64, D.cus)), dim3(256), 0, s, D.d_p2.as<float>(), B, H1, W1, 64, 64, 0,
                       D.d_cat1.as<float>(), 192, 128);
    // e1b: 64 -> 64 (output in cat1[128:192))
    relu_grad(D.d_cat1.as<float>(), 192, 128, D.cat1, 192, 128, 1, P1, 64, 64);
    wgrad(s, D.cus, 64, D.a1.as<uint16_t>(), B, H1, W1, 64, 0, 64, G, 64, 0, 64, dw(3), db(3));
    conv(s, 64, G, B, H1, W1, 64, 64, 0, bw(3), nullptr, 64, 0, nullptr, 0, 0, D.d_a1.as<float>(), 64, 0, 0);
    // e1a: 32 -> 64
    relu_grad(D.d_a1.as<float>(), 64, 0, D.a1, 64, 0, 1, P1, 64, 64);
    wgrad(s, D.cus, 64, D.p1.as<uint16_t>(), B, H1, W1, 32, 0, 32, G, 64, 0, 64, dw(2), db(2));
    conv(s, 32, G, B, H1, W1, 64, 64, 0, bw(2), nullptr, 32, 0, nullptr, 0, 0, D.d_p1.as<float>(), 32, 0, 0);
    // pool(skip 0): d cat0[64:96) += d p1 / 4
    hipLaunchKernelGGL(k_pool2_grad, dim3(grid1(P * 32, D.cus)), dim3(256), 0, s, D.d_p1.as<float>(), B, H, W, 32, 32, 0,
                       D.d_cat0.as<float>(), 96, 64);
    // e0b: 32 -> 32 (output in cat0[64:96))
    relu_grad(D.d_cat0.as<float>(), 96, 64, D.cat0, 96, 64, 1, P, 32, 32);
    wgrad(s, D.cus, 32, D.a0.as<uint16_t>(), B, H, W, 32, 0, 32, G, 32, 0, 32, dw(1), db(1));
    conv(s, 32, G, B, H, W, 32, 32, 0, bw(1), nullptr, 32, 0, nullptr, 0, 0, D.d_a0.as<float>(), 32, 0, 0);
    // e0a: 10 -> 32
    relu_grad(D.d_a0.as<float>(), 32, 0, D.a0, 32, 0, 1, P, 32, 32);
    wgrad(s, D.cus, 32, D.x0.as<uint16_t>(), B, H, W, 32, 0, 10, G, 32, 0, 32, dw(0), db(0));
    HIPCHECK(hipGetLastError());
}

// inputs: host or device fp32 arrays -> device pointers
static void stage_inputs(lrt_denoiser& D, const float* rad, const float* aux, const float* tgt, size_t P, int device_ptrs,
                         const float*& drad, const float*& daux, const float*& dtgt) {
    if (device_ptrs) {
        after_caller(D.stream, D.caller);
        drad = rad;
        daux = aux;
        dtgt = tgt;
        return;
    }
    upload(D.in_rad, rad, P * 12, D.stream);
    upload(D.in_aux, aux, P * 28, D.stream);
    drad = D.in_rad.as<float>();
    daux = D.in_aux.as<float>();
    dtgt = nullptr;
    if (tgt) {
        upload(D.in_tgt, tgt, P * 12, D.stream);
        dtgt = D.in_tgt.as<float>();
    }
}

}  // namespace lrt

extern "C" {

int lrt_denoiser_create(int device, uint64_t seed, lrt_denoiser** out) {
    return guarded([&] {
        require(out != nullptr, "null pointer");
        HIPCHECK(hipSetDevice(device));
        auto D = std::make_unique<lrt_denoiser>();
        D->device = device;
        HIPCHECK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&D->caller, hipEventDisableTiming));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        D->cus = prop.multiProcessorCount;
        size_t off = 0;
        for (int li = 0; li < 10; ++li) {
            const int* k = kNet[li];
            NnLayer& l = D->L[li];
            l.cin = k[0];
            l.cout = k[1];
            l.relu = k[2];
            l.cinp = round_up(l.cin, 32);
            l.coutp = round_up(l.cout, 16);
            l.woff = off;
            off += (size_t)9 * l.cin * l.cout;
            l.boff = off;
            off += (size_t)l.cout;
        }
        D->n_params = off;
        for (NnLayer& l : D->L) {
            l.fwd.alloc((size_t)9 * (l.cinp / 32) * (l.coutp / 16) * 64 * 16);
            l.bwd.alloc((size_t)9 * (round_up(l.coutp, 32) / 32) * (l.cinp / 16) * 64 * 16);
        }
        // He-normal initial weights (the output layer scaled down), zero biases
        std::vector<float> p(off, 0.0f);
        uint64_t st = seed * 0x9e3779b97f4a7c15ull + 1;
        auto uni = [&]() {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        for (int li = 0; li < 10; ++li) {
            const NnLayer& l = D->L[li];
            const double sd = std::sqrt(2.0 / (9.0 * l.cin)) * (li == 9 ? 0.1 : 1.0);
            for (size_t i = 0; i < (size_t)9 * l.cin * l.cout; ++i) {
                const double u1 = uni(), u2 = uni();
                p[l.woff + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
        }
        upload(D->params, p.data(), off * 4, D->stream);
        D->grads.alloc(off * 4);
        D->m.alloc(off * 4);
        D->v.alloc(off * 4);
        hipLaunchKernelGGL(k_zero_f32, dim3(grid1((long long)off, D->cus)), dim3(256), 0, D->stream, D->m.as<float>(), (long long)off);
        hipLaunchKernelGGL(k_zero_f32, dim3(grid1((long long)off, D->cus)), dim3(256), 0, D->stream, D->v.as<float>(), (long long)off);
        D->loss.alloc(256);
        pack_all(*D);
        HIPCHECK(hipStreamSynchronize(D->stream));
        *out = D.release();
    });
}

int lrt_denoiser_release(lrt_denoiser* d) {
    return guarded([&] {
        if (d) {
            HIPCHECK(hipSetDevice(d->device));
            delete d;
        }
    });
}

int lrt_denoiser_param_count(const lrt_denoiser* d, uint64_t* n) {
    return guarded([&] {
        require(d && n, "null pointer");
        *n = d->n_params;
    });
}

int lrt_denoiser_params(lrt_denoiser* d, float* get, const float* set) {
    return guarded([&] {
        require(d != nullptr, "null pointer");
        HIPCHECK(hipSetDevice(d->device));
        if (get) {
            HIPCHECK(hipMemcpyAsync(get, d->params.p, d->n_params * 4, hipMemcpyDeviceToHost, d->stream));
            HIPCHECK(hipStreamSynchronize(d->stream));
        }
        if (set) {
            HIPCHECK(hipMemcpyAsync(d->params.p, set, d->n_params * 4, hipMemcpyHostToDevice, d->stream));
            pack_all(*d);
            d->adam_t = 0;
            const long long n = (long long)d->n_params;
            hipLaunchKernelGGL(k_zero_f32, dim3(grid1(n, d->cus)), dim3(256), 0, d->stream, d->m.as<float>(), n);
            hipLaunchKernelGGL(k_zero_f32, dim3(grid1(n, d->cus)), dim3(256), 0, d->stream, d->v.as<float>(), n);
            HIPCHECK(hipStreamSynchronize(d->stream));
        }
    });
}

int lrt_denoiser_forward(lrt_denoiser* d, const float* rad, const float* aux, uint32_t B, uint32_t H, uint32_t W,
                         float* out, int device_ptrs) {
    return guarded([&] {
        require(d && rad && aux && out, "null pointer");
        HIPCHECK(hipSetDevice(d->device));
        shape(*d, (int)B, (int)H, (int)W);
        const size_t P = (size_t)B * H * W;
        const float *drad, *daux, *dtgt;
        stage_inputs(*d, rad, aux, nullptr, P, device_ptrs, drad, daux, dtgt);
        float* dout = out;
        if (!device_ptrs) {
            d->res.alloc(P * 12);
            dout = d->res.as<float>();
        }
        forward(*d, drad, daux, false, dout);
        if (!device_ptrs) HIPCHECK(hipMemcpyAsync(out, dout, P * 12, hipMemcpyDeviceToHost, d->stream));
        HIPCHECK(hipStreamSynchronize(d->stream));
        if (d->timing) {
            for (int k = 0; k < 10; ++k) {
                float ms = 0.0f;
                HIPCHECK(hipEventElapsedTime(&ms, d->ev[k], d->ev[k + 1]));
                d->layer_ms[k] += ms;
            }
            d->timed_calls++;
        }
    });
}

int lrt_denoiser_loss_grad(lrt_denoiser* d, const float* rad, const float* aux, const float* target, uint32_t B,
                           uint32_t H, uint32_t W, float* loss, float* grads, int device_ptrs) {
    return guarded([&] {
        require(d && rad && aux && target, "null pointer");
        HIPCHECK(hipSetDevice(d->device));
        shape(*d, (int)B, (int)H, (int)W);
        shape_grad(*d);
        const size_t P = (size_t)B * H * W;
        const float *drad, *daux, *dtgt;
        stage_inputs(*d, rad, aux, target, P, device_ptrs, drad, daux, dtgt);
        forward(*d, drad, daux, true, nullptr);
        const long long n = (long long)d->n_params;
        hipLaunchKernelGGL(k_zero_f32, dim3(grid1(n, d->cus)), dim3(256), 0, d->stream, d->grads.as<float>(), n);
        hipLaunchKernelGGL(k_zero_f32, dim3(grid1((long long)P * 16, d->cus)), dim3(256), 0, d->stream, d->d_out.as<float>(),
                           (long long)P * 16);
        hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, d->stream, d->loss.as<float>(), 1ll);
        hipLaunchKernelGGL(k_nn_loss, dim3(grid1((long long)P, d->cus)), dim3(256), 0, d->stream, d->out.as<float>(), 16, drad,
                           dtgt, (long long)P, d->d_out.as<float>(), 16, d->loss.as<float>());
        HIPCHECK(hipGetLastError());
        backward(*d);
        float l = 0.0f;
        HIPCHECK(hipMemcpyAsync(&l, d->loss.p, 4, hipMemcpyDeviceToHost, d->stream));
        if (grads) HIPCHECK(hipMemcpyAsync(grads, d->grads.p, d->n_params * 4, hipMemcpyDeviceToHost, d->stream));
        HIPCHECK(hipStreamSynchronize(d->stream));
        if (loss) *loss = l;
    });
}

int lrt_denoiser_timing(lrt_denoiser* d, int enable, double* layer_ms, uint64_t* calls) {
    return guarded([&] {
        require(d != nullptr, "null pointer");
        HIPCHECK(hipSetDevice(d->device));
        if (layer_ms)
            for (int k = 0; k < 10; ++k) layer_ms[k] = d->layer_ms[k];
        if (calls) *calls = d->timed_calls;
        for (double& v : d->layer_ms) v = 0.0;
        d->timed_calls = 0;
        if (enable && !d->ev[0])
            for (hipEvent_t& e : d->ev) HIPCHECK(hipEventCreate(&e));
        d->timing = enable != 0;
    });
}

int lrt_denoiser_adam_step(lrt_denoiser* d, float lr) {
    return guarded([&] {
        require(d != nullptr, "null pointer");
        HIPCHECK(hipSetDevice(d->device));
        d->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - (float)std::pow(b1, d->adam_t), c2 = 1.0f - (float)std::pow(b2, d->adam_t);
        const long long n = (long long)d->n_params;
        hipLaunchKernelGGL(k_adam, dim3(grid1(n, d->cus)), dim3(256), 0, d->stream, d->params.as<float>(),
                           d->grads.as<float>(), d->m.as<float>(), d->v.as<float>(), n, lr, b1, b2, 1e-8f, c1, c2);
        pack_all(*d);
        HIPCHECK(hipStreamSynchronize(d->stream));
    });
}

// one layer on host fp32 arrays (inputs rounded to bf16 on the device)
int lrt_conv3x3(int device, int op, const float* x, const float* w, const float* b, const float* g, uint32_t B,
                uint32_t H, uint32_t W, uint32_t cin, uint32_t cout, int relu, float* out) {
    return guarded([&] {
        require(out && w, "null pointer");
        require(op >= 0 && op <= 2, "op must be 0 (forward), 1 (data gradient) or 2 (weight gradient)");
        require(cin >= 1 && cin <= 192 && cout >= 1 && cout <= 128, "channel counts out of range");
        HIPCHECK(hipSetDevice(device));
        hipStream_t s = nullptr;
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        const int cus = prop.multiProcessorCount;
        const int cinp = round_up((int)cin, 32), coutp = round_up((int)cout, 16), gcs = round_up(coutp, 32);
        const size_t P = (size_t)B * H * W;
        DevBuf dw, dfrag, dx, dg, dout, db;
        upload(dw, w, (size_t)9 * cin * cout * 4, s);
        if (b) upload(db, b, (size_t)cout * 4, s);
        auto to_bf = [&](const float* src, int c, int cs, DevBuf& dst) {  // host fp32 [P][c] -> device bf16 [P][cs]
            std::vector<float> pad(P * cs, 0.0f);
            for (size_t q = 0; q < P; ++q)
                for (int k = 0; k < c; ++k) pad[q * cs + k] = src[q * c + k];
            DevBuf tmp;
            upload(tmp, pad.data(), pad.size() * 4, s);
            dst.alloc(pad.size() * 2);
            hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid1((long long)pad.size(), cus)), dim3(256), 0, s, tmp.as<float>(),
                               (long long)pad.size(), dst.as<uint16_t>());
            HIPCHECK(hipStreamSynchronize(s));
        };
        if (op == 0) {
            require(x != nullptr, "null pointer");
            to_bf(x, (int)cin, cinp, dx);
            dfrag.alloc((size_t)9 * (cinp / 32) * (coutp / 16) * 64 * 16);
            hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, s, dw.as<float>(), (int)cin, (int)cout, 0, cinp, coutp,
                               dfrag.as<bf16x8>());
            dout.alloc(P * cout * 4);
            conv_fwd(s, coutp, in1(src(dx.as<uint16_t>(), cinp, 0, cinp)), (int)B, (int)H, (int)W, cinp, dfrag.p,
                     b ? db.as<float>() : nullptr, (int)cout, relu, out_f32(dout.as<float>(), (int)cout, 0, 0));
            HIPCHECK(hipMemcpy(out, dout.p, P * cout * 4, hipMemcpyDeviceToHost));
        } else if (op == 1) {
            require(g != nullptr, "null pointer");
            to_bf(g, (int)cout, gcs, dg);
            dfrag.alloc((size_t)9 * (gcs / 32) * (cinp / 16) * 64 * 16);
            hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, s, dw.as<float>(), (int)cin, (int)cout, 1, gcs, cinp,
                               dfrag.as<bf16x8>());
            dout.alloc(P * cin * 4);
            conv(s, cinp, dg.as<uint16_t>(), (int)B, (int)H, (int)W, gcs, gcs, 0, dfrag.p, nullptr, (int)cin, 0, nullptr, 0, 0,
                 dout.as<float>(), (int)cin, 0, 0);
            HIPCHECK(hipMemcpy(out, dout.p, P * cin * 4, hipMemcpyDeviceToHost));
        } else {
            require(x && g, "null pointer");
            to_bf(x, (int)cin, cinp, dx);
            to_bf(g, (int)cout, gcs, dg);
            const size_t nw = (size_t)9 * cin * cout;
            dout.alloc((nw + cout) * 4);
            hipLaunchKernelGGL(k_zero_f32, dim3(grid1((long long)(nw + cout), cus)), dim3(256), 0, s, dout.as<float>(),
                               (long long)(nw + cout));
            wgrad(s, cus, coutp, dx.as<uint16_t>(), (int)B, (int)H, (int)W, cinp, 0, (int)cin, dg.as<uint16_t>(), gcs, 0,
                  (int)cout, dout.as<float>(), dout.as<float>() + nw);
            HIPCHECK(hipMemcpy(out, dout.p, (nw + cout) * 4, hipMemcpyDeviceToHost));
        }
        HIPCHECK(hipDeviceSynchronize());
    });
}

}  // extern "C"
