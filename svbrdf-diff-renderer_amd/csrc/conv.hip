// conv.hip -- fp32 3x3 convolutions on the gfx950 matrix cores for the
// networks of the MaterialGAN caller (src/materialgan.py:21-173): the
// modulated / demodulated convolutions of the StyleGAN2 synthesis network
// (src/higan_models/stylegan2_generator_network.py:585-760, ModulateConvBlock)
// and the 3x3 convolutions of the VGG-19 feature loss (src/descriptor.py:7-79).
// Both compute in fp32, as the reference's torch convolutions do.
//
// Kernels
//   conv3x3_pack_kernel   [Cout, Cin, 3, 3] weights -> the GEMM layout
//                         [9 taps][Cin][Cout]; optionally scaled per input
//                         channel by the style (modulation), demodulated per
//                         output channel (1 / sqrt(sum w^2 + 1e-8)), or
//                         flipped and transposed ([9][Cout][Cin] with tap
//                         8 - t) for the input gradient.
//   conv3x3_mfma_kernel   y = act(conv3x3(x, w) + noise_strength noise + bias),
//                         stride 1, zero padding 1, as an implicit GEMM:
//                         M = output channels, N = a 2-D tile of output
//                         pixels, K = input channels x 9 taps, accumulated
//                         by v_mfma_f32_32x32x2f32.
//   upsample2_blur_kernel x2 upsampling by zero insertion and the [1, 3, 3, 1]
//                         binomial filter (gain 4): out(2p) = (x(p-1) + 3 x(p))/4,
//                         out(2p+1) = (3 x(p) + x(p+1))/4 per axis, zero outside
//                         the image; and its adjoint.
//
// The convolution stages, per chunk of 8 input channels, the chunk's
// weights (8 x 9 x BM floats) and the input tile with a one-pixel halo in
// LDS; every wave then runs 9 taps x 4 K-steps of (BM/64) x (BN/64) MFMA
// tiles of 32 x 32 from LDS.  The input gradient of a convolution is the same
// convolution of dL/dy with the flipped, transposed weights.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "../../include/svbrdf_hip.h"

using svbrdf_internal::check_launch;
using svbrdf_internal::fail;

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;   // four waves
constexpr int kKC = 8;          // input channels per staged K chunk
constexpr int kMaxChannels = 4096;
constexpr int kMaxSide = 8192;

// ---------------------------------------------------------------------------
// Weight packing.  One block per output channel: the demodulation factor is a
// block reduction over its Cin x 9 weights.
__global__ __launch_bounds__(kThreads) void conv3x3_pack_kernel(const float *__restrict__ w, int cout, int cin,
                                                                const float *__restrict__ style, float scale,
                                                                int demodulate, int transpose,
                                                                float *__restrict__ packed) {
    __shared__ float red[kThreads / 64];
    __shared__ float s_factor;
    const int co = blockIdx.x;
    const int n = cin * 9;
    const float *wc = w + (size_t)co * n;
    float factor = 1.0f;
    if (demodulate) {
        float acc = 0.0f;
        for (int e = threadIdx.x; e < n; e += kThreads) {
            const float v = wc[e] * scale * (style ? style[e / 9] : 1.0f);
            acc += v * v;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) s_factor = rsqrtf(((red[0] + red[1]) + red[2]) + red[3] + 1e-8f);
        __syncthreads();
        factor = s_factor;
    }
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const int ci = e / 9, t = e - ci * 9;
        const float v = wc[e] * scale * (style ? style[ci] : 1.0f) * factor;
        if (transpose)
            packed[((size_t)(8 - t) * cout + co) * cin + ci] = v;   // [9][Cout][Cin], taps flipped
        else
            packed[((size_t)t * cin + ci) * cout + co] = v;         // [9][Cin][Cout]
    }
}

// ---------------------------------------------------------------------------
// The convolution.
struct ConvArgs {
    const float *x;       // [N, Cin, H, W]
    const float *w;       // [9][Cin][Cout]
    const float *bias;    // [Cout] or null
    const float *noise;   // [H, W] or null
    float noise_strength;
    int act;              // SVBRDF_ACT_*
    float *y;             // [N, Cout, H, W]
    int cin, cout, h, w_;
    int tw, th, tiles_x;  // pixel tile: th rows x tw columns, tw * th = BN
};

// LDS floats per staged channel of the input tile: (th + 2) (tw + 2) =
// BN + 2 (th + tw) + 4 with 4 <= tw <= 32 (so th + tw <= 32 + BN / 4).
template <int BN>
constexpr int halo_floats() { return BN + 2 * (32 + BN / 4) + 4; }

__device__ __forceinline__ float activate(float v, int act) {
    if (act == SVBRDF_ACT_RELU) return v > 0.0f ? v : 0.0f;
    if (act == SVBRDF_ACT_LRELU) return (v > 0.0f ? v : v * 0.2f) * 1.41421356237309515f;
    return v;
}

template <int BM, int BN>
__global__ __launch_bounds__(kThreads) void conv3x3_mfma_kernel(ConvArgs a) {
    constexpr int TM = BM / 64, TN = BN / 64;   // 32 x 32 MFMA tiles per wave along M and N
    constexpr int HF = halo_floats<BN>();
    __shared__ float As[kKC][9][BM];
    __shared__ float Bs[kKC][HF];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;    // the wave's quarter of the block tile
    const int img = blockIdx.z;
    const int co0 = blockIdx.y * BM;
    const int tyi = blockIdx.x / a.tiles_x, txi = blockIdx.x - tyi * a.tiles_x;
    const int y0 = tyi * a.th, x0 = txi * a.tw;
    const int tw2 = a.tw + 2, hs = (a.th + 2) * tw2;
    const int H = a.h, W = a.w_;
    const float *xi = a.x + (size_t)img * a.cin * H * W;

    // the halo-tile offset of the pixel each lane feeds to MFMA tile j
    int pix[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int nn = wn * (BN / 2) + j * 32 + (lane & 31);
        const int ty = nn / a.tw, tx = nn - ty * a.tw;
        pix[j] = ty * tw2 + tx;
    }
    v16f acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int kc_lane = lane >> 5;   // which of the two K rows of a K-step this lane carries
    for (int ci0 = 0; ci0 < a.cin; ci0 += kKC) {
        // stage the chunk's weights: As[kc][tap][m] = w[tap][ci0 + kc][co0 + m]
        for (int e = threadIdx.x; e < kKC * 9 * BM; e += kThreads) {
            const int kc = e / (9 * BM), rem = e - kc * (9 * BM);
            const int t = rem / BM, m = rem - t * BM;
            const int ci = ci0 + kc, co = co0 + m;
            As[kc][t][m] = (ci < a.cin && co < a.cout) ? a.w[((size_t)t * a.cin + ci) * a.cout + co] : 0.0f;
        }
        // stage the input tile with its halo (zero outside the image)
        for (int e = threadIdx.x; e < kKC * hs; e += kThreads) {
            const int kc = e / hs, rem = e - kc * hs;
            const int r = rem / tw2, c = rem - r * tw2;
            const int yy = y0 - 1 + r, xx = x0 - 1 + c, ci = ci0 + kc;
            Bs[kc][rem] = (ci < a.cin && yy >= 0 && yy < H && xx >= 0 && xx < W)
                              ? xi[((size_t)ci * H + yy) * W + xx] : 0.0f;
        }
        __syncthreads();
        // the chunk's 72 products per output are summed in a fresh accumulator
        // and then added to the running sum: the rounding error grows with
        // 72 + Cin / 8 terms instead of 9 Cin
        v16f part[TM][TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) part[i][j][r] = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int boff = (t / 3) * tw2 + (t % 3);
#pragma unroll
            for (int ks = 0; ks < kKC / 2; ++ks) {
                const int kc = 2 * ks + kc_lane;
                float av[TM], bv[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) av[i] = As[kc][t][wm * (BM / 2) + i * 32 + (lane & 31)];
#pragma unroll
                for (int j = 0; j < TN; ++j) bv[j] = Bs[kc][pix[j] + boff];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        part[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i], bv[j], part[i][j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] += part[i][j];
        __syncthreads();
    }

    // epilogue: element r of tile (i, j) is row m = 8 (r / 4) + 4 (lane / 32) + r % 4
    // and column n = lane % 32 of that 32 x 32 tile
    float *yi = a.y + (size_t)img * a.cout * H * W;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int nn = wn * (BN / 2) + j * 32 + (lane & 31);
        const int ty = nn / a.tw, tx = nn - ty * a.tw;
        const int yy = y0 + ty, xx = x0 + tx;
        if (yy >= H || xx >= W) continue;
        const size_t p = (size_t)yy * W + xx;
        const float nz = a.noise ? a.noise_strength * a.noise[p] : 0.0f;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int co = co0 + wm * (BM / 2) + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (co >= a.cout) continue;
                float v = acc[i][j][r];
                if (a.noise) v = v + nz;
                if (a.bias) v = v + a.bias[co];
                yi[(size_t)co * H * W + p] = activate(v, a.act);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// x2 upsampling with the [1, 3, 3, 1] filter and its adjoint.  Zero insertion
// (x(p) at 2p), padding 2 before and 1 after, and the filter [1, 3, 3, 1]/4
// per axis give, per axis,
//   out(2p) = x(p - 1)/4 + 3 x(p)/4,   out(2p + 1) = 3 x(p)/4 + x(p + 1)/4,
// with x = 0 outside the image; rows are filtered after columns.
__device__ __forceinline__ void up_taps(int o, int &i0, float &w0, int &i1, float &w1) {
    const int p = o >> 1;
    if (o & 1) {
        i0 = p; w0 = 0.75f; i1 = p + 1; w1 = 0.25f;
    } else {
        i0 = p - 1; w0 = 0.25f; i1 = p; w1 = 0.75f;
    }
}

__global__ __launch_bounds__(kThreads) void upsample2_blur_kernel(const float *__restrict__ x, int planes, int h,
                                                                  int w, float *__restrict__ y) {
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    const int H2 = 2 * h, W2 = 2 * w;
    if (e >= (size_t)planes * H2 * W2) return;
    const size_t pl = e / ((size_t)H2 * W2);
    const int rem = (int)(e - pl * (size_t)H2 * W2);
    const int oy = rem / W2, ox = rem - oy * W2;
    const float *s = x + pl * (size_t)h * w;
    int r0, r1, c0, c1;
    float wr0, wr1, wc0, wc1;
    up_taps(oy, r0, wr0, r1, wr1);
    up_taps(ox, c0, wc0, c1, wc1);
    auto at = [&](int yy, int xx) { return (yy >= 0 && yy < h && xx >= 0 && xx < w) ? s[(size_t)yy * w + xx] : 0.0f; };
    const float h0 = wc0 * at(r0, c0) + wc1 * at(r0, c1);
    const float h1 = wc0 * at(r1, c0) + wc1 * at(r1, c1);
    y[e] = wr0 * h0 + wr1 * h1;
}

// The adjoint: per axis dL/dx(p) = g(2p - 1)/4 + 3 g(2p)/4 + 3 g(2p + 1)/4 + g(2p + 2)/4,
// terms outside the output dropped.
__global__ __launch_bounds__(kThreads) void upsample2_blur_adjoint_kernel(const float *__restrict__ gy, int planes,
                                                                          int h, int w, float *__restrict__ gx) {
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= (size_t)planes * h * w) return;
    const size_t pl = e / ((size_t)h * w);
    const int rem = (int)(e - pl * (size_t)h * w);
    const int p = rem / w, q = rem - p * w;
    const int H2 = 2 * h, W2 = 2 * w;
    const float *g = gy + pl * (size_t)H2 * W2;
    const float wts[4] = {0.25f, 0.75f, 0.75f, 0.25f};
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int oy = 2 * p - 1 + a;
        if (oy < 0 || oy >= H2) continue;
        float racc = 0.0f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int ox = 2 * q - 1 + b;
            if (ox < 0 || ox >= W2) continue;
            racc += wts[b] * g[(size_t)oy * W2 + ox];
        }
        acc += wts[a] * racc;
    }
    gx[e] = acc;
}

int pick_tile_width(int w) {
    int tw = 4;
    while (tw < w && tw < 32) tw *= 2;
    return tw;
}

template <int BM, int BN>
int launch_conv(ConvArgs a, int n, hipStream_t s) {
    a.tw = pick_tile_width(a.w_);
    a.th = BN / a.tw;
    a.tiles_x = (a.w_ + a.tw - 1) / a.tw;
    const int tiles_y = (a.h + a.th - 1) / a.th;
    const dim3 grid((unsigned)(a.tiles_x * tiles_y), (unsigned)((a.cout + BM - 1) / BM), (unsigned)n);
    hipLaunchKernelGGL((conv3x3_mfma_kernel<BM, BN>), grid, dim3(kThreads), 0, s, a);
    return check_launch("svbrdf_conv3x3");
}

}  // namespace

extern "C" {

int svbrdf_conv3x3_pack(const float *weight, int cout, int cin, const float *style, float scale, int demodulate,
                        int transpose, float *packed, void *stream) {
    const char *what = "svbrdf_conv3x3_pack";
    if (!weight || !packed) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (cout < 1 || cin < 1 || cout > kMaxChannels || cin > kMaxChannels)
        return fail(SVBRDF_ESHAPE, "%s: cout=%d cin=%d outside [1, %d]", what, cout, cin, kMaxChannels);
    hipLaunchKernelGGL(conv3x3_pack_kernel, dim3((unsigned)cout), dim3(kThreads), 0, (hipStream_t)stream, weight,
                       cout, cin, style, scale, demodulate ? 1 : 0, transpose ? 1 : 0, packed);
    return check_launch(what);
}

int svbrdf_conv3x3(const float *x, int n, int cin, int height, int width, const float *packed, int cout,
                   const float *bias, const float *noise, float noise_strength, int activation, float *y,
                   void *stream) {
    const char *what = "svbrdf_conv3x3";
    if (!x || !packed || !y) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (n < 1 || n > 65535 || cin < 1 || cout < 1 || cin > kMaxChannels || cout > kMaxChannels || height < 1 ||
        width < 1 || height > kMaxSide || width > kMaxSide)
        return fail(SVBRDF_ESHAPE, "%s: n=%d cin=%d cout=%d height=%d width=%d", what, n, cin, cout, height, width);
    if (activation < SVBRDF_ACT_NONE || activation > SVBRDF_ACT_LRELU)
        return fail(SVBRDF_EINVAL, "%s: unknown activation %d", what, activation);
    ConvArgs a{};
    a.x = x; a.w = packed; a.bias = bias; a.noise = noise; a.noise_strength = noise_strength; a.act = activation;
    a.y = y; a.cin = cin; a.cout = cout; a.h = height; a.w_ = width;
    hipStream_t s = (hipStream_t)stream;
    // 128 output channels per block where there are that many; 64-pixel tiles
    // when 128-pixel tiles would leave the device's compute units short of
    // blocks
    const long long pixels = (long long)height * width * n;
    if (cout >= 128) {
        if (pixels * ((cout + 127) / 128) >= 128LL * 512) return launch_conv<128, 128>(a, n, s);
        return launch_conv<128, 64>(a, n, s);
    }
    if (pixels * ((cout + 63) / 64) >= 128LL * 512) return launch_conv<64, 128>(a, n, s);
    return launch_conv<64, 64>(a, n, s);
}

int svbrdf_upsample2_blur(const float *x, int planes, int height, int width, float *y, void *stream) {
    const char *what = "svbrdf_upsample2_blur";
    if (!x || !y) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (planes < 1 || height < 1 || width < 1 || height > kMaxSide / 2 || width > kMaxSide / 2 ||
        (size_t)planes * 4 * height * width > ((size_t)1 << 34))
        return fail(SVBRDF_ESHAPE, "%s: planes=%d height=%d width=%d", what, planes, height, width);
    const size_t total = (size_t)planes * 4 * height * width;
    hipLaunchKernelGGL(upsample2_blur_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       (hipStream_t)stream, x, planes, height, width, y);
    return check_launch(what);
}

int svbrdf_upsample2_blur_adjoint(const float *grad_y, int planes, int height, int width, float *grad_x,
                                  void *stream) {
    const char *what = "svbrdf_upsample2_blur_adjoint";
    if (!grad_y || !grad_x) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (planes < 1 || height < 1 || width < 1 || height > kMaxSide / 2 || width > kMaxSide / 2 ||
        (size_t)planes * 4 * height * width > ((size_t)1 << 34))
        return fail(SVBRDF_ESHAPE, "%s: planes=%d height=%d width=%d", what, planes, height, width);
    const size_t total = (size_t)planes * height * width;
    hipLaunchKernelGGL(upsample2_blur_adjoint_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, (hipStream_t)stream, grad_y, planes, height, width, grad_x);
    return check_launch(what);
}

}  // extern "C"
