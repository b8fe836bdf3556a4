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
//                         by v_mfma_f32_32x32x2f32.  Where the tiles alone
//                         leave the device short of blocks (small images,
//                         many channels) K is split into slices whose
//                         partial sums go to the caller's workspace;
//   conv_split_sum_kernel then adds the slices in order and applies the
//                         epilogue.
//   conv3x3_wgrad_kernel  dL/dw of the convolution: per tap a GEMM over the
//                         pixels (M = Cout, N = Cin), the nine taps sharing
//                         the dL/dy operand; wgrad_sum_kernel adds its
//                         pixel slices.
//   conv3x3_direct_kernel the same convolution for at most 8 output channels
//                         (the input gradient of a layer that reads the RGB
//                         image): input tiles staged in LDS, two pixels per
//                         thread, all output channels in registers, on the
//                         vector ALUs -- an MFMA tile of 64 output channels
//                         would be mostly padding.
//   upsample2_blur_kernel x2 upsampling by zero insertion and the [1, 3, 3, 1]
//                         binomial filter (gain 4): out(2p) = (x(p-1) + 3 x(p))/4,
//                         out(2p+1) = (3 x(p) + x(p+1))/4 per axis, zero outside
//                         the image; and its adjoint.
//
// The convolution stages, per chunk of KC (4) input channels, the chunk's
// weights (KC x 9 x BM floats) and the input tile with a one-pixel halo in
// one of two LDS buffers (the next chunk's global loads in flight meanwhile);
// every wave then runs 9 taps x KC/2 K-steps of (BM/64) x (BN/64) MFMA tiles
// of 32 x 32 from LDS.  At KC = 4 a block holds 45 KB of LDS, so two blocks
// (the register limit of the 128 x 128 tile) share a compute unit.  The
// input gradient of a convolution is the same convolution of dL/dy with the
// flipped, transposed weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "../../include/svbrdf_hip.h"

using svbrdf_internal::check_launch;
using svbrdf_internal::fail;

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;   // four waves
constexpr int kMaxChannels = 4096;
constexpr int kMaxSide = 8192;

// ---------------------------------------------------------------------------
// Weight packing.  The demodulation factor of an output channel is a block
// reduction over its Cin x 9 scaled weights, in the same per-thread order in
// every form below, so the factors and the packed values are the same bits.
// The packed layouts hold consecutive output channels ([9][Cin][Cout]) or,
// for the input gradient, consecutive input channels ([9][Cout][Cin]) side by
// side; where that count is a multiple of 4 a thread writes four of them as
// one 16-byte store (a block per 4 output channels, or per output channel
// with 4 input channels per step) instead of 4-byte stores a row apart.
__device__ float pack_factor(const float *__restrict__ wc, int n, const float *__restrict__ style, float scale,
                             float *red) {
    float acc = 0.0f;
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const float v = wc[e] * scale * (style ? style[e / 9] : 1.0f);
        acc += v * v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    const float f = rsqrtf(((red[0] + red[1]) + red[2]) + red[3] + 1e-8f);
    __syncthreads();
    return f;
}

// MODE 0: one block per output channel, any shape, 4-byte stores (either
//         layout, by `transpose`).
// MODE 1: [9][Cin][Cout], one block per 4 output channels (Cout % 4 == 0).
// MODE 2: [9][Cout][Cin] with the taps flipped, one block per output channel,
//         4 input channels per step (Cin % 4 == 0).
template <int MODE>
__global__ __launch_bounds__(kThreads) void conv3x3_pack_kernel(const float *__restrict__ w, int cout, int cin,
                                                                const float *__restrict__ style, float scale,
                                                                int demodulate, int transpose,
                                                                float *__restrict__ packed) {
    __shared__ float red[kThreads / 64];
    const int n = cin * 9;
    if constexpr (MODE == 1) {
        const int co0 = blockIdx.x * 4;
        float factor[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        if (demodulate) {
            // the four channels' sums in one pass, each in pack_factor's order
            __shared__ float red4[4][kThreads / 64];
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
            for (int e = threadIdx.x; e < n; e += kThreads) {
                const float sv = style ? style[e / 9] : 1.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float v = w[(size_t)(co0 + j) * n + e] * scale * sv;
                    acc[j] += v * v;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
                if ((threadIdx.x & 63) == 0) red4[j][threadIdx.x >> 6] = acc[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 4; ++j)
                factor[j] = rsqrtf(((red4[j][0] + red4[j][1]) + red4[j][2]) + red4[j][3] + 1e-8f);
        }
#pragma unroll 2
        for (int e = threadIdx.x; e < n; e += kThreads) {
            const int ci = e / 9, t = e - ci * 9;
            const float sv = style ? style[ci] : 1.0f;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = w[(size_t)(co0 + j) * n + e] * scale * sv * factor[j];
            *reinterpret_cast<float4 *>(packed + ((size_t)t * cin + ci) * cout + co0) = make_float4(v[0], v[1], v[2], v[3]);
        }
    } else {
        const int co = blockIdx.x;
        const float *wc = w + (size_t)co * n;
        const float factor = demodulate ? pack_factor(wc, n, style, scale, red) : 1.0f;
        if constexpr (MODE == 2) {
            for (int q = threadIdx.x; q < n / 4; q += kThreads) {
                const int ci0 = (q / 9) * 4, t = q - (q / 9) * 9;
                float v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v[j] = wc[(ci0 + j) * 9 + t] * scale * (style ? style[ci0 + j] : 1.0f) * factor;
                *reinterpret_cast<float4 *>(packed + ((size_t)(8 - t) * cout + co) * cin + ci0) =
                    make_float4(v[0], v[1], v[2], v[3]);
            }
        } else {
            for (int e = threadIdx.x; e < n; e += kThreads) {
                const int ci = e / 9, t = e - ci * 9;
                const float v = wc[e] * scale * (style ? style[ci] : 1.0f) * factor;
                if (transpose)
                    packed[((size_t)(8 - t) * cout + co) * cin + ci] = v;   // [9][Cout][Cin], taps flipped
                else
                    packed[((size_t)t * cin + ci) * cout + co] = v;         // [9][Cin][Cout]
            }
        }
    }
}

// dL/d(pre-activation) = dL/dy x (y > 0 ? pos : neg): the derivative of a
// sign-keeping activation read off its output (ReLU: 1 / 0; leaky ReLU x
// sqrt(2): sqrt(2) / 0.2 sqrt(2)), four elements per thread.
__global__ __launch_bounds__(kThreads) void act_grad_kernel(const float4 *__restrict__ g, const float4 *__restrict__ y,
                                                            size_t n4, float pos, float neg, float4 *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n4) return;
    const float4 gv = g[i], yv = y[i];
    out[i] = make_float4(gv.x * (yv.x > 0.0f ? pos : neg), gv.y * (yv.y > 0.0f ? pos : neg),
                         gv.z * (yv.z > 0.0f ? pos : neg), gv.w * (yv.w > 0.0f ? pos : neg));
}

__global__ __launch_bounds__(kThreads) void act_grad_tail_kernel(const float *__restrict__ g, const float *__restrict__ y,
                                                                 size_t n, float pos, float neg,
                                                                 float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = g[i] * (y[i] > 0.0f ? pos : neg);
}

// Gradient of the modulated, demodulated weights with respect to the style:
// with u = scale style[ci] w[co][ci][t], d_co = 1/sqrt(sum_{ci,t} u^2 + 1e-8)
// and the packed weight u d_co, given g = dL/d(u d_co) in [Cout, Cin, 3, 3],
//   dL/du = g d_co - u d_co^3 sum_{ci,t} g u        (e_co = d_co^3 sum g u)
//   dL/dstyle[ci] = scale sum_{co,t} w (g d_co - e_co u)
// (d = 1, e = 0 without demodulation).  Two passes with fixed summation
// orders: one block per output channel for (d, e), one per input channel for
// the style gradient.
template <int BLOCK>
__device__ __forceinline__ float block_sum_f(float v, float *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0)
        for (int k = 0; k < BLOCK / 64; ++k) r += red[k];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kThreads) void modulation_rows_kernel(const float *__restrict__ w, int cin,
                                                                   const float *__restrict__ style, float scale,
                                                                   int demodulate, const float *__restrict__ g,
                                                                   float2 *__restrict__ de) {
    __shared__ float red[kThreads / 64];
    const int co = blockIdx.x, n = cin * 9;
    const float *wc = w + (size_t)co * n, *gc = g + (size_t)co * n;
    float ss = 0.0f, gu = 0.0f;
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const float u = wc[e] * scale * style[e / 9];
        ss += u * u;
        gu += gc[e] * u;
    }
    ss = block_sum_f<kThreads>(ss, red);
    gu = block_sum_f<kThreads>(gu, red);
    if (threadIdx.x == 0) {
        const float d = demodulate ? rsqrtf(ss + 1e-8f) : 1.0f;
        de[co] = make_float2(d, demodulate ? d * d * d * gu : 0.0f);
    }
}

__global__ __launch_bounds__(kThreads) void modulation_cols_kernel(const float *__restrict__ w, int cout, int cin,
                                                                   const float *__restrict__ style, float scale,
                                                                   const float *__restrict__ g,
                                                                   const float2 *__restrict__ de,
                                                                   float *__restrict__ grad_style) {
    __shared__ float red[kThreads / 64];
    const int ci = blockIdx.x;
    const float sc = scale * style[ci];
    float acc = 0.0f;
    for (int co = threadIdx.x; co < cout; co += kThreads) {
        const size_t base = ((size_t)co * cin + ci) * 9;
        const float2 f = de[co];
        float r = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const float wv = w[base + t];
            r += wv * (g[base + t] * f.x - f.y * sc * wv);
        }
        acc += r;
    }
    acc = block_sum_f<kThreads>(acc, red);
    if (threadIdx.x == 0) grad_style[ci] = scale * acc;
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
    // input gate: x is read as x * (gate > 0 ? gate_pos : gate_neg) where gate
    // is not null ([N, Cin, H, W], the forward output of the layer whose input
    // gradient this convolution computes: the activation's derivative)
    const float *gate;
    float gate_pos, gate_neg;
    float *y;             // [N, Cout, H, W]
    float *part;          // [slices][N, Cout, H, W] partial sums when K is split, else null
    int n, cin, cout, h, w_;
    int tw, th, tiles_x;  // pixel tile: th rows x tw columns, tw * th = BN
    int slices, cps;      // K slices (blockIdx.z = image x slices + slice) of cps chunks each
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

// the epilogue of output channel co at pixel p: noise, bias, activation
__device__ __forceinline__ float finish(const ConvArgs &a, float v, int co, size_t p) {
    if (a.noise) v = v + a.noise_strength * a.noise[p];
    if (a.bias) v = v + a.bias[co];
    return activate(v, a.act);
}

// Block tile BM output channels x BN pixels (th rows x tw columns), four
// waves of (BM/2) x (BN/2).  K runs over chunks of KC input channels; each
// chunk stages its weights (KC x 9 x BM floats) and its input tile with a
// one-pixel halo (KC x (th+2)(tw+2) floats) in one of two LDS buffers.  The
// global loads of chunk c + 1 are issued into registers before the MFMAs of
// chunk c and written to the other buffer after them, so their latency is
// hidden by the chunk's 9 taps x KC/2 K-steps of MFMA work and one barrier
// per chunk separates the buffers' uses.  VEC_A loads the weights as float4
// (Cout a multiple of 4).
template <int BM, int BN, int KC, bool VEC_A, bool GATE>
__global__ __launch_bounds__(kThreads) void conv3x3_mfma_kernel(ConvArgs a) {
    constexpr int TM = BM / 64, TN = BN / 64;   // 32 x 32 MFMA tiles per wave along M and N
    constexpr int HF = halo_floats<BN>();
    constexpr int AV = VEC_A ? 4 : 1;                       // floats per weight load
    constexpr int A_LOADS = KC * 9 * BM / AV;               // weight loads per chunk
    constexpr int A_PER = (A_LOADS + kThreads - 1) / kThreads;
    constexpr int B_PER = (KC * HF + kThreads - 1) / kThreads;
    __shared__ float As[2][KC * 9 * BM];
    __shared__ float Bs[2][KC * HF];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;    // the wave's quarter of the block tile
    const int img = blockIdx.z / a.slices, slice = blockIdx.z - img * a.slices;
    const int co0 = blockIdx.y * BM;
    const int tyi = blockIdx.x / a.tiles_x, txi = blockIdx.x - tyi * a.tiles_x;
    const int y0 = tyi * a.th, x0 = txi * a.tw;
    const int tw2 = a.tw + 2, hs = (a.th + 2) * tw2;
    const int H = a.h, W = a.w_, HW = H * W;
    const float *xi = a.x + (size_t)img * a.cin * HW;

    // The staging loads of this thread, fixed for every chunk: offsets from
    // the chunk's first weight row / input plane (-1: nothing there) and the
    // chunk channel they belong to.  Loads without data read a valid address
    // (the chunk's first element) and select 0, so the staging has no branches.
    int aoff[A_PER], akc[A_PER];
#pragma unroll
    for (int k = 0; k < A_PER; ++k) {
        const int q = tid + k * kThreads;
        const int e = q * AV;                           // float index within the chunk: (kc, t, m)
        const int kc = e / (9 * BM), rem = e - kc * (9 * BM);
        const int t = rem / BM, m = rem - t * BM;
        const bool live = q < A_LOADS && co0 + m < a.cout;
        aoff[k] = live ? (t * a.cin + kc) * a.cout + co0 + m : -1;
        akc[k] = kc;
    }
    int boff[B_PER], bkc[B_PER];
#pragma unroll
    for (int k = 0; k < B_PER; ++k) {
        const int q = tid + k * kThreads;
        const int kc = q / hs, rem = q - kc * hs;
        const int r = rem / tw2, c = rem - r * tw2;
        const int yy = y0 - 1 + r, xx = x0 - 1 + c;
        const bool live = q < KC * hs && yy >= 0 && yy < H && xx >= 0 && xx < W;
        boff[k] = live ? kc * HW + yy * W + xx : -1;
        bkc[k] = q < KC * hs ? kc : KC;         // KC: nothing to store
    }
    // load_chunk only issues the loads (into registers); store_chunk, after
    // the chunk's MFMAs, applies the zero selects and the gate and writes LDS,
    // so no wait on these loads falls before the MFMAs
    float4 areg[A_PER];
    float breg[B_PER], greg[B_PER];
    bool aok[A_PER], bok[B_PER];
    auto load_chunk = [&](int ci0) {
        const int lim = a.cin - ci0;            // channels left: the last chunk may be partial
        const float *wc = a.w + (size_t)ci0 * a.cout;
        const float *xc = xi + (size_t)ci0 * HW;
        const float *gc = GATE ? a.gate + (size_t)img * a.cin * HW + (size_t)ci0 * HW : nullptr;
#pragma unroll
        for (int k = 0; k < A_PER; ++k) {
            aok[k] = aoff[k] >= 0 && akc[k] < lim;
            const float *src = wc + (aok[k] ? aoff[k] : 0);
            if constexpr (VEC_A) areg[k] = *reinterpret_cast<const float4 *>(src);
            else areg[k].x = *src;
        }
#pragma unroll
        for (int k = 0; k < B_PER; ++k) {
            bok[k] = boff[k] >= 0 && bkc[k] < lim;
            const int o = bok[k] ? boff[k] : 0;
            breg[k] = xc[o];
            if constexpr (GATE) greg[k] = gc[o];
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int k = 0; k < A_PER; ++k) {
            const int q = tid + k * kThreads;
            if (q < A_LOADS) {
                if constexpr (VEC_A)
                    *reinterpret_cast<float4 *>(&As[buf][q * 4]) = aok[k] ? areg[k] : make_float4(0.f, 0.f, 0.f, 0.f);
                else
                    As[buf][q] = aok[k] ? areg[k].x : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < B_PER; ++k) {
            const int q = tid + k * kThreads;
            float v = breg[k];
            if constexpr (GATE) v = v * (greg[k] > 0.0f ? a.gate_pos : a.gate_neg);
            if (bkc[k] < KC) Bs[buf][q + bkc[k] * (HF - hs)] = bok[k] ? v : 0.0f;   // row kc of the buffer
        }
    };

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
    const int n_chunks = (a.cin + KC - 1) / KC;
    const int c_begin = slice * a.cps;
    const int c_end = min(n_chunks, c_begin + a.cps);
    load_chunk(c_begin * KC);
    store_chunk(0);
    __syncthreads();
    for (int c = c_begin; c < c_end; ++c) {
        const int buf = (c - c_begin) & 1;
        if (c + 1 < c_end) load_chunk((c + 1) * KC);
        const float *A = As[buf] + wm * (BM / 2) + (lane & 31);
        const float *B = Bs[buf];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int bo = (t / 3) * tw2 + (t % 3);
#pragma unroll
            for (int ks = 0; ks < KC / 2; ++ks) {
                const int kc = 2 * ks + kc_lane;
                float av[TM], bv[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) av[i] = A[(kc * 9 + t) * BM + i * 32];
#pragma unroll
                for (int j = 0; j < TN; ++j) bv[j] = B[kc * HF + pix[j] + bo];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i], bv[j], acc[i][j], 0, 0, 0);
            }
        }
        if (c + 1 < c_end) store_chunk(buf ^ 1);
        __syncthreads();
    }

    // epilogue: element r of tile (i, j) is row m = 8 (r / 4) + 4 (lane / 32) + r % 4
    // and column n = lane % 32 of that 32 x 32 tile.  A K slice stores its raw
    // partial sums instead.
    float *yi = a.part ? a.part + ((size_t)slice * a.n + img) * a.cout * HW : a.y + (size_t)img * a.cout * HW;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int nn = wn * (BN / 2) + j * 32 + (lane & 31);
        const int ty = nn / a.tw, tx = nn - ty * a.tw;
        const int yy = y0 + ty, xx = x0 + tx;
        if (yy >= H || xx >= W) continue;
        const size_t p = (size_t)yy * W + xx;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int co = co0 + wm * (BM / 2) + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (co >= a.cout) continue;
                yi[(size_t)co * HW + p] = a.part ? acc[i][j][r] : finish(a, acc[i][j][r], co, p);
            }
        }
    }
}

// Sum of the K slices' partial sums, in slice order, and the epilogue; one
// output element per thread.
__global__ __launch_bounds__(kThreads) void conv_split_sum_kernel(ConvArgs a) {
    const size_t HW = (size_t)a.h * a.w_;
    const size_t total = (size_t)a.n * a.cout * HW;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= total) return;
    float v = a.part[e];
    for (int s = 1; s < a.slices; ++s) v += a.part[(size_t)s * total + e];
    const size_t p = e % HW;
    const int co = (int)((e / HW) % a.cout);
    a.y[e] = finish(a, v, co, p);
}

// Direct convolution for Cout = CO <= 8: a 64 x 8 pixel block, two
// vertically adjacent pixels per thread, the 2 x CO sums in registers.  Per
// chunk of 8 input channels the block stages the 66 x 10 input tiles (with
// the one-pixel halo, zero outside the image; times the gate, when there is
// one) in LDS with coalesced loads, and every thread reads its 4 x 3 window
// per channel from there.  The weights of one input channel and tap are the
// same for every thread: scalar loads.
template <int CO, bool GATE>
__global__ __launch_bounds__(kThreads) void conv3x3_direct_kernel(ConvArgs a) {
    constexpr int TX = 64, TY = 8, HX = TX + 2, HY = TY + 2, CC = 8;
    __shared__ float tile[CC][HY * HX];
    const int H = a.h, W = a.w_;
    const int bx0 = blockIdx.x * TX, by0 = blockIdx.y * TY;
    const int tx = threadIdx.x & 63, ty = (threadIdx.x >> 6) * 2;   // column, first of the two rows
    const int img = blockIdx.z;
    const size_t HW = (size_t)H * W;
    const float *xi = a.x + (size_t)img * a.cin * HW;
    const float *gi = GATE ? a.gate + (size_t)img * a.cin * HW : nullptr;
    const size_t tap_stride = (size_t)a.cin * CO;
    float acc0[CO], acc1[CO];
#pragma unroll
    for (int co = 0; co < CO; ++co) acc0[co] = acc1[co] = 0.0f;
    constexpr int NE = CC * HY * HX, PER = (NE + kThreads - 1) / kThreads;
    for (int c0 = 0; c0 < a.cin; c0 += CC) {
        // every load of the chunk first (branch-free: a clamped address and a
        // select), then the LDS stores
        float xv[PER], gv[PER];
        bool ok[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + k * kThreads;
            const int cc = e / (HY * HX), rem = e - cc * (HY * HX);
            const int r = rem / HX, c = rem - r * HX;
            const int ci = c0 + cc, yy = by0 - 1 + r, xx = bx0 - 1 + c;
            ok[k] = e < NE && ci < a.cin && yy >= 0 && yy < H && xx >= 0 && xx < W;
            const size_t o = ok[k] ? (size_t)ci * HW + (size_t)yy * W + xx : 0;
            xv[k] = xi[o];
            if constexpr (GATE) gv[k] = gi[o];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + k * kThreads;
            float v = xv[k];
            if constexpr (GATE) v = v * (gv[k] > 0.0f ? a.gate_pos : a.gate_neg);
            if (e < NE) (&tile[0][0])[e] = ok[k] ? v : 0.0f;
        }
        __syncthreads();
        const int nc = min(CC, a.cin - c0);
        for (int cc = 0; cc < nc; ++cc) {
            float v[4][3];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) v[r][c] = tile[cc][(ty + r) * HX + tx + c];
            const float *wc = a.w + (size_t)(c0 + cc) * CO;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
#pragma unroll
                for (int co = 0; co < CO; ++co) {
                    const float wv = wc[t * tap_stride + co];
                    acc0[co] += v[t / 3][t % 3] * wv;
                    acc1[co] += v[t / 3 + 1][t % 3] * wv;
                }
            }
        }
        __syncthreads();
    }
    const int xx = bx0 + tx, y0 = by0 + ty;
    if (xx >= W || y0 >= H) return;
    float *yi = a.y + (size_t)img * CO * HW;
    const size_t p0 = (size_t)y0 * W + xx;
#pragma unroll
    for (int co = 0; co < CO; ++co) yi[(size_t)co * HW + p0] = finish(a, acc0[co], co, p0);
    if (y0 + 1 < H) {
#pragma unroll
        for (int co = 0; co < CO; ++co) yi[(size_t)co * HW + p0 + W] = finish(a, acc1[co], co, p0 + W);
    }
}

template <int CO>
void launch_direct(const ConvArgs &a, hipStream_t s) {
    const dim3 grid((unsigned)((a.w_ + 63) / 64), (unsigned)((a.h + 7) / 8), (unsigned)a.n);
    // the gate is a separate instantiation: a null test inside the unrolled
    // loop puts a branch (and a wait) between the loads
    if (a.gate) hipLaunchKernelGGL((conv3x3_direct_kernel<CO, true>), grid, dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL((conv3x3_direct_kernel<CO, false>), grid, dim3(kThreads), 0, s, a);
}

// ---------------------------------------------------------------------------
// Weight gradient: dL/dw[co][ci][t] = sum over images and pixels p of
// gz[co][p] x[ci][p + shift(t)] (gz = dL/d(pre-activation), zero padding).
// Per tap a GEMM with M = output channels, N = input channels, K = pixels;
// the nine taps share the gz operand.  A block owns 64 x 64 (co, ci) for all
// nine taps (each wave a 32 x 32 quarter: 9 MFMA accumulator tiles), and K
// runs over chunks of 16 pixels (an 8 x 2 tile of one image, 4 x 4 for
// narrower images): the chunk's gz
// values and its input tile with a one-pixel halo go to LDS transposed
// (pixel-major, channels contiguous; a row pitch of 97 = 33 mod 64 floats
// keeps the two half-waves' rows, one pixel apart, on disjoint banks and the
// stores of consecutive pixels on distinct ones), double-buffered with the next chunk's loads in flight.  Many
// pixels and few channel tiles: the chunks are split into slices whose
// partial sums a second kernel adds in slice order.
constexpr int kWgM = 64, kWgN = 64, kWgP = 16, kWgPitch = 97;

struct WgradArgs {
    const float *x;    // [N, Cin, H, W]
    const float *gz;   // [N, Cout, H, W]
    float *out;        // [Cout][Cin][9], or [slices][Cout][Cin][9] partial sums
    int n, cin, cout, h, w_;
    int tiles_x, tiles_per_img, n_chunks, cps;
};

// Two waves per SIMD: the nine 32 x 32 accumulator tiles take 144 of a
// wave's 256 registers.
template <int TW>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void conv3x3_wgrad_kernel(WgradArgs a) {
    constexpr int TH = kWgP / TW, TW2 = TW + 2, HS = (TH + 2) * TW2;   // the chunk's pixel tile and halo
    constexpr int G_PER = kWgM * kWgP / kThreads;
    constexpr int X_PER = (kWgN * HS + kThreads - 1) / kThreads;
    __shared__ float Gs[2][kWgP * kWgPitch];
    __shared__ float Xs[2][HS * kWgPitch];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;   // the wave's 32 output and 32 input channels
    const int co0 = blockIdx.x * kWgM, ci0 = blockIdx.y * kWgN;
    const int slice = blockIdx.z;
    const int H = a.h, W = a.w_, HW = H * W;
    const int c_begin = slice * a.cps, c_end = min(a.n_chunks, c_begin + a.cps);

    float greg[G_PER], xreg[X_PER];
    auto load_chunk = [&](int c) {
        const int img = c / a.tiles_per_img, tile = c - img * a.tiles_per_img;
        const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
        const int y0 = ty * TH, x0 = tx * TW;
        const float *gi = a.gz + (size_t)img * a.cout * HW;
        const float *xi = a.x + (size_t)img * a.cin * HW;
#pragma unroll
        for (int k = 0; k < G_PER; ++k) {
            const int q = tid + k * kThreads;
            const int co = q / kWgP, pp = q - co * kWgP;
            const int yy = y0 + pp / TW, xx = x0 + pp % TW;
            const bool ok = co0 + co < a.cout && yy < H && xx < W;
            const float v = gi[ok ? (size_t)(co0 + co) * HW + yy * W + xx : 0];
            greg[k] = ok ? v : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < X_PER; ++k) {
            const int q = tid + k * kThreads;
            const int ci = q / HS, hp = q - ci * HS;
            const int yy = y0 - 1 + hp / TW2, xx = x0 - 1 + hp % TW2;
            const bool ok = q < kWgN * HS && ci0 + ci < a.cin && yy >= 0 && yy < H && xx >= 0 && xx < W;
            const float v = xi[ok ? (size_t)(ci0 + ci) * HW + yy * W + xx : 0];
            xreg[k] = ok ? v : 0.0f;
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int k = 0; k < G_PER; ++k) {
            const int q = tid + k * kThreads;
            const int co = q / kWgP, pp = q - co * kWgP;
            Gs[buf][pp * kWgPitch + co] = greg[k];
        }
#pragma unroll
        for (int k = 0; k < X_PER; ++k) {
            const int q = tid + k * kThreads;
            if (q < kWgN * HS) {
                const int ci = q / HS, hp = q - ci * HS;
                Xs[buf][hp * kWgPitch + ci] = xreg[k];
            }
        }
    };

    v16f acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    if (c_begin < c_end) {
        load_chunk(c_begin);
        store_chunk(0);
    }
    __syncthreads();
    const int kp = lane >> 5;       // which pixel of a K-step pair this lane carries
    for (int c = c_begin; c < c_end; ++c) {
        const int buf = (c - c_begin) & 1;
        if (c + 1 < c_end) load_chunk(c + 1);
        const float *G = Gs[buf] + wm * 32 + (lane & 31);
        const float *X = Xs[buf] + wn * 32 + (lane & 31);
#pragma unroll 2
        for (int ks = 0; ks < kWgP / 2; ++ks) {
            const int pp = 2 * ks + kp;
            const int py = pp / TW, px = pp - py * TW;
            const float av = G[pp * kWgPitch];
            const float *xb = X + (py * TW2 + px) * kWgPitch;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const float bv = xb[((t / 3) * TW2 + (t % 3)) * kWgPitch];
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[t], 0, 0, 0);
            }
        }
        if (c + 1 < c_end) store_chunk(buf ^ 1);
        __syncthreads();
    }

    // element r of tile t: co row 8 (r / 4) + 4 (lane / 32) + r % 4, ci column lane % 32
    float *o = a.out + (size_t)slice * a.cout * a.cin * 9;
    const int ci = ci0 + wn * 32 + (lane & 31);
    if (ci >= a.cin) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int co = co0 + wm * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (co >= a.cout) continue;
#pragma unroll
        for (int t = 0; t < 9; ++t) o[((size_t)co * a.cin + ci) * 9 + t] = acc[t][r];
    }
}

// Sum of the weight-gradient slices: eight running sums over the slices
// s = j mod 8 (eight loads in flight per trip), then combined in a fixed
// order -- the same bits on every run.
__global__ __launch_bounds__(kThreads) void wgrad_sum_kernel(const float *__restrict__ part, int slices, size_t total,
                                                             float *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= total) return;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = 0.0f;
    int s = 0;
    for (; s + 8 <= slices; s += 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] += part[(size_t)(s + j) * total + e];
    }
    for (int j = 0; s + j < slices; ++j) v[j] += part[(size_t)(s + j) * total + e];
    out[e] = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// ---------------------------------------------------------------------------
// 2 x 2 average pooling, stride 2 (torch.nn.AvgPool2d(kernel_size=2), which
// the VGG feature loss puts in place of max pooling), and its adjoint.  One
// thread per output pixel; the window is summed row by row, left to right,
// then divided by 4, as torch's kernel does, so the outputs are the same bits.
// A trailing odd row or column is dropped (its gradient is zero).
__global__ __launch_bounds__(kThreads) void avgpool2_kernel(const float *__restrict__ x, int planes, int h, int w,
                                                            float *__restrict__ y) {
    const int ho = h / 2, wo = w / 2;
    const size_t total = (size_t)planes * ho * wo;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= total) return;
    const int j = (int)(e % wo), i = (int)((e / wo) % ho);
    const size_t pl = e / ((size_t)ho * wo);
    const float *r0 = x + (pl * h + 2 * i) * (size_t)w + 2 * j;
    const float *r1 = r0 + w;
    float v = r0[0];
    v += r0[1];
    v += r1[0];
    v += r1[1];
    y[e] = v / 4.0f;
}

__global__ __launch_bounds__(kThreads) void avgpool2_adjoint_kernel(const float *__restrict__ gy, int planes, int h,
                                                                    int w, float *__restrict__ gx) {
    // one thread per input pair (2 columns of one row): g / 4 where the pair
    // lies in a window, 0 in a dropped row or column
    const int ho = h / 2, wo = w / 2, wp = (w + 1) / 2;
    const size_t total = (size_t)planes * h * wp;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= total) return;
    const int jp = (int)(e % wp), row = (int)((e / wp) % h);
    const size_t pl = e / ((size_t)h * wp);
    const int i = row / 2;
    const float g = (i < ho && jp < wo) ? gy[(pl * ho + i) * (size_t)wo + jp] / 4.0f : 0.0f;
    float *o = gx + (pl * h + row) * (size_t)w + 2 * jp;
    o[0] = g;
    if (2 * jp + 1 < w) o[1] = g;
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

#ifndef SVBRDF_CONV_KC
#define SVBRDF_CONV_KC 4   // input channels per staged chunk
#endif
constexpr int kKC = SVBRDF_CONV_KC;
constexpr int kDirectMaxCout = 8;          // conv3x3_direct_kernel up to this many output channels
constexpr int kMinChunksPerSlice = 2;
constexpr int kMaxSlices = 64;
constexpr int kComputeUnits = 256;   // MI355X

// How one convolution runs: the tile shape, the tiling of the image and the
// K split.  The workspace query and the launch share it.
struct ConvPlan {
    bool direct;
    int bm, bn, tw, th, tiles_x, tiles_y, co_tiles;
    int slices, cps;
    size_t part_floats;   // workspace floats (0 without a K split)
};

// Estimated time (us) of a tiling with `slices` K slices.  Blocks are dealt
// to the compute units evenly, so a unit runs ceil(blocks / CUs) of them one
// after the other (co-resident blocks share its matrix cores); a block costs
// its chunks plus about one chunk of prologue and epilogue, a chunk of a
// BM x BN tile 2 BM BN 9 KC flops at ~0.45 TFLOP/s per compute unit (115
// TFLOP/s over the device, what the 128 x 128 tile sustains) over the tile's
// relative efficiency, and 15 % more where a unit holds a single block (one
// wave per matrix core leaves its pipeline idle between dependent issues);
// a split adds the sum kernel's pass over the partial
// sums (slices reads and one write of the output at ~8 TB/s).  Measured on
// MI355X, this ranks the tilings of the VGG-19 and StyleGAN2 layers as their
// timings do, within ~10 %.
double estimate_us(const ConvPlan &p, int n, size_t out_bytes) {
    const double eff = p.bm == 128 ? (p.bn == 128 ? 1.0 : 0.95) : (p.bn == 128 ? 0.97 : 0.93);
    const long long blocks = (long long)p.tiles_x * p.tiles_y * p.co_tiles * n * p.slices;
    const double rounds = (double)((blocks + kComputeUnits - 1) / kComputeUnits);
    const double chunk_us = 2.0 * p.bm * p.bn * 9 * kKC / 0.45e6 / eff;
    double t = rounds * (p.cps + 1) * chunk_us;
    if (blocks <= kComputeUnits) t /= 0.85;   // one block per unit: one wave per matrix core
    if (p.slices > 1) t += (double)(p.slices + 1) * out_bytes / 8.0e6;
    return t;
}

ConvPlan plan_conv(int n, int cin, int h, int w, int cout) {
    ConvPlan best{};
    best.slices = 1;
    if (cout <= kDirectMaxCout) {
        best.direct = true;
        return best;
    }
    const int n_chunks = (cin + kKC - 1) / kKC;
    const size_t out_bytes = (size_t)n * cout * h * w * sizeof(float);
    const int max_slices = std::max(1, std::min({kMaxSlices, n_chunks / kMinChunksPerSlice, 65535 / n}));
    double best_t = 0.0;
    // 128-channel tiles only where there are that many channels; fewer slices
    // win ties (less workspace)
    for (int bm = 128; bm >= 64; bm -= 64) {
        if (bm == 128 && cout < 128) continue;
        for (int bn = 128; bn >= 64; bn -= 64) {
            ConvPlan p{};
            p.bm = bm;
            p.bn = bn;
            p.co_tiles = (cout + bm - 1) / bm;
            p.tw = pick_tile_width(w);
            p.th = bn / p.tw;
            p.tiles_x = (w + p.tw - 1) / p.tw;
            p.tiles_y = (h + p.th - 1) / p.th;
            for (int s = 1; s <= max_slices; ++s) {
                p.cps = (n_chunks + s - 1) / s;
                p.slices = (n_chunks + p.cps - 1) / p.cps;   // no empty slice
                if (p.slices != s) continue;
                const double t = estimate_us(p, n, out_bytes);
                if (best.bm == 0 || t < 0.98 * best_t) {
                    best = p;
                    best_t = t;
                }
            }
        }
    }
    best.part_floats = best.slices > 1 ? (size_t)best.slices * n * cout * h * w : 0;
    return best;
}

// Tiling of the weight gradient: 64 x 64 (co, ci) blocks, pixel chunks of
// 16 split into slices by the same cost reasoning as plan_conv (rounds of
// blocks per compute unit x chunks per slice, plus the sum pass).
struct WgradPlan {
    int tw, th, tiles_x, tiles_per_img, n_chunks, co_tiles, ci_tiles, slices, cps;
    size_t part_floats;
};

WgradPlan plan_wgrad(int n, int cin, int h, int w, int cout) {
    WgradPlan p{};
    p.tw = std::min(pick_tile_width(w), 8);   // 8 x 2 pixel tiles: the smallest halo, no spills
    p.th = kWgP / p.tw;
    p.tiles_x = (w + p.tw - 1) / p.tw;
    p.tiles_per_img = p.tiles_x * ((h + p.th - 1) / p.th);
    p.n_chunks = n * p.tiles_per_img;
    p.co_tiles = (cout + kWgM - 1) / kWgM;
    p.ci_tiles = (cin + kWgN - 1) / kWgN;
    const long long blocks = (long long)p.co_tiles * p.ci_tiles;
    const double chunk_us = 2.0 * kWgM * kWgN * 9 * kWgP / 0.45e6;
    const double out_bytes = 9.0 * cin * cout * sizeof(float);
    double best_t = 0.0;
    p.slices = 1;
    p.cps = p.n_chunks;
    const int max_slices = std::max(1, std::min(256, p.n_chunks / kMinChunksPerSlice));
    for (int s = 1; s <= max_slices; ++s) {
        const int cps = (p.n_chunks + s - 1) / s;
        if ((p.n_chunks + cps - 1) / cps != s) continue;   // no empty slice
        const long long b = blocks * s;
        double t = (double)((b + kComputeUnits - 1) / kComputeUnits) * (cps + 1) * chunk_us;
        if (b <= kComputeUnits) t /= 0.85;
        // the slices' partial sums: written by the blocks, read by the sum
        // pass (measured ~5 TB/s at 150 MB)
        if (s > 1) t += 2.0 * s * out_bytes / 5.0e6 + 2.0;
        if (s == 1 || t < 0.98 * best_t) {
            best_t = t;
            p.slices = s;
            p.cps = cps;
        }
    }
    p.part_floats = p.slices > 1 ? (size_t)p.slices * 9 * cin * cout : 0;
    return p;
}

template <int BM, int BN>
void launch_mfma(const ConvArgs &a, const ConvPlan &p, hipStream_t s) {
    const dim3 grid((unsigned)(p.tiles_x * p.tiles_y), (unsigned)p.co_tiles, (unsigned)(a.n * p.slices));
    // float4 weight loads need 16-byte aligned rows of Cout floats; the input
    // gate is its own instantiation (no null test between the loads)
    const bool vec = a.cout % 4 == 0 && ((uintptr_t)a.w & 15u) == 0;
    if (a.gate) {
        if (vec) hipLaunchKernelGGL((conv3x3_mfma_kernel<BM, BN, kKC, true, true>), grid, dim3(kThreads), 0, s, a);
        else hipLaunchKernelGGL((conv3x3_mfma_kernel<BM, BN, kKC, false, true>), grid, dim3(kThreads), 0, s, a);
    } else {
        if (vec) hipLaunchKernelGGL((conv3x3_mfma_kernel<BM, BN, kKC, true, false>), grid, dim3(kThreads), 0, s, a);
        else hipLaunchKernelGGL((conv3x3_mfma_kernel<BM, BN, kKC, false, false>), grid, dim3(kThreads), 0, s, a);
    }
}

}  // namespace

extern "C" {

int svbrdf_conv3x3_pack(const float *weight, int cout, int cin, const float *style, float scale, int demodulate,
                        int transpose, float *packed, void *stream) {
    const char *what = "svbrdf_conv3x3_pack";
    if (!weight || !packed) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (cout < 1 || cin < 1 || cout > kMaxChannels || cin > kMaxChannels)
        return fail(SVBRDF_ESHAPE, "%s: cout=%d cin=%d outside [1, %d]", what, cout, cin, kMaxChannels);
    hipStream_t s = (hipStream_t)stream;
    const int dm = demodulate ? 1 : 0, tr = transpose ? 1 : 0;
    const bool aligned = ((uintptr_t)packed & 15u) == 0;
    if (!transpose && aligned && cout % 4 == 0)
        hipLaunchKernelGGL(conv3x3_pack_kernel<1>, dim3((unsigned)(cout / 4)), dim3(kThreads), 0, s, weight, cout, cin,
                           style, scale, dm, tr, packed);
    else if (transpose && aligned && cin % 4 == 0)
        hipLaunchKernelGGL(conv3x3_pack_kernel<2>, dim3((unsigned)cout), dim3(kThreads), 0, s, weight, cout, cin, style,
                           scale, dm, tr, packed);
    else
        hipLaunchKernelGGL(conv3x3_pack_kernel<0>, dim3((unsigned)cout), dim3(kThreads), 0, s, weight, cout, cin, style,
                           scale, dm, tr, packed);
    return check_launch(what);
}

int svbrdf_act_grad(const float *grad_y, const float *y, size_t n, float pos, float neg, float *out,
                    void *stream) {
    const char *what = "svbrdf_act_grad";
    if (!grad_y || !y || !out) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (n == 0) return SVBRDF_OK;
    hipStream_t s = (hipStream_t)stream;
    const bool vec = n % 4 == 0 && ((uintptr_t)grad_y & 15u) == 0 && ((uintptr_t)y & 15u) == 0 &&
                     ((uintptr_t)out & 15u) == 0;
    if (vec) {
        const size_t n4 = n / 4;
        hipLaunchKernelGGL(act_grad_kernel, dim3((unsigned)((n4 + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                           (const float4 *)grad_y, (const float4 *)y, n4, pos, neg, (float4 *)out);
    } else {
        hipLaunchKernelGGL(act_grad_tail_kernel, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                           grad_y, y, n, pos, neg, out);
    }
    return check_launch(what);
}

int svbrdf_conv3x3_modulation_grad(const float *weight, int cout, int cin, const float *style, float scale,
                                   int demodulate, const float *grad_weight, float *grad_style, void *workspace,
                                   size_t workspace_bytes, void *stream) {
    const char *what = "svbrdf_conv3x3_modulation_grad";
    if (!weight || !style || !grad_weight || !grad_style || !workspace)
        return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (cout < 1 || cin < 1 || cout > kMaxChannels || cin > kMaxChannels)
        return fail(SVBRDF_ESHAPE, "%s: cout=%d cin=%d outside [1, %d]", what, cout, cin, kMaxChannels);
    if (workspace_bytes < (size_t)cout * sizeof(float2))
        return fail(SVBRDF_ESPACE, "%s: workspace of %zu bytes < %zu needed", what, workspace_bytes,
                    (size_t)cout * sizeof(float2));
    if ((uintptr_t)workspace % sizeof(float2))
        return fail(SVBRDF_EINVAL, "%s: workspace must be 8-byte aligned", what);
    hipStream_t s = (hipStream_t)stream;
    float2 *de = static_cast<float2 *>(workspace);
    hipLaunchKernelGGL(modulation_rows_kernel, dim3((unsigned)cout), dim3(kThreads), 0, s, weight, cin, style, scale,
                       demodulate ? 1 : 0, grad_weight, de);
    if (int rc = check_launch(what)) return rc;
    hipLaunchKernelGGL(modulation_cols_kernel, dim3((unsigned)cin), dim3(kThreads), 0, s, weight, cout, cin, style,
                       scale, grad_weight, (const float2 *)de, grad_style);
    return check_launch(what);
}

static bool conv_shape_ok(int n, int cin, int height, int width, int cout) {
    return n >= 1 && n <= 65535 && cin >= 1 && cout >= 1 && cin <= kMaxChannels && cout <= kMaxChannels &&
           height >= 1 && width >= 1 && height <= kMaxSide && width <= kMaxSide;
}

size_t svbrdf_conv3x3_workspace_bytes(int n, int cin, int height, int width, int cout) {
    if (!conv_shape_ok(n, cin, height, width, cout)) return 0;
    return plan_conv(n, cin, height, width, cout).part_floats * sizeof(float);
}

static int run_conv(const char *what, ConvArgs a, void *workspace, size_t workspace_bytes, hipStream_t s) {
    if (!conv_shape_ok(a.n, a.cin, a.h, a.w_, a.cout))
        return fail(SVBRDF_ESHAPE, "%s: n=%d cin=%d cout=%d height=%d width=%d", what, a.n, a.cin, a.cout, a.h,
                    a.w_);
    if (a.act < SVBRDF_ACT_NONE || a.act > SVBRDF_ACT_LRELU)
        return fail(SVBRDF_EINVAL, "%s: unknown activation %d", what, a.act);
    const ConvPlan p = plan_conv(a.n, a.cin, a.h, a.w_, a.cout);
    const size_t need = p.part_floats * sizeof(float);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SVBRDF_ESPACE, "%s: workspace of %zu bytes, this shape needs %zu "
                    "(svbrdf_conv3x3_workspace_bytes)", what, workspace ? workspace_bytes : (size_t)0, need);
    a.part = p.slices > 1 ? (float *)workspace : nullptr;
    a.tw = p.tw; a.th = p.th; a.tiles_x = p.tiles_x; a.slices = p.slices; a.cps = p.cps;
    if (p.direct) {
        switch (a.cout) {
            case 1: launch_direct<1>(a, s); break;
            case 2: launch_direct<2>(a, s); break;
            case 3: launch_direct<3>(a, s); break;
            case 4: launch_direct<4>(a, s); break;
            case 5: launch_direct<5>(a, s); break;
            case 6: launch_direct<6>(a, s); break;
            case 7: launch_direct<7>(a, s); break;
            default: launch_direct<8>(a, s); break;
        }
        return check_launch(what);
    }
    if (p.bm == 128 && p.bn == 128) launch_mfma<128, 128>(a, p, s);
    else if (p.bm == 128) launch_mfma<128, 64>(a, p, s);
    else if (p.bn == 128) launch_mfma<64, 128>(a, p, s);
    else launch_mfma<64, 64>(a, p, s);
    if (p.slices > 1) {
        const size_t total = (size_t)a.n * a.cout * a.h * a.w_;
        hipLaunchKernelGGL(conv_split_sum_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, s, a);
    }
    return check_launch(what);
}

int svbrdf_conv3x3(const float *x, int n, int cin, int height, int width, const float *packed, int cout,
                   const float *bias, const float *noise, float noise_strength, int activation, float *y,
                   void *workspace, size_t workspace_bytes, void *stream) {
    const char *what = "svbrdf_conv3x3";
    if (!x || !packed || !y) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    ConvArgs a{};
    a.x = x; a.w = packed; a.bias = bias; a.noise = noise; a.noise_strength = noise_strength; a.act = activation;
    a.y = y; a.n = n; a.cin = cin; a.cout = cout; a.h = height; a.w_ = width;
    return run_conv(what, a, workspace, workspace_bytes, (hipStream_t)stream);
}

int svbrdf_conv3x3_input_grad(const float *grad_y, const float *y, int activation, int n, int cout, int height,
                              int width, const float *packed_t, int cin, float *grad_x, void *workspace,
                              size_t workspace_bytes, void *stream) {
    const char *what = "svbrdf_conv3x3_input_grad";
    if (!grad_y || !y || !packed_t || !grad_x) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (activation < SVBRDF_ACT_NONE || activation > SVBRDF_ACT_LRELU)
        return fail(SVBRDF_EINVAL, "%s: unknown activation %d", what, activation);
    ConvArgs a{};
    // the convolution of dL/dy (cout channels) with the flipped, transposed
    // weights, its input gated by the activation's derivative at y
    a.x = grad_y; a.w = packed_t; a.act = SVBRDF_ACT_NONE; a.y = grad_x;
    a.n = n; a.cin = cout; a.cout = cin; a.h = height; a.w_ = width;
    if (activation != SVBRDF_ACT_NONE) {
        a.gate = y;
        a.gate_pos = activation == SVBRDF_ACT_LRELU ? 1.41421356237309515f : 1.0f;
        a.gate_neg = activation == SVBRDF_ACT_LRELU ? 0.2f * 1.41421356237309515f : 0.0f;
    }
    return run_conv(what, a, workspace, workspace_bytes, (hipStream_t)stream);
}

size_t svbrdf_conv3x3_wgrad_workspace_bytes(int n, int cin, int height, int width, int cout) {
    if (!conv_shape_ok(n, cin, height, width, cout)) return 0;
    return plan_wgrad(n, cin, height, width, cout).part_floats * sizeof(float);
}

int svbrdf_conv3x3_wgrad(const float *x, const float *grad_z, int n, int cin, int height, int width, int cout,
                         float *grad_w, void *workspace, size_t workspace_bytes, void *stream) {
    const char *what = "svbrdf_conv3x3_wgrad";
    if (!x || !grad_z || !grad_w) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (!conv_shape_ok(n, cin, height, width, cout))
        return fail(SVBRDF_ESHAPE, "%s: n=%d cin=%d cout=%d height=%d width=%d", what, n, cin, cout, height, width);
    const WgradPlan p = plan_wgrad(n, cin, height, width, cout);
    const size_t need = p.part_floats * sizeof(float);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SVBRDF_ESPACE, "%s: workspace of %zu bytes, this shape needs %zu "
                    "(svbrdf_conv3x3_wgrad_workspace_bytes)", what, workspace ? workspace_bytes : (size_t)0, need);
    WgradArgs a{};
    a.x = x; a.gz = grad_z; a.out = p.slices > 1 ? (float *)workspace : grad_w;
    a.n = n; a.cin = cin; a.cout = cout; a.h = height; a.w_ = width;
    a.tiles_x = p.tiles_x; a.tiles_per_img = p.tiles_per_img; a.n_chunks = p.n_chunks;
    a.cps = p.cps;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)p.co_tiles, (unsigned)p.ci_tiles, (unsigned)p.slices);
    if (p.tw == 8) hipLaunchKernelGGL(conv3x3_wgrad_kernel<8>, grid, dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL(conv3x3_wgrad_kernel<4>, grid, dim3(kThreads), 0, s, a);
    if (p.slices > 1) {
        const size_t total = (size_t)9 * cin * cout;
        hipLaunchKernelGGL(wgrad_sum_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           s, (const float *)workspace, p.slices, total, grad_w);
    }
    return check_launch(what);
}

int svbrdf_avgpool2(const float *x, int planes, int height, int width, float *y, void *stream) {
    const char *what = "svbrdf_avgpool2";
    if (!x || !y) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (planes < 1 || height < 2 || width < 2 || height > kMaxSide || width > kMaxSide ||
        (size_t)planes * height * width > ((size_t)1 << 34))
        return fail(SVBRDF_ESHAPE, "%s: planes=%d height=%d width=%d", what, planes, height, width);
    const size_t total = (size_t)planes * (height / 2) * (width / 2);
    hipLaunchKernelGGL(avgpool2_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       (hipStream_t)stream, x, planes, height, width, y);
    return check_launch(what);
}

int svbrdf_avgpool2_adjoint(const float *grad_y, int planes, int height, int width, float *grad_x, void *stream) {
    const char *what = "svbrdf_avgpool2_adjoint";
    if (!grad_y || !grad_x) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (planes < 1 || height < 2 || width < 2 || height > kMaxSide || width > kMaxSide ||
        (size_t)planes * height * width > ((size_t)1 << 34))
        return fail(SVBRDF_ESHAPE, "%s: planes=%d height=%d width=%d", what, planes, height, width);
    const size_t total = (size_t)planes * height * ((width + 1) / 2);
    hipLaunchKernelGGL(avgpool2_adjoint_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, (hipStream_t)stream, grad_y, planes, height, width, grad_x);
    return check_launch(what);
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
