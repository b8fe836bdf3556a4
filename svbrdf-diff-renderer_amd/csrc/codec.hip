// codec.hip -- gfx950 kernels for the map / target codec on either side of the
// optimisation path, and their extern "C" entry points (include/svbrdf_hip.h).
//
// Reference behaviour (src/imageio.py:11-76, src/svbrdf.py:126-132,150-221):
// PNG files hold 8-bit maps and images; loading resizes the 8-bit image with
// cv2.INTER_LANCZOS4, divides by 255 and applies the per-map flag; saving
// clips, scales by 255 and truncates to 8 bits.  The PNG container itself
// (inflate / deflate, scanline filters) is sequential byte work and stays on
// the host; every per-pixel step is here.
//
// Kernels (byte and float elementwise work, HBM-bound):
//   resize_lanczos4_kernel  8-bit Lanczos-4 resample, 8x8 taps, integer-exact
//   images_from_u8_kernel   [n, h, w, C] u8 -> [n, C', h, w] f32   C B read, 4 C' B written / px
//   textures_from_u8_kernel 4 maps u8 -> [9, R, R] f32             10-12 B read, 36 B written / px
//   textures_to_u8_kernel   [9, R, R] f32 -> 4 maps u8             36 B read, 10 B written / px
//   images_to_u8_kernel     [n, C, h, w] f32 -> [n, h, w, C] u8    4 C B read, C B written / px
//   upsample2x_kernel       [P, R, R] f32 -> [P, 2R, 2R] f32 bilinear (stage chaining in memory);
//   upsample2x_rows_kernel  the same, a strip of 4 source rows x 2 columns (8 x 4 outputs) per thread
//                           (even R, aligned rows)
//
// Float arithmetic follows the reference's float32 numpy / torch op order one
// rounding per op: contraction into fused multiply-adds is off in this file,
// and `/` and sqrtf are correctly rounded (hipcc's default fp32 divide / sqrt
// lowering; the __fsqrt_rn intrinsic is the native approximation instead).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "../../include/svbrdf_hip.h"

#pragma clang fp contract(off)

using svbrdf_internal::check_launch;
using svbrdf_internal::fail;

namespace {

constexpr int kBlock = 256;
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kTile = 16;                 // resize: 16 x 16 output pixels per block
constexpr int kMaxSide = 16384;           // largest 8-bit image side accepted
constexpr int kMaxRes = 7680;             // as the renderer's maximum resolution
constexpr int kCoefBits = 11;             // fixed-point bits of the resize weights

// ---------------------------------------------------------------------------
// Lanczos-4 resize of 8-bit images, as OpenCV's INTER_LANCZOS4 for 8U data
// (modules/imgproc/src/resize.cpp).  Destination index d samples source
// coordinate f = (d + 0.5) * scale - 0.5, scale = 1 / (dst / src) in double,
// rounded to float; taps floor(f) - 3 .. floor(f) + 4, border replicated.
// Weights: the window sinc(t) sinc(t/4) at t_i = frac + 3 - i, which is
// (-1)^i sin(a - i pi/4) / t_i^2 up to a factor common to all taps
// (a = (frac + 3) pi / 4); normalised in float, scaled by 2^11, rounded
// half-to-even.  Sums are exact integers; the result is (sum + 2^21) >> 22
// saturated to [0, 255].
struct AxisPlan {
    int first;   // source index of tap 0 (before clamping)
    int w[8];
};

constexpr int kPlans = 2 * kTile;          // 16 column plans + 16 row plans per tile
constexpr int kMaxSpan = 64;              // source rows a tile may stage in LDS

// Source coordinate of destination index d: floor(f) and the fraction.
__device__ __forceinline__ void source_coord(int d, int n_src, int n_dst, int &first, float &frac) {
    const double scale = 1.0 / ((double)n_dst / (double)n_src);
    const float f = (float)(((double)d + 0.5) * scale - 0.5);
    const float s = floorf(f);
    frac = f - s;
    first = (int)s - 3;
}

// Float weights of the 8 taps for fraction x, before normalisation.  The
// angle a0 = -(x + 3) pi / 4 is common to all taps (tap i's sine is a signed
// combination of sin(a0) and cos(a0)), so one sin / cos pair serves a plan.
__device__ __forceinline__ void lanczos4_taps(float x, float (&w)[8]) {
    constexpr double kS45 = 0.70710678118654752440;
    // (-1)^(i+1) (cos(i pi/4), sin(i pi/4))
    constexpr double kA[8] = {-1.0, kS45, 0.0, -kS45, 1.0, -kS45, 0.0, kS45};
    constexpr double kB[8] = {0.0, kS45, -1.0, kS45, 0.0, -kS45, 1.0, -kS45};
    const float xp3 = x + 3.0f;
    const double a0 = -(double)xp3 * M_PI * 0.25;
    const double sa = sin(a0), ca = cos(a0);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float t = xp3 - (float)i;
        const double y = -(double)t * M_PI * 0.25;
        // |t| < 1e-6: the unit impulse after normalisation
        w[i] = fabsf(t) < 1e-6f ? 1e30f : (float)((kA[i] * sa + kB[i] * ca) / (y * y));
    }
}

// The tile's 16 column and 16 row plans, one plan per thread of the first 32:
// float weights, their sequential float sum and its reciprocal, then the
// rounded 11-bit weights.
__device__ void tile_plans(int x0, int y0, int sh, int sw, int dh, int dw, AxisPlan *plans) {
    const int plan = threadIdx.x;
    if (plan < kPlans) {
        const bool is_x = plan < kTile;
        const int d = is_x ? x0 + plan : y0 + plan - kTile;
        if (d < (is_x ? dw : dh)) {
            int first = 0;
            float frac = 0.0f;
            source_coord(d, is_x ? sw : sh, is_x ? dw : dh, first, frac);
            float w[8];
            lanczos4_taps(frac, w);
            float total = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; ++k) total = total + w[k];
            const float inv = 1.0f / total;
            plans[plan].first = first;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float v = rintf((w[i] * inv) * (float)(1 << kCoefBits));
                plans[plan].w[i] = (int)fminf(fmaxf(v, -32768.0f), 32767.0f);
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ uint8_t finish(int acc) {
    const int v = (acc + (1 << (2 * kCoefBits - 1))) >> (2 * kCoefBits);
    return (uint8_t)min(max(v, 0), 255);
}

// One 16 x 16 output tile per block.  When the tile's source rows fit
// kMaxSpan, the horizontal 8-tap sums of those rows are staged in LDS and the
// vertical pass reads them (separable: 8 + 8 taps per output instead of 64);
// otherwise (strong downscaling) each output sums its 8 x 8 taps directly.
// Integer sums make both orders exact and equal.
template <int C>
__global__ __launch_bounds__(kBlock) void resize_lanczos4_kernel(const uint8_t *__restrict__ src, int sh, int sw,
                                                                 uint8_t *__restrict__ dst, int dh, int dw) {
    __shared__ AxisPlan plans[kPlans];
    __shared__ int hsum[kMaxSpan][kTile][C];
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    tile_plans(x0, y0, sh, sw, dh, dw, plans);
    const AxisPlan *px = plans, *py = plans + kTile;
    const int rows_live = min(kTile, dh - y0), cols_live = min(kTile, dw - x0);
    const int r0 = py[0].first;
    const int span = py[rows_live - 1].first + 8 - r0;   // uniform across the block
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    if (span <= kMaxSpan) {
        for (int idx = threadIdx.x; idx < span * kTile; idx += kBlock) {
            const int r = idx / kTile, xx = idx % kTile;
            if (xx >= cols_live) continue;
            const uint8_t *row = src + (size_t)min(max(r0 + r, 0), sh - 1) * sw * C;
            const AxisPlan &ax = px[xx];
            int h[C];
#pragma unroll
            for (int c = 0; c < C; ++c) h[c] = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint8_t *q = row + (size_t)min(max(ax.first + j, 0), sw - 1) * C;
#pragma unroll
                for (int c = 0; c < C; ++c) h[c] += (int)q[c] * ax.w[j];
            }
#pragma unroll
            for (int c = 0; c < C; ++c) hsum[r][xx][c] = h[c];
        }
        __syncthreads();
        if (tx >= cols_live || ty >= rows_live) return;
        const AxisPlan &ay = py[ty];
        int acc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] += hsum[ay.first + k - r0][tx][c] * ay.w[k];
        uint8_t *o = dst + ((size_t)(y0 + ty) * dw + x0 + tx) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) o[c] = finish(acc[c]);
        return;
    }
    if (tx >= cols_live || ty >= rows_live) return;
    const AxisPlan &ax = px[tx], &ay = py[ty];
    int cols[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) cols[j] = min(max(ax.first + j, 0), sw - 1);
    int acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint8_t *r = src + (size_t)min(max(ay.first + k, 0), sh - 1) * sw * C;
        int h[C];
#pragma unroll
        for (int c = 0; c < C; ++c) h[c] = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int c = 0; c < C; ++c) h[c] += (int)r[(size_t)cols[j] * C + c] * ax.w[j];
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] += h[c] * ay.w[k];
    }
    uint8_t *o = dst + ((size_t)(y0 + ty) * dw + x0 + tx) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) o[c] = finish(acc[c]);
}

// ---------------------------------------------------------------------------
// Per-pixel conversions.  u / 255 is the correctly rounded float quotient, as
// numpy's float32 division (src/imageio.py:17-18).
__device__ __forceinline__ float unit(uint8_t u) { return (float)u / 255.0f; }

// (v * 255).astype("uint8") of a value already clipped to [0, 1]: truncation.
__device__ __forceinline__ uint8_t to_u8(float v) { return (uint8_t)(int)(v * 255.0f); }

// NaN-preserving clip, as numpy's clip.
__device__ __forceinline__ float clip(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// imread's per-flag decode (src/imageio.py:27-47) of one pixel of an 8-bit
// image already at its final size.
__device__ __forceinline__ float decode_rough(const uint8_t *px, int channels) {
    if (channels == 3) {
        // np.mean over the channel axis: (a + b) + c, then / 3
        const float s = (unit(px[0]) + unit(px[1])) + unit(px[2]);
        return s / 3.0f;
    }
    return unit(px[0]);
}

// "normal": im * 2 - 1 divided by np.linalg.norm (sqrt of the sequential sum
// of squares)
__device__ __forceinline__ void decode_normal(const uint8_t *px, float n[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = unit(px[c]) * 2.0f - 1.0f;
    const float norm = sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = n[c] / norm;
}

// imread after the resize, for n images of h x w pixels: [n, h, w, C] bytes ->
// [n, C', h, w] float32 with C' = C (no flag, "srgb"), 1 ("rough") or 3
// ("normal").
__global__ __launch_bounds__(kBlock) void images_from_u8_kernel(const uint8_t *__restrict__ in, size_t total,
                                                                unsigned npix, int channels, int flag,
                                                                float *__restrict__ out) {
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= total) return;
    const size_t img = p / npix, q = p - img * npix;
    const uint8_t *src = in + p * (size_t)channels;
    if (flag == SVBRDF_IMREAD_ROUGH) {
        out[img * npix + q] = decode_rough(src, channels);
    } else if (flag == SVBRDF_IMREAD_NORMAL) {
        float n[3];
        decode_normal(src, n);
        float *dst = out + img * 3 * (size_t)npix + q;
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[(size_t)c * npix] = n[c];
    } else {
        float *dst = out + img * (size_t)channels * npix + q;
        for (int c = 0; c < channels; ++c) dst[(size_t)c * npix] = unit(src[c]);
    }
}

// load_textures_th (src/svbrdf.py:150-166) with imread flags (src/imageio.py:
// 27-47): maps already at res x res, RGB bytes (roughness 1 or 3 channels).
__global__ __launch_bounds__(kBlock) void textures_from_u8_kernel(const uint8_t *__restrict__ dif,
                                                                  const uint8_t *__restrict__ nom,
                                                                  const uint8_t *__restrict__ spe,
                                                                  const uint8_t *__restrict__ rgh, int rgh_channels,
                                                                  unsigned npix, float *__restrict__ tex) {
    const unsigned p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float out[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        out[c] = unit(dif[3 * (size_t)p + c]) * 2.0f - 1.0f;
        out[6 + c] = unit(spe[3 * (size_t)p + c]) * 2.0f - 1.0f;
    }
    out[5] = decode_rough(rgh + (size_t)rgh_channels * p, rgh_channels) * 2.0f - 1.0f;
    float n[3];
    decode_normal(nom + 3 * (size_t)p, n);
    out[3] = n[0];
    out[4] = n[1];
#pragma unroll
    for (int c = 0; c < 9; ++c) tex[(size_t)c * npix + p] = out[c];
}

// save_textures_th (src/svbrdf.py:168-189) with imwrite flags (src/imageio.py:
// 52-72) for one texel: [9] floats -> dif / nom / spe [3] and rgh bytes.
// The normal is SvbrdfIO.reconstruct_normal (src/svbrdf.py:126-132): z from
// clamp(x^2 + y^2, 0, 1), divided by the 2-norm summed as (x^2 + y^2) + z^2.
__device__ __forceinline__ void texel_to_u8(const float t[9], uint8_t dif[3], uint8_t nom[3], uint8_t spe[3],
                                            uint8_t &rgh) {
    auto half = [](float v) { return (v + 1.0f) / 2.0f; };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dif[c] = to_u8(clip(half(t[c]), 0.0f, 1.0f));
        spe[c] = to_u8(clip(half(t[6 + c]), 0.0f, 1.0f));
    }
    rgh = to_u8(clip(half(t[5]), 0.0f, 1.0f));
    const float x = clip(t[3], -1.0f, 1.0f);
    const float y = clip(t[4], -1.0f, 1.0f);
    const float xy = x * x + y * y;
    const float z = sqrtf(1.0f - clip(xy, 0.0f, 1.0f));
    const float norm = sqrtf(xy + z * z);
    const float n[3] = {x / norm, y / norm, z / norm};
#pragma unroll
    for (int c = 0; c < 3; ++c) nom[c] = to_u8(half(clip(n[c], -1.0f, 1.0f)));
}

// [9, R, R] float32 -> dif / nom / spe [R, R, 3] and rgh [R, R] bytes.
__global__ __launch_bounds__(kBlock) void textures_to_u8_kernel(const float *__restrict__ tex, unsigned npix,
                                                                uint8_t *__restrict__ dif, uint8_t *__restrict__ nom,
                                                                uint8_t *__restrict__ spe, uint8_t *__restrict__ rgh) {
    const unsigned p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float t[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) t[c] = tex[(size_t)c * npix + p];
    uint8_t d[3], n[3], sp[3], r;
    texel_to_u8(t, d, n, sp, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dif[3 * (size_t)p + c] = d[c];
        nom[3 * (size_t)p + c] = n[c];
        spe[3 * (size_t)p + c] = sp[c];
    }
    rgh[p] = r;
}

// imwrite before the PNG encode (src/imageio.py:54-72), for n images of
// h x w pixels: [n, C, h, w] float32 -> [n, h, w, C] bytes.  "srgb" and
// "rough" clip to [0, 1]; "normal" clips to [-1, 1] and maps (n + 1) / 2;
// without a flag the reference casts unclipped values, which is only defined
// inside [0, 1], so values are clipped there too.
__global__ __launch_bounds__(kBlock) void images_to_u8_kernel(const float *__restrict__ in, size_t total,
                                                              unsigned npix, int channels, int flag,
                                                              uint8_t *__restrict__ out) {
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= total) return;
    const size_t img = p / npix, q = p - img * npix;
    const float *src = in + img * (size_t)channels * npix + q;
    for (int c = 0; c < channels; ++c) {
        const float v = src[(size_t)c * npix];
        out[p * channels + c] = flag == SVBRDF_IMREAD_NORMAL ? to_u8((clip(v, -1.0f, 1.0f) + 1.0f) / 2.0f)
                                                             : to_u8(clip(v, 0.0f, 1.0f));
    }
}

// ---------------------------------------------------------------------------
// Four pixels per thread: the common layouts (3-channel images, pixel counts
// divisible by 4, 4-byte aligned bytes and 16-byte aligned floats) move as
// whole dwords and float4s instead of single bytes and floats.  The per-pixel
// arithmetic is the functions above, so the results are the same bits.
struct Px4 {
    uint8_t b[12];   // 4 pixels x 3 channels, as stored
};
__device__ __forceinline__ Px4 load_px4(const uint8_t *base, size_t group) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(base) + 3 * group;
    Px4 p;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = w[k];
#pragma unroll
        for (int i = 0; i < 4; ++i) p.b[4 * k + i] = (uint8_t)(v >> (8 * i));
    }
    return p;
}
__device__ __forceinline__ void store_px4(uint8_t *base, size_t group, const uint8_t (&b)[12]) {
    uint32_t *w = reinterpret_cast<uint32_t *>(base) + 3 * group;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        w[k] = (uint32_t)b[4 * k] | ((uint32_t)b[4 * k + 1] << 8) | ((uint32_t)b[4 * k + 2] << 16) |
               ((uint32_t)b[4 * k + 3] << 24);
}
// Non-temporal: the decoded planes stream past the caches to their consumer
// (images_from_u8 22.4 -> 20.1 us, textures_from_u8 11.6 -> 9.4 us at 1024^2).
__device__ __forceinline__ void store4(float *plane, size_t q, float a, float b, float c, float d) {
    const v4f v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(plane + q));
}
__device__ __forceinline__ float4 load4(const float *plane, size_t q) {
    return *reinterpret_cast<const float4 *>(plane + q);
}

// images_from_u8_kernel for 3-channel images, npix % 4 == 0.
__global__ __launch_bounds__(kBlock) void images_from_u8_x4_kernel(const uint8_t *__restrict__ in, size_t groups,
                                                                   unsigned npix, int flag, float *__restrict__ out) {
    const size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= groups) return;
    const size_t p = 4 * g, img = p / npix, q = p - img * npix;   // the 4 pixels share one image
    const Px4 x = load_px4(in, g);
    if (flag == SVBRDF_IMREAD_ROUGH) {
        store4(out + img * npix, q, decode_rough(x.b, 3), decode_rough(x.b + 3, 3), decode_rough(x.b + 6, 3),
               decode_rough(x.b + 9, 3));
    } else if (flag == SVBRDF_IMREAD_NORMAL) {
        float n[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) decode_normal(x.b + 3 * k, n[k]);
#pragma unroll
        for (int c = 0; c < 3; ++c) store4(out + (img * 3 + c) * npix, q, n[0][c], n[1][c], n[2][c], n[3][c]);
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            store4(out + (img * 3 + c) * npix, q, unit(x.b[c]), unit(x.b[3 + c]), unit(x.b[6 + c]),
                   unit(x.b[9 + c]));
    }
}

// textures_from_u8_kernel, npix % 4 == 0.
__global__ __launch_bounds__(kBlock) void textures_from_u8_x4_kernel(const uint8_t *__restrict__ dif,
                                                                     const uint8_t *__restrict__ nom,
                                                                     const uint8_t *__restrict__ spe,
                                                                     const uint8_t *__restrict__ rgh,
                                                                     int rgh_channels, unsigned npix,
                                                                     float *__restrict__ tex) {
    const unsigned g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= npix / 4) return;
    const unsigned q = 4 * g;
    const Px4 d = load_px4(dif, g), s = load_px4(spe, g), n = load_px4(nom, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        store4(tex + (size_t)c * npix, q, unit(d.b[c]) * 2.0f - 1.0f, unit(d.b[3 + c]) * 2.0f - 1.0f,
               unit(d.b[6 + c]) * 2.0f - 1.0f, unit(d.b[9 + c]) * 2.0f - 1.0f);
        store4(tex + (size_t)(6 + c) * npix, q, unit(s.b[c]) * 2.0f - 1.0f, unit(s.b[3 + c]) * 2.0f - 1.0f,
               unit(s.b[6 + c]) * 2.0f - 1.0f, unit(s.b[9 + c]) * 2.0f - 1.0f);
    }
    float r[4];
    if (rgh_channels == 3) {
        const Px4 x = load_px4(rgh, g);
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = decode_rough(x.b + 3 * k, 3) * 2.0f - 1.0f;
    } else {
        const uint32_t w = reinterpret_cast<const uint32_t *>(rgh)[g];
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = unit((uint8_t)(w >> (8 * k))) * 2.0f - 1.0f;
    }
    store4(tex + (size_t)5 * npix, q, r[0], r[1], r[2], r[3]);
    float nv[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) decode_normal(n.b + 3 * k, nv[k]);
    store4(tex + (size_t)3 * npix, q, nv[0][0], nv[1][0], nv[2][0], nv[3][0]);
    store4(tex + (size_t)4 * npix, q, nv[0][1], nv[1][1], nv[2][1], nv[3][1]);
}

// textures_to_u8_kernel, npix % 4 == 0.
__global__ __launch_bounds__(kBlock) void textures_to_u8_x4_kernel(const float *__restrict__ tex, unsigned npix,
                                                                   uint8_t *__restrict__ dif, uint8_t *__restrict__ nom,
                                                                   uint8_t *__restrict__ spe, uint8_t *__restrict__ rgh) {
    const unsigned g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= npix / 4) return;
    const unsigned q = 4 * g;
    float4 v[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) v[c] = load4(tex + (size_t)c * npix, q);
    uint8_t d[12], n[12], s[12], r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float t[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) t[c] = k == 0 ? v[c].x : k == 1 ? v[c].y : k == 2 ? v[c].z : v[c].w;
        texel_to_u8(t, d + 3 * k, n + 3 * k, s + 3 * k, r[k]);
    }
    store_px4(dif, g, d);
    store_px4(nom, g, n);
    store_px4(spe, g, s);
    reinterpret_cast<uint32_t *>(rgh)[g] =
        (uint32_t)r[0] | ((uint32_t)r[1] << 8) | ((uint32_t)r[2] << 16) | ((uint32_t)r[3] << 24);
}

// images_to_u8_kernel for 3-channel images, npix % 4 == 0.
__global__ __launch_bounds__(kBlock) void images_to_u8_x4_kernel(const float *__restrict__ in, size_t groups,
                                                                 unsigned npix, int flag, uint8_t *__restrict__ out) {
    const size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= groups) return;
    const size_t p = 4 * g, img = p / npix, q = p - img * npix;
    float4 v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = load4(in + (img * 3 + c) * npix, q);
    uint8_t b[12];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float x = k == 0 ? v[c].x : k == 1 ? v[c].y : k == 2 ? v[c].z : v[c].w;
            b[3 * k + c] = flag == SVBRDF_IMREAD_NORMAL ? to_u8((clip(x, -1.0f, 1.0f) + 1.0f) / 2.0f)
                                                        : to_u8(clip(x, 0.0f, 1.0f));
        }
    store_px4(out, g, b);
}

// x2 bilinear upsample of n_planes planes, as torch's upsample_bilinear2d with
// align_corners=False and scale_factor=2 computes it: source coordinate
// max(0.5 (d + 0.5) - 0.5, 0) in float, its integer part and fraction as the
// row / column and weight, the neighbour clamped at the last row / column,
// and out = h0 (w0 a + w1 b) + h1 (w0 c + w1 d).  One thread per output pixel
// walks the planes.
__device__ __forceinline__ void up2_axis(unsigned d, unsigned n_src, unsigned &i0, unsigned &i1, float &l0,
                                         float &l1) {
    const float r = fmaxf(0.5f * ((float)d + 0.5f) - 0.5f, 0.0f);
    i0 = (unsigned)r;
    i1 = i0 + (i0 < n_src - 1u ? 1u : 0u);
    l1 = r - (float)i0;
    l0 = 1.0f - l1;
}

__global__ __launch_bounds__(kBlock) void upsample2x_kernel(const float *__restrict__ src, unsigned res,
                                                            size_t src_stride, float *__restrict__ dst,
                                                            size_t dst_stride, int n_planes) {
    const unsigned side = 2u * res;
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= (size_t)side * side) return;
    const unsigned i = (unsigned)(p / side), j = (unsigned)(p - (size_t)i * side);
    unsigned h0, h1, w0, w1;
    float hl0, hl1, wl0, wl1;
    up2_axis(i, res, h0, h1, hl0, hl1);
    up2_axis(j, res, w0, w1, wl0, wl1);
    const size_t r0 = (size_t)h0 * res, r1 = (size_t)h1 * res;
    for (int c = 0; c < n_planes; ++c) {
        const float *s = src + (size_t)c * src_stride;
        const float a = s[r0 + w0], b = s[r0 + w1], e = s[r1 + w0], f = s[r1 + w1];
        dst[(size_t)c * dst_stride + p] = hl0 * (wl0 * a + wl1 * b) + hl1 * (wl0 * e + wl1 * f);
    }
}

// The same upsample for an even res with 16-byte aligned destination rows:
// one plane per thread (blockIdx.y), source columns 2q, 2q + 1 (output
// columns 4q .. 4q + 3) and a strip of kUpRows source rows (output rows 2i,
// 2i + 1 for each).  With align_corners=False and scale 2 the weights are 1/4
// and 3/4 everywhere except output row / column 0 (weights 1 and 0) and the
// neighbour index is clamped at the last row / column, so the strip needs
// source rows i0 - 1 .. i0 + kUpRows and columns 2q - 1 .. 2q + 2 (clamped):
// they are loaded once (a float2 and two scalars per row, all before the first
// store) and each row's horizontal interpolation serves the two output rows
// either side of it, so a thread issues (kUpRows + 2) x 3 loads for
// 2 kUpRows float4 stores.  Every output is the same expression, in the same
// order, as upsample2x_kernel evaluates it (h0 (w0 a + w1 b) + h1 (w0 c + w1 d),
// one rounding per operation), so the two kernels agree bit for bit.
// kUpRows = 4 measured best at 512^2 -> 1024^2 x 9 planes (9.6 us against
// 10.1 for 2 rows, 10.3 for 8, and 12.6 for one row pair per thread).
constexpr int kUpRows = 4;

__global__ __launch_bounds__(kBlock) void upsample2x_rows_kernel(const float *__restrict__ src, unsigned res,
                                                                 size_t src_stride, float *__restrict__ dst,
                                                                 size_t dst_stride) {
    const unsigned half = res / 2u;
    const unsigned strips = (res + kUpRows - 1u) / kUpRows;
    const unsigned t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= strips * half) return;
    const unsigned strip = t / half, q = t - strip * half;
    const unsigned i0 = strip * kUpRows, last = res - 1u;
    const unsigned c0 = q == 0u ? 0u : 2u * q - 1u, c1 = 2u * q, c3 = 2u * q + 2u > last ? last : 2u * q + 2u;
    const bool left = q == 0u;
    const float *s = src + (size_t)blockIdx.y * src_stride;
    // source rows i0 - 1 .. i0 + kUpRows, clamped to the image
    float win[kUpRows + 2][4];
#pragma unroll
    for (int r = 0; r < kUpRows + 2; ++r) {
        const int ri = (int)i0 - 1 + r;
        const unsigned row_i = ri < 0 ? 0u : ((unsigned)ri > last ? last : (unsigned)ri);
        const float *row = s + (size_t)row_i * res;
        const float2 mid = *reinterpret_cast<const float2 *>(row + c1);
        win[r][0] = row[c0];
        win[r][1] = mid.x;
        win[r][2] = mid.y;
        win[r][3] = row[c3];
    }
    float h[kUpRows + 2][4];
#pragma unroll
    for (int r = 0; r < kUpRows + 2; ++r) {
        const float w0 = win[r][0], w1 = win[r][1], w2 = win[r][2], w3 = win[r][3];
        h[r][0] = left ? 1.0f * w1 + 0.0f * w2 : 0.25f * w0 + 0.75f * w1;
        h[r][1] = 0.75f * w1 + 0.25f * w2;
        h[r][2] = 0.25f * w1 + 0.75f * w2;
        h[r][3] = 0.75f * w2 + 0.25f * w3;
    }
    float *d = dst + (size_t)blockIdx.y * dst_stride;
#pragma unroll
    for (int r = 0; r < kUpRows; ++r) {
        const unsigned i = i0 + (unsigned)r;
        if (i > last) break;
        const bool top = i == 0u;
        float a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a[k] = top ? 1.0f * h[r + 1][k] + 0.0f * h[r + 2][k] : 0.25f * h[r][k] + 0.75f * h[r + 1][k];
            b[k] = 0.75f * h[r + 1][k] + 0.25f * h[r + 2][k];
        }
        const size_t o = (size_t)(2u * i) * (2u * res) + 4u * q;
        // non-temporal: the output streams past the caches (8.2-8.5 against
        // 9.4-10.1 us at 512^2 -> 1024^2 x 9)
        const v4f va = {a[0], a[1], a[2], a[3]}, vb = {b[0], b[1], b[2], b[3]};
        __builtin_nontemporal_store(va, reinterpret_cast<v4f *>(d + o));
        __builtin_nontemporal_store(vb, reinterpret_cast<v4f *>(d + o + 2u * res));
    }
}

bool aligned(const void *p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

unsigned blocks_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int check_image_args(const char *what, const void *a, const void *b, int n, int height, int width, int channels,
                     int flag) {
    if (!a || !b) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (n < 1 || height < 1 || width < 1 || height > kMaxSide || width > kMaxSide)
        return fail(SVBRDF_ESHAPE, "%s: n=%d height=%d width=%d", what, n, height, width);
    if (flag < SVBRDF_IMREAD_NONE || flag > SVBRDF_IMREAD_NORMAL)
        return fail(SVBRDF_EINVAL, "%s: unknown flag %d", what, flag);
    if (channels < 1 || channels > 4) return fail(SVBRDF_EINVAL, "%s: channels=%d outside [1, 4]", what, channels);
    if ((flag == SVBRDF_IMREAD_SRGB || flag == SVBRDF_IMREAD_NORMAL) && channels != 3)
        return fail(SVBRDF_EINVAL, "%s: this flag needs a 3 channel image, got %d", what, channels);
    if (flag == SVBRDF_IMREAD_ROUGH && channels != 1 && channels != 3)
        return fail(SVBRDF_EINVAL, "%s: a roughness image has 1 or 3 channels, got %d", what, channels);
    return SVBRDF_OK;
}

}  // namespace

extern "C" {

int svbrdf_resize_lanczos4_u8(const unsigned char *src, int src_h, int src_w, int channels, unsigned char *dst,
                              int dst_h, int dst_w, void *stream) {
    const char *what = "svbrdf_resize_lanczos4_u8";
    if (!src || !dst) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (channels < 1 || channels > 4) return fail(SVBRDF_EINVAL, "%s: channels=%d outside [1, 4]", what, channels);
    if (src_h < 1 || src_w < 1 || dst_h < 1 || dst_w < 1 || src_h > kMaxSide || src_w > kMaxSide ||
        dst_h > kMaxSide || dst_w > kMaxSide)
        return fail(SVBRDF_ESHAPE, "%s: sizes %dx%d -> %dx%d outside [1, %d]", what, src_h, src_w, dst_h, dst_w,
                    kMaxSide);
    hipStream_t s = (hipStream_t)stream;
    if (src_h == dst_h && src_w == dst_w) {
        // cv2.resize returns the source unchanged at equal size
        if (hipMemcpyAsync(dst, src, (size_t)src_h * src_w * channels, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return fail(SVBRDF_EHIP, "%s: copy failed", what);
        return SVBRDF_OK;
    }
    const dim3 grid((dst_w + kTile - 1) / kTile, (dst_h + kTile - 1) / kTile);
    switch (channels) {
        case 1: hipLaunchKernelGGL(resize_lanczos4_kernel<1>, grid, dim3(kBlock), 0, s, src, src_h, src_w, dst, dst_h, dst_w); break;
        case 2: hipLaunchKernelGGL(resize_lanczos4_kernel<2>, grid, dim3(kBlock), 0, s, src, src_h, src_w, dst, dst_h, dst_w); break;
        case 3: hipLaunchKernelGGL(resize_lanczos4_kernel<3>, grid, dim3(kBlock), 0, s, src, src_h, src_w, dst, dst_h, dst_w); break;
        default: hipLaunchKernelGGL(resize_lanczos4_kernel<4>, grid, dim3(kBlock), 0, s, src, src_h, src_w, dst, dst_h, dst_w); break;
    }
    return check_launch(what);
}

int svbrdf_upsample2x_bilinear(const float *src, int res, size_t src_stride, float *dst, size_t dst_stride,
                               int n_planes, void *stream) {
    const char *what = "svbrdf_upsample2x_bilinear";
    if (!src || !dst) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (res < 1 || res > kMaxSide / 2) return fail(SVBRDF_ESHAPE, "%s: res=%d outside [1, %d]", what, res, kMaxSide / 2);
    if (n_planes < 1 || n_planes > 64) return fail(SVBRDF_EINVAL, "%s: n_planes=%d outside [1, 64]", what, n_planes);
    const size_t n_in = (size_t)res * res, n_out = 4 * n_in;
    if (src_stride == 0) src_stride = n_in;
    if (dst_stride == 0) dst_stride = n_out;
    if (src_stride < n_in || dst_stride < n_out)
        return fail(SVBRDF_ESHAPE, "%s: plane strides %zu / %zu below %zu / %zu", what, src_stride, dst_stride, n_in,
                    n_out);
    // the kernels read src while they write dst: the two ranges must be disjoint
    const uintptr_t s0 = (uintptr_t)src, s1 = s0 + ((n_planes - 1) * src_stride + n_in) * sizeof(float);
    const uintptr_t d0 = (uintptr_t)dst, d1 = d0 + ((n_planes - 1) * dst_stride + n_out) * sizeof(float);
    if (s0 < d1 && d0 < s1) return fail(SVBRDF_EINVAL, "%s: source and destination overlap", what);
    hipStream_t s = (hipStream_t)stream;
    if (res % 2 == 0 && dst_stride % 4 == 0 && aligned(dst, 16) && src_stride % 2 == 0 && aligned(src, 8)) {
        const unsigned threads = (unsigned)((res + kUpRows - 1) / kUpRows) * (unsigned)(res / 2);
        hipLaunchKernelGGL(upsample2x_rows_kernel, dim3(blocks_for(threads), (unsigned)n_planes), dim3(kBlock), 0, s,
                           src, (unsigned)res, src_stride, dst, dst_stride);
    } else {
        hipLaunchKernelGGL(upsample2x_kernel, dim3(blocks_for(n_out)), dim3(kBlock), 0, s, src, (unsigned)res,
                           src_stride, dst, dst_stride, n_planes);
    }
    return check_launch(what);
}

int svbrdf_images_from_u8(const unsigned char *images, int n, int height, int width, int channels, int flag,
                          float *out, void *stream) {
    const char *what = "svbrdf_images_from_u8";
    int rc = check_image_args(what, images, out, n, height, width, channels, flag);
    if (rc) return rc;
    const unsigned npix = (unsigned)height * (unsigned)width;
    const size_t total = (size_t)n * npix;
    if (channels == 3 && npix % 4 == 0 && aligned(images, 4) && aligned(out, 16))
        hipLaunchKernelGGL(images_from_u8_x4_kernel, dim3(blocks_for(total / 4)), dim3(kBlock), 0,
                           (hipStream_t)stream, images, total / 4, npix, flag, out);
    else
        hipLaunchKernelGGL(images_from_u8_kernel, dim3(blocks_for(total)), dim3(kBlock), 0, (hipStream_t)stream,
                           images, total, npix, channels, flag, out);
    return check_launch(what);
}

int svbrdf_textures_from_u8(const unsigned char *dif, const unsigned char *nom, const unsigned char *spe,
                            const unsigned char *rgh, int rgh_channels, int res, float *textures, void *stream) {
    const char *what = "svbrdf_textures_from_u8";
    if (!dif || !nom || !spe || !rgh || !textures) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (rgh_channels != 1 && rgh_channels != 3)
        return fail(SVBRDF_EINVAL, "%s: roughness map must have 1 or 3 channels, got %d", what, rgh_channels);
    if (res < 1 || res > kMaxRes) return fail(SVBRDF_ESHAPE, "%s: res=%d outside [1, %d]", what, res, kMaxRes);
    const unsigned npix = (unsigned)res * (unsigned)res;
    if (npix % 4 == 0 && aligned(dif, 4) && aligned(nom, 4) && aligned(spe, 4) && aligned(rgh, 4) &&
        aligned(textures, 16))
        hipLaunchKernelGGL(textures_from_u8_x4_kernel, dim3(blocks_for(npix / 4)), dim3(kBlock), 0,
                           (hipStream_t)stream, dif, nom, spe, rgh, rgh_channels, npix, textures);
    else
        hipLaunchKernelGGL(textures_from_u8_kernel, dim3(blocks_for(npix)), dim3(kBlock), 0, (hipStream_t)stream,
                           dif, nom, spe, rgh, rgh_channels, npix, textures);
    return check_launch(what);
}

int svbrdf_textures_to_u8(const float *textures, int res, unsigned char *dif, unsigned char *nom,
                          unsigned char *spe, unsigned char *rgh, void *stream) {
    const char *what = "svbrdf_textures_to_u8";
    if (!textures || !dif || !nom || !spe || !rgh) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (res < 1 || res > kMaxRes) return fail(SVBRDF_ESHAPE, "%s: res=%d outside [1, %d]", what, res, kMaxRes);
    const unsigned npix = (unsigned)res * (unsigned)res;
    if (npix % 4 == 0 && aligned(textures, 16) && aligned(dif, 4) && aligned(nom, 4) && aligned(spe, 4) &&
        aligned(rgh, 4))
        hipLaunchKernelGGL(textures_to_u8_x4_kernel, dim3(blocks_for(npix / 4)), dim3(kBlock), 0,
                           (hipStream_t)stream, textures, npix, dif, nom, spe, rgh);
    else
        hipLaunchKernelGGL(textures_to_u8_kernel, dim3(blocks_for(npix)), dim3(kBlock), 0, (hipStream_t)stream,
                           textures, npix, dif, nom, spe, rgh);
    return check_launch(what);
}

int svbrdf_images_to_u8(const float *images, int n, int height, int width, int channels, int flag,
                        unsigned char *out, void *stream) {
    const char *what = "svbrdf_images_to_u8";
    int rc = check_image_args(what, images, out, n, height, width, channels, flag);
    if (rc) return rc;
    if (flag == SVBRDF_IMREAD_ROUGH && channels != 1)
        return fail(SVBRDF_EINVAL, "%s: a roughness image is written with 1 channel", what);
    const unsigned npix = (unsigned)height * (unsigned)width;
    const size_t total = (size_t)n * npix;
    if (channels == 3 && npix % 4 == 0 && aligned(images, 16) && aligned(out, 4))
        hipLaunchKernelGGL(images_to_u8_x4_kernel, dim3(blocks_for(total / 4)), dim3(kBlock), 0,
                           (hipStream_t)stream, images, total / 4, npix, flag, out);
    else
        hipLaunchKernelGGL(images_to_u8_kernel, dim3(blocks_for(total)), dim3(kBlock), 0, (hipStream_t)stream,
                           images, total, npix, channels, flag, out);
    return check_launch(what);
}

}  // extern "C"
