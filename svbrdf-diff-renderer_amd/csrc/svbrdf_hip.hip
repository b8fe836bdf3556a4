// svbrdf_hip.hip -- gfx950 kernels and the extern "C" entry points declared in
// include/svbrdf_hip.h.
//
// Kernels (all elementwise over pixels, HBM-bound, no MFMA):
//   render_kernel        forward render                   36 B read + 12 V B written / pixel
//   backward_kernel      adjoint of the render            36 + 12 V B read, 36 B written / pixel
//   fused_step_kernel    clamp + render + loss + adjoint + Adam, in place:
//                        params, exp_avg, exp_avg_sq read and written once
//                        (216 B) + targets read once (12 V B) per pixel
//   prepare_targets_kernel, state_init_kernel, reduce_light_kernel: small helpers.
//
// Layout: planar NCHW as the reference keeps its tensors; thread i owns pixel
// i, so every plane access of a wavefront is one contiguous 256-byte run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "brdf.hpp"
#include "common.hpp"
#include "../../include/svbrdf_hip.h"

using namespace svbrdf;

namespace {

constexpr int kBlock = 256;
// The forward render and the adjoint launch one wave per block.  Their waves
// run long (one pixel per lane through every view), and a workgroup starts on
// a CU only when all of its waves fit there at once: with 256-thread blocks
// the slots freed by the first waves of a block to finish stay empty until
// enough others free up, while one-wave blocks refill each slot as it frees.
// Measured at 1024^2 x 9, bit-identical results: adjoint 38.5 -> 31.2 us,
// forward 26.5 -> 24.7 us.  The fused step keeps 256-thread blocks: its
// per-block arrival atomics and the next step's reduction of the per-block
// partials grow with the block count (64-thread blocks: 53 -> 88 us; 128: no
// change).
constexpr int kRenderBlock = 64;
constexpr int kAdjointBlock = 64;
constexpr int kMaxRes = 7680;   // 9 planes of res^2 floats (and every plane offset) below 2^31 bytes
constexpr int kMaxViews = 1024;

using svbrdf_internal::check_launch;
using svbrdf_internal::fail;

// Plane stride of the fused step's buffers: 0 selects res^2; otherwise at
// least res^2, and the 9 planes of one tensor must stay inside one 2 GiB
// buffer range (offsets are 32-bit).
int check_plane_stride(const char *what, int res, size_t plane_stride, unsigned &ps) {
    const size_t npix = (size_t)res * res;
    const size_t v = plane_stride ? plane_stride : npix;
    if (v < npix) return fail(SVBRDF_EINVAL, "%s: plane_stride=%zu < res^2=%zu", what, v, npix);
    if (v * 9u * 4u > 0x7fffffffu)
        return fail(SVBRDF_ESHAPE, "%s: plane_stride=%zu puts 9 planes beyond 2 GiB", what, v);
    ps = (unsigned)v;
    return SVBRDF_OK;
}

int check_common(const char *what, const void *tex, int res, int n_views, const void *cam,
                 const void *light, const void *pow_) {
    if (!tex || !cam || !light || !pow_) return fail(SVBRDF_EINVAL, "%s: null pointer argument", what);
    if (res < 1 || res > kMaxRes) return fail(SVBRDF_ESHAPE, "%s: res=%d outside [1, %d]", what, res, kMaxRes);
    if (n_views < 1 || n_views > kMaxViews)
        return fail(SVBRDF_ESHAPE, "%s: n_views=%d outside [1, %d]", what, n_views, kMaxViews);
    return SVBRDF_OK;
}

inline unsigned n_blocks(int res, int npix_per_thread = 1, int block = kBlock) {
    const size_t npix = (size_t)res * res;
    const size_t per_block = (size_t)block * npix_per_thread;
    return (unsigned)((npix + per_block - 1) / per_block);
}

// Geometry of view k, read with wave-uniform addresses (scalar loads).
// Camera / light positions and light power are read-only for every kernel
// that reads them and wave-uniform: read through the constant address space,
// so they come in by scalar loads even when the pointer itself was loaded
// from memory (the batched kernel's material table), where the generic
// pointer would otherwise give a vector flat load per view.
typedef const __attribute__((address_space(4))) float const_float;
__device__ __forceinline__ float ldu(const float *p, int i) { return ((const_float *)p)[i]; }

struct ViewPos {
    float cx, cy, cz, lx, ly, lz;
};

__device__ __forceinline__ ViewPos load_view(const float *__restrict__ cam, const float *__restrict__ light,
                                             int k) {
    ViewPos p;
    p.cx = ldu(cam, 3 * k + 0); p.cy = ldu(cam, 3 * k + 1); p.cz = ldu(cam, 3 * k + 2);
    p.lx = ldu(light, 3 * k + 0); p.ly = ldu(light, 3 * k + 1); p.lz = ldu(light, 3 * k + 2);
    return p;
}

// True when every view's light sits exactly at its camera (the flash setup of
// the reference's captures): then the shading uses the co-located forms.
// Each lane compares one coordinate and a wave vote combines them: one round
// trip of vector loads per 64 coordinates at the start of a block (a loop of
// scalar loads that stops at the first difference waits on 3 V dependent
// round trips there).  The whole wave must be active: every caller reaches it
// with all 64 lanes on (full waves, outside any divergent branch).
__device__ __forceinline__ bool all_colocated(const float *__restrict__ cam, const float *__restrict__ light,
                                              int n_views) {
    const int n = 3 * n_views, lane = (int)(threadIdx.x & 63u);
    bool differ = false;
    for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        // a NaN coordinate compares unequal: not co-located
        if (e < n) differ = differ || !(cam[e] == light[e]);
    }
    return __builtin_amdgcn_ballot_w64(differ) == 0;
}

// The pixel a thread owns.  Threads past the end keep working on the last
// pixel (so every load stays in bounds and no branch splits the wave) but
// store nothing and contribute nothing to the sums.
struct Pix {
    unsigned p;      // clamped pixel index
    unsigned off;    // its byte offset within a plane (npix <= 2^28, so it fits)
    bool valid;
    float px, py;    // position on the plane
};

// Plane access through buffer instructions: ONE resource descriptor per
// planar tensor (base, byte size of all its planes; four SGPRs, wave-uniform),
// the plane chosen by the instruction's scalar offset (c * plane bytes) and the
// pixel by the lane's 32-bit byte offset, which every plane shares.  One
// descriptor per tensor instead of one per plane keeps the step kernel's
// scalar registers within budget (27 per-plane descriptors spilled to vector
// lanes).  Word 3 = 0x00020000 selects the 32-bit data format that untyped
// gfx9 buffer accesses need.  All planes of a tensor must fit the 32-bit
// range: 9 * res^2 * 4 bytes < 2^31, hence kMaxRes = 7680 (offsets and sizes
// then fit a signed 32-bit value as well).
struct Tensor {
    __amdgpu_buffer_rsrc_t r;
    unsigned stride;   // bytes per plane
};
__device__ __forceinline__ Tensor tensor(const float *base, unsigned n_planes, unsigned npix) {
    Tensor t;
    t.stride = npix * 4u;
    t.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), 0, (int)(n_planes * t.stride), 0x00020000);
    return t;
}
// Planes `stride` floats apart (>= the pixel count): buffers the caller lays
// out with svbrdf_plane_stride() padding between planes.
__device__ __forceinline__ Tensor tensor_ps(const float *base, unsigned n_planes, unsigned stride) {
    Tensor t;
    t.stride = stride * 4u;
    t.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), 0, (int)(n_planes * t.stride), 0x00020000);
    return t;
}
__device__ __forceinline__ float ld(const Tensor &t, unsigned c, unsigned byte_off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(t.r, byte_off, (int)(c * t.stride), 0));
}
__device__ __forceinline__ void st(const Tensor &t, unsigned c, unsigned byte_off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), t.r, byte_off, (int)(c * t.stride), 0);
}

__device__ __forceinline__ Pix pixel_at(unsigned p, unsigned npix, int res, float size) {
    Pix x;
    x.valid = p < npix;
    x.p = x.valid ? p : npix - 1u;
    x.off = x.p * 4u;
    if ((res & (res - 1)) == 0) {
        // power-of-two R (a wave-uniform branch): the row is a shift and the
        // division by R an exponent change, both exact, so the coordinates
        // are the IEEE-division ones bit for bit
        const int k = __builtin_ctz((unsigned)res);
        const unsigned i = x.p >> k, j = x.p & (unsigned)(res - 1);
        x.px = (ldexpf((float)j + 0.5f, -k) - 0.5f) * size;
        x.py = -((ldexpf((float)i + 0.5f, -k) - 0.5f) * size);
        return x;
    }
    const unsigned i = x.p / (unsigned)res, j = x.p - i * (unsigned)res;
    x.px = pixel_coord((int)j, res, size);
    x.py = -pixel_coord((int)i, res, size);
    return x;
}

// Wave64 + block reduction of a float over BLOCK threads; result valid in
// thread 0 (a one-wave block needs no LDS round).
template <int BLOCK = kBlock>
__device__ __forceinline__ float block_sum(float x, float *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if constexpr (BLOCK == 64) return x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = x;
    __syncthreads();
    float r = 0.f;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) r += lds[w];
    }
    __syncthreads();
    return r;
}

// The same for a double over a 256-thread block.
__device__ __forceinline__ double block_sum_d(double x, double *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = x;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) r += lds[w];
    }
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Forward render: NPIX pixels per thread (block-strided, so each plane access
// of a wave stays one contiguous 256-byte run).
template <bool COLOC, int NPIX, int TERMS = 0>
__device__ __forceinline__ void render_pixels(const Pix (&px)[NPIX], const float *__restrict__ tex, unsigned npix,
                                              int n_views, const float *__restrict__ cam,
                                              const float *__restrict__ light, const float P[3],
                                              float *__restrict__ out) {
    Maps m[NPIX];
    const Tensor T = tensor(tex, 9, npix);
    // co-located: img_c = (P_c f1_c + DGq P_c F_c) (n.l)/d^2, the view-
    // independent products formed once per pixel
    float Pf1[NPIX][3], PF[NPIX][3];
#pragma unroll
    for (int q = 0; q < NPIX; ++q) {
        float t[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) t[c] = ld(T, c, px[q].off);
        decode(t, m[q]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Pf1[q][c] = P[c] * m[q].f1[c];
            PF[q][c] = P[c] * m[q].Fc[c];
        }
    }
    for (int k = 0; k < n_views; ++k) {
        const ViewPos vp = load_view(cam, light, k);
        const Tensor O = tensor(out + (size_t)(3 * k) * npix, 3, npix);
#pragma unroll
        for (int q = 0; q < NPIX; ++q) {
            float img[3];
            if constexpr (COLOC) {
                float DGq, nl_d2;
                shade_coloc_fwd(m[q], px[q].px, px[q].py, vp.cx, vp.cy, vp.cz, DGq, nl_d2);
#pragma unroll
                for (int c = 0; c < 3; ++c) img[c] = (Pf1[q][c] + DGq * PF[q][c]) * nl_d2;
            } else {
                Geom g;
                geometry<false>(px[q].px, px[q].py, vp.cx, vp.cy, vp.cz, vp.lx, vp.ly, vp.lz, g);
                Shade s;
                shade<false, TERMS>(m[q], g, P, s);
#pragma unroll
                for (int c = 0; c < 3; ++c) img[c] = s.img[c];
            }
            if (px[q].valid) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float ic = tclamp_med3(img[c], kEps, 1.0f);
                    st(O, c, px[q].off, fexp2(flog2(ic) * kInvGamma));
                }
            }
        }
    }
}

// TERMS != 0 (alternate BRDF terms) always takes the general path: the
// co-located specialisation hard-codes GGX and the spherical-Gaussian Fresnel.
template <int NPIX, int TERMS = 0>
__global__ __launch_bounds__(kRenderBlock) void render_kernel(
    const float *__restrict__ tex, int res, int n_views, float size, const float *__restrict__ cam,
    const float *__restrict__ light, const float *__restrict__ lpow, float *__restrict__ out) {
    const unsigned npix = (unsigned)res * (unsigned)res;
    Pix px[NPIX];
#pragma unroll
    for (int q = 0; q < NPIX; ++q)
        px[q] = pixel_at(blockIdx.x * (kRenderBlock * NPIX) + q * kRenderBlock + threadIdx.x, npix, res, size);
    const float P[3] = {lpow[0], lpow[1], lpow[2]};
    if (TERMS == 0 && all_colocated(cam, light, n_views))
        render_pixels<true, NPIX>(px, tex, npix, n_views, cam, light, P, out);
    else
        render_pixels<false, NPIX, TERMS>(px, tex, npix, n_views, cam, light, P, out);
}

// ---------------------------------------------------------------------------
// Adjoint of the render for an arbitrary dL/dimages.
template <bool COLOC, bool LIGHT_GRAD, int TERMS = 0>
__device__ __forceinline__ void backward_pixel(const Pix &px, const float *__restrict__ tex, unsigned npix,
                                               int n_views, const float *__restrict__ cam,
                                               const float *__restrict__ light, const float P[3],
                                               const float *__restrict__ gout, float *__restrict__ gtex,
                                               Accum &a) {
    float t[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) t[c] = ld(tensor(tex, 9, npix), c, px.off);
    Maps m;
    decode(t, m);
    // co-located: img_c = (P_c f1_c + DGq P_c F_c) (n.l)/d^2, the light power
    // folded into the view-independent factors.  The loop accumulates
    // dL/d(P_c f_c) without the constant 1/2.2 of d out/d img; P_c and 1/2.2
    // are applied once after the loop (every accumulator is linear in them).
    float Pf1[3], PF[3];
    if constexpr (COLOC) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Pf1[c] = P[c] * m.f1[c];
            PF[c] = P[c] * m.Fc[c];
        }
    }
    // dL/dimages of view k + 1 are loaded while view k is shaded.  The loads
    // are unconditional (no exec-mask branches around them): a lane past the
    // end reads the last pixel, which it owns as a copy and never stores, and
    // the last view re-reads its own values.
    auto load_gout = [&](int k, float (&g)[3]) {
        const Tensor G = tensor(gout + (size_t)(3 * k) * npix, 3, npix);
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] = ld(G, c, px.off);
    };
    float gnext[3];
    load_gout(0, gnext);
#pragma unroll 2
    for (int k = 0; k < n_views; ++k) {
        const ViewPos vp = load_view(cam, light, k);
        float go[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) go[c] = gnext[c];
        load_gout(k + 1 < n_views ? k + 1 : k, gnext);
        // d/dimg of clamp(img)^(1/2.2) = (1/2.2) out / img inside [eps, 1]
        auto dimg = [&](float img, float g) {
            const float ic = tclamp(img, kEps, 1.0f);
            const float o = fexp2(flog2(ic) * kInvGamma);
            return (img >= kEps && img <= 1.0f) ? g * kInvGamma * o * frcp(ic) : 0.0f;
        };
        if constexpr (COLOC) {
            ShadeC s;
            shade_coloc(m, px.px, px.py, vp.cx, vp.cy, vp.cz, s);
            // dL/d(P_c f_c) = gimg_c nl_d2 and dL/d(P_c F_c) = gimg_c nl_d2 DGq
            // (a.S, a.T), W = nl_d2 sum_c gimg_c P_c F_c and the direct
            // dL/d(n.l) = sum_c gimg_c P_c f_c / d^2
            const float ndg = s.nl_d2 * s.DGq;
            float U, gnl;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float pf = Pf1[c] + s.DGq * PF[c];
                const float img = pf * s.nl_d2;
                const float ic = fminf(fmaxf(img, kEps), 1.0f);
                // 2.2 d out/d img = ic^(1/2.2 - 1) inside the clamp, 0 outside
                // (a NaN img is outside, as torch's clamp backward masks it)
                const float gimg = ic == img ? go[c] * fexp2(flog2(ic) * (kInvGamma - 1.0f)) : 0.0f;
                a.S[c] += gimg * s.nl_d2;
                a.T[c] += gimg * ndg;
                U = c == 0 ? gimg * PF[c] : U + gimg * PF[c];
                gnl = c == 0 ? gimg * pf : gnl + gimg * pf;
                // dL/dP_c = sum gimg f_c nl_d2 = (sum gimg img_c) / P_c
                if (LIGHT_GRAD) a.gP[c] += gimg * img;
            }
            coloc_gradients(m, s, U * s.nl_d2, gnl * s.inv_d2, a);
        } else {
            Geom g;
            geometry<false>(px.px, px.py, vp.cx, vp.cy, vp.cz, vp.lx, vp.ly, vp.lz, g);
            Shade s;
            shade<false, TERMS>(m, g, P, s);
            float gimg[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) gimg[c] = dimg(s.img[c], go[c]);
            shade_backward<LIGHT_GRAD, TERMS>(m, g, s, P, gimg, a);
        }
    }
    if constexpr (COLOC) {
        coloc_finish(a);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a.S[c] *= kInvGamma * P[c];
            a.T[c] *= kInvGamma * P[c];
            a.gn[c] *= kInvGamma;
            // P_c = 0 renders img = 0 < eps everywhere: the clamp masks every
            // view, and the gradient is 0
            if (LIGHT_GRAD) a.gP[c] = P[c] != 0.0f ? a.gP[c] * kInvGamma / P[c] : 0.0f;
        }
        a.A *= kInvGamma;
    }
    float gt[9];
    decode_backward(t, m, a, gt);
    if (px.valid) {
#pragma unroll
        for (int c = 0; c < 9; ++c) st(tensor(gtex, 9, npix), c, px.off, gt[c]);
    }
}

#ifndef SVBRDF_BACKWARD_WAVES
#define SVBRDF_BACKWARD_WAVES 5      // minimum waves per SIMD for the adjoint (91-96 VGPRs; 8-28 B of spills in
                                     // some variants, none inside the view loops)
#endif
template <bool LIGHT_GRAD, int TERMS = 0>
__global__ __launch_bounds__(kAdjointBlock) __attribute__((amdgpu_waves_per_eu(SVBRDF_BACKWARD_WAVES))) void backward_kernel(
    const float *__restrict__ tex, int res, int n_views, float size, const float *__restrict__ cam,
    const float *__restrict__ light, const float *__restrict__ lpow, const float *__restrict__ gout,
    float *__restrict__ gtex, float4 *__restrict__ light_partials) {
    __shared__ float lds[kAdjointBlock / 64];
    const unsigned npix = (unsigned)res * (unsigned)res;
    const Pix px = pixel_at(blockIdx.x * kAdjointBlock + threadIdx.x, npix, res, size);
    const float P[3] = {lpow[0], lpow[1], lpow[2]};
    Accum a;
    accum_zero(a);
    if (TERMS == 0 && all_colocated(cam, light, n_views))
        backward_pixel<true, LIGHT_GRAD>(px, tex, npix, n_views, cam, light, P, gout, gtex, a);
    else
        backward_pixel<false, LIGHT_GRAD, TERMS>(px, tex, npix, n_views, cam, light, P, gout, gtex, a);
    if (LIGHT_GRAD) {
        float s[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)   // a lane past the end shaded a copy of the last pixel: no contribution
            s[c] = block_sum<kAdjointBlock>(px.valid ? a.gP[c] : 0.0f, lds);
        if (threadIdx.x == 0) light_partials[blockIdx.x] = make_float4(s[0], s[1], s[2], 0.0f);
    }
}

// Sums the adjoint's per-block light-power partials (float4: dL/dP rgb, 0) in
// a fixed order (deterministic): thread i adds partials i, i + 1024, ... in
// double, four loads in flight, then the waves combine in index order.  1,024
// threads take the 16,384 partials of a 1024^2 adjoint in four rounds.
constexpr int kLightReduceBlock = 1024;
__global__ __launch_bounds__(kLightReduceBlock) void reduce_light_kernel(const float4 *__restrict__ partials,
                                                                         unsigned n, float *__restrict__ out) {
    __shared__ double lds[3][kLightReduceBlock / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    for (unsigned b0 = threadIdx.x; b0 < n; b0 += 4 * kLightReduceBlock) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned b = b0 + u * kLightReduceBlock;
            q[u] = b < n ? partials[b] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[0] += q[u].x;
            acc[1] += q[u].y;
            acc[2] += q[u].z;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double x = acc[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) lds[c][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double t = 0.0;
            for (int w = 0; w < kLightReduceBlock / 64; ++w) t += lds[c][w];
            out[c] = (float)t;
        }
    }
}

// ---------------------------------------------------------------------------
// Targets: ln(clamp(t, floor, 1)) for the log-space loss, or a copy.
__global__ __launch_bounds__(kBlock) void prepare_targets_kernel(const float *__restrict__ in, size_t n,
                                                                 int log_mode, float floor_v,
                                                                 float *__restrict__ out) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
        const float x = in[e];
        out[e] = log_mode ? flog2(tclamp(x, floor_v, 1.0f)) * kLn2 : x;
    }
}

// Bias corrections of step t.  torch computes 1 - beta^t in double with pow
// from the float step count (torch/optim/adam.py:519-524); here beta^t is
// carried in double and multiplied once per step, which gives the same
// fp32-rounded step size -lr / (1 - beta1^t) for every t tested (2e5 steps).
// The kernel reads 1/(1 - beta1^t) in double and 1/sqrt(1 - beta2^t) in fp32.
__device__ void set_bias_corrections(svbrdf_step_state *st) {
    st->inv_bias_correction1 = 1.0 / (1.0 - st->beta1_pow);
    st->inv_bias_correction2_sqrt = (float)(1.0 / sqrt(1.0 - st->beta2_pow));
}

__global__ void state_init_kernel(svbrdf_step_state *st, double beta1, double beta2) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->beta1_pow = beta1;
    st->beta2_pow = beta2;
    set_bias_corrections(st);
    st->step = 0.0f;
    st->loss = 0.0f;
    st->pending_step = 0u;
    st->reserved[0] = st->reserved[1] = 0u;
    for (int c = 0; c < 4; ++c) {
        st->light_exp_avg[c] = 0.0f;
        st->light_exp_avg_sq[c] = 0.0f;
        st->light_grad[c] = 0.0f;
    }
}

// One Adam update of a single fp32 element in torch's single-tensor order
// (torch/optim/adam.py:431-547): lerp, mul+addcmul, sqrt/bc2_sqrt + eps,
// addcdiv with value -lr/bc1.  Three operations use the hardware's
// approximate forms instead of the correctly rounded ones torch calls:
// v_sqrt_f32 (within 1 ulp) for sqrt, a multiply by the fp32-rounded
// 1/sqrt(bc2) for the division by bc2_sqrt, and v_rcp_f32 (within 1 ulp) for
// addcdiv's division.  The step m/denom then differs from torch's by at most
// ~4 ulp relative (about 2.4e-9 absolute at lr 0.01, where Adam's step is
// ~lr), below half an ulp of a parameter of magnitude >= 0.08, so the stored
// parameter differs by at most one ulp where it differs at all.  The IEEE
// forms cost ~25 instructions each (18 per pixel), about 15 % more VALU
// issue for the step.
__device__ __forceinline__ void adam_update(float &p, float &m, float &v, float g, float w1, float beta2,
                                            float om_beta2, float inv_bc2_sqrt, float eps, float neg_step) {
    m = m + w1 * (g - m);
    v = v * beta2 + om_beta2 * g * g;
    const float denom = fsqrt(v) * inv_bc2_sqrt + eps;
    p = p + neg_step * m * frcp(denom);
}

struct StepScalars {
    float w1, beta2, om_beta2, inv_bc2_sqrt, eps, neg_step;
};

// Launch constants of a step, computed once on the host in double (as
// torch's Python floats are) and rounded once to fp32.  In the kernel every
// wave would redo these double divisions and conversions: on the VALU, at
// the double rate.
struct StepConsts {
    double lr, beta1, beta2, eps;   // as given (state advance, light-power Adam)
    double count;                   // image-loss elements, V * 3 * R^2
    float K2;                       // d mean / d element times 1/2.2: (2 / count) / 2.2
    float w1, beta2_f, om_beta2, eps_f;
};

StepConsts make_consts(double lr, double beta1, double beta2, double eps, double count) {
    StepConsts c;
    c.lr = lr; c.beta1 = beta1; c.beta2 = beta2; c.eps = eps; c.count = count;
    c.K2 = count > 0.0 ? (float)(2.0 / count / 2.2) : 0.0f;
    c.w1 = (float)(1.0 - beta1);
    c.beta2_f = (float)beta2;
    c.om_beta2 = (float)(1.0 - beta2);
    c.eps_f = (float)eps;
    return c;
}

// K2 of one material of a batch (host side, stored in its table entry).
float loss_scale(int res, int n_views) {
    return (float)(2.0 / ((double)n_views * 3.0 * (double)res * (double)res) / 2.2);
}

__device__ __forceinline__ StepScalars step_scalars(const svbrdf_step_state *st, const StepConsts &c) {
    StepScalars s;
    s.w1 = c.w1;
    s.beta2 = c.beta2_f;
    s.om_beta2 = c.om_beta2;
    s.inv_bc2_sqrt = st->inv_bias_correction2_sqrt;
    s.eps = c.eps_f;
    s.neg_step = (float)(-c.lr * st->inv_bias_correction1);
    return s;
}

// ---------------------------------------------------------------------------
// Fused per-pixel optimisation step.  Every block writes its partial loss (and
// light-power gradient); the last block to finish reduces them in a fixed
// order, updates light_pow when it is optimised, and advances the step state.
#ifndef SVBRDF_FUSED_NPIX
#define SVBRDF_FUSED_NPIX 1       // pixels per thread
#endif
#ifndef SVBRDF_FUSED_UNROLL
#define SVBRDF_FUSED_UNROLL 1     // views per loop trip
#endif
#ifndef SVBRDF_FUSED_WAVES
#define SVBRDF_FUSED_WAVES 5      // minimum waves per SIMD the register allocator must allow
#endif
#ifndef SVBRDF_MOMENT_PREFETCH
// When the Adam moments are loaded: 0 = per channel next to its stores (the
// compiler cannot reorder a load past a store through another descriptor, so
// the nine round trips serialise), 1 = all 18 after the view loop and before
// the first store (one round trip; measured 69 vs 76 us per step at
// 1024^2 x 9), 2 = all 18 before the view loop (18 more VGPRs through the
// loop: 126, slower).
#define SVBRDF_MOMENT_PREFETCH 1
#endif
#ifndef SVBRDF_TARGET_PREFETCH
#define SVBRDF_TARGET_PREFETCH 0  // 1: load view k+1's targets while view k is shaded (no gain measured)
#endif

// Loss residual and dL/dimg of one channel (general geometry): r = ln out -
// ln target' (log-L2) or out - target (L2), with torch's inclusive clamp mask
// on img.  K2 = (2 / element count) / 2.2.
template <int LOSS>
__device__ __forceinline__ float loss_channel(float img, float y, float K2, float &loss) {
    const float ic = tclamp(img, kEps, 1.0f);
    const float lic = flog2(ic);
    const bool inside = (img >= kEps && img <= 1.0f);
    float r, gi;
    if (LOSS == SVBRDF_LOSS_LOG_L2) {
        r = lic * (kInvGamma * kLn2) - y;
        gi = K2 * r * frcp(ic);
    } else {
        const float o = fexp2(lic * kInvGamma);
        r = o - y;
        gi = K2 * r * o * frcp(ic);
    }
    loss += r * r;
    return inside ? gi : 0.0f;
}

// Co-located views.  Inside the clamp range (ic = img = P f nl/d2) the chain
// dL/dimg -> f, nl, P collapses: with rm = r (log-L2) or r out (L2),
//   dL/df_c = K2 rm / f_c,  dL/dnl = K2 sum_c rm_c / nl,  dL/dP_c = K2 rm / P_c.
// Outside the range rm = 0 (torch's clamp mask); f and nl are floored at
// 1e-30 so that a masked zero never meets an infinite reciprocal.  The
// accumulators are linear in dL/df and dL/dnl, so the loop accumulates them
// without K2 and scales them once per pixel.
template <int LOSS, bool OPT_LIGHT, int UNROLL>
__device__ __forceinline__ void fused_views_coloc(const Pix &px, const Maps &m, const float *__restrict__ tgt,
                                                  unsigned ps, int n_views, const float *__restrict__ cam,
                                                  const float P[3], float K2, float &loss, float (&rP)[3],
                                                  Accum &a) {
    float yn[3];
    if (SVBRDF_TARGET_PREFETCH) {
#pragma unroll
        for (int c = 0; c < 3; ++c) yn[c] = ld(tensor_ps(tgt, 3, ps), c, px.off);
    }
    // P_c f_c = P_c f1_c + DGq P_c F_c: the light power folded into the
    // view-independent factors.  The loop then accumulates dL/d(P_c f_c)
    // (= dL/df_c / P_c) and P_c is applied once with K2 after the loop.
    float Pf1[3], PF[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Pf1[c] = P[c] * m.f1[c];
        PF[c] = P[c] * m.Fc[c];
    }
#pragma unroll UNROLL
    for (int k = 0; k < n_views; ++k) {
        const float cx = ldu(cam, 3 * k + 0), cy = ldu(cam, 3 * k + 1), cz = ldu(cam, 3 * k + 2);
        float y[3];
        if (SVBRDF_TARGET_PREFETCH) {
            // issue view k+1's loads now (the last view re-reads its own, an L1/L2 hit)
            const int kn = k + 1 < n_views ? k + 1 : k;
            const Tensor TN = tensor_ps(tgt + (size_t)(3 * kn) * ps, 3, ps);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                y[c] = yn[c];
                yn[c] = ld(TN, c, px.off);
            }
        } else {
            const Tensor TK = tensor_ps(tgt + (size_t)(3 * k) * ps, 3, ps);
#pragma unroll
            for (int c = 0; c < 3; ++c) y[c] = ld(TK, c, px.off);
        }
        ShadeC s;
        shade_coloc(m, px.px, px.py, cx, cy, cz, s);
        float gf[3], rsum;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float pf = Pf1[c] + s.DGq * PF[c];
            const float img = pf * s.nl_d2;
            const float ic = fminf(fmaxf(img, kEps), 1.0f);
            const bool inside = (ic == img);
            const float lic = flog2(ic);
            float r, rm;
            if (LOSS == SVBRDF_LOSS_LOG_L2) {
                r = lic * (kInvGamma * kLn2) - y[c];
                rm = inside ? r : 0.0f;
            } else {
                const float o = fexp2(lic * kInvGamma);
                r = o - y[c];
                rm = inside ? r * o : 0.0f;
            }
            loss += r * r;
            // dL/d(P_c f_c) / K2: every accumulator is linear in these, so
            // K2 (and P_c) is applied once per pixel after the loop
            gf[c] = rm * frcp(fmaxf(pf, 1e-30f));
            rsum = c == 0 ? rm : rsum + rm;   // no "0 + x": it is not x for x = -0
            if (OPT_LIGHT) rP[c] += rm;
        }
        const float rnv = frcp(fmaxf(s.nv, 1e-30f));
        backward_coloc(m, s, gf, rsum * rnv, PF, a);
    }
    coloc_finish(a);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.S[c] *= K2 * P[c];
        a.T[c] *= K2 * P[c];
        a.gn[c] *= K2;
    }
    a.A *= K2;
}

template <int LOSS, bool OPT_LIGHT, bool COLOC, int NPIX, int UNROLL>
__device__ __forceinline__ void fused_pixels(const Pix (&px)[NPIX], float *__restrict__ tex,
                                             float *__restrict__ exp_avg, float *__restrict__ exp_avg_sq,
                                             const float *__restrict__ tgt, unsigned tps, unsigned ps,
                                             int n_views, const float *__restrict__ cam,
                                             const float *__restrict__ light, const float P[3], float K2,
                                             const StepScalars &sc, float &loss, float (&gP)[3]) {
    const Tensor TX = tensor_ps(tex, 9, tps), TM = tensor_ps(exp_avg, 9, ps), TV = tensor_ps(exp_avg_sq, 9, ps);
    float traw[NPIX][9];
    float mm[NPIX][9], vv[NPIX][9];
    Maps m[NPIX];
    Accum a[NPIX];
    float lq[NPIX];
    float rP[NPIX][3];
#pragma unroll
    for (int q = 0; q < NPIX; ++q) {
#pragma unroll
        for (int c = 0; c < 9; ++c) traw[q][c] = ld(TX, c, px[q].off);
        if (SVBRDF_MOMENT_PREFETCH == 2) {
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                mm[q][c] = ld(TM, c, px[q].off);
                vv[q][c] = ld(TV, c, px[q].off);
            }
        }
        float t[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) t[c] = tclamp(traw[q][c], -1.0f, 1.0f);   // textures.clamp(-1, 1), svbrdf.py:60
        decode(t, m[q]);
        accum_zero(a[q]);
        lq[q] = 0.f;
        rP[q][0] = rP[q][1] = rP[q][2] = 0.f;
    }
    if constexpr (COLOC) {
#pragma unroll
        for (int q = 0; q < NPIX; ++q)
            fused_views_coloc<LOSS, OPT_LIGHT, UNROLL>(px[q], m[q], tgt, ps, n_views, cam, P, K2, lq[q], rP[q],
                                                       a[q]);
    } else {
        float yn[NPIX][3];
        if (SVBRDF_TARGET_PREFETCH) {
#pragma unroll
            for (int q = 0; q < NPIX; ++q)
#pragma unroll
                for (int c = 0; c < 3; ++c) yn[q][c] = ld(tensor_ps(tgt, 3, ps), c, px[q].off);
        }
#pragma unroll UNROLL
        for (int k = 0; k < n_views; ++k) {
            const ViewPos vp = load_view(cam, light, k);
            const Tensor TK = tensor_ps(tgt + (size_t)(3 * k) * ps, 3, ps);
            const Tensor TN = tensor_ps(tgt + (size_t)(3 * (k + 1 < n_views ? k + 1 : k)) * ps, 3, ps);
#pragma unroll
            for (int q = 0; q < NPIX; ++q) {
                float y[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (SVBRDF_TARGET_PREFETCH) {
                        y[c] = yn[q][c];
                        yn[q][c] = ld(TN, c, px[q].off);
                    } else {
                        y[c] = ld(TK, c, px[q].off);
                    }
                }
                Geom g;
                geometry<false>(px[q].px, px[q].py, vp.cx, vp.cy, vp.cz, vp.lx, vp.ly, vp.lz, g);
                Shade s;
                shade<false>(m[q], g, P, s);
                float gimg[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) gimg[c] = loss_channel<LOSS>(s.img[c], y[c], K2, lq[q]);
                shade_backward<OPT_LIGHT>(m[q], g, s, P, gimg, a[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NPIX; ++q) {
        if (!px[q].valid) continue;
        // parameters or a light power that are NaN make the loss NaN, as in
        // the reference
        bool finite = P[0] == P[0] && P[1] == P[1] && P[2] == P[2];
#pragma unroll
        for (int c = 0; c < 9; ++c) finite = finite && (traw[q][c] == traw[q][c]);
        loss += finite ? lq[q] : __builtin_nanf("");
        if (OPT_LIGHT) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                // dL/dP_c = K2 sum rm / P_c (co-located) or the accumulated sum (general)
                gP[c] += COLOC ? (P[c] != 0.0f ? K2 * rP[q][c] / P[c] : 0.0f) : a[q].gP[c];
            }
        }
        float t[9], gt[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) t[c] = tclamp(traw[q][c], -1.0f, 1.0f);
        decode_backward(t, m[q], a[q], gt);
        const unsigned o = px[q].off;
        if (SVBRDF_MOMENT_PREFETCH == 1) {
            // every load ahead of the first store: one memory latency, not nine
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                mm[q][c] = ld(TM, c, o);
                vv[q][c] = ld(TV, c, o);
            }
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            float pp = traw[q][c], m1, v1;
            if (SVBRDF_MOMENT_PREFETCH == 0) {
                m1 = ld(TM, c, o);
                v1 = ld(TV, c, o);
            } else {
                m1 = mm[q][c];
                v1 = vv[q][c];
            }
            // the clamp on read passes the gradient only inside [-1, 1] (inclusive)
            const float g = (traw[q][c] >= -1.0f && traw[q][c] <= 1.0f) ? gt[c] : 0.0f;
            adam_update(pp, m1, v1, g, sc.w1, sc.beta2, sc.om_beta2, sc.inv_bc2_sqrt, sc.eps, sc.neg_step);
            st(TX, c, o, pp);
            st(TM, c, o, m1);
            st(TV, c, o, v1);
        }
    }
}

// Workspace of the fused step: arrival counters, then two buffers of per-block
// partial sums (float4: loss, dL/dP rgb), used by alternate steps.  Step t
// writes buffer t & 1 and the reduction of step t's partials happens in a
// later launch, where they are visible without any cross-block release inside
// a launch (an agent-scope release per block writes back the whole L2 of the
// XCD on gfx950).
//
// Arrival is counted in two levels so that no address takes more than
// ceil(blocks / kArrivalGroups) atomics: block b counts itself in group
// b % kArrivalGroups; the block completing a group counts the group in the top
// counter.  Each counter has its own 128-byte line.
constexpr unsigned kArrivalGroups = 64;
constexpr unsigned kCounterStride = 128 / sizeof(unsigned);
constexpr size_t kCounterBytes = ((kArrivalGroups + 1) * 128 + 255) / 256 * 256;
inline size_t partial_buffer_bytes(unsigned nb) { return ((size_t)nb * sizeof(float4) + 255) / 256 * 256; }
inline size_t fused_workspace_bytes(unsigned nb) { return kCounterBytes + 2 * partial_buffer_bytes(nb); }
__device__ __forceinline__ float4 *partial_buffer(void *ws, unsigned nb, unsigned parity) {
    const size_t stride = ((size_t)nb * sizeof(float4) + 255) / 256 * 256;
    return reinterpret_cast<float4 *>(static_cast<char *>(ws) + kCounterBytes + parity * stride);
}

// Fixed-order sum of n float4 partials (member .x, and .y .z .w when NQ == 4)
// by one block of kBlock threads: thread i adds i, i + kBlock, ... in double,
// then the wave and the block combine in index order.  Every reader of a
// step's partials (the next step's block 0, the flush, the light finalize)
// uses this routine, so they agree to the bit.  Result valid in thread 0.
template <int NQ, int BLOCK = kBlock>
__device__ void reduce_partials(const float4 *__restrict__ parts, unsigned n, double (&tot)[NQ],
                                double (*lds)[BLOCK / 64]) {
    double acc[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) acc[c] = 0.0;
    for (unsigned b0 = threadIdx.x; b0 < n; b0 += 4 * BLOCK) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned b = b0 + u * BLOCK;
            q[u] = b < n ? parts[b] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[0] += q[u].x;
            if constexpr (NQ == 4) { acc[1] += q[u].y; acc[2] += q[u].z; acc[3] += q[u].w; }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        double x = acc[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) lds[c][wave] = x;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        tot[c] = 0.0;
        if (threadIdx.x == 0)
            for (int w = 0; w < BLOCK / 64; ++w) tot[c] += lds[c][w];
    }
    __syncthreads();
}

// Records the loss of step t (1-based) from its partials: state->loss and
// loss_history[t - 1].  Thread 0 only.
__device__ __forceinline__ void record_loss(svbrdf_step_state *st, unsigned t, double total, double count,
                                            float *loss_history, int history_len) {
    const float L = (float)(total / count);
    st->loss = L;
    if (loss_history && t >= 1 && t <= (unsigned)max(history_len, 0)) loss_history[t - 1] = L;
}

// Step count, beta powers and bias corrections of the next step.
__device__ __forceinline__ void advance_step(svbrdf_step_state *st, float step, double beta1, double beta2) {
    st->step = step;
    st->beta1_pow *= beta1;
    st->beta2_pow *= beta2;
    set_bias_corrections(st);
}

// One fused iteration over all pixels.
//  * Every block reads the step state once (Adam scalars, completed steps t),
//    writes its partial sums to buffer t & 1 and then, from thread 0, counts
//    itself in the arrival counter.
//  * Block 0 first reduces the partials of the previous step when they are
//    pending (state->pending_step == t) and records that step's loss.
//  * Without light optimisation the block that arrives last (all blocks have
//    read the state by then) advances the state for the next launch, marks
//    step t + 1 pending and re-arms the counter: one launch per iteration.
//  * With light optimisation the light power of the next step depends on the
//    full dL/dP of this one, so a one-block finalize follows instead.
template <int LOSS, bool OPT_LIGHT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SVBRDF_FUSED_WAVES))) void fused_step_kernel(
    float *__restrict__ tex, float *__restrict__ exp_avg, float *__restrict__ exp_avg_sq,
    const float *__restrict__ tgt, int res, unsigned tps, unsigned ps, int n_views, float size,
    const float *__restrict__ cam,
    const float *__restrict__ light, float *__restrict__ lpow, StepConsts c, svbrdf_step_state *st, void *ws,
    float *loss_history, int history_len) {
    constexpr int NPIX = SVBRDF_FUSED_NPIX;
    __shared__ float lds[kBlock / 64];
    const unsigned npix = (unsigned)res * (unsigned)res;
    const double count = c.count;
    // d mean / d element, times d out^... / d ln img = 1/2.2 (folded once)
    const float K2 = c.K2;
    const StepScalars sc = step_scalars(st, c);
    const float step_f = st->step;
    const unsigned t = (unsigned)step_f;
    if (blockIdx.x == 0 && t >= 1 && st->pending_step == t) {
        __shared__ double ldsd[1][kBlock / 64];
        double tot[1];
        reduce_partials<1>(partial_buffer(ws, gridDim.x, (t - 1) & 1u), gridDim.x, tot, ldsd);
        if (threadIdx.x == 0) record_loss(st, t, tot[0], count, loss_history, history_len);
    }
    const float P[3] = {lpow[0], lpow[1], lpow[2]};
    Pix px[NPIX];
#pragma unroll
    for (int q = 0; q < NPIX; ++q)
        px[q] = pixel_at(blockIdx.x * (kBlock * NPIX) + q * kBlock + threadIdx.x, npix, res, size);
    float loss = 0.f;
    float gPs[3] = {0.f, 0.f, 0.f};
    if (all_colocated(cam, light, n_views))
        fused_pixels<LOSS, OPT_LIGHT, true, NPIX, SVBRDF_FUSED_UNROLL>(px, tex, exp_avg, exp_avg_sq, tgt, tps, ps,
                                                                        n_views, cam, light, P, K2, sc, loss, gPs);
    else
        fused_pixels<LOSS, OPT_LIGHT, false, NPIX, SVBRDF_FUSED_UNROLL>(px, tex, exp_avg, exp_avg_sq, tgt, tps, ps,
                                                                         n_views, cam, light, P, K2, sc, loss, gPs);
    const float bl = block_sum(loss, lds);
    float bp[3] = {0.f, 0.f, 0.f};
    if (OPT_LIGHT) {
#pragma unroll
        for (int c = 0; c < 3; ++c) bp[c] = block_sum(gPs[c], lds);
    }
    if (threadIdx.x != 0) return;
    partial_buffer(ws, gridDim.x, t & 1u)[blockIdx.x] = make_float4(bl, bp[0], bp[1], bp[2]);
    if constexpr (!OPT_LIGHT) {
        unsigned *counters = static_cast<unsigned *>(ws);
        const unsigned nb = gridDim.x, g = blockIdx.x % kArrivalGroups;
        const unsigned group_size = (nb - g + kArrivalGroups - 1) / kArrivalGroups;
        const unsigned n_groups = nb < kArrivalGroups ? nb : kArrivalGroups;
        unsigned *group = counters + (1 + g) * kCounterStride;
        if (atomicAdd(group, 1u) != group_size - 1) return;
        __hip_atomic_store(group, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (atomicAdd(counters, 1u) != n_groups - 1) return;
        // every block of this launch has read the state
        __hip_atomic_store(counters, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        advance_step(st, step_f + 1.0f, c.beta1, c.beta2);
        st->pending_step = t + 1;
    }
}

// One block of kFinalizeBlock threads, after a light-optimising step:
// reduces that step's partials in the fixed order of reduce_partials (at this
// block size: the light path's only reader), records the loss, takes the
// light-power Adam step (one param group with the maps, src/svbrdf.py:50)
// and advances the state.  The next step needs the new light power in every
// block, so this sits between two step launches; 1,024 threads take the
// 4,096 partials of a 1024^2 step in one round of loads.
constexpr int kFinalizeBlock = 1024;
template <bool OPT_LIGHT>
__global__ __launch_bounds__(kFinalizeBlock) void finalize_step_kernel(void *ws, unsigned n_partials, StepConsts k,
                                                                      float *__restrict__ lpow,
                                                                      svbrdf_step_state *st, float *loss_history,
                                                                      int history_len) {
    constexpr int NQ = OPT_LIGHT ? 4 : 1;
    __shared__ double lds[NQ][kFinalizeBlock / 64];
    const float step_f = st->step;
    const unsigned t = (unsigned)step_f;
    double tot[NQ];
    reduce_partials<NQ, kFinalizeBlock>(partial_buffer(ws, n_partials, t & 1u), n_partials, tot, lds);
    if (threadIdx.x != 0) return;
    record_loss(st, t + 1, tot[0], k.count, loss_history, history_len);
    if constexpr (OPT_LIGHT) {
        // Adam on light_pow with the same scalars as the maps
        const StepScalars sc = step_scalars(st, k);
        for (int c = 0; c < 3; ++c) {
            float pp = lpow[c], mm = st->light_exp_avg[c], vv = st->light_exp_avg_sq[c];
            const float g = (float)tot[1 + c];
            adam_update(pp, mm, vv, g, sc.w1, sc.beta2, sc.om_beta2, sc.inv_bc2_sqrt, sc.eps, sc.neg_step);
            lpow[c] = pp;
            st->light_exp_avg[c] = mm;
            st->light_exp_avg_sq[c] = vv;
            st->light_grad[c] = g;
        }
    }
    advance_step(st, step_f + 1.0f, k.beta1, k.beta2);
    st->pending_step = 0u;
}

// One block: records the loss of the last completed step when its partials
// are still pending, and clears the mark.
__global__ __launch_bounds__(kBlock) void flush_step_kernel(const void *ws, unsigned n_partials, double count,
                                                           svbrdf_step_state *st, float *loss_history,
                                                           int history_len) {
    __shared__ double lds[1][kBlock / 64];
    const unsigned t = (unsigned)st->step;
    if (t == 0 || st->pending_step != t) return;
    double tot[1];
    reduce_partials<1>(partial_buffer(const_cast<void *>(ws), n_partials, (t - 1) & 1u), n_partials, tot, lds);
    if (threadIdx.x != 0) return;
    record_loss(st, t, tot[0], count, loss_history, history_len);
    st->pending_step = 0u;
}

// ---------------------------------------------------------------------------
// Batched persistent step.  A grid sized to the device pulls tiles from an
// atomic ticket counter; the counter is the only state blocks share, so no
// fence is needed, and every block leaves once its ticket passes n_tiles.
#ifndef SVBRDF_BATCH_WAVES
#define SVBRDF_BATCH_WAVES 4      // minimum waves per SIMD for the batched kernel's register allocation (128 VGPRs, no spills at one pixel per thread)
#endif
template <int LOSS, int NPIX, int UNROLL, bool OPT_LIGHT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NPIX == 1 ? SVBRDF_BATCH_WAVES : 1))) void batch_step_kernel(
    const svbrdf_material *__restrict__ mats, const unsigned *__restrict__ order,
    const unsigned *__restrict__ tile_mat, unsigned n_tiles, int tile_pixels, unsigned *__restrict__ queue,
    float4 *__restrict__ partials, StepConsts c, const svbrdf_step_state *__restrict__ st) {
    __shared__ float lds[kBlock / 64];
    __shared__ unsigned s_slot;
    const StepScalars sc = step_scalars(st, c);
    // The ticket is read back through readfirstlane so that the loop exit is a
    // scalar (wave-uniform) branch: with a divergent exit the control-flow
    // structurizer may keep lanes re-running a tile while lane 0 waits to draw
    // the next ticket.  The trip count is also bounded by n_tiles + 1.
    if (threadIdx.x == 0) s_slot = atomicAdd(queue, 1u);
    __syncthreads();
    unsigned slot = __builtin_amdgcn_readfirstlane(s_slot);
    for (unsigned trip = 0; slot < n_tiles && trip <= n_tiles; ++trip) {
        const unsigned id = order[slot];
        const svbrdf_material mt = mats[tile_mat[id]];
        const unsigned npix = (unsigned)mt.res * (unsigned)mt.res;
        const unsigned first = (id - mt.tile_begin) * (unsigned)tile_pixels;
        const unsigned ps = mt.plane_stride ? mt.plane_stride : npix;
        const unsigned tps = mt.texture_stride ? mt.texture_stride : npix;
        const float K2 = mt.loss_scale;
        const float P[3] = {ldu(mt.light_pow, 0), ldu(mt.light_pow, 1), ldu(mt.light_pow, 2)};
        const bool coloc = all_colocated(mt.camera_pos, mt.light_pos, mt.n_views);
        float loss = 0.f;
        float gPs[3] = {0.f, 0.f, 0.f};
        for (unsigned base = 0; base < (unsigned)tile_pixels && first + base < npix; base += kBlock * NPIX) {
            Pix px[NPIX];
#pragma unroll
            for (int q = 0; q < NPIX; ++q)
                px[q] = pixel_at(first + base + q * kBlock + threadIdx.x, npix, mt.res, mt.size);
            if (coloc)
                fused_pixels<LOSS, OPT_LIGHT, true, NPIX, UNROLL>(px, mt.textures, mt.exp_avg, mt.exp_avg_sq,
                                                                  mt.targets_prepared, tps, ps, mt.n_views,
                                                                  mt.camera_pos, mt.light_pos, P, K2, sc, loss, gPs);
            else
                fused_pixels<LOSS, OPT_LIGHT, false, NPIX, UNROLL>(px, mt.textures, mt.exp_avg, mt.exp_avg_sq,
                                                                   mt.targets_prepared, tps, ps, mt.n_views,
                                                                   mt.camera_pos, mt.light_pos, P, K2, sc, loss, gPs);
        }
        // block_sum ends with a barrier, so every thread has read s_slot before
        // thread 0 overwrites it with the next ticket
        const float bl = block_sum(loss, lds);
        float bp[3] = {0.f, 0.f, 0.f};
        if (OPT_LIGHT) {
#pragma unroll
            for (int c = 0; c < 3; ++c) bp[c] = block_sum(gPs[c], lds);
        }
        if (threadIdx.x == 0) {
            partials[id] = make_float4(bl, bp[0], bp[1], bp[2]);
            s_slot = atomicAdd(queue, 1u);
        }
        __syncthreads();
        slot = __builtin_amdgcn_readfirstlane(s_slot);
    }
}

// One block per tile instead of a persistent grid: block b takes the tile in
// cost order b (the dispatcher hands blocks out in index order, so the
// costliest tiles start first), with one pixel per thread (256-pixel tiles).
// No ticket counter and no loop; the register budget is the single-material
// kernel's.
template <int LOSS, int UNROLL, bool OPT_LIGHT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SVBRDF_FUSED_WAVES))) void batch_grid_kernel(
    const svbrdf_material *__restrict__ mats, const unsigned *__restrict__ order,
    const unsigned *__restrict__ tile_mat, float4 *__restrict__ partials, StepConsts c,
    const svbrdf_step_state *__restrict__ st) {
    __shared__ float lds[kBlock / 64];
    const StepScalars sc = step_scalars(st, c);
    const unsigned id = order[blockIdx.x];
    const svbrdf_material mt = mats[tile_mat[id]];
    const unsigned npix = (unsigned)mt.res * (unsigned)mt.res;
    const unsigned ps = mt.plane_stride ? mt.plane_stride : npix;
    const unsigned tps = mt.texture_stride ? mt.texture_stride : npix;
    const float K2 = mt.loss_scale;
    const float P[3] = {ldu(mt.light_pow, 0), ldu(mt.light_pow, 1), ldu(mt.light_pow, 2)};
    float loss = 0.f;
    float gPs[3] = {0.f, 0.f, 0.f};
    Pix px[1];
    px[0] = pixel_at((id - mt.tile_begin) * (unsigned)kBlock + threadIdx.x, npix, mt.res, mt.size);
    if (all_colocated(mt.camera_pos, mt.light_pos, mt.n_views))
        fused_pixels<LOSS, OPT_LIGHT, true, 1, UNROLL>(px, mt.textures, mt.exp_avg, mt.exp_avg_sq,
                                                       mt.targets_prepared, tps, ps, mt.n_views, mt.camera_pos,
                                                       mt.light_pos, P, K2, sc, loss, gPs);
    else
        fused_pixels<LOSS, OPT_LIGHT, false, 1, UNROLL>(px, mt.textures, mt.exp_avg, mt.exp_avg_sq,
                                                        mt.targets_prepared, tps, ps, mt.n_views, mt.camera_pos,
                                                        mt.light_pos, P, K2, sc, loss, gPs);
    const float bl = block_sum(loss, lds);
    float bp[3] = {0.f, 0.f, 0.f};
    if (OPT_LIGHT) {
#pragma unroll
        for (int c = 0; c < 3; ++c) bp[c] = block_sum(gPs[c], lds);
    }
    if (threadIdx.x == 0) partials[id] = make_float4(bl, bp[0], bp[1], bp[2]);
}

// Block m sums material m's tile partials in tile order (fixed, so the loss is
// bit-reproducible) and records its loss; reads the step count only.  With
// OPT_LIGHT it also sums the tiles' dL/dlight_pow and takes material m's
// Adam step on its light power (one param group with the maps,
// src/svbrdf.py:50: the maps' step scalars), moments in mt.light_adam.
template <bool OPT_LIGHT>
__global__ __launch_bounds__(kBlock) void batch_reduce_kernel(const svbrdf_material *__restrict__ mats,
                                                              const float4 *__restrict__ partials, StepConsts k,
                                                              const svbrdf_step_state *__restrict__ st,
                                                              float *__restrict__ losses,
                                                              float *__restrict__ loss_history, int history_len,
                                                              int n_materials) {
    constexpr int NQ = OPT_LIGHT ? 4 : 1;
    __shared__ double ldsd[kBlock / 64];
    const svbrdf_material mt = mats[blockIdx.x];
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (unsigned b = threadIdx.x; b < mt.n_tiles; b += kBlock) {
        const float4 v = partials[mt.tile_begin + b];
        acc[0] += (double)v.x;
        if constexpr (OPT_LIGHT) {
            acc[1] += (double)v.y;
            acc[2] += (double)v.z;
            acc[3] += (double)v.w;
        }
    }
    double tot[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) tot[q] = block_sum_d(acc[q], ldsd);   // each ends with a barrier
    if (threadIdx.x != 0) return;
    const double count = (double)mt.n_views * 3.0 * (double)mt.res * (double)mt.res;
    const float L = (float)(tot[0] / count);
    losses[blockIdx.x] = L;
    const int t = (int)st->step;   // completed steps before this one
    if (loss_history && t < history_len) loss_history[(size_t)t * n_materials + blockIdx.x] = L;
    if constexpr (OPT_LIGHT) {
        if (!mt.light_adam) return;
        const StepScalars sc = step_scalars(st, k);
        for (int c = 0; c < 3; ++c) {
            float pp = mt.light_pow[c], mm = mt.light_adam[c], vv = mt.light_adam[3 + c];
            adam_update(pp, mm, vv, (float)tot[1 + c], sc.w1, sc.beta2, sc.om_beta2, sc.inv_bc2_sqrt, sc.eps,
                        sc.neg_step);
            mt.light_pow[c] = pp;
            mt.light_adam[c] = mm;
            mt.light_adam[3 + c] = vv;
        }
    }
}

// Advances the shared step state and re-arms the ticket counter.
__global__ void batch_advance_kernel(svbrdf_step_state *__restrict__ st, unsigned *__restrict__ queue, double beta1,
                                     double beta2) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->step = st->step + 1.0f;
    st->beta1_pow *= beta1;
    st->beta2_pow *= beta2;
    set_bias_corrections(st);
    *queue = 0u;
}

size_t batch_partials_bytes(size_t n_tiles) { return (n_tiles * sizeof(float4) + 255) / 256 * 256; }

template <int L, bool LIGHT>
const void *batch_kernel_ptr(int npix, int unroll) {
    if (npix == 1)
        return unroll == 3 ? (const void *)batch_step_kernel<L, 1, 3, LIGHT> : (const void *)batch_step_kernel<L, 1, 1, LIGHT>;
    return unroll == 3 ? (const void *)batch_step_kernel<L, 2, 3, LIGHT> : (const void *)batch_step_kernel<L, 2, 1, LIGHT>;
}

template <int L, bool LIGHT>
hipError_t launch_batch_step(const svbrdf_material *materials, const unsigned *order, const unsigned *tile_material,
                       size_t n_tiles, int tile_pixels, int pixels_per_thread, int views_per_trip, int grid_blocks,
                       unsigned *queue, float4 *partials, const StepConsts &consts, svbrdf_step_state *st,
                       hipStream_t s) {
    if (grid_blocks == 0) {
        if (views_per_trip == 3)
            hipLaunchKernelGGL((batch_grid_kernel<L, 3, LIGHT>), dim3((unsigned)n_tiles), dim3(kBlock), 0, s,
                               materials, order, tile_material, partials, consts, st);
        else
            hipLaunchKernelGGL((batch_grid_kernel<L, 1, LIGHT>), dim3((unsigned)n_tiles), dim3(kBlock), 0, s,
                               materials, order, tile_material, partials, consts, st);
        return hipSuccess;
    }
    const void *k = batch_kernel_ptr<L, LIGHT>(pixels_per_thread, views_per_trip);
    const unsigned nt = (unsigned)n_tiles;
    void *args[] = {(void *)&materials, (void *)&order, (void *)&tile_material, (void *)&nt, (void *)&tile_pixels,
                    (void *)&queue, (void *)&partials, (void *)&consts, (void *)&st};
    return hipLaunchKernel(k, dim3(grid_blocks), dim3(kBlock), args, 0, s);
}

int check_variant(const char *what, int loss_mode, int npix, int unroll) {
    if (loss_mode != SVBRDF_LOSS_L2 && loss_mode != SVBRDF_LOSS_LOG_L2)
        return fail(SVBRDF_EINVAL, "%s: unknown loss_mode %d", what, loss_mode);
    if (npix != 1 && npix != 2) return fail(SVBRDF_EINVAL, "%s: pixels_per_thread must be 1 or 2", what);
    if (unroll != 1 && unroll != 3) return fail(SVBRDF_EINVAL, "%s: views_per_trip must be 1 or 3", what);
    return SVBRDF_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free adjoint.  The light-power partial sums (one float4 slot per
// 64-pixel block) of svbrdf_render_backward_noalloc live in this file-scope
// device array instead of a caller workspace: 262,144 slots, 4 MiB, enough for
// res <= 4096.  Each GPU holds its own copy for the life of the process.  In
// host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress, once per device (with that
// device current), and only that address reaches kernels.
constexpr int kScratchMaxRes = 4096;
constexpr unsigned kScratchSlots = (unsigned)(((size_t)kScratchMaxRes * kScratchMaxRes + kAdjointBlock - 1) / kAdjointBlock);
constexpr size_t kScratchBytes = (size_t)kScratchSlots * 4 * sizeof(float);
}  // namespace

static __device__ __align__(256) float scratch[kScratchSlots * 4];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_address[kMaxDevices];

// Device address of the current device's `scratch`, resolved on the first
// call on that device.
int scratch_address(const char *what, void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return fail(SVBRDF_EHIP, "%s: no current HIP device (or device index >= %d)", what, kMaxDevices);
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(SVBRDF_EHIP, "%s: hipGetSymbolAddress(scratch) on device %d failed: %s", what, dev,
                        hipGetErrorString(e));
        }
        g_scratch_address[dev] = p;
    }
    *out = g_scratch_address[dev];
    return SVBRDF_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

int svbrdf_abi_version(void) { return SVBRDF_ABI_VERSION; }
const char *svbrdf_last_error(void) { return svbrdf_internal::last_error(); }
size_t svbrdf_step_state_bytes(void) { return sizeof(svbrdf_step_state); }
int svbrdf_max_res(void) { return kMaxRes; }
int svbrdf_max_views(void) { return kMaxViews; }

int svbrdf_render_terms(const float *textures, int res, int n_views, float size, const float *camera_pos,
                        const float *light_pos, const float *light_pow, int terms, float *images, void *stream) {
    int rc = check_common("svbrdf_render", textures, res, n_views, camera_pos, light_pos, light_pow);
    if (rc) return rc;
    if (!images) return fail(SVBRDF_EINVAL, "svbrdf_render: null images");
    if (terms < 0 || terms > (SVBRDF_NDF_BECKMANN | SVBRDF_FRESNEL_SCHLICK))
        return fail(SVBRDF_EINVAL, "svbrdf_render: unknown terms %d", terms);
    hipStream_t s = (hipStream_t)stream;
#ifndef SVBRDF_RENDER_NPIX
#define SVBRDF_RENDER_NPIX 1      // pixels per thread of the forward render (2 measured equal at 1024^2 x 9)
#endif
#define SVBRDF_LAUNCH_RENDER(T)                                                                               \
    hipLaunchKernelGGL((render_kernel<SVBRDF_RENDER_NPIX, T>), dim3(n_blocks(res, SVBRDF_RENDER_NPIX, kRenderBlock)), \
                       dim3(kRenderBlock), 0, s, textures, res, n_views, size, camera_pos, light_pos, light_pow,       \
                       images)
    switch (terms) {
        case 0: SVBRDF_LAUNCH_RENDER(0); break;
        case 1: SVBRDF_LAUNCH_RENDER(1); break;
        case 2: SVBRDF_LAUNCH_RENDER(2); break;
        default: SVBRDF_LAUNCH_RENDER(3); break;
    }
#undef SVBRDF_LAUNCH_RENDER
    return check_launch("svbrdf_render");
}

int svbrdf_render(const float *textures, int res, int n_views, float size, const float *camera_pos,
                  const float *light_pos, const float *light_pow, float *images, void *stream) {
    return svbrdf_render_terms(textures, res, n_views, size, camera_pos, light_pos, light_pow, SVBRDF_TERMS_EVAL,
                               images, stream);
}

size_t svbrdf_render_backward_workspace_bytes(int res) {
    if (res < 1 || res > kMaxRes) return 0;
    return (size_t)n_blocks(res, 1, kAdjointBlock) * sizeof(float4);
}

int svbrdf_render_backward_terms(const float *textures, int res, int n_views, float size, const float *camera_pos,
                                 const float *light_pos, const float *light_pow, int terms,
                                 const float *grad_images, float *grad_textures, float *grad_light_pow,
                                 void *workspace, size_t workspace_bytes, void *stream) {
    int rc = check_common("svbrdf_render_backward", textures, res, n_views, camera_pos, light_pos, light_pow);
    if (rc) return rc;
    if (!grad_images || !grad_textures)
        return fail(SVBRDF_EINVAL, "svbrdf_render_backward: null gradient pointer");
    if (terms < 0 || terms > (SVBRDF_NDF_BECKMANN | SVBRDF_FRESNEL_SCHLICK))
        return fail(SVBRDF_EINVAL, "svbrdf_render_backward: unknown terms %d", terms);
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = n_blocks(res, 1, kAdjointBlock);
    if (grad_light_pow) {
        const size_t need = svbrdf_render_backward_workspace_bytes(res);
        if (!workspace || workspace_bytes < need)
            return fail(SVBRDF_ESPACE, "svbrdf_render_backward: workspace of %zu bytes < %zu needed",
                        workspace_bytes, need);
        if ((uintptr_t)workspace % 16)
            return fail(SVBRDF_EINVAL, "svbrdf_render_backward: workspace must be 16-byte aligned");
    }
    float4 *parts = grad_light_pow ? (float4 *)workspace : nullptr;
#define SVBRDF_LAUNCH_BWD(L, T)                                                                                \
    hipLaunchKernelGGL((backward_kernel<L, T>), dim3(nb), dim3(kAdjointBlock), 0, s, textures, res, n_views, size,      \
                       camera_pos, light_pos, light_pow, grad_images, grad_textures, parts)
#define SVBRDF_LAUNCH_BWD_T(L)                                                                                 \
    switch (terms) {                                                                                            \
        case 0: SVBRDF_LAUNCH_BWD(L, 0); break;                                                                 \
        case 1: SVBRDF_LAUNCH_BWD(L, 1); break;                                                                 \
        case 2: SVBRDF_LAUNCH_BWD(L, 2); break;                                                                 \
        default: SVBRDF_LAUNCH_BWD(L, 3); break;                                                                \
    }
    if (grad_light_pow) {
        SVBRDF_LAUNCH_BWD_T(true);
    } else {
        SVBRDF_LAUNCH_BWD_T(false);
    }
#undef SVBRDF_LAUNCH_BWD_T
#undef SVBRDF_LAUNCH_BWD
    if ((rc = check_launch("svbrdf_render_backward"))) return rc;
    if (grad_light_pow)
        hipLaunchKernelGGL(reduce_light_kernel, dim3(1), dim3(kLightReduceBlock), 0, s, (const float4 *)workspace,
                           nb, grad_light_pow);
    return check_launch("svbrdf_render_backward");
}

int svbrdf_render_backward(const float *textures, int res, int n_views, float size, const float *camera_pos,
                           const float *light_pos, const float *light_pow, const float *grad_images,
                           float *grad_textures, float *grad_light_pow, void *workspace,
                           size_t workspace_bytes, void *stream) {
    return svbrdf_render_backward_terms(textures, res, n_views, size, camera_pos, light_pos, light_pow,
                                        SVBRDF_TERMS_EVAL, grad_images, grad_textures, grad_light_pow, workspace,
                                        workspace_bytes, stream);
}

int svbrdf_render_backward_noalloc(const float *textures, int res, int n_views, float size, const float *camera_pos,
                                   const float *light_pos, const float *light_pow, const float *grad_images,
                                   float *grad_textures, float *grad_light_pow, void *stream) {
    const char *what = "svbrdf_render_backward_noalloc";
    if (res > kScratchMaxRes)
        return fail(SVBRDF_ESHAPE, "%s: res=%d above %d, the largest its %zu-byte scratch array serves (use "
                    "svbrdf_render_backward with a workspace)", what, res, kScratchMaxRes, kScratchBytes);
    void *ws = nullptr;
    if (int rc = scratch_address(what, &ws)) return rc;
    return svbrdf_render_backward_terms(textures, res, n_views, size, camera_pos, light_pos, light_pow,
                                        SVBRDF_TERMS_EVAL, grad_images, grad_textures, grad_light_pow, ws,
                                        kScratchBytes, stream);
}

int svbrdf_prepare_targets(const float *targets, size_t count, int loss_mode, float *prepared, void *stream) {
    if (!targets || !prepared) return fail(SVBRDF_EINVAL, "svbrdf_prepare_targets: null pointer");
    if (loss_mode != SVBRDF_LOSS_L2 && loss_mode != SVBRDF_LOSS_LOG_L2)
        return fail(SVBRDF_EINVAL, "svbrdf_prepare_targets: unknown loss_mode %d", loss_mode);
    if (count == 0) return SVBRDF_OK;
    const float floor_v = powf(1e-6f, 1.0f / 2.2f);
    size_t blocks = (count + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(prepare_targets_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                       targets, count, loss_mode == SVBRDF_LOSS_LOG_L2 ? 1 : 0, floor_v, prepared);
    return check_launch("svbrdf_prepare_targets");
}

int svbrdf_step_state_init(void *state, double beta1, double beta2, void *stream) {
    if (!state) return fail(SVBRDF_EINVAL, "svbrdf_step_state_init: null state");
    if (!(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0))
        return fail(SVBRDF_EINVAL, "svbrdf_step_state_init: betas must lie in [0, 1)");
    hipLaunchKernelGGL(state_init_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, (svbrdf_step_state *)state,
                       beta1, beta2);
    return check_launch("svbrdf_step_state_init");
}

size_t svbrdf_plane_stride(int res) {
    if (res < 1 || res > kMaxRes) return 0;
    const size_t npix = (size_t)res * res;
    return (npix + 63) / 64 * 64 + 448;
}

size_t svbrdf_fused_step_workspace_bytes(int res) {
    if (res < 1 || res > kMaxRes) return 0;
    return fused_workspace_bytes(n_blocks(res, SVBRDF_FUSED_NPIX));
}

int svbrdf_fused_step(float *textures, float *exp_avg, float *exp_avg_sq, const float *targets_prepared, int res,
                      size_t texture_stride, size_t plane_stride, int n_views, float size,
                      const float *camera_pos, const float *light_pos, float *light_pow,
                      int loss_mode, int optim_light, double lr, double beta1, double beta2, double eps,
                      void *state, float *loss_history, int history_len, void *workspace,
                      size_t workspace_bytes, void *stream) {
    int rc = check_common("svbrdf_fused_step", textures, res, n_views, camera_pos, light_pos, light_pow);
    if (rc) return rc;
    if (!exp_avg || !exp_avg_sq || !targets_prepared || !state)
        return fail(SVBRDF_EINVAL, "svbrdf_fused_step: null pointer argument");
    if (loss_mode != SVBRDF_LOSS_L2 && loss_mode != SVBRDF_LOSS_LOG_L2)
        return fail(SVBRDF_EINVAL, "svbrdf_fused_step: unknown loss_mode %d", loss_mode);
    if (!(lr >= 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps >= 0.0))
        return fail(SVBRDF_EINVAL, "svbrdf_fused_step: invalid Adam hyper-parameters");
    const size_t need = svbrdf_fused_step_workspace_bytes(res);
    if (!workspace || workspace_bytes < need)
        return fail(SVBRDF_ESPACE, "svbrdf_fused_step: workspace of %zu bytes < %zu needed", workspace_bytes, need);
    if (((uintptr_t)workspace & 255u) || ((uintptr_t)state & 15u))
        return fail(SVBRDF_EINVAL, "svbrdf_fused_step: workspace must be 256-byte and state 16-byte aligned");
    unsigned tps = 0, ps = 0;
    if ((rc = check_plane_stride("svbrdf_fused_step", res, texture_stride, tps))) return rc;
    if ((rc = check_plane_stride("svbrdf_fused_step", res, plane_stride, ps))) return rc;
    const unsigned nb = n_blocks(res, SVBRDF_FUSED_NPIX);
    hipStream_t s = (hipStream_t)stream;
    svbrdf_step_state *st = (svbrdf_step_state *)state;
    const StepConsts consts = make_consts(lr, beta1, beta2, eps, (double)n_views * 3.0 * (double)res * (double)res);
#define SVBRDF_LAUNCH_STEP(L, O)                                                                             \
    hipLaunchKernelGGL((fused_step_kernel<L, O>), dim3(nb), dim3(kBlock), 0, s, textures, exp_avg, exp_avg_sq, \
                       targets_prepared, res, tps, ps, n_views, size, camera_pos, light_pos, light_pow, consts, \
                       st, workspace, loss_history, history_len)
    if (loss_mode == SVBRDF_LOSS_LOG_L2) {
        if (optim_light) SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_LOG_L2, true);
        else SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_LOG_L2, false);
    } else {
        if (optim_light) SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_L2, true);
        else SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_L2, false);
    }
#undef SVBRDF_LAUNCH_STEP
    if ((rc = check_launch("svbrdf_fused_step"))) return rc;
    if (optim_light) {
        hipLaunchKernelGGL(finalize_step_kernel<true>, dim3(1), dim3(kFinalizeBlock), 0, s, workspace, nb, consts,
                           light_pow, st, loss_history, history_len);
    }
    return check_launch("svbrdf_fused_step");
}

int svbrdf_fused_step_flush(void *state, int res, int n_views, float *loss_history, int history_len,
                            const void *workspace, size_t workspace_bytes, void *stream) {
    if (!state || !workspace) return fail(SVBRDF_EINVAL, "svbrdf_fused_step_flush: null pointer argument");
    if (res < 1 || res > kMaxRes) return fail(SVBRDF_ESHAPE, "svbrdf_fused_step_flush: res=%d", res);
    if (n_views < 1 || n_views > kMaxViews)
        return fail(SVBRDF_ESHAPE, "svbrdf_fused_step_flush: n_views=%d", n_views);
    const size_t need = svbrdf_fused_step_workspace_bytes(res);
    if (workspace_bytes < need)
        return fail(SVBRDF_ESPACE, "svbrdf_fused_step_flush: workspace of %zu bytes < %zu needed", workspace_bytes,
                    need);
    const double count = (double)n_views * 3.0 * (double)res * (double)res;
    hipLaunchKernelGGL(flush_step_kernel, dim3(1), dim3(kBlock), 0, (hipStream_t)stream, workspace,
                       n_blocks(res, SVBRDF_FUSED_NPIX), count, (svbrdf_step_state *)state, loss_history,
                       history_len);
    return check_launch("svbrdf_fused_step_flush");
}

// ---------------------------------------------------------------------------
// Batched step

size_t svbrdf_batch_tile_count(const svbrdf_material *materials, int n_materials, int tile_pixels) {
    if (!materials || n_materials < 1 || tile_pixels < kBlock || tile_pixels % kBlock) return 0;
    size_t n = 0;
    for (int i = 0; i < n_materials; ++i) {
        const int r = materials[i].res;
        if (r < 1 || r > kMaxRes) return 0;
        n += ((size_t)r * r + tile_pixels - 1) / tile_pixels;
    }
    return n;
}

int svbrdf_batch_plan(svbrdf_material *materials, int n_materials, int tile_pixels, unsigned int *order,
                      unsigned int *tile_material, size_t n_tiles) {
    if (!materials || !order || !tile_material || n_materials < 1)
        return fail(SVBRDF_EINVAL, "svbrdf_batch_plan: null pointer or empty batch");
    if (tile_pixels < kBlock || tile_pixels % kBlock)
        return fail(SVBRDF_EINVAL, "svbrdf_batch_plan: tile_pixels=%d must be a positive multiple of %d",
                    tile_pixels, kBlock);
    for (int i = 0; i < n_materials; ++i) {
        const svbrdf_material &m = materials[i];
        if (!m.textures || !m.exp_avg || !m.exp_avg_sq || !m.targets_prepared || !m.camera_pos || !m.light_pos ||
            !m.light_pow)
            return fail(SVBRDF_EINVAL, "svbrdf_batch_plan: material %d has a null pointer", i);
        if (m.res < 1 || m.res > kMaxRes || m.n_views < 1 || m.n_views > kMaxViews)
            return fail(SVBRDF_ESHAPE, "svbrdf_batch_plan: material %d has res=%d n_views=%d", i, m.res, m.n_views);
        unsigned ps = 0;
        if (int rc = check_plane_stride("svbrdf_batch_plan", m.res, m.plane_stride, ps)) return rc;
        materials[i].loss_scale = loss_scale(m.res, m.n_views);
        if (int rc = check_plane_stride("svbrdf_batch_plan", m.res, m.texture_stride, ps)) return rc;
    }
    const size_t need = svbrdf_batch_tile_count(materials, n_materials, tile_pixels);
    if (need != n_tiles)
        return fail(SVBRDF_ESPACE, "svbrdf_batch_plan: tile arrays hold %zu entries, the batch has %zu tiles",
                    n_tiles, need);
    if (n_tiles > 0xffffffffull) return fail(SVBRDF_ESHAPE, "svbrdf_batch_plan: too many tiles");
    std::vector<double> cost(n_tiles);
    unsigned t = 0;
    for (int i = 0; i < n_materials; ++i) {
        svbrdf_material &m = materials[i];
        const size_t npix = (size_t)m.res * m.res;
        m.tile_begin = t;
        m.n_tiles = (unsigned)((npix + tile_pixels - 1) / tile_pixels);
        for (unsigned k = 0; k < m.n_tiles; ++k, ++t) {
            tile_material[t] = (unsigned)i;
            const size_t px = std::min<size_t>(tile_pixels, npix - (size_t)k * tile_pixels);
            cost[t] = (double)px * (double)m.n_views;
        }
    }
    for (size_t k = 0; k < n_tiles; ++k) order[k] = (unsigned)k;
    std::stable_sort(order, order + n_tiles, [&](unsigned a, unsigned b) { return cost[a] > cost[b]; });
    return SVBRDF_OK;
}

size_t svbrdf_batch_workspace_bytes(size_t n_tiles) { return batch_partials_bytes(n_tiles) + 256; }

int svbrdf_batch_max_blocks_per_cu(int loss_mode, int optim_light, int pixels_per_thread, int views_per_trip) {
    if (check_variant("svbrdf_batch_max_blocks_per_cu", loss_mode, pixels_per_thread, views_per_trip)) return 0;
    const bool log_l2 = loss_mode == SVBRDF_LOSS_LOG_L2;
    const void *k = optim_light
                        ? (log_l2 ? batch_kernel_ptr<SVBRDF_LOSS_LOG_L2, true>(pixels_per_thread, views_per_trip)
                                  : batch_kernel_ptr<SVBRDF_LOSS_L2, true>(pixels_per_thread, views_per_trip))
                        : (log_l2 ? batch_kernel_ptr<SVBRDF_LOSS_LOG_L2, false>(pixels_per_thread, views_per_trip)
                                  : batch_kernel_ptr<SVBRDF_LOSS_L2, false>(pixels_per_thread, views_per_trip));
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kBlock, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int svbrdf_batch_step(const svbrdf_material *materials, int n_materials, const unsigned int *order,
                      const unsigned int *tile_material, size_t n_tiles, int tile_pixels, int pixels_per_thread,
                      int views_per_trip, int grid_blocks, int loss_mode, int optim_light, double lr, double beta1,
                      double beta2, double eps, void *state, float *losses, float *loss_history, int history_len,
                      void *workspace, size_t workspace_bytes, void *stream) {
    const char *what = "svbrdf_batch_step";
    int rc = check_variant(what, loss_mode, pixels_per_thread, views_per_trip);
    if (rc) return rc;
    if (!materials || !order || !tile_material || !state || !losses || !workspace || n_materials < 1)
        return fail(SVBRDF_EINVAL, "%s: null pointer argument or empty batch", what);
    if (n_tiles == 0 || n_tiles > 0xffffffffull) return fail(SVBRDF_ESHAPE, "%s: n_tiles=%zu", what, n_tiles);
    if (tile_pixels < kBlock * pixels_per_thread || tile_pixels % (kBlock * pixels_per_thread))
        return fail(SVBRDF_EINVAL, "%s: tile_pixels=%d must be a multiple of %d", what, tile_pixels,
                    kBlock * pixels_per_thread);
    if (grid_blocks < 0 || grid_blocks > (1 << 20)) return fail(SVBRDF_EINVAL, "%s: grid_blocks=%d", what, grid_blocks);
    if (grid_blocks == 0 && (tile_pixels != kBlock || pixels_per_thread != 1 || n_tiles > 0x7fffffffull))
        return fail(SVBRDF_EINVAL, "%s: one block per tile (grid_blocks=0) needs tile_pixels=%d, one pixel per "
                    "thread and fewer than 2^31 tiles", what, kBlock);
    if (!(lr >= 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps >= 0.0))
        return fail(SVBRDF_EINVAL, "%s: invalid Adam hyper-parameters", what);
    if (workspace_bytes < svbrdf_batch_workspace_bytes(n_tiles))
        return fail(SVBRDF_ESPACE, "%s: workspace of %zu bytes < %zu needed", what, workspace_bytes,
                    svbrdf_batch_workspace_bytes(n_tiles));
    if (((uintptr_t)workspace & 255u) || ((uintptr_t)state & 15u))
        return fail(SVBRDF_EINVAL, "%s: workspace must be 256-byte and state 16-byte aligned", what);
    hipStream_t s = (hipStream_t)stream;
    float4 *partials = (float4 *)workspace;
    const StepConsts consts = make_consts(lr, beta1, beta2, eps, 0.0);   // K2 is per material
    unsigned *queue = (unsigned *)((char *)workspace + batch_partials_bytes(n_tiles));
    svbrdf_step_state *st = (svbrdf_step_state *)state;
    const bool log_l2 = loss_mode == SVBRDF_LOSS_LOG_L2;
    const auto launch = optim_light ? (log_l2 ? launch_batch_step<SVBRDF_LOSS_LOG_L2, true>
                                              : launch_batch_step<SVBRDF_LOSS_L2, true>)
                                    : (log_l2 ? launch_batch_step<SVBRDF_LOSS_LOG_L2, false>
                                              : launch_batch_step<SVBRDF_LOSS_L2, false>);
    if (launch(materials, order, tile_material, n_tiles, tile_pixels, pixels_per_thread, views_per_trip, grid_blocks,
               queue, partials, consts, st, s) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SVBRDF_EHIP, "%s: step launch failed", what);
    }
    if ((rc = check_launch(what))) return rc;
    if (optim_light)
        hipLaunchKernelGGL(batch_reduce_kernel<true>, dim3(n_materials), dim3(kBlock), 0, s, materials, partials,
                           consts, st, losses, loss_history, history_len, n_materials);
    else
        hipLaunchKernelGGL(batch_reduce_kernel<false>, dim3(n_materials), dim3(kBlock), 0, s, materials, partials,
                           consts, st, losses, loss_history, history_len, n_materials);
    if ((rc = check_launch(what))) return rc;
    hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, s, st, queue, beta1, beta2);
    return check_launch(what);
}

}  // extern "C"
