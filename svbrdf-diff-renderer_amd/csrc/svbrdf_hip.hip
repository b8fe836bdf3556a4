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

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "brdf.hpp"
#include "../../include/svbrdf_hip.h"

using namespace svbrdf;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxRes = 16384;
constexpr int kMaxViews = 1024;

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SVBRDF_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
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

inline unsigned n_blocks(int res) {
    const size_t npix = (size_t)res * res;
    return (unsigned)((npix + kBlock - 1) / kBlock);
}

// Geometry of view k, read with wave-uniform addresses (scalar loads).
struct ViewPos {
    float cx, cy, cz, lx, ly, lz;
    bool coloc;
};

__device__ __forceinline__ ViewPos load_view(const float *__restrict__ cam, const float *__restrict__ light,
                                             int k) {
    ViewPos p;
    p.cx = cam[3 * k + 0]; p.cy = cam[3 * k + 1]; p.cz = cam[3 * k + 2];
    p.lx = light[3 * k + 0]; p.ly = light[3 * k + 1]; p.lz = light[3 * k + 2];
    p.coloc = (p.cx == p.lx) && (p.cy == p.ly) && (p.cz == p.lz);
    return p;
}

// Wave64 + block reduction of a float; result valid in thread 0.
__device__ __forceinline__ float block_sum(float x, float *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = x;
    __syncthreads();
    float r = 0.f;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) r += lds[w];
    }
    __syncthreads();
    return r;
}

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
// Forward render.
__global__ __launch_bounds__(kBlock) void render_kernel(
    const float *__restrict__ tex, int res, int n_views, float size, const float *__restrict__ cam,
    const float *__restrict__ light, const float *__restrict__ lpow, float *__restrict__ out) {
    const size_t npix = (size_t)res * res;
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    const int i = (int)(p / (unsigned)res), j = (int)(p - (size_t)i * res);
    const float px = pixel_coord(j, res, size), py = -pixel_coord(i, res, size);
    float t[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) t[c] = tex[c * npix + p];
    Maps m;
    decode(t, m);
    const float P[3] = {lpow[0], lpow[1], lpow[2]};
    for (int k = 0; k < n_views; ++k) {
        const ViewPos vp = load_view(cam, light, k);
        Geom g;
        geometry(px, py, vp.cx, vp.cy, vp.cz, vp.lx, vp.ly, vp.lz, vp.coloc, g);
        Shade s;
        shade(m, g, P, s);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float ic = tclamp(s.img[c], kEps, 1.0f);
            out[(size_t)(3 * k + c) * npix + p] = fexp2(flog2(ic) * kInvGamma);
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the render for an arbitrary dL/dimages.
template <bool LIGHT_GRAD>
__global__ __launch_bounds__(kBlock) void backward_kernel(
    const float *__restrict__ tex, int res, int n_views, float size, const float *__restrict__ cam,
    const float *__restrict__ light, const float *__restrict__ lpow, const float *__restrict__ gout,
    float *__restrict__ gtex, float *__restrict__ light_partials) {
    __shared__ float lds[kBlock / 64];
    const size_t npix = (size_t)res * res;
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    Accum a;
    accum_zero(a);
    if (p < npix) {
        const int i = (int)(p / (unsigned)res), j = (int)(p - (size_t)i * res);
        const float px = pixel_coord(j, res, size), py = -pixel_coord(i, res, size);
        float t[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) t[c] = tex[c * npix + p];
        Maps m;
        decode(t, m);
        const float P[3] = {lpow[0], lpow[1], lpow[2]};
        for (int k = 0; k < n_views; ++k) {
            const ViewPos vp = load_view(cam, light, k);
            Geom g;
            geometry(px, py, vp.cx, vp.cy, vp.cz, vp.lx, vp.ly, vp.lz, vp.coloc, g);
            Shade s;
            shade(m, g, P, s);
            float gimg[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float img = s.img[c];
                const float ic = tclamp(img, kEps, 1.0f);
                const float o = fexp2(flog2(ic) * kInvGamma);
                const float go = gout[(size_t)(3 * k + c) * npix + p];
                // d/dimg of clamp(img)^(1/2.2) = (1/2.2) out / img inside [eps, 1]
                gimg[c] = (img >= kEps && img <= 1.0f) ? go * kInvGamma * o * frcp(ic) : 0.0f;
            }
            shade_backward<LIGHT_GRAD>(m, g, s, P, gimg, a);
        }
        float gt[9];
        decode_backward(t, m, a, gt);
#pragma unroll
        for (int c = 0; c < 9; ++c) gtex[c * npix + p] = gt[c];
    }
    if (LIGHT_GRAD) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float s = block_sum(a.gP[c], lds);
            if (threadIdx.x == 0) light_partials[(size_t)blockIdx.x * 4 + c] = s;
        }
    }
}

// Sums per-block light-power partials in a fixed order (deterministic).
__global__ __launch_bounds__(kBlock) void reduce_light_kernel(const float *__restrict__ partials,
                                                              unsigned n, float *__restrict__ out) {
    __shared__ double lds[kBlock / 64];
    for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (unsigned b = threadIdx.x; b < n; b += kBlock) acc += (double)partials[(size_t)b * 4 + c];
        const double s = block_sum_d(acc, lds);
        if (threadIdx.x == 0) out[c] = (float)s;
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

__global__ void state_init_kernel(svbrdf_step_state *st, double beta1, double beta2) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->bias_correction1 = 1.0 - beta1;
    st->bias_correction2 = 1.0 - beta2;
    st->step = 0.0f;
    st->loss = 0.0f;
    st->blocks_done = 0u;
    st->n_partials = 0u;
    for (int c = 0; c < 4; ++c) {
        st->light_exp_avg[c] = 0.0f;
        st->light_exp_avg_sq[c] = 0.0f;
        st->light_grad[c] = 0.0f;
    }
}

// One Adam update of a single fp32 element in torch's single-tensor order
// (torch/optim/adam.py:431-547): lerp, mul+addcmul, sqrt/bc2_sqrt + eps,
// addcdiv with value -lr/bc1.
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

__device__ __forceinline__ StepScalars step_scalars(const svbrdf_step_state *st, double lr, double beta1,
                                                    double beta2, double eps) {
    StepScalars s;
    s.w1 = (float)(1.0 - beta1);
    s.beta2 = (float)beta2;
    s.om_beta2 = (float)(1.0 - beta2);
    s.inv_bc2_sqrt = 1.0f / (float)sqrt(st->bias_correction2);
    s.eps = (float)eps;
    s.neg_step = (float)(-(lr / st->bias_correction1));
    return s;
}

// ---------------------------------------------------------------------------
// Fused per-pixel optimisation step.  Every block writes its partial loss (and
// light-power gradient); the last block to finish reduces them in a fixed
// order, updates light_pow when it is optimised, and advances the step state.
template <int LOSS, bool OPT_LIGHT>
__global__ __launch_bounds__(kBlock) void fused_step_kernel(
    float *__restrict__ tex, float *__restrict__ exp_avg, float *__restrict__ exp_avg_sq,
    const float *__restrict__ tgt, int res, int n_views, float size, const float *__restrict__ cam,
    const float *__restrict__ light, float *__restrict__ lpow, double lr, double beta1, double beta2,
    double eps, svbrdf_step_state *__restrict__ st, float *__restrict__ loss_history, int history_len,
    float *__restrict__ partials) {
    __shared__ float lds[kBlock / 64];
    __shared__ double ldsd[kBlock / 64];
    __shared__ int am_last;
    const size_t npix = (size_t)res * res;
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const float K = 2.0f / (float)((double)n_views * 3.0 * (double)npix);   // d mean / d elem
    const StepScalars sc = step_scalars(st, lr, beta1, beta2, eps);
    const float P[3] = {lpow[0], lpow[1], lpow[2]};
    float loss = 0.f;
    Accum a;
    accum_zero(a);
    if (p < npix) {
        const int i = (int)(p / (unsigned)res), j = (int)(p - (size_t)i * res);
        const float px = pixel_coord(j, res, size), py = -pixel_coord(i, res, size);
        float traw[9], t[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            traw[c] = tex[c * npix + p];
            t[c] = tclamp(traw[c], -1.0f, 1.0f);   // textures.clamp(-1, 1), svbrdf.py:60
        }
        Maps m;
        decode(t, m);
        for (int k = 0; k < n_views; ++k) {
            const ViewPos vp = load_view(cam, light, k);
            Geom g;
            geometry(px, py, vp.cx, vp.cy, vp.cz, vp.lx, vp.ly, vp.lz, vp.coloc, g);
            Shade s;
            shade(m, g, P, s);
            float gimg[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float img = s.img[c];
                const float ic = tclamp(img, kEps, 1.0f);
                const float lic = flog2(ic);
                const float y = tgt[(size_t)(3 * k + c) * npix + p];
                const bool inside = (img >= kEps && img <= 1.0f);
                float r, gi;
                if (LOSS == SVBRDF_LOSS_LOG_L2) {
                    r = lic * (kInvGamma * kLn2) - y;           // ln out - ln target'
                    gi = K * kInvGamma * r * frcp(ic);
                } else {
                    const float o = fexp2(lic * kInvGamma);
                    r = o - y;
                    gi = K * kInvGamma * r * o * frcp(ic);
                }
                loss += r * r;
                gimg[c] = inside ? gi : 0.0f;
            }
            shade_backward<OPT_LIGHT>(m, g, s, P, gimg, a);
        }
        float gt[9];
        decode_backward(t, m, a, gt);
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const size_t e = c * npix + p;
            float pp = traw[c], mm = exp_avg[e], vv = exp_avg_sq[e];
            // the clamp on read passes the gradient only inside [-1, 1] (inclusive)
            const float g = (traw[c] >= -1.0f && traw[c] <= 1.0f) ? gt[c] : 0.0f;
            adam_update(pp, mm, vv, g, sc.w1, sc.beta2, sc.om_beta2, sc.inv_bc2_sqrt, sc.eps, sc.neg_step);
            tex[e] = pp;
            exp_avg[e] = mm;
            exp_avg_sq[e] = vv;
        }
    }
    // per-block partials: loss, dL/dP
    const float bl = block_sum(loss, lds);
    float bp[3] = {0.f, 0.f, 0.f};
    if (OPT_LIGHT) {
#pragma unroll
        for (int c = 0; c < 3; ++c) bp[c] = block_sum(a.gP[c], lds);
    }
    if (threadIdx.x == 0) {
        float4 *dst = reinterpret_cast<float4 *>(partials) + blockIdx.x;
        *dst = make_float4(bl, bp[0], bp[1], bp[2]);
        __threadfence();
        const unsigned done = atomicAdd(&st->blocks_done, 1u);
        am_last = (done == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!am_last) return;
    __threadfence();
    // last block: deterministic reduction of all partials
    const float4 *parts = reinterpret_cast<const float4 *>(partials);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (unsigned b = threadIdx.x; b < gridDim.x; b += kBlock) {
        const float4 q = parts[b];
        acc[0] += q.x; acc[1] += q.y; acc[2] += q.z; acc[3] += q.w;
    }
    double tot[4];
    tot[0] = block_sum_d(acc[0], ldsd);
    if (OPT_LIGHT) {
        for (int c = 1; c < 4; ++c) tot[c] = block_sum_d(acc[c], ldsd);
    }
    if (threadIdx.x == 0) {
        const double step = (double)st->step + 1.0;
        const float L = (float)(tot[0] / ((double)n_views * 3.0 * (double)npix));
        st->loss = L;
        if (loss_history && (double)history_len >= step) loss_history[(int)step - 1] = L;
        if (OPT_LIGHT) {
            // Adam on light_pow with the same scalars as the maps (one param group)
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
        st->step = (float)step;
        st->bias_correction1 = 1.0 - pow(beta1, step + 1.0);
        st->bias_correction2 = 1.0 - pow(beta2, step + 1.0);
        st->n_partials = gridDim.x;
        st->blocks_done = 0u;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

int svbrdf_abi_version(void) { return SVBRDF_ABI_VERSION; }
const char *svbrdf_last_error(void) { return g_last_error.c_str(); }
size_t svbrdf_step_state_bytes(void) { return sizeof(svbrdf_step_state); }
int svbrdf_max_res(void) { return kMaxRes; }
int svbrdf_max_views(void) { return kMaxViews; }

int svbrdf_render(const float *textures, int res, int n_views, float size, const float *camera_pos,
                  const float *light_pos, const float *light_pow, float *images, void *stream) {
    int rc = check_common("svbrdf_render", textures, res, n_views, camera_pos, light_pos, light_pow);
    if (rc) return rc;
    if (!images) return fail(SVBRDF_EINVAL, "svbrdf_render: null images");
    hipLaunchKernelGGL(render_kernel, dim3(n_blocks(res)), dim3(kBlock), 0, (hipStream_t)stream, textures,
                       res, n_views, size, camera_pos, light_pos, light_pow, images);
    return check_launch("svbrdf_render");
}

size_t svbrdf_render_backward_workspace_bytes(int res) {
    if (res < 1 || res > kMaxRes) return 0;
    return (size_t)n_blocks(res) * 4 * sizeof(float);
}

int svbrdf_render_backward(const float *textures, int res, int n_views, float size, const float *camera_pos,
                           const float *light_pos, const float *light_pow, const float *grad_images,
                           float *grad_textures, float *grad_light_pow, void *workspace,
                           size_t workspace_bytes, void *stream) {
    int rc = check_common("svbrdf_render_backward", textures, res, n_views, camera_pos, light_pos, light_pow);
    if (rc) return rc;
    if (!grad_images || !grad_textures)
        return fail(SVBRDF_EINVAL, "svbrdf_render_backward: null gradient pointer");
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = n_blocks(res);
    if (grad_light_pow) {
        const size_t need = svbrdf_render_backward_workspace_bytes(res);
        if (!workspace || workspace_bytes < need)
            return fail(SVBRDF_ESPACE, "svbrdf_render_backward: workspace of %zu bytes < %zu needed",
                        workspace_bytes, need);
        hipLaunchKernelGGL(backward_kernel<true>, dim3(nb), dim3(kBlock), 0, s, textures, res, n_views, size,
                           camera_pos, light_pos, light_pow, grad_images, grad_textures, (float *)workspace);
        if ((rc = check_launch("svbrdf_render_backward"))) return rc;
        hipLaunchKernelGGL(reduce_light_kernel, dim3(1), dim3(kBlock), 0, s, (const float *)workspace, nb,
                           grad_light_pow);
    } else {
        hipLaunchKernelGGL(backward_kernel<false>, dim3(nb), dim3(kBlock), 0, s, textures, res, n_views, size,
                           camera_pos, light_pos, light_pow, grad_images, grad_textures, (float *)nullptr);
    }
    return check_launch("svbrdf_render_backward");
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

size_t svbrdf_fused_step_workspace_bytes(int res) {
    if (res < 1 || res > kMaxRes) return 0;
    return (size_t)n_blocks(res) * 4 * sizeof(float);
}

int svbrdf_fused_step(float *textures, float *exp_avg, float *exp_avg_sq, const float *targets_prepared, int res,
                      int n_views, float size, const float *camera_pos, const float *light_pos, float *light_pow,
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
    if (((uintptr_t)workspace & 15u) || ((uintptr_t)state & 15u))
        return fail(SVBRDF_EINVAL, "svbrdf_fused_step: workspace and state must be 16-byte aligned");
    const unsigned nb = n_blocks(res);
    hipStream_t s = (hipStream_t)stream;
    svbrdf_step_state *st = (svbrdf_step_state *)state;
    float *ws = (float *)workspace;
#define SVBRDF_LAUNCH_STEP(L, O)                                                                             \
    hipLaunchKernelGGL((fused_step_kernel<L, O>), dim3(nb), dim3(kBlock), 0, s, textures, exp_avg, exp_avg_sq, \
                       targets_prepared, res, n_views, size, camera_pos, light_pos, light_pow, lr, beta1,      \
                       beta2, eps, st, loss_history, history_len, ws)
    if (loss_mode == SVBRDF_LOSS_LOG_L2) {
        if (optim_light) SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_LOG_L2, true);
        else SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_LOG_L2, false);
    } else {
        if (optim_light) SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_L2, true);
        else SVBRDF_LAUNCH_STEP(SVBRDF_LOSS_L2, false);
    }
#undef SVBRDF_LAUNCH_STEP
    return check_launch("svbrdf_fused_step");
}

}  // extern "C"
