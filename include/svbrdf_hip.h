/*
 * svbrdf_hip.h -- C ABI of the MI355X (gfx950) microfacet renderer and
 * per-pixel SVBRDF optimisation step.
 *
 * The reference has no FFI: its boundary is the duck-typed Python renderer
 * object held by the optimiser (src/optimization.py:15).  Each entry point
 * below replaces one operation of that object or of the optimiser loop that
 * drives it; the Python wrapper in svbrdf-diff-renderer_amd/ binds them with
 * ctypes and keeps the reference's class surface.
 *
 * Conventions
 *   - All tensor pointers are DEVICE pointers to contiguous fp32 arrays in
 *     the reference's layouts: textures [9,R,R] (channels: diffuse 0:3,
 *     normal xy 3:5, roughness 5, specular 6:9; src/svbrdf.py:166),
 *     images [V,3,R,R], camera_pos / light_pos [V,3] in cm, light_pow [3].
 *   - The library borrows every pointer for the duration of a stream-ordered
 *     call and allocates nothing.  Work that needs scratch memory takes a
 *     caller-owned workspace whose size the *_workspace_bytes() query gives.
 *   - Every call is asynchronous on `stream` (a hipStream_t passed as void*;
 *     NULL = the default stream) and contains no host synchronisation, so it
 *     can be captured into a HIP graph.
 *   - Return value: SVBRDF_OK or a negative status; svbrdf_last_error()
 *     returns a message for the calling thread's last failure.  Arguments are
 *     validated on the host before anything is launched.
 */
#ifndef SVBRDF_HIP_H
#define SVBRDF_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVBRDF_ABI_VERSION 12

enum svbrdf_status {
    SVBRDF_OK = 0,
    SVBRDF_EINVAL = -1,   /* null pointer, bad size or bad enum value         */
    SVBRDF_ESHAPE = -2,   /* resolution / view count out of range             */
    SVBRDF_ESPACE = -3,   /* workspace too small                              */
    SVBRDF_EHIP = -4,     /* a HIP runtime call failed (message has details)  */
};

enum svbrdf_loss {
    SVBRDF_LOSS_L2 = 0,      /* mean((out - target)^2): th.nn.MSELoss, src/optimization.py:14,29 */
    SVBRDF_LOSS_LOG_L2 = 1,  /* mean((ln out - ln clamp(target, eps^(1/2.2), 1))^2)             */
};

/* Device-resident optimiser state of the fused step.  Lives in device
 * memory (svbrdf_step_state_bytes() bytes, 16-byte aligned), is set up by
 * svbrdf_step_state_init() and advanced by every svbrdf_fused_step() call,
 * so a captured graph replays correctly without host-side scalars. */
typedef struct svbrdf_step_state {
    double beta1_pow;            /* beta1^t for the step t the next call performs   */
    double beta2_pow;            /* beta2^t                                         */
    double inv_bias_correction1; /* 1 / (1 - beta1^t)                               */
    float inv_bias_correction2_sqrt; /* 1 / sqrt(1 - beta2^t), rounded to fp32      */
    float step;                  /* completed steps (torch keeps it as a float)     */
    float loss;                  /* image loss of the last recorded step            */
    unsigned int pending_step;   /* step whose partial sums await reduction, or 0   */
    unsigned int reserved[2];
    float light_exp_avg[4];      /* Adam state of light_pow (when optimised)        */
    float light_exp_avg_sq[4];
    float light_grad[4];         /* dL/dlight_pow of the last step (when optimised) */
} svbrdf_step_state;

int svbrdf_abi_version(void);
const char *svbrdf_last_error(void);
size_t svbrdf_step_state_bytes(void);

/* Largest resolution and view count accepted (7680: the 9 planes of a map
 * must fit one buffer range below 2^31 bytes). */
int svbrdf_max_res(void);
int svbrdf_max_views(void);

/* Forward render.  Replaces Microfacet.eval(textures) (src/microfacet.py:84-120)
 * with the light power set by Microfacet.update_light (src/microfacet.py:81-82)
 * and the geometry of Microfacet.__init__ (src/microfacet.py:10-24).
 * images[v,c,i,j] = clamp(P_c f_c(n.l)/|l|^2, 1e-6, 1)^(1/2.2). */
int svbrdf_render(const float *textures, int res, int n_views, float size,
                  const float *camera_pos, const float *light_pos,
                  const float *light_pow, float *images, void *stream);

/* BRDF terms of the render (bit set).  SVBRDF_TERMS_EVAL is what
 * Microfacet.eval uses: GGX (src/microfacet.py:28-32) and the
 * spherical-Gaussian Schlick Fresnel (:43-45).  The reference also defines
 * Beckmann (:34-38) and Fresnel_f0 (:40-41, here with f0 = the specular map)
 * without calling them from eval. */
enum svbrdf_terms {
    SVBRDF_TERMS_EVAL = 0,
    SVBRDF_NDF_BECKMANN = 1,       /* D = exp(-t^2/a^2) / (pi a^2 (n.h)^4), t^2 = (1 - (n.h)^2)/(n.h)^2 */
    SVBRDF_FRESNEL_SCHLICK = 2,    /* F = spe + (1 - spe)(1 - v.h)^5                                    */
};

/* svbrdf_render / svbrdf_render_backward with a choice of terms (any
 * combination of the bits above; SVBRDF_TERMS_EVAL gives exactly
 * svbrdf_render / svbrdf_render_backward).  The fused step always uses the
 * eval terms. */
int svbrdf_render_terms(const float *textures, int res, int n_views, float size,
                        const float *camera_pos, const float *light_pos,
                        const float *light_pow, int terms, float *images, void *stream);
int svbrdf_render_backward_terms(const float *textures, int res, int n_views, float size,
                                 const float *camera_pos, const float *light_pos,
                                 const float *light_pow, int terms, const float *grad_images,
                                 float *grad_textures, float *grad_light_pow,
                                 void *workspace, size_t workspace_bytes, void *stream);

/* Adjoint of svbrdf_render: the backward of Microfacet.eval under torch
 * autograd (src/microfacet.py:64-120).  Given dL/dimages it writes
 * dL/dtextures (overwritten, not accumulated) and, when grad_light_pow is not
 * NULL, dL/dlight_pow[3] (needs a 16-byte aligned workspace of
 * svbrdf_render_backward_workspace_bytes(res) bytes; otherwise workspace may
 * be NULL).  The BRDF terms are recomputed, not stored. */
size_t svbrdf_render_backward_workspace_bytes(int res);
int svbrdf_render_backward(const float *textures, int res, int n_views, float size,
                           const float *camera_pos, const float *light_pos,
                           const float *light_pow, const float *grad_images,
                           float *grad_textures, float *grad_light_pow,
                           void *workspace, size_t workspace_bytes, void *stream);

/* svbrdf_render_backward without a workspace argument and without any
 * allocation: no counterpart in the reference.  It is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The light-power partial
 * sums go to a 4 MiB device array inside the library (one per GPU, held
 * until the process ends), so res is limited to 4096 (larger res returns
 * SVBRDF_ESHAPE: use svbrdf_render_backward with a workspace).  The first
 * call on a device resolves that array's address (hipGetSymbolAddress, no
 * allocation).  All calls on one GPU share the array: calls that overlap on
 * one GPU (concurrent streams) race on it and must be ordered by the caller.
 * Results are bit-identical to svbrdf_render_backward. */
int svbrdf_render_backward_noalloc(const float *textures, int res, int n_views, float size,
                                   const float *camera_pos, const float *light_pos,
                                   const float *light_pow, const float *grad_images,
                                   float *grad_textures, float *grad_light_pow, void *stream);

/* Target preparation for the fused step (once per target set).  For
 * SVBRDF_LOSS_LOG_L2 writes ln(clamp(targets, 1e-6^(1/2.2), 1)); for
 * SVBRDF_LOSS_L2 copies.  `prepared` may alias `targets`. */
int svbrdf_prepare_targets(const float *targets, size_t count, int loss_mode,
                           float *prepared, void *stream);

/* Resets a svbrdf_step_state to "no step taken" for the given betas (as
 * torch.optim.Adam's lazily created state, torch/optim/adam.py:142-164).
 * beta^t is advanced by one multiplication per step, so the betas passed to
 * svbrdf_fused_step must stay the ones given here. */
int svbrdf_step_state_init(void *state, double beta1, double beta2, void *stream);

/* One fused per-pixel optimisation iteration: the body of the loop of
 * SvbrdfOptim.optim (src/svbrdf.py:58-71):
 *   clamp(textures,-1,1) -> render -> image loss -> backward -> Adam step,
 * with Adam = torch.optim.Adam(lr, betas, eps), no weight decay, no amsgrad.
 * textures / exp_avg / exp_avg_sq ([9,R,R]) are updated in place; the stored
 * textures stay unclamped (the clamp is applied on read, with torch's
 * inclusive-bound gradient mask).  targets_prepared comes from
 * svbrdf_prepare_targets with the same loss_mode.  When optim_light is
 * non-zero light_pow[3] is optimised too (src/svbrdf.py:48-50,58-59) with
 * its Adam state in `state`.
 *
 * Losses: the loss of step t (1-based) goes to state->loss and, when
 * loss_history is not NULL and t <= history_len, to loss_history[t - 1]
 * (device memory), so a loop keeps every loss without a host
 * synchronisation per iteration (the reference reads loss.item() each
 * epoch, src/svbrdf.py:64).  Without light optimisation the call is ONE
 * kernel launch and the loss of step t is reduced by the launch of step
 * t + 1 or by svbrdf_fused_step_flush(); with optim_light a one-block
 * finalize follows the step kernel (the next step's light power needs this
 * step's full dL/dP) and records the loss at once.
 *
 * Layout: every plane of textures ([9] planes), exp_avg, exp_avg_sq ([9]) and
 * targets_prepared ([3 V]) holds R*R floats; consecutive planes of textures
 * start texture_stride floats apart, those of the other three plane_stride
 * floats apart (0 means R*R: dense [.., R, R] tensors as torch holds them).
 * Buffers the caller allocates for the fused step alone should use
 * svbrdf_plane_stride(res): at R = 1024 the dense plane spacing is exactly
 * 4 MiB, and the 54 planes a wave touches then fall on the same HBM
 * channels (about 7 % slower).
 *
 * Workspace: svbrdf_fused_step_workspace_bytes(res) bytes, 256-byte aligned,
 * zero-filled before the first call and then kept with `state` (it holds
 * arrival counters, which every launch leaves at zero, and the partial sums
 * of the last step).  One state + workspace pair serves one problem (one
 * res); calls that share a pair must be stream-ordered. */
size_t svbrdf_fused_step_workspace_bytes(int res);
/* Plane stride (in floats) for buffers laid out for the fused step: R*R
 * rounded up to 64 floats (256-byte aligned planes) plus 448 floats. */
size_t svbrdf_plane_stride(int res);
int svbrdf_fused_step(float *textures, float *exp_avg, float *exp_avg_sq,
                      const float *targets_prepared, int res, size_t texture_stride,
                      size_t plane_stride, int n_views, float size,
                      const float *camera_pos, const float *light_pos, float *light_pow,
                      int loss_mode, int optim_light, double lr, double beta1,
                      double beta2, double eps, void *state, float *loss_history,
                      int history_len, void *workspace, size_t workspace_bytes,
                      void *stream);

/* Records the loss of the last step taken by svbrdf_fused_step when it is
 * still pending (no-op otherwise): state->loss and loss_history[t - 1].
 * Call it before reading losses on the host (no counterpart in the
 * reference, where the loss is read synchronously each epoch). */
int svbrdf_fused_step_flush(void *state, int res, int n_views, float *loss_history,
                            int history_len, const void *workspace,
                            size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------
 * Map / target codec: the per-pixel work between PNG files and the tensors
 * above (src/imageio.py:11-76, src/svbrdf.py:126-132,150-221).  The PNG
 * container (inflate / deflate, scanline filters) stays on the host; these
 * take and produce 8-bit images in device memory, RGB-ordered and
 * channel-interleaved ([H, W, C] bytes), as the files store them.
 * ------------------------------------------------------------------------ */

/* imresize (src/imageio.py:75-76): cv2.resize(..., INTER_LANCZOS4) of an
 * 8-bit [src_h, src_w, channels] image (channels 1..4) to [dst_h, dst_w,
 * channels]: 8x8 taps, replicated border, 11-bit fixed-point weights,
 * rounded and saturated; a same-size resize is a copy.  Sides up to 16384. */
int svbrdf_resize_lanczos4_u8(const unsigned char *src, int src_h, int src_w, int channels,
                              unsigned char *dst, int dst_h, int dst_w, void *stream);

/* Coarse-to-fine stage chaining in memory: the reference chains stages
 * through 8-bit PNG files and a Lanczos-4 resize (src/scripts.py:83-85,
 * src/svbrdf.py:150-166, src/imageio.py:75-76; svbrdf_resize_lanczos4_u8
 * above).  This is the float alternative with the semantics of
 * torch.nn.functional.interpolate(scale_factor=2, mode="bilinear",
 * align_corners=False): n_planes planes of res x res floats, src_stride
 * floats apart (0: res^2), to planes of 2res x 2res, dst_stride floats apart
 * (0: 4 res^2; the next stage's svbrdf_plane_stride(2 res) layout fits).
 * res up to 8192, n_planes up to 64; the buffers must not overlap. */
int svbrdf_upsample2x_bilinear(const float *src, int res, size_t src_stride, float *dst, size_t dst_stride,
                               int n_planes, void *stream);

/* imread flags (src/imageio.py:26-47). */
enum svbrdf_imread_flag {
    SVBRDF_IMREAD_NONE = 0,    /* u / 255 per channel                               */
    SVBRDF_IMREAD_SRGB = 1,    /* 3 channels, u / 255                               */
    SVBRDF_IMREAD_ROUGH = 2,   /* 1 channel, or the float mean of 3, of u / 255      */
    SVBRDF_IMREAD_NORMAL = 3,  /* 3 channels, u / 255 * 2 - 1 divided by its 2-norm  */
};

/* imread after the resize (src/imageio.py:11-51; SvbrdfIO.load_images_th,
 * src/svbrdf.py:191-204, is n "srgb" images): n images [n, height, width,
 * channels] bytes -> [n, C', height, width] float32, C' = channels, 1 for
 * SVBRDF_IMREAD_ROUGH, 3 for SVBRDF_IMREAD_NORMAL. */
int svbrdf_images_from_u8(const unsigned char *images, int n, int height, int width, int channels,
                          int flag, float *out, void *stream);

/* SvbrdfIO.load_textures_th (src/svbrdf.py:150-166) after the resizes: the
 * diffuse, normal and specular maps [res, res, 3] and the roughness map
 * [res, res, rgh_channels] (1, or 3 averaged) -> textures [9, res, res]:
 * u / 255 * 2 - 1 for diffuse / specular / roughness, the normal decoded as
 * u / 255 * 2 - 1 and divided by its 2-norm, channels 0:3 diffuse, 3:5
 * normal xy, 5 roughness, 6:9 specular. */
int svbrdf_textures_from_u8(const unsigned char *dif, const unsigned char *nom,
                            const unsigned char *spe, const unsigned char *rgh,
                            int rgh_channels, int res, float *textures, void *stream);

/* SvbrdfIO.save_textures_th (src/svbrdf.py:168-189) before the PNG encode:
 * textures [9, res, res] -> dif / nom / spe [res, res, 3] and rgh
 * [res, res] bytes: (t + 1) / 2 clipped to [0, 1] (normal: reconstructed
 * as SvbrdfIO.reconstruct_normal, clipped to [-1, 1], then (n + 1) / 2),
 * times 255, truncated (src/imageio.py:54-72). */
int svbrdf_textures_to_u8(const float *textures, int res, unsigned char *dif,
                          unsigned char *nom, unsigned char *spe, unsigned char *rgh,
                          void *stream);

/* imwrite before the PNG encode (src/imageio.py:54-72; SvbrdfIO.save_images_th,
 * src/svbrdf.py:206-221, is n "srgb" images): [n, channels, height, width]
 * float32 -> [n, height, width, channels] bytes = (v * 255) truncated, with v
 * clipped to [0, 1] (SRGB, ROUGH with 1 channel, and NONE, whose unclipped
 * reference cast is only defined there) or (clip(v, -1, 1) + 1) / 2 (NORMAL). */
int svbrdf_images_to_u8(const float *images, int n, int height, int width, int channels,
                        int flag, unsigned char *out, void *stream);

/* ------------------------------------------------------------------------
 * Batched step: many independent materials (uneven resolution and view count)
 * in one launch.  The reference runs one material per process
 * (src/scripts.py:67-97); this is the same per-pixel iteration applied to
 * every material of a batch at once.  Each material is one svbrdf_material
 * (device pointers, host-side sizes); the host plans tiles once, copies the
 * table and the two tile arrays to device memory, and then calls
 * svbrdf_batch_step once per iteration.  Tiles run in cost order (views x
 * pixels, largest first), either one block per 256-pixel tile or from an
 * atomic ticket counter in a persistent grid; per-material losses are reduced
 * in a fixed order (bit-reproducible).
 * Light-power optimisation is not available in batched mode.
 */
typedef struct svbrdf_material {
    float *textures;                /* [9,R,R], updated in place                   */
    float *exp_avg;                 /* [9,R,R], zero before the first step          */
    float *exp_avg_sq;              /* [9,R,R], zero before the first step          */
    const float *targets_prepared;  /* [V,3,R,R] from svbrdf_prepare_targets        */
    const float *camera_pos;        /* [V,3] cm                                     */
    const float *light_pos;         /* [V,3] cm                                     */
    float *light_pow;               /* [3]; updated in place when the batch
                                       optimises the light power             */
    int res;                        /* R                                            */
    int n_views;                    /* V                                            */
    float size;                     /* plane width, cm                              */
    unsigned int tile_begin;        /* set by svbrdf_batch_plan                     */
    unsigned int n_tiles;           /* set by svbrdf_batch_plan                     */
    unsigned int plane_stride;      /* floats between the planes of exp_avg,
                                       exp_avg_sq and targets_prepared; 0 = R*R  */
    unsigned int texture_stride;    /* floats between the planes of textures;
                                       0 = R*R                                   */
    float loss_scale;               /* set by svbrdf_batch_plan                   */
    float *light_adam;              /* [6]: Adam's exp_avg[3], exp_avg_sq[3] of
                                       light_pow, zero before the first step;
                                       read when the batch optimises the light
                                       power (NULL: this material's light power
                                       stays fixed)                          */
} svbrdf_material;

/* Number of tiles of `tile_pixels` pixels (a multiple of 256) the batch splits
 * into; 0 on invalid input. */
size_t svbrdf_batch_tile_count(const svbrdf_material *materials, int n_materials, int tile_pixels);

/* Host-side plan: fills tile_begin / n_tiles of every material (host array)
 * and writes the execution order (tile ids, costliest first) and the tile ->
 * material map, each of svbrdf_batch_tile_count() entries. */
int svbrdf_batch_plan(svbrdf_material *materials, int n_materials, int tile_pixels, unsigned int *order,
                      unsigned int *tile_material, size_t n_tiles);

/* Workspace: per-tile partial sums plus the ticket counter; must be zeroed
 * once before the first step (each step leaves the counter at zero). */
size_t svbrdf_batch_workspace_bytes(size_t n_tiles);

/* Number of blocks that fit on the device at once for a kernel variant
 * (pixels_per_thread in {1, 2}, views_per_trip in {1, 3}, with or without the
 * light-power gradient): the grid size of a persistent launch is
 * blocks_per_cu x compute units. */
int svbrdf_batch_max_blocks_per_cu(int loss_mode, int optim_light, int pixels_per_thread, int views_per_trip);

/* One optimisation iteration of every material of the batch (device copies of
 * the table and tile arrays).  losses[m] receives material m's loss;
 * loss_history (optional) is [history_len][n_materials].  state is shared by
 * the batch (all materials step together).  grid_blocks > 0 launches that
 * many persistent blocks drawing tiles from the ticket counter; grid_blocks
 * = 0 launches one block per tile instead (tile_pixels = 256, one pixel per
 * thread), in cost order, and leaves the scheduling to the dispatcher.
 * optim_light != 0 also takes each material's Adam step on its light power
 * (SvbrdfOptim.optim(optim_light=True), src/svbrdf.py:48-50,58-59, per
 * material), with the moments in its light_adam. */
int svbrdf_batch_step(const svbrdf_material *materials, int n_materials, const unsigned int *order,
                      const unsigned int *tile_material, size_t n_tiles, int tile_pixels,
                      int pixels_per_thread, int views_per_trip, int grid_blocks, int loss_mode,
                      int optim_light, double lr, double beta1, double beta2, double eps, void *state, float *losses,
                      float *loss_history, int history_len, void *workspace, size_t workspace_bytes,
                      void *stream);

/* ------------------------------------------------------------------------
 * Convolutions of the MaterialGAN caller's networks (src/materialgan.py:21-173)
 * on the matrix cores, fp32 as the reference computes them: the 3x3
 * modulated / demodulated convolutions of the StyleGAN2 synthesis network
 * (ModulateConvBlock, src/higan_models/stylegan2_generator_network.py:585-760)
 * and the 3x3 convolutions of the VGG-19 feature loss (src/descriptor.py:7-79).
 * ------------------------------------------------------------------------ */

enum svbrdf_activation {
    SVBRDF_ACT_NONE = 0,
    SVBRDF_ACT_RELU = 1,    /* max(v, 0): VGG (src/descriptor.py:7-79)                              */
    SVBRDF_ACT_LRELU = 2,   /* leaky ReLU (slope 0.2) x sqrt(2): ModulateConvBlock's activation     */
};

/* Weights [cout, cin, 3, 3] -> the layout svbrdf_conv3x3 reads, [9][cin][cout]
 * (tap = 3 kh + kw), each value scale * style[ci] * w (style may be NULL: 1)
 * and, when demodulate is non-zero, divided per output channel by
 * sqrt(sum over ci, taps of that value^2 + 1e-8) (the style modulation and
 * demodulation of ModulateConvBlock.forward, stylegan2_generator_network.py:694-706).
 * transpose != 0 writes the input-gradient weights instead: [9][cout][cin]
 * with the taps flipped (8 - tap), so that svbrdf_conv3x3 of dL/dy with them
 * (cin and cout exchanged) is dL/dx. */
int svbrdf_conv3x3_pack(const float *weight, int cout, int cin, const float *style, float scale, int demodulate,
                        int transpose, float *packed, void *stream);

/* dL/d(pre-activation) of a sign-keeping activation from dL/dy and the
 * activation's output y: out = grad_y * (y > 0 ? pos : neg), n elements (ReLU:
 * pos 1, neg 0; leaky ReLU 0.2 x sqrt(2): sqrt(2), 0.2 sqrt(2)) -- the
 * backward of the activations the reference's conv blocks apply
 * (src/higan_models/stylegan2_generator_network.py:738-755,
 * src/descriptor.py:7-79), which it
 * leaves to torch autograd.  out may alias grad_y. */
int svbrdf_act_grad(const float *grad_y, const float *y, size_t n, float pos, float neg, float *out,
                    void *stream);

/* dL/dstyle of the weights svbrdf_conv3x3_pack writes with a style and
 * scale (and demodulation when demodulate != 0), given grad_weight = dL/d(the
 * packed weight) in the natural [cout, cin, 3, 3] layout (what
 * svbrdf_conv3x3_wgrad returns): the backward of the modulation and
 * demodulation of ModulateConvBlock.forward,
 * src/higan_models/stylegan2_generator_network.py:694-706, which the
 * reference leaves to torch autograd.  grad_style [cin] is
 * overwritten.  workspace: 8-byte aligned, at least 2 cout floats. */
int svbrdf_conv3x3_modulation_grad(const float *weight, int cout, int cin, const float *style, float scale,
                                   int demodulate, const float *grad_weight, float *grad_style, void *workspace,
                                   size_t workspace_bytes, void *stream);

/* y = act(conv3x3(x, w) + noise_strength * noise + bias), stride 1, zero
 * padding 1 (F.conv2d(x, w, padding=1) as the reference's conv layers call
 * it): x [n, cin, height, width], packed from svbrdf_conv3x3_pack, bias
 * [cout] or NULL, noise [height, width] (shared by the n images) or NULL, y
 * [n, cout, height, width].  The noise and bias terms and the activation are
 * ModulateConvBlock's (stylegan2_generator_network.py:738-755) or VGG's
 * (bias, ReLU).  Shapes with few output tiles (small images, many input
 * channels) split the channel sum into slices whose partial sums need a
 * workspace of svbrdf_conv3x3_workspace_bytes(n, cin, height, width, cout)
 * bytes (0 for most shapes: workspace may then be NULL; a smaller one
 * returns SVBRDF_ESPACE).  The slices are added in a fixed order: results do
 * not depend on scheduling. */
size_t svbrdf_conv3x3_workspace_bytes(int n, int cin, int height, int width, int cout);
int svbrdf_conv3x3(const float *x, int n, int cin, int height, int width, const float *packed, int cout,
                   const float *bias, const float *noise, float noise_strength, int activation, float *y,
                   void *workspace, size_t workspace_bytes, void *stream);

/* dL/dx of y = act(conv3x3(x, w) + ...): the convolution of dL/dy [n, cout,
 * height, width] gated by the activation's derivative at y (the layer's
 * forward output: 1 / 0 for ReLU, sqrt(2) / 0.2 sqrt(2) for leaky ReLU by the
 * sign of y; act = SVBRDF_ACT_NONE ignores y) with the flipped, transposed
 * weights (svbrdf_conv3x3_pack with transpose != 0), written to grad_x [n,
 * cin, height, width].  torch autograd's backward of F.conv2d + the
 * activation (src/descriptor.py:7-79 for VGG's ReLU layers) in one pass:
 * the gate is applied while the kernel stages its input.  Workspace:
 * svbrdf_conv3x3_workspace_bytes(n, cout, height, width, cin). */
int svbrdf_conv3x3_input_grad(const float *grad_y, const float *y, int activation, int n, int cout, int height,
                              int width, const float *packed_t, int cin, float *grad_x, void *workspace,
                              size_t workspace_bytes, void *stream);

/* dL/dw [cout, cin, 3, 3] of a 3x3 convolution (stride 1, zero padding 1)
 * from its input x [n, cin, height, width] and grad_z [n, cout, height,
 * width], the gradient of the loss with respect to the convolution's output
 * before bias and activation (torch.nn.grad.conv2d_weight, the weight
 * gradient torch autograd takes through F.conv2d; for the modulated weights
 * of ModulateConvBlock, stylegan2_generator_network.py:694-755, it then flows
 * to the style and the latent).  Overwrites grad_w.  Many pixels per channel
 * tile split the pixel sum into slices: workspace of
 * svbrdf_conv3x3_wgrad_workspace_bytes(n, cin, height, width, cout) bytes (0:
 * workspace may be NULL); the slices are added in a fixed order. */
size_t svbrdf_conv3x3_wgrad_workspace_bytes(int n, int cin, int height, int width, int cout);
int svbrdf_conv3x3_wgrad(const float *x, const float *grad_z, int n, int cin, int height, int width, int cout,
                         float *grad_w, void *workspace, size_t workspace_bytes, void *stream);

/* 2 x 2 average pooling with stride 2 of `planes` planes of height x width
 * (y [planes, height / 2, width / 2], a trailing odd row / column dropped):
 * the AvgPool2d(kernel_size=2) the VGG feature loss uses in place of max
 * pooling (src/descriptor.py:15-18), bit-identical to torch's (window summed
 * row by row, then / 4); and its adjoint (grad_x [planes, height, width]). */
int svbrdf_avgpool2(const float *x, int planes, int height, int width, float *y, void *stream);
int svbrdf_avgpool2_adjoint(const float *grad_y, int planes, int height, int width, float *grad_x, void *stream);

/* x2 upsampling of `planes` planes of height x width: zero insertion and the
 * [1, 3, 3, 1] binomial filter with gain 4 (UpsamplingLayer,
 * stylegan2_generator_network.py:419-479, as the synthesis network's x2
 * resolution steps use it), y [planes, 2 height, 2 width]; and its adjoint
 * (grad_y [planes, 2 height, 2 width] -> grad_x [planes, height, width]). */
int svbrdf_upsample2_blur(const float *x, int planes, int height, int width, float *y, void *stream);
int svbrdf_upsample2_blur_adjoint(const float *grad_y, int planes, int height, int width, float *grad_x,
                                  void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SVBRDF_HIP_H */
