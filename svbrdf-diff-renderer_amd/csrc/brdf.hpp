// brdf.hpp -- per-pixel microfacet shading and its adjoint, gfx950 device code.
//
// One thread owns one pixel: it decodes the 9 texture channels once, walks
// the views, and keeps every per-view sum (gradients over views and RGB) in
// registers, so no [V,3,R,R] intermediate ever reaches memory.
//
// Math follows src/microfacet.py:
//   decode      :64-79   u = (t+1)/2, map = u^2.2; normal from (nx, ny)
//   geometry    :16-24, :57-62, :91-93 (v, l, |l|^2, h)
//   dots        :54-55, :96-99 (clamped at 0)
//   GGX         :28-32   D = a2 / (pi den^2 + eps), a2 = rgh^4
//   Fresnel     :43-45   F = spe + (1-spe) 2^((-5.55473 c - 6.98316) c)
//   Smith       :47-52   G = G1(n.v) G1(n.l), k = rgh^2/2 + eps
//   radiance    :101-117 img = P (dif/pi (1-spe) + D F G/(4 nv nl + eps)) nl / d2
//   output      :120     clamp(img, eps, 1)^(1/2.2)
// and the backward follows torch autograd's conventions for that op
// sequence (inclusive clamp masks, sqrt' = 1/(2 sqrt), full Jacobian of the
// normal normalisation, expand-backward sums over views and channels).
//
// Transcendentals use the hardware instructions (v_rcp/v_rsq/v_sqrt/v_exp/
// v_log, ~1 ulp); every input that reaches them is either 0 or a normal
// number (u >= 2^-25 or 0, img >= 1e-6), so no denormal handling is needed.
#pragma once
#include <hip/hip_runtime.h>

namespace svbrdf {

constexpr float kEps = 1e-6f;                       // microfacet.py:14
constexpr float kOneMinusEps = 0.999999f;           // 1 - eps, microfacet.py:67
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kGamma = 2.2f;                      // microfacet.py:74-77
constexpr float kGammaM1 = 1.2f;
constexpr float kInvGamma = 0.454545454545454545f;  // 1/2.2, microfacet.py:120
constexpr float kLn2 = 0.693147180559945309f;
constexpr float kFresA = -5.55473f;                 // microfacet.py:44
constexpr float kFresB = -6.98316f;
constexpr float kLog2E = 1.44269504088896340736f;

// BRDF term selection (bit set).  0 = the terms Microfacet.eval uses: GGX
// (microfacet.py:28-32) and the spherical-Gaussian Fresnel (:43-45).
constexpr int kTermsBeckmann = 1;   // Beckmann NDF instead of GGX (:34-38)
constexpr int kTermsSchlick = 2;    // Fresnel_f0(v.h, specular) = spe + (1-spe)(1-v.h)^5 (:40-41)

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }
// torch.clamp semantics: NaN passes through (fminf/fmaxf would drop it).
__device__ __forceinline__ float tclamp(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ float tmax0(float x) { return x < 0.0f ? 0.0f : x; }
// The same clamp (NaN passes through, as torch.clamp lets it) as a median of
// three plus a NaN select: three instructions instead of four.
__device__ __forceinline__ float tclamp_med3(float x, float lo, float hi) {
    const float r = __builtin_amdgcn_fmed3f(x, lo, hi);
    return x != x ? x : r;
}

// Pixel-centre position, microfacet.py:16-19: ((j + 0.5) / R - 0.5) * size,
// with an IEEE division so non-power-of-two R rounds as the reference does.
__device__ __forceinline__ float pixel_coord(int j, int res, float size) {
    return (((float)j + 0.5f) / (float)res - 0.5f) * size;
}

// Decoded maps of one pixel (values the per-view loop reads).
struct Maps {
    float dif[3], spe[3], omspe[3], f1[3];
    float rgh, alpha, a2, k, omk;
    float n[3];
    float Fc[3];   // Fresnel term when light and camera coincide (v.h = 1)
};

// 2^((A v.h + B) v.h) at v.h = 1 (co-located light and camera): 2 to the
// fp32 sum -12.5378895 of the two constants, correctly rounded (the value
// torch.pow gives for the reference's expression at v.h = 1), a constant
// instead of one v_exp_f32 per pixel.
constexpr float kSphgColoc = 0x1.60a77cp-13f;
__device__ __forceinline__ float sphg_coloc() { return kSphgColoc; }

// u = (t+1)/2 as one fma: the halving is exact, so fma(t, 1/2, 1/2) rounds
// t/2 + 1/2 once, to the value of the reference's (t + 1)/2.
__device__ __forceinline__ float map_unit(float t) { return fmaf(t, 0.5f, 0.5f); }

// u^2.2 of u = (t+1)/2; u in {0} or >= 2^-25, so log2 is finite or -inf.
__device__ __forceinline__ float map_pow(float t) {
    const float u = map_unit(t);
    return fexp2(kGamma * flog2(u));
}

// d(u^2.2)/dt = 2.2 u^1.2 / 2 (torch pow_backward: exponent * x^(exponent-1)).
__device__ __forceinline__ float map_pow_grad(float t) {
    const float u = map_unit(t);
    return 0.5f * kGamma * fexp2(kGammaM1 * flog2(u));
}

__device__ __forceinline__ void decode(const float t[9], Maps &m) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        m.dif[c] = map_pow(t[c]);
        m.spe[c] = map_pow(t[6 + c]);
        m.omspe[c] = 1.0f - m.spe[c];
        m.f1[c] = m.dif[c] * kInvPi * m.omspe[c];
    }
    const float sc = sphg_coloc();
#pragma unroll
    for (int c = 0; c < 3; ++c) m.Fc[c] = m.spe[c] + m.omspe[c] * sc;
    m.rgh = map_pow(t[5]);
    m.alpha = m.rgh * m.rgh;
    m.a2 = m.alpha * m.alpha;
    m.k = m.alpha * 0.5f + kEps;
    m.omk = 1.0f - m.k;
    // reconstruct_normal, microfacet.py:64-70
    const float nx = tclamp(t[3], -1.0f, 1.0f);
    const float ny = tclamp(t[4], -1.0f, 1.0f);
    const float s = tclamp(nx * nx + ny * ny, 0.0f, kOneMinusEps);
    const float nz = fsqrt(1.0f - s);
    const float il = frsq(nx * nx + ny * ny + nz * nz);
    m.n[0] = nx * il;
    m.n[1] = ny * il;
    m.n[2] = nz * il;
}

// Per-view geometry of one pixel: v, l, h, 1/|l|^2 and clamp(v.h, 0).
struct Geom {
    float v[3], l[3], h[3];
    float inv_d2, vh;
};

// COLOC (camera == light, true for every view of a launch, decided on the
// device before the view loop) makes l = v and h = v/|v| exactly as the
// general formulas would compute them, with fewer ops.
template <bool COLOC>
__device__ __forceinline__ void geometry(float px, float py, float cx, float cy, float cz,
                                         float lx, float ly, float lz, Geom &g) {
    const float vx = cx - px, vy = cy - py, vz = cz;
    const float sv = vx * vx + vy * vy + vz * vz;
    const float rv = frsq(sv);
    g.v[0] = vx * rv; g.v[1] = vy * rv; g.v[2] = vz * rv;
    if (COLOC) {
        g.l[0] = g.v[0]; g.l[1] = g.v[1]; g.l[2] = g.v[2];
        g.inv_d2 = frcp(sv);
        const float w = g.v[0] * g.v[0] + g.v[1] * g.v[1] + g.v[2] * g.v[2];
        const float rh = frsq(w);
        g.h[0] = g.v[0] * rh; g.h[1] = g.v[1] * rh; g.h[2] = g.v[2] * rh;
        g.vh = tmax0(w * rh);
    } else {
        const float ax = lx - px, ay = ly - py, az = lz;
        const float sl = ax * ax + ay * ay + az * az;
        const float rl = frsq(sl);
        g.l[0] = ax * rl; g.l[1] = ay * rl; g.l[2] = az * rl;
        g.inv_d2 = frcp(sl);
        const float hx = g.l[0] + g.v[0], hy = g.l[1] + g.v[1], hz = g.l[2] + g.v[2];
        const float rh = frsq(hx * hx + hy * hy + hz * hz);
        g.h[0] = hx * rh; g.h[1] = hy * rh; g.h[2] = hz * rh;
        g.vh = tmax0(g.v[0] * g.h[0] + g.v[1] * g.h[1] + g.v[2] * g.h[2]);
    }
}

// Forward BRDF terms of one (pixel, view).
struct Shade {
    float nv_raw, nl_raw, nh_raw, nv, nl, nh;
    float c2, den, iqd, D;
    float omsphg, F[3];
    float idv, idl, g1v, g1l, G, iq, DGq;
    float nl_d2, f[3], img[3];
};

template <bool COLOC, int TERMS = 0>
__device__ __forceinline__ void shade(const Maps &m, const Geom &g, const float P[3], Shade &s) {
    s.nv_raw = m.n[0] * g.v[0] + m.n[1] * g.v[1] + m.n[2] * g.v[2];
    if (COLOC) {
        s.nl_raw = s.nv_raw;
    } else {
        s.nl_raw = m.n[0] * g.l[0] + m.n[1] * g.l[1] + m.n[2] * g.l[2];
    }
    s.nh_raw = m.n[0] * g.h[0] + m.n[1] * g.h[1] + m.n[2] * g.h[2];
    s.nv = tmax0(s.nv_raw);
    s.nl = tmax0(s.nl_raw);
    s.nh = tmax0(s.nh_raw);
    // GGX(n.h, rgh^2): den = c2 a2 + (1 - c2).  Near the highlight n.h -> 1
    // and 1 - c2 cancels catastrophically (a 1-ulp error in n.h becomes a
    // ~1e-4 relative error of D at n.h = 0.9994); for unit n, h it equals
    // |n x h|^2, which has no cancellation.
    s.c2 = s.nh * s.nh;
    const float xx = m.n[1] * g.h[2] - m.n[2] * g.h[1];
    const float xy = m.n[2] * g.h[0] - m.n[0] * g.h[2];
    const float xz = m.n[0] * g.h[1] - m.n[1] * g.h[0];
    const float one_m_c2 = s.nh_raw < 0.0f ? 1.0f : xx * xx + xy * xy + xz * xz;
    if constexpr (TERMS & kTermsBeckmann) {
        // Beckmann: exp(-t2/a2) / (pi a2 c2^2), t2 = (1 - c2)/c2 (den holds t2,
        // iqd the reciprocal of the denominator)
        s.den = one_m_c2 * frcp(s.c2);
        s.iqd = frcp(kPi * m.a2 * (s.c2 * s.c2));
        s.D = fexp2(-s.den * frcp(m.a2) * kLog2E) * s.iqd;
    } else {
        s.den = s.c2 * m.a2 + one_m_c2;
        s.iqd = frcp(kPi * (s.den * s.den) + kEps);
        s.D = m.a2 * s.iqd;
    }
    // Fresnel(v.h, spe): spe + (1 - spe) sphg
    float sphg;
    if constexpr (TERMS & kTermsSchlick) {
        const float o = 1.0f - g.vh;
        const float o2 = o * o;
        sphg = o2 * o2 * o;
    } else {
        sphg = fexp2((kFresA * g.vh + kFresB) * g.vh);
    }
    s.omsphg = 1.0f - sphg;
    // Smith(n.v, n.l, rgh^2)
    s.idv = frcp(s.nv * m.omk + m.k);
    s.g1v = s.nv * s.idv;
    if (COLOC) {
        s.idl = s.idv;
        s.g1l = s.g1v;
    } else {
        s.idl = frcp(s.nl * m.omk + m.k);
        s.g1l = s.nl * s.idl;
    }
    s.G = s.g1v * s.g1l;
    s.iq = frcp(4.0f * s.nv * s.nl + kEps);
    s.DGq = s.D * s.G * s.iq;
    s.nl_d2 = s.nl * g.inv_d2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.F[c] = m.spe[c] + m.omspe[c] * sphg;
        s.f[c] = m.f1[c] + s.DGq * s.F[c];
        s.img[c] = P[c] * s.f[c] * s.nl_d2;
    }
}

// Per-pixel gradient accumulators (sums over views).
struct Accum {
    float S[3];    // sum_v dL/df_c            -> diffuse and the f1 part of specular
    float T[3];    // sum_v dL/dF_c (1 - sphg) -> Fresnel part of specular
    float A;       // sum_v dL/dalpha          -> roughness
    float gn[3];   // sum_v dL/dn              -> normal
    float gP[3];   // sum_v dL/dP_c            -> light power
};

__device__ __forceinline__ void accum_zero(Accum &a) {
#pragma unroll
    for (int c = 0; c < 3; ++c) { a.S[c] = 0.f; a.T[c] = 0.f; a.gn[c] = 0.f; a.gP[c] = 0.f; }
    a.A = 0.f;
}

// Backward of shade() for one view given dL/dimg_c (already masked by the
// output clamp).  Accumulates into `a`.
template <bool LIGHT_GRAD, int TERMS = 0>
__device__ __forceinline__ void shade_backward(const Maps &m, const Geom &g, const Shade &s,
                                               const float P[3], const float gimg[3], Accum &a) {
    float W = 0.f, gnl = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float gf = gimg[c] * P[c] * s.nl_d2;
        a.S[c] += gf;
        a.T[c] += gf * s.DGq * s.omsphg;
        W += gf * s.F[c];
        gnl += gimg[c] * P[c] * s.f[c];
        if (LIGHT_GRAD) a.gP[c] += gimg[c] * s.f[c] * s.nl_d2;
    }
    gnl *= g.inv_d2;
    // f2 = D F G / q, q = 4 nv nl + eps
    const float gD = W * s.G * s.iq;
    const float gG = W * s.D * s.iq;
    const float gq = -W * s.DGq * s.iq;
    // G = G1(nv) G1(nl), G1(c) = c / (c (1-k) + k)
    const float dg1v_dc = m.k * s.idv * s.idv;
    const float dg1l_dc = m.k * s.idl * s.idl;
    float gnv = gq * 4.0f * s.nl + gG * s.g1l * dg1v_dc;
    gnl += gq * 4.0f * s.nv + gG * s.g1v * dg1l_dc;
    const float gk = -gG * (s.g1l * s.nv * (1.0f - s.nv) * s.idv * s.idv +
                            s.g1v * s.nl * (1.0f - s.nl) * s.idl * s.idl);
    float ga2, gc2;
    if constexpr (TERMS & kTermsBeckmann) {
        // D = exp(-t2/a2) / (pi a2 c2^2), t2 = (1 - c2)/c2:
        // dD/dc2 = D (1/(a2 c2^2) - 2/c2), dD/da2 = D (t2/a2^2 - 1/a2)
        const float ic2 = frcp(s.c2), ia2 = frcp(m.a2);
        gc2 = gD * s.D * (ia2 * ic2 * ic2 - 2.0f * ic2);
        ga2 = gD * s.D * (s.den * ia2 * ia2 - ia2);
    } else {
        // D = a2 / (pi den^2 + eps), den = c2 a2 + 1 - c2
        const float dD_dden = -2.0f * kPi * m.a2 * s.den * s.iqd * s.iqd;
        ga2 = gD * (s.iqd + dD_dden * s.c2);
        gc2 = gD * dD_dden * (m.a2 - 1.0f);
    }
    float gnh = gc2 * 2.0f * s.nh;
    // alpha = rgh^2 feeds k = alpha/2 + eps and a2 = alpha^2
    a.A += gk * 0.5f + ga2 * 2.0f * m.alpha;
    // clamp_min(0) masks (inclusive)
    gnv = s.nv_raw >= 0.0f ? gnv : 0.0f;
    gnl = s.nl_raw >= 0.0f ? gnl : 0.0f;
    gnh = s.nh_raw >= 0.0f ? gnh : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) a.gn[c] += gnv * g.v[c] + gnl * g.l[c] + gnh * g.h[c];
}

// ---------------------------------------------------------------------------
// Co-located light and camera (the flash setup of every capture the reference
// handles): l = v and h = (l + v)/|l + v| = v, so n.l = n.h = n.v and
// v.h = |v| = 1 up to the rounding of the normalisation (~1e-7, below the FP32
// noise of the reference's own evaluation).  The Fresnel factor is then the
// pixel constant Maps::Fc, and the three dot-product gradients act along v.
struct ShadeC {
    float v[3];
    float nv_raw, nv, c2, den, iqd, D, idv, g1, iq, Dq1, DGq, inv_d2, nl_d2;
};

__device__ __forceinline__ void shade_coloc(const Maps &m, float px, float py, float cx, float cy, float cz,
                                            ShadeC &s) {
    const float vx = cx - px, vy = cy - py, vz = cz;
    const float sv = vx * vx + vy * vy + vz * vz;
    const float rv = frsq(sv);
    s.v[0] = vx * rv; s.v[1] = vy * rv; s.v[2] = vz * rv;
    // 1/d^2 as the square of the 1/d that normalises v, as the forward
    // render computes it: the adjoint's recomputed image, and so its clamp
    // mask, follows the forward's rounding
    s.inv_d2 = rv * rv;
    s.nv_raw = m.n[0] * s.v[0] + m.n[1] * s.v[1] + m.n[2] * s.v[2];
    s.nv = tmax0(s.nv_raw);
    // GGX with 1 - (n.h)^2 = |v - (n.v) n|^2 for unit n, v: the components
    // are O(sin) with O(ulp) absolute error, so there is no cancellation at
    // the highlight (1 - c^2 itself would lose all digits as c -> 1).  For
    // n.v < 0 the clamped c = 0 makes 1 - c^2 = 1, but every use of den is
    // then multiplied by c^2 = 0 (D G / q through g1 = 0, the gradient
    // through the masked dL/dimg = 0), so |w|^2 < 1 there changes nothing.
    s.c2 = s.nv * s.nv;
    const float wx = s.v[0] - s.nv_raw * m.n[0];
    const float wy = s.v[1] - s.nv_raw * m.n[1];
    const float wz = s.v[2] - s.nv_raw * m.n[2];
    s.den = s.c2 * m.a2 + (wx * wx + wy * wy + wz * wz);
    s.iqd = frcp(kPi * (s.den * s.den) + kEps);
    s.D = m.a2 * s.iqd;
    // Smith: G = G1(n.v)^2
    s.idv = frcp(s.nv * m.omk + m.k);
    s.g1 = s.nv * s.idv;
    s.iq = frcp(4.0f * s.c2 + kEps);
    // D G / q as (D g1 / q) g1: the first factor is the adjoint's D G /(q n.v)
    s.Dq1 = s.D * s.iq * s.g1;
    s.DGq = s.Dq1 * s.g1;
    s.nl_d2 = s.nv * s.inv_d2;
}

// Forward-only co-located shading: DGq = D G / (4 (n.v)^2 + eps) and
// (n.l) / d^2, for a render that needs no derivatives.  The three quotients
// of GGX, Smith and the specular denominator become one:
//   DGq = a2 c2 / ((pi den^2 + eps) (nv (1 - k) + k)^2 (4 c2 + eps)),
// whose denominator factors lie in [1e-6, 3.2], [1e-12, 1] and [1e-6, 4]
// for maps in [-1, 1] (no underflow), and 1/d^2 is the square of the 1/d
// that normalises v.
__device__ __forceinline__ void shade_coloc_fwd(const Maps &m, float px, float py, float cx, float cy, float cz,
                                                float &DGq, float &nl_d2) {
    const float vx = cx - px, vy = cy - py, vz = cz;
    const float sv = vx * vx + vy * vy + vz * vz;
    const float rv = frsq(sv);
    const float v0 = vx * rv, v1 = vy * rv, v2 = vz * rv;
    const float nv_raw = m.n[0] * v0 + m.n[1] * v1 + m.n[2] * v2;
    const float nv = tmax0(nv_raw);
    const float c2 = nv * nv;
    // |w|^2 = 1 - c^2 also for n.v < 0, where c2 = 0 makes DGq = 0 (shade_coloc)
    const float wx = v0 - nv_raw * m.n[0], wy = v1 - nv_raw * m.n[1], wz = v2 - nv_raw * m.n[2];
    const float den = c2 * m.a2 + (wx * wx + wy * wy + wz * wz);
    const float sd = nv * m.omk + m.k;
    const float q = (kPi * (den * den) + kEps) * (sd * sd) * (4.0f * c2 + kEps);
    DGq = (m.a2 * c2) * frcp(q);
    nl_d2 = nv * (rv * rv);
}

// Backward of shade_coloc given dL/df_c (masked) and the direct dL/d(n.l) of
// img = P f (n.l) / d^2.  Accumulates into `a`; a.T collects dL/dF_c, to be
// scaled by dF/dspe = 1 - 2^(A+B) once per pixel (coloc_finish).  Fw are the
// Fresnel factors paired with gf: F_c, or P_c F_c when the caller passes
// gf / P_c (and scales a.S and a.T by P_c after the view loop).
// coloc_gradients is its second half: the normal and roughness terms from
// W = sum_c dL/df_c Fw_c and the direct dL/d(n.l), for callers that
// accumulate a.S and a.T themselves.
__device__ __forceinline__ void coloc_gradients(const Maps &m, const ShadeC &s, float W, float gnl, Accum &a);

__device__ __forceinline__ void backward_coloc(const Maps &m, const ShadeC &s, const float gf[3], float gnl,
                                               const float Fw[3], Accum &a) {
    float W;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.S[c] += gf[c];
        a.T[c] += gf[c] * s.DGq;
        W = c == 0 ? gf[c] * Fw[c] : W + gf[c] * Fw[c];
    }
    coloc_gradients(m, s, W, gnl, a);
}

__device__ __forceinline__ void coloc_gradients(const Maps &m, const ShadeC &s, float W, float gnl, Accum &a) {
    // f2 = D G F / q with D, G, q functions of n.v (= n.l = n.h), W =
    // sum_c dL/df_c F_c and WD = W D G / q.  With den and iqd of GGX and idv
    // of Smith:
    //   d ln D / d nv = 4 pi (1 - a2) den iqd nv     (dden/dnv = 2 (a2 - 1) nv)
    //   d ln G / d nv = 2 k idv / nv                 (G = g1^2, g1 = nv idv)
    //   d ln q / d nv = 8 nv iq
    // and WD nv = W Dq1 c2 idv, WD / nv = W Dq1 idv (Dq1 = D g1 / q), so
    //   dL/dnv = W Dq1 idv (c2 (4 pi (1 - a2) den iqd - 8 iq) + 2 k idv)
    // with no 1/n.v.  Roughness: alpha feeds a2 = alpha^2 and k = alpha/2 + eps,
    //   d ln D / d a2 = 1/a2 - 2 pi den iqd c2       (dden/da2 = c2)
    //   d ln G / d k  = -2 (1 - nv) idv
    //   dL/dalpha = WD (2/alpha - 4 pi alpha den iqd c2 - (1 - nv) idv),
    // 2/alpha for 2 alpha/a2, taken as 0 at alpha = 0 (there a2 = 0, D = 0
    // and WD = 0; for alpha > 0 it is at most 2^111).  n.v < 0 needs no
    // clamp mask: dL/dimg is masked to 0 for every channel there (img = 0 <
    // eps), so W = 0, gnl = 0 and dL/dnv = 0 with every factor finite.
    const float di = s.den * s.iqd;
    const float WDq1i = W * s.Dq1 * s.idv;
    const float WD = W * s.DGq;
    const float bn = s.c2 * ((4.0f * kPi) * (1.0f - m.a2) * di - 8.0f * s.iq) + (2.0f * m.k) * s.idv;
    const float gnv = gnl + WDq1i * bn;
    const float two_ia = m.alpha > 0.0f ? 2.0f * frcp(m.alpha) : 0.0f;
    const float ba = two_ia - (4.0f * kPi) * m.alpha * (di * s.c2) - (1.0f - s.nv) * s.idv;
    a.A += WD * ba;
#pragma unroll
    for (int c = 0; c < 3; ++c) a.gn[c] += gnv * s.v[c];
}

__device__ __forceinline__ void coloc_finish(Accum &a) {
    const float om = 1.0f - sphg_coloc();
#pragma unroll
    for (int c = 0; c < 3; ++c) a.T[c] *= om;
}

// Texture gradient of one pixel from the view sums: the backward of decode().
// `t` are the values the forward decoded (already clamped for the fused step).
__device__ __forceinline__ void decode_backward(const float t[9], const Maps &m, const Accum &a,
                                                float gt[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float gdif = a.S[c] * m.omspe[c] * kInvPi;
        const float gspe = -a.S[c] * m.dif[c] * kInvPi + a.T[c];
        gt[c] = gdif * map_pow_grad(t[c]);
        gt[6 + c] = gspe * map_pow_grad(t[6 + c]);
    }
    gt[5] = a.A * 2.0f * m.rgh * map_pow_grad(t[5]);
    // normal: n = vec/|vec|, vec = (nx, ny, sqrt(1 - clamp(nx^2+ny^2, 0, 1-eps)))
    const float nx = tclamp(t[3], -1.0f, 1.0f);
    const float ny = tclamp(t[4], -1.0f, 1.0f);
    const float s_raw = nx * nx + ny * ny;
    const float s = tclamp(s_raw, 0.0f, kOneMinusEps);
    const float nz = fsqrt(1.0f - s);
    const float il = frsq(nx * nx + ny * ny + nz * nz);
    const float il3 = il * il * il;
    const float proj = a.gn[0] * nx + a.gn[1] * ny + a.gn[2] * nz;
    const float gvx = a.gn[0] * il - proj * nx * il3;
    const float gvy = a.gn[1] * il - proj * ny * il3;
    const float gvz = a.gn[2] * il - proj * nz * il3;
    float gs = -gvz * 0.5f * frcp(nz);
    gs = (s_raw >= 0.0f && s_raw <= kOneMinusEps) ? gs : 0.0f;
    const float gnx = gvx + gs * 2.0f * nx;
    const float gny = gvy + gs * 2.0f * ny;
    gt[3] = (t[3] >= -1.0f && t[3] <= 1.0f) ? gnx : 0.0f;
    gt[4] = (t[4] >= -1.0f && t[4] <= 1.0f) ? gny : 0.0f;
}

}  // namespace svbrdf
