"""NumPy restatement of the reference microfacet renderer, its adjoint, the
image losses and the Adam step.  TEST INFRASTRUCTURE ONLY (see oracle/__init__).

Every function takes a ``dtype`` (np.float32 or np.float64).  In float32 each
elementwise op rounds once, in the same order as the reference's torch op
sequence, so this is the "reference FP32 render"; in float64 it is the FP64
render the error reports are measured against.

Layout follows the reference: textures ``[9, R, R]`` in [-1, 1] with channel
order diffuse 0:3, normal (nx, ny) 3:5, roughness 5, specular 6:9
(src/svbrdf.py:166, src/microfacet.py:74-77); images ``[V, 3, R, R]``.
Channel-invariant terms are kept as ``[V, 1, R, R]``; the reference expands
them to 3 channels (src/microfacet.py:55,76), which changes no value.

Pinned by: tests/golden/*.npz written by the reference's own Microfacet
(forward and vector-Jacobian product, FP32 and FP64), closed-form known
answers (flat normal under a co-located light, all-black textures,
saturation, back-facing), central finite differences of the FP64 forward and
torch.autograd of the eager restatement (tests/test_oracle.py), and
torch.optim.Adam itself for the optimiser step.
"""
import numpy as np

EPS = 1e-6                 # src/microfacet.py:14
GAMMA = 2.2                # src/microfacet.py:74-77 (maps), :120 (1/2.2 output)
FRESNEL_A = -5.55473       # src/microfacet.py:44
FRESNEL_B = -6.98316       # src/microfacet.py:44

LOSS_L2 = 0                # mean((out - target)^2)          src/optimization.py:14,29
LOSS_LOG_L2 = 1            # mean((ln out - ln target')^2)   defined by this project


def _c(x, dtype):
    return dtype(x)


def pixel_positions(res, size, dtype=np.float32):
    """Pixel-centre positions on the z=0 plane, ``[3, R, R]``.

    src/microfacet.py:16-19: ``t = ((j + 0.5) / res - 0.5) * size``; the
    grid is ``(x=t[j], y=-t[i], 0)`` so row 0 lies at +y.  The reference
    builds the grid in float32 whatever the texture dtype (``arange(...,
    dtype=float32)`` at :16), so it is computed in float32 and then cast.
    """
    f = np.float32
    t = np.arange(res, dtype=f)
    t = ((t + f(0.5)) / f(res) - f(0.5)) * f(size)
    t = t.astype(dtype)
    x = np.broadcast_to(t[None, :], (res, res))
    y = np.broadcast_to(t[:, None], (res, res))
    return np.stack((x, -y, np.zeros_like(x)), 0)


def _norm(vec):
    # src/microfacet.py:58  vec.norm(2.0, 1, keepdim=True)
    return np.sqrt((vec * vec).sum(1, keepdims=True))


def _dot(a, b):
    # src/microfacet.py:55  (a * b).sum(1, keepdim=True)  (the expand to 3 ch is implicit)
    return (a * b).sum(1, keepdims=True)


def geometry(res, size, cam, light, dtype=np.float32, rows=None):
    """View/light directions for every (view, pixel).

    src/microfacet.py:60-62 (get_dir), :91-93 (v, l, |l|^2, h), :99 (v.h).
    Returns dict of [V,3,R,R] directions and [V,1,R,R] scalars (only the
    image rows ``rows`` -- a slice -- when given).
    """
    pos = pixel_positions(res, size, dtype)[None]
    if rows is not None:
        pos = pos[:, :, rows]
    cam = np.asarray(cam, dtype=dtype)[:, :, None, None]
    light = np.asarray(light, dtype=dtype)[:, :, None, None]
    vvec = cam - pos
    v = vvec / _norm(vvec)
    lvec = light - pos
    l = lvec / _norm(lvec)
    d2 = _dot(lvec, lvec)
    hsum = l + v
    h = hsum / _norm(hsum)
    vh_raw = _dot(v, h)
    return dict(v=v, l=l, h=h, d2=d2, vh=np.maximum(vh_raw, _c(0, dtype)))


def decode(tex, dtype=np.float32):
    """Texture decode, src/microfacet.py:64-79 (tex2map + reconstruct_normal).

    Returns dict with dif [3,R,R], spe [3,R,R], rgh [1,R,R], the unit normal
    n [3,R,R] and the intermediates the adjoint needs.
    """
    t = np.asarray(tex, dtype=dtype).reshape(9, *tex.shape[-2:])
    one, two = _c(1, dtype), _c(2, dtype)
    g = _c(GAMMA, dtype)
    u_dif = (t[0:3] + one) / two
    u_rgh = (t[5:6] + one) / two
    u_spe = (t[6:9] + one) / two
    dif = u_dif ** g
    rgh = u_rgh ** g
    spe = u_spe ** g
    nx = np.clip(t[3], -one, one)
    ny = np.clip(t[4], -one, one)
    s_raw = nx ** 2 + ny ** 2
    s = np.clip(s_raw, _c(0, dtype), _c(1 - EPS, dtype))
    nz = np.sqrt(one - s)
    nvec = np.stack((nx, ny, nz), 0)[None]
    nlen = _norm(nvec)
    n = (nvec / nlen)[0]
    return dict(t=t, u_dif=u_dif, u_rgh=u_rgh, u_spe=u_spe, dif=dif, rgh=rgh,
                spe=spe, nx=nx, ny=ny, s_raw=s_raw, nz=nz, nvec=nvec[0],
                nlen=nlen[0], n=n)


NDFS = ("ggx", "beckmann")          # src/microfacet.py:28-32, :34-38
FRESNELS = ("sg", "schlick")        # src/microfacet.py:43-45, :40-41


def _terms(mp, geo, light_pow, dtype, dot_scale=None, ndf="ggx", fresnel="sg"):
    """Forward BRDF terms shared by render() and the adjoint.

    src/microfacet.py:96-117 with GGX :28-32, Fresnel :43-45, Smith :47-52.
    ``ndf="beckmann"`` uses Beckmann (:34-38) instead of GGX, and
    ``fresnel="schlick"`` uses Fresnel_f0(v.h, specular) (:40-41) instead of
    Fresnel(v.h, specular): the reference defines both but eval calls GGX and
    Fresnel.  ``dot_scale`` = (sv, sl, sh) multiplies n.v, n.l, n.h (used only
    by :func:`fp32_noise_scale`).
    """
    assert ndf in NDFS and fresnel in FRESNELS
    one, eps = _c(1, dtype), _c(EPS, dtype)
    pi = _c(np.pi, dtype)
    n = mp["n"][None]
    nv_raw = _dot(n, geo["v"])
    nl_raw = _dot(n, geo["l"])
    nh_raw = _dot(n, geo["h"])
    if dot_scale is not None:
        nv_raw = nv_raw * dtype(dot_scale[0])
        nl_raw = nl_raw * dtype(dot_scale[1])
        nh_raw = nh_raw * dtype(dot_scale[2])
    zero = _c(0, dtype)
    nv = np.maximum(nv_raw, zero)
    nl = np.maximum(nl_raw, zero)
    nh = np.maximum(nh_raw, zero)
    vh = geo["vh"]
    dif, spe, rgh = mp["dif"][None], mp["spe"][None], mp["rgh"][None]
    f1 = dif / pi
    f1 = f1 * (one - spe)
    alpha = rgh ** 2
    c2 = nh ** 2
    a2 = alpha ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        if ndf == "ggx":
            # GGX(n_dot_h, roughness**2)
            den = c2 * a2 + (one - c2)
            qd = pi * den ** 2 + eps
            D = a2 / qd
        else:
            # Beckmann(n_dot_h, roughness**2): exp(-t2/a2) / (pi a2 c2^2)
            den = (one - c2) / c2                 # t2
            qd = pi * a2 * c2 ** 2
            D = np.exp(-den / a2) / qd
    if fresnel == "sg":
        # Fresnel(v_dot_h, specular)
        sphg = np.power(_c(2, dtype), ((_c(FRESNEL_A, dtype) * vh) + _c(FRESNEL_B, dtype)) * vh)
    else:
        # Fresnel_f0(v_dot_h, specular): f0 + (1 - f0) (1 - cos)^5
        sphg = (one - vh) ** 5
    F = spe + (one - spe) * sphg
    # Smith(n_dot_v, n_dot_l, roughness**2)
    k = alpha * _c(0.5, dtype) + eps
    dv = nv * (one - k) + k
    dl = nl * (one - k) + k
    g1v = nv / dv
    g1l = nl / dl
    G = g1v * g1l
    q = _c(4, dtype) * nv * nl + eps
    f2 = D * F * G / q
    f = f1 + f2
    P = np.asarray(light_pow, dtype=dtype)[None, :, None, None]
    img = P * f * nl / geo["d2"]
    return dict(nv_raw=nv_raw, nl_raw=nl_raw, nh_raw=nh_raw, nv=nv, nl=nl, nh=nh,
                f1=f1, alpha=alpha, c2=c2, a2=a2, den=den, qd=qd, D=D, sphg=sphg,
                F=F, k=k, dv=dv, dl=dl, g1v=g1v, g1l=g1l, G=G, q=q, f2=f2, f=f,
                P=P, img=img)


def render(tex, res, size, cam, light, light_pow, dtype=np.float32, return_all=False,
           dot_scale=None, rows=None, ndf="ggx", fresnel="sg"):
    """Forward render, src/microfacet.py:84-120.  Returns ``[V,3,R,R]``
    (or ``[V,3,len(rows),R]`` for a slice ``rows`` of image rows).

    ``img.clamp(eps, 1) ** (1/2.2)``: gamma-encoded, in [eps^(1/2.2), 1].
    """
    tex = np.asarray(tex)
    assert tex.shape[-1] == tex.shape[-2] == res
    if rows is not None:
        tex = tex.reshape(9, res, res)[:, rows]
    mp = decode(tex, dtype)
    geo = geometry(res, size, cam, light, dtype, rows)
    tm = _terms(mp, geo, light_pow, dtype, dot_scale, ndf, fresnel)
    ic = np.clip(tm["img"], _c(EPS, dtype), _c(1, dtype))
    out = ic ** _c(1 / GAMMA, dtype)
    if return_all:
        return out, dict(mp=mp, geo=geo, tm=tm, ic=ic)
    return out


def fp32_noise_scale(tex, res, size, cam, light, light_pow, ulps=1.0, rows=None, ndf="ggx", fresnel="sg"):
    """Per-element FP32 rounding sensitivity of :func:`render`, in FP64.

    The ill-conditioned intermediates of the renderer are the three dot
    products n.v, n.l, n.h: GGX forms 1 - (n.h)^2 (src/microfacet.py:31) and
    Smith/the cosine terms divide by small n.v, n.l near grazing angles.  Any
    FP32 implementation rounds them with an error of about one ulp, so the
    output of two correct FP32 implementations can differ by about
    sum_x |out(x (1 + u)) - out(x)| with u = ulps * 2^-24.  Returns that sum
    (FP64, shape [V,3,R,R]).
    """
    u = ulps * 2.0 ** -24
    base = render(tex, res, size, cam, light, light_pow, np.float64, rows=rows, ndf=ndf, fresnel=fresnel)
    s = np.zeros_like(base)
    for k in range(3):
        for sign in (1.0, -1.0):
            sc = [1.0, 1.0, 1.0]
            sc[k] = 1.0 + sign * u
            d = np.abs(render(tex, res, size, cam, light, light_pow, np.float64,
                              dot_scale=sc, rows=rows, ndf=ndf, fresnel=fresnel) - base)
            s_k = d if sign > 0 else np.maximum(s_k, d)
        s += s_k
    return s


def target_floor(dtype=np.float32):
    """Smallest value the renderer can output: eps^(1/2.2)."""
    return _c(EPS, dtype) ** _c(1 / GAMMA, dtype)


def loss(out, target, mode, dtype=np.float32):
    """Image loss.  L2: src/optimization.py:14,29 (MSELoss, mean reduction).

    LOG_L2 (not in the reference; the benchmark workloads use it): mean of
    (ln out - ln clamp(target, eps^(1/2.2), 1))^2.  The clamp keeps a
    target pixel the renderer cannot reach (e.g. a black photo pixel) finite.
    """
    out = np.asarray(out, dtype=dtype)
    target = np.asarray(target, dtype=dtype)
    if mode == LOSS_L2:
        r = out - target
    else:
        r = np.log(out) - np.log(np.clip(target, target_floor(dtype), _c(1, dtype)))
    return float(np.mean((r * r).astype(np.float64)))


def loss_grad(out, target, mode, dtype=np.float32):
    """dL/dout of :func:`loss` (mean reduction over all V*3*R*R elements)."""
    out = np.asarray(out, dtype=dtype)
    target = np.asarray(target, dtype=dtype)
    N = _c(out.size, dtype)
    if mode == LOSS_L2:
        return _c(2, dtype) * (out - target) / N
    r = np.log(out) - np.log(np.clip(target, target_floor(dtype), _c(1, dtype)))
    return _c(2, dtype) * r / N / out


def render_vjp(tex, res, size, cam, light, light_pow, gout, dtype=np.float32, rows=None, ndf="ggx",
               fresnel="sg"):
    """Vector-Jacobian product of :func:`render`: returns (dtex [9,R,R], dP [3]).

    Hand-derived adjoint of src/microfacet.py:64-120 with torch's gradient
    conventions: clamp/clamp_min pass the gradient at the bounds (inclusive
    masks), sqrt' = 1/(2 sqrt), the normalize of :70 contributes its full
    Jacobian (I - n n^T)/|n| even when |n| = 1, and expanded tensors sum
    their gradients over views and channels.  Positions carry no gradient.
    """
    out, st = render(tex, res, size, cam, light, light_pow, dtype, return_all=True, rows=rows, ndf=ndf,
                     fresnel=fresnel)
    mp, geo, tm = st["mp"], st["geo"], st["tm"]
    one, eps, zero = _c(1, dtype), _c(EPS, dtype), _c(0, dtype)
    pi = _c(np.pi, dtype)
    gout = np.asarray(gout, dtype=dtype)
    img = tm["img"]
    # out = clamp(img, eps, 1) ** (1/2.2)
    m_out = (img >= eps) & (img <= one)
    e = _c(1 / GAMMA, dtype)
    g_img = gout * (e * st["ic"] ** (e - one)) * m_out
    # img = ((P * f) * nl) / d2
    d2 = geo["d2"]
    P, f, nl, nv = tm["P"], tm["f"], tm["nl"], tm["nv"]
    g_f = g_img * P * nl / d2
    g_nl = (g_img * P * f / d2).sum(1, keepdims=True)
    dP = (g_img * f * nl / d2).sum(axis=(0, 2, 3))
    # f = f1 + f2 ; f1 = dif/pi*(1-spe) ; f2 = D*F*G/q
    dif, spe = mp["dif"][None], mp["spe"][None]
    D, F, G, q = tm["D"], tm["F"], tm["G"], tm["q"]
    g_dif = (g_f * (one - spe) / pi).sum(0)
    g_spe = (-g_f * dif / pi).sum(0)
    g_D = (g_f * F * G / q).sum(1, keepdims=True)
    g_F = g_f * D * G / q
    g_G = (g_f * D * F / q).sum(1, keepdims=True)
    g_q = (-g_f * D * F * G / (q * q)).sum(1, keepdims=True)
    g_nv = g_q * _c(4, dtype) * nl
    g_nl = g_nl + g_q * _c(4, dtype) * nv
    # F = spe + (1 - spe) * sphg     (sphg depends on v.h only: no gradient)
    g_spe = g_spe + (g_F * (one - tm["sphg"])).sum(0)
    # G = G1(nv) G1(nl),  G1(c) = c / (c(1-k) + k)
    k, dv, dl, g1v, g1l = tm["k"], tm["dv"], tm["dl"], tm["g1v"], tm["g1l"]
    g_nv = g_nv + g_G * g1l * k / (dv * dv)
    g_nl = g_nl + g_G * g1v * k / (dl * dl)
    g_k = g_G * (g1l * (-nv * (one - nv)) / (dv * dv) + g1v * (-nl * (one - nl)) / (dl * dl))
    # k = alpha/2 + eps ; D = a2 / (pi den^2 + eps), a2 = alpha^2, den = c2 a2 + 1 - c2
    alpha, a2, den, qd, c2, nh = tm["alpha"], tm["a2"], tm["den"], tm["qd"], tm["c2"], tm["nh"]
    if ndf == "ggx":
        dD_dden = -a2 * _c(2, dtype) * pi * den / (qd * qd)
        g_a2 = g_D * (one / qd + dD_dden * c2)
        g_c2 = g_D * dD_dden * (a2 - one)
    else:
        # D = exp(-t2/a2) / (pi a2 c2^2), t2 = (1 - c2)/c2 (den holds t2):
        # dD/dc2 = D (1/(a2 c2^2) - 2/c2), dD/da2 = D (t2/a2^2 - 1/a2)
        with np.errstate(divide="ignore", invalid="ignore"):
            g_c2 = g_D * D * (one / (a2 * c2 * c2) - _c(2, dtype) / c2)
            g_a2 = g_D * D * (den / (a2 * a2) - one / a2)
    g_nh = g_c2 * _c(2, dtype) * nh
    g_alpha = g_k * _c(0.5, dtype) + g_a2 * _c(2, dtype) * alpha
    g_rgh = (g_alpha * _c(2, dtype) * mp["rgh"][None]).sum(0)   # [1,R,R]: summed over views
    # clamp_min(0) on the dots (inclusive) then dot -> normal, summed over views
    g_nv = g_nv * (tm["nv_raw"] >= zero)
    g_nl = g_nl * (tm["nl_raw"] >= zero)
    g_nh = g_nh * (tm["nh_raw"] >= zero)
    g_n = (g_nv * geo["v"] + g_nl * geo["l"] + g_nh * geo["h"]).sum(0)   # [3,R,R]
    # normalize: n = vec / |vec|
    nvec, nlen = mp["nvec"], mp["nlen"]
    g_vec = g_n / nlen - (g_n * nvec).sum(0, keepdims=True) * nvec / (nlen * nlen * nlen)
    # nz = sqrt(1 - s), s = clamp(nx^2 + ny^2, 0, 1-eps)
    g_s = -g_vec[2] / (_c(2, dtype) * mp["nz"])
    s_raw = mp["s_raw"]
    g_s = g_s * ((s_raw >= zero) & (s_raw <= _c(1 - EPS, dtype)))
    t = mp["t"]
    g_nx = (g_vec[0] + g_s * _c(2, dtype) * mp["nx"]) * ((t[3] >= -one) & (t[3] <= one))
    g_ny = (g_vec[1] + g_s * _c(2, dtype) * mp["ny"]) * ((t[4] >= -one) & (t[4] <= one))
    # u ** 2.2 with u = (t + 1) / 2
    g = _c(GAMMA, dtype)
    half = _c(0.5, dtype)
    dtex = np.empty_like(t)
    dtex[0:3] = g_dif * (g * mp["u_dif"] ** (g - one)) * half
    dtex[3] = g_nx
    dtex[4] = g_ny
    dtex[5] = (g_rgh * (g * mp["u_rgh"] ** (g - one)) * half)[0]
    dtex[6:9] = g_spe * (g * mp["u_spe"] ** (g - one)) * half
    return dtex, dP


def clamp_grad_mask(tex, dtype=np.float32):
    """Inclusive mask of ``textures.clamp(-1, 1)`` (src/svbrdf.py:60)."""
    t = np.asarray(tex, dtype=dtype)
    return (t >= _c(-1, dtype)) & (t <= _c(1, dtype))


def _fma(a, b, c, dtype):
    """a*b + c with one rounding (emulated in float64 for float32 inputs)."""
    if dtype == np.float64:
        return a * b + c
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(dtype)


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8, dtype=np.float32):
    """One torch.optim.Adam update (single-tensor order, no weight decay,
    no amsgrad), torch/optim/adam.py:431-547.  ``step`` is the new step
    count (>= 1).  Bias-correction scalars are computed in double, as the
    Python floats torch computes them, and rounded once to ``dtype``.
    torch's CPU lerp_ and addcmul_ round once (fused multiply-add):
    m = fma(1-beta1, g - m, m), v = fma((1-beta2) g, g, v beta2).
    Returns (p, m, v) as new arrays.
    """
    p = np.asarray(p, dtype=dtype)
    g = np.asarray(g, dtype=dtype)
    m = np.asarray(m, dtype=dtype)
    v = np.asarray(v, dtype=dtype)
    m = _fma(_c(1 - beta1, dtype), g - m, m, dtype)                  # exp_avg.lerp_(grad, 1-beta1)
    v = _fma(_c(1 - beta2, dtype) * g, g, v * _c(beta2, dtype), dtype)  # mul_(b2).addcmul_(g, g, 1-b2)
    bc1 = 1 - beta1 ** float(step)
    bc2 = 1 - beta2 ** float(step)
    step_size = lr / bc1
    bc2_sqrt = bc2 ** 0.5
    denom = np.sqrt(v) / _c(bc2_sqrt, dtype) + _c(eps, dtype)
    p = p + _c(-step_size, dtype) * m / denom            # addcdiv_(m, denom, value=-step_size)
    return p, m, v


def fused_step(tex, m, v, target, step, res, size, cam, light, light_pow, mode,
               lr, beta1=0.9, beta2=0.999, eps=1e-8, dtype=np.float32, light_state=None):
    """One per-pixel optimisation iteration, src/svbrdf.py:56-71:
    clamp on read -> render -> loss -> backward -> Adam.  The stored
    parameters stay unclamped.  With ``light_state = (m_P, v_P)`` the light
    power is optimised too (one Adam param group, src/svbrdf.py:48-50).
    Returns (tex, m, v, loss) or, with light_state, (tex, m, v, loss,
    (light_pow, m_P, v_P)).
    """
    t = np.asarray(tex, dtype=dtype)
    tc = np.clip(t, _c(-1, dtype), _c(1, dtype))
    out = render(tc, res, size, cam, light, light_pow, dtype)
    L = loss(out, target, mode, dtype)
    gout = loss_grad(out, target, mode, dtype)
    dtex, dP = render_vjp(tc, res, size, cam, light, light_pow, gout, dtype)
    dtex = dtex * clamp_grad_mask(t, dtype)
    p, m, v = adam_step(t, dtex, m, v, step, lr, beta1, beta2, eps, dtype)
    if light_state is None:
        return p, m, v, L
    lp, lm, lv = adam_step(light_pow, dP, light_state[0], light_state[1], step, lr, beta1, beta2, eps, dtype)
    return p, m, v, L, (lp, lm, lv)
