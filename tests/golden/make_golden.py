"""Write the golden vectors in this directory from the reference's own renderer.

Runs the reference's ``Microfacet`` class (src/microfacet.py, which needs
only numpy and torch) on small seeded inputs and stores inputs and outputs
as .npz fixtures.  Only this script touches the reference, and only where
the reference tree is present; the fixtures it writes are data (inputs and
the outputs the reference computed) and are what the tests read.

    python tests/golden/make_golden.py [/path/to/reference]

Fixtures
--------
fwd_vjp_<case>.npz
    tex [9,R,R] f32, cam/light [V,3], light_pow [3], size, res;
    out32 / out64: Microfacet.eval in float32 / float64;
    gout: a seeded cotangent [V,3,R,R];
    dtex32/dtex64, dpow32/dpow64: gradients of sum(gout * eval(tex)) with
    respect to tex and light_pow (light_pow given via update_light).
optim_<loss>_r8.npz
    5 iterations of the loop of src/svbrdf.py:56-71 (clamp on read, render,
    loss, backward, torch.optim.Adam(lr, betas=(0.9, 0.999)), single-tensor
    update) driven by the reference Microfacet, with and without light
    optimisation: the textures, light power and loss after every step.
"""
import importlib.util
import pathlib
import sys

import numpy as np
import torch

HERE = pathlib.Path(__file__).resolve().parent


def load_reference_microfacet(ref_root):
    path = pathlib.Path(ref_root) / "src" / "microfacet.py"
    spec = importlib.util.spec_from_file_location("reference_microfacet", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.Microfacet


def seeded_tex(res, seed, edge=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.empty((9, res, res), np.float64)
    t[0:3] = rng.uniform(0.05, 0.9, (3, res, res)) * 2 - 1
    ang = rng.uniform(0, 2 * np.pi, (res, res))
    rad = rng.uniform(0.0, 0.45, (res, res))
    t[3] = rad * np.cos(ang)
    t[4] = rad * np.sin(ang)
    t[5] = rng.uniform(0.1, 0.9, (res, res)) * 2 - 1
    t[6:9] = rng.uniform(0.02, 0.3, (3, res, res)) * 2 - 1
    if edge:
        # exact bounds, out-of-disk normals, black / white maps
        t[0, 0, :] = -1.0
        t[1, 1, :] = 1.0
        t[3, 2, :] = 1.0
        t[4, 2, :] = 0.5          # nx^2 + ny^2 > 1: clamped to 1 - eps
        t[3, 3, :] = -1.0
        t[4, 3, :] = 0.0
        t[5, 4, :] = -1.0         # zero roughness
        t[6:9, 5, :] = -1.0       # zero specular
        t[5, 6, :] = 1.0
        t[:, 7, 0] = -1.0         # fully black texel
    return t.astype(np.float32)


def vjp(Microfacet, tex, cam, light, pow_, size, gout, dtype):
    res = tex.shape[-1]
    V = cam.shape[0]
    cl = [torch.tensor(cam, dtype=dtype), torch.tensor(light, dtype=dtype),
          torch.tensor(pow_, dtype=dtype)]
    r = Microfacet(res, V, size, cl, "cpu")
    r.pos = r.pos.to(dtype)
    p = torch.tensor(pow_, dtype=dtype, requires_grad=True)
    r.update_light(p)
    t = torch.tensor(tex, dtype=dtype, requires_grad=True)
    out = r.eval(t[None])
    (out * torch.tensor(gout, dtype=dtype)).sum().backward()
    return (out.detach().numpy(), t.grad.numpy(), p.grad.numpy())


def write_fwd_vjp(Microfacet, name, tex, cam, light, pow_, size, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    V, res = cam.shape[0], tex.shape[-1]
    gout = rng.standard_normal((V, 3, res, res)).astype(np.float32)
    o32, d32, p32 = vjp(Microfacet, tex, cam, light, pow_, size, gout, torch.float32)
    o64, d64, p64 = vjp(Microfacet, tex, cam, light, pow_, size, gout, torch.float64)
    np.savez_compressed(HERE / f"fwd_vjp_{name}.npz", tex=tex, cam=cam, light=light,
                        light_pow=pow_, size=np.float64(size), res=np.int64(res),
                        out32=o32, out64=o64, gout=gout, dtex32=d32, dtex64=d64,
                        dpow32=p32, dpow64=p64)


def write_optim(Microfacet, mode):
    res, steps, lr = 8, 5, 0.01
    size = 17.0
    cam = np.array([[x, y, 30.0] for y in (8.0, 0.0, -8.0) for x in (-8.0, 0.0, 8.0)], np.float32)
    gt = seeded_tex(res, 11)
    rng = np.random.Generator(np.random.PCG64(12))
    init = np.clip(gt + rng.normal(0, 0.1, gt.shape), -1, 1).astype(np.float32)
    pow_ = np.full(3, 1500.0, np.float32)
    eps_floor = 1e-6 ** (1 / 2.2)
    result = {}
    for optim_light in (False, True):
        cl = [torch.tensor(cam), torch.tensor(cam), torch.tensor(pow_)]
        r = Microfacet(res, 9, size, cl, "cpu")
        with torch.no_grad():
            target = r.eval(torch.tensor(gt)[None])
        # a darker start for the light so that its gradient is non-trivial
        light = torch.tensor(pow_ * np.float32(0.8), requires_grad=True)
        r.update_light(light)
        tex = torch.tensor(init)[None].requires_grad_(True)
        params = [tex] + ([light] if optim_light else [])
        opt = torch.optim.Adam(params, lr=lr, betas=(0.9, 0.999), foreach=False)
        texs, lights, losses = [], [], []
        for _ in range(steps):
            if optim_light:
                r.update_light(light)
            out = r.eval(tex.clamp(-1, 1))
            if mode == "l2":
                loss = torch.nn.MSELoss()(out, target)
            else:
                loss = torch.nn.MSELoss()(torch.log(out), torch.log(target.clamp(eps_floor, 1)))
            losses.append(loss.item())
            opt.zero_grad()
            loss.backward()
            opt.step()
            texs.append(tex.detach()[0].numpy().copy())
            lights.append(light.detach().numpy().copy())
        key = "light" if optim_light else "tex"
        result[f"{key}_textures"] = np.stack(texs)
        result[f"{key}_light_pow"] = np.stack(lights)
        result[f"{key}_loss"] = np.array(losses)
        result["target"] = target.numpy()[:, :, :, :]
    np.savez_compressed(HERE / f"optim_{mode}_r8.npz", init=init, gt=gt, cam=cam,
                        light=cam, light_pow0=pow_ * np.float32(0.8), size=np.float64(size),
                        res=np.int64(res), lr=np.float64(lr), **result)


def main(ref_root):
    Microfacet = load_reference_microfacet(ref_root)
    torch.manual_seed(0)
    # render.json geometry (data/random/render.json:5-25)
    rj = np.array([[-3, 3, 16], [0, 3, 16], [3, 3, 16], [-3, 0, 16], [0, 0, 16],
                   [3, 0, 16], [-3, -3, 16], [0, -3, 16], [3, -3, 16]], np.float32)
    pow_ = np.full(3, 1500.0, np.float32)
    write_fwd_vjp(Microfacet, "renderjson_r8", seeded_tex(8, 1), rj, rj, pow_, 6.848, 101)
    grid = np.array([[x, y, 30.0] for y in (8.0, 0.0, -8.0) for x in (-8.0, 0.0, 8.0)], np.float32)
    write_fwd_vjp(Microfacet, "grid_r16", seeded_tex(16, 2), grid, grid, pow_, 17.0, 102)
    # separate camera and light, a light below the plane (back-facing), and
    # a very bright light that saturates part of the image
    cam = np.array([[-4, 2, 20], [3, -1, 25], [0, 0, 10], [1, 1, 40]], np.float32)
    light = np.array([[2, 2, 18], [-3, 1, 22], [0, 0, -5], [1, 1, 40]], np.float32)
    pow_e = np.array([20000.0, 900.0, 300.0], np.float32)
    write_fwd_vjp(Microfacet, "edge_r8", seeded_tex(8, 3, edge=True), cam, light, pow_e, 9.0, 103)
    # non-power-of-two resolution
    write_fwd_vjp(Microfacet, "grid_r13", seeded_tex(13, 4), grid[:5], grid[:5], pow_, 17.0, 104)
    write_optim(Microfacet, "l2")
    write_optim(Microfacet, "logl2")
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
