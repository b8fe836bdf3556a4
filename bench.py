"""Benchmark: per-pixel SVBRDF optimisation iterations/sec at 1024^2 x 9 views.

One step = one fused iteration of the reference loop body (src/svbrdf.py:58-71):
clamp -> render -> log-space L2 loss -> backward -> Adam, on a 1024x1024
material seen from 9 co-located flash views, all fp32, as one kernel launch
per iteration, replayed from HIP graphs of 100 iterations.  Inputs are
synthetic (seeded maps, targets rendered from the ground truth) and resident
in HBM before timing starts.

Printed JSON line (one):
  value             iterations/sec of the fused step over the K timed
                    iterations (HIP events around the whole region)
  roofline          algorithmic HBM bytes per iteration (324 B/pixel at V=9:
                    params, exp_avg, exp_avg_sq read+written once, targets read
                    once) / the kernel's average duration (HIP events around
                    each graph replay of back-to-back launches, median over
                    50 replays), against 8 TB/s; step_frac is the same over
                    the whole-job time per iteration; traffic is the
                    profiler-measured HBM bytes per launch
  per_replay_us_per_step  median and IQR over 50 individually timed graph replays
  eager_launch_ms   events around one eager launch (includes the events' own
                    end-of-kernel cache write-back)
  with_sync         it/s with the loss read on the host every iteration, as the
                    reference loop does (src/svbrdf.py:64)
  torch_gpu_baseline  the reference's eager op sequence + torch.optim.Adam on
                    the same GPU (it/s), and the speedup over it
  cpu_baseline      the same eager restatement on the host CPU cores (oracle/);
                    cpu_baseline_256 the same at 256^2 x 9
  forward           forward-render throughput at the same shape
  adjoint           the render adjoint alone (dL/dimages -> dL/dtextures, 180 B/pixel)
  optim_light       the fused step with the light power optimised as well
  cache_cold        the fused step with its working set kept out of the
                    256 MB Infinity Cache (two problems stepping alternately)

  batched           64 materials of uneven size in one persistent
                    launch (autotuned), against a per-material loop
  composed          render adjoint under a VGG-shaped feature loss (256^2 x 9)
  codec             the map / target codec kernels (8-bit maps and images <->
                    tensors, Lanczos-4 resize, stage transition) at the same shape
  materialgan       the MaterialGAN caller: generator -> render -> feature loss ->
                    latent Adam per iteration at 256^2 x 9, and the render share
  driver            scripts.render + scripts.optim_perpixel end to end (1000 epochs
                    with the reference's PNG / plot side outputs), wall clock

    python bench.py [--steps K] [--warmup W] [--res 1024] [--views 9]
                    [--no-cpu] [--no-torch] [--no-batch] [--no-compose] [--no-codec]
                    [--no-materialgan]
                    [--cpu-seconds 20] [--graph-steps 100 | --eager] [--no-eager-probes]
"""
import argparse
import ctypes
import json
import math
import os
import pathlib
import re
import sys
import time

import numpy as np
import torch

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

METRIC = "per-pixel optim iters/sec at 1024²×9 views (fused vs torch autograd); HBM GB/s"
# synthetic materials of every workload: one albedo / specular / roughness draw
# per channel for the whole map, detail in the height-field normals
ALBEDO = "material"
HBM_PEAK_GBS = 8000.0


# HBM bytes per fused-step launch, from the rocprofv3 summary committed under
# profiles/ for the profiled shape (FETCH_SIZE and WRITE_SIZE in KiB per
# dispatch, each scaled by the factor a plain device copy of known size gives
# the same counter; the summary states both factors).
TRAFFIC_PROFILES = {(1024, 9): ROOT / "profiles" / "fused_step_traffic_1024x9.md"}


def measured_traffic(res, views):
    """(bytes per launch, source) from the committed traffic profile of this
    shape, or (None, reason)."""
    path = TRAFFIC_PROFILES.get((res, views))
    if path is None or not path.exists():
        return None, f"no traffic profile for {res}^2 x {views}"
    text = path.read_text()
    row = re.search(r"\|\s*`fused_step_kernel<1, false>`\s*\|\s*([\d,]+)\s*\|\s*([\d,]+)\s*\|", text)
    scale = re.search(r"FETCH_SIZE is scaled \S(\d+(?:\.\d+)?) and WRITE_SIZE \S(\d+(?:\.\d+)?)", text)
    if not row or not scale:
        return None, f"{path.name}: no fused_step_kernel row or scale factors"
    fetch, write = (int(x.replace(",", "")) for x in row.groups())
    fs, ws = (float(x) for x in scale.groups())
    nbytes = int(round((fetch * fs + write * ws) * 1024))
    return nbytes, (f"profiles/{path.name}: FETCH_SIZE {fetch:,} KiB x {fs:g} + WRITE_SIZE {write:,} KiB x {ws:g} "
                    "per dispatch (rocprofv3 --pmc)")


def algorithmic_bytes_step(res, views):
    # params, exp_avg, exp_avg_sq: 9 fp32 each, read and written once (216 B)
    # targets: 3 fp32 per view read once (12 V B)
    return res * res * (216 + 12 * views)


def algorithmic_bytes_render(res, views):
    return res * res * (36 + 12 * views)


INFINITY_CACHE_BYTES = 256 * 2**20   # MI355X last-level (memory-side) cache
WARM_KERNEL_LAUNCHES = 1500           # untimed launches before the forward / adjoint legs are timed


def working_set_bytes(fs, views):
    """Device bytes a FusedStep reads and writes every iteration: its maps
    (at texture_stride), two Adam moments and the prepared targets (at
    plane_stride)."""
    npix = fs.res * fs.res
    tps, ps = fs.texture_stride or npix, fs.plane_stride or npix   # 0: dense planes
    return 4 * (9 * tps + 18 * ps + 3 * views * ps)


def algorithmic_bytes_adjoint(res, views):
    # textures and dL/dimages read once, dL/dtextures written once
    return res * res * (36 + 12 * views + 36)


def make_problem(pkg, res, views, dev, loss="log_l2"):
    syn = pkg.synthetic
    gt = syn.ground_truth_textures(res, seed=0, albedo=ALBEDO)
    init = syn.init_textures(gt, seed=1)
    cam, light = syn.flash_grid(n_views=views, jitter=3.0 if views != 9 else 0.0,
                                rng=np.random.default_rng(3))
    P = syn.light_power()
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, views, syn.PLANE_SIZE_CM, cl, dev)
    with torch.no_grad():
        target = r.eval(torch.tensor(gt, device=dev)[None])
    tex = torch.tensor(init, device=dev)[None].contiguous()
    # the maps at the library's padded plane stride, as SvbrdfOptim.optim keeps them
    fs = pkg.FusedStep(r, tex, target, lr=0.01, loss=loss, history=0, own_textures=True)
    return dict(renderer=r, fused=fs, tex=tex, target=target, gt=gt, init=init, cam=cam, light=light, P=P)


def kernel_duration_ms(launch, n):
    """Average duration of one launch: HIP events recorded right before and
    after each launch on the launching stream (torch's current stream), summed
    over n launches; the gaps between launches are excluded."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    launch()
    torch.cuda.synchronize()
    for a, b in ev:
        a.record()
        launch()
        b.record()
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / n


def replay_spread(run, per, n):
    """Median and interquartile range (us per iteration) over n single
    timed units of `per` iterations (one graph replay, or one eager launch)."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    torch.cuda.synchronize()
    for a, b in ev:
        a.record()
        run(per)
        b.record()
    torch.cuda.synchronize()
    us = np.array([a.elapsed_time(b) * 1000.0 / per for a, b in ev])
    q1, med, q3 = np.percentile(us, [25, 50, 75])
    return {"median": round(float(med), 2), "iqr": round(float(q3 - q1), 2), "n": n, "iterations_each": per}


def with_sync_rate(fs, n):
    """Iterations/sec when every iteration's loss is read on the host, as the
    reference loop does (loss.item(), src/svbrdf.py:64): one launch, the
    flush and a device-to-host copy of the state per iteration."""
    for _ in range(3):
        fs.step()
        fs.loss()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        fs.step()
        fs.loss()
    dt = (time.perf_counter() - t0) / n
    return {"value": round(1.0 / dt, 1), "unit": "it/s", "ms_per_step": round(dt * 1000.0, 4),
            "what": "eager launch + loss flush + host read of the loss every iteration (src/svbrdf.py:64)"}


def time_events(fn, n, stream=None):
    """Average ms per call of fn() over n calls, with HIP events on the stream
    the kernels are launched on (torch's current stream)."""
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(n):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / n


# --------------------------------------------------------------------------
# The reference's eager op sequence on the GPU (torch autograd + Adam): the
# denominator of the speedup.  Same ops as src/microfacet.py:84-120, one
# torch op per term on expanded [V,3,R,R] tensors.
class EagerReferencePath:
    def __init__(self, res, n, size, cam, light, P, dev):
        t = torch.arange(res, dtype=torch.float32, device=dev)
        t = ((t + 0.5) / res - 0.5) * size
        x, y = torch.meshgrid(t, t, indexing="xy")
        self.pos = torch.stack((x, -y, torch.zeros_like(x)), 2).permute(2, 0, 1)[None].expand(n, -1, -1, -1)
        self.cam = torch.tensor(cam, device=dev)[:, :, None, None].expand(-1, -1, res, res)
        self.light = torch.tensor(light, device=dev)[:, :, None, None].expand(-1, -1, res, res)
        self.P = torch.tensor(P, device=dev)[None, :, None, None].expand(n, -1, res, res)
        self.n = n

    @staticmethod
    def dot(a, b):
        return (a * b).sum(1, keepdim=True).expand(-1, 3, -1, -1)

    @staticmethod
    def normalize(v):
        return v / v.norm(2.0, 1, keepdim=True)

    def eval(self, tex):
        n = self.n
        dif = (((tex[:, 0:3] + 1) / 2) ** 2.2).expand(n, -1, -1, -1)
        nx, ny = tex[:, 3].clamp(-1, 1), tex[:, 4].clamp(-1, 1)
        nz = (1 - (nx ** 2 + ny ** 2).clamp(0, 1 - 1e-6)).sqrt()
        nrm = self.normalize(torch.stack((nx, ny, nz), 1)).expand(n, -1, -1, -1)
        rgh = (((tex[:, 5] + 1) / 2) ** 2.2).expand(n, 3, -1, -1)
        spe = (((tex[:, 6:9] + 1) / 2) ** 2.2).expand(n, -1, -1, -1)
        vv = self.cam - self.pos
        v = self.normalize(vv)
        lv = self.light - self.pos
        l, d2 = self.normalize(lv), self.dot(lv, lv)
        h = self.normalize(l + v)
        nv = self.dot(nrm, v).clamp(min=0)
        nl = self.dot(nrm, l).clamp(min=0)
        nh = self.dot(nrm, h).clamp(min=0)
        vh = self.dot(v, h).clamp(min=0)
        f1 = dif / math.pi * (1 - spe)
        alpha = rgh ** 2
        c2 = nh ** 2
        a2 = alpha ** 2
        D = a2 / (math.pi * (c2 * a2 + (1 - c2)) ** 2 + 1e-6)
        F = spe + (1.0 - spe) * torch.pow(2.0, ((-5.55473 * vh) - 6.98316) * vh)
        k = rgh ** 2 * 0.5 + 1e-6
        G = nv / (nv * (1.0 - k) + k) * (nl / (nl * (1.0 - k) + k))
        f = f1 + D * F * G / (4 * nv * nl + 1e-6)
        img = self.P * f * nl / d2
        return img.clamp(1e-6, 1) ** (1 / 2.2)


def torch_gpu_baseline(prob, res, views, dev, steps=20, warmup=3, sync=True):
    path = EagerReferencePath(res, views, 17.0, prob["cam"], prob["light"], prob["P"], dev)
    tex = torch.tensor(prob["init"], device=dev)[None].requires_grad_(True)
    log_tgt = torch.log(prob["target"].detach().clamp(1e-6 ** (1 / 2.2), 1))
    opt = torch.optim.Adam([tex], lr=0.01, betas=(0.9, 0.999))

    def it():
        out = path.eval(tex.clamp(-1, 1))
        loss = torch.nn.functional.mse_loss(torch.log(out), log_tgt)
        if sync:
            loss.item()                   # the reference reads the loss every epoch (src/svbrdf.py:64)
        opt.zero_grad()
        loss.backward()
        opt.step()

    for _ in range(warmup):
        it()
    ms = time_events(it, steps)
    return {"value": 1000.0 / ms, "unit": "it/s", "ms_per_step": ms, "steps": steps,
            "what": "reference op sequence (src/microfacet.py:84-120) + log-L2 + torch.optim.Adam, eager, "
                    "autograd" + (", per-iteration loss.item() as src/svbrdf.py:64" if sync else
                                  ", no per-iteration host read")}


def batch_materials(pkg, dev, seed=3):
    """64 materials (24 at 256^2, 24 at 512^2, 16 at 1024^2), views
    drawn uniformly from 4..16, flash positions jittered +-3 cm around the
    3x3 grid; maps seeded per material, targets rendered from the truth."""
    syn = pkg.synthetic
    rng = np.random.default_rng(seed)
    sizes = [256] * 24 + [512] * 24 + [1024] * 16
    views = rng.integers(4, 17, len(sizes))
    mats = []
    for i, (res, V) in enumerate(zip(sizes, views)):
        gt = syn.ground_truth_textures(res, seed=1000 + i, albedo=ALBEDO)
        cam, light = syn.flash_grid(jitter=3.0, rng=rng, n_views=int(V))
        mats.append((res, int(V), gt, syn.init_textures(gt, seed=2000 + i), cam, light))
    return mats


def batch_leg(pkg, dev, steps=20, warmup=3):
    syn = pkg.synthetic
    mats = batch_materials(pkg, dev)
    P = syn.light_power()

    def problems():
        out = []
        for res, V, gt, init, cam, light in mats:
            cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
            r = pkg.Microfacet(res, V, syn.PLANE_SIZE_CM, cl, dev)
            with torch.no_grad():
                tgt = r.eval(torch.tensor(gt, device=dev)[None])
            out.append((r, torch.tensor(init, device=dev)[None].contiguous(), tgt))
        return out

    nbytes = sum(res * res * (216 + 12 * V) for res, V, *_ in mats)
    # maps at the padded plane stride in both the batch and the per-material loop
    bs = pkg.BatchedStep(problems(), 0.01, loss="log_l2", own_textures=True)
    cfg, timings = bs.autotune()
    per = 10
    bs.capture(per)
    for _ in range(warmup):
        bs.replay()
    ms = time_events(bs.replay, max(1, steps // per)) / per
    # baseline: the unbatched step once per material, the whole loop in one graph
    singles = [pkg.FusedStep(r, t, y, 0.01, loss="log_l2", own_textures=True) for r, t, y in problems()]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        g.capture_begin()
        for f in singles:
            f.step()
        g.capture_end()
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(warmup):
        g.replay()
    ms_loop = time_events(g.replay, steps)
    del singles, g
    # the same batch with every material's light power optimised too
    # (optim_light=True), in the launch form tuned above
    bl = pkg.BatchedStep(problems(), 0.01, loss="log_l2", own_textures=True, optim_light=True, config=cfg)
    bl.capture(per)
    for _ in range(warmup):
        bl.replay()
    ms_light = time_events(bl.replay, max(1, steps // per)) / per
    total_px = sum(res * res for res, *_ in mats)
    if cfg["blocks_per_cu"] < 0:
        form = "one block per 256-pixel tile in cost order (the dispatcher balances the load)"
    else:
        form = (f"persistent grid of {cfg['blocks_per_cu']} blocks per CU drawing {cfg['tile_pixels']}-pixel tiles "
                "from a ticket counter")
    return {"workload": f"{len(mats)} materials (24x256^2, 24x512^2, 16x1024^2 = {total_px} px), V in 4..16, "
                        f"fused log-L2 iteration; autotuned form: {form}; three launches per iteration (the step, "
                        "the per-material loss reduction, the state advance), HIP graphs of 10 iterations",
            "value": round(1000.0 / ms, 2), "unit": "batched it/s", "ms_per_step": round(ms, 4),
            "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "algorithmic_bytes": nbytes,
            "autotuned": cfg, "candidates_timed": len(timings),
            "per_material_loop": {"ms_per_step": round(ms_loop, 4), "what": "FusedStep per material, 64 steps "
                                  "(one launch each) in one HIP graph"},
            "speedup_vs_loop": round(ms_loop / ms, 3),
            "optim_light": {"ms_per_step": round(ms_light, 4), "value": round(1000.0 / ms_light, 2),
                            "what": "the same batch and launch form with each material's light power optimised "
                                    "(per-tile dL/dP partials, Adam in the per-material reduction)"}}


def copy_bandwidth(dev, mib=1024, n=10):
    """Attainable HBM bandwidth: a plain device copy of `mib` MiB (read and
    written once), HIP events around n copies."""
    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    ms = time_events(lambda: dst.copy_(src), n)
    gbs = 2 * src.numel() * 4 / (ms * 1e-3) / 1e9
    return {"what": f"torch copy of {mib} MiB (read + write)", "ms": round(ms, 4), "achieved_GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def staged_leg(pkg, dev, views, stages=(256, 512, 1024), iters=20):
    """Coarse to fine: 20 fused iterations at each of 256^2, 512^2 and 1024^2, each
    stage starting from the previous one's maps upsampled x2 (bilinear,
    align_corners=False); one graph of `iters` steps per stage.  Per-stage
    ms and the whole run, with the final loss."""
    syn = pkg.synthetic
    cam, light = syn.flash_grid(n_views=views)
    P = syn.light_power()
    tex = None
    out = {"stages": {}}
    total = 0.0
    # one material: the finest ground truth, box-averaged for the coarse stages
    finest = max(stages)
    gt_fine = syn.ground_truth_textures(finest, seed=0, albedo=ALBEDO)
    for res in stages:
        cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
        r = pkg.Microfacet(res, views, syn.PLANE_SIZE_CM, cl, dev)
        gt = torch.tensor(syn.downsample(gt_fine, finest // res), device=dev)[None]
        with torch.no_grad():
            target = r.eval(gt)
        if tex is None:
            tex = torch.tensor(syn.init_textures(gt[0].cpu().numpy(), seed=1), device=dev)[None].contiguous()
        else:
            tex = pkg.upsample_textures(tex)
        fs = pkg.FusedStep(r, tex, target, 0.01, loss="log_l2", history=iters, own_textures=True)
        g = fs._capture_graph(iters)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        fs.steps_launched = iters
        fs.write_back()
        losses = fs.losses()
        out["stages"][f"{res}^2"] = {"ms": round(ms, 4), "it_per_s": round(iters * 1000.0 / ms, 1),
                                     "loss_first": float(losses[0]), "loss_last": float(losses[-1])}
        first = out["stages"][f"{stages[0]}^2"]["loss_first"]
        if res != stages[0] and not losses[0] < first:
            raise RuntimeError(f"staged leg: the {res}^2 stage starts at loss {losses[0]:.4g}, not below the "
                               f"{stages[0]}^2 initial loss {first:.4g}: the stages are not one material")
        total += ms
    out["total_ms"] = round(total, 4)
    out["what"] = (f"{iters} fused log-L2 iterations per stage at {', '.join(f'{r}^2' for r in stages)} x {views} "
                   f"views on one material ({finest}^2 ground truth, box-averaged for the coarse stages), "
                   "bilinear x2 between stages, one graph replay per stage (first replay: includes the graph "
                   "launch); each later stage starts below the first stage's initial loss")
    return out


def light_leg(pkg, dev, res, views, per=100, n=20):
    """The fused step with the light power optimised too (src/svbrdf.py:47-50,
    58-59): one step launch plus a one-block finalize per iteration, replayed
    from graphs of `per` iterations; median us per iteration."""
    prob = make_problem(pkg, res, views, dev)
    r = prob["renderer"]
    r.update_light(r.light_pow.detach().to(torch.float32).reshape(3).clone().contiguous())
    fs = pkg.FusedStep(r, prob["tex"], prob["target"], lr=0.01, loss="log_l2", optim_light=True,
                       own_textures=True)
    g = fs._capture_graph(per)
    g.replay()
    sp = replay_spread(lambda k: g.replay(), per, n)
    nbytes = algorithmic_bytes_step(res, views)
    return {"us_per_step": sp["median"], "iqr": sp["iqr"], "value": round(1e6 / sp["median"], 1), "unit": "it/s",
            "frac": round(nbytes / (sp["median"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "light_pow_after": [round(float(x), 3) for x in fs.light_pow.cpu()]}


def cache_cold_leg(pkg, dev, res, views, per=50, n=10):
    """The fused step of the bench shape with its working set kept out of
    the Infinity Cache: two independent problems (2 x 226.6 MB at 1024^2 x 9)
    step alternately inside one graph, so each step finds the other one's
    data in the cache rather than its own.  Median us per step over n
    replays of 2 x `per` steps."""
    steps = [make_problem(pkg, res, views, dev)["fused"] for _ in range(2)]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        g.capture_begin()
        for _ in range(per):
            for fs in steps:
                fs._launch(ctypes.c_void_p(s.cuda_stream))
        g.capture_end()
    torch.cuda.current_stream().wait_stream(s)
    g.replay()
    sp = replay_spread(lambda k: g.replay(), 2 * per, n)
    nbytes = algorithmic_bytes_step(res, views)
    return {"us_per_step": sp["median"], "iqr": sp["iqr"], "value": round(1e6 / sp["median"], 1), "unit": "it/s",
            "achieved_GBs": round(nbytes / (sp["median"] * 1e-6) / 1e9, 1),
            "frac": round(nbytes / (sp["median"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "working_set_bytes": sum(working_set_bytes(fs, views) for fs in steps),
            "loss_after": [fs.loss() for fs in steps]}


def driver_leg(pkg, dev, res=1024, epochs=1000):
    """The reference's per-pixel workflow end to end (src/scripts.py:31-41,
    67-97; run.py:37's 1024^2 / lr 0.01 / 1000 epochs from random maps): a
    JSON configuration and 9 synthetic views on disk, render the targets,
    then optim_perpixel with every side output the reference writes (maps,
    re-renders and the loss plot after epoch 1, every 100th and the last).
    Wall-clock seconds, files written, and the optimisation loop alone."""
    import tempfile
    syn = pkg.synthetic
    cam, light = syn.flash_grid(n_views=9)
    with tempfile.TemporaryDirectory() as tmp:
        tmp = pathlib.Path(tmp)
        cfg = tmp / "cfg.json"
        cfg.write_text(json.dumps({"reference_dir": "ref", "target_dir": "target", "optimize_dir": "opt",
                                   "rerender_dir": "rr", "im_size": float(syn.PLANE_SIZE_CM), "idx": list(range(9)),
                                   "camera_pos": cam.tolist(), "light_pos": light.tolist(),
                                   "light_pow": syn.light_power().tolist()}))
        sio = pkg.SvbrdfIO(cfg, dev)
        sio.save_textures_th(torch.tensor(syn.ground_truth_textures(res, seed=0, albedo=ALBEDO), device=dev)[None],
                             sio.reference_dir)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pkg.scripts.render(cfg, res)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        opt = pkg.scripts.optim_perpixel(cfg, res, 0.01, epochs, "random")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        files = sum(1 for f in tmp.rglob("*") if f.is_file())
        o2 = pkg.SvbrdfOptim(dev, opt.renderer_obj)
        o2.init_from_randn()
        o2.load_targets(opt.targets)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        o2.optim(epochs, 0.01)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
    return {"workload": f"scripts.render + scripts.optim_perpixel at {res}^2 x 9, lr 0.01, {epochs} epochs from "
                        "random maps, with the reference's side outputs (PNG maps, re-renders, loss plot)",
            "render_s": round(t1 - t0, 3), "optim_perpixel_s": round(t2 - t1, 3),
            "epochs_per_s": round(epochs / (t2 - t1), 1), "files_written": files,
            "loss_first": opt.loss_image_list[0], "loss_last": opt.loss_image_list[-1],
            "loop_alone_s": round(t4 - t3, 4)}


def codec_leg(pkg, dev, res=1024, views=9, iters=20):
    """The map / target codec around the path (src/imageio.py, src/svbrdf.py:
    168-237) at the bench shape: bytes moved / time, against 8 TB/s."""
    from svbrdf_amd import imageio as io
    from svbrdf_amd import svbrdf_io as sio
    rng = np.random.default_rng(9)
    u8 = torch.from_numpy(rng.integers(0, 256, (views, res, res, 3), dtype=np.uint8)).to(dev)
    tex = (torch.rand((1, 9, res, res), generator=torch.Generator().manual_seed(9)) * 2 - 1).to(dev)
    half = tex[:, :, ::2, ::2].contiguous()
    maps = sio.maps_u8_from_textures(tex)
    small = sio.maps_u8_from_textures(half)
    npx = res * res
    L = pkg._lib
    st = L.stream_handle(dev)
    imgs = torch.empty((views, 3, res, res), dtype=torch.float32, device=dev)
    tex_out = torch.empty_like(tex)
    outs = [torch.empty_like(m) for m in maps]
    big = torch.empty((res, res, 3), dtype=torch.uint8, device=dev)
    # kernel-level calls through the C ABI into preallocated buffers (what
    # the roofline fraction is taken on), and the same work through the Python
    # mirror (allocations and host overhead included, "api_ms")
    cases = {
        "load_images (svbrdf_images_from_u8, srgb)": (
            lambda: L.call("svbrdf_images_from_u8", L.ptr(u8), views, res, res, 3, L.IMREAD_SRGB, L.ptr(imgs), st),
            lambda: io.decode(u8, "srgb"), views * npx * (3 + 12)),
        "load_textures (svbrdf_textures_from_u8)": (
            lambda: L.call("svbrdf_textures_from_u8", L.ptr(maps[0]), L.ptr(maps[1]), L.ptr(maps[2]),
                           L.ptr(maps[3]), 1, res, L.ptr(tex_out), st),
            lambda: sio.textures_from_maps_u8(*maps, res), npx * (10 + 36)),
        "save_textures (svbrdf_textures_to_u8)": (
            lambda: L.call("svbrdf_textures_to_u8", L.ptr(tex), res, *(L.ptr(o) for o in outs), st),
            lambda: sio.maps_u8_from_textures(tex), npx * (36 + 10)),
        f"resize {res // 2}->{res} x3 (svbrdf_resize_lanczos4_u8)": (
            lambda: L.call("svbrdf_resize_lanczos4_u8", L.ptr(small[0]), res // 2, res // 2, 3, L.ptr(big), res, res,
                           st),
            lambda: io.imresize(small[0], (res, res)), (npx // 4) * 3 + npx * 3),
        f"stage transition {res // 2}->{res} (save + 4 resizes + load)": (
            None, lambda: sio.reload_textures(half, res), None),
        f"stage transition in memory {res // 2}->{res} (svbrdf_upsample2x_bilinear)": (
            lambda: L.call("svbrdf_upsample2x_bilinear", L.ptr(half), res // 2, 0, L.ptr(tex_out), 0, 9, st),
            lambda: pkg.upsample_textures(half), 9 * 4 * (npx // 4 + npx)),
        # the ceiling of a kernel that writes these bytes: torch's fill of the
        # transition's output (write-only, the same 37.7 MB at 1024^2)
        f"write-only reference: fill of the {res}^2 x 9 transition output (torch fill_)": (
            lambda: tex_out.fill_(0.5), lambda: tex_out.fill_(0.5), 9 * 4 * npx),
    }
    out = {}
    for name, (kernel, api, nbytes) in cases.items():
        rec = {}
        if kernel is not None:
            for _ in range(3):
                kernel()
            ms = time_events(kernel, iters)
            rec["ms"] = round(ms, 5)
            rec.update(bytes=nbytes, achieved_GBs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                       frac=round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        for _ in range(3):
            api()
        rec["api_ms"] = round(time_events(api, iters), 5)
        out[name] = rec
    return out


def materialgan_leg(pkg, dev, res=256, views=9, iters=20):
    """The MaterialGAN caller (src/materialgan.py:101-153) at 256^2 x 9:
    StyleGAN2-shaped generator -> HIP render -> image MSE + 0.1 x VGG feature
    loss -> backward to the latent -> Adam, per iteration; and the share of
    the render forward + adjoint in it."""
    from svbrdf_amd import materialgan as MG
    syn = pkg.synthetic
    cam, light = syn.flash_grid(n_views=views)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(syn.light_power())]
    r = pkg.Microfacet(res, views, syn.PLANE_SIZE_CM, cl, dev)
    opt = MG.MaterialGANOptim(dev, r, seed=7)
    opt.init_from(seed=123)
    with torch.no_grad():
        targets = r.eval(opt.latent_to_textures(opt.latent))
    opt.init_from(seed=5)
    opt.load_targets(targets)
    opt.optim((3, 2, 1), 0.02)                                      # warm-up (records both epoch graphs)
    ms = time_events(lambda: opt.optim((iters, iters // 2, iters // 2), 0.02), 1) / iters
    # the same epochs launched op by op (torch.optim.Adam per epoch, as the
    # reference's loop): the host's launch time included
    opt.optim((2, 1, 1), 0.02, graph=False)
    ms_eager = time_events(lambda: opt.optim((iters, iters // 2, iters // 2), 0.02, graph=False), 1) / iters
    tex = opt.textures.clone().requires_grad_(True)
    go = torch.randn((views, 3, res, res), device=dev)

    def fb():
        out = r.eval(tex)
        (out * go).sum().backward()
    for _ in range(3):
        fb()
    ms_render = time_events(fb, 20)
    return {"workload": f"{res}^2 x {views} views, StyleGAN2-shaped generator (W+ 14 x 512, seeded weights), "
                        "VGG-19 feature loss r11/r12/r32/r42 (avg pool, seeded weights), alternating latent / noise "
                        "Adam, HIP render adjoint",
            "iters": iters, "value": round(1000.0 / ms, 2), "unit": "it/s", "ms_per_step": round(ms, 3),
            "epochs_from": "HIP graphs (one per epoch kind), bit-identical to the eager loop",
            "eager_ms_per_step": round(ms_eager, 3),
            "render_fwd_bwd_ms": round(ms_render, 4), "render_share": round(ms_render / ms, 4),
            "loss_image_last": opt.loss_image_list[-1]}


def compose_leg(pkg, dev, res=256, views=9, iters=100):
    """Render -> VGG-19-shaped random-init conv stack -> Gram L2 on
    5 layers + pixel L2 -> backward (HIP adjoint) -> torch Adam, 100 steps."""
    from svbrdf_amd import compose
    syn = pkg.synthetic
    gt = syn.ground_truth_textures(res, seed=0, albedo=ALBEDO)
    cam, light = syn.flash_grid(n_views=views)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(syn.light_power())]
    r = pkg.Microfacet(res, views, syn.PLANE_SIZE_CM, cl, dev)
    net = compose.VGGShapedFeatures(seed=4, device=dev)
    with torch.no_grad():
        target = r.eval(torch.tensor(gt, device=dev)[None])
    fl = compose.FeatureLoss(net, target)
    tex = torch.tensor(syn.init_textures(gt, seed=1), device=dev)[None].requires_grad_(True)
    opt = torch.optim.Adam([tex], lr=0.01)

    def it():
        loss = fl(r.eval(tex.clamp(-1, 1)))
        opt.zero_grad()
        loss.backward()
        opt.step()
        return loss

    first = it().item()
    for _ in range(2):
        it()
    ms = time_events(it, iters)
    last = it().item()
    # render forward + adjoint alone, same shapes and inputs
    t2 = tex.detach().clone().requires_grad_(True)
    gout = torch.randn(views, 3, res, res, device=dev)

    def fb():
        out = r.eval(t2)
        out.backward(gout)

    for _ in range(3):
        fb()
    ms_render = time_events(fb, 50)
    return {"workload": f"{res}^2 x {views} views, VGG-19-shaped conv1_1..conv5_1 (He init, seed 4), Gram L2 on "
                        "5 layers + pixel L2, torch Adam", "iters": iters,
            "value": round(1000.0 / ms, 2), "unit": "it/s", "ms_per_step": round(ms, 3),
            "render_fwd_bwd_ms": round(ms_render, 4), "render_share": round(ms_render / ms, 4),
            "loss_first": first, "loss_last": last}


def conv_layers_leg(pkg, dev, reps=20):
    """Every 3x3 layer shape of the two networks at the legs' sizes -- the
    VGG-19-shaped stack on 9 images of 256^2 (forward and input-gradient
    shapes) and the 256^2 generator at batch 1 (plus its weight gradient) --
    on the HIP kernels (svbrdf_conv3x3 with a bias + ReLU epilogue,
    svbrdf_conv3x3_wgrad) and on F.conv2d + ReLU / conv2d_weight (MIOpen),
    HIP events around `reps` launches after 3 warm-ups; TFLOP/s at
    2 * 9 * Cin * Cout * pixels."""
    from svbrdf_amd import convnet
    import torch.nn.functional as F
    shapes = []
    for cin, cout, r in [(3, 64, 256), (64, 64, 256), (64, 128, 128), (128, 128, 128), (128, 256, 64),
                         (256, 256, 64), (256, 512, 32), (512, 512, 32), (512, 512, 16)]:
        shapes += [("vgg", 9, cin, cout, r), ("vgg dL/dx", 9, cout, cin, r)]
    for cin, cout, r in [(512, 512, 4), (512, 512, 8), (512, 512, 16), (512, 512, 32), (512, 512, 64),
                         (512, 256, 128), (256, 256, 128), (256, 128, 256), (128, 128, 256)]:
        shapes.append(("generator", 1, cin, cout, r))
    g = torch.Generator().manual_seed(11)
    rows, tot_hip, tot_lib, tot_whip, tot_wlib = [], 0.0, 0.0, 0.0, 0.0
    for net, n, cin, cout, r in shapes:
        x = torch.randn((n, cin, r, r), generator=g).to(dev)
        w = (torch.randn((cout, cin, 3, 3), generator=g) / (cin * 9) ** 0.5).to(dev)
        b = torch.zeros(cout, device=dev)
        packed = convnet.pack3x3(w)

        def hip():
            convnet.conv3x3_packed(x, packed, cout, b, None, 0.0, "relu")

        def lib():
            torch.relu(F.conv2d(x, w, b, padding=1))
        for fn in (hip, lib):
            for _ in range(3):
                fn()
        t_hip = time_events(hip, reps) * 1e3
        t_lib = time_events(lib, reps) * 1e3
        flops = 2.0 * 9 * cin * cout * n * r * r
        tot_hip += t_hip
        tot_lib += t_lib
        row = {"net": net, "n": n, "cin": cin, "cout": cout, "side": r, "hip_us": round(t_hip, 1),
               "hip_tflops": round(flops / t_hip / 1e6, 1), "miopen_us": round(t_lib, 1),
               "miopen_tflops": round(flops / t_lib / 1e6, 1)}
        if net == "generator":
            # the modulated weights' gradient (latent epochs): svbrdf_conv3x3_wgrad
            # against torch.nn.grad.conv2d_weight
            gz = torch.randn((n, cout, r, r), generator=g).to(dev)

            def hip_w():
                convnet.conv3x3_wgrad(x, gz, cout)

            def lib_w():
                torch.nn.grad.conv2d_weight(x, (cout, cin, 3, 3), gz, padding=1)
            for fn in (hip_w, lib_w):
                for _ in range(3):
                    fn()
            tw_hip = time_events(hip_w, reps) * 1e3
            tw_lib = time_events(lib_w, reps) * 1e3
            tot_whip += tw_hip
            tot_wlib += tw_lib
            row.update(hip_wgrad_us=round(tw_hip, 1), hip_wgrad_tflops=round(flops / tw_hip / 1e6, 1),
                       miopen_wgrad_us=round(tw_lib, 1), miopen_wgrad_tflops=round(flops / tw_lib / 1e6, 1))
        rows.append(row)
    return {"workload": "3x3 layers of the VGG-19-shaped stack (9 x 256^2, forward and dL/dx shapes) and the 256^2 "
                        "StyleGAN2-shaped generator (batch 1), fp32; HIP vs F.conv2d (MIOpen)",
            "total_hip_us": round(tot_hip, 1), "total_miopen_us": round(tot_lib, 1),
            "generator_wgrad_total_hip_us": round(tot_whip, 1), "generator_wgrad_total_miopen_us": round(tot_wlib, 1),
            "layers": rows}


def host_cpu():
    """(threads to use, description) for the CPU legs: every core of this
    process's affinity set, capped at OMP_NUM_THREADS when the environment
    sets it (the GPU box allots each GPU a share of its host cores that
    way), and the CPU model from /proc/cpuinfo."""
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, {"cpu_model": model, "affinity_cores": affinity, "omp_num_threads": omp}


def cpu_forward(res, views, n=3):
    """One forward render of the eager restatement on the host cores at the
    bench shape (the CPU counterpart of the forward line): median of n."""
    from oracle import microfacet_torch as MT   # the CPU baseline leg only
    pkg = __graft_entry__.load_package()
    syn = pkg.synthetic
    threads, host = host_cpu()
    torch.set_num_threads(threads)
    cam, light = syn.flash_grid(n_views=views)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(syn.light_power())]
    r = MT.MicrofacetEager(res, views, syn.PLANE_SIZE_CM, cl, "cpu")
    gt = torch.tensor(syn.ground_truth_textures(res, seed=0, albedo=ALBEDO))[None]
    times = []
    with torch.no_grad():
        for _ in range(n):
            t0 = time.perf_counter()
            r.eval(gt)
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(1.0 / med, 3), "unit": "renders/s", "cores": threads, "kind": "port", **host,
            "sample": f"{n} forward renders of the eager restatement at {res}^2 x {views} views, median s = "
                      f"{med:.3f}"}


def cpu_baseline(res, views, seconds):
    """The oracle's eager restatement on the host cores: the reference's CPU
    path (src/scripts.py:68 falls back to device 'cpu')."""
    from oracle import microfacet_torch as MT   # the CPU baseline leg only
    pkg = __graft_entry__.load_package()
    syn = pkg.synthetic
    threads, host = host_cpu()
    torch.set_num_threads(threads)
    gt = syn.ground_truth_textures(res, seed=0, albedo=ALBEDO)
    init = syn.init_textures(gt, seed=1)
    cam, light = syn.flash_grid(n_views=views)
    P = syn.light_power()
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = MT.MicrofacetEager(res, views, syn.PLANE_SIZE_CM, cl, "cpu")
    with torch.no_grad():
        target = r.eval(torch.tensor(gt)[None])
    tex = torch.tensor(init)[None].requires_grad_(True)
    opt = torch.optim.Adam([tex], lr=0.01, betas=(0.9, 0.999))
    log_tgt = torch.log(target.clamp(MT.target_floor(), 1))
    times = []
    t_end = time.perf_counter() + seconds
    while True:
        t0 = time.perf_counter()
        out = r.eval(tex.clamp(-1, 1))
        loss = torch.nn.functional.mse_loss(torch.log(out), log_tgt)
        loss.item()
        opt.zero_grad()
        loss.backward()
        opt.step()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 50:
            break
    timed = times[1:] if len(times) > 1 else times
    s = float(np.median(timed))
    return {"value": 1.0 / s, "unit": "it/s", "cores": threads, "kind": "port", **host,
            "sample": f"{len(timed)} timed iteration(s) (+1 warm-up) of the eager restatement at {res}^2 x "
                      f"{views} views, log-L2 + torch.optim.Adam, median s/it = {s:.3f}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    # a 4,000-launch trace of the step settles at 52.4-52.5 us after ~500
    # back-to-back launches (56.5 us average over the first 250)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--views", type=int, default=9)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-torch", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-material batched leg")
    ap.add_argument("--no-compose", action="store_true", help="skip the feature-loss composition leg (256^2 x 9)")
    ap.add_argument("--no-codec", action="store_true", help="skip the map / target codec leg")
    ap.add_argument("--no-materialgan", action="store_true", help="skip the MaterialGAN caller leg")
    ap.add_argument("--no-conv-layers", action="store_true",
                    help="skip the per-layer HIP vs MIOpen timing of the two networks' 3x3 convolutions")
    ap.add_argument("--no-light", action="store_true", help="skip the light-power optimisation leg")
    ap.add_argument("--no-cache-cold", action="store_true",
                    help="skip the leg that keeps the step's working set out of the Infinity Cache")
    ap.add_argument("--no-staged", action="store_true", help="skip the 256^2 -> 512^2 -> 1024^2 staged leg")
    ap.add_argument("--no-driver", action="store_true",
                    help="skip the end-to-end optim_perpixel leg (files in a temporary directory)")
    ap.add_argument("--eager", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--no-eager-probes", action="store_true",
                    help="skip the single-launch and per-iteration-sync timings (so that a profiler's kernel "
                         "statistics cover the graph-replayed steps only)")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="iterations captured per HIP graph (each graph launch costs ~35 us on the GPU, so "
                         "a long optimisation loop is captured in large pieces)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the bench, run 5 plain device copies of MIB MiB (a known-size kernel that "
                         "calibrates the profiler's FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the fused step runs on a single MI355X here. Across GPUs it would "
              "shard by row bands of the maps, with no exchange on the gradient path.", file=sys.stderr)
        sys.exit(2)

    pkg = __graft_entry__.load_package()
    dev = torch.device("cuda:0")
    res, views = args.res, args.views
    prob = make_problem(pkg, res, views, dev)
    fs = prob["fused"]

    # fused step: K iterations as eager launches, or as replays of a graph of
    # GRAPH_STEPS iterations (plus one graph for the remainder), so exactly K
    # iterations run in the timed region
    if args.eager:
        def run(n):
            for _ in range(n):
                fs.step()
    else:
        per = max(1, min(args.graph_steps, max(args.steps, 1)))
        graphs = {per: fs.capture(per)}

        def run(n):
            for _ in range(n // per):
                graphs[per].replay()
            rem = n % per
            if rem:
                if rem not in graphs:
                    graphs[rem] = fs.capture(rem)
                graphs[rem].replay()
        run(per + args.steps % per)          # captures every graph before the timed region
    run(args.warmup)
    torch.cuda.synchronize()
    ms = time_events(lambda: run(args.steps), 1) / args.steps
    loss_after = fs.loss()
    value = 1000.0 / ms
    spread = replay_spread(run, args.graph_steps if not args.eager else 1, 50)
    # the kernel's average duration in the timed configuration: HIP events
    # around each graph replay (graph_steps back-to-back launches, nothing
    # else), median over 50 replays; with --eager, around single launches
    kernel_ms = spread["median"] / 1000.0
    probes = {}
    if not args.no_eager_probes:
        # one eager launch between two events: the events' end-of-kernel cache
        # write-back lands inside the interval, so this reads longer
        probes["eager_launch_ms"] = round(kernel_duration_ms(fs.step, 50), 5)
        probes["with_sync"] = with_sync_rate(fs, 50)

    nbytes = algorithmic_bytes_step(res, views)
    achieved = nbytes / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_source = measured_traffic(res, views)
    ws_bytes = working_set_bytes(fs, views)
    regime = ("infinity-cache-resident: the working set between iterations fits the 256 MiB Infinity Cache, so "
              "part of each iteration's reads come from it; hbm_resident_frac is the same step with its working "
              "set kept out of the cache (cache_cold)" if ws_bytes <= INFINITY_CACHE_BYTES else
              "hbm-resident: the working set exceeds the 256 MiB Infinity Cache")
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "it/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"fused per-pixel iteration (clamp+render+log-L2+adjoint+Adam) {res}^2 x {views} "
                               f"views, one kernel per step, "
                               + ("eager launches" if args.eager else f"HIP graphs of {args.graph_steps} steps"),
                   "res": res, "views": views, "loss": "log_l2", "graph": not args.eager,
                   "graph_steps": None if args.eager else args.graph_steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_source,
                     "algorithmic_bytes": nbytes, "kernel": "fused_step_kernel",
                     "kernel_ms": round(kernel_ms, 5),
                     "step_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     # bytes the step keeps between iterations (maps, moments, prepared
                     # targets at their plane strides): at or below the 256 MB
                     # Infinity Cache, part of each iteration's reads are served
                     # from it, and `traffic` (L2-to-fabric bytes) cannot tell them apart
                     "working_set_bytes": ws_bytes,
                     "infinity_cache_bytes": INFINITY_CACHE_BYTES,
                     "regime": regime, "hbm_resident_frac": None},
        "loss_after": loss_after,
        "per_replay_us_per_step": spread,
    }
    out.update(probes)

    # forward render at the same shape (1024^2 x 9 views)
    r = prob["renderer"]
    gt = torch.tensor(prob["gt"], device=dev)[None]
    # untimed launches first: the clock settles over the first ~1,000
    # back-to-back launches of these kernels (a 1,500-launch trace of the
    # adjoint falls from 38 to 30.6 us per launch, then holds; the forward
    # render holds from the start)
    with torch.no_grad():
        for _ in range(WARM_KERNEL_LAUNCHES):
            r.eval(gt)
        fms = time_events(lambda: r.eval(gt), 200)
    fb = algorithmic_bytes_render(res, views)
    out["forward"] = {"value": round(1000.0 / fms, 1), "unit": "renders/s", "ms": round(fms, 5),
                      "achieved_GBs": round(fb / (fms * 1e-3) / 1e9, 1),
                      "frac": round(fb / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    # the render adjoint on its own (RenderFunction.backward: dL/dimages ->
    # dL/dtextures), as a torch-composed loss drives it
    gout = torch.randn((views, 3, res, res), generator=torch.Generator().manual_seed(5)).to(dev)
    gtex = torch.empty_like(gt)
    L = pkg._lib
    stream = L.stream_handle(dev)

    def adjoint():
        L.call("svbrdf_render_backward_terms", L.ptr(gt), res, views, r.size, L.ptr(r.camera_pos),
               L.ptr(r.light_pos), L.ptr(r.light_pow), 0, L.ptr(gout), L.ptr(gtex), None, None, 0, stream)
    for _ in range(WARM_KERNEL_LAUNCHES):
        adjoint()
    bms = time_events(adjoint, 200)
    bb = algorithmic_bytes_adjoint(res, views)
    out["adjoint"] = {"value": round(1000.0 / bms, 1), "unit": "adjoints/s", "ms": round(bms, 5),
                      "achieved_GBs": round(bb / (bms * 1e-3) / 1e9, 1),
                      "frac": round(bb / (bms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    if not args.no_light:
        out["optim_light"] = light_leg(pkg, dev, res, views)
    if not args.no_cache_cold:
        out["cache_cold"] = cache_cold_leg(pkg, dev, res, views)
        out["roofline"]["hbm_resident_frac"] = out["cache_cold"]["frac"]
    out["hbm_copy"] = copy_bandwidth(dev)
    if not args.no_staged:
        out["staged"] = staged_leg(pkg, dev, views)
    if not args.no_batch:
        out["batched"] = batch_leg(pkg, dev)
    if not args.no_compose:
        out["composed"] = compose_leg(pkg, dev)
    if not args.no_codec:
        out["codec"] = codec_leg(pkg, dev, res, views)
    if not args.no_materialgan:
        out["materialgan"] = materialgan_leg(pkg, dev)
    if not args.no_conv_layers:
        out["conv_layers"] = conv_layers_leg(pkg, dev)
    if not args.no_driver:
        out["driver"] = driver_leg(pkg, dev, res)
    if not args.no_torch:
        try:
            tb = torch_gpu_baseline(prob, res, views, dev)
            out["torch_gpu_baseline"] = tb
            out["speedup_vs_torch_gpu"] = round(value / tb["value"], 1)
            out["torch_gpu_baseline_nosync"] = torch_gpu_baseline(prob, res, views, dev, sync=False)
        except torch.cuda.OutOfMemoryError as e:     # pragma: no cover
            out["torch_gpu_baseline"] = {"error": str(e)[:200]}
    if not args.no_cpu:
        cb = cpu_baseline(res, views, args.cpu_seconds)
        out["cpu_baseline"] = cb
        # 256^2 x 9 (the reference's CPU-runnable case) as well
        out["cpu_baseline_256"] = cpu_baseline(256, views, min(args.cpu_seconds, 5.0))
        out["cpu_forward"] = cpu_forward(res, views)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = n * 4
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
