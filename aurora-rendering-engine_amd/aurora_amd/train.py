"""Train the denoiser on images the path tracer renders itself.

Dataset: Cornell-box variants (scenes.cornell_variant, seeds 0..n-1) at
S x S pixels; for every scene, noisy inputs at several sample counts (sums
of radiance, albedo and normal, AOVs on) and a high-sample-count target
rendered from an independent sample range.  Training: random crops, one
input sample count per step, Adam on the library's forward/backward kernels.
Evaluation on held-out seeds: relMSE of the plain input mean and of the
denoised image against the target, per sample count.

    python -m aurora_amd.train --scenes 256 --size 256 --steps 3000 --out weights.f32
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def render_sums(scene, cam, spp, offset, aov, max_depth=50, seed=0):
    import torch

    from .core import RenderSettings
    H, W = cam.image_height, cam.image_width
    rgb = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    alb = torch.zeros_like(rgb) if aov else None
    nrm = torch.zeros_like(rgb) if aov else None
    scene.render_device(cam, RenderSettings(spp=spp, sample_offset=offset, max_depth=max_depth, seed=seed, aov=aov),
                        rgb, alb, nrm)
    return rgb, alb, nrm


def make_dataset(ctx, seeds, size, in_spp, target_spp, progress=True):
    import torch

    from . import scenes
    n = len(seeds)
    data = {"in_spp": list(in_spp), "size": size, "seeds": list(seeds)}
    data["target"] = torch.zeros((n, size, size, 3), dtype=torch.float32, device="cuda")
    for k in in_spp:
        data[k] = [torch.zeros((n, size, size, 3), dtype=torch.float32, device="cuda") for _ in range(3)]
    cam = scenes.cornell_camera(size)
    t0 = t_log = time.time()
    for i, sd in enumerate(seeds):
        scene = ctx.upload(scenes.cornell_variant(sd))
        scene.build_bvh()
        rgb, _, _ = render_sums(scene, cam, target_spp, 0, aov=False)
        data["target"][i] = rgb / target_spp
        off = target_spp
        for k in in_spp:
            r, a, nm = render_sums(scene, cam, k, off, aov=True)
            off += k
            data[k][0][i], data[k][1][i], data[k][2][i] = r, a, nm
        scene.close()
        if progress and (i % 16 == 0 or i == n - 1 or time.time() - t_log > 30):
            t_log = time.time()
            log(f"dataset: scene {i + 1}/{n} at {size}x{size} rendered ({t_log - t0:.1f} s)")
    return data


def train(ctx, datasets, steps, batch=8, crop=128, lr=2e-3, seed=0, net=None, log_every=100):
    """datasets: one dataset or a list; each step draws one dataset
    uniformly, then one input sample count and `batch` random crops."""
    import torch

    from .denoiser import Denoiser
    net = net or Denoiser(ctx, seed=seed)
    g = torch.Generator(device="cpu").manual_seed(seed)
    if isinstance(datasets, dict):
        datasets = [datasets]
    t0 = time.time()
    hist = []
    for it in range(steps):
        data = datasets[int(torch.randint(len(datasets), (1,), generator=g))]
        n = data["target"].shape[0]
        size = data["size"]
        levels = data["in_spp"]
        k = levels[int(torch.randint(len(levels), (1,), generator=g))]
        idx = torch.randint(n, (batch,), generator=g)
        oy = torch.randint(0, size - crop + 1, (batch,), generator=g) // 8 * 8
        ox = torch.randint(0, size - crop + 1, (batch,), generator=g) // 8 * 8
        # random flips keep the Cornell statistics but multiply the data
        fl = torch.randint(2, (batch,), generator=g)

        def crops(t):
            out = torch.stack([t[int(idx[b]), int(oy[b]):int(oy[b]) + crop, int(ox[b]):int(ox[b]) + crop]
                               for b in range(batch)])
            flip = fl.bool().to(out.device)
            out[flip] = out[flip].flip(2)
            return out.contiguous()

        r, a, nm = (crops(t) for t in data[k])
        nm[fl.bool().to(nm.device), ..., 0] *= -1.0  # mirrored normals
        tg = crops(data["target"])
        cur_lr = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * it / max(1, steps))))
        loss = net.train_step(r, a, nm, tg, k, cur_lr)
        hist.append(loss)
        if (it + 1) % log_every == 0 or it == 0:
            log(f"train: step {it + 1}/{steps} loss {np.mean(hist[-log_every:]):.5f} lr {cur_lr:.2e} "
                f"({time.time() - t0:.1f} s)")
    return net, hist


def evaluate(net, data):
    """relMSE (against the target) of the plain mean and of the denoised
    image, averaged over the scenes of `data`, per input sample count."""
    from .metric import relmse
    res = {}
    n = data["target"].shape[0]
    for k in data["in_spp"]:
        pl, dn = [], []
        for i in range(n):
            r, a, nm = (t[i].contiguous() for t in data[k])
            tgt = data["target"][i]
            pl.append(relmse(r / k, tgt))
            dn.append(relmse(net.denoise(r, a, nm, k), tgt))
        res[k] = {"plain_relmse": float(np.mean(pl)), "denoised_relmse": float(np.mean(dn))}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--in-spp", type=str, default="1,4,16,64,256")
    ap.add_argument("--target-spp", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--hires-scenes", type=int, default=0, help="extra variants rendered at --hires-size")
    ap.add_argument("--hires-size", type=int, default=1024)
    ap.add_argument("--hires-in-spp", type=str, default="256,1024")
    ap.add_argument("--hires-target-spp", type=int, default=2048)
    ap.add_argument("--init", type=str, default=None, help="start from these weights")
    ap.add_argument("--loss", choices=["rel", "log"], default="rel",
                    help="relative L2 (the metric's form) or L2 in log(1 + radiance)")
    ap.add_argument("--out", type=str, required=True)
    ap.add_argument("--report", type=str, default=None)
    args = ap.parse_args()
    import aurora_amd
    from .denoiser import Denoiser
    ctx = aurora_amd.Context(0)
    in_spp = [int(x) for x in args.in_spp.split(",")]
    t0 = time.time()
    data = make_dataset(ctx, list(range(args.scenes)), args.size, in_spp, args.target_spp)
    held = make_dataset(ctx, list(range(100000, 100000 + args.heldout)), args.size, in_spp, args.target_spp)
    sets = [data]
    if args.hires_scenes:
        hin = [int(x) for x in args.hires_in_spp.split(",")]
        sets.append(make_dataset(ctx, list(range(200000, 200000 + args.hires_scenes)), args.hires_size, hin,
                                 args.hires_target_spp))
    t_data = time.time() - t0
    net = Denoiser(ctx, seed=0)
    net.set_loss(args.loss)
    if args.init:
        net.load(args.init)
    before = evaluate(net, held)
    t1 = time.time()
    net, hist = train(ctx, sets, args.steps, args.batch, args.crop, args.lr, net=net)
    t_train = time.time() - t1
    net.save(args.out)
    after = evaluate(net, held)
    rep = {"scenes": args.scenes, "size": args.size, "in_spp": in_spp, "target_spp": args.target_spp,
           "hires_scenes": args.hires_scenes, "hires_size": args.hires_size,
           "steps": args.steps, "batch": args.batch, "loss": args.loss, "init": args.init, "crop": args.crop, "dataset_s": t_data, "train_s": t_train,
           "final_loss": float(np.mean(hist[-100:])), "heldout_before": before, "heldout_after": after}
    log(json.dumps(rep))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
