"""Benchmark: paths/s of the wavefront path tracer on the Cornell box at
1024x1024 (max depth 50, Philox seed 0) on one MI355X.

A step is one wavefront pass of S samples per pixel over the whole image
(raygen, up to 50 x {extend, shade}, accumulate), with the scene, LBVH and
accumulation buffer already resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "aurora-rendering-engine_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "paths/sec (1 GPU); time to relMSE<=1e-3, denoised vs plain, Cornell 1024²"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(arr, cam, spp, seconds=12.0):
    """The oracle (same algorithm, FP32) on the host cores, on a bounded
    sample of the same workload: whole-image passes of `spp` samples at
    successive sample offsets until ~`seconds` of wall time."""
    from aurora_amd.core import RenderSettings
    from oracle import orc

    threads = min(16, os.cpu_count() or 1)
    o = orc.OracleScene(arr)
    o.build_bvh()
    passes, paths, elapsed = 0, 0, 0.0
    while elapsed < seconds:
        rs = RenderSettings(spp=spp, sample_offset=passes * spp, max_depth=50, seed=0)
        t0 = time.perf_counter()
        o.render(cam, rs, threads=threads)
        elapsed += time.perf_counter() - t0
        paths += cam.image_width * cam.image_height * spp
        passes += 1
    return {"value": paths / elapsed, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"{passes} whole-image passes of the {cam.image_width}x{cam.image_height} Cornell box at "
                      f"{spp} spp, depth 50 ({paths} paths, {elapsed:.1f} s, FP32 oracle)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--spp-per-step", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tail-paths", type=int, default=None, help="fused-tail switch (default: library's)")
    ap.add_argument("--sweep-tail", action="store_true", help="time several fused-tail switch points")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the render would shard by image tiles or sample ranges with one final gather)", file=sys.stderr)
        sys.exit(2)

    import torch

    import aurora_amd
    from aurora_amd import scenes
    from aurora_amd.core import RenderSettings

    ctx = aurora_amd.Context(0)
    if args.tail_paths is not None:
        ctx.set_launch(tail_paths=args.tail_paths)
    arr = scenes.cornell_box()
    scene = ctx.upload(arr)
    bvh = scene.build_bvh()
    W = H = args.width
    cam = scenes.cornell_camera(W, H)
    S = args.spp_per_step
    rgb = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")

    def step(i, count=False):
        rs = RenderSettings(spp=S, sample_offset=i * S, max_depth=50, seed=0, spp_per_pass=S,
                            count_fetches=count)
        return scene.render_device(cam, rs, rgb)

    if args.sweep_tail:
        for tp in (0, 65536, 262144, 1 << 20, 4 << 20, 1 << 30):
            ctx.set_launch(tail_paths=tp)
            step(0)
            ctx.synchronize()
            t0 = time.perf_counter()
            for i in range(5):
                step(1 + i)
            ctx.synchronize()
            dt = (time.perf_counter() - t0) / 5
            print(f"sweep tail_paths={tp}: {dt * 1e3:.2f} ms/step, {W * H * S / dt / 1e6:.1f} Mpaths/s",
                  file=sys.stderr, flush=True)
        ctx.set_launch(tail_paths=args.tail_paths if args.tail_paths is not None else 65536)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    ctx.synchronize()
    steps = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        steps.append(step(args.warmup + i))
    ctx.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    paths = sum(s["paths"] for s in steps)
    segments = sum(s["segments"] for s in steps)
    ext_ms = sum(s["extend_ms"] for s in steps)
    launches = sum(s["extend_launches"] for s in steps)
    # algorithmic traversal bytes of the first timed step: 32 B per node,
    # 16 B per sphere, 48 B per quad record fetched (counted by the same
    # traversal, rerun untimed with counting on)
    cnt = step(args.warmup, count=True)
    alg_bytes = 32 * cnt["nodes_fetched"] + 16 * cnt["spheres_fetched"] + 48 * cnt["quads_fetched"]
    first = steps[0]
    achieved = alg_bytes / (first["extend_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": paths / elapsed,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"cornell_box_{W}x{H}_depth50", "width": W, "height": H, "spp_per_step": S,
                   "max_depth": 50, "seed": 0, "n_quads": int(arr.n_prims), "bvh_nodes": bvh["n_nodes"]},
        "rays_per_s": segments / elapsed,
        "segments_per_path": segments / paths,
        "extend_ms_per_step": ext_ms / args.steps,
        "extend_share": ext_ms / (elapsed * 1e3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_extend", "alg_bytes_per_launch": alg_bytes / max(1, first["extend_launches"]),
                     "avg_launch_ms": first["extend_ms"] / max(1, first["extend_launches"])},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(arr, cam, S)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
