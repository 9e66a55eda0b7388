"""Batched rendering (many scenes in shared passes) and hipGraph-captured
passes: every image equals its own single-scene render bit for bit."""
import numpy as np
import pytest
import torch

from aurora_amd import scenes
from aurora_amd.core import RenderSettings, render_batch

pytestmark = pytest.mark.gpu


def zeros(cam):
    return torch.zeros((cam.image_height, cam.image_width, 3), dtype=torch.float32, device="cuda")


def workload(ctx):
    items = []
    specs = [
        (scenes.cornell_box(), scenes.cornell_camera(40), RenderSettings(spp=5, max_depth=50)),
        (scenes.random_spheres(0), scenes.random_spheres_camera(48, 27),
         RenderSettings(spp=3, max_depth=50, background=(0.7, 0.8, 1.0), aov=True)),
        (scenes.stress_scene(3000, 3000, seed=9), scenes.stress_camera(32, 18),
         RenderSettings(spp=2, max_depth=16, background=(0.7, 0.8, 1.0), tmin=1e-4)),
    ]
    specs.append((scenes.textured_scene(1), scenes.textured_camera(32, 24),
                  RenderSettings(spp=2, max_depth=12, background=(0.6, 0.7, 0.9), aov=True)))
    for seed in (1, 2, 3):
        arr, cam, spp = scenes.batch_scene(seed, max_prims=5000)
        cam.image_width = cam.image_height = 32
        specs.append((arr, cam, RenderSettings(spp=spp, max_depth=20, background=(0.5, 0.6, 0.8))))
    for arr, cam, rs in specs:
        sc = ctx.upload(arr)
        sc.build_bvh()
        items.append((sc, cam, rs))
    return items


# fused-tail switch points: never (wavefront kernels only), part of the
# pass, and always (one persistent kernel carries every path)
TAILS = [0, 4096, (1 << 32) - 1]


def test_batch_equals_single_renders(ctx):
    print("batch vs single", flush=True)
    items = workload(ctx)
    singles, seg = [], 0
    ctx.set_launch(tail_paths=0)
    try:
        for sc, cam, rs in items:
            rgb, alb, nrm = zeros(cam), zeros(cam), zeros(cam)
            st = sc.render_device(cam, rs, rgb, alb if rs.aov else None, nrm if rs.aov else None)
            singles.append((rgb.cpu().numpy(), alb.cpu().numpy() if rs.aov else None))
            seg += st["segments"]
        for tail, wide in [(t, False) for t in TAILS] + [(TAILS[-1], True)]:
            ctx.set_launch(tail_paths=tail)
            ctx.set_render_tree(wide)
            for graph in (False, True):
                bufs = [(zeros(cam), zeros(cam), zeros(cam)) for _, cam, _ in items]
                st = render_batch(ctx, [(sc, cam, rs, b[0], b[1] if rs.aov else None, b[2] if rs.aov else None)
                                        for (sc, cam, rs), b in zip(items, bufs)], graph=graph)
                print("tail", tail, "wide", wide, "graph", graph, "paths", st["paths"], "segments", st["segments"],
                      "single", seg, flush=True)
                assert st["segments"] == seg
                for k, ((rgb, alb, _), (s_rgb, s_alb)) in enumerate(zip(bufs, singles)):
                    assert np.array_equal(rgb.cpu().numpy(), s_rgb), (tail, graph, k)
                    if s_alb is not None:
                        assert np.array_equal(alb.cpu().numpy(), s_alb), (tail, graph, k)
    finally:
        ctx.set_launch()
        ctx.set_render_tree(False)


@pytest.mark.parametrize("tail", TAILS)
def test_graph_render_equals_eager(ctx, tail):
    print("graph vs eager, tail", tail, flush=True)
    sc = ctx.upload(scenes.cornell_box())
    sc.build_bvh()
    cam = scenes.cornell_camera(64)
    a, b = zeros(cam), zeros(cam)
    ctx.set_launch(tail_paths=tail)
    try:
        sc.render_device(cam, RenderSettings(spp=12, spp_per_pass=2), a)
        st = sc.render_device(cam, RenderSettings(spp=12, spp_per_pass=2, graph=True), b)
    finally:
        ctx.set_launch()
    assert st["passes"] == 6
    assert torch.equal(a, b)
    # replays across calls with eager work and synchronisations between them:
    # each graph-mode call captures its first pass and replays it twice; the
    # image is checked against the eager one after every call
    ctx.set_launch(tail_paths=tail)
    try:
        for i in range(4):
            sc.render_device(cam, RenderSettings(spp=6, spp_per_pass=2, sample_offset=12 + 6 * i), a)
            scratch = (a * 0.5 + 1.0).sum()  # unrelated eager work on the default stream
            torch.cuda.synchronize()
            sc.render_device(cam, RenderSettings(spp=6, spp_per_pass=2, sample_offset=12 + 6 * i, graph=True), b,
                             stats=False)
            ctx.synchronize()
            assert torch.isfinite(scratch)
            assert torch.equal(a, b), (tail, i)
    finally:
        ctx.set_launch()


def test_pass_splits_do_not_change_images(ctx):
    """Samples per pass, explicit (1, 2 or 3) or automatic, mixed within one
    batch: every image equals its single render in one pass bit for bit
    (sums are added in sample order whatever the split)."""
    import dataclasses
    print("pass splits", flush=True)
    items = workload(ctx)
    singles = []
    for sc, cam, rs in items:
        rgb = zeros(cam)
        sc.render_device(cam, dataclasses.replace(rs, spp_per_pass=rs.spp, aov=False), rgb)
        singles.append(rgb.cpu().numpy())
    for graph in (False, True):
        bufs = [zeros(cam) for _, cam, _ in items]
        its = [(sc, cam, dataclasses.replace(rs, spp_per_pass=(0, 1, 2, 3)[k % 4], aov=False), b, None, None)
               for k, ((sc, cam, rs), b) in enumerate(zip(items, bufs))]
        render_batch(ctx, its, graph=graph)
        for k, (b, s) in enumerate(zip(bufs, singles)):
            assert np.array_equal(b.cpu().numpy(), s), (graph, k)


def test_queue_sizes_do_not_change_images(ctx):
    """aurora_set_queue: the most paths per pass (here 2^16, one sample of a
    256^2 image per pass) and the paths per work-pool claim (16 and 1024)
    change speed only: the sums equal the default render's bit for bit."""
    from aurora_amd import scenes

    print("queue sizes", flush=True)
    g = ctx.upload(scenes.cornell_box())
    g.build_bvh()
    cam = scenes.cornell_camera(256)
    rs = RenderSettings(spp=6, max_depth=50, seed=3, aov=True)
    ctx.set_queue()
    ref, ra, rn, st0 = g.render(cam, rs)
    try:
        for pass_paths, claim in [(1 << 16, 16), (1 << 16, 1024), (0, 16), (1 << 17, 0)]:
            ctx.set_queue(pass_paths, claim)
            rgb, alb, nrm, st = g.render(cam, rs)
            assert np.array_equal(rgb, ref) and np.array_equal(alb, ra) and np.array_equal(nrm, rn)
            assert st["segments"] == st0["segments"]
            if pass_paths == 1 << 16:
                assert st["passes"] == 6
        with pytest.raises(ValueError, match="pass_paths"):
            ctx.set_queue(100, 0)
        with pytest.raises(ValueError, match="claim_paths"):
            ctx.set_queue(0, 2000)
    finally:
        ctx.set_queue()
        g.close()
