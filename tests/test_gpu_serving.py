"""The batched-serving workload at its full size: 64 seeded scenes of uneven
size (10 to 200k spheres and quads, 256^2 to 1024^2, 1 to 16 spp, depth 16,
first-hit AOVs), rendered in shared passes under every launch shape the
autotuner chooses among (traversal grid 8 or 16 blocks per CU; fused-tail
switch at 2^20 live paths or the whole pass in the persistent kernel;
eager or hipGraph-captured passes; binary or four-wide render tree).  Every
image, albedo and normal sum equals its own single-scene render bit for bit
in every shape, and 256 seeded pixels of every scene equal the CPU oracle."""
import numpy as np
import pytest
import torch

from aurora_amd import autotune, scenes
from aurora_amd.core import RenderSettings, render_batch
from oracle import orc

pytestmark = pytest.mark.gpu

TAIL_ALWAYS = (1 << 32) - 1
SHAPES = [dict(extend_blocks_per_cu=e, shade_blocks_per_cu=8, tail_paths=t, graph=g, wide=w)
          for e in (8, 16) for t in (1 << 20, TAIL_ALWAYS) for g in (False, True) for w in (False, True)]


def zeros3(cam):
    return tuple(torch.zeros((cam.image_height, cam.image_width, 3), dtype=torch.float32, device="cuda")
                 for _ in range(3))


def test_serving_batch_of_64_scenes(ctx):
    items, arrs = [], []
    for s in range(64):
        arr, cam, spp = scenes.batch_scene(s)
        sc = ctx.upload(arr)
        sc.build_bvh()
        items.append((sc, cam, RenderSettings(spp=spp, max_depth=16, background=(0.6, 0.7, 0.9), aov=True)))
        arrs.append(arr)
    prims = [a.n_prims for a in arrs]
    res = [c.image_width for _, c, _ in items]
    spps = [rs.spp for _, _, rs in items]
    print(f"serving: prims {min(prims)}..{max(prims)}, res {min(res)}..{max(res)}, spp {min(spps)}..{max(spps)}",
          flush=True)
    assert min(prims) < 100 and max(prims) > 150_000
    assert min(res) == 256 and max(res) == 1024 and min(spps) == 1 and max(spps) >= 15
    try:
        # single renders in the library's default shape
        ctx.set_launch()
        ctx.set_render_tree(False)
        singles, seg = [], 0
        for sc, cam, rs in items:
            b = zeros3(cam)
            seg += sc.render_device(cam, rs, *b)["segments"]
            singles.append(b)
        # the oracle on 256 seeded pixels of every scene
        for k, ((sc, cam, rs), arr) in enumerate(zip(items, arrs)):
            o = orc.OracleScene(arr)
            o.build_bvh()
            W, H = cam.image_width, cam.image_height
            pix = np.random.default_rng(1000 + k).choice(W * H, 256, replace=False).astype(np.uint32)
            ro, ra, rn, _ = o.render(cam, rs, pixels=pix)
            rgb, alb, nrm = (t.reshape(-1, 3)[torch.from_numpy(pix.astype(np.int64)).cuda()].cpu().numpy()
                             for t in singles[k])
            assert np.array_equal(rgb.view(np.uint32), ro.view(np.uint32)), ("oracle rgb", k)
            assert np.array_equal(alb.view(np.uint32), ra.view(np.uint32)), ("oracle albedo", k)
            assert np.array_equal(nrm.view(np.uint32), rn.view(np.uint32)), ("oracle normal", k)
        print("serving: singles equal the oracle on 64 x 256 pixels", flush=True)
        # every launch shape of the batch
        bufs = [zeros3(cam) for _, cam, _ in items]
        for shape in SHAPES:
            autotune.apply(ctx, shape)
            for b in bufs:
                for t in b:
                    t.zero_()
            st = render_batch(ctx, [(sc, cam, rs, *b) for (sc, cam, rs), b in zip(items, bufs)],
                              graph=shape["graph"])
            print("serving shape", shape, "passes", st["passes"], "segments", st["segments"], flush=True)
            assert st["segments"] == seg
            assert st["paths"] == sum(c.image_width * c.image_height * rs.spp for _, c, rs in items)
            for k, (b, s) in enumerate(zip(bufs, singles)):
                for t, u, what in zip(b, s, ("rgb", "albedo", "normal")):
                    assert torch.equal(t, u), (shape, k, what)
    finally:
        ctx.set_launch()
        ctx.set_render_tree(False)
        for sc, _, _ in items:
            sc.close()
