"""The GPU patch-projection renderer (aurora_patch_render, after
experiments/rt10.cpp) against the oracle's restatement, bit for bit in
double, through the C ABI; and against the image file the reference
program itself writes for its own scene (tests/golden), byte for byte
through the reference's 8-bit rule."""
import json
import os
import zlib

import numpy as np
import pytest

from oracle import orc
from aurora_amd import patch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "rt10_cornell_900x650.ppm.zlib")


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def cases():
    yield "rt10_cornell_300x200", patch.rt10_cornell(300, 200)
    for seed in range(6):
        yield f"random{seed}", patch.random_scene(30 + 5 * seed, seed, max_depth=2 + seed % 4)
    yield "mirror_hall_deep", patch.mirror_hall(tiles=3, width=200, height=150, max_depth=6, min_tex_res=4)
    yield "mirror_hall_wide", patch.mirror_hall(tiles=4, width=96, height=72, max_depth=4, min_area_px=1.0,
                                                max_tex_res=40, min_tex_res=2)
    yield "mirror_hall_640", patch.mirror_hall(tiles=4, width=640, height=480, max_depth=6)
    t, m, c, g = patch.random_scene(40, 100, reflective_share=0.9, max_depth=6)
    g["min_tex_res"] = 1
    g["max_tex_res"] = 24
    g["min_area_px"] = 0.0
    yield "deep_small_textures", (t, m, c, g)
    t, m, c, g = patch.random_scene(25, 101)
    t["p"][3] = t["p"][3][0]
    t["p"][4][2] = t["p"][4][0] + 2.0 * (t["p"][4][1] - t["p"][4][0])
    t["uv"][5] = [[-0.5, 1.5], [2.0, -1.0], [0.3, 0.3]]
    yield "degenerate", (t, m, c, g)
    yield "tiny_2x2", patch.rt10_cornell(2, 2)
    yield "odd_17x13", patch.rt10_cornell(17, 13)
    t, m = patch.SceneBuilder().arrays()
    yield "empty", (t, m, patch.camera((0.0, 1.0, -3.0), -2.0, 1.0, 1.6, 33, 21), patch.config())


@pytest.mark.parametrize("name,case", list(cases()), ids=[n for n, _ in cases()])
def test_patch_render_matches_oracle(ctx, name, case):
    t, m, c, g = case
    print(f"patch case {name}: {t.size} triangles, {int(c['width'].item())}x{int(c['height'].item())}", flush=True)
    img, st = ctx.patch_render(t, m, c, g)
    ref, n_tex = orc.patch_render(t, m, c, g)
    assert st["textures"] == n_tex
    assert same_bits(img, ref), f"max |diff| {np.abs(img - ref).max()}, {np.sum(img != ref)} values differ"


def test_patch_render_writes_the_reference_programs_file(ctx):
    t, m, c, g = patch.rt10_cornell()
    img, st = ctx.patch_render(t, m, c, g)
    golden = zlib.decompress(open(GOLDEN, "rb").read())
    meta = json.load(open(GOLDEN.replace(".ppm.zlib", ".json")))
    assert len(golden) == meta["bytes"]
    assert patch.ppm_bytes(img) == golden
    assert st["raster_textures"] > 0 and st["levels"] >= 1 and st["kernel_ms"] > 0


def test_patch_render_device_equals_host(ctx):
    import torch

    t, m, c, g = patch.mirror_hall(tiles=3, width=160, height=120, max_depth=5)
    img, _ = ctx.patch_render(t, m, c, g)
    out = torch.full((120, 160, 3), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = ctx.patch_render_device(t, m, c, g, out)
    assert st["texels"] >= 160 * 120
    assert same_bits(out.cpu().numpy(), img)
    # a second, smaller render reuses the context's tables
    t2, m2, c2, g2 = patch.rt10_cornell(40, 30)
    img2, _ = ctx.patch_render(t2, m2, c2, g2)
    assert same_bits(img2, orc.patch_render(t2, m2, c2, g2)[0])


def test_patch_render_rejects_bad_descriptors(ctx):
    t, m, c, g = patch.rt10_cornell(20, 20)
    bad = t.copy()
    bad["material"][0] = 99
    with pytest.raises(ValueError, match="material index"):
        ctx.patch_render(bad, m, c, g)
    c1 = c.copy()
    c1["width"] = 1
    with pytest.raises(ValueError, match="image sides"):
        ctx.patch_render(t, m, c1, g)
    g1 = g.copy()
    g1["max_depth"] = patch.MAX_DEPTH + 1
    with pytest.raises(ValueError, match="max_depth"):
        ctx.patch_render(t, m, c, g1)
    g2 = g.copy()
    g2["min_tex_res"] = 0
    with pytest.raises(ValueError, match="min_tex_res"):
        ctx.patch_render(t, m, c, g2)
    m2 = m.copy()
    m2["type"][0] = 7
    with pytest.raises(ValueError, match="material type"):
        ctx.patch_render(t, m2, c, g)
    t3 = t.copy()
    t3["p"][0][0][0] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        ctx.patch_render(t3, m, c, g)
    # the context is still usable
    img, _ = ctx.patch_render(t, m, c, g)
    assert same_bits(img, orc.patch_render(t, m, c, g)[0])
