"""The oracle's restatement of the reference's patch-projection renderer
(oracle/orc_patch.cpp, after experiments/rt10.cpp) pinned three ways:

- against rt10.cpp itself, compiled from its source (oracle/_ref/librt10.so,
  Camera::render wrapped by oracle/ref/rt10_driver.cpp): the double images
  are equal bit for bit on the reference's own scene, on two facing mirrors
  (recursion to depth 6, child textures above the minimum size), and on
  seeded scenes with repeated ids (the recursion guard compares ids),
  triangles without a material, degenerate triangles and texture sizes down
  to one texel;
- against the image file the unchanged reference program writes for its own
  scene (tests/golden/rt10_cornell_900x650.ppm.zlib), byte for byte through
  the reference's 8-bit rule (aurora_amd.patch.to_rgb8);
- by closed forms that hold without the reference: an empty scene shows the
  environment inside each viewport triangle and twice it on their shared
  diagonal; a diffuse square facing the eye shows its albedo; a mirror
  facing the eye shows the wall behind the eye tinted and mixed by its
  metalness; the recursion stops at max_depth.
"""
import ctypes as C
import json
import os
import subprocess
import tempfile
import zlib

import numpy as np
import pytest

from oracle import orc
from aurora_amd import patch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "librt10.so")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "rt10")
GOLDEN = os.path.join(ROOT, "tests", "golden", "rt10_cornell_900x650.ppm.zlib")
need_ref = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (no reference sources)")


def ref_render(t, m, c, g):
    L = C.CDLL(REF_LIB)
    img = np.zeros((int(c["height"].item()), int(c["width"].item()), 3))
    vp = C.c_void_p
    rc = L.ref_rt10_render(vp(t.ctypes.data if t.size else 0), C.c_uint32(t.size), vp(m.ctypes.data if m.size else 0),
                           C.c_uint32(m.size), vp(c.ctypes.data), vp(g.ctypes.data), vp(img.ctypes.data))
    assert rc == 0
    return img


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def seeded_cases():
    yield "rt10_cornell_300x200", patch.rt10_cornell(300, 200)
    for seed in range(8):
        yield f"random{seed}", patch.random_scene(30 + 5 * seed, seed, max_depth=2 + seed % 4)
    # mirrors facing each other: textures at every depth, above the minimum size
    yield "mirror_hall_deep", patch.mirror_hall(tiles=3, width=200, height=150, max_depth=6, min_tex_res=4)
    yield "mirror_hall_wide", patch.mirror_hall(tiles=4, width=96, height=72, max_depth=4, min_area_px=1.0,
                                                max_tex_res=40, min_tex_res=2)
    t, m, c, g = patch.random_scene(40, 100, reflective_share=0.9, max_depth=6)
    g["min_tex_res"] = 1
    g["max_tex_res"] = 24
    g["min_area_px"] = 0.0
    yield "deep_small_textures", (t, m, c, g)
    t, m, c, g = patch.random_scene(25, 101)
    t["p"][3] = t["p"][3][0]  # a point-like triangle
    t["p"][4][2] = t["p"][4][0] + 2.0 * (t["p"][4][1] - t["p"][4][0])  # collinear vertices
    t["uv"][5] = [[-0.5, 1.5], [2.0, -1.0], [0.3, 0.3]]  # uv outside [0,1]
    yield "degenerate", (t, m, c, g)


@need_ref
@pytest.mark.parametrize("name,case", list(seeded_cases()), ids=[n for n, _ in seeded_cases()])
def test_oracle_matches_compiled_reference(name, case):
    t, m, c, g = case
    ours, n = orc.patch_render(t, m, c, g)
    ref = ref_render(t, m, c, g)
    assert same_bits(ours, ref), np.abs(ours - ref).max()
    assert n > 0 or t.size == 0


@need_ref
def test_reference_program_output_is_the_golden_file():
    with tempfile.TemporaryDirectory() as d:
        subprocess.check_call([REF_EXE], cwd=d, stderr=subprocess.DEVNULL)
        data = open(os.path.join(d, "output_rt10.ppm"), "rb").read()
    assert data == zlib.decompress(open(GOLDEN, "rb").read())


def test_oracle_writes_the_reference_programs_file():
    t, m, c, g = patch.rt10_cornell()
    img, n = orc.patch_render(t, m, c, g)
    golden = zlib.decompress(open(GOLDEN, "rb").read())
    meta = json.load(open(GOLDEN.replace(".ppm.zlib", ".json")))
    assert len(golden) == meta["bytes"]
    assert patch.ppm_bytes(img) == golden
    # the reference's scene recurses: some textures are mirrored views
    assert n > 24


def test_to_rgb8_rounds_half_away_from_zero():
    # 255 c^(1/2.2) lands on k + 0.5 exactly only by accident; build the
    # values from the inverse so the rounding rule is what decides
    k = np.arange(0, 255, dtype=np.float64)
    lin_half_up = ((k + 0.5 + 1e-9) / 255.0) ** 2.2
    lin_half_dn = ((k + 0.5 - 1e-9) / 255.0) ** 2.2
    assert np.array_equal(patch.to_rgb8(lin_half_up), (k + 1).astype(np.uint8))
    assert np.array_equal(patch.to_rgb8(lin_half_dn), k.astype(np.uint8))
    assert patch.to_rgb8(np.array([-1.0, 2.0, np.float64(0.0)])).tolist() == [0, 255, 0]


def uv_grid(w, h):
    y, x = np.mgrid[0:h, 0:w]
    return x / (w - 1), y / (h - 1)


def test_empty_scene_shows_environment_and_the_diagonal_twice():
    c = patch.camera((0.0, 1.0, -3.0), -2.0, 1.0, 1.6, 41, 31)
    g = patch.config(env=(0.1, 0.2, 0.45))
    t, m = patch.SceneBuilder().arrays()
    img, n = orc.patch_render(t, m, c, g)
    assert n == 0
    u, v = uv_grid(41, 31)
    s = u + v
    on_diag = np.abs(s - 1.0) < 1e-9
    env = np.array([0.1, 0.2, 0.45])
    assert np.all(img[~on_diag] == env)
    assert np.allclose(img[on_diag], np.minimum(2 * env, 1.0), rtol=0, atol=1e-15)
    assert on_diag.sum() >= 2  # the corners (1,0) and (0,1) at least


def facing_quad(b, z, half, mat):
    # a square at depth z facing the eye at (0,1,-3)
    b.quad((-half, 1 - half, z), (half, 1 - half, z), (half, 1 + half, z), (-half, 1 + half, z), mat)


def interior(w, h):
    # pixels inside a facing square's projection and away from the seams:
    # the image diagonal u + v = 1 (the reference masks a viewport at
    # x / (w - 1) but rasterises at the pixel centre x + 0.5, so a seam about
    # a pixel wide runs along it) and the square's own diagonal (bilinear
    # samples there blend in texels outside its triangle)
    u, v = uv_grid(w, h)
    return (np.abs(u + v - 1) > 0.1) & (np.abs(u - v) > 0.1) & (u > 0.2) & (u < 0.8) & (v > 0.1) & (v < 0.9)


# A triangle is painted only when one of its vertices projects inside the
# viewport (rt10.cpp:462-484), so the closed-form scenes keep a vertex of
# each triangle in view: a facing square of half-size 3 at z = 1 projects to
# u in [0.145, 0.855], v in [0.031, 0.969] of a 33 x 25 image (the viewport
# is 1.6 high at distance 1 from the eye).
def test_diffuse_square_shows_its_albedo():
    b = patch.SceneBuilder()
    facing_quad(b, 1.0, 3.0, b.material(patch.DIFFUSE, (0.3, 0.6, 0.9)))
    t, m = b.arrays()
    c = patch.camera((0.0, 1.0, -3.0), -2.0, 1.0, 1.6, 33, 25)
    img, n = orc.patch_render(t, m, c, patch.config())
    sel = interior(33, 25)
    assert sel.sum() > 100 and n == 4  # two triangles, seen through both viewport triangles
    # a constant texture sampled bilinearly: exact up to the lerp's rounding
    assert np.allclose(img[sel], [0.3, 0.6, 0.9], rtol=0, atol=1e-15)
    u, v = uv_grid(33, 25)
    assert np.all(img[(u < 0.12) & (np.abs(u + v - 1) > 1e-9)] == [0.08, 0.08, 0.10])  # the environment


def test_mirror_shows_the_wall_behind_the_eye():
    b = patch.SceneBuilder()
    wall = (0.2, 0.7, 0.4)
    mirror_albedo = (0.9, 0.5, 1.0)
    metal = 0.75
    facing_quad(b, 1.0, 3.0, b.material(patch.REFLECTIVE, mirror_albedo, metal))
    # behind the eye: a fan of four triangles around (0, 1, -5), whose centre
    # vertex the mirrored eye (0, 1, 5) sees through the mirror's centre
    wm = b.material(patch.DIFFUSE, wall)
    ctr, R = (0.0, 1.0, -5.0), 60.0
    ring = [(-R, 1 - R, -5.0), (R, 1 - R, -5.0), (R, 1 + R, -5.0), (-R, 1 + R, -5.0)]
    for k in range(4):
        b.triangle(ctr, ring[k], ring[(k + 1) % 4], [(0.5, 0.5), (0, 0), (1, 1)], wm)
    t, m = b.arrays()
    c = patch.camera((0.0, 1.0, -3.0), -2.0, 1.0, 1.6, 33, 25)
    img, n = orc.patch_render(t, m, c, patch.config(min_area_px=0.0))
    sel = interior(33, 25)
    a, w = np.array(mirror_albedo), np.array(wall)
    expect = a * (1 - metal) + (w * a) * metal
    env = np.array([0.08, 0.08, 0.10])
    env_seen = a * (1 - metal) + (env * a) * metal
    got = img[sel]
    exact = np.all(np.abs(got - expect) <= 1e-12, axis=1)
    # the reference rasterises a texture at texel centres x + 0.5 but samples
    # it at u (w - 1): seams half a texel wide run along the reflected fan's
    # edges, where the sample blends in texels the wall does not cover
    # (environment); everything else is the tinted wall exactly
    assert exact.mean() > 0.8, exact.mean()
    lo, hi = np.minimum(expect, env_seen) - 1e-12, np.maximum(expect, env_seen) + 1e-12
    assert np.all((got >= lo) & (got <= hi))
    # with max_depth 0 the mirror shows its bare albedo and nothing recurses
    img0, n0 = orc.patch_render(t, m, c, patch.config(max_depth=0))
    assert np.allclose(img0[sel], a, rtol=0, atol=1e-15) and n0 == 4 < n


def test_bad_descriptors_are_rejected():
    t, m, c, g = patch.rt10_cornell(20, 20)
    bad = t.copy()
    bad["material"][0] = 99
    with pytest.raises(ValueError):
        orc.patch_render(bad, m, c, g)
    c1 = c.copy()
    c1["width"] = 1
    with pytest.raises(ValueError):
        orc.patch_render(t, m, c1, g)
