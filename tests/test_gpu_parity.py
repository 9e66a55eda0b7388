"""GPU (libaurora, gfx950) against the CPU oracle, through the C ABI.

Bar: bit-exact for the LBVH (Morton codes, order, every node word), for
closest-hit primitive ids, hit distances and fetch counts, and for
per-pixel radiance sums (FP32, same Philox streams, same operation order)."""
import numpy as np
import pytest
import torch

from oracle import orc
from aurora_amd import scenes
from aurora_amd.core import Camera, RenderSettings, SceneArrays

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def both(ctx, arr):
    g = ctx.upload(arr)
    gs = g.build_bvh()
    o = orc.OracleScene(arr)
    os_ = o.build_bvh()
    return g, gs, o, os_


SCENES = {
    "cornell": lambda: scenes.cornell_box(),
    "spheres": lambda: scenes.random_spheres(0),
    "stress20k": lambda: scenes.stress_scene(10000, 10000, seed=3),
    "stress200k": lambda: scenes.stress_scene(100000, 100000, seed=4),
    # greedy collapse too deep for the traversal stack: the pairwise form
    "deep_chain": lambda: scenes.deep_chain_scene(),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_lbvh_bit_exact(ctx, name):
    say("lbvh", name)
    arr = SCENES[name]()
    g, gs, o, os_ = both(ctx, arr)
    for k in ("n_prims", "n_nodes", "n_wide", "depth", "wide_depth", "wide_form", "scene_lo", "scene_hi",
              "centroid_lo", "centroid_hi"):
        assert gs[k] == os_[k], k
    assert 3 * (gs["wide_depth"] - 1) <= 96
    gm, go = g.order()
    om, oo = o.order()
    assert np.array_equal(gm, om)
    assert np.array_equal(go, oo)
    assert np.array_equal(g.nodes(), o.nodes())
    assert np.array_equal(g.wide(), o.wide())


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 100, 2048])
def test_lbvh_small_and_duplicates(ctx, n):
    say("lbvh small", n)
    rng = np.random.default_rng(n)
    b = scenes.Builder()
    m = b.lambertian((0.5, 0.5, 0.5))
    for c in rng.uniform(-1, 1, (n, 3)):
        b.sphere(c, 0.1, m)
    for c in np.zeros((n, 3)):  # exact duplicates
        b.sphere(c, 0.2, m)
    arr = b.arrays() if n else SceneArrays()
    g, gs, o, os_ = both(ctx, arr)
    assert gs["n_nodes"] == os_["n_nodes"] and gs["n_wide"] == os_["n_wide"]
    if arr.n_prims:
        assert np.array_equal(g.nodes(), o.nodes())
        assert np.array_equal(g.wide(), o.wide())
        assert np.array_equal(g.order()[1], o.order()[1])


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("name,nrays", [("cornell", 1 << 18), ("spheres", 1 << 18), ("stress200k", 1 << 20),
                                       ("deep_chain", 1 << 16)])
def test_closest_hit_bit_exact(ctx, name, nrays):
    say("trace", name, nrays)
    arr = SCENES[name]()
    g, gs, o, os_ = both(ctx, arr)
    lo = np.array(gs["scene_lo"]) - 0.05
    hi = np.array(gs["scene_hi"]) + 0.05
    if name == "spheres":
        lo, hi = np.array([-12, 0.01, -12]), np.array([12, 3, 12])
    org, dirs = scenes.random_rays(nrays, lo, hi, seed=11)
    op, ot, oc = o.trace(org, dirs, tmin=1e-4, counts=True)
    dp = torch.zeros(nrays, dtype=torch.int32, device="cuda")
    dt = torch.zeros(nrays, dtype=torch.float32, device="cuda")
    dc = torch.zeros((nrays, 3), dtype=torch.int32, device="cuda")
    g.trace_device(dev(org), dev(dirs), dp, dt, tmin=1e-4, counts=dc)
    ctx.synchronize()
    gp, gt, gc = dp.cpu().numpy(), dt.cpu().numpy(), dc.cpu().numpy().astype(np.uint32)
    assert (op >= 0).mean() > 0.1
    assert np.array_equal(gp, op), f"{(gp != op).sum()} prim mismatches"
    assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(gc, oc)
    # traversal in sorted ray order (origin Morton code, octant) and in the
    # given order give the same hits and fetch counts
    ctx.set_ray_sort(2)
    try:
        dp.zero_(); dt.zero_(); dc.zero_()
        g.trace_device(dev(org), dev(dirs), dp, dt, tmin=1e-4, counts=dc)
        ctx.synchronize()
    finally:
        ctx.set_ray_sort(1)
    assert np.array_equal(dp.cpu().numpy(), op)
    assert np.array_equal(dt.cpu().numpy().view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(dc.cpu().numpy().astype(np.uint32), oc)
    # host-buffer entry point agrees with the device one
    hp, ht = g.trace(org[:4096], dirs[:4096], tmin=1e-4)
    assert np.array_equal(hp, op[:4096]) and np.array_equal(ht, ot[:4096])
    # brute force on the GPU gives the same closest hits
    if arr.n_prims <= 20000:
        g.trace_device(dev(org), dev(dirs), dp, dt, tmin=1e-4, brute=True)
        ctx.synchronize()
        assert np.array_equal(dp.cpu().numpy(), op)
        assert np.array_equal(dt.cpu().numpy(), ot)


@pytest.mark.parametrize("name", ["cornell", "spheres"])
def test_closest_hit_2_24_rays_against_brute_force(ctx, name):
    """2^24 seeded rays on a small scene: the four-wide traversal (sorted,
    as a query of this size runs) equals the GPU brute-force scan on every
    ray, bit for bit; every 1024th ray is checked against the oracle."""
    say("trace 2^24 vs brute", name)
    arr = SCENES[name]()
    g = ctx.upload(arr)
    gs = g.build_bvh()
    lo = np.array(gs["scene_lo"]) - 0.05 * (np.array(gs["scene_hi"]) - np.array(gs["scene_lo"]))
    hi = np.array(gs["scene_hi"]) + 0.05 * (np.array(gs["scene_hi"]) - np.array(gs["scene_lo"]))
    if name == "spheres":
        lo, hi = [-12, 0.01, -12], [12, 3, 12]
    n = 1 << 24
    od, dd = scenes.random_rays_torch(n, list(lo), list(hi), seed=31)
    bp = torch.empty(n, dtype=torch.int32, device="cuda")
    bt = torch.empty(n, dtype=torch.float32, device="cuda")
    fp = torch.empty_like(bp)
    ft = torch.empty_like(bt)
    ctx.set_ray_sort(2)
    try:
        g.trace_device(od, dd, bp, bt, tmin=1e-4)
    finally:
        ctx.set_ray_sort(1)
    g.trace_device(od, dd, fp, ft, tmin=1e-4, brute=True)
    ctx.synchronize()
    assert (bp >= 0).float().mean().item() > 0.2
    assert torch.equal(bp, fp)
    assert torch.equal(bt.view(torch.int32), ft.view(torch.int32))
    sub = torch.arange(0, n, 1024, device="cuda")
    o = orc.OracleScene(arr)
    o.build_bvh()
    op, ot = o.trace(od[sub].cpu().numpy(), dd[sub].cpu().numpy(), tmin=1e-4)
    assert np.array_equal(bp[sub].cpu().numpy(), op)
    assert np.array_equal(bt[sub].cpu().numpy().view(np.uint32), ot.view(np.uint32))


def test_full_size_stress_tree_and_hits(ctx):
    """The traversal stress scene at its full size (2M spheres + 2M
    triangles, 4M leaves, the greedy four-wide collapse): the GPU LBVH equals
    the oracle's word for word (order, binary nodes, four-wide nodes, stats),
    and the closest hits of 16384 incoherent rays over it (ids, t bits and
    per-ray fetch counts) equal the oracle's."""
    say("full-size stress tree")
    arr = scenes.stress_scene()
    g, gs, o, os_ = both(ctx, arr)
    for k in ("n_prims", "n_nodes", "n_wide", "depth", "wide_depth", "wide_form", "scene_lo", "scene_hi",
              "centroid_lo", "centroid_hi"):
        assert gs[k] == os_[k], k
    assert 3 * (gs["wide_depth"] - 1) <= 96
    assert gs["n_prims"] == 4_000_000
    assert np.array_equal(g.order()[1], o.order()[1])
    assert np.array_equal(g.nodes(), o.nodes())
    assert np.array_equal(g.wide(), o.wide())
    org, dirs = scenes.random_rays(16384, [-0.05] * 3, [1.05] * 3, seed=77)
    say("full-size stress hits")
    op, ot, oc = o.trace(org, dirs, tmin=1e-4, counts=True, threads=16)
    n = len(org)
    dp = torch.zeros(n, dtype=torch.int32, device="cuda")
    dt = torch.zeros(n, dtype=torch.float32, device="cuda")
    cnt = torch.zeros((n, 3), dtype=torch.int32, device="cuda")
    g.trace_device(dev(org), dev(dirs), dp, dt, tmin=1e-4, counts=cnt)
    ctx.synchronize()
    assert (op >= 0).mean() > 0.99
    assert np.array_equal(dp.cpu().numpy(), op)
    assert np.array_equal(dt.cpu().numpy().view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(cnt.cpu().numpy().astype(np.uint32), oc)


def test_closest_hit_on_silhouettes_independent_of_tree(ctx):
    """Rays grazing small distant spheres among crossing quads, where a
    primitive's computed hit can fall outside its box: GPU queries over the
    four-wide tree, GPU brute force and the oracle agree bit for bit (hits,
    t and per-ray fetch counts), and a render of the scene is the same with
    the binary and the four-wide tree and equals the oracle."""
    say("silhouette rays")
    arr, org, dirs = scenes.silhouette_scene(1)
    g, gs, o, os_ = both(ctx, arr)
    op, ot, oc = o.trace(org, dirs, tmin=1e-3, counts=True)
    n = len(org)
    dp = torch.zeros(n, dtype=torch.int32, device="cuda")
    dt = torch.zeros(n, dtype=torch.float32, device="cuda")
    cnt = torch.zeros((n, 3), dtype=torch.int32, device="cuda")
    g.trace_device(dev(org), dev(dirs), dp, dt, tmin=1e-3, counts=cnt)
    bp = torch.zeros_like(dp)
    bt = torch.zeros_like(dt)
    g.trace_device(dev(org), dev(dirs), bp, bt, tmin=1e-3, brute=True)
    ctx.synchronize()
    assert (op >= 0).mean() > 0.9
    assert np.array_equal(dp.cpu().numpy(), op)
    assert np.array_equal(dt.cpu().numpy().view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(cnt.cpu().numpy().astype(np.uint32), oc)
    assert np.array_equal(bp.cpu().numpy(), op)
    assert np.array_equal(bt.cpu().numpy().view(np.uint32), ot.view(np.uint32))
    cam = Camera(64, 48, 12.0, (0, 0.2, 4), (0, 0, 0), (0, 1, 0), 0.0, 1.0)
    rs = RenderSettings(spp=4, max_depth=8, background=(0.7, 0.8, 1.0))
    orgb, _, _, segs = o.render(cam, rs)
    for wide in (False, True):
        ctx.set_render_tree(wide)
        try:
            grgb, _, _, st = g.render(cam, rs)
        finally:
            ctx.set_render_tree(False)
        assert st["segments"] == segs
        assert np.array_equal(grgb.view(np.uint32), orgb.view(np.uint32)), wide


def test_closest_hit_axis_rays_and_rays_on_faces(ctx):
    """Directions along the axes (1/d infinite, 0*inf in the slabs) and
    origins on the planes of the Cornell walls and box faces: GPU queries,
    GPU brute force and the oracle agree bit for bit."""
    say("trace axis rays")
    arr = scenes.cornell_box()
    g, gs, o, os_ = both(ctx, arr)
    rng = np.random.default_rng(12)
    dirs, orgs = [], []
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    for k in range(6000):
        d = axes[k % 6]
        p = rng.uniform(-10, 565, 3).astype(np.float32)
        # snap one coordinate onto a wall or box-face plane
        p[rng.integers(3)] = np.float32(rng.choice([0.0, 555.0, 165.0, 330.0, 265.0, 295.0, 130.0, 65.0]))
        orgs.append(p)
        dirs.append(d)
    for k in range(2000):  # axis-parallel in one coordinate only
        d = rng.normal(size=3).astype(np.float32)
        d[k % 3] = 0.0
        d /= np.linalg.norm(d)
        orgs.append(rng.uniform(-10, 565, 3).astype(np.float32))
        dirs.append(d.astype(np.float32))
    org = np.array(orgs, np.float32)
    dirs = np.array(dirs, np.float32)
    op, ot = o.trace(org, dirs, tmin=1e-4)
    n = len(org)
    dp = torch.zeros(n, dtype=torch.int32, device="cuda")
    dt = torch.zeros(n, dtype=torch.float32, device="cuda")
    g.trace_device(dev(org), dev(dirs), dp, dt, tmin=1e-4)
    bp = torch.zeros_like(dp)
    bt = torch.zeros_like(dt)
    g.trace_device(dev(org), dev(dirs), bp, bt, tmin=1e-4, brute=True)
    ctx.synchronize()
    assert (op >= 0).mean() > 0.3
    assert np.array_equal(dp.cpu().numpy(), op)
    assert np.array_equal(dt.cpu().numpy().view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(bp.cpu().numpy(), op)
    assert np.array_equal(bt.cpu().numpy().view(np.uint32), ot.view(np.uint32))


RENDERS = [
    ("cornell", scenes.cornell_box, lambda: scenes.cornell_camera(48), RenderSettings(spp=12, max_depth=50)),
    ("spheres", lambda: scenes.random_spheres(0), lambda: scenes.random_spheres_camera(64, 36),
     RenderSettings(spp=8, max_depth=50, background=(0.7, 0.8, 1.0))),
    ("stress20k", SCENES["stress20k"], lambda: scenes.stress_camera(48, 27),
     RenderSettings(spp=4, max_depth=16, background=(0.7, 0.8, 1.0), tmin=1e-4)),
]


@pytest.mark.parametrize("name,mk,mkcam,rs", RENDERS, ids=[r[0] for r in RENDERS])
def test_render_radiance_bit_exact(ctx, name, mk, mkcam, rs):
    say("render", name)
    arr = mk()
    cam = mkcam()
    g, gs, o, os_ = both(ctx, arr)
    orgb, _, _, segs = o.render(cam, rs)
    # pure wavefront, wavefront then fused tail, fused from the first bounce;
    # binary, four-wide and flat render trees (flat: scenes of at most 256
    # primitives in the fused tail; the four-wide tree elsewhere)
    for tail, wide in ((0, 0), (2048, 0), (1 << 30, 0), (0, 1), (1 << 30, 1), (2048, 2), (1 << 30, 2)):
        ctx.set_launch(tail_paths=tail)
        ctx.set_render_tree(wide)
        try:
            grgb, _, _, st = g.render(cam, rs)
        finally:
            ctx.set_launch()
            ctx.set_render_tree(False)
        eq = (grgb.view(np.uint32) == orgb.view(np.uint32)).all(-1)
        say(name, "tail", tail, "wide", wide, "pixels bit-equal:", eq.mean(), "segments gpu/cpu", st["segments"], segs)
        assert st["paths"] == cam.image_width * cam.image_height * rs.spp
        assert st["segments"] == segs
        assert eq.all()


# Flat-walk scene sizes: one sphere, one triangle, spheres and triangles
# mixed up to the flat array's limit of 256 primitives, and one past it
# (257: no flat array, the four-wide tree stands in).  Primitives five times
# the stress scene's size (a lone triangle forty times), so that a 40x24
# image hits them; checker
# and Perlin textures.
FLAT_SIZES = [(1, 0), (0, 1), (200, 56), (0, 256), (256, 0), (129, 128)]


@pytest.mark.parametrize("ns,nt", FLAT_SIZES, ids=[f"s{a}_t{b}" for a, b in FLAT_SIZES])
def test_flat_walk_sizes_bit_exact(ctx, ns, nt):
    say("flat walk", ns, "spheres", nt, "triangles")
    arr = scenes.stress_scene(ns, nt, seed=11)
    arr.spheres[:, 3] *= 5.0
    arr.quads[:, 3:] *= 40.0 if nt == 1 else 5.0  # a lone triangle large enough to be seen
    cam = scenes.stress_camera(40, 24)
    rs = RenderSettings(spp=4, max_depth=16, background=(0.7, 0.8, 1.0), tmin=1e-4)
    g, gs, o, os_ = both(ctx, arr)
    orgb, _, _, segs = o.render(cam, rs)
    hits = (orgb != orgb[0, 0]).any(-1).mean()
    for tail, tree in ((1 << 30, 2), (2048, 2), (1 << 30, 1)):
        ctx.set_launch(tail_paths=tail)
        ctx.set_render_tree(tree)
        try:
            grgb, _, _, st = g.render(cam, rs)
        finally:
            ctx.set_launch()
            ctx.set_render_tree(0)
        say("flat walk", ns, nt, "tail", tail, "tree", tree, "segments gpu/cpu", st["segments"], segs,
            "pixels off the background", hits)
        assert st["segments"] == segs
        assert np.array_equal(grgb.view(np.uint32), orgb.view(np.uint32))
    assert segs > cam.image_width * cam.image_height * rs.spp  # some paths bounce


def test_render_deep_tree_bit_exact(ctx):
    """Paths through nested glass spheres over a tree 21 four-wide levels
    deep (the pairwise collapse: three pushes per level reach the stack's
    global spill past its 16 LDS levels): radiance and segment counts of the
    resumable four-wide walk, the binary walk and the wavefront kernels
    against the oracle."""
    say("render deep tree")
    arr = scenes.deep_chain_scene(glass=True)
    cam = Camera(32, 24, 40.0, (1.5, 1.2, 2.0), (0.0, 0.0, 0.0))
    rs = RenderSettings(spp=4, max_depth=16, background=(0.7, 0.8, 1.0), tmin=1e-4)
    g, gs, o, os_ = both(ctx, arr)
    assert gs["wide_depth"] > 16
    orgb, _, _, segs = o.render(cam, rs)
    assert len(np.unique(orgb.reshape(-1, 3), axis=0)) > 20
    for tail, tree in ((1 << 30, 1), (1 << 30, 0), (2048, 1), (0, 1)):
        ctx.set_launch(tail_paths=tail)
        ctx.set_render_tree(tree)
        try:
            grgb, _, _, st = g.render(cam, rs)
        finally:
            ctx.set_launch()
            ctx.set_render_tree(0)
        say("deep tree tail", tail, "tree", tree, "segments gpu/cpu", st["segments"], segs)
        assert st["segments"] == segs
        assert np.array_equal(grgb.view(np.uint32), orgb.view(np.uint32))


def test_render_textures_bit_exact(ctx):
    """Image textures on a sphere (spherical u, v from the series atan2 and
    acos), a parallelogram, a triangle and an emitter, with checker, Perlin,
    fuzzed metal and glass: radiance and AOVs bit-exact against the oracle
    in the wavefront-only and persistent-tail paths."""
    say("render textures")
    arr = scenes.textured_scene(0)
    cam = scenes.textured_camera(64, 48)
    g, gs, o, os_ = both(ctx, arr)
    rs = RenderSettings(spp=4, max_depth=20, background=(0.6, 0.7, 0.9), aov=True)
    orgb, oalb, onrm, segs = o.render(cam, rs)
    assert len(np.unique(oalb.reshape(-1, 3), axis=0)) > 500  # the textures vary across the image
    assert arr.n_prims <= 256  # small enough for the flat form
    for tail, tree in ((0, 0), ((1 << 32) - 1, 0), ((1 << 32) - 1, 1), ((1 << 32) - 1, 2)):
        ctx.set_launch(tail_paths=tail)
        ctx.set_render_tree(tree)
        try:
            grgb, galb, gnrm, st = g.render(cam, rs)
        finally:
            ctx.set_launch()
            ctx.set_render_tree(0)
        say("textures tail", tail, "tree", tree, "segments gpu/cpu", st["segments"], segs)
        assert st["segments"] == segs
        assert np.array_equal(grgb.view(np.uint32), orgb.view(np.uint32))
        assert np.array_equal(galb.view(np.uint32), oalb.view(np.uint32))
        assert np.array_equal(gnrm.view(np.uint32), onrm.view(np.uint32))


def test_render_aov_and_progressive(ctx):
    say("render aov/progressive")
    arr = scenes.cornell_box()
    cam = scenes.cornell_camera(32)
    g, gs, o, os_ = both(ctx, arr)
    rs = RenderSettings(spp=6, aov=True, spp_per_pass=4)
    orgb, oalb, onrm, _ = o.render(cam, rs)
    for tail, tree in ((0, 0), (1 << 30, 0), (1 << 30, 2)):
        ctx.set_launch(tail_paths=tail)
        ctx.set_render_tree(tree)
        try:
            grgb, galb, gnrm, _ = g.render(cam, rs)
        finally:
            ctx.set_launch()
            ctx.set_render_tree(0)
        assert np.array_equal(grgb, orgb)
        assert np.array_equal(galb, oalb)
        assert np.array_equal(gnrm, onrm)
    # two calls with sample_offset == one call (sums continue in sample order)
    a, _, _, _ = g.render(cam, RenderSettings(spp=2))
    a, _, _, _ = g.render(cam, RenderSettings(spp=4, sample_offset=2), rgb=a)
    b, _, _, _ = g.render(cam, RenderSettings(spp=6, spp_per_pass=1))
    assert np.array_equal(a, b)
    assert np.array_equal(a, orgb)


def test_render_defocus_and_odd_sizes(ctx):
    """A thin-lens camera (defocus disk sampled by Philox draws with
    rejection, as the oracle does), a 1x1 image, a 7x3 image, a single
    bounce and an explicit ray offset: radiance bit-exact against the oracle
    with both render trees."""
    say("render defocus and odd sizes")
    arr = scenes.random_spheres(0)
    g, gs, o, os_ = both(ctx, arr)
    cases = [
        (Camera(40, 24, 20.0, (13, 2, 3), (0, 0, 0), (0, 1, 0), 0.6, 10.0),
         RenderSettings(spp=4, max_depth=12, background=(0.7, 0.8, 1.0))),
        (Camera(1, 1, 20.0, (13, 2, 3), (0, 0, 0), (0, 1, 0), 0.0, 10.0),
         RenderSettings(spp=9, max_depth=50, background=(0.7, 0.8, 1.0))),
        (Camera(7, 3, 30.0, (13, 2, 3), (0, 0, 0), (0, 1, 0), 0.0, 10.0),
         RenderSettings(spp=3, max_depth=1, background=(0.7, 0.8, 1.0), tmin=1e-3)),
    ]
    for cam, rs in cases:
        orgb, _, _, segs = o.render(cam, rs)
        for wide in (False, True):
            ctx.set_render_tree(wide)
            try:
                grgb, _, _, st = g.render(cam, rs)
            finally:
                ctx.set_render_tree(False)
            assert st["segments"] == segs, (cam.image_width, wide)
            assert np.array_equal(grgb.view(np.uint32), orgb.view(np.uint32)), (cam.image_width, wide)


def test_camera_errors_are_reported(ctx):
    say("camera errors")
    g = ctx.upload(scenes.cornell_box())
    g.build_bvh()
    rs = RenderSettings(spp=1)
    for cam, msg in [(Camera(8, 8, 40.0, (1, 2, 3), (1, 2, 3), (0, 1, 0), 0.0, 1.0), "coincide"),
                     (Camera(8, 8, 40.0, (0, 0, 0), (0, 1, 0), (0, 1, 0), 0.0, 1.0), "parallel"),
                     (Camera(8, 8, 180.0, (0, 0, 0), (0, 0, 1), (0, 1, 0), 0.0, 1.0), "vfov"),
                     (Camera(8, 0, 40.0, (0, 0, 0), (0, 0, 1), (0, 1, 0), 0.0, 1.0), "positive")]:
        with pytest.raises(ValueError, match=msg):
            g.render(cam, rs)


def test_render_closed_form_light(ctx):
    say("render closed form")
    b = scenes.Builder()
    b.quad((-100, -100, 5), (200, 0, 0), (0, 200, 0), b.light((15.0, 15.0, 15.0)))
    g = ctx.upload(b.arrays())
    g.build_bvh()
    cam = Camera(16, 12, 60.0, (0, 0, 0), (0, 0, 1), (0, 1, 0), 0.0, 1.0)
    rgb, _, _, st = g.render(cam, RenderSettings(spp=7))
    assert np.all(rgb == np.float32(105.0))
    assert st["segments"] == 16 * 12 * 7


def test_errors_are_reported(ctx):
    say("errors")
    b = scenes.Builder()
    m = b.lambertian((0.5, 0.5, 0.5))
    b.quad((0, 0, 0), (1, 0, 0), (2, 0, 0), m)  # collinear edges
    with pytest.raises(ValueError, match="collinear"):
        ctx.upload(b.arrays())
    b = scenes.Builder()
    b.sphere((0, 0, 0), 1.0, 3)  # material out of range
    with pytest.raises(ValueError, match="material"):
        ctx.upload(b.arrays())
    g = ctx.upload(scenes.cornell_box())
    with pytest.raises(RuntimeError, match="BVH"):
        g.render(scenes.cornell_camera(8), RenderSettings(spp=1))
    g.build_bvh()
    with pytest.raises(ValueError, match="max_depth"):
        g.render(scenes.cornell_camera(8), RenderSettings(spp=1, max_depth=0))


def test_stress_full_scale_traversal(ctx):
    """Full traversal-stress scene (2M spheres + 2M triangles): the GPU LBVH
    equals the oracle's bit for bit; 2^24 seeded rays are traced on the GPU;
    every 256th ray (65536) is checked bit-exact against the oracle, the rest
    by size-independent properties."""
    say("stress full scale: scene")
    arr = scenes.stress_scene()
    g, gs, o, os_ = both(ctx, arr)
    say("stress full scale: lbvh compare")
    assert np.array_equal(g.order()[1], o.order()[1])
    assert np.array_equal(g.nodes(), o.nodes())
    assert gs["n_wide"] == os_["n_wide"]
    assert np.array_equal(g.wide(), o.wide())
    n = 1 << 24
    sub = np.arange(0, n, 256)
    say("stress full scale: rays")
    org, dirs = scenes.random_rays(len(sub), [-0.05] * 3, [1.05] * 3, seed=24)
    od, dd = scenes.random_rays_torch(n, [-0.05] * 3, [1.05] * 3, seed=24)
    od[sub] = dev(org)
    dd[sub] = dev(dirs)
    dp = torch.empty(n, dtype=torch.int32, device="cuda")
    dt = torch.empty(n, dtype=torch.float32, device="cuda")
    say("stress full scale: gpu trace 2^24")
    g.trace_device(od, dd, dp, dt, tmin=1e-4)
    ctx.synchronize()
    gp = dp.cpu().numpy()
    gt = dt.cpu().numpy()
    assert np.all((gp >= -1) & (gp < arr.n_prims))
    assert np.all((gp >= 0) == np.isfinite(gt)) and np.all(gt[gp >= 0] > 1e-4)
    say("stress full scale: oracle subset")
    op, ot = o.trace(org, dirs, tmin=1e-4)
    assert np.array_equal(gp[sub], op)
    assert np.array_equal(gt[sub].view(np.uint32), ot.view(np.uint32))
