"""Workload scenes, flattened to SceneArrays.

Every random draw comes from Philox4x32-10 keyed (seed, STREAM_SCENE), in a
documented order, so a scene is a pure function of its seed.

  cornell_box()      Cornell box of "Ray Tracing: The Next Week": 555-unit
                     room, red/green/white lambertian walls, a 130x105 quad
                     light emitting 15, two white boxes rotated 15 and -18
                     degrees.  Material constants match the reference's
                     experiments/rt9.cpp:1446-1451; geometry the 555-unit
                     room of experiments/rt12.cpp:508-601.
  random_spheres()   "Ray Tracing in One Weekend" final scene: r=1000 ground,
                     22x22 jittered grid of r=0.2 spheres (80% lambertian,
                     15% metal, 5% glass), three r=1 spheres.
  stress_scene()     traversal stress: N spheres with log-uniform radii and N
                     random triangles in the unit cube, checker and Perlin
                     textures.
  random_rays()      origins uniform in a box, directions uniform on the sphere.
"""
import math

import numpy as np

from . import abi
from .core import Camera, SceneArrays, checker, image, noise, solid

STREAM_RENDER = 0
STREAM_SCENE = 1

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 on uint32 arrays; returns 4 uint32 arrays."""
    x = [np.asarray(c, dtype=np.uint64) & _MASK for c in (c0, c1, c2, c3)]
    x = list(np.broadcast_arrays(*x))
    x = [np.array(v, dtype=np.uint64) for v in x]
    ka, kb = int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * x[0]
        p1 = _M1 * x[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        x = [hi1 ^ x[1] ^ np.uint64(ka), lo1, hi0 ^ x[3] ^ np.uint64(kb), lo0]
        ka = (ka + _W0) & 0xFFFFFFFF
        kb = (kb + _W1) & 0xFFFFFFFF
    return [v.astype(np.uint32) for v in x]


def u01(x):
    """24-bit uniform in [0,1) as float32, exactly as the kernels compute it."""
    return (np.asarray(x, np.uint32) >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)


class Draws:
    """Sequential uniforms: draw k is word k%4 of Philox(ctr=(k//4, 0, 0, 0),
    key=(seed, STREAM_SCENE)), as float64 from the 24-bit float32 value."""

    def __init__(self, seed):
        self.seed = seed
        self.k = 0
        self.buf = []

    def __call__(self, lo=0.0, hi=1.0):
        if not self.buf:
            b = self.k // 4
            r = philox4x32_10(b, 0, 0, 0, self.seed, STREAM_SCENE)
            self.buf = [float(u01(v)) for v in r]
        self.k += 1
        u = self.buf.pop(0)
        return lo + (hi - lo) * u

    def array(self, n, lo=0.0, hi=1.0):
        """n draws at once (continues the same sequence)."""
        if n == 0:
            return np.zeros(0)
        assert not self.buf and self.k % 4 == 0, "array() needs a block boundary"
        nb = (n + 3) // 4
        b0 = self.k // 4
        r = philox4x32_10(np.arange(b0, b0 + nb, dtype=np.uint64), 0, 0, 0, self.seed, STREAM_SCENE)
        u = np.stack([u01(v) for v in r], axis=1).reshape(-1)[:n].astype(np.float64)
        self.k += 4 * nb
        return lo + (hi - lo) * u


class Builder:
    """Accumulates primitives; quads carry (Q, u, v) rows."""

    def __init__(self):
        self.spheres, self.sphere_mat = [], []
        self.quads, self.quad_mat, self.quad_kind = [], [], []
        self.materials, self.textures = [], []

    def texture(self, t):
        self.textures.append(t)
        return len(self.textures) - 1

    def material(self, mtype, tex=0, fuzz=0.0, ior=1.0):
        self.materials.append((mtype, tex, fuzz, ior))
        return len(self.materials) - 1

    def lambertian(self, rgb):
        return self.material(abi.MAT_LAMBERTIAN, self.texture(solid(*rgb)))

    def metal(self, rgb, fuzz):
        return self.material(abi.MAT_METAL, self.texture(solid(*rgb)), fuzz=fuzz)

    def dielectric(self, ior):
        return self.material(abi.MAT_DIELECTRIC, 0, ior=ior)

    def light(self, rgb):
        return self.material(abi.MAT_DIFFUSE_LIGHT, self.texture(solid(*rgb)))

    def sphere(self, c, r, m):
        self.spheres.append((c[0], c[1], c[2], r))
        self.sphere_mat.append(m)

    def quad(self, Q, u, v, m, kind=abi.QUAD_PARALLELOGRAM):
        self.quads.append(tuple(Q) + tuple(u) + tuple(v))
        self.quad_mat.append(m)
        self.quad_kind.append(kind)

    def box(self, a, b, m, rot_y_deg=0.0, offset=(0.0, 0.0, 0.0)):
        """Six quads of the box [a, b], rotated about +y then translated
        (rotate_y / translate baked in double precision)."""
        lo = np.minimum(a, b).astype(np.float64)
        hi = np.maximum(a, b).astype(np.float64)
        dx = np.array([hi[0] - lo[0], 0, 0])
        dy = np.array([0, hi[1] - lo[1], 0])
        dz = np.array([0, 0, hi[2] - lo[2]])
        sides = [
            ((lo[0], lo[1], hi[2]), dx, dy),
            ((hi[0], lo[1], hi[2]), -dz, dy),
            ((hi[0], lo[1], lo[2]), -dx, dy),
            ((lo[0], lo[1], lo[2]), dz, dy),
            ((lo[0], hi[1], hi[2]), dx, -dz),
            ((lo[0], lo[1], lo[2]), dx, dz),
        ]
        th = math.radians(rot_y_deg)
        c, s = math.cos(th), math.sin(th)

        def rot(p):
            return np.array([c * p[0] + s * p[2], p[1], -s * p[0] + c * p[2]])

        off = np.asarray(offset, np.float64)
        for Q, u, v in sides:
            self.quad(rot(np.asarray(Q)) + off, rot(u), rot(v), m)

    def arrays(self):
        return SceneArrays(
            spheres=np.asarray(self.spheres, np.float32).reshape(-1, 4),
            sphere_mat=np.asarray(self.sphere_mat, np.uint32),
            quads=np.asarray(self.quads, np.float32).reshape(-1, 9),
            quad_mat=np.asarray(self.quad_mat, np.uint32),
            quad_kind=np.asarray(self.quad_kind, np.uint32),
            materials=list(self.materials),
            textures=list(self.textures),
        )


def cornell_box(box_rot=(15.0, -18.0), box_size=((165, 330, 165), (165, 165, 165)),
                albedo_white=(0.73, 0.73, 0.73), albedo_red=(0.65, 0.05, 0.05), albedo_green=(0.12, 0.45, 0.15),
                light_size=(130.0, 105.0), emit=15.0):
    b = Builder()
    red = b.lambertian(albedo_red)
    white = b.lambertian(albedo_white)
    green = b.lambertian(albedo_green)
    light = b.light((emit, emit, emit))
    lw, ld = light_size
    b.quad((555, 0, 0), (0, 555, 0), (0, 0, 555), green)
    b.quad((0, 0, 0), (0, 555, 0), (0, 0, 555), red)
    b.quad((278 + lw / 2, 554, 279.5 + ld / 2), (-lw, 0, 0), (0, 0, -ld), light)
    b.quad((0, 0, 0), (555, 0, 0), (0, 0, 555), white)
    b.quad((555, 555, 555), (-555, 0, 0), (0, 0, -555), white)
    b.quad((0, 0, 555), (555, 0, 0), (0, 555, 0), white)
    b.box((0, 0, 0), box_size[0], white, box_rot[0], (265, 0, 295))
    b.box((0, 0, 0), box_size[1], white, box_rot[1], (130, 0, 65))
    return b.arrays()


def cornell_camera(width=600, height=None):
    return Camera(width, height or width, 40.0, (278.0, 278.0, -800.0), (278.0, 278.0, 0.0), (0.0, 1.0, 0.0),
                  0.0, 10.0)


def random_spheres(seed=0):
    rnd = Draws(seed)
    b = Builder()
    b.sphere((0, -1000, 0), 1000, b.lambertian((0.5, 0.5, 0.5)))
    for a in range(-11, 11):
        for bb in range(-11, 11):
            choose = rnd()
            cx, cz = a + 0.9 * rnd(), bb + 0.9 * rnd()
            if math.sqrt((cx - 4) ** 2 + 0 + cz ** 2) <= 0.9:
                continue
            if choose < 0.8:
                c1 = (rnd(), rnd(), rnd())
                c2 = (rnd(), rnd(), rnd())
                m = b.lambertian(tuple(x * y for x, y in zip(c1, c2)))
            elif choose < 0.95:
                alb = (rnd(0.5, 1), rnd(0.5, 1), rnd(0.5, 1))
                m = b.metal(alb, rnd(0, 0.5))
            else:
                m = b.dielectric(1.5)
            b.sphere((cx, 0.2, cz), 0.2, m)
    b.sphere((0, 1, 0), 1.0, b.dielectric(1.5))
    b.sphere((-4, 1, 0), 1.0, b.lambertian((0.4, 0.2, 0.1)))
    b.sphere((4, 1, 0), 1.0, b.metal((0.7, 0.6, 0.5), 0.0))
    return b.arrays()


def random_spheres_camera(width=1200, height=675):
    return Camera(width, height, 20.0, (13.0, 2.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.6, 10.0)


def stress_scene(n_spheres=2_000_000, n_tris=2_000_000, seed=0):
    """Traversal stress test.  Draw order: sphere centres (3n), log10 radii
    (n), triangle origins (3m), u directions (2m), u lengths (m), v
    directions (2m), v lengths (m), material choice (n+m).  Edge directions
    are uniform on the unit sphere: z = 2a - 1 and azimuth 2*pi*b from two
    draws (a, b) (Archimedes: z of a uniform direction is uniform)."""
    rnd = Draws(seed)
    cen = rnd.array(3 * n_spheres).reshape(-1, 3)
    rad = 10.0 ** rnd.array(n_spheres, -3.0, -1.0)
    Q = rnd.array(3 * n_tris).reshape(-1, 3)

    def dirs(m):
        ab = rnd.array(2 * m).reshape(-1, 2)
        z = 2.0 * ab[:, 0] - 1.0
        phi = 2.0 * np.pi * ab[:, 1]
        s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)

    ud = dirs(n_tris)
    ul = rnd.array(n_tris, 0.001, 0.05)
    vd = dirs(n_tris)
    vl = rnd.array(n_tris, 0.001, 0.05)
    u = ud * ul[:, None]
    v = vd * vl[:, None]
    # reject near-collinear edges: replace v by u x (previous v) direction
    cr = np.cross(u, v)
    bad = np.linalg.norm(cr, axis=1) < 1e-7
    if bad.any():
        alt = np.cross(u[bad], np.array([0.0, 0.0, 1.0]))
        alt[np.linalg.norm(alt, axis=1) < 1e-9] = np.array([0.0, 1.0, 0.0]) * 1e-2
        v[bad] = alt / np.linalg.norm(alt, axis=1, keepdims=True) * vl[bad, None]
    choice = rnd.array(n_spheres + n_tris)
    b = Builder()
    t_even = b.texture(solid(0.2, 0.3, 0.1))
    t_odd = b.texture(solid(0.9, 0.9, 0.9))
    t_chk = b.texture(checker(0.02, t_even, t_odd))
    t_noise = b.texture(noise(40.0, seed=0))
    m_chk = b.material(abi.MAT_LAMBERTIAN, t_chk)
    m_noise = b.material(abi.MAT_LAMBERTIAN, t_noise)
    mats = np.where(choice < 0.5, m_chk, m_noise).astype(np.uint32)
    arr = b.arrays()
    arr.spheres = np.concatenate([cen, rad[:, None]], axis=1).astype(np.float32)
    arr.sphere_mat = mats[:n_spheres]
    arr.quads = np.concatenate([Q, u, v], axis=1).astype(np.float32)
    arr.quad_mat = mats[n_spheres:]
    arr.quad_kind = np.full(n_tris, abi.QUAD_TRIANGLE, np.uint32)
    return arr


def textured_scene(seed=0):
    """Every texture and material kind in one small scene: image textures
    (texels drawn from Philox) on a sphere (spherical u, v), a parallelogram
    and a triangle (barycentric u, v) and on an emitter; checker and Perlin
    spheres; fuzzed metal; glass; a lambertian ground under a sky."""
    rnd = Draws(seed)
    b = Builder()
    t1 = rnd.array(3 * 64 * 32).reshape(-1, 3)          # 64 x 32 image
    t2 = rnd.array(3 * 16 * 16).reshape(-1, 3) * 4.0    # 16 x 16 emitter image
    img1 = b.texture(image(64, 32, 0))
    img2 = b.texture(image(16, 16, 64 * 32))
    chk = b.texture(checker(0.3, b.texture(solid(0.9, 0.2, 0.1)), b.texture(solid(0.1, 0.1, 0.8))))
    per = b.texture(noise(3.0, seed=seed))
    m_img = b.material(abi.MAT_LAMBERTIAN, img1)
    b.quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), m_img)
    b.sphere((0, 1, 0), 1.0, m_img)
    b.sphere((-2.2, 0.8, 0.5), 0.8, b.material(abi.MAT_LAMBERTIAN, chk))
    b.sphere((2.2, 0.8, 0.5), 0.8, b.material(abi.MAT_LAMBERTIAN, per))
    b.sphere((1.0, 0.4, 2.0), 0.4, b.metal((0.8, 0.7, 0.6), 0.3))
    b.sphere((-1.0, 0.4, 2.0), 0.4, b.dielectric(1.5))
    b.quad((-3, 0.2, -2), (2.0, 2.5, 0.0), (1.5, 0.0, -0.5), m_img, kind=abi.QUAD_TRIANGLE)
    b.quad((-1.5, 3.5, -1.0), (3.0, 0.0, 0.0), (0.0, 0.0, 2.0), b.material(abi.MAT_DIFFUSE_LIGHT, img2))
    arr = b.arrays()
    arr.texels = np.concatenate([t1, t2]).astype(np.float32)
    return arr


def textured_camera(width=64, height=48):
    return Camera(width, height, 40.0, (0.0, 2.5, 7.0), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), 0.0, 7.0)


def stress_camera(width=1920, height=1080):
    return Camera(width, height, 45.0, (0.5, 0.5, -1.2), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.0, 1.7)


def cornell_triangles():
    """The scene of host/examples/ingest_triangles.cpp as the ingestion of a
    reference-style object set produces it: a 555-unit Cornell room, its
    light and two axis-aligned boxes (one a mirror: metal of albedo 0.9, fuzz
    0), each quad (Q, u, v) given as the triangles (Q, Q+u, Q+v) and
    (Q+u+v, Q+v, Q+u) and ingested from those vertices: Q' = v0,
    u' = v1 - v0, v' = v2 - v0 in double, then rounded to float."""
    quads = [((555, 0, 0), (0, 555, 0), (0, 0, 555), "green"), ((0, 0, 0), (0, 555, 0), (0, 0, 555), "red"),
             ((343, 554, 332), (-130, 0, 0), (0, 0, -105), "light"), ((0, 0, 0), (555, 0, 0), (0, 0, 555), "white"),
             ((555, 555, 555), (-555, 0, 0), (0, 0, -555), "white"), ((0, 0, 555), (555, 0, 0), (0, 555, 0), "white")]

    def box(a, b, m):
        dx, dy, dz = b[0] - a[0], b[1] - a[1], b[2] - a[2]
        return [((a[0], a[1], b[2]), (dx, 0, 0), (0, dy, 0), m), ((b[0], a[1], b[2]), (0, 0, -dz), (0, dy, 0), m),
                ((b[0], a[1], a[2]), (-dx, 0, 0), (0, dy, 0), m), ((a[0], a[1], a[2]), (0, 0, dz), (0, dy, 0), m),
                ((a[0], b[1], b[2]), (dx, 0, 0), (0, 0, -dz), m), ((a[0], a[1], a[2]), (dx, 0, 0), (0, 0, dz), m)]

    quads += box((130, 0, 65), (295, 165, 230), "mirror") + box((265, 0, 295), (430, 330, 460), "diffuse")
    b = Builder()
    mats = {"red": b.lambertian((0.65, 0.05, 0.05)), "green": b.lambertian((0.12, 0.45, 0.15)),
            "white": b.lambertian((0.73, 0.73, 0.73)), "light": b.light((15, 15, 15)),
            "mirror": b.metal((0.9, 0.9, 0.9), 0.0), "diffuse": b.lambertian((0.73, 0.73, 0.73))}
    for Q, u, v, m in quads:
        Q, u, v = (np.array(x, np.float64) for x in (Q, u, v))
        for v0, v1, v2 in ((Q, Q + u, Q + v), (Q + u + v, Q + v, Q + u)):
            b.quad(tuple(v0), tuple(v1 - v0), tuple(v2 - v0), mats[m], kind=abi.QUAD_TRIANGLE)
    return b.arrays()


def deep_chain_scene(levels=30, cluster=4096, glass=False):
    """A degenerate LBVH: a spine of `levels` binary levels, one per Morton
    bit (from the top), each peeling off a group of three coincident spheres
    whose box is larger than everything below, then `cluster` coincident tiny
    spheres.  Centroids span exactly [0, 1]^3 (a tiny sphere at each corner),
    so group k sits in the cell whose Morton code has only bit 29-k set.  The
    greedy four-wide collapse opens each group before the spine, so it
    advances one binary level per wide level and its walk's worst-case
    stack need (three pushes per level) exceeds 96 entries.  glass: the
    groups alternate glass (ior 1.5) and grey diffuse and the cluster emits,
    so that rendered paths enter the nested spheres and walk deep into the
    tree (same geometry, same tree)."""
    b = Builder()
    m = b.lambertian((0.5, 0.5, 0.5))
    g = b.dielectric(1.5) if glass else m
    e = b.light((4.0, 3.0, 2.0)) if glass else m
    b.sphere((1.0, 1.0, 1.0), 1e-4, e)
    for k in range(levels):
        axis, bit = k % 3, 9 - k // 3
        c = [0.5 / 1024.0] * 3
        c[axis] = (2.0 ** bit + 0.5) / 1024.0
        r = 2.0 * 2.0 ** (-k / 3.0)
        for _ in range(3):
            b.sphere(tuple(c), r, g if k % 2 == 0 else m)
    for _ in range(cluster):
        b.sphere((0.0, 0.0, 0.0), 1e-4, e)
    return b.arrays()


def random_rays(n, lo, hi, seed=0):
    """n rays, origins uniform in [lo, hi]^3 and directions uniform on the
    unit sphere (Philox(i, 1, 0, 0) / (i, 2, 0, 0) with key (seed,
    STREAM_SCENE); normal deviates via Box-Muller in float64)."""
    i = np.arange(n, dtype=np.uint64)
    a = philox4x32_10(i, 1, 0, 0, seed, STREAM_SCENE)
    c = philox4x32_10(i, 2, 0, 0, seed, STREAM_SCENE)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    o = lo + (hi - lo) * np.stack([u01(a[0]), u01(a[1]), u01(a[2])], axis=1).astype(np.float64)
    u1 = 1.0 - u01(c[0]).astype(np.float64)
    u2 = u01(c[1]).astype(np.float64)
    u3 = 1.0 - u01(c[2]).astype(np.float64)
    u4 = u01(c[3]).astype(np.float64)
    r1 = np.sqrt(-2 * np.log(u1))
    r2 = np.sqrt(-2 * np.log(u3))
    g = np.stack([r1 * np.cos(2 * np.pi * u2), r1 * np.sin(2 * np.pi * u2), r2 * np.cos(2 * np.pi * u4)], axis=1)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return o.astype(np.float32), g.astype(np.float32)


def silhouette_scene(seed=0, n_spheres=3000, n_quads=3000, n_rays=40000):
    """Small spheres (r in [0.004, 0.02]) and small quads crossing each other
    in [-0.3, 0.3]^3, seen from 4 units away, and rays aimed at the sphere
    silhouettes (within 0.5% inside to 0.05% outside the rim).  This is where
    the textbook ray/sphere quadratic loses its discriminant and where the
    closest hit would depend on the visit order without the leaf-box rule.
    Returns (arrays, origins, directions) as float32."""
    rng = np.random.default_rng(seed)
    b = Builder()
    m = b.lambertian((0.5, 0.5, 0.5))
    c = rng.uniform(-0.3, 0.3, (n_spheres, 3))
    r = rng.uniform(0.004, 0.02, n_spheres)
    for k in range(n_spheres):
        b.sphere(tuple(c[k]), float(r[k]), m)
    Q = rng.uniform(-0.3, 0.3, (n_quads, 3))
    u = rng.uniform(-0.03, 0.03, (n_quads, 3))
    v = rng.uniform(-0.03, 0.03, (n_quads, 3))
    for k in range(n_quads):
        b.quad(tuple(Q[k]), tuple(u[k]), tuple(v[k]), m)
    org = np.array([0.0, 0.2, 4.0])
    ks = rng.integers(0, n_spheres, n_rays)
    w = rng.normal(size=(n_rays, 3))
    to = c[ks] - org
    to /= np.linalg.norm(to, axis=1, keepdims=True)
    perp = w - (w * to).sum(1, keepdims=True) * to
    perp /= np.linalg.norm(perp, axis=1, keepdims=True)
    tgt = c[ks] + perp * (r[ks] * rng.uniform(0.995, 1.0005, n_rays))[:, None]
    d = (tgt - org).astype(np.float32)
    o = np.broadcast_to(org.astype(np.float32), d.shape).copy()
    return b.arrays(), o, d


def random_rays_torch(n, lo, hi, seed=0, device="cuda"):
    """Device-side ray set for timing runs: origins uniform in [lo, hi]^3,
    directions normalised Gaussian triples (torch generator seeded `seed`)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    lo = torch.tensor(lo, dtype=torch.float32, device=device)
    hi = torch.tensor(hi, dtype=torch.float32, device=device)
    o = lo + (hi - lo) * torch.rand((n, 3), generator=g, device=device)
    d = torch.randn((n, 3), generator=g, device=device)
    d = d / d.norm(dim=1, keepdim=True).clamp_min(1e-12)
    return o.contiguous(), d.contiguous()


def cornell_variant(seed):
    """Denoiser training scene: a Cornell box with box sizes U(100, 200)^3,
    rotations U(-45, 45) degrees, wall albedos U(0.05, 0.8)^3 and a light of
    side lengths with area in U(50^2, 200^2).  Draw order: box 1 size (3),
    box 2 size (3), rotations (2), red, white, green albedos (9), light area
    (1), light aspect (1)."""
    rnd = Draws(seed)
    s1 = tuple(rnd(100, 200) for _ in range(3))
    s2 = tuple(rnd(100, 200) for _ in range(3))
    r1, r2 = rnd(-45, 45), rnd(-45, 45)
    red = tuple(rnd(0.05, 0.8) for _ in range(3))
    white = tuple(rnd(0.05, 0.8) for _ in range(3))
    green = tuple(rnd(0.05, 0.8) for _ in range(3))
    area = rnd(50.0 ** 2, 200.0 ** 2)
    aspect = rnd(0.6, 1.6)
    lw = min(500.0, math.sqrt(area * aspect))
    ld = min(500.0, area / lw)
    return cornell_box(box_rot=(r1, r2), box_size=(s1, s2), albedo_white=white, albedo_red=red,
                       albedo_green=green, light_size=(lw, ld))


def batch_scene(seed, min_prims=10, max_prims=200_000):
    """Batched-serving scene: 10 * (2e4)^U primitives (spheres and quads in
    a random ratio) scattered in [-1, 1]^3 with lambertian / metal / glass
    materials and checker / noise textures, under a sky; resolution from
    {256, 384, 512, 768, 1024}^2 and 1..16 spp.  Returns (arrays, camera,
    spp)."""
    rnd = Draws(seed)
    n = int(min(max_prims, max(min_prims, round(10 * (2e4) ** rnd()))))
    frac_sph = rnd()
    res = [256, 384, 512, 768, 1024][min(4, int(rnd() * 5))]
    spp = 1 + min(15, int(rnd() * 16))
    b = Builder()
    t0 = b.texture(solid(0.8, 0.3, 0.2))
    t1 = b.texture(solid(0.9, 0.9, 0.9))
    chk = b.texture(checker(0.25, t0, t1))
    mats = [b.material(abi.MAT_LAMBERTIAN, chk), b.lambertian((0.4, 0.6, 0.8)),
            b.material(abi.MAT_LAMBERTIAN, b.texture(noise(4.0, seed=seed))), b.metal((0.8, 0.8, 0.7), 0.1),
            b.dielectric(1.5)]
    ns = int(n * frac_sph)
    nq = n - ns
    # vectorised draws keep large scenes quick to build
    rnd.k += (-rnd.k) % 4
    rnd.buf = []
    scale = 1.0 / max(1.0, n ** (1.0 / 3.0))
    if ns:
        c = rnd.array(3 * ns, -1.0, 1.0).reshape(-1, 3)
        r = rnd.array(ns, 0.2, 1.0) * 0.5 * scale
        m = (rnd.array(ns) * len(mats)).astype(np.int64).clip(0, len(mats) - 1)
        b.spheres.extend(map(tuple, np.concatenate([c, r[:, None]], 1)))
        b.sphere_mat.extend(int(mats[i]) for i in m)
    if nq:
        Q = rnd.array(3 * nq, -1.0, 1.0).reshape(-1, 3)
        u = rnd.array(3 * nq, -1.0, 1.0).reshape(-1, 3) * scale
        v = rnd.array(3 * nq, -1.0, 1.0).reshape(-1, 3) * scale
        bad = np.linalg.norm(np.cross(u, v), axis=1) < 1e-6 * scale * scale
        u[bad] = np.array([scale, 0, 0])
        v[bad] = np.array([0, scale, 0])
        m = (rnd.array(nq) * len(mats)).astype(np.int64).clip(0, len(mats) - 1)
        b.quads.extend(map(tuple, np.concatenate([Q, u, v], 1)))
        b.quad_mat.extend(int(mats[i]) for i in m)
        b.quad_kind.extend([abi.QUAD_PARALLELOGRAM] * nq)
    cam = Camera(res, res, 40.0, (0.0, 0.5, 4.0), (0.0, 0.0, 0.0))
    return b.arrays(), cam, spp
