"""Patch-projection renderer descriptors (aurora_patch_* of include/aurora.h)
as numpy structured dtypes, scene builders in the reference's terms, the
reference's own Cornell scene, and its 8-bit writer.

The renderer is the recursive mirror renderer of experiments/rt10.cpp:
Context.patch_render runs it on the GPU, bit for bit the reference's
double image.
"""
import numpy as np

from . import abi

DIFFUSE = 0
REFLECTIVE = 1
NO_MATERIAL = 0xFFFFFFFF
MAX_DEPTH = 64
MAX_TEX_RES = 4096

MATERIAL = np.dtype([("type", "<u4"), ("reserved", "<u4"), ("albedo", "<f8", (3,)), ("metalness", "<f8")])
TRIANGLE = np.dtype([("p", "<f8", (3, 3)), ("uv", "<f8", (3, 2)), ("id", "<i4"), ("material", "<u4")])
CAMERA = np.dtype([("origin", "<f8", (3,)), ("viewport", TRIANGLE, (2,)), ("width", "<u4"), ("height", "<u4")])
CONFIG = np.dtype([("max_depth", "<i4"), ("max_tex_res", "<i4"), ("min_tex_res", "<i4"), ("reserved", "<u4"),
                   ("min_area_px", "<f8"), ("env", "<f8", (3,))])


class Stats(abi.C.Structure):
    _fields_ = [
        ("textures", abi.C.c_uint32),
        ("raster_textures", abi.C.c_uint32),
        ("levels", abi.C.c_uint32),
        ("warps", abi.C.c_uint32),
        ("texels", abi.C.c_uint64),
        ("warp_texels", abi.C.c_uint64),
        ("plan_ms", abi.C.c_double),
        ("kernel_ms", abi.C.c_double),
    ]


class SceneBuilder:
    """Triangles and materials in the reference's terms (rt10.cpp:789-830:
    a quad is two triangles with uv (0,0)-(1,1); a box is six quads); ids
    count from 1 as the reference's do."""

    def __init__(self):
        self.tris = []
        self.mats = []
        self.next_id = 1

    def material(self, kind, albedo, metalness=0.0):
        m = np.zeros((), MATERIAL)
        m["type"] = kind
        m["albedo"] = albedo
        m["metalness"] = metalness
        self.mats.append(m)
        return len(self.mats) - 1

    def triangle(self, p0, p1, p2, uv, mat):
        t = np.zeros((), TRIANGLE)
        t["p"] = [p0, p1, p2]
        t["uv"] = uv
        t["id"] = self.next_id
        t["material"] = mat
        self.next_id += 1
        self.tris.append(t)

    def quad(self, p00, p10, p11, p01, mat):
        self.triangle(p00, p10, p01, [(0, 0), (1, 0), (0, 1)], mat)
        self.triangle(p10, p11, p01, [(1, 0), (1, 1), (0, 1)], mat)

    def box(self, lo, hi, mat):
        (x0, y0, z0), (x1, y1, z1) = lo, hi
        c = {(i, j, k): (x1 if i else x0, y1 if j else y0, z1 if k else z0) for i in (0, 1) for j in (0, 1)
             for k in (0, 1)}
        for a, b, cc, d in [((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)), ((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),
                            ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)), ((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),
                            ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)), ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))]:
            self.quad(c[a], c[b], c[cc], c[d], mat)

    def arrays(self):
        tris = np.array(self.tris, TRIANGLE) if self.tris else np.zeros(0, TRIANGLE)
        mats = np.array(self.mats, MATERIAL) if self.mats else np.zeros(0, MATERIAL)
        return tris, mats


def camera(origin, viewport_z, center_y, view_h, width, height):
    """The reference's camera (rt10.cpp:888-915): a viewport rectangle of
    height view_h at z = viewport_z, centred at (0, center_y), split into
    triangles A = (TL, TR, BL) and B = (TR, BR, BL) with uv u right, v down."""
    c = np.zeros((), CAMERA)
    c["origin"] = origin
    aspect = float(width) / float(height)
    vw = view_h * aspect
    tl = (0.0 - vw * 0.5, center_y - view_h * 0.5, viewport_z)
    tr = (0.0 + vw * 0.5, center_y - view_h * 0.5, viewport_z)
    bl = (0.0 - vw * 0.5, center_y + view_h * 0.5, viewport_z)
    br = (0.0 + vw * 0.5, center_y + view_h * 0.5, viewport_z)
    c["viewport"][0]["p"] = [tl, tr, bl]
    c["viewport"][0]["uv"] = [(0, 0), (1, 0), (0, 1)]
    c["viewport"][0]["id"] = -100
    c["viewport"][1]["p"] = [tr, br, bl]
    c["viewport"][1]["uv"] = [(1, 0), (1, 1), (0, 1)]
    c["viewport"][1]["id"] = -101
    c["viewport"]["material"] = NO_MATERIAL
    c["width"] = width
    c["height"] = height
    return c


def config(max_depth=4, min_area_px=4.0, max_tex_res=256, min_tex_res=16, env=(0.08, 0.08, 0.10)):
    """RenderConfig (rt10.cpp:543-549); defaults are the reference's."""
    g = np.zeros((), CONFIG)
    g["max_depth"] = max_depth
    g["min_area_px"] = min_area_px
    g["max_tex_res"] = max_tex_res
    g["min_tex_res"] = min_tex_res
    g["env"] = env
    return g


def rt10_cornell(width=900, height=650):
    """The reference program's scene (rt10.cpp:837-922): an open-front
    Cornell room x in [-1,1], y in [0,2], z in [0,2] with white floor,
    ceiling and back wall, red left and green right walls, a blue and a
    yellow reflective box; the eye at (0,1,-3) looking down +z through a
    1.6-high viewport at z = -2; max depth 4, no recursion below 6 texels,
    textures 16..256, environment (0.06, 0.07, 0.09)."""
    b = SceneBuilder()
    white = b.material(DIFFUSE, (0.85, 0.85, 0.85))
    red = b.material(DIFFUSE, (0.85, 0.25, 0.25))
    green = b.material(DIFFUSE, (0.25, 0.85, 0.25))
    blue = b.material(REFLECTIVE, (0.25, 0.45, 1.0), 0.90)
    yellow = b.material(REFLECTIVE, (1.0, 0.92, 0.20), 0.92)
    b.quad((-1, 0, 0), (1, 0, 0), (1, 0, 2), (-1, 0, 2), white)
    b.quad((-1, 2, 0), (-1, 2, 2), (1, 2, 2), (1, 2, 0), white)
    b.quad((-1, 0, 2), (1, 0, 2), (1, 2, 2), (-1, 2, 2), white)
    b.quad((-1, 0, 0), (-1, 0, 2), (-1, 2, 2), (-1, 2, 0), red)
    b.quad((1, 0, 0), (1, 2, 0), (1, 2, 2), (1, 0, 2), green)
    b.box((-0.70, 0.0, 0.80), (-0.15, 0.60, 1.30), blue)
    b.box((0.15, 0.0, 1.00), (0.70, 1.10, 1.65), yellow)
    tris, mats = b.arrays()
    cam = camera((0.0, 1.0, -3.0), -2.0, 1.0, 1.6, width, height)
    cfg = config(max_depth=4, min_area_px=6.0, max_tex_res=256, min_tex_res=16, env=(0.06, 0.07, 0.09))
    return tris, mats, cam, cfg


def mirror_hall(tiles=4, width=640, height=480, max_depth=6, min_area_px=4.0, max_tex_res=256, min_tex_res=8):
    """A corridor of two facing mirrors: the reference's room (x in [-1,1],
    y in [0,2], z in [0,2]) with mirrors for its back wall (z = 2) and its
    front wall (z = 0), red and green side walls, a white floor and
    ceiling, and a diffuse box between the mirrors; every wall split into
    tiles x tiles quads.  The eye stands inside the room at (0.15, 1.1,
    0.3), looking down +z at the back mirror through a viewport 0.4 ahead
    (the reference's field of view).  The back mirror shows the front one,
    whose textures show the back one again, and so on down to max_depth or
    until the textures shrink below min_area_px; the tessellation keeps
    triangle vertices inside every tile's view, which the reference's
    visibility test (a vertex projecting inside) needs."""
    b = SceneBuilder()
    white = b.material(DIFFUSE, (0.80, 0.78, 0.72))
    red = b.material(DIFFUSE, (0.80, 0.25, 0.22))
    green = b.material(DIFFUSE, (0.25, 0.75, 0.30))
    box = b.material(DIFFUSE, (0.30, 0.45, 0.85))
    back = b.material(REFLECTIVE, (0.95, 0.95, 0.90), 0.85)
    front = b.material(REFLECTIVE, (0.90, 0.95, 1.00), 0.80)
    g = [2.0 * k / tiles for k in range(tiles + 1)]
    x = [-1.0 + 2.0 * k / tiles for k in range(tiles + 1)]
    for i in range(tiles):
        for j in range(tiles):
            y0, y1, z0, z1, x0, x1 = g[i], g[i + 1], g[j], g[j + 1], x[j], x[j + 1]
            b.quad((x0, y0, 2), (x1, y0, 2), (x1, y1, 2), (x0, y1, 2), back)
            b.quad((x0, y0, 0), (x0, y1, 0), (x1, y1, 0), (x1, y0, 0), front)
            b.quad((-1, y0, z0), (-1, y0, z1), (-1, y1, z1), (-1, y1, z0), red)
            b.quad((1, y0, z0), (1, y1, z0), (1, y1, z1), (1, y0, z1), green)
            xa, xb = x[i], x[i + 1]
            b.quad((xa, 0, z0), (xb, 0, z0), (xb, 0, z1), (xa, 0, z1), white)
            b.quad((xa, 2, z0), (xa, 2, z1), (xb, 2, z1), (xb, 2, z0), white)
    b.box((-0.55, 0.0, 1.1), (-0.05, 0.6, 1.5), box)
    tris, mats = b.arrays()
    cam = camera((0.15, 1.1, 0.3), 0.7, 1.1, 0.64, width, height)
    cfg = config(max_depth=max_depth, min_area_px=min_area_px, max_tex_res=max_tex_res, min_tex_res=min_tex_res,
                 env=(0.05, 0.05, 0.07))
    return tris, mats, cam, cfg


def random_scene(n_tris, seed, reflective_share=0.5, width=64, height=48, max_depth=3):
    """Seeded triangles in front of the reference camera: random vertices in
    the room, random uv, materials drawn from a few diffuse and reflective
    ones (some ids repeated, some triangles without a material)."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    mats = [b.material(DIFFUSE, rng.uniform(0.1, 1.0, 3)) for _ in range(3)]
    mats += [b.material(REFLECTIVE, rng.uniform(0.2, 1.0, 3), rng.uniform(0.0, 1.2)) for _ in range(3)]
    for i in range(n_tris):
        c = rng.uniform((-1, 0, 0), (1, 2, 2))
        p = [c + rng.normal(0, 0.4, 3) for _ in range(3)]
        uv = rng.uniform(0, 1, (3, 2))
        if rng.uniform() < 0.05:
            mat = NO_MATERIAL
        else:
            mat = mats[3 + rng.integers(3)] if rng.uniform() < reflective_share else mats[rng.integers(3)]
        b.triangle(p[0], p[1], p[2], uv, mat)
        if i % 7 == 3:
            b.next_id -= 1  # a repeated id
    tris, marr = b.arrays()
    cam = camera((rng.uniform(-0.3, 0.3), 1.0 + rng.uniform(-0.2, 0.2), -3.0), -2.0, 1.0, 1.6, width, height)
    cfg = config(max_depth=max_depth, min_area_px=4.0, max_tex_res=64, min_tex_res=4,
                 env=tuple(rng.uniform(0, 0.2, 3)))
    return tris, marr, cam, cfg


def to_rgb8(img, gamma=2.2):
    """The reference's 8-bit writer (rt10.cpp:118-137): clamp to [0,1],
    c^(1/gamma), round half away from zero of 255 c, clamp to [0,255]."""
    c = np.clip(np.asarray(img, np.float64), 0.0, 1.0)
    c = np.power(c, 1.0 / gamma) * 255.0
    fl = np.floor(c)
    r = fl + (c - fl >= 0.5)
    return np.clip(r, 0, 255).astype(np.uint8)


def ppm_bytes(img, gamma=2.2):
    h, w = img.shape[:2]
    return b"P6\n%d %d\n255\n" % (w, h) + to_rgb8(img, gamma).tobytes()
