"""Python handles over libaurora's C ABI.

`SceneArrays` is the flattened scene (structure-of-arrays numpy buffers, the
layout of aurora_scene_desc).  `Context` owns a device and stream; `Scene`
owns an uploaded scene and its LBVH.  Errors from the library raise
ValueError (AURORA_ERR_INVALID) or RuntimeError (everything else), the
Python analogues of the std::invalid_argument / std::runtime_error the
reference's constructors throw (src/object/triangle.cpp:13-36,
src/texture.cpp:11-43).
"""
import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import abi


# fused-tail switch above any queue length: every path of a pass runs in the
# persistent tail kernel (the default; 0 = wavefront kernels only)
TAIL_ALWAYS = (1 << 32) - 1


@dataclass
class SceneArrays:
    spheres: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    sphere_mat: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint32))
    quads: np.ndarray = field(default_factory=lambda: np.zeros((0, 9), np.float32))
    quad_mat: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint32))
    quad_kind: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint32))
    materials: list = field(default_factory=list)  # (type, texture, fuzz, ior)
    textures: list = field(default_factory=list)   # dicts, see texture()
    texels: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))

    @property
    def n_prims(self):
        return len(self.spheres) + len(self.quads)

    def desc(self):
        """Build an aurora_scene_desc; returns (desc, keepalive)."""
        sp = np.ascontiguousarray(self.spheres, np.float32).reshape(-1, 4)
        sm = np.ascontiguousarray(self.sphere_mat, np.uint32).reshape(-1)
        qd = np.ascontiguousarray(self.quads, np.float32).reshape(-1, 9)
        qm = np.ascontiguousarray(self.quad_mat, np.uint32).reshape(-1)
        qk = np.ascontiguousarray(self.quad_kind, np.uint32).reshape(-1)
        if len(qk) != len(qd):
            qk = np.zeros(len(qd), np.uint32)
        mats = (abi.Material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = abi.Material(int(m[0]), int(m[1]), float(m[2]), float(m[3]))
        texs = (abi.Texture * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            texs[i] = abi.Texture(
                int(t.get("type", 0)), int(t.get("even", 0)), int(t.get("odd", 0)), int(t.get("seed", 0)),
                (C.c_float * 3)(*[float(x) for x in t.get("color", (0, 0, 0))]), float(t.get("scale", 1.0)),
                int(t.get("width", 0)), int(t.get("height", 0)), int(t.get("texel_offset", 0)))
        tx = np.ascontiguousarray(self.texels, np.float32).reshape(-1, 3)
        d = abi.SceneDesc()
        d.n_spheres = len(sp)
        d.spheres = sp.ctypes.data if len(sp) else None
        d.sphere_material = sm.ctypes.data if len(sm) else None
        d.n_quads = len(qd)
        d.quads = qd.ctypes.data if len(qd) else None
        d.quad_material = qm.ctypes.data if len(qm) else None
        d.quad_kind = qk.ctypes.data if len(qk) else None
        d.n_materials = len(self.materials)
        d.materials = C.addressof(mats)
        d.n_textures = len(self.textures)
        d.textures = C.addressof(texs)
        d.n_texels = len(tx)
        d.texels = tx.ctypes.data if len(tx) else None
        return d, (sp, sm, qd, qm, qk, mats, texs, tx)


def solid(r, g, b):
    return {"type": abi.TEX_SOLID, "color": (r, g, b)}


def checker(scale, even, odd):
    return {"type": abi.TEX_CHECKER, "scale": scale, "even": even, "odd": odd}


def noise(scale, seed=0):
    return {"type": abi.TEX_NOISE, "scale": scale, "seed": seed}


def image(width, height, texel_offset):
    return {"type": abi.TEX_IMAGE, "width": width, "height": height, "texel_offset": texel_offset}


@dataclass
class Camera:
    image_width: int
    image_height: int
    vfov: float
    lookfrom: tuple
    lookat: tuple
    vup: tuple = (0.0, 1.0, 0.0)
    defocus_angle: float = 0.0
    focus_dist: float = 10.0

    def desc(self):
        c = abi.CameraDesc()
        c.image_width = self.image_width
        c.image_height = self.image_height
        c.vfov = self.vfov
        c.lookfrom = (C.c_float * 3)(*self.lookfrom)
        c.lookat = (C.c_float * 3)(*self.lookat)
        c.vup = (C.c_float * 3)(*self.vup)
        c.defocus_angle = self.defocus_angle
        c.focus_dist = self.focus_dist
        return c


@dataclass
class RenderSettings:
    spp: int
    max_depth: int = 50
    seed: int = 0
    sample_offset: int = 0
    background: tuple = (0.0, 0.0, 0.0)
    aov: bool = False
    count_fetches: bool = False
    graph: bool = False
    spp_per_pass: int = 0
    tmin: float = 0.0

    def desc(self):
        r = abi.RenderDesc()
        r.spp = self.spp
        r.sample_offset = self.sample_offset
        r.max_depth = self.max_depth
        r.seed = self.seed
        r.background = (C.c_float * 3)(*self.background)
        r.flags = ((abi.RENDER_AOV if self.aov else 0) | (abi.RENDER_COUNT_FETCHES if self.count_fetches else 0)
                   | (abi.RENDER_GRAPH if self.graph else 0))
        r.spp_per_pass = self.spp_per_pass
        r.tmin = self.tmin
        return r


def stats_dict(st):
    return {k: getattr(st, k) if not isinstance(getattr(st, k), C.Array) else list(getattr(st, k))
            for k, _ in st._fields_}


class Context:
    def __init__(self, device=0):
        self.lib = abi.load_library()
        h = C.c_void_p()
        rc = self.lib.aurora_create(device, C.byref(h))
        if rc != abi.AURORA_OK:
            raise RuntimeError(f"aurora_create failed ({rc}): {self.lib.aurora_last_error(None).decode()}")
        self.h = h
        self._children = weakref.WeakSet()  # scenes / denoisers: closed before the context

    def adopt(self, child):
        self._children.add(child)
        return child

    def check(self, rc):
        if rc != abi.AURORA_OK:
            msg = self.lib.aurora_last_error(self.h).decode()
            raise (ValueError if rc == abi.AURORA_ERR_INVALID else RuntimeError)(f"libaurora error {rc}: {msg}")

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.close()
            self.lib.aurora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, extend_blocks_per_cu=8, shade_blocks_per_cu=8, tail_paths=TAIL_ALWAYS):
        self.check(self.lib.aurora_set_launch(self.h, extend_blocks_per_cu, shade_blocks_per_cu, tail_paths))

    def set_render_tree(self, tree=0):
        """Tree the render kernels walk: 0 (or False) the binary LBVH
        (default), 1 (or True) the four-wide tree, 2 the flat leaf array of a
        small scene (single-scene renders; four-wide elsewhere).  Speed only;
        the images are the same."""
        self.check(self.lib.aurora_set_render_tree(self.h, int(tree)))

    def selftest_rcp(self, lo, hi):
        """(mismatches, first mismatching bit patterns) of the kernels' fast
        reciprocal against IEEE division over float32 bit patterns [lo, hi)."""
        n = C.c_uint64(0)
        first = (C.c_uint32 * 16)()
        self.check(self.lib.aurora_selftest_rcp(self.h, int(lo), int(hi), C.byref(n), first))
        return int(n.value), [int(first[k]) for k in range(min(16, int(n.value)))]

    def set_queue(self, pass_paths=0, claim_paths=0):
        """Queue sizes: most paths per automatic pass (0: 2^27) and paths per
        work-pool claim in the fused tail (0: by pass size).  Speed only."""
        self.check(self.lib.aurora_set_queue(self.h, int(pass_paths), int(claim_paths)))

    def set_ray_sort(self, mode=1):
        """Ray queries: 0 trace in the given order, 1 sort large queries on
        large trees (default), 2 always sort.  Hits are the same."""
        self.check(self.lib.aurora_set_ray_sort(self.h, mode))

    def synchronize(self):
        self.check(self.lib.aurora_synchronize(self.h))

    def stream_handle(self):
        return self.lib.aurora_stream(self.h)

    def upload(self, arrays: SceneArrays):
        return Scene(self, arrays)

    def camera_rays(self, cam, pixels, samples, seed=0):
        """First ray of each (pixel, sample) of a render with this camera:
        (origins, unit directions) as (n, 3) float32.  Host buffers."""
        p = np.ascontiguousarray(pixels, np.uint32)
        s = np.ascontiguousarray(samples, np.uint32)
        n = len(p)
        o = np.zeros((max(n, 1), 3), np.float32)
        d = np.zeros((max(n, 1), 3), np.float32)
        c = cam.desc()
        self.check(self.lib.aurora_camera_rays(self.h, C.byref(c), seed, p.ctypes.data, s.ctypes.data, n,
                                               o.ctypes.data, d.ctypes.data))
        return o[:n], d[:n]

    def camera_rays_device(self, cam, pixels, samples, origins, dirs, seed=0):
        """Device form: pixels/samples int32 or uint32 tensors (n,), origins and
        dirs float32 (n, 3).  Asynchronous on the context stream."""
        c = cam.desc()
        self.check(self.lib.aurora_camera_rays_device(self.h, C.byref(c), seed, pixels.data_ptr(),
                                                      samples.data_ptr(), pixels.shape[0], origins.data_ptr(),
                                                      dirs.data_ptr()))

    def paste(self, src, dst, corners):
        """are::Texture::paste on the GPU (host buffers): src (sh, sw, 3) is
        warped into the quad corners = (LT, RT, LB, RB) (x, y pairs) of a
        copy of dst (dh, dw, 3); returns the new destination."""
        s = np.ascontiguousarray(src, np.float32)
        d = np.array(dst, np.float32, copy=True)
        c = np.ascontiguousarray(corners, np.int32).reshape(8)
        self.check(self.lib.aurora_paste(self.h, s.ctypes.data, s.shape[1], s.shape[0], d.ctypes.data, d.shape[1],
                                         d.shape[0], c.ctypes.data))
        return d

    def paste_device(self, src, dst, corners):
        """Device form on float32 tensors (sh, sw, 3) and (dh, dw, 3); dst is
        updated in place, asynchronously on the context stream."""
        c = np.ascontiguousarray(corners, np.int32).reshape(8)
        self.check(self.lib.aurora_paste_device(self.h, src.data_ptr(), src.shape[1], src.shape[0], dst.data_ptr(),
                                                dst.shape[1], dst.shape[0], c.ctypes.data))

    def patch_render(self, tris, mats, cam, cfg):
        """experiments/rt10.cpp's patch-projection renderer on the GPU
        (aurora_patch_render): tris, mats, cam, cfg are aurora_amd.patch
        structured arrays; returns (image (h, w, 3) float64, stats dict)."""
        from . import patch

        t, m, c, g = (np.ascontiguousarray(x) for x in (tris, mats, cam, cfg))
        assert t.dtype == patch.TRIANGLE and m.dtype == patch.MATERIAL
        assert c.dtype == patch.CAMERA and g.dtype == patch.CONFIG
        img = np.zeros((int(c["height"].item()), int(c["width"].item()), 3), np.float64)
        st = patch.Stats()
        self.check(self.lib.aurora_patch_render(self.h, t.ctypes.data if t.size else None, t.size,
                                                m.ctypes.data if m.size else None, m.size, c.ctypes.data,
                                                g.ctypes.data, img.ctypes.data, C.byref(st)))
        return img, {k: getattr(st, k) for k, _ in st._fields_}

    def patch_render_device(self, tris, mats, cam, cfg, out):
        """Same into a float64 device tensor out (h, w, 3); returns stats."""
        from . import patch

        t, m, c, g = (np.ascontiguousarray(x) for x in (tris, mats, cam, cfg))
        st = patch.Stats()
        self.check(self.lib.aurora_patch_render_device(self.h, t.ctypes.data if t.size else None, t.size,
                                                       m.ctypes.data if m.size else None, m.size, c.ctypes.data,
                                                       g.ctypes.data, out.data_ptr(), C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def resolve_rgb8_device(self, rgb_sum, spp, out, transfer=abi.TRANSFER_GAMMA22):
        """8-bit resolve of a device radiance sum (H, W, 3) float32 into a
        uint8 tensor of the same shape.  Asynchronous on the context stream."""
        H, W = rgb_sum.shape[:2]
        self.check(self.lib.aurora_resolve_rgb8_device(self.h, rgb_sum.data_ptr(), W, H, spp, transfer,
                                                       out.data_ptr()))


class Scene:
    def __init__(self, ctx: Context, arrays: SceneArrays):
        self.ctx = ctx
        self.lib = ctx.lib
        self.arrays = arrays
        d, keep = arrays.desc()
        h = C.c_void_p()
        ctx.check(self.lib.aurora_scene_upload(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.bvh_stats = None
        ctx.adopt(self)

    def close(self):
        if self.h:
            self.lib.aurora_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_bvh(self):
        st = abi.BvhStats()
        self.ctx.check(self.lib.aurora_bvh_build(self.h, C.byref(st)))
        self.bvh_stats = stats_dict(st)
        return self.bvh_stats

    def nodes(self):
        n = self.bvh_stats["n_nodes"]
        buf = np.zeros((max(n, 1), 8), np.uint32)
        self.ctx.check(self.lib.aurora_bvh_download(self.h, buf.ctypes.data, n))
        return buf[:n]

    def wide(self):
        """Four-wide traversal nodes as (n_wide, 32) uint32 words."""
        n = self.bvh_stats["n_wide"]
        buf = np.zeros((max(n, 1), 32), np.uint32)
        self.ctx.check(self.lib.aurora_bvh_download_wide(self.h, buf.ctypes.data, n))
        return buf[:n]

    def order(self):
        n = self.arrays.n_prims
        m = np.zeros(max(n, 1), np.uint32)
        o = np.zeros(max(n, 1), np.uint32)
        self.ctx.check(self.lib.aurora_bvh_download_order(self.h, m.ctypes.data, o.ctypes.data, n))
        return m[:n], o[:n]

    def trace(self, origins, dirs, tmin=0.0, tmax=np.inf):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(o)
        prim = np.zeros(max(n, 1), np.int32)
        t = np.zeros(max(n, 1), np.float32)
        self.ctx.check(self.lib.aurora_trace_closest(self.h, o.ctypes.data, d.ctypes.data, n, tmin, tmax,
                                                     prim.ctypes.data, t.ctypes.data))
        return prim[:n], t[:n]

    def trace_device(self, o, d, prim, t, tmin=0.0, tmax=float("inf"), counts=None, brute=False):
        """o, d: device tensors (n,3) float32; prim int32 (n,), t float32 (n,);
        counts optional uint32/int32 (n,3).  Asynchronous on the context stream."""
        n = o.shape[0]
        if brute:
            rc = self.lib.aurora_trace_brute_device(self.h, o.data_ptr(), d.data_ptr(), n, tmin, tmax,
                                                    prim.data_ptr(), t.data_ptr())
        else:
            rc = self.lib.aurora_trace_closest_device(self.h, o.data_ptr(), d.data_ptr(), n, tmin, tmax,
                                                      prim.data_ptr(), t.data_ptr(),
                                                      counts.data_ptr() if counts is not None else None)
        self.ctx.check(rc)

    def trace_device_noalloc(self, o, d, prim, t, tmin=0.0, tmax=float("inf"), counts=None):
        """aurora_trace_closest_device_noalloc: as trace_device, with every
        intermediate buffer in the library's static per-device scratch array
        (at most abi.NOALLOC_MAX_RAYS rays).  Asynchronous."""
        self.ctx.check(self.lib.aurora_trace_closest_device_noalloc(
            self.h, o.data_ptr(), d.data_ptr(), o.shape[0], tmin, tmax, prim.data_ptr(), t.data_ptr(),
            counts.data_ptr() if counts is not None else None))

    def trace_device_timed(self, o, d, prim, t, tmin=0.0, tmax=float("inf"), count_fetches=False):
        """Device trace with the traversal kernel timed by HIP events;
        returns the stats dict (extend_ms, fetch totals when counted)."""
        st = abi.TraceStats()
        self.ctx.check(self.lib.aurora_trace_closest_device_stats(
            self.h, o.data_ptr(), d.data_ptr(), o.shape[0], tmin, tmax, prim.data_ptr(), t.data_ptr(),
            1 if count_fetches else 0, C.byref(st)))
        return stats_dict(st)

    def render(self, cam: Camera, rs: RenderSettings, rgb=None, albedo=None, normal=None):
        """Host-buffer render; returns (rgb_sum, albedo_sum, normal_sum, stats)."""
        W, H = cam.image_width, cam.image_height
        rgb = np.zeros((H, W, 3), np.float32) if rgb is None else np.ascontiguousarray(rgb, np.float32)
        if rs.aov:
            albedo = np.zeros((H, W, 3), np.float32) if albedo is None else np.ascontiguousarray(albedo, np.float32)
            normal = np.zeros((H, W, 3), np.float32) if normal is None else np.ascontiguousarray(normal, np.float32)
        st = abi.RenderStats()
        c, r = cam.desc(), rs.desc()
        self.ctx.check(self.lib.aurora_render(self.h, C.byref(c), C.byref(r), rgb.ctypes.data,
                                              albedo.ctypes.data if rs.aov else None,
                                              normal.ctypes.data if rs.aov else None, C.byref(st)))
        return rgb, albedo, normal, stats_dict(st)

    def render_device(self, cam: Camera, rs: RenderSettings, rgb, albedo=None, normal=None, stats=True):
        """Device render into torch tensors (H, W, 3) float32 on the context's
        device; synchronous on return when stats are requested."""
        st = abi.RenderStats()
        c, r = cam.desc(), rs.desc()
        self.ctx.check(self.lib.aurora_render_device(
            self.h, C.byref(c), C.byref(r), rgb.data_ptr(),
            albedo.data_ptr() if albedo is not None else None,
            normal.data_ptr() if normal is not None else None, C.byref(st) if stats else None))
        if stats:
            self.ctx.synchronize()
        return stats_dict(st) if stats else None


def render_batch(ctx, items, graph=False):
    """Render many scenes in shared passes.  items: (scene, camera,
    settings, rgb, albedo, normal) with device tensors (albedo/normal may be
    None).  Returns the stats dict; synchronous."""
    arr = (abi.BatchItem * max(1, len(items)))()
    for k, (sc, cam, rs, rgb, alb, nrm) in enumerate(items):
        arr[k].scene = sc.h
        arr[k].camera = cam.desc()
        arr[k].render = rs.desc()
        arr[k].d_rgb_sum = rgb.data_ptr()
        arr[k].d_albedo_sum = alb.data_ptr() if alb is not None else None
        arr[k].d_normal_sum = nrm.data_ptr() if nrm is not None else None
    st = abi.RenderStats()
    ctx.check(ctx.lib.aurora_render_batch_device(ctx.h, arr, len(items), abi.RENDER_GRAPH if graph else 0,
                                                 C.byref(st)))
    ctx.synchronize()
    return stats_dict(st)


def resolve_rgb8(rgb_sum, spp, transfer=abi.TRANSFER_GAMMA22):
    """8-bit bytes (H, W, 3) of a radiance sum over spp samples, with the
    transfer curve of experiments/rt.cpp:377-392 by default (abi.TRANSFER_*)."""
    lib = abi.load_library()
    a = np.ascontiguousarray(rgb_sum, np.float32)
    H, W = a.shape[:2]
    out = np.zeros((H, W, 3), np.uint8)
    rc = lib.aurora_resolve_rgb8(a.ctypes.data, W, H, spp, transfer, out.ctypes.data)
    if rc != abi.AURORA_OK:
        raise ValueError("aurora_resolve_rgb8: invalid arguments")
    return out


def write_ppm(path, rgb8):
    """P6 writer with the byte layout of are::Texture::save_texture
    (src/texture.cpp:370-391): header 'P6\\n<w> <h>\\n255\\n', rows top-down."""
    H, W = rgb8.shape[:2]
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (W, H))
        f.write(np.ascontiguousarray(rgb8, np.uint8).tobytes())
