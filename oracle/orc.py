"""ctypes binding of the CPU oracle (oracle/build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg to check / time the GPU path.  It reuses the
ABI structs of aurora_amd.abi because the oracle reads the same scene,
camera and render descriptors as libaurora."""
import ctypes as C
import os

import numpy as np

from aurora_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liborc.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, u32, f = C.c_void_p, C.c_uint32, C.c_float
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_philox": (None, [vp, vp, vp]),
            "orc_philox_batch": (None, [vp, u32, u32, u32, vp]),
            "orc_perlin": (None, [u32, vp, vp]),
            "orc_sin": (C.c_double, [C.c_double]),
            "orc_atan2": (C.c_double, [C.c_double, C.c_double]),
            "orc_acos": (C.c_double, [C.c_double]),
            "orc_reflect": (None, [vp, vp, vp]),
            "orc_refract": (None, [vp, vp, f, vp]),
            "orc_schlick": (f, [f, f]),
            "orc_random_unit": (None, [u32, u32, u32, u32, vp]),
            "orc_scene_create": (C.c_int, [C.POINTER(abi.SceneDesc), C.POINTER(vp)]),
            "orc_scene_destroy": (None, [vp]),
            "orc_bvh_build": (C.c_int, [vp, C.POINTER(abi.BvhStats)]),
            "orc_bvh_nodes": (C.c_int, [vp, vp, u32]),
            "orc_bvh_order": (C.c_int, [vp, vp, vp, u32]),
            "orc_bvh_wide": (C.c_int, [vp, vp, u32]),
            "orc_prim_box": (None, [vp, u32, vp, vp]),
            "orc_trace": (C.c_int, [vp, vp, vp, u32, f, f, C.c_int, vp, vp, vp, C.c_int]),
            "orc_wide_order": (C.c_int, [vp, vp, vp]),
            "orc_set_render_tree": (None, [C.c_int]),
            "orc_path_segments": (C.c_int, [vp, C.POINTER(abi.CameraDesc), C.POINTER(abi.RenderDesc), u32, u32,
                                            vp, u32]),
            "orc_hit_sphere": (C.c_int, [vp, vp, vp, f, f, vp]),
            "orc_hit_quad": (C.c_int, [vp, u32, vp, vp, f, f, vp]),
            "orc_texture_value": (None, [vp, u32, f, f, vp, vp]),
            "orc_sphere_uv": (None, [vp, vp]),
            "orc_camera": (C.c_int, [C.POINTER(abi.CameraDesc), vp]),
            "orc_camera_rays": (C.c_int, [C.POINTER(abi.CameraDesc), u32, vp, vp, u32, vp, vp]),
            "orc_resolve_rgb8": (C.c_int, [vp, C.c_uint64, u32, u32, vp]),
            "orc_paste": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp]),
            "orc_patch_render": (C.c_int, [vp, C.c_uint32, vp, C.c_uint32, vp, vp, vp, vp]),
            "orc_patch_last_error": (C.c_char_p, []),
            "orc_render_pixels": (C.c_int, [vp, C.POINTER(abi.CameraDesc), C.POINTER(abi.RenderDesc), vp, u32,
                                            C.c_int, vp, vp, vp, vp]),
            "orc_render_pixels_f64": (C.c_int, [vp, C.POINTER(abi.CameraDesc), C.POINTER(abi.RenderDesc), vp,
                                                u32, C.c_int, vp, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _err():
    return lib().orc_last_error().decode()


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def philox_batch(ctrs, k0, k1):
    c = np.ascontiguousarray(ctrs, np.uint32).reshape(-1, 4)
    out = np.zeros_like(c)
    lib().orc_philox_batch(c.ctypes.data, len(c), k0, k1, out.ctypes.data)
    return out


def perlin(seed):
    v = np.zeros((256, 3), np.float32)
    p = np.zeros((3, 256), np.int32)
    lib().orc_perlin(seed, v.ctypes.data, p.ctypes.data)
    return v, p


def sphere_uv(n):
    nn = np.asarray(n, np.float32)
    uv = np.zeros(2, np.float32)
    lib().orc_sphere_uv(nn.ctypes.data, uv.ctypes.data)
    return uv


def reflect(v, n):
    a, b = np.asarray(v, np.float32), np.asarray(n, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_reflect(a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def refract(v, n, eta):
    a, b = np.asarray(v, np.float32), np.asarray(n, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_refract(a.ctypes.data, b.ctypes.data, eta, out.ctypes.data)
    return out


def schlick(cosine, ri):
    return lib().orc_schlick(cosine, ri)


def random_unit(pixel, sample, bounce, seed=0):
    out = np.zeros(3, np.float32)
    lib().orc_random_unit(pixel, sample, bounce, seed, out.ctypes.data)
    return out


def wide_order(t, hit):
    """Descent order of a four-wide step: (m, ord) with ord[:m] the hit
    slots nearest first."""
    tt = np.ascontiguousarray(t, np.float32)
    hh = np.ascontiguousarray(hit, np.int32)
    out = np.zeros(4, np.int32)
    m = lib().orc_wide_order(tt.ctypes.data, hh.ctypes.data, out.ctypes.data)
    return m, out


def camera_frame(cam):
    out = np.zeros(18, np.float32)
    c = cam.desc()
    rc = lib().orc_camera(C.byref(c), out.ctypes.data)
    if rc:
        raise ValueError(_err())
    return out.reshape(6, 3)


def camera_rays(cam, pixels, samples, seed=0):
    """First ray of each (pixel, sample): (origins, unit directions), n x 3."""
    p = np.ascontiguousarray(pixels, np.uint32)
    s = np.ascontiguousarray(samples, np.uint32)
    n = len(p)
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    c = cam.desc()
    rc = lib().orc_camera_rays(C.byref(c), seed, p.ctypes.data, s.ctypes.data, n, o.ctypes.data, d.ctypes.data)
    if rc:
        raise ValueError(_err())
    return o, d


def resolve_rgb8(rgb_sum, spp, transfer):
    a = np.ascontiguousarray(rgb_sum, np.float32)
    out = np.zeros(a.shape, np.uint8)
    rc = lib().orc_resolve_rgb8(a.ctypes.data, a.size, spp, transfer, out.ctypes.data)
    if rc:
        raise ValueError(_err())
    return out


def paste(src, dst, corners):
    """are::Texture::paste restated: returns (new dst, pixels painted).
    src (sh, sw, 3) and dst (dh, dw, 3) float32; corners LT, RT, LB, RB."""
    s = np.ascontiguousarray(src, np.float32)
    d = np.array(dst, np.float32, copy=True)
    c = np.ascontiguousarray(corners, np.int32).reshape(8)
    n = lib().orc_paste(s.ctypes.data, s.shape[1], s.shape[0], d.ctypes.data, d.shape[1], d.shape[0], c.ctypes.data)
    return d, n


def patch_render(tris, mats, cam, cfg):
    """experiments/rt10.cpp's patch-projection renderer restated: returns
    (image (h, w, 3) float64, triangle textures of the recursion tree).
    Arguments are aurora_amd.patch structured arrays."""
    t = np.ascontiguousarray(tris)
    m = np.ascontiguousarray(mats)
    c = np.ascontiguousarray(cam)
    g = np.ascontiguousarray(cfg)
    img = np.zeros((int(c["height"].item()), int(c["width"].item()), 3), np.float64)
    n = C.c_uint64(0)
    rc = lib().orc_patch_render(t.ctypes.data if t.size else None, t.size, m.ctypes.data if m.size else None, m.size,
                                c.ctypes.data, g.ctypes.data, img.ctypes.data, C.byref(n))
    if rc != 0:
        raise ValueError(lib().orc_patch_last_error().decode())
    return img, n.value


def hit_sphere(sphere, o, d, tmin=0.0, tmax=np.inf):
    s = np.asarray(sphere, np.float32)
    oo = np.asarray(o, np.float32)
    dd = np.asarray(d, np.float32)
    t = C.c_float(0)
    h = lib().orc_hit_sphere(s.ctypes.data, oo.ctypes.data, dd.ctypes.data, tmin, tmax, C.byref(t))
    return bool(h), t.value


def hit_quad(quad, kind, o, d, tmin=0.0, tmax=np.inf):
    q = np.asarray(quad, np.float32)
    oo = np.asarray(o, np.float32)
    dd = np.asarray(d, np.float32)
    tab = np.zeros(3, np.float32)
    h = lib().orc_hit_quad(q.ctypes.data, kind, oo.ctypes.data, dd.ctypes.data, tmin, tmax, tab.ctypes.data)
    return bool(h), tab


class OracleScene:
    def __init__(self, arrays):
        self.arrays = arrays
        d, self._keep = arrays.desc()
        h = C.c_void_p()
        rc = lib().orc_scene_create(C.byref(d), C.byref(h))
        if rc == abi.AURORA_ERR_INVALID:
            raise ValueError(_err())
        if rc:
            raise RuntimeError(_err())
        self.h = h
        self.stats = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def build_bvh(self):
        st = abi.BvhStats()
        if lib().orc_bvh_build(self.h, C.byref(st)):
            raise RuntimeError(_err())
        from aurora_amd.core import stats_dict
        self.stats = stats_dict(st)
        return self.stats

    def nodes(self):
        n = self.stats["n_nodes"]
        buf = np.zeros((max(n, 1), 8), np.uint32)
        lib().orc_bvh_nodes(self.h, buf.ctypes.data, n)
        return buf[:n]

    def wide(self):
        """Four-wide traversal nodes as (n_wide, 32) uint32 words."""
        n = self.stats["n_wide"]
        buf = np.zeros((max(n, 1), 32), np.uint32)
        lib().orc_bvh_wide(self.h, buf.ctypes.data, n)
        return buf[:n]

    def order(self):
        n = self.arrays.n_prims
        m = np.zeros(max(n, 1), np.uint32)
        o = np.zeros(max(n, 1), np.uint32)
        lib().orc_bvh_order(self.h, m.ctypes.data, o.ctypes.data, n)
        return m[:n], o[:n]

    def prim_box(self, i):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        lib().orc_prim_box(self.h, i, lo.ctypes.data, hi.ctypes.data)
        return lo, hi

    def trace(self, origins, dirs, tmin=0.0, tmax=np.inf, brute=False, counts=False, threads=0, binary=False):
        """Closest hits over the four-wide tree (default), the binary tree
        (binary=True) or every primitive (brute=True)."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(o)
        prim = np.zeros(max(n, 1), np.int32)
        t = np.zeros(max(n, 1), np.float32)
        cnt = np.zeros((max(n, 1), 3), np.uint32) if counts else None
        rc = lib().orc_trace(self.h, o.ctypes.data, d.ctypes.data, n, tmin, tmax, 1 if brute else 2 if binary else 0,
                             prim.ctypes.data, t.ctypes.data, cnt.ctypes.data if counts else None, threads)
        if rc:
            raise RuntimeError(_err())
        if counts:
            return prim[:n], t[:n], cnt[:n]
        return prim[:n], t[:n]

    def texture_value(self, tex, u, v, p):
        pp = np.asarray(p, np.float32)
        out = np.zeros(3, np.float32)
        lib().orc_texture_value(self.h, tex, u, v, pp.ctypes.data, out.ctypes.data)
        return out

    def path_segments(self, cam, rs, pixel, sample, wide=False, cap=256):
        """The segments of one path: (origins, directions, prims, ts)."""
        out = np.zeros((cap, 8), np.float32)
        c, r = cam.desc(), rs.desc()
        lib().orc_set_render_tree(1 if wide else 0)
        try:
            n = lib().orc_path_segments(self.h, C.byref(c), C.byref(r), pixel, sample, out.ctypes.data, cap)
        finally:
            lib().orc_set_render_tree(0)
        if n < 0:
            raise RuntimeError(_err())
        seg = out[:min(n, cap)]
        return seg[:, 0:3], seg[:, 3:6], seg[:, 6].view(np.int32), seg[:, 7]

    def render(self, cam, rs, pixels=None, f64=False, threads=0, aov=None, wide=False):
        """Sum of rs.spp samples per pixel (row-major pixel ids or all);
        wide: walk the four-wide tree instead of the binary one."""
        W, H = cam.image_width, cam.image_height
        n = W * H if pixels is None else len(pixels)
        dt = np.float64 if f64 else np.float32
        rgb = np.zeros((n, 3), dt)
        use_aov = rs.aov if aov is None else aov
        alb = np.zeros((n, 3), dt) if use_aov else None
        nrm = np.zeros((n, 3), dt) if use_aov else None
        px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
        segs = C.c_uint64(0)
        fn = lib().orc_render_pixels_f64 if f64 else lib().orc_render_pixels
        c, r = cam.desc(), rs.desc()
        lib().orc_set_render_tree(1 if wide else 0)
        try:
            rc = fn(self.h, C.byref(c), C.byref(r), px.ctypes.data if px is not None else None, n, threads,
                    rgb.ctypes.data, alb.ctypes.data if use_aov else None, nrm.ctypes.data if use_aov else None,
                    C.byref(segs))
        finally:
            lib().orc_set_render_tree(0)
        if rc == abi.AURORA_ERR_INVALID:
            raise ValueError(_err())
        if rc:
            raise RuntimeError(_err())
        if pixels is None:
            rgb = rgb.reshape(H, W, 3)
            if use_aov:
                alb = alb.reshape(H, W, 3)
                nrm = nrm.reshape(H, W, 3)
        return rgb, alb, nrm, segs.value
