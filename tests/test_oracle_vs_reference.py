"""The oracle against routines of the reference renderer itself, compiled
from its sources into oracle/_ref (oracle/ref/Makefile): reflect / refract
(src/basic/vec3.cpp:182-199), the plane intersection of
src/basic/plane.cpp:13-27 (pins the quad test's hit distance), and the
bilinear sampling inside are::Texture::paste (src/texture.cpp:310-330, pins
the image texture).  Skipped where the reference was not available to
build."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import orc
from aurora_amd.core import SceneArrays, image

LIB = os.path.join(os.path.dirname(orc.HERE), "oracle", "_ref", "libareref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref not built (no reference sources)")


@pytest.fixture(scope="module")
def ref():
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.ref_reflect.argtypes = [vp, vp, vp]
    L.ref_refract.argtypes = [vp, vp, C.c_double, vp]
    L.ref_plane_hit.argtypes = [vp, vp, vp, vp, vp]
    L.ref_plane_hit.restype = C.c_int
    L.ref_paste.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, C.c_double, vp]
    return L


def d3(x):
    return np.ascontiguousarray(x, np.float64)


def test_reflect_matches_reference(ref):
    rng = np.random.default_rng(0)
    for _ in range(500):
        v = rng.normal(size=3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v32, n32 = v.astype(np.float32), n.astype(np.float32)
        out = np.zeros(3)
        a, b = d3(v32), d3(n32)
        ref.ref_reflect(a.ctypes.data, b.ctypes.data, out.ctypes.data)
        assert np.allclose(orc.reflect(v32, n32), out, rtol=1e-5, atol=1e-5)


def test_refract_matches_reference(ref):
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(500):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if v @ n > 0:
            v = -v
        eta = float(rng.choice([1 / 1.5, 1.5, 1 / 1.33, 1.0]))
        cos_t = -v @ n
        if eta * np.sqrt(max(0.0, 1 - cos_t ** 2)) >= 0.999:
            continue  # total internal reflection: the caller reflects instead
        v32, n32 = v.astype(np.float32), n.astype(np.float32)
        out = np.zeros(3)
        a, b = d3(v32), d3(n32)
        ref.ref_refract(a.ctypes.data, b.ctypes.data, float(np.float32(eta)), out.ctypes.data)
        assert np.allclose(orc.refract(v32, n32, np.float32(eta)), out, rtol=1e-4, atol=1e-5)
        checked += 1
    assert checked > 300


def test_quad_hit_distance_matches_reference_plane(ref):
    # quad in the plane through Q with normal u x v; the reference normalises
    # the ray direction, so its hit point is o + t|d| d/|d| = o + t d
    rng = np.random.default_rng(2)
    for _ in range(300):
        Q = rng.uniform(-5, 5, 3)
        u = rng.normal(size=3)
        v = rng.normal(size=3)
        a, b = rng.uniform(0.05, 0.95, 2)
        target = Q + a * u + b * v
        o = target + rng.normal(size=3) * 3
        d = target - o
        quad = np.concatenate([Q, u, v]).astype(np.float32)
        ok, tab = orc.hit_quad(quad, 0, o.astype(np.float32), d.astype(np.float32))
        n = np.cross(quad[3:6].astype(np.float64), quad[6:9].astype(np.float64))
        hit = np.zeros(3)
        args = [d3(quad[:3]), d3(n), d3(o.astype(np.float32)), d3(d.astype(np.float32))]
        assert ref.ref_plane_hit(*[x.ctypes.data for x in args], hit.ctypes.data) == 1
        assert ok
        p = o.astype(np.float32).astype(np.float64) + float(tab[0]) * d.astype(np.float32).astype(np.float64)
        assert np.allclose(p, hit, rtol=1e-4, atol=1e-4)


def test_image_texture_matches_reference_bilinear(ref):
    rng = np.random.default_rng(3)
    sw, sh, s = 5, 4, 4
    src = rng.uniform(0, 1, (sh, sw, 3)).astype(np.float32)
    dw, dh = s * (sw - 1) + 1, s * (sh - 1) + 1
    corners = np.array([0, 0, dw - 1, 0, 0, dh - 1, dw - 1, dh - 1], np.int32)
    dst = np.zeros((dh, dw, 3))
    src64 = d3(src)
    ref.ref_paste(sw, sh, src64.ctypes.data, dw, dh, corners.ctypes.data, -1.0, dst.ctypes.data)
    sc = orc.OracleScene(SceneArrays(textures=[image(sw, sh, 0)], texels=src.reshape(-1, 3)))
    written = 0
    for y in range(dh):
        for x in range(dw):
            if dst[y, x, 0] < 0:
                continue  # pixel centre outside the pasted quad
            # the homography is the scaling by s: source point ((x+.5)/s, (y+.5)/s)
            sx, sy = (x + 0.5) / s, (y + 0.5) / s
            u = np.float32(sx / (sw - 1))
            v = np.float32(1.0 - sy / (sh - 1))
            got = sc.texture_value(0, u, v, [0, 0, 0])
            assert np.allclose(got, dst[y, x], atol=2e-6), (x, y, got, dst[y, x])
            written += 1
    assert written > 0.5 * dw * dh
