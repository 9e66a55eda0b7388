"""Closed forms of the camera, shared by the CPU and GPU camera tests.

The pixel-centre direction is the one of experiments/rt.cpp:339-343 (basis:
forward = normalize(target - position), right = normalize(forward x up),
up' = right x forward) and :364-366 (pixel (x, y) -> forward + fx*right +
fy*up' with fx = (2(x + 1/2)/W - 1) * aspect * tan(fov/2) and
fy = (1 - 2(y + 1/2)/H) * tan(fov/2)), i.e. fov is the vertical angle, the
aspect is W/H and row 0 is the top.  Everything here is float64 numpy and
independent of the oracle's and the library's own frame code.
"""
import numpy as np


def basis(cam):
    lf = np.asarray(cam.lookfrom, np.float64)
    la = np.asarray(cam.lookat, np.float64)
    vup = np.asarray(cam.vup, np.float64)
    fwd = la - lf
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, vup)
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    scale = np.tan(np.radians(cam.vfov) / 2.0)
    aspect = cam.image_width / cam.image_height
    return lf, fwd, right, up, scale, aspect


def closed_form_dirs(cam, x, y):
    """Unit directions through continuous image positions (x, y) in pixels
    (pixel i spans [i, i+1]; its centre is i + 1/2)."""
    lf, fwd, right, up, scale, aspect = basis(cam)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    fx = (2.0 * x / cam.image_width - 1.0) * aspect * scale
    fy = (1.0 - 2.0 * y / cam.image_height) * scale
    d = fwd[None] + fx[:, None] * right[None] + fy[:, None] * up[None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def image_position(cam, dirs):
    """Inverse of closed_form_dirs: continuous pixel position (x, y) where a
    ray from the centre of projection with direction d crosses the image."""
    lf, fwd, right, up, scale, aspect = basis(cam)
    d = np.asarray(dirs, np.float64)
    f = d @ fwd
    fx = (d @ right) / f
    fy = (d @ up) / f
    x = (fx / (aspect * scale) + 1.0) * cam.image_width / 2.0
    y = (1.0 - fy / scale) * cam.image_height / 2.0
    return x, y


def frame_dirs(frame, i, j):
    """Pixel-centre directions implied by a frame (centre, pixel00, delta_u,
    delta_v, ...) as the oracle and the library lay it out."""
    c, p00, du, dv = (np.asarray(v, np.float64) for v in frame[:4])
    i = np.asarray(i, np.float64)
    j = np.asarray(j, np.float64)
    p = p00[None] + i[:, None] * du[None] + j[:, None] * dv[None]
    d = p - c[None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def sample_pixels(cam, n=256, seed=0):
    """The four corners, the centre and n random pixels (column, row)."""
    W, H = cam.image_width, cam.image_height
    rng = np.random.default_rng(seed)
    i = np.concatenate([[0, W - 1, 0, W - 1, W // 2], rng.integers(0, W, n)])
    j = np.concatenate([[0, 0, H - 1, H - 1, H // 2], rng.integers(0, H, n)])
    return i, j


def frame_error_px(frame, cam, n=256, seed=0):
    """Largest distance, in pixels, between the frame's pixel centres and the
    closed form's i + 1/2, j + 1/2."""
    i, j = sample_pixels(cam, n, seed)
    x, y = image_position(cam, frame_dirs(frame, i, j))
    return float(max(np.max(np.abs(x - (i + 0.5))), np.max(np.abs(y - (j + 0.5)))))


def random_cameras(n, seed=0, defocus=False):
    """Seeded cameras of varied pose, vertical field of view, aspect and size."""
    from aurora_amd.core import Camera
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        lf = rng.uniform(-20, 20, 3)
        la = rng.uniform(-20, 20, 3)
        vup = rng.normal(size=3)
        fwd = (la - lf) / np.linalg.norm(la - lf)
        if np.linalg.norm(la - lf) < 1.0 or np.linalg.norm(np.cross(fwd, vup / np.linalg.norm(vup))) < 0.2:
            continue
        W = int(rng.integers(1, 1025))
        H = int(rng.integers(1, 1025))
        vfov = float(rng.uniform(5.0, 150.0))
        dfa = float(rng.uniform(0.5, 10.0)) if defocus else 0.0
        fd = float(rng.uniform(1.0, 40.0))
        out.append(Camera(W, H, vfov, tuple(map(float, lf)), tuple(map(float, la)), tuple(map(float, vup)), dfa, fd))
    return out
