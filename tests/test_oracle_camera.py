"""The oracle's camera (oracle/orc_render.cpp camera_setup / camera_ray)
against the closed form of experiments/rt.cpp:339-343,364-366: vfov is the
vertical angle, the aspect is W/H, row 0 is the top and pixel centres sit at
+1/2.  Planted mistakes in a frame (a half-pixel shift, a flipped v axis, the
field of view taken as horizontal or as a half angle) are caught by the same
check, so the check pins those choices rather than restating them."""
import numpy as np
import pytest

from aurora_amd.core import Camera
from oracle import orc
from aurora_amd import scenes

from camera_closed_form import (closed_form_dirs, frame_dirs, frame_error_px, image_position, random_cameras,
                                basis, sample_pixels)

TOL_PX = 0.05  # the frame is rounded to float: a few thousandths of a pixel


def u01(x):
    return (np.asarray(x, np.uint64) >> np.uint64(8)).astype(np.float64) / 16777216.0


def test_cornell_frame_matches_closed_form():
    for w, h in [(600, 600), (1024, 1024), (1200, 675), (7, 3), (1, 1)]:
        cam = scenes.cornell_camera(w, h)
        assert frame_error_px(orc.camera_frame(cam), cam) < TOL_PX


def test_random_frames_match_closed_form():
    for cam in random_cameras(60, seed=1):
        assert frame_error_px(orc.camera_frame(cam), cam) < TOL_PX, cam


def test_closed_form_directions_agree():
    # the two closed-form helpers invert each other and put row 0 at the top
    cam = Camera(64, 32, 50.0, (0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
    i, j = sample_pixels(cam)
    x, y = image_position(cam, closed_form_dirs(cam, i + 0.5, j + 0.5))
    assert np.allclose(x, i + 0.5, atol=1e-9) and np.allclose(y, j + 0.5, atol=1e-9)
    d = closed_form_dirs(cam, [32.0], [0.0])[0]
    assert d[1] > 0  # top edge looks up (vup = +y)
    # the top edge of the image is vfov/2 above the axis
    assert np.isclose(np.degrees(np.arctan2(d[1], -d[2])), 25.0)
    # the right edge is atan(aspect * tan(vfov/2)) to the right
    d = closed_form_dirs(cam, [64.0], [16.0])[0]
    assert np.isclose(np.arctan2(d[0], -d[2]), np.arctan(2.0 * np.tan(np.radians(25.0))))


def mutants(cam):
    """Frames with planted mistakes, built from the oracle's frame."""
    f = orc.camera_frame(cam).astype(np.float64)
    c, p00, du, dv = f[0], f[1], f[2], f[3]
    W, H = cam.image_width, cam.image_height
    out = {
        "half_pixel_shift": np.stack([c, p00 - 0.5 * (du + dv), du, dv]),
        "flipped_v": np.stack([c, p00 + (H - 1) * dv, du, -dv]),
        "flipped_u": np.stack([c, p00 + (W - 1) * du, -du, dv]),
    }
    a = W / H
    hv = 2.0 * np.degrees(np.arctan(np.tan(np.radians(cam.vfov) / 2.0) / a))
    out["fov_horizontal"] = orc.camera_frame(Camera(W, H, hv, cam.lookfrom, cam.lookat, cam.vup, 0.0,
                                                    cam.focus_dist))[:4]
    out["fov_half_angle"] = orc.camera_frame(Camera(W, H, 2.0 * cam.vfov if cam.vfov < 90 else cam.vfov / 2.0,
                                                    cam.lookfrom, cam.lookat, cam.vup, 0.0, cam.focus_dist))[:4]
    return out


@pytest.mark.parametrize("which", ["half_pixel_shift", "flipped_v", "flipped_u", "fov_horizontal", "fov_half_angle"])
def test_planted_frame_mistakes_are_caught(which):
    cams = [scenes.cornell_camera(600, 400), Camera(640, 360, 30.0, (1.0, 2.0, 3.0), (0.0, 0.5, -4.0))]
    for cam in cams:
        assert frame_error_px(orc.camera_frame(cam), cam) < TOL_PX
        assert frame_error_px(mutants(cam)[which], cam) > 0.4, which


def test_jittered_rays_land_in_their_pixel():
    """Each sample's ray crosses the image at pixel + (1/2 + jitter), the
    jitter being U[-1/2, 1/2)^2 from Philox(pixel, sample, 0, 0)."""
    for cam in [scenes.cornell_camera(96, 64)] + random_cameras(8, seed=3):
        W, H = cam.image_width, cam.image_height
        rng = np.random.default_rng(7)
        n = 400
        pix = rng.integers(0, W * H, n).astype(np.uint32)
        smp = rng.integers(0, 1 << 20, n).astype(np.uint32)
        seed = 5
        o, d = orc.camera_rays(cam, pix, smp, seed)
        assert np.allclose(o, np.float32(cam.lookfrom))
        assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=1), 1.0, atol=1e-6)
        x, y = image_position(cam, d)
        q = orc.philox_batch(np.stack([pix, smp, 0 * pix, 0 * pix], 1), seed, 0)
        ex = (pix % W) + u01(q[:, 0])
        ey = (pix // W) + u01(q[:, 1])
        assert np.max(np.abs(x - ex)) < TOL_PX and np.max(np.abs(y - ey)) < TOL_PX


def test_defocus_disk_radius_and_focus_plane():
    """With depth of field the origin lies on the lens disk of radius
    focus_dist * tan(defocus_angle / 2) across the view axis, and every ray
    of a sample passes through that sample's point on the focus plane."""
    for cam in random_cameras(8, seed=4, defocus=True):
        W, H = cam.image_width, cam.image_height
        f = orc.camera_frame(cam).astype(np.float64)
        lf, fwd, right, up, scale, aspect = basis(cam)
        rad = cam.focus_dist * np.tan(np.radians(cam.defocus_angle) / 2.0)
        assert np.isclose(np.linalg.norm(f[4]), rad, rtol=1e-6) and np.isclose(np.linalg.norm(f[5]), rad, rtol=1e-6)
        assert abs(f[4] @ right) > 0.999999 * rad and abs(f[5] @ up) > 0.999999 * rad
        rng = np.random.default_rng(11)
        n = 300
        pix = rng.integers(0, W * H, n).astype(np.uint32)
        smp = rng.integers(0, 1 << 20, n).astype(np.uint32)
        o, d = orc.camera_rays(cam, pix, smp, 0)
        off = o.astype(np.float64) - lf
        assert np.all(np.linalg.norm(off, axis=1) <= rad * (1 + 1e-5) + 1e-5)
        assert np.all(np.abs(off @ fwd) <= 1e-4 * (1 + rad))
        assert np.linalg.norm(off, axis=1).max() > 0.3 * rad  # the disk is used
        # point on the focus plane
        t = (cam.focus_dist - off @ fwd) / (d.astype(np.float64) @ fwd)
        P = o + t[:, None] * d
        x, y = image_position(cam, P - lf)
        q = orc.philox_batch(np.stack([pix, smp, 0 * pix, 0 * pix], 1), 0, 0)
        ex = (pix % W) + u01(q[:, 0])
        ey = (pix // W) + u01(q[:, 1])
        assert np.max(np.abs(x - ex)) < TOL_PX and np.max(np.abs(y - ey)) < TOL_PX


def test_camera_errors():
    bad = [Camera(0, 4, 40.0, (0, 0, 0), (0, 0, -1)),
           Camera(4, 4, 0.0, (0, 0, 0), (0, 0, -1)),
           Camera(4, 4, 180.0, (0, 0, 0), (0, 0, -1)),
           Camera(4, 4, 40.0, (1, 1, 1), (1, 1, 1)),
           Camera(4, 4, 40.0, (0, 0, 0), (0, 1, 0), (0, 1, 0)),
           Camera(4, 4, 40.0, (0, 0, 0), (0, 0, -1), focus_dist=0.0)]
    for cam in bad:
        with pytest.raises(ValueError):
            orc.camera_rays(cam, [0], [0])
