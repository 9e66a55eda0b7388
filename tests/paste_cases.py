"""Seeded cases of are::Texture::paste (src/texture.cpp:85-360), shared by
the CPU and GPU paste tests: sources of float texels, destinations filled
with a constant, and corner sets LT, RT, LB, RB (x, y) covering the
identity, scalings, rotations, perspective quads, quads partly or wholly
outside the destination, mirrored (self-crossing) quads, a degenerate quad
and an empty source."""
import numpy as np


def cases(seed=0, n_random=24, big=False):
    rng = np.random.default_rng(seed)
    out = []

    def src_img(w, h):
        return rng.uniform(0.0, 1.0, (h, w, 3)).astype(np.float32)

    # identity: corners on the corner pixels of a same-size destination
    out.append(("identity", src_img(17, 11), 17, 11, [0, 0, 16, 0, 0, 10, 16, 10]))
    out.append(("scale2", src_img(8, 6), 40, 30, [2, 3, 31, 3, 2, 25, 31, 25]))
    out.append(("outside", src_img(8, 8), 20, 20, [30, 30, 40, 30, 30, 40, 40, 40]))
    out.append(("partly_outside", src_img(16, 16), 24, 24, [-10, -6, 20, -4, -8, 18, 30, 28]))
    out.append(("mirrored", src_img(9, 7), 32, 32, [25, 2, 3, 4, 27, 26, 2, 29]))
    out.append(("degenerate", src_img(5, 5), 16, 16, [1, 1, 8, 8, 4, 4, 12, 12]))
    out.append(("empty_source", np.zeros((0, 0, 3), np.float32), 16, 16, [0, 0, 10, 0, 0, 10, 10, 10]))
    out.append(("single_texel", src_img(1, 1), 16, 16, [2, 2, 12, 2, 2, 12, 12, 12]))
    for k in range(n_random):
        sw, sh = int(rng.integers(2, 64)), int(rng.integers(2, 64))
        dw, dh = int(rng.integers(8, 160)), int(rng.integers(8, 160))
        # a convex-ish quad: jittered corners of a box inside (or near) the destination
        cx, cy = rng.uniform(-0.2, 1.2) * dw, rng.uniform(-0.2, 1.2) * dh
        hw, hh = rng.uniform(3, dw), rng.uniform(3, dh)
        ang = rng.uniform(0, 2 * np.pi)
        base = np.array([[-hw, -hh], [hw, -hh], [-hw, hh], [hw, hh]]) / 2
        rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
        pts = base @ rot.T + np.array([cx, cy]) + rng.normal(0, 0.15, (4, 2)) * np.array([hw, hh])
        out.append((f"random{k}", src_img(sw, sh), dw, dh, list(np.round(pts).astype(int).reshape(-1))))
    if big:
        out.append(("big", src_img(512, 384), 2048, 1536, [100, 50, 1900, 200, 300, 1400, 1700, 1500]))
    return out
