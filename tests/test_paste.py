"""are::Texture::paste (src/texture.cpp:85-360): the oracle's restatement
(oracle/orc_paste.cpp) against the reference's own routine built from its
sources (oracle/_ref), bit for bit (the reference's double result rounded
to float), and against closed forms; on the GPU (aurora_paste and
aurora_paste_device) against the oracle, bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import orc

from paste_cases import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libareref.so")
FILL = 0.25


def ref_paste(src, dw, dh, corners):
    L = C.CDLL(REF)
    vp = C.c_void_p
    L.ref_paste.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, C.c_double, vp]
    s = np.ascontiguousarray(src, np.float64)
    c = np.ascontiguousarray(corners, np.int32)
    out = np.zeros((dh, dw, 3), np.float64)
    sh, sw = src.shape[:2]
    L.ref_paste(sw, sh, s.ctypes.data if s.size else None, dw, dh, c.ctypes.data, FILL, out.ctypes.data)
    return out


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (no reference sources)")
@pytest.mark.parametrize("name,src,dw,dh,corners", cases(), ids=[c[0] for c in cases()])
def test_oracle_paste_equals_reference(name, src, dw, dh, corners):
    dst = np.full((dh, dw, 3), FILL, np.float32)
    ours, n = orc.paste(src, dst, corners)
    if src.size == 0:
        # the reference cannot build an empty texture (its constructor
        # throws, src/texture.cpp:52-55); paste returns at once for one
        # (src/texture.cpp:93-95)
        assert n == 0 and np.array_equal(ours, dst)
        return
    ref = ref_paste(src, dw, dh, corners)
    assert np.array_equal(ours.view(np.uint32), ref.astype(np.float32).view(np.uint32))
    painted = (ours != np.float32(FILL)).any(-1).sum()
    assert painted <= n
    if name in ("identity", "scale2", "partly_outside", "random0"):
        assert n > 0
    if name in ("outside", "degenerate", "empty_source"):
        assert n == 0


def test_identity_paste_is_the_two_by_two_average():
    """Corners on the corner pixels of a same-size image: pixel (x, y) maps to
    source (x + 1/2, y + 1/2), the mean of the 2x2 texels around it; pixels
    whose centre falls outside the quad (last row and column) keep the fill."""
    src = np.random.default_rng(1).uniform(0, 1, (11, 17, 3)).astype(np.float32)
    dst = np.full((11, 17, 3), FILL, np.float32)
    out, n = orc.paste(src, dst, [0, 0, 16, 0, 0, 10, 16, 10])
    s = src.astype(np.float64)
    avg = (((s[:-1, :-1] * 0.25 + s[:-1, 1:] * 0.25) + s[1:, :-1] * 0.25) + s[1:, 1:] * 0.25).astype(np.float32)
    assert n == 16 * 10
    assert np.array_equal(out[:-1, :-1], avg)
    assert np.all(out[-1] == np.float32(FILL)) and np.all(out[:, -1] == np.float32(FILL))


def test_translation_paste_copies_texels():
    """An integer translation by (dx, dy) with the quad one pixel larger than
    the source corners: a pixel centre maps to a texel's midpoint, so every
    painted value is the mean of two or four texels at equal weights."""
    src = np.random.default_rng(2).uniform(0, 1, (6, 9, 3)).astype(np.float32)
    dst = np.full((20, 20, 3), FILL, np.float32)
    out, n = orc.paste(src, dst, [4, 3, 12, 3, 4, 8, 12, 8])
    s = src.astype(np.float64)
    exp = (((s[:-1, :-1] * 0.25 + s[:-1, 1:] * 0.25) + s[1:, :-1] * 0.25) + s[1:, 1:] * 0.25).astype(np.float32)
    assert n == 8 * 5
    assert np.array_equal(out[3:8, 4:12], exp)
