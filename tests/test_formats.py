"""FP32 image dumps (PFM) and the accumulation checkpoint of a progressive
render: the Python module (aurora_amd/io.py) and the C++ host API
(host/src/aurora_io.cpp, driven through host/examples/progressive.cpp)
write the same bytes, read each other's files, follow the published PFM
layout (colour "PF", negative scale = little-endian, bottom row first) and
report broken files with the reference loader's kind of message
(src/texture.cpp:9-50).  On the GPU: stop, save, load and resume gives
the sums of an uninterrupted render bit for bit, in Python and in C++, and
across the two."""
import os
import struct
import subprocess

import numpy as np
import pytest

from aurora_amd import io, scenes
from aurora_amd.core import RenderSettings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "aurora-rendering-engine_amd")
TOOL = os.path.join(PKG, "progressive")


@pytest.fixture(scope="module", autouse=True)
def built():
    from conftest import build_host_programs
    build_host_programs()


def tool(*args):
    r = subprocess.run([TOOL, *map(str, args)], capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout.strip()


def image(h, w, seed=0):
    return np.random.default_rng(seed).normal(size=(h, w, 3)).astype(np.float32)


def test_pfm_layout_closed_form(tmp_path):
    img = image(2, 3)
    p = tmp_path / "a.pfm"
    io.write_pfm(p, img)
    body = b"".join(struct.pack("<3f", *img[y, x]) for y in (1, 0) for x in range(3))
    assert p.read_bytes() == b"PF\n3 2\n-1.0\n" + body
    assert np.array_equal(io.read_pfm(p), img)


def test_pfm_python_and_cpp_agree(tmp_path):
    for h, w in [(1, 1), (5, 7), (64, 33)]:
        img = image(h, w, h + w)
        p = tmp_path / f"py_{h}_{w}.pfm"
        io.write_pfm(p, img)
        raw, again = tmp_path / "raw.f32", tmp_path / f"cpp_{h}_{w}.pfm"
        rc, out = tool("pfm", p, raw, again)
        assert rc == 0 and out == f"ok {w} {h}"
        assert np.array_equal(np.fromfile(raw, np.float32).reshape(h, w, 3), img)
        assert again.read_bytes() == p.read_bytes()


def test_pfm_big_endian_is_read(tmp_path):
    img = image(3, 4, 9)
    p = tmp_path / "be.pfm"
    p.write_bytes(b"PF\n4 3\n1.0\n" + np.ascontiguousarray(img[::-1]).astype(">f4").tobytes())
    assert np.array_equal(io.read_pfm(p), img)
    rc, _ = tool("pfm", p, tmp_path / "raw.f32", tmp_path / "le.pfm")
    assert rc == 0
    assert np.array_equal(np.fromfile(tmp_path / "raw.f32", np.float32).reshape(3, 4, 3), img)


@pytest.mark.parametrize("case,msg", [("missing", "Failed to open image file"),
                                      ("ppm", "Unsupported PFM format in file"),
                                      ("header", "Invalid PFM header in file"),
                                      ("truncated", "Unexpected end of file while reading pixel data in file")])
def test_pfm_errors(tmp_path, case, msg):
    p = tmp_path / f"{case}.pfm"
    if case == "ppm":
        p.write_bytes(b"P6\n1 1\n255\n" + bytes(3))
    elif case == "header":
        p.write_bytes(b"PF\n4 x\n-1.0\n")
    elif case == "truncated":
        p.write_bytes(b"PF\n2 2\n-1.0\n" + bytes(20))
    with pytest.raises(RuntimeError, match=msg):
        io.read_pfm(p)
    rc, out = tool("pfm", p, tmp_path / "x.f32", tmp_path / "x.pfm")
    assert rc == 2 and out == f"error: {msg}: {p}"


def accumulation(aov, spp=5, seed=0):
    cam = scenes.cornell_camera(8, 6)
    rng = np.random.default_rng(seed)
    sums = [rng.uniform(0, spp, (6, 8, 3)).astype(np.float32) for _ in range(3)]
    rs = RenderSettings(spp=0, max_depth=50, seed=3, sample_offset=2, background=(0.1, 0.2, 0.3), aov=aov)
    return io.Accumulation(cam, rs, spp, sums[0], sums[1] if aov else None, sums[2] if aov else None)


@pytest.mark.parametrize("aov", [False, True])
def test_checkpoint_round_trip_python_and_cpp(tmp_path, aov):
    acc = accumulation(aov)
    p = tmp_path / "a.ckpt"
    io.save_checkpoint(p, acc)
    data = p.read_bytes()
    assert data[:8] == b"AURACC01" and len(data) == 128 + 4 * 8 * 6 * 3 * (3 if aov else 1)
    back = io.load_checkpoint(p)
    assert back.spp == 5 and back.settings.sample_offset == 2 and back.settings.seed == 3
    assert back.camera.image_width == 8 and back.camera.image_height == 6
    assert np.array_equal(back.rgb_sum, acc.rgb_sum)
    if aov:
        assert np.array_equal(back.albedo_sum, acc.albedo_sum) and np.array_equal(back.normal_sum, acc.normal_sum)
    q = tmp_path / "b.ckpt"
    rc, out = tool("copy", p, q)
    assert rc == 0 and out == "ok 5 samples"
    assert q.read_bytes() == data


@pytest.mark.parametrize("case,msg", [("magic", "Not an accumulation checkpoint"),
                                      ("truncated", "Unexpected end of file while reading sums in file"),
                                      ("trailing", "Trailing data in checkpoint file"),
                                      ("flipped", "Checksum mismatch in file")])
def test_checkpoint_errors(tmp_path, case, msg):
    p = tmp_path / "a.ckpt"
    io.save_checkpoint(p, accumulation(True))
    data = bytearray(p.read_bytes())
    if case == "magic":
        data[0:8] = b"AURACC99"
    elif case == "truncated":
        data = data[:-10]
    elif case == "trailing":
        data += b"\0"
    else:
        data[200] ^= 0x10
    p.write_bytes(bytes(data))
    with pytest.raises(RuntimeError, match=msg):
        io.load_checkpoint(p)
    rc, out = tool("copy", p, tmp_path / "b.ckpt")
    assert rc == 2 and out.startswith(f"error: {msg}")


@pytest.mark.gpu
def test_resumed_render_equals_uninterrupted(ctx, tmp_path):
    """Cornell box at 64^2: 5 samples, checkpoint, 7 more, with AOVs, in
    Python; the same in C++ (a fresh process per step); and a Python
    checkpoint resumed by C++: every sum equals one 12-sample render."""
    print("checkpoint resume", flush=True)
    g = ctx.upload(scenes.cornell_box())
    g.build_bvh()
    cam = scenes.cornell_camera(64)
    full = g.render(cam, RenderSettings(spp=12, max_depth=50, aov=True))
    z = np.zeros((64, 64, 3), np.float32)
    acc = io.Accumulation(cam, RenderSettings(spp=0, max_depth=50), 0, z.copy(), z.copy(), z.copy())
    io.accumulate(g, acc, 5)
    p = tmp_path / "py.ckpt"
    io.save_checkpoint(p, acc)
    back = io.accumulate(g, io.load_checkpoint(p), 7)
    assert back.spp == 12
    for a, b in zip((back.rgb_sum, back.albedo_sum, back.normal_sum), full[:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # C++: start (aov) and resume in separate processes
    c1, c2, pfm = tmp_path / "c1.ckpt", tmp_path / "c2.ckpt", tmp_path / "mean.pfm"
    rc, out = tool("start", 64, 5, c1, "aov")
    assert rc == 0, out
    rc, out = tool("resume", c1, 7, c2, pfm)
    assert rc == 0 and out == "ok 12 samples", out
    cpp = io.load_checkpoint(c2)
    for a, b in zip((cpp.rgb_sum, cpp.albedo_sum, cpp.normal_sum), full[:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(io.read_pfm(pfm).view(np.uint32), (full[0] / np.float32(12)).view(np.uint32))
    # the Python checkpoint after 5 samples, resumed by C++
    rc, out = tool("resume", p, 7, tmp_path / "x.ckpt", tmp_path / "x.pfm")
    assert rc == 0 and out == "ok 12 samples", out
    x = io.load_checkpoint(tmp_path / "x.ckpt")
    assert np.array_equal(x.rgb_sum.view(np.uint32), full[0].view(np.uint32))
    g.close()
