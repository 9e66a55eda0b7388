"""The are:: C++ host API: builds, links against libaurora, and flattens the
"Ray Tracing: The Next Week" Cornell box into exactly the records the Python
scene builder produces (so both feed the kernels identical scenes)."""
import os
import subprocess

import numpy as np
import pytest

from aurora_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "aurora-rendering-engine_amd")
EXE = os.path.join(PKG, "cornell_example")


@pytest.fixture(scope="module", autouse=True)
def built():
    from conftest import build_host_programs
    build_host_programs()


def test_library_and_example_built():
    assert os.path.exists(os.path.join(PKG, "libare.so"))
    assert os.path.exists(EXE)
    assert os.path.exists(os.path.join(PKG, "ppm_tool"))


def test_flattened_cornell_matches_python_builder(tmp_path):
    out = tmp_path / "quads.f32"
    subprocess.check_call([EXE, "--dump", str(out)])
    q = np.fromfile(out, np.float32).reshape(-1, 9)
    ref = scenes.cornell_box().quads
    assert q.shape == ref.shape == (18, 9)
    assert np.array_equal(q, ref)


@pytest.mark.gpu
def test_cpp_render_matches_python_path(tmp_path, ctx):
    from aurora_amd.core import RenderSettings
    from aurora_amd import abi
    from oracle import orc
    out = tmp_path / "img.f32"
    ppm = tmp_path / "img.ppm"
    r = subprocess.run([EXE, "64", "8", str(out), str(ppm)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    img = np.fromfile(out, np.float32).reshape(64, 64, 3)
    g = ctx.upload(scenes.cornell_box())
    g.build_bvh()
    rgb, _, _, _ = g.render(scenes.cornell_camera(64), RenderSettings(spp=8, max_depth=50))
    assert np.array_equal(img, rgb * np.float32(1.0 / 8))
    # camera::render_ppm: a P6 file with the transfer of experiments/rt.cpp:377-392
    body = orc.resolve_rgb8(rgb.reshape(-1), 8, abi.TRANSFER_GAMMA22).tobytes()
    assert ppm.read_bytes() == b"P6\n64 64\n255\n" + body
    # the ray query of the example hits the same primitive as the Python trace
    prim, t = g.trace(np.array([[278, 278, -800]], np.float32), np.array([[0, 0, 1]], np.float32), 0.001)
    assert f"hit prim {prim[0]} " in r.stdout


RT10 = os.path.join(PKG, "rt10_scene")


def test_patch_api_example_built():
    # are::patch (host/include/are/patch.hpp) and the example that builds
    # experiments/rt10.cpp's scene with it
    assert os.path.exists(RT10)


@pytest.mark.gpu
def test_patch_api_writes_the_reference_programs_file(tmp_path):
    import zlib

    out = tmp_path / "output_rt10.ppm"
    r = subprocess.run([RT10, str(out), "--stats"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    golden = zlib.decompress(open(os.path.join(ROOT, "tests", "golden", "rt10_cornell_900x650.ppm.zlib"), "rb").read())
    assert out.read_bytes() == golden
    assert "textures" in r.stdout
