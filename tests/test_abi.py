"""The C-ABI library loads, exports every entry point include/aurora.h
declares, and the ctypes mirror matches the C struct layouts.  No compute
calls (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from aurora_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aurora.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aurora_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert {e[0] for e in abi.EXPORTS} == set(names)
    assert lib.aurora_abi_version() == abi.ABI_VERSION
    assert f"#define AURORA_ABI_VERSION {abi.ABI_VERSION}" in open(HEADER).read()


def test_struct_sizes_match_c():
    src = """
#include <stdio.h>
#include <stddef.h>
#include "aurora.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(aurora_material), sizeof(aurora_texture),
    sizeof(aurora_scene_desc), sizeof(aurora_bvh_node), sizeof(aurora_bvh_stats),
    sizeof(aurora_camera_desc), sizeof(aurora_render_desc), sizeof(aurora_render_stats));
  printf("%zu %zu %zu\\n", offsetof(aurora_texture, texel_offset), offsetof(aurora_scene_desc, texels),
    offsetof(aurora_render_stats, nodes_fetched));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "s.c")
        open(p, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), p, "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
    got = [int(x) for x in out]
    exp = [C.sizeof(abi.Material), C.sizeof(abi.Texture), C.sizeof(abi.SceneDesc), C.sizeof(abi.BvhNode),
           C.sizeof(abi.BvhStats), C.sizeof(abi.CameraDesc), C.sizeof(abi.RenderDesc), C.sizeof(abi.RenderStats),
           abi.Texture.texel_offset.offset, abi.SceneDesc.texels.offset, abi.RenderStats.nodes_fetched.offset]
    assert got == exp
    assert C.sizeof(abi.BvhNode) == 32


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        abi.load_library(str(tmp_path / "libaurora.so"))


def test_create_without_device_reports_error():
    # in a container without a GPU, aurora_create fails with a message, no crash
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = abi.load_library()
    h = C.c_void_p()
    rc = lib.aurora_create(0, C.byref(h))
    assert rc != abi.AURORA_OK
    assert len(lib.aurora_last_error(None)) > 0


def test_patch_descriptor_layouts_match_c():
    from aurora_amd import patch

    src = """
#include <stdio.h>
#include <stddef.h>
#include "aurora.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\\n", sizeof(aurora_patch_material), sizeof(aurora_patch_triangle),
    sizeof(aurora_patch_camera), sizeof(aurora_patch_config), sizeof(aurora_patch_stats));
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", offsetof(aurora_patch_material, albedo),
    offsetof(aurora_patch_triangle, uv), offsetof(aurora_patch_triangle, id),
    offsetof(aurora_patch_camera, viewport), offsetof(aurora_patch_camera, width),
    offsetof(aurora_patch_config, min_area_px), offsetof(aurora_patch_stats, plan_ms));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "s.c")
        open(p, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), p, "-o", exe])
        got = [int(x) for x in subprocess.check_output([exe]).decode().split()]
    exp = [patch.MATERIAL.itemsize, patch.TRIANGLE.itemsize, patch.CAMERA.itemsize, patch.CONFIG.itemsize,
           C.sizeof(patch.Stats), patch.MATERIAL.fields["albedo"][1], patch.TRIANGLE.fields["uv"][1],
           patch.TRIANGLE.fields["id"][1], patch.CAMERA.fields["viewport"][1], patch.CAMERA.fields["width"][1],
           patch.CONFIG.fields["min_area_px"][1], patch.Stats.plan_ms.offset]
    assert got == exp
