"""aurora_paste / aurora_paste_device (the homography warp of
are::Texture::paste, src/texture.cpp:85-360, as a HIP kernel) against the
oracle's restatement, bit for bit, through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import orc

from paste_cases import cases

pytestmark = pytest.mark.gpu
FILL = 0.25


def test_paste_matches_oracle(ctx):
    for name, src, dw, dh, corners in cases(seed=3, big=True):
        print("paste", name, src.shape, (dh, dw), flush=True)
        dst = np.random.default_rng(len(name)).uniform(0, 1, (dh, dw, 3)).astype(np.float32)
        ref, n = orc.paste(src, dst, corners)
        got = ctx.paste(src, dst, corners)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), name
        if src.size:
            ts = torch.from_numpy(src).cuda()
            td = torch.from_numpy(dst).cuda()
            ctx.paste_device(ts, td, corners)
            ctx.synchronize()
            assert np.array_equal(td.cpu().numpy().view(np.uint32), ref.view(np.uint32)), name


def test_paste_rejects_bad_arguments(ctx):
    src = np.zeros((4, 4, 3), np.float32)
    with pytest.raises(ValueError):
        ctx.paste(src, np.zeros((0, 4, 3), np.float32), [0, 0, 3, 0, 0, 3, 3, 3])
