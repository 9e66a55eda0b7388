"""aurora_trace_closest_device_noalloc (the allocation-free form of
aurora_trace_closest_device, for callers that cannot allocate) against the
allocating entry point, through the C ABI: identical primitive ids, hit
distances and fetch counts on two calls in a row, the second with fewer
rays, at the maximum size and past it (rejected)."""
import numpy as np
import pytest
import torch

from aurora_amd import abi, scenes

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def query(g, o, d, noalloc, counts):
    n = o.shape[0]
    p = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    t = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
    c = torch.zeros((n, 3), dtype=torch.int32, device="cuda") if counts else None
    if noalloc:
        g.trace_device_noalloc(o, d, p, t, tmin=1e-4, counts=c)
    else:
        g.trace_device(o, d, p, t, tmin=1e-4, counts=c)
    g.ctx.synchronize()
    return p.cpu().numpy(), t.cpu().numpy(), (c.cpu().numpy() if counts else None)


@pytest.mark.parametrize("name", ["cornell", "stress"])
def test_noalloc_equals_allocating_query(ctx, name):
    arr = scenes.cornell_box() if name == "cornell" else scenes.stress_scene(50000, 50000, seed=12)
    g = ctx.upload(arr)
    st = g.build_bvh()
    lo = np.array(st["scene_lo"]) - 0.05
    hi = np.array(st["scene_hi"]) + 0.05
    for k, n in enumerate([abi.NOALLOC_MAX_RAYS, 1000, 1]):
        print("noalloc", name, "call", k, "rays", n, flush=True)
        org, dirs = scenes.random_rays(n, lo, hi, seed=40 + k)
        o, d = dev(org), dev(dirs)
        for counts in (False, True):
            ap, at, ac = query(g, o, d, False, counts)
            bp, bt, bc = query(g, o, d, True, counts)
            assert (ap >= 0).mean() > 0.1 or n < 10
            assert np.array_equal(ap, bp) and np.array_equal(at.view(np.uint32), bt.view(np.uint32))
            if counts:
                assert np.array_equal(ac, bc)
    g.close()


def test_noalloc_rejects_more_rays_than_its_scratch_holds(ctx):
    g = ctx.upload(scenes.cornell_box())
    g.build_bvh()
    n = abi.NOALLOC_MAX_RAYS + 1
    o = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    p = torch.zeros((n,), dtype=torch.int32, device="cuda")
    t = torch.zeros((n,), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="allocation-free maximum"):
        g.trace_device_noalloc(o, o, p, t)
    g.close()
