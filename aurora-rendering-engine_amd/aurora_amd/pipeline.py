"""Render-to-error-budget pipelines: plain progressive rendering, and
denoised rendering with the plain render as fallback.

Both estimate their own error without a reference: an image rendered as two
independent halves (disjoint sample ranges, so independent Philox streams)
gives E[(a - b)^2] = 4 Var(mean) for a plain mean, and
relMSE(D(a), D(b)) / 4 is the same estimate for the denoised output (its
variance part; bias is not seen, which is why the plain render stays the
reference: `render_denoised` hands the image over to plain rendering whenever
the estimate exceeds the budget).
"""
import time

from .core import RenderSettings
from .metric import plain_error_estimate, relmse


def _zeros(cam, n=1):
    import torch
    return [torch.zeros((cam.image_height, cam.image_width, 3), dtype=torch.float32, device="cuda")
            for _ in range(n)]


def render_plain_to_budget(scene, cam, budget, chunk_spp=1024, max_spp=1 << 16, max_depth=50, seed=0):
    """Progressive plain rendering in chunks of 2 x chunk_spp samples until
    the two-half error estimate is within `budget`.  Returns dict(image,
    spp, estimate, render_s)."""
    a, b = _zeros(cam, 2)
    spp_half, t = 0, 0.0
    est = float("inf")
    while spp_half * 2 < max_spp:
        t0 = time.perf_counter()
        scene.render_device(cam, RenderSettings(spp=chunk_spp, sample_offset=spp_half, max_depth=max_depth,
                                                seed=seed), a)
        scene.render_device(cam, RenderSettings(spp=chunk_spp, sample_offset=(1 << 30) + spp_half,
                                                max_depth=max_depth, seed=seed), b)
        scene.ctx.synchronize()
        t += time.perf_counter() - t0
        spp_half += chunk_spp
        est = plain_error_estimate(a / spp_half, b / spp_half)
        if est <= budget:
            break
    return {"image": (a + b) / (2 * spp_half), "spp": 2 * spp_half, "estimate": est, "render_s": t}


def render_denoised(scene, cam, denoiser, spp, budget, max_depth=50, seed=0, fallback_max_spp=1 << 16):
    """Render spp samples (two halves, AOVs on), denoise, estimate the
    denoised error from the halves; fall back to plain rendering when the
    estimate exceeds `budget`."""
    ra, aa, na, rb, ab, nb = _zeros(cam, 6)
    half = spp // 2
    t0 = time.perf_counter()
    scene.render_device(cam, RenderSettings(spp=half, sample_offset=0, max_depth=max_depth, seed=seed, aov=True),
                        ra, aa, na)
    scene.render_device(cam, RenderSettings(spp=half, sample_offset=1 << 30, max_depth=max_depth, seed=seed,
                                            aov=True), rb, ab, nb)
    scene.ctx.synchronize()
    t_render = time.perf_counter() - t0
    t1 = time.perf_counter()
    out = denoiser.denoise(ra + rb, aa + ab, na + nb, 2 * half)
    t_denoise = time.perf_counter() - t1
    est = relmse(denoiser.denoise(ra, aa, na, half), denoiser.denoise(rb, ab, nb, half)) / 4.0
    res = {"image": out, "spp": 2 * half, "estimate": est, "render_s": t_render, "denoise_s": t_denoise,
           "fallback": False}
    if est > budget:
        plain = render_plain_to_budget(scene, cam, budget, max_spp=fallback_max_spp, max_depth=max_depth, seed=seed)
        res.update({"image": plain["image"], "fallback": True, "fallback_spp": plain["spp"],
                    "fallback_estimate": plain["estimate"], "fallback_render_s": plain["render_s"]})
    return res
