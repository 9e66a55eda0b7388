"""Image error metric and the plain-versus-denoised error decisions.

relMSE(x, r) = mean over pixels and channels of (x - r)^2 / (r^2 + 0.01).
With r itself a Monte-Carlo estimate, E[(x - r)^2] = MSE(x) + Var(r); the
reference's own variance is estimated from two independent halves
(E[(a - b)^2] = 4 Var of their mean) and subtracted (`unbiased_relmse`).
"""
import numpy as np


def relmse(x, ref):
    import torch
    if isinstance(x, torch.Tensor):
        x64 = x.to(torch.float64)
        r64 = ref.to(torch.float64)
        return float(((x64 - r64) ** 2 / (r64 ** 2 + 0.01)).mean())
    x = np.asarray(x, np.float64)
    r = np.asarray(ref, np.float64)
    return float(np.mean((x - r) ** 2 / (r ** 2 + 0.01)))


def unbiased_relmse(x, ref_a, ref_b):
    """relMSE of x against the truth, from a reference given as two
    independent halves (means over equal sample counts): the reference
    variance term (a - b)^2 / 4 is removed per pixel."""
    import torch
    a = ref_a.to(torch.float64)
    b = ref_b.to(torch.float64)
    r = 0.5 * (a + b)
    x64 = x.to(torch.float64)
    num = (x64 - r) ** 2 - 0.25 * (a - b) ** 2
    return float((num / (r ** 2 + 0.01)).mean())


def plain_error_estimate(half_a, half_b):
    """Estimated relMSE of the mean of two independent halves (means)."""
    import torch
    a = half_a.to(torch.float64)
    b = half_b.to(torch.float64)
    m = 0.5 * (a + b)
    return float((0.25 * (a - b) ** 2 / (m ** 2 + 0.01)).mean())
