"""relMSE and the reference-free error estimates on synthetic images with a
known noise level (closed forms)."""
import numpy as np
import torch

from aurora_amd.metric import plain_error_estimate, relmse, unbiased_relmse


def test_relmse_closed_form():
    r = np.full((4, 4, 3), 0.5)
    x = r + 0.1
    assert abs(relmse(x, r) - 0.01 / (0.25 + 0.01)) < 1e-12
    assert relmse(r, r) == 0.0


def test_unbiased_relmse_removes_reference_noise():
    g = torch.Generator().manual_seed(0)
    truth = torch.rand((256, 256, 3), generator=g, dtype=torch.float64) + 0.5
    sx, sr = 0.05, 0.08  # noise of the test image and of each reference half
    x = truth + sx * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    a = truth + sr * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    b = truth + sr * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    exact = relmse(x, truth)
    naive = relmse(x, 0.5 * (a + b))
    est = unbiased_relmse(x, a, b)
    assert naive > 1.5 * exact  # the noisy reference inflates the naive figure
    assert abs(est - exact) < 0.1 * exact


def test_plain_error_estimate_from_halves():
    g = torch.Generator().manual_seed(1)
    truth = torch.rand((256, 256, 3), generator=g, dtype=torch.float64) + 0.5
    s = 0.1  # noise of each half's mean
    a = truth + s * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    b = truth + s * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    exact = relmse(0.5 * (a + b), truth)
    est = plain_error_estimate(a, b)
    assert abs(est - exact) < 0.1 * exact


def halves_and_image(truth, sx, sr, g, heavy=False):
    """A test image with noise sx and two reference halves with noise sr
    (per pixel and channel, zero mean); heavy: a skewed, heavy-tailed noise
    (centred exponential cubed), as path tracing's rare bright paths give."""
    def noise(s):
        if not heavy:
            return s * torch.randn(truth.shape, generator=g, dtype=torch.float64)
        e = torch.empty(truth.shape, dtype=torch.float64).exponential_(generator=g) ** 3
        return s * (e - 6.0) / np.sqrt(720.0 - 36.0)  # E[E^3] = 6, Var[E^3] = 720 - 36
    return truth + noise(sx), truth + noise(sr), truth + noise(sr)


def test_unbiased_relmse_on_dark_pixels():
    """Truth far below the 0.1 of the denominator's 0.01 (the Cornell box's
    shadowed corners), where (x - r)^2 / (r^2 + 0.01) weighs noise most."""
    g = torch.Generator().manual_seed(2)
    truth = torch.rand((256, 256, 3), generator=g, dtype=torch.float64) * 0.03
    for sx, sr in ((0.01, 0.02), (0.003, 0.003), (0.02, 0.005)):
        x, a, b = halves_and_image(truth, sx, sr, g)
        exact = relmse(x, truth)
        est = unbiased_relmse(x, a, b)
        assert abs(est - exact) < 0.1 * exact, (sx, sr, est, exact)


def test_unbiased_relmse_with_heavy_tailed_noise():
    """Skewed, heavy-tailed noise on a mixed image (dark and bright halves):
    the correction still recovers the error against the truth."""
    g = torch.Generator().manual_seed(3)
    truth = torch.rand((512, 512, 3), generator=g, dtype=torch.float64)
    truth[:256] *= 0.02
    x, a, b = halves_and_image(truth, 0.01, 0.02, g, heavy=True)
    exact = relmse(x, truth)
    naive = relmse(x, 0.5 * (a + b))
    est = unbiased_relmse(x, a, b)
    assert naive > 1.5 * exact
    assert abs(est - exact) < 0.15 * exact, (est, exact)


def test_plain_error_estimate_on_dark_pixels():
    g = torch.Generator().manual_seed(4)
    truth = torch.rand((256, 256, 3), generator=g, dtype=torch.float64) * 0.03
    a = truth + 0.01 * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    b = truth + 0.01 * torch.randn(truth.shape, generator=g, dtype=torch.float64)
    exact = relmse(0.5 * (a + b), truth)
    est = plain_error_estimate(a, b)
    # the estimate's denominator uses the noisy mean, not the truth: within 20% here
    assert abs(est - exact) < 0.2 * exact, (est, exact)
