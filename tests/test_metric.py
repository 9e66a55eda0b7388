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
