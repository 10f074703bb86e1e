"""The pass-variance estimate of GuidedProgressive (renderer_amd/progressive.py),
on the CPU: it is the unbiased sample variance of a pass, relative to
mean^2 + eps and averaged over the pixels."""
import numpy as np
import pytest
import torch

from renderer_amd.progressive import GuidedProgressive, relative_pass_variance


def test_relative_pass_variance_matches_definition():
    rng = np.random.default_rng(1)
    passes = [torch.from_numpy(rng.normal(2.0, 0.5, (6, 5, 3)).astype(np.float32)) for _ in range(5)]
    a = np.stack([p.numpy().astype(np.float64) for p in passes])
    m = a.mean(0)
    want = (a.var(0, ddof=1) / (m * m + 1e-2)).mean()
    assert relative_pass_variance(passes) == pytest.approx(want, rel=1e-12)


def test_relative_pass_variance_of_known_noise():
    # passes of constant mean 3 with noise of variance 0.25: relative variance 0.25 / (9 + eps)
    rng = np.random.default_rng(2)
    passes = [torch.from_numpy(3.0 + 0.5 * rng.standard_normal((64, 64, 3))) for _ in range(8)]
    assert relative_pass_variance(passes) == pytest.approx(0.25 / 9.01, rel=0.05)


def test_rejects_bad_schedules():
    with pytest.raises(ValueError):
        relative_pass_variance([torch.zeros(2)])
    with pytest.raises(ValueError):
        GuidedProgressive(None, None, probe_passes=1)
    with pytest.raises(ValueError):
        GuidedProgressive(None, None, pass_spp=0)


def test_smooth_image_is_a_box_average():
    from renderer_amd.progressive import smooth_image

    img = torch.zeros((9, 9, 3), dtype=torch.float64)
    img[4, 4] = torch.tensor([49.0, 98.0, 0.0])
    out = smooth_image(img)
    assert torch.allclose(out[4, 4], torch.tensor([1.0, 2.0, 0.0], dtype=torch.float64))
    # at (1, 1) the window holds 5 x 5 pixels of the image
    assert torch.allclose(out[1, 1], torch.tensor([49.0 / 25, 98.0 / 25, 0.0], dtype=torch.float64))
    assert float(out[0, 0, 0]) == 0.0  # outside the 7 x 7 window
    flat = torch.full((5, 6, 3), 2.5, dtype=torch.float64)
    assert torch.allclose(smooth_image(flat), flat)  # edges average the pixels inside only
