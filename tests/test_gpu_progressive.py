"""GuidedProgressive on the GPU: the guided renderer with plain path tracing
as its fallback and error judge.  On the box scene it trains, probes both
estimators, picks one, fills the budget, and its image (the mean of all its
passes) equals a high-sample plain render within 4 standard errors."""
import numpy as np
import pytest

from renderer_amd import MODE_UNBIASED, scenes
from renderer_amd.progressive import GuidedProgressive

pytestmark = pytest.mark.gpu


def test_progressive_guided_with_fallback_is_unbiased(gpu_ctx):
    import torch

    sc = scenes.box_scene(48, 48)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.guide_init(3)
    seen = []
    gp = GuidedProgressive(gpu_ctx, sc.camera, pass_spp=16, train_passes=4, probe_passes=3, lr=1e-3, max_depth=8)
    img, info = gp.run(1.0, seed0=100, on_pass=lambda t, spp, im: seen.append((t, spp)))
    print(f"progressive: {info}", flush=True)
    assert info["choice"] in ("guided", "plain")
    assert info["passes"] >= 4 + 2 * 3 and info["seconds"] >= 1.0
    assert set(info["probe"]) == {"guided_var", "plain_var", "guided_pass_s", "plain_pass_s"}
    assert all(v > 0 for v in info["probe"].values())
    assert len(seen) == info["passes"] and seen[-1][1] == info["spp"]
    assert [s for _, s in seen] == [16 * (k + 1) for k in range(len(seen))]
    # reference: plain PT, 64 passes of 16 spp (other seeds)
    ref = np.zeros((48, 48, 3))
    ref2 = np.zeros((48, 48, 3))
    K = 64
    for k in range(K):
        f = gpu_ctx.render([(sc.camera, 4, 4)], seed=1000 + k, mode=MODE_UNBIASED, max_depth=8, rr_start=3,
                           rr_prob=0.5)[0].astype(np.float64)
        ref += f
        ref2 += f * f
    ref /= K
    var_pass = np.maximum(ref2 / K - ref * ref, 0) * K / (K - 1)
    got = img.cpu().numpy()
    # channel means over the image: standard error from the plain per-pass
    # variance, doubled for the progressive image (its training passes may be
    # noisier than plain ones)
    se = np.sqrt(var_pass.sum((0, 1)) / (48 * 48) ** 2 * (2.0 / info["passes"] + 1.0 / K))
    diff = got.mean((0, 1)) - ref.mean((0, 1))
    print(f"means {got.mean((0, 1))} vs {ref.mean((0, 1))}, se {se}", flush=True)
    assert np.all(np.abs(diff) < 4 * se + 1e-6)
