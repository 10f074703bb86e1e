"""The finishing launch (rt_launch_config.finish_depth: the remaining
bounces of a pass traced, shaded and shadow-tested per path in one launch)
against the per-bounce launches: every path changes its state in the same
order either way, so the films agree bit for bit; the ray counts agree, and
those the finishing launch reports are the late bounces' share."""
import numpy as np
import pytest

from renderer_amd import MODE_COMPAT, MODE_UNBIASED, scenes
from tests.test_gpu_csg import _mixed_scene

pytestmark = pytest.mark.gpu


def _render(ctx, view, finish, **params):
    ctx.set_launch_config(finish_depth=finish)
    try:
        img = ctx.render([view], **params)[0]
        return img, ctx.stats()
    finally:
        ctx.set_launch_config(finish_depth=4)


def _scene(name):
    if name == "box":
        return scenes.box_scene(64, 48)
    if name == "mesh":
        return scenes.mesh_scene(48, 48, n=24, seed=2)
    return _mixed_scene(48)


@pytest.mark.parametrize("scene", ["box", "mesh", "objects"])
@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_finish_same_film(gpu_ctx, scene, mode):
    sc = _scene(scene)
    gpu_ctx.upload(sc.flatten())
    view = (sc.camera, 4, 4)
    params = dict(mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=5)
    ref, rs = _render(gpu_ctx, view, -1, **params)
    assert rs["finish_rays"] == 0 and rs["finish_shadow_rays"] == 0
    for finish in (1, 2, 3, 6):
        img, st = _render(gpu_ctx, view, finish, **params)
        print(f"{scene} mode {mode} finish {finish}: finish rays {st['finish_rays']} / {st['extension_rays']}, "
              f"shadow {st['finish_shadow_rays']} / {st['shadow_rays']}", flush=True)
        assert np.array_equal(img, ref), f"finish_depth {finish}: films differ"
        for k in ("camera_rays", "extension_rays", "shadow_rays"):
            assert st[k] == rs[k], k
        assert 0 < st["finish_rays"] <= st["extension_rays"]
        assert st["finish_shadow_rays"] <= st["shadow_rays"]
        assert st["extend_launches"] == finish * st["passes"]


def test_finish_rays_are_the_late_bounces(gpu_ctx):
    """The rays the finishing launch reports are those the per-bounce
    launches trace from that bounce on: the counts of a render cut at
    max_depth = finish_depth subtracted from the full render's."""
    sc = scenes.mesh_scene(48, 48, n=24, seed=2)
    gpu_ctx.upload(sc.flatten())
    view = (sc.camera, 4, 4)
    params = dict(mode=MODE_UNBIASED, rr_start=3, rr_prob=0.5, seed=9)
    _, full = _render(gpu_ctx, view, -1, max_depth=8, **params)
    _, cut = _render(gpu_ctx, view, -1, max_depth=3, **params)
    _, fin = _render(gpu_ctx, view, 3, max_depth=8, **params)
    assert fin["finish_rays"] == full["extension_rays"] - cut["extension_rays"]
    assert fin["finish_shadow_rays"] == full["shadow_rays"] - cut["shadow_rays"]


def test_finish_depth_zero_keeps_setting(gpu_ctx):
    gpu_ctx.set_launch_config(finish_depth=2)
    assert gpu_ctx.launch_config()["finish_depth"] == 2
    gpu_ctx.set_launch_config()
    assert gpu_ctx.launch_config()["finish_depth"] == 2
    gpu_ctx.set_launch_config(finish_depth=4)
