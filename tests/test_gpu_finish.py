"""The finishing launch (rt_launch_config.finish_depth, finish_paths: a
pass's remaining paths traced, shaded and shadow-tested per path in one
launch, from the first bounce >= finish_depth whose queue holds at most
finish_paths paths, decided on the device) against the per-bounce launches:
every path changes its state in the same order either way, so the films
agree bit for bit; the ray counts agree, and those the finishing launch
reports are the late bounces' share."""
import numpy as np
import pytest

from renderer_amd import MODE_COMPAT, MODE_UNBIASED, scenes
from tests.test_gpu_csg import _mixed_scene

pytestmark = pytest.mark.gpu

BIG = 1 << 30


def _render(ctx, view, finish, paths=BIG, **params):
    old = ctx.launch_config()
    ctx.set_launch_config(finish_depth=finish, finish_paths=paths)
    try:
        img = ctx.render([view], **params)[0]
        return img, ctx.stats()
    finally:
        ctx.set_launch_config(finish_depth=old["finish_depth"], finish_paths=old["finish_paths"])


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
    assert rs["finish_rays"] == 0 and rs["finish_shadow_rays"] == 0 and rs["finish_bounce"] == -1
    assert rs["extend_launches"] == 8 * rs["passes"]
    for finish in (1, 2, 3, 6):
        img, st = _render(gpu_ctx, view, finish, **params)
        print(f"{scene} mode {mode} finish {finish}: finish rays {st['finish_rays']} / {st['extension_rays']}, "
              f"shadow {st['finish_shadow_rays']} / {st['shadow_rays']}", flush=True)
        assert np.array_equal(img, ref), f"finish_depth {finish}: films differ"
        for k in ("camera_rays", "extension_rays", "shadow_rays"):
            assert st[k] == rs[k], k
        assert st["finish_bounce"] == finish  # every queue of these scenes is under the bound
        assert 0 < st["finish_rays"] <= st["extension_rays"]
        assert st["finish_shadow_rays"] <= st["shadow_rays"]
        assert st["extend_launches"] == finish * st["passes"]


def _queue_lengths(ctx, view, max_depth, **params):
    """Paths in the queue of each bounce 1 .. max_depth - 1: the extension
    rays a render cut at depth d + 1 traces beyond one cut at depth d."""
    ext = []
    for d in range(1, max_depth + 1):
        _, st = _render(ctx, view, -1, max_depth=d, **params)
        ext.append(st["extension_rays"])
    return [ext[d] - ext[d - 1] for d in range(1, max_depth)]


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_finish_from_the_first_short_queue(gpu_ctx, mode):
    """With a queue bound the render moves to the finishing launch at the
    first bounce whose queue holds at most that many paths; the film is
    the same as with every bounce in per-bounce launches."""
    sc = scenes.mesh_scene(48, 48, n=24, seed=2)
    gpu_ctx.upload(sc.flatten())
    view = (sc.camera, 4, 4)
    params = dict(mode=mode, rr_start=3, rr_prob=0.5, seed=7)
    q = _queue_lengths(gpu_ctx, view, 8, **params)  # q[b - 1]: queue of bounce b
    ref, rs = _render(gpu_ctx, view, -1, max_depth=8, **params)
    print(f"mode {mode}: queues {q}", flush=True)
    for bound in (q[0], q[1] - 1, q[2] - 1, max(q[3], 1)):
        expect = next((b for b in range(1, 8) if q[b - 1] <= bound), -1)
        img, st = _render(gpu_ctx, view, 1, paths=bound, max_depth=8, **params)
        assert st["finish_bounce"] == expect, (bound, expect, st["finish_bounce"])
        assert np.array_equal(img, ref)
        for k in ("camera_rays", "extension_rays", "shadow_rays"):
            assert st[k] == rs[k], k
        if expect > 0:
            assert st["extend_launches"] == expect * st["passes"]


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
    assert fin["finish_bounce"] == 3
    assert fin["finish_rays"] == full["extension_rays"] - cut["extension_rays"]
    assert fin["finish_shadow_rays"] == full["shadow_rays"] - cut["shadow_rays"]


def test_finish_config_zero_keeps_setting(gpu_ctx):
    old = gpu_ctx.launch_config()
    assert old["finish_depth"] == 1 and old["finish_paths"] % 4096 == 0 and old["finish_paths"] > 0
    gpu_ctx.set_launch_config(finish_depth=2, finish_paths=1000)
    gpu_ctx.set_launch_config()
    cfg = gpu_ctx.launch_config()
    assert cfg["finish_depth"] == 2 and cfg["finish_paths"] == 1000
    gpu_ctx.set_launch_config(finish_depth=old["finish_depth"], finish_paths=old["finish_paths"])
