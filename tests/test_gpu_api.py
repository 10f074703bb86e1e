"""Error handling, cancellation, the fault hook, the autotuner and the
environment light through the C ABI on the GPU."""
import time

import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_COMPAT, MODE_GUIDED, MODE_UNBIASED, GpuContext, RtError, scenes
from renderer_amd.autotune import autotune

pytestmark = pytest.mark.gpu


def test_fault_hook_fails_the_next_call_only(gpu_ctx):
    sc = scenes.box_scene(16, 16)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.set_fault(True)
    with pytest.raises(RtError) as e:
        gpu_ctx.render([(sc.camera, 1, 1)], mode=MODE_UNBIASED)
    assert e.value.status == -6
    img = gpu_ctx.render([(sc.camera, 1, 1)], mode=MODE_UNBIASED)[0]
    assert img.mean() > 0


def test_cancel_between_passes(gpu_ctx):
    sc = scenes.box_scene(16, 16)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.request_stop(True)
    try:
        with pytest.raises(RtError) as e:
            gpu_ctx.render([(sc.camera, 2, 2)], mode=MODE_UNBIASED)
        assert e.value.status == -5
    finally:
        gpu_ctx.request_stop(False)
    assert gpu_ctx.render([(sc.camera, 1, 1)], mode=MODE_UNBIASED)[0].mean() > 0


def _render_in_thread(fn):
    import threading
    out = {}

    def run():
        t0 = time.perf_counter()
        try:
            fn()
            out["status"] = 0
        except RtError as e:
            out["status"] = e.status
        out["seconds"] = time.perf_counter() - t0

    th = threading.Thread(target=run)
    th.start()
    return th, out


def test_stop_from_another_thread_ends_a_render_early():
    """rt_request_stop from a second thread while rt_render runs: passes are
    queued at most two ahead of the GPU, so the call returns
    RT_ERR_CANCELLED long before the render would have finished."""
    sc = scenes.box_scene(256, 256)
    with GpuContext(0) as ctx:
        ctx.upload(sc.flatten())
        ctx.set_launch_config(path_slots=1 << 20)  # 64 passes of 1M paths
        view = [(sc.camera, 32, 32)]
        ctx.render(view, mode=MODE_UNBIASED, seed=1)  # warm-up
        t0 = time.perf_counter()
        ctx.render(view, mode=MODE_UNBIASED, seed=1)
        full = time.perf_counter() - t0
        print(f"full render {full:.3f} s, {ctx.stats()['passes']} passes", flush=True)
        th, out = _render_in_thread(lambda: ctx.render(view, mode=MODE_UNBIASED, seed=2))
        time.sleep(0.2 * full)
        ctx.request_stop(True)
        th.join(60)
        ctx.request_stop(False)
        print(f"cancelled render: status {out.get('status')} after {out.get('seconds', 0):.3f} s", flush=True)
        assert out["status"] == -5
        assert out["seconds"] < 0.6 * full
        assert ctx.render([(sc.camera, 1, 1)], mode=MODE_UNBIASED)[0].mean() > 0


def test_stop_from_another_thread_ends_metropolis_early():
    from renderer_amd import MLT_METROPOLIS

    sc = scenes.box_scene(128, 128)
    with GpuContext(0) as ctx:
        ctx.upload(sc.flatten())
        kw = dict(kind=MLT_METROPOLIS, mutations=2048, chains=1 << 16, bootstrap=1 << 16, seed=3)
        t0 = time.perf_counter()
        ctx.render_mlt(sc.camera, **kw)
        full = time.perf_counter() - t0
        print(f"full Metropolis render {full:.3f} s", flush=True)
        th, out = _render_in_thread(lambda: ctx.render_mlt(sc.camera, **kw))
        time.sleep(0.2 * full)
        ctx.request_stop(True)
        th.join(60)
        ctx.request_stop(False)
        print(f"cancelled: status {out.get('status')} after {out.get('seconds', 0):.3f} s", flush=True)
        assert out["status"] == -5
        assert out["seconds"] < 0.6 * full


def test_invalid_arguments_are_rejected(gpu_ctx):
    sc = scenes.box_scene(16, 16)
    gpu_ctx.upload(sc.flatten())
    for bad in (dict(max_depth=0), dict(max_depth=100), dict(rr_prob=1.0), dict(rr_prob=-0.1), dict(mode=7)):
        with pytest.raises(RtError) as e:
            gpu_ctx.render([(sc.camera, 1, 1)], **bad)
        assert e.value.status == -1


def test_guided_needs_a_network():
    with GpuContext(0) as ctx:
        sc = scenes.box_scene(16, 16)
        ctx.upload(sc.flatten())
        with pytest.raises(RtError) as e:
            ctx.render([(sc.camera, 1, 1)], mode=MODE_GUIDED)
        assert e.value.status == -7


def test_render_without_scene():
    with GpuContext(0) as ctx:
        sc = scenes.box_scene(16, 16)
        with pytest.raises(RtError) as e:
            ctx.render([(sc.camera, 1, 1)], mode=MODE_UNBIASED)
        assert e.value.status == -2


def test_autotuner_choice_does_not_change_the_image(gpu_ctx):
    sc = scenes.mesh_scene(64, 48, n=32, seed=3)
    gpu_ctx.upload(sc.flatten())
    views = [(sc.camera, 4, 4), (scenes.batched_views(1, seed=2, resolutions=((40, 30),))[0][0], 2, 3)]
    params = dict(mode=MODE_UNBIASED, max_depth=6)
    base = gpu_ctx.render(views, seed=5, **params)
    res = autotune(gpu_ctx, views, params, slots=(1 << 12, 1 << 14, 1 << 16), blocks_per_cu=(2, 6), spp_divisor=2,
                   repeats=1)
    assert (res["path_slots"], res["blocks_per_cu"]) in {(t["path_slots"], t["blocks_per_cu"]) for t in res["trials"]}
    tuned = gpu_ctx.render(views, seed=5, **params)
    for a, b in zip(base, tuned):
        assert np.array_equal(a, b)
    gpu_ctx.set_launch_config(path_slots=1 << 24, blocks_per_cu=12)


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_environment_light_pixelwise(gpu_ctx, mode):
    """Mesh lit only by the procedural sky: same-seed GPU and oracle images."""
    sc = scenes.mesh_scene(40, 40, n=24, seed=4, env=True)
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    orc = Oracle(arr)
    view = (sc.camera, 2, 2)
    ref, _ = orc.render(view, mode=mode, max_depth=6, seed=8)
    img = gpu_ctx.render([view], mode=mode, max_depth=6, seed=8)[0]
    d = np.abs(img - ref)
    agree = np.mean(np.all(d <= 1e-3 * np.abs(ref) + 1e-3, axis=2))
    print(f"env mode {mode}: agreement {agree:.4f} means {img.mean():.4f} {ref.mean():.4f}", flush=True)
    assert agree > 0.97
    assert abs(img.mean() - ref.mean()) <= 0.02 * ref.mean()
