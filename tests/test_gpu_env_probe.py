"""Light-probe environment maps on the GPU against the oracle
(EnvironmentMap::get, src/environmentmap.cpp:35-66): a synthetic probe with
a small bright sun over a smooth sky, seen directly and lighting a diffuse
and a Blinn sphere; the same Philox streams on both sides, so the films
agree pixel by pixel.  A constant probe is a white furnace, and the guided
estimator stays unbiased under the probe."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_COMPAT, MODE_GUIDED, MODE_UNBIASED
from renderer_amd.scene import (BlinnMaterial, Camera, DiffuseMaterial, EnvironmentMap, IntersectableList, Scene,
                                Sphere)

pytestmark = pytest.mark.gpu


def _probe(n=96):
    y, x = np.mgrid[0:n, 0:n].astype(np.float32)
    img = np.zeros((n, n, 3), np.float32)
    img[..., 0] = 0.3 + 0.5 * x / n
    img[..., 1] = 0.4 + 0.3 * y / n
    img[..., 2] = 0.8 - 0.2 * x / n
    r = np.hypot(x - 0.62 * n, y - 0.30 * n)
    img[r < 3] = (120.0, 110.0, 90.0)
    return img


def _scene(res=64, probe=None):
    env = EnvironmentMap(_probe() if probe is None else probe, (1.0, 0.9, 1.1), mapping="probe")
    objs = [Sphere((-0.55, 0, 0), 0.5, DiffuseMaterial((0.7, 0.6, 0.5))),
            Sphere((0.55, 0, 0), 0.5, BlinnMaterial((0.3, 0.3, 0.3), (0.5, 0.5, 0.5), 30.0))]
    cam = Camera((0, 0.4, 3), (0, 0, 0), (0, 1, 0), 0.7, res, res)
    return Scene(cam, [env], IntersectableList(objs))


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_probe_lit_scene_pixelwise(gpu_ctx, mode):
    sc = _scene()
    arr = sc.flatten()
    print(f"probe scene, mode {mode}: 64x64, 16 spp", flush=True)
    gpu_ctx.upload(arr)
    view = (sc.camera, 4, 4)
    img = gpu_ctx.render([view], mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=5)[0]
    ref, _ = Oracle(arr, accel=1).render(view, mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=5)
    d = np.abs(img - ref)
    agree = np.all(d <= 1e-3 * np.abs(ref) + 1e-4, axis=2).mean()
    print(f"agreement {agree:.4f}, means {img.mean():.5f} {ref.mean():.5f}", flush=True)
    assert agree > 0.97
    assert abs(img.mean() - ref.mean()) <= 0.01 * ref.mean()


def test_constant_probe_furnace_on_gpu(gpu_ctx):
    env = EnvironmentMap(np.ones((32, 32, 3), np.float32), mapping="probe")
    cam = Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.45, 48, 48)
    sc = Scene(cam, [env], IntersectableList([Sphere((0, 0, 0), 0.5, DiffuseMaterial((0.6, 0.6, 0.6)))]))
    gpu_ctx.upload(sc.flatten())
    inside = np.zeros((48, 48), bool)
    inside[19:29, 19:29] = True
    for mode, expect in ((MODE_UNBIASED, 0.6), (MODE_COMPAT, 2 * 0.6 / (3 * math.pi))):
        img = gpu_ctx.render([(cam, 8, 8)], mode=mode, max_depth=5 if mode == MODE_COMPAT else 4, seed=9)[0]
        m = img[inside].mean()
        print(f"probe furnace mode {mode}: {m:.5f} expect {expect:.5f}", flush=True)
        assert m == pytest.approx(expect, rel=0.01)


def test_guided_under_the_probe_without_bias(gpu_ctx):
    sc = _scene(48)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.guide_init(5)
    view = (sc.camera, 16, 16)
    ref = gpu_ctx.render([view], mode=MODE_UNBIASED, max_depth=6, seed=31)[0]
    img = gpu_ctx.render([view], mode=MODE_GUIDED, max_depth=6, seed=32, guide_train=1)[0]
    print(f"guided mean {img.mean():.5f} unbiased {ref.mean():.5f}", flush=True)
    assert np.all(np.isfinite(img))
    assert img.mean() == pytest.approx(ref.mean(), rel=0.03)
