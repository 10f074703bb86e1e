"""Blinn and refractive materials and the cone light on the GPU against the
oracle: the same Philox streams on both sides, so the films agree pixel by
pixel up to floating-point divergence of single paths; the glass sphere in
a white furnace stays invisible; the guided estimator renders the same
scene without bias against the unbiased one."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_COMPAT, MODE_GUIDED, MODE_UNBIASED, scenes
from renderer_amd.scene import (AreaLight, BlinnMaterial, Camera, ConeLight, DiffuseMaterial, EnvironmentMap,
                                IntersectableList, Quad, Scene, Sphere, TransparentMaterial)

pytestmark = pytest.mark.gpu


def _scene(res=64):
    grey = DiffuseMaterial((0.75, 0.75, 0.75))
    walls = [
        Quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), grey),
        Quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), grey),
        Quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), grey),
        Quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), DiffuseMaterial((0.63, 0.06, 0.05))),
        Quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), DiffuseMaterial((0.14, 0.45, 0.09))),
    ]
    objs = [Sphere((0.3, 0.2, 0.62), 0.2, BlinnMaterial((0.4, 0.3, 0.2), (0.4, 0.4, 0.4), 40.0)),
            Sphere((0.7, 0.2, 0.35), 0.2, TransparentMaterial(1.5))]
    lights = [AreaLight((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), (17, 12, 4)),
              ConeLight((0.5, 0.95, 0.8), (0, -1, -0.3), 0.6, (1.5, 1.5, 1.5))]
    cam = Camera((0.5, 0.5, 2.4), (0.5, 0.5, 0.0), (0, 1, 0), 0.6981317, res, res)
    return Scene(cam, lights, IntersectableList(walls + objs))


def _agreement(a, b):
    d = np.abs(a - b)
    return np.all(d <= 1e-3 * np.abs(b) + 1e-4, axis=2).mean()


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_materials_and_cone_light_pixelwise(gpu_ctx, mode):
    sc = _scene()
    arr = sc.flatten()
    print(f"materials scene, mode {mode}: 64x64, 16 spp", flush=True)
    gpu_ctx.upload(arr)
    view = (sc.camera, 4, 4)
    img = gpu_ctx.render([view], mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=13)[0]
    ref, _ = Oracle(arr, accel=1).render(view, mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=13)
    agree = _agreement(img, ref)
    print(f"agreement {agree:.4f}, means {img.mean():.5f} {ref.mean():.5f}", flush=True)
    assert agree > 0.97
    assert abs(img.mean() - ref.mean()) <= 0.01 * ref.mean()


def test_glass_furnace_on_gpu(gpu_ctx):
    env = EnvironmentMap(np.ones((8, 16, 3), np.float32))
    cam = Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.45, 48, 48)
    sc = Scene(cam, [env], IntersectableList([Sphere((0, 0, 0), 0.5, TransparentMaterial(1.5))]))
    gpu_ctx.upload(sc.flatten())
    for mode in (MODE_COMPAT, MODE_UNBIASED):
        img = gpu_ctx.render([(cam, 4, 4)], mode=mode, max_depth=64, seed=3)[0]
        print(f"glass furnace mode {mode}: mean {img.mean():.6f}", flush=True)
        assert img.mean() == pytest.approx(1.0, abs=2e-3)
        assert np.all(img <= 1 + 1e-5)


def test_guided_renders_the_new_materials_without_bias(gpu_ctx):
    sc = _scene(48)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.guide_init(5)
    view = (sc.camera, 16, 16)
    ref = gpu_ctx.render([view], mode=MODE_UNBIASED, max_depth=8, seed=21)[0]
    img = gpu_ctx.render([view], mode=MODE_GUIDED, max_depth=8, seed=22, guide_train=1)[0]
    print(f"guided mean {img.mean():.5f} unbiased {ref.mean():.5f}", flush=True)
    assert np.all(np.isfinite(img))
    assert img.mean() == pytest.approx(ref.mean(), rel=0.03)
