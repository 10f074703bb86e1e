"""Texture materials on the GPU against the oracle: quads (default texture
coordinates), a triangle with its own coordinates, a sphere and a CSG
sphere (coordinates from the normal) under a checker and a gradient image,
rendered in both estimators pixel by pixel; a scene file naming a PNG
texture renders bit for bit like the host model."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_COMPAT, MODE_UNBIASED
from renderer_amd.film import write_image
from renderer_amd.scene import (AreaLight, Box, Camera, CSGDifference, DiffuseMaterial, IntersectableList, Quad, Scene,
                                Sphere, TextureMaterial, Triangle)
from renderer_amd.scene_file import SceneFile

pytestmark = pytest.mark.gpu


def _images():
    y, x = np.mgrid[0:16, 0:16]
    checker = np.where(((x // 4) + (y // 4)) % 2 == 0, 0.9, 0.15).astype(np.float32)
    chk = np.dstack([checker, checker * 0.8, checker * 0.5]).astype(np.float32)
    grad = np.dstack([x / 15.0, y / 15.0, 0.5 + 0 * x]).astype(np.float32)
    return chk, grad


def _scene(res=64):
    chk, grad = _images()
    t1, t2 = TextureMaterial(chk, (1.0, 0.9, 0.8), 1.0), TextureMaterial(grad, (0.9, 0.9, 0.9), 2.2)
    grey = DiffuseMaterial((0.7, 0.7, 0.7))
    objs = [Quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), t1),
            Quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), t2),
            Quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), grey),
            Triangle((1, 0, 0), (1, 1, 0), (1, 0.5, 1), t1, texcoords=[[0.1, 0.1], [0.9, 0.2], [0.5, 0.95]]),
            Sphere((0.3, 0.2, 0.6), 0.2, t2),
            CSGDifference(Sphere((0.7, 0.25, 0.45), 0.22, t1), Box((0.6, 0.3, 0.3), (0.9, 0.6, 0.6), grey))]
    lights = [AreaLight((0.35, 0.999, 0.35), (0.3, 0, 0), (0, 0, 0.3), (14, 14, 14))]
    cam = Camera((0.5, 0.5, 2.4), (0.5, 0.45, 0.0), (0, 1, 0), 0.7, res, res)
    return Scene(cam, lights, IntersectableList(objs))


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_textured_scene_renders_like_the_oracle(gpu_ctx, mode):
    sc = _scene()
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    view = (sc.camera, 4, 4)
    print(f"textured scene, mode {mode}", flush=True)
    img = gpu_ctx.render([view], mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=23)[0]
    ref, _ = Oracle(arr).render(view, mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=23)
    agree = np.all(np.abs(img - ref) <= 1e-3 * np.abs(ref) + 1e-4, axis=2).mean()
    print(f"agreement {agree:.4f}, means {img.mean():.5f} {ref.mean():.5f}", flush=True)
    assert agree > 0.97
    assert abs(img.mean() - ref.mean()) <= 0.01 * ref.mean()


def test_texture_scene_file_renders_like_the_host_model(gpu_ctx, tmp_path):
    chk, _ = _images()
    png = tmp_path / "checker.png"
    write_image(str(png), np.round(chk * 255).astype(np.uint8))
    text = f"""
camera = camera([0.5 0.5 2.4], [0.5 0.45 0], [0 1 0], 0.7, 48, 48);
lights = lights{{ arealight([0.35 0.999 0.35], [0.3 0 0], [0 0 0.3], [14 14 14]) }};
floor = texture("{png}", [1 0.9 0.8], 2.2);
intersectable = intersectables{{
  quad([0 0 0], [1 0 0], [1 0 1], [0 0 1], floor)
  quad([0 0 0], [0 1 0], [1 1 0], [1 0 0], texture("{png}", [0.5 0.5 0.5], 3))
  t([1 0 0], [1 1 0], [1 0.5 1], [0 0 0], [0 0 0], [0 0 0], [0.1 0.1], [0.9 0.2], [0.5 0.95], floor)
  sphere([0.3 0.2 0.6], 0.2, floor)
}};
"""
    a = SceneFile.parse(text).arrays()
    img8 = TextureMaterial.load(png).image
    # the grammar's texture rule keeps only the file name (src/sceneparser.cpp:109,
    # ast_texture_material): coefficient and gamma are the defaults (1, 1, 1) and 1
    floor = TextureMaterial(img8)
    host = Scene(Camera((0.5, 0.5, 2.4), (0.5, 0.45, 0), (0, 1, 0), 0.7, 48, 48),
                 [AreaLight((0.35, 0.999, 0.35), (0.3, 0, 0), (0, 0, 0.3), (14, 14, 14))],
                 IntersectableList([Quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), floor),
                                    Quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), floor),
                                    Triangle((1, 0, 0), (1, 1, 0), (1, 0.5, 1), floor,
                                             texcoords=[[0.1, 0.1], [0.9, 0.2], [0.5, 0.95]]),
                                    Sphere((0.3, 0.2, 0.6), 0.2, floor)])).flatten()
    views = [(a.camera, 2, 2)]
    gpu_ctx.upload(a)
    img_file = gpu_ctx.render(views, mode=MODE_UNBIASED, max_depth=6, seed=3)[0]
    gpu_ctx.upload(host)
    img_host = gpu_ctx.render(views, mode=MODE_UNBIASED, max_depth=6, seed=3)[0]
    assert img_file.mean() > 0
    assert np.array_equal(img_file, img_host)
