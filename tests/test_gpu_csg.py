"""Planes, boxes and CSG objects on the GPU against the oracle: closest and
any hits of random rays (same arithmetic on both sides, so t agrees bit for
bit), and renders of a scene with a ground plane, a box and CSG objects in
both estimators, pixel by pixel."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_COMPAT, MODE_UNBIASED
from renderer_amd.scene import (AreaLight, Box, Camera, CSGDifference, CSGIntersection, CSGUnion, CSGXor,
                                DiffuseMaterial, IntersectableList, MirrorMaterial, PhongMaterial, Plane, PointLight,
                                Quad, Scene, Sphere)
from tests.test_oracle_csg import _build, _random_tree

pytestmark = pytest.mark.gpu


def _mixed_scene(res=64):
    grey, red = DiffuseMaterial((0.7, 0.7, 0.7)), DiffuseMaterial((0.7, 0.2, 0.1))
    blue, shiny = DiffuseMaterial((0.1, 0.2, 0.7)), PhongMaterial((0.4, 0.4, 0.2), (0.3, 0.3, 0.3), 30.0)
    objs = [
        Plane([0, 1, 0, 1], grey),                                                      # ground y = -1
        Box((-2.0, -1.0, -1.5), (-0.9, 0.2, -0.4), red),
        CSGDifference(Sphere((0.6, -0.2, -0.6), 0.8, shiny), Box((0.1, -0.3, -0.1), (1.5, 1.0, 0.5), blue)),
        CSGUnion(CSGIntersection(Box((-0.6, -1.0, 0.6), (0.2, -0.2, 1.4), red), Sphere((-0.2, -0.6, 1.0), 0.5, blue)),
                 Sphere((1.4, -0.7, 1.0), 0.3, MirrorMaterial(1))),
        CSGXor(Sphere((-1.5, 0.9, 0.5), 0.45, grey), Sphere((-1.2, 0.9, 0.5), 0.45, red)),
        Quad((-3, -1, -3), (3, -1, -3), (3, 3, -3), (-3, 3, -3), grey),
        Sphere((1.6, 0.8, -1.2), 0.35, shiny),
    ]
    lights = [AreaLight((-0.5, 2.9, -0.5), (1.0, 0, 0), (0, 0, 1.0), (9, 9, 8)), PointLight((2, 2, 2), (2, 2, 2))]
    cam = Camera((0.3, 0.8, 4.5), (0, -0.3, 0), (0, 1, 0), 0.75, res, res)
    return Scene(cam, lights, IntersectableList(objs))


def _random_rays(rng, n, lo=-3.0, hi=3.0):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o4 = np.concatenate([o, np.full((n, 1), 1e-3, np.float32)], 1)
    d4 = np.concatenate([d, np.full((n, 1), np.inf, np.float32)], 1)
    return o4, d4


def _compare_hits(gpu_ctx, arr, o, d):
    gp, gt = gpu_ctx.trace_closest(o, d)
    orc = Oracle(arr)
    op, ot = orc.trace_closest(o, d)
    same = (gp == op) & ((gt == ot) | (np.isnan(gt) & np.isnan(ot)) | (np.isinf(gt) & np.isinf(ot)))
    ga, oa = gpu_ctx.trace_any(o, d), orc.trace_any(o, d)
    return same.mean(), (ga == oa).mean(), (op >= 0).mean()


@pytest.mark.parametrize("seed", range(4))
def test_random_csg_trees_hit_like_the_oracle(gpu_ctx, seed):
    rng = np.random.default_rng(100 + seed)
    objs = [_build(_random_tree(rng, 3)) for _ in range(3)]
    sc = Scene(Camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 0.8, 8, 8), [PointLight((0, 5, 0), (1, 1, 1))],
               IntersectableList(objs))
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    o, d = _random_rays(rng, 50000)
    print(f"random CSG trees, seed {seed}: 50000 rays", flush=True)
    closest, anyhit, frac = _compare_hits(gpu_ctx, arr, o, d)
    print(f"closest agreement {closest:.6f}, any {anyhit:.6f}, hit fraction {frac:.3f}", flush=True)
    assert closest > 0.9995 and anyhit > 0.9995


def test_mixed_scene_hits_like_the_oracle(gpu_ctx):
    sc = _mixed_scene()
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    o, d = _random_rays(np.random.default_rng(3), 100000)
    closest, anyhit, frac = _compare_hits(gpu_ctx, arr, o, d)
    print(f"mixed scene: closest agreement {closest:.6f}, any {anyhit:.6f}, hit fraction {frac:.3f}", flush=True)
    assert closest > 0.9995 and anyhit > 0.9995 and frac > 0.5


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_mixed_scene_renders_like_the_oracle(gpu_ctx, mode):
    sc = _mixed_scene()
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    view = (sc.camera, 4, 4)
    print(f"mixed scene render, mode {mode}", flush=True)
    img = gpu_ctx.render([view], mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=17)[0]
    ref, _ = Oracle(arr).render(view, mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=17)
    agree = np.all(np.abs(img - ref) <= 1e-3 * np.abs(ref) + 1e-4, axis=2).mean()
    print(f"agreement {agree:.4f}, means {img.mean():.5f} {ref.mean():.5f}", flush=True)
    assert agree > 0.97
    assert abs(img.mean() - ref.mean()) <= 0.01 * ref.mean()


INSTANCED = """camera = camera([0.3 0.8 4.5],[0 -0.3 0],[0 1 0],0.75,64,64);
lights = lights{ arealight([-0.5 2.9 -0.5],[1 0 0],[0 0 1],[9 9 8]) pointlight([2 2 2],[2 2 2]) };
intersectable = intersectables{
  plane([0 1 0 1], diffuse([.7 .7 .7]))
  instance(translate([-1.2 -0.3 -0.6]) * scale([1.5 0.6 0.9]), sphere([0 0 0], 0.6, phong([.4 .4 .2],[.3 .3 .3],30)))
  instance(translate([0.8 -0.4 -0.4]) * rotate(35, [0 1 1]), box([-0.4 -0.4 -0.4],[0.4 0.4 0.4], diffuse([.7 .2 .1])))
  instance(translate([0 -0.3 1]) * rotate(25, [1 0 0]) * scale([1 1.5 0.7]),
           diff(box([-0.5 -0.5 -0.5],[0.5 0.5 0.5], diffuse([.1 .2 .7])), sphere([0 0 0], 0.6, mirror(1))))
  quad([-3 -1 -3],[3 -1 -3],[3 3 -3],[-3 3 -3], diffuse([.7 .7 .7]))
};
"""


def test_instanced_objects_hit_like_the_oracle(gpu_ctx):
    """An ellipsoid, a rotated box and a rotated, scaled CSG difference (the
    ray mapped into object space on both sides with the same arithmetic, so
    t agrees bit for bit)."""
    from renderer_amd.scene_file import SceneFile

    sf = SceneFile.parse(INSTANCED)
    arr = sf.arrays()
    assert arr.object_xform is not None
    gpu_ctx.upload(arr)
    o, d = _random_rays(np.random.default_rng(5), 100000)
    closest, anyhit, frac = _compare_hits(gpu_ctx, arr, o, d)
    print(f"instanced objects: closest agreement {closest:.6f}, any {anyhit:.6f}, hit fraction {frac:.3f}",
          flush=True)
    assert closest > 0.9995 and anyhit > 0.9995 and frac > 0.5


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_instanced_objects_render_like_the_oracle(gpu_ctx, mode):
    from renderer_amd.scene_file import SceneFile

    sf = SceneFile.parse(INSTANCED)
    arr = sf.arrays()
    gpu_ctx.upload(arr)
    view = (arr.camera, 4, 4)
    print(f"instanced scene render, mode {mode}", flush=True)
    img = gpu_ctx.render([view], mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=23)[0]
    ref, _ = Oracle(arr).render(view, mode=mode, max_depth=6, rr_start=3, rr_prob=0.5, seed=23)
    agree = np.all(np.abs(img - ref) <= 1e-3 * np.abs(ref) + 1e-4, axis=2).mean()
    print(f"agreement {agree:.4f}, means {img.mean():.5f} {ref.mean():.5f}", flush=True)
    assert agree > 0.97
    assert abs(img.mean() - ref.mean()) <= 0.01 * ref.mean()
