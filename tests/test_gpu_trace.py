"""GPU ray-batch parity: librt_hip.so closest / any hit against the oracle's
reference BVH (BVHNode::intersect, src/bvh.cpp:135-149) on seeded batches."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import scenes
from renderer_amd.scene import IntersectableList, Mesh, Scene, Sphere, DiffuseMaterial, Camera
from tests.helpers import camera_rays, compare_hits, random_rays

pytestmark = pytest.mark.gpu


def _check(ctx, arrays, o4, d4, max_bad_frac=0.0):
    orc = Oracle(arrays, accel=1)
    ctx.upload(arrays)
    pg, tg = ctx.trace_closest(o4, d4)
    pr, tr = orc.trace_closest(o4, d4)
    exact, ties, bad = compare_hits(orc, o4, d4, pg, tg, pr, tr)
    n = len(o4)
    print(f"rays {n}: exact {exact} ties {ties} bad {len(bad)} hits {int((pr >= 0).sum())}", flush=True)
    assert len(bad) <= max_bad_frac * n, (bad[:10], pg[bad[:10]], pr[bad[:10]], tg[bad[:10]], tr[bad[:10]])
    occ_g = ctx.trace_any(o4, d4)
    occ_r = orc.trace_any(o4, d4)
    mism = np.nonzero(occ_g != occ_r)[0]
    assert len(mism) <= max(len(bad), int(max_bad_frac * n)), mism[:10]
    return pg, pr


def test_box_random_rays(gpu_ctx):
    arr = scenes.box_scene(64, 64).flatten()
    o4, d4 = random_rays(200_000, (0, 0, 0), (1, 1, 1), seed=1)
    pg, pr = _check(gpu_ctx, arr, o4, d4)
    assert (pr >= 0).mean() > 0.5


def test_box_camera_rays_unnormalised(gpu_ctx):
    sc = scenes.box_scene(128, 128)
    o4, d4 = camera_rays(sc, spp=2, seed=2)
    _check(gpu_ctx, sc.flatten(), o4, d4)


def test_box_finite_ranges(gpu_ctx):
    arr = scenes.box_scene(64, 64).flatten()
    o4, d4 = random_rays(100_000, (0, 0, 0), (1, 1, 1), seed=3, normalize=False)
    rng = np.random.default_rng(4)
    d4[:, 3] = rng.uniform(0.0, 2.0, len(d4))
    o4[:, 3] = rng.uniform(0.0, 0.5, len(d4))
    _check(gpu_ctx, arr, o4, d4)


def test_mesh_random_rays(gpu_ctx):
    sc = scenes.mesh_scene(64, 64, n=64, seed=5)
    o4, d4 = random_rays(100_000, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2), seed=6)
    pg, pr = _check(gpu_ctx, sc.flatten(), o4, d4, max_bad_frac=1e-4)
    assert (pr >= 0).mean() > 0.3


def test_spheres_inside_and_outside(gpu_ctx):
    m = DiffuseMaterial((0.5, 0.5, 0.5))
    rng = np.random.default_rng(7)
    sph = [Sphere(rng.uniform(-1, 1, 3), rng.uniform(0.05, 0.4), m) for _ in range(200)]
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.7, 8, 8), [], IntersectableList(sph))
    o4, d4 = random_rays(100_000, (-1, -1, -1), (1, 1, 1), seed=8)
    _check(gpu_ctx, sc.flatten(), o4, d4)


def test_empty_and_degenerate_batches(gpu_ctx):
    arr = scenes.box_scene(8, 8).flatten()
    gpu_ctx.upload(arr)
    p, t = gpu_ctx.trace_closest(np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32))
    assert len(p) == 0
    # empty range (tmax < tmin) never hits
    o4, d4 = random_rays(1000, (0, 0, 0), (1, 1, 1), seed=9)
    d4[:, 3] = 0.0
    o4[:, 3] = 0.5
    p, t = gpu_ctx.trace_closest(o4, d4)
    assert (p < 0).all() and np.isnan(t).all()


def test_scene_without_primitives(gpu_ctx):
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.7, 8, 8), [], IntersectableList([]))
    gpu_ctx.upload(sc.flatten())
    o4, d4 = random_rays(100, (0, 0, 0), (1, 1, 1), seed=10)
    p, t = gpu_ctx.trace_closest(o4, d4)
    assert (p < 0).all()
