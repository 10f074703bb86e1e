"""Pins the CPU oracle to what fixes it independently of its own code:
published known-answer vectors, closed-form geometry, brute force, and
closed-form expectations of both estimators (the reference holds no tests,
golden images or fixtures of its own)."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, philox
from renderer_amd import scenes
from renderer_amd.scene import (AreaLight, Camera, DiffuseMaterial, EnvironmentMap, IntersectableList, Mesh,
                                MirrorMaterial, PointLight, Quad, Scene, Sphere, Triangle)
from tests.helpers import random_rays

INF = np.inf


# ---------------------------------------------------------------------------
# Philox-4x32-10: Random123 known-answer vectors (Salmon et al., SC'11)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert philox(ctr, key).tolist() == expect


def test_philox_uniformity():
    out = np.array([philox([i, 7, 3, 0], [42, 1]) for i in range(4000)]).reshape(-1)
    u = (out >> 8).astype(np.float64) / 2 ** 24
    hist, _ = np.histogram(u, bins=16, range=(0, 1))
    expected = len(u) / 16
    chi2 = ((hist - expected) ** 2 / expected).sum()
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4


# ---------------------------------------------------------------------------
# primitive tests against closed forms
# ---------------------------------------------------------------------------
def _ray(o, d, tmin=0.001, tmax=INF):
    return np.array([*o, tmin], np.float32), np.array([*d, tmax], np.float32)


def _single(obj, lights=()):
    sc = Scene(Camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 0.5, 4, 4), list(lights), IntersectableList([obj]))
    return Oracle(sc.flatten(), accel=0)


def test_triangle_closed_form():
    m = DiffuseMaterial((1, 1, 1))
    orc = _single(Triangle((0, 0, 0), (1, 0, 0), (0, 1, 0), m))
    assert orc.intersect_prim(0, *_ray((0.2, 0.2, 1), (0, 0, -1))) == pytest.approx(1.0, abs=1e-7)
    assert orc.intersect_prim(0, *_ray((0.2, 0.2, -1), (0, 0, 1))) == pytest.approx(1.0, abs=1e-7)  # two-sided
    assert orc.intersect_prim(0, *_ray((0.2, 0.2, 2), (0, 0, -2))) == pytest.approx(1.0, abs=1e-7)  # |d| = 2
    assert orc.intersect_prim(0, *_ray((0.5, 0.5, 1), (0, 0, -1))) == pytest.approx(1.0)  # u + v = 1, inclusive
    assert math.isnan(orc.intersect_prim(0, *_ray((0.6, 0.6, 1), (0, 0, -1))))  # u + v > 1
    assert math.isnan(orc.intersect_prim(0, *_ray((-0.1, 0.2, 1), (0, 0, -1))))  # u < 0
    assert math.isnan(orc.intersect_prim(0, *_ray((0.2, 0.2, 1), (1, 0, 0))))  # parallel: w = 0
    assert math.isnan(orc.intersect_prim(0, *_ray((0.2, 0.2, 1), (0, 0, -1), tmax=0.5)))  # beyond range
    assert math.isnan(orc.intersect_prim(0, *_ray((0.2, 0.2, 1), (0, 0, 1))))  # behind the origin
    t = orc.intersect_prim(0, *_ray((0.1, 0.3, 1), (0, 0, -1), tmin=1.0, tmax=1.0))  # inclusive range ends
    assert t == pytest.approx(1.0)


def test_quad_and_parallelogram_closed_form():
    m = DiffuseMaterial((1, 1, 1))
    orc = _single(Quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), m))
    for x, y in [(0.1, 0.1), (0.9, 0.9), (0.9, 0.1), (0.1, 0.9), (0.5, 0.5)]:
        assert orc.intersect_prim(0, *_ray((x, y, 2), (0, 0, -1))) == pytest.approx(2.0)
    assert math.isnan(orc.intersect_prim(0, *_ray((1.1, 0.5, 2), (0, 0, -1))))
    # area light: u, v in [0, 1] each, not u + v <= 1
    L = AreaLight((0, 0, 0), (2, 0, 0), (0, 1, 0), (1, 1, 1))
    sc = Scene(Camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 0.5, 4, 4), [L], IntersectableList([]))
    orc = Oracle(sc.flatten(), accel=0)
    assert orc.intersect_prim(0, *_ray((1.9, 0.9, 1), (0, 0, -1))) == pytest.approx(1.0)
    assert math.isnan(orc.intersect_prim(0, *_ray((2.1, 0.9, 1), (0, 0, -1))))


def test_sphere_closed_form():
    m = DiffuseMaterial((1, 1, 1))
    orc = _single(Sphere((0, 0, 0), 1.0, m))
    assert orc.intersect_prim(0, *_ray((0, 0, 3), (0, 0, -1))) == pytest.approx(2.0)
    assert orc.intersect_prim(0, *_ray((0, 0, 3), (0, 0, -2))) == pytest.approx(1.0)
    assert orc.intersect_prim(0, *_ray((0, 0, 0), (0, 0, 1))) == pytest.approx(1.0)  # inside: far root
    assert orc.intersect_prim(0, *_ray((1, 0, 3), (0, 0, -1))) == pytest.approx(3.0)  # tangent
    assert math.isnan(orc.intersect_prim(0, *_ray((0, 0, 3), (0, 0, 1))))  # behind
    assert math.isnan(orc.intersect_prim(0, *_ray((0, 0, 3), (0, 0, -1), tmax=1.5)))
    # from inside the range start past the near root picks the far root
    assert orc.intersect_prim(0, *_ray((0, 0, 3), (0, 0, -1), tmin=2.5)) == pytest.approx(4.0)


# ---------------------------------------------------------------------------
# reference BVH (leaf = 1, both children) == brute-force list
# ---------------------------------------------------------------------------
def _bvh_vs_list(arrays, o4, d4):
    ob, ol = Oracle(arrays, accel=1), Oracle(arrays, accel=0)
    pb, tb = ob.trace_closest(o4, d4)
    pl, tl = ol.trace_closest(o4, d4)
    hit = pl >= 0
    assert np.array_equal(pb >= 0, hit)
    assert np.array_equal(tb[hit], tl[hit])  # same arithmetic: bitwise equal t
    diff = np.nonzero(pb != pl)[0]
    for i in diff:  # only exact ties may differ (BVH: right child, list: later component)
        assert ob.intersect_prim(pb[i], o4[i], d4[i]) == tl[i]
    return hit.mean()


def test_bvh_matches_brute_force_box():
    arr = scenes.box_scene(8, 8).flatten()
    o4, d4 = random_rays(30000, (0, 0, 0), (1, 1, 1), seed=11, axis_fraction=0.3)
    assert _bvh_vs_list(arr, o4, d4) > 0.5


def test_bvh_matches_brute_force_mesh_and_spheres():
    rng = np.random.default_rng(3)
    m = DiffuseMaterial((0.5, 0.5, 0.5))
    P, F = scenes.displaced_sphere("octa", 8, 0.1, seed=2)
    items = [Mesh(P, F, m)] + [Sphere(rng.uniform(-1.5, 1.5, 3), rng.uniform(0.05, 0.3), m) for _ in range(20)]
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.7, 8, 8),
               [AreaLight((-0.5, 2, -0.5), (1, 0, 0), (0, 0, 1), (1, 1, 1)), PointLight((0, 3, 0), (1, 1, 1))],
               IntersectableList(items))
    o4, d4 = random_rays(30000, (-1.5, -1.5, -1.5), (1.5, 2, 1.5), seed=12)
    assert _bvh_vs_list(sc.flatten(), o4, d4) > 0.2


# ---------------------------------------------------------------------------
# estimators against closed forms
# ---------------------------------------------------------------------------
def _furnace_scene(rho, res=24):
    env = EnvironmentMap(np.ones((8, 16, 3), np.float32))
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.45, res, res), [env],
               IntersectableList([Sphere((0, 0, 0), 0.5, DiffuseMaterial((rho, rho, rho)))]))
    return sc


def _disk_pixels(res, cam_dist, fov, radius, frac):
    """Pixels whose whole footprint (with the half-pixel camera shift) lies
    well inside the sphere's silhouette."""
    t = math.tan(fov / 2)
    ang = math.asin(radius / cam_dist)
    rs = math.tan(ang) / t * res / 2 * frac
    yy, xx = np.mgrid[0:res, 0:res]
    cx = cy = res / 2 - 1.0  # centre of samples [j+0.5, j+1.5) -> pixel j+1 at the image centre
    return (xx - cx) ** 2 + (yy - cy) ** 2 < rs ** 2


def test_white_furnace_unbiased():
    """Convex diffuse sphere (albedo rho) under a constant environment L = 1:
    radiance rho on the sphere, 1 on the background."""
    rho = 0.6
    sc = _furnace_scene(rho)
    orc = Oracle(sc.flatten())
    img, _ = orc.render((sc.camera, 16, 16), mode=1, max_depth=4, seed=1)
    inside = _disk_pixels(24, 3.0, 0.45, 0.5, 0.75)
    assert img[inside].mean() == pytest.approx(rho, rel=0.01)
    assert img[0, 0] == pytest.approx([1, 1, 1], rel=1e-6)


def test_white_furnace_reference_estimator():
    """Reference estimator (src/unidipathtracingintegrator.cpp:70-82): the
    escaping bounce hits the environment, which replaces the next-event
    estimate by emission * brdf * |cos| without the pdf, so the sphere shows
    rho / pi * E[cos] = 2 rho / (3 pi) under cosine sampling."""
    rho = 0.6
    sc = _furnace_scene(rho)
    orc = Oracle(sc.flatten())
    img, _ = orc.render((sc.camera, 16, 16), mode=0, max_depth=5, seed=2)
    inside = _disk_pixels(24, 3.0, 0.45, 0.5, 0.75)
    assert img[inside].mean() == pytest.approx(2 * rho / (3 * math.pi), rel=0.01)
    assert img[0, 0] == pytest.approx([1, 1, 1], rel=1e-6)


def _rect_form_factor(a, b, c):
    """Differential area to a parallel a x b rectangle at height c, the point
    below one corner (Howell's catalogue, configuration B-3)."""
    X, Y = a / c, b / c
    return (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X)) +
            Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))) / (2 * math.pi)


def _floor_scene(light):
    floor = Quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5), DiffuseMaterial((0.5, 0.5, 0.5)))
    cam = Camera((0, 0.5, 0), (0, 0, 0), (0, 0, -1), 0.02, 8, 8)
    return Scene(cam, [light], IntersectableList([floor]))


def test_area_light_irradiance_closed_form():
    L = AreaLight((-0.5, 1, -0.5), (1, 0, 0), (0, 0, 1), (1, 1, 1))
    sc = _floor_scene(L)
    orc = Oracle(sc.flatten())
    img, _ = orc.render((sc.camera, 32, 32), mode=1, max_depth=3, seed=3)
    F = 4 * _rect_form_factor(0.5, 0.5, 1.0)
    assert img.mean() == pytest.approx(0.5 * F, rel=0.01)


def test_split_area_light_closed_form():
    """The same unit light as two halves: next-event estimation then picks a
    light with probability 1/2, and the MIS weights on both strategies must
    carry that choice or the sum is off the form factor."""
    halves = [AreaLight((-0.5, 1, -0.5), (0.5, 0, 0), (0, 0, 1), (1, 1, 1)),
              AreaLight((0.0, 1, -0.5), (0.5, 0, 0), (0, 0, 1), (1, 1, 1))]
    floor = Quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5), DiffuseMaterial((0.5, 0.5, 0.5)))
    cam = Camera((0, 0.5, 0), (0, 0, 0), (0, 0, -1), 0.02, 8, 8)
    orc = Oracle(Scene(cam, halves, IntersectableList([floor])).flatten())
    img, _ = orc.render((cam, 32, 32), mode=1, max_depth=3, seed=6)
    F = 4 * _rect_form_factor(0.5, 0.5, 1.0)
    assert img.mean() == pytest.approx(0.5 * F, rel=0.01)


def test_point_light_closed_form_both_estimators():
    sc = _floor_scene(PointLight((0, 1, 0), (2, 2, 2)))
    orc = Oracle(sc.flatten())
    expect = 0.5 / math.pi * 2 / 1.0
    for mode in (0, 1):
        img, _ = orc.render((sc.camera, 4, 4), mode=mode, max_depth=5, seed=4)
        assert img.mean() == pytest.approx(expect, rel=2e-3)


def test_mirror_reflects_environment():
    env = EnvironmentMap(np.full((8, 16, 3), 0.7, np.float32))
    mirror = Quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5), MirrorMaterial(1))
    cam = Camera((0, 1, 2), (0, 0, 0), (0, 1, 0), 0.3, 8, 8)
    sc = Scene(cam, [env], IntersectableList([mirror]))
    orc = Oracle(sc.flatten())
    for mode in (0, 1):
        img, _ = orc.render((cam, 2, 2), mode=mode, max_depth=4, seed=5)
        assert np.allclose(img, 0.7, rtol=1e-5)


def test_env_importance_sampling_is_a_density():
    img = scenes.sky_image(64, 32, seed=3, sun_radiance=500.0)
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.5, 4, 4), [EnvironmentMap(img)], IntersectableList([]))
    orc = Oracle(sc.flatten())
    rng = np.random.default_rng(0)
    # the pdf a sample reports is the pdf of its direction
    for u0, u1 in rng.random((200, 2)):
        d, p = orc.env_sample(u0, u1)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        assert orc.env_pdf(d) == pytest.approx(p, rel=2e-3)
    # integral of the pdf over the sphere is 1 (uniform-sphere Monte Carlo)
    v = rng.normal(size=(40000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    vals = np.array([orc.env_pdf(x.astype(np.float32)) for x in v])
    assert vals.mean() * 4 * math.pi == pytest.approx(1.0, rel=0.05)
