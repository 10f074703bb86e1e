"""The oracle's Blinn and transparent materials and cone light
(src/blinnmaterial.cpp, src/transparentmaterial.cpp, src/conelight.cpp),
pinned by textbook forms: the Fresnel equations and Snell's law for the
refraction choice, energy conservation of a glass sphere in a white
furnace, the Blinn sampling density integrating to the probability of a
valid sample and agreeing with its samples, reciprocity of the Blinn BRDF, and a cone light that equals a
point light inside its cone and is dark outside."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, transparent_dir
from renderer_amd.scene import (BlinnMaterial, Camera, ConeLight, DiffuseMaterial, EnvironmentMap,
                                IntersectableList, PointLight, Quad, Scene, Sphere, TransparentMaterial)


def _fresnel(cos_i, n1, n2):
    sin_t = n1 / n2 * math.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t > 1:
        return 1.0, None
    cos_t = math.sqrt(1 - sin_t * sin_t)
    rs = ((n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)) ** 2
    rp = ((n1 * cos_t - n2 * cos_i) / (n1 * cos_t + n2 * cos_i)) ** 2
    return 0.5 * (rs + rp), cos_t


@pytest.mark.parametrize("deg", [0, 20, 45, 70, 85])
@pytest.mark.parametrize("inside", [False, True])
def test_refraction_fresnel_and_snell(deg, inside):
    ior = 1.5
    n = np.array([0, 0, 1.0])                      # outward normal of the glass
    th = math.radians(deg)
    if inside:   # travelling outwards from inside the glass
        d = np.array([math.sin(th), 0, math.cos(th)])
        n1, n2 = ior, 1.0
    else:
        d = np.array([math.sin(th), 0, -math.cos(th)])
        n1, n2 = 1.0, ior
    R, cos_t = _fresnel(math.cos(th), n1, n2)
    refl = d - 2 * np.dot(d, n) * n
    if R < 1:
        below = transparent_dir(d, n, ior, R * (1 - 1e-4))
        assert np.allclose(below, refl, atol=1e-6)
        above = transparent_dir(d, n, ior, R + 1e-4 * (1 - R))
        # Snell: the tangential component scales by n1 / n2, the normal one keeps its side
        assert np.linalg.norm(above) == pytest.approx(1, abs=1e-5)
        assert above[0] == pytest.approx(n1 / n2 * d[0], abs=1e-6)
        assert above[1] == pytest.approx(0, abs=1e-7)
        assert above[2] == pytest.approx(math.copysign(cos_t, d[2]), abs=1e-5)
    else:  # total internal reflection for any u
        assert np.allclose(transparent_dir(d, n, ior, 0.9999), refl, atol=1e-6)


def test_glass_sphere_in_a_white_furnace_is_invisible():
    """Specular refraction and reflection keep the throughput; every path
    leaves the sphere and sees the constant environment L = 1."""
    env = EnvironmentMap(np.ones((8, 16, 3), np.float32))
    cam = Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.45, 24, 24)
    sc = Scene(cam, [env], IntersectableList([Sphere((0, 0, 0), 0.5, TransparentMaterial(1.5))]))
    orc = Oracle(sc.flatten())
    for mode in (0, 1):
        img, _ = orc.render((cam, 4, 4), mode=mode, max_depth=64, seed=6)
        assert img.mean() == pytest.approx(1.0, abs=2e-3)
        assert np.all(img <= 1 + 1e-6)


def _blinn_oracle(kd=0.3, ks=0.5, e=30.0):
    m = BlinnMaterial((kd, kd, kd), (ks, ks, ks), e)
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.5, 4, 4), [PointLight((0, 2, 0), (1, 1, 1))],
               IntersectableList([Sphere((0, 0, 0), 0.5, m)]))
    return Oracle(sc.flatten())


def _sphere_dirs(n, rng):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("theta_o", [0.1, 0.7, 1.3])
def test_blinn_density_integrates_to_one_and_matches_its_samples(theta_o):
    orc = _blinn_oracle()
    rng = np.random.default_rng(1)
    nf = np.array([0, 0, 1.0])
    wo = np.array([math.sin(theta_o), 0, math.cos(theta_o)])
    dirs = _sphere_dirs(60000, rng)
    vals = [orc.bsdf(0, wo, w, nf) for w in dirs]
    pdf = np.array([p for _, p in vals])
    fcos = np.array([f[0] * max(0.0, w[2]) for (f, _), w in zip(vals, dirs)])
    # albedo: uniform-sphere estimate vs the sampler's own estimate
    albedo_uniform = (4 * math.pi * fcos).mean()
    est, ok = [], 0
    for u in rng.random((60000, 3)):
        wi = orc.bsdf_sample(0, wo, nf, *u)
        if wi is None:
            est.append(0.0)
            continue
        ok += 1
        f, p = orc.bsdf(0, wo, wi, nf)
        est.append(f[0] * wi[2] / p)
    # the density integrates to the probability that a sample is valid (a
    # half vector can reflect wo below the surface; such a sample ends the path)
    assert (4 * math.pi * pdf).mean() == pytest.approx(ok / 60000, abs=0.03)
    assert np.mean(est) == pytest.approx(albedo_uniform, rel=0.03)


def test_blinn_is_reciprocal():
    orc = _blinn_oracle()
    rng = np.random.default_rng(2)
    nf = np.array([0, 0, 1.0])
    for _ in range(200):
        a, b = _sphere_dirs(2, rng)
        a[2], b[2] = abs(a[2]), abs(b[2])
        fa, _ = orc.bsdf(0, a, b, nf)
        fb, _ = orc.bsdf(0, b, a, nf)
        assert fa[0] == pytest.approx(fb[0], rel=1e-5)


def test_cone_light_is_a_point_light_inside_its_cone():
    floor = Quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5), DiffuseMaterial((0.5, 0.5, 0.5)))
    cam = Camera((0, 2, 0.001), (0, 0, 0), (0, 0, -1), 1.2, 40, 40)
    angle = 0.5
    imgs = {}
    for name, light in (("cone", ConeLight((0, 1, 0), (0, -1, 0), angle, (2, 2, 2))),
                        ("point", PointLight((0, 1, 0), (2, 2, 2)))):
        orc = Oracle(Scene(cam, [light], IntersectableList([floor])).flatten())
        imgs[name], _ = orc.render((cam, 2, 2), mode=1, max_depth=1, seed=7)
    # floor point of each pixel's sample region (camera looking down the y axis)
    t = math.tan(1.2 / 2)
    j = np.arange(40) + 1.0                            # sample positions [j + .5, j + 1.5)
    x = (2 * j / 40 - 1) * t * 2.0
    X, Z = np.meshgrid(x, x)
    ang = np.arctan(np.hypot(X, Z) / 1.0)
    half_pixel = 2 * t * 2.0 / 40
    inside = np.arctan((np.hypot(X, Z) + half_pixel) / 1.0) < angle
    outside = np.arctan(np.maximum(np.hypot(X, Z) - half_pixel, 0) / 1.0) > angle
    assert inside.sum() > 50 and outside.sum() > 50
    assert np.array_equal(imgs["cone"][inside], imgs["point"][inside])
    assert np.all(imgs["cone"][outside] == 0)
    assert np.all(imgs["point"][outside] > 0)
