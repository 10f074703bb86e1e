"""The Markov-chain renderers of the oracle (src/metropolisrenderer.cpp,
src/energyredistributionrenderer.cpp, src/metropolissample.cpp) pinned by
what the mathematics fixes: the perturbation's log-uniform step law and
wrap-around, the sample-driven integrator equal to the pixel sampler when
both are fed the same numbers, conservation of the deposited energy
(Metropolis: luminance b per pixel; ERPT: e_d per chain step), the
Metropolis image converging to the path-traced image when the two target
the same integrand, and ERPT reproducing a constant background."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, mlt_perturb, philox
from renderer_amd import MODE_COMPAT, scenes
from renderer_amd._abi import MLT_ERPT, MLT_METROPOLIS
from renderer_amd.scene import (AreaLight, Camera, DiffuseMaterial, EnvironmentMap, IntersectableList, PointLight,
                                Quad, Scene)

A, B = 1.0 / 1024, 1.0 / 64


def _lum(img):
    return 0.299 * img[..., 0].astype(np.float64) + 0.587 * img[..., 1] + 0.114 * img[..., 2]


def test_perturbation_step_law():
    rng = np.random.default_rng(1)
    n = 200000
    s = rng.random(n).astype(np.float32)
    u = rng.random(n).astype(np.float32)
    v = rng.random(n).astype(np.float32)
    out = mlt_perturb(s, u, v)
    assert out.min() >= 0 and out.max() <= 1
    d = (out.astype(np.float64) - s + 0.5) % 1.0 - 0.5   # the step, undoing the wrap
    assert np.all((d < 0) == (v < 0.5))
    m = np.abs(d)
    assert m.min() >= A * (1 - 1e-4) and m.max() <= B * (1 + 1e-4)
    # log-uniform on [a, b]: log|d| is uniform, the mean step is (b - a) / ln(b / a)
    t = (np.log(m) - math.log(A)) / math.log(B / A)
    hist = np.histogram(t, bins=20, range=(0, 1))[0] / n
    assert np.abs(hist - 0.05).max() < 0.003
    assert m.mean() == pytest.approx((B - A) / math.log(B / A), rel=0.01)
    # u = 0 is the longest step, u -> 1 the shortest; wrap-around both ways
    edge = mlt_perturb(np.float32([0.001, 0.999, 0.5, 0.5]), np.float32([0, 0, 0, 0.9999999]),
                       np.float32([0.2, 0.7, 0.7, 0.7]))
    assert edge[0] == pytest.approx(0.001 - B + 1, abs=1e-7)
    assert edge[1] == pytest.approx(0.999 + B - 1, abs=1e-7)
    assert edge[2] == pytest.approx(0.5 + B, abs=1e-7)
    assert edge[3] == pytest.approx(0.5 + A, abs=1e-7)


def _uniforms(seed, view, pixel, sample, group):
    x = philox([pixel, sample, group, 0], [seed, view])
    return (x >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)


def test_sample_driven_integrator_equals_pixel_sampler():
    """A 1x1 camera with one sample per pixel puts the camera sample at the
    jitter itself, so the primary-sample path built from the pixel sampler's
    own numbers must be the same path, bit for bit."""
    sc = scenes.box_scene(16, 16)
    cam1 = Camera(sc.camera.cop, sc.camera.look_at, sc.camera.up, sc.camera.fov, 1, 1)
    orc = Oracle(sc.flatten())
    depth = 5
    G = 2 * depth + 1
    seeds = range(300)
    x = np.array([[_uniforms(k, 0, 0, 0, g) for g in range(G)] for k in seeds], np.float32)
    got = orc.integrate_samples(cam1, x, max_depth=depth, rr_start=3, termination_prob=0.5)
    want = np.array([orc.render_sample((cam1, 1, 1), 0, 0, 0, mode=MODE_COMPAT, max_depth=depth, rr_start=3,
                                       rr_prob=0.5, seed=k) for k in seeds])
    assert np.array_equal(got, want)
    assert np.count_nonzero(want.sum(axis=1)) > 50


def _cornell(res, eye=(0.5, 0.5, 2.2), look_at=(0.5, 0.5, 0.0)):
    grey = DiffuseMaterial((0.7, 0.7, 0.7))
    walls = [Quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), grey),
             Quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), grey),
             Quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), grey),
             Quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), DiffuseMaterial((0.6, 0.1, 0.1))),
             Quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), DiffuseMaterial((0.1, 0.5, 0.1)))]
    lights = [AreaLight((0.35, 0.999, 0.35), (0.3, 0, 0), (0, 0, 0.3), (12, 12, 12)),
              PointLight((0.5, 0.5, 0.9), (0.05, 0.05, 0.05))]
    cam = Camera(eye, look_at, (0, 1, 0), 0.6, res, res)
    return Scene(cam, lights, IntersectableList(walls))


def test_metropolis_conserves_luminance():
    sc = _cornell(16)
    orc = Oracle(sc.flatten())
    film, res = orc.render_mlt(sc.camera, kind=MLT_METROPOLIS, mutations=32, chains=512, bootstrap=4000, seed=3)
    area = 16 * 16
    assert res["chains"] == 512 and res["mutations"] == 512 * 16 and res["seed_paths"] == 4000
    assert 0 < res["accepted"] < res["mutations"]
    assert _lum(film).sum() == pytest.approx(res["b"] * area, rel=1e-4)


def test_metropolis_converges_to_path_tracing():
    """With termination probability 0 the bootstrap and the chains sample
    full-length paths, so the chains' stationary image is the path-traced
    image of the same estimator (PSSMLT, Kelemen et al. 2002)."""
    sc = _cornell(16)
    orc = Oracle(sc.flatten())
    pt, _ = orc.render((sc.camera, 64, 64), mode=MODE_COMPAT, max_depth=5, rr_start=3, rr_prob=0.0, seed=6)
    blk = lambda im: im.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    errs = []
    for muts in (1024, 8192):
        mlt, res = orc.render_mlt(sc.camera, kind=MLT_METROPOLIS, termination_prob=0.0, mutations=muts, chains=4096,
                                  bootstrap=400000, seed=5)
        rel = np.abs(blk(mlt) - blk(pt)) / blk(pt)
        errs.append(rel.mean())
        print(muts, "block relative errors: max", rel.max(), "mean", rel.mean(), "image mean ratio", mlt.mean() / pt.mean())
        assert mlt.mean() == pytest.approx(pt.mean(), rel=0.005)
    assert rel.max() < 0.05 and rel.mean() < 0.02
    assert errs[1] < 0.7 * errs[0]  # the error shrinks with the chain length


def test_erpt_reproduces_a_constant_background():
    """Every path sees only a constant environment: all values are equal, so
    each initial sample starts one chain and the deposits add up to the
    environment's radiance per pixel."""
    env = EnvironmentMap(np.full((8, 16, 3), 0.5, np.float32), (1.0, 0.8, 0.6))
    cam = Camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 0.8, 12, 10)
    sc = Scene(cam, [env], IntersectableList([]))
    film, res = Oracle(sc.flatten()).render_mlt(cam, kind=MLT_ERPT, mutations=20, erpt_spp=(2, 2), seed=2)
    assert res["seed_paths"] == 120 + 480
    assert res["ed"] == pytest.approx(math.sqrt(0.5 ** 2 + 0.4 ** 2 + 0.3 ** 2) / 20, rel=1e-6)
    assert res["chains"] == res["chain_seeds"] and abs(res["chains"] - 480) <= 2
    assert film.reshape(-1, 3).mean(axis=0) == pytest.approx([0.5, 0.4, 0.3], rel=0.01)


def test_erpt_deposits_e_d_per_chain_step():
    """In a grey scene every deposit points along (1, 1, 1), so norms add:
    the film's summed norm is e_d * mutations / spp per initial sample that
    started chains, whatever the chain count (the reference divides e_d
    among the chains of one sample)."""
    sc = _cornell(12)
    for q in sc.object.items:
        q.material = DiffuseMaterial((0.6, 0.6, 0.6))
    orc = Oracle(sc.flatten())
    M, spp = 8, 4
    film, res = orc.render_mlt(sc.camera, kind=MLT_ERPT, mutations=M, erpt_spp=(2, 2), seed=4)
    assert res["chains"] >= res["chain_seeds"] > 0
    assert res["mutations"] == res["chains"] * M
    total = np.sqrt((film.astype(np.float64) ** 2).sum(axis=2)).sum()
    assert total == pytest.approx(res["chain_seeds"] * M * res["ed"] / spp, rel=1e-4)


def test_symmetric_filter_keeps_flat_regions_and_edges():
    """SymmetricFilter (src/symmetricfilter.cpp): with zero variance a flat
    region is kept, and so is a step edge whose two sides differ far more
    than the noise (every weight across it vanishes), away from the border
    where the zero padding enters the patch distances."""
    from oracle.oracle import sym_filter
    H, W, R, P = 40, 44, 10, 7
    m = R + P  # interior margin
    img = np.full((H, W, 3), 0.7, np.float32)
    img[:, W // 2:] = (0.1, 0.2, 0.3)
    zero = np.zeros_like(img)
    out = sym_filter(img, img, zero, P, R)
    assert np.allclose(out[m:H - m, m:W - m], img[m:H - m, m:W - m], rtol=1e-5, atol=0)   # summation rounding
    # noise of known variance: the filter lowers it where the guide is flat
    rng = np.random.default_rng(1)
    noisy = (0.5 + rng.normal(0, 0.05, (H, W, 3))).astype(np.float32)
    var = np.full_like(noisy, 0.05 ** 2)
    f = sym_filter(noisy, noisy, var, P, R)
    assert f[m:H - m, m:W - m].var() < 0.3 * noisy[m:H - m, m:W - m].var()
    assert f[m:H - m, m:W - m].mean() == pytest.approx(0.5, abs=0.01)


def test_filtered_metropolis_conserves_norm_per_film():
    """MetropolisFltRenderer without filtering, one pass: the reference gives
    each film b * pixels / 2 of summed deposit per pass (deposits of
    b * passes / mutations-per-pixel over mutations-per-pixel * pixels /
    passes / 2 mutations, src/metropolisfltrenderer.cpp:139,158-165), so in a
    grey scene (norms add) the output, the mean of the two films, has summed
    norm b * pixels / 2.  300 chains do not divide the 2048 mutations per
    film: they run 7 steps each (2100), and the total must not follow the
    step count."""
    from renderer_amd._abi import FLT_NO_FILTER, MLT_FILTERED
    sc = _cornell(16)
    for q in sc.object.items:
        q.material = DiffuseMaterial((0.6, 0.6, 0.6))
    orc = Oracle(sc.flatten())
    film, res = orc.render_mlt(sc.camera, kind=MLT_FILTERED, mutations=16, chains=300, bootstrap=4000, seed=2,
                               flt_passes=1, flt_flags=FLT_NO_FILTER)
    total = np.sqrt((film.astype(np.float64) ** 2).sum(axis=2)).sum()
    assert res["mutations"] == 2 * 300 * 7
    assert total == pytest.approx(res["b"] * 16 * 16 / 2, rel=1e-4)


def test_filtered_metropolis_default_passes_run():
    from renderer_amd._abi import MLT_FILTERED
    sc = _cornell(24)
    film, res = Oracle(sc.flatten()).render_mlt(sc.camera, kind=MLT_FILTERED, mutations=16, bootstrap=1000, seed=3)
    assert np.all(np.isfinite(film)) and film.mean() > 0
    assert res["accepted"] > 0
