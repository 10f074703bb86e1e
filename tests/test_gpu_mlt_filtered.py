"""The filtered Metropolis renderer (src/metropolisfltrenderer.cpp) and its
symmetric filter (src/symmetricfilter.cpp) on the GPU against the oracle:
the filter on the same images (weights below 0.05 are cut, so a rounding
difference can drop one weight: agreement is counted per pixel), one pass
without filtering chain by chain with its norm conservation, and the
default three passes by their image mean."""
import numpy as np
import pytest

from oracle.oracle import Oracle, sym_filter
from renderer_amd import (FLT_NO_CROSSFILTERING, FLT_NO_FILTER, FLT_NO_VARIANCE_FILTER, FLT_SIMPLE_VARIANCE_FILTER,
                          MLT_FILTERED)
from renderer_amd.scene import DiffuseMaterial
from tests.test_oracle_mlt import _cornell

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("patch,radius,H,W", [(7, 10, 37, 45), (1, 2, 37, 45), (3, 4, 37, 45), (7, 10, 220, 300),
                                                (11, 3, 37, 45), (7, 3, 6, 5)])
def test_symmetric_filter_matches_oracle(gpu_ctx, patch, radius, H, W):
    # 300 x 220: the filter's scratch (16 floats per pixel) is the last
    # region of its buffer, so a short scratch would write 256 KiB past it;
    # patch 11 takes the multi-pass form (blur radius above the one-launch
    # kernel's), 5 x 6 under patch 7 reflects indices more than once
    rng = np.random.default_rng(patch)
    base = np.zeros((H, W, 3), np.float32)
    base[:, W // 3:] = (0.8, 0.5, 0.2)
    img = (base + rng.normal(0, 0.05, (H, W, 3))).astype(np.float32)
    guide = (base + rng.normal(0, 0.05, (H, W, 3))).astype(np.float32)
    var = rng.uniform(0, 0.005, (H, W, 3)).astype(np.float32)
    print(f"symmetric filter {W}x{H} patch {patch} radius {radius}", flush=True)
    got = gpu_ctx.symmetric_filter(img, guide, var, patch, radius)
    want = sym_filter(img, guide, var, patch, radius)
    close = np.all(np.abs(got - want) <= 1e-4 * np.abs(want) + 1e-6, axis=2).mean()
    print(f"agreement {close:.4f}", flush=True)
    assert close > 0.99
    assert np.abs(got - want).max() < 0.05


def test_filtered_one_pass_matches_oracle(gpu_ctx):
    sc = _cornell(24, eye=(0.46, 0.53, 2.2), look_at=(0.52, 0.48, 0.0))
    for q in sc.object.items:
        q.material = DiffuseMaterial((0.6, 0.6, 0.6))
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    kw = dict(kind=MLT_FILTERED, mutations=16, chains=1024, bootstrap=8000, seed=5, flt_passes=1,
              flt_flags=FLT_NO_FILTER)
    print("filtered metropolis, one pass, no filter: 24x24", flush=True)
    img, res = gpu_ctx.render_mlt(sc.camera, **kw)
    ref, rref = Oracle(arr).render_mlt(sc.camera, **kw)
    print("gpu", res, "\noracle", rref, flush=True)
    for k in ("chains", "mutations", "seed_paths"):
        assert res[k] == rref[k], k
    assert res["b"] == pytest.approx(rref["b"], rel=1e-5)
    assert abs(res["accepted"] - rref["accepted"]) <= 0.01 * rref["accepted"]
    d = np.abs(img - ref)
    agree = np.all(d <= 1e-3 * np.abs(ref) + 1e-6, axis=2).mean()
    print(f"pixel agreement {agree:.4f}", flush=True)
    assert agree > 0.95
    # grey scene: norms add; each film receives b * pixels / 2 per pass (the
    # reference's normalisation), however many steps the chains take: 1024
    # chains run 5 steps each, 5120 mutations per film for a nominal 4608
    total = np.sqrt((img.astype(np.float64) ** 2).sum(axis=2)).sum()
    assert res["mutations"] == 2 * 1024 * 5
    assert total == pytest.approx(res["b"] * 24 * 24 / 2, rel=1e-4)
    again, _ = gpu_ctx.render_mlt(sc.camera, **kw)
    assert np.array_equal(img, again)


def test_filtered_default_passes_match_oracle_mean(gpu_ctx):
    sc = _cornell(24)
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    kw = dict(kind=MLT_FILTERED, mutations=32, chains=1024, bootstrap=8000, seed=9)
    print("filtered metropolis, three passes: 24x24", flush=True)
    img, res = gpu_ctx.render_mlt(sc.camera, **kw)
    ref, rref = Oracle(arr).render_mlt(sc.camera, **kw)
    print("gpu", res, "\noracle", rref, flush=True)
    assert res["mutations"] == rref["mutations"] and res["chains"] == rref["chains"]
    assert abs(res["accepted"] - rref["accepted"]) <= 0.01 * rref["accepted"]
    assert np.all(np.isfinite(img))
    assert img.mean() == pytest.approx(ref.mean(), rel=0.05)


def _agreement(img, ref, rtol=1e-3):
    return np.all(np.abs(img - ref) <= rtol * np.abs(ref) + 1e-6, axis=2).mean()


def test_filtered_three_passes_without_filter_match_oracle_pixels(gpu_ctx):
    """Three passes with FLT_NO_FILTER: the importance map of the previous
    pass's relative variance and its normalisation, the rescale of both films
    by summed / new summed importance, b from the first film, and the running
    mean and variance merge, compared pixel by pixel (no weight cut of the
    filter is involved)."""
    sc = _cornell(24, eye=(0.46, 0.53, 2.2), look_at=(0.52, 0.48, 0.0))
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    kw = dict(kind=MLT_FILTERED, mutations=24, chains=1000, bootstrap=8000, seed=11, flt_passes=3,
              flt_flags=FLT_NO_FILTER)
    print("filtered metropolis, three passes, no filter: 24x24", flush=True)
    img, res = gpu_ctx.render_mlt(sc.camera, **kw)
    ref, rref = Oracle(arr).render_mlt(sc.camera, **kw)
    print("gpu", res, "\noracle", rref, flush=True)
    for k in ("chains", "mutations", "seed_paths"):
        assert res[k] == rref[k], k
    assert res["b"] == pytest.approx(rref["b"], rel=1e-5)
    assert abs(res["accepted"] - rref["accepted"]) <= 0.01 * rref["accepted"]
    agree = _agreement(img, ref)
    print(f"pixel agreement {agree:.4f}", flush=True)
    assert agree > 0.95
    assert img.sum() == pytest.approx(ref.sum(), rel=2e-3)


@pytest.mark.parametrize("flag", [FLT_NO_CROSSFILTERING, FLT_NO_VARIANCE_FILTER, FLT_SIMPLE_VARIANCE_FILTER])
def test_filtered_flags_match_oracle(gpu_ctx, flag):
    """Each remaining --metflt-* switch, two passes: the filter paths it
    selects (FLT_SIMPLE_VARIANCE_FILTER runs the separable 3-tap variance
    blur, k_sep3) against the oracle's."""
    sc = _cornell(24, eye=(0.46, 0.53, 2.2), look_at=(0.52, 0.48, 0.0))
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    kw = dict(kind=MLT_FILTERED, mutations=16, chains=1000, bootstrap=8000, seed=13, flt_passes=2, flt_flags=flag)
    print(f"filtered metropolis, flags {flag}: 24x24", flush=True)
    img, res = gpu_ctx.render_mlt(sc.camera, **kw)
    ref, rref = Oracle(arr).render_mlt(sc.camera, **kw)
    print("gpu", res, "\noracle", rref, flush=True)
    assert res["mutations"] == rref["mutations"]
    assert abs(res["accepted"] - rref["accepted"]) <= 0.01 * rref["accepted"]
    assert np.all(np.isfinite(img))
    agree = _agreement(img, ref, 1e-2)
    print(f"pixel agreement {agree:.4f}", flush=True)
    assert agree > 0.9
    assert img.mean() == pytest.approx(ref.mean(), rel=0.01)
