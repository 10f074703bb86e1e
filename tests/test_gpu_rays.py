"""rt_integrate_rays, the GPU form of UniDiPathTracingIntegrator::integrate(
const Ray&, ...) (src/unidipathtracingintegrator.cpp:34-46), against the
oracle's per-ray entry on the same rays and keys: random rays inside and
around the box scene (axis-parallel ones included, unnormalised
directions, some with a finite range end) and camera rays of the mesh
scene, in both estimators.  Both sides draw the same Philox numbers, so a
ray agrees unless floating-point divergence makes its path part."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_COMPAT, MODE_UNBIASED, scenes
from tests.helpers import camera_rays, random_rays

pytestmark = pytest.mark.gpu


def _agreement(a, b):
    return np.all(np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5, axis=1).mean()


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_box_rays_match_oracle(gpu_ctx, mode):
    sc = scenes.box_scene(32, 32)
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    orc = Oracle(arr)
    o4, d4 = random_rays(40000, (0.02, 0.02, 0.02), (0.98, 0.98, 1.5), seed=3, normalize=False)
    d4[:5000, 3] = np.random.default_rng(4).uniform(0.05, 2.0, 5000)  # finite range ends
    kw = dict(mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=9, key0=1000)
    print(f"integrate_rays: {len(o4)} box rays, mode {mode}", flush=True)
    got = gpu_ctx.integrate_rays(o4, d4, **kw)
    want = orc.integrate_rays(o4, d4, threads=16, **kw)
    agree = _agreement(got, want)
    print(f"agreement {agree:.4f}, means {got.mean(0)} {want.mean(0)}", flush=True)
    assert agree > 0.99
    assert np.allclose(got.mean(0), want.mean(0), rtol=0.02, atol=1e-4)
    # deterministic, and the key addresses the stream
    assert np.array_equal(gpu_ctx.integrate_rays(o4, d4, **kw), got)
    kw2 = dict(kw, key0=1000 + 7000)
    assert np.array_equal(gpu_ctx.integrate_rays(o4[7000:], d4[7000:], **kw2), got[7000:])


def test_mesh_camera_rays_match_oracle(gpu_ctx):
    sc = scenes.mesh_scene(96, 96, n=64, seed=2)
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    orc = Oracle(arr)
    o4, d4 = camera_rays(sc, 2, seed=5)
    for mode in (MODE_COMPAT, MODE_UNBIASED):
        kw = dict(mode=mode, max_depth=8, seed=4, key0=0)
        got = gpu_ctx.integrate_rays(o4, d4, **kw)
        want = orc.integrate_rays(o4, d4, threads=16, **kw)
        agree = _agreement(got, want)
        print(f"mesh mode {mode}: agreement {agree:.4f}", flush=True)
        assert agree > 0.98
        assert np.allclose(got.mean(0), want.mean(0), rtol=0.02)


def test_integrate_rays_rejects_guided_and_bad_parameters(gpu_ctx):
    from renderer_amd import MODE_GUIDED, RtError
    sc = scenes.box_scene(8, 8)
    gpu_ctx.upload(sc.flatten())
    o4, d4 = random_rays(10, (0, 0, 0), (1, 1, 1), seed=1)
    for bad in (dict(mode=MODE_GUIDED), dict(max_depth=0), dict(rr_prob=1.0)):
        with pytest.raises(RtError) as e:
            gpu_ctx.integrate_rays(o4, d4, **bad)
        assert e.value.status == -1
    assert gpu_ctx.integrate_rays(o4[:0], d4[:0]).shape == (0, 3)
