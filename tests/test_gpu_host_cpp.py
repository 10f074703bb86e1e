"""The C++ host interface (renderer_amd/host/gpu_integrator.h): the example
program builds the box scene with SceneBuilder and renders it through
GpuUniDiPathTracingIntegrator + PerPixelRenderer; its film must equal the
Python host path's film bit for bit (same scene arrays, same seed)."""
import os
import subprocess

import numpy as np
import pytest

from renderer_amd import MODE_COMPAT, MODE_UNBIASED, scenes
from renderer_amd.build import EXAMPLE

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_cpp_host_matches_python_host(gpu_ctx, tmp_path, mode):
    if not os.path.exists(EXAMPLE):
        pytest.fail("renderer_amd/host/example_box is not built (run __graft_entry__.build())")
    out = tmp_path / "film.raw"
    r = subprocess.run([EXAMPLE, str(out), "48", "40", "2", "3", str(mode), "8", "11"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    cpp = np.fromfile(out, np.float32).reshape(40, 48, 3)
    sc = scenes.box_scene(48, 40)
    gpu_ctx.upload(sc.flatten())
    py = gpu_ctx.render([(sc.camera, 2, 3)], mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=11)[0]
    assert np.array_equal(cpp, py)
    assert cpp.mean() > 0


def test_cpp_scene_file_matches_python_host(gpu_ctx, tmp_path):
    """example_box --scene: the box read from its scene file by the C++ host
    (SceneFile) renders the same film as the Python host's box."""
    if not os.path.exists(EXAMPLE):
        pytest.fail("renderer_amd/host/example_box is not built (run __graft_entry__.build())")
    scene = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "box.scene")
    out = tmp_path / "film.raw"
    r = subprocess.run([EXAMPLE, str(out), "--scene", scene, "1", "2", str(MODE_UNBIASED), "8", "7"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cpp = np.fromfile(out, np.float32).reshape(256, 256, 3)
    sc = scenes.box_scene(256, 256)
    gpu_ctx.upload(sc.flatten())
    py = gpu_ctx.render([(sc.camera, 1, 2)], mode=MODE_UNBIASED, max_depth=8, rr_start=3, rr_prob=0.5, seed=7)[0]
    assert np.array_equal(cpp, py)


@pytest.mark.parametrize("kind", ["metropolis", "energyredist"])
def test_cpp_markov_chain_renderers_match_python_host(gpu_ctx, tmp_path, kind):
    """example_box --mlt: MetropolisRenderer / EnergyRedistributionRenderer of
    the C++ host render the box exactly as the Python host does."""
    from renderer_amd import EnergyRedistributionRenderer, MetropolisRenderer

    if not os.path.exists(EXAMPLE):
        pytest.fail("renderer_amd/host/example_box is not built (run __graft_entry__.build())")
    out = tmp_path / "film.raw"
    muts = 16 if kind == "metropolis" else 12
    r = subprocess.run([EXAMPLE, str(out), "--mlt", kind, "40", "32", str(muts), "9"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    cpp = np.fromfile(out, np.float32).reshape(32, 40, 3)
    sc = scenes.box_scene(40, 32)
    if kind == "metropolis":
        rend = MetropolisRenderer(mutations=muts, seed=9, context=gpu_ctx)
    else:
        rend = EnergyRedistributionRenderer(mutations=muts, seed=9, context=gpu_ctx)
    py = rend.render(sc)
    assert np.array_equal(cpp, py)
    assert cpp.mean() > 0


@pytest.mark.parametrize("mode", [MODE_COMPAT, MODE_UNBIASED])
def test_cpp_per_ray_integrate_matches_python_batch(gpu_ctx, tmp_path, mode):
    """example_box --rays: GpuUniDiPathTracingIntegrator::integrate(ray, key)
    called once per ray from C++ gives, ray for ray, the values of one
    batched rt_integrate_rays call with keys key0 + i (the same stream)."""
    from tests.helpers import random_rays

    if not os.path.exists(EXAMPLE):
        pytest.fail("renderer_amd/host/example_box is not built (run __graft_entry__.build())")
    o4, d4 = random_rays(64, (0.05, 0.05, 0.05), (0.95, 0.95, 1.5), seed=8, normalize=False)
    rays = tmp_path / "rays.raw"
    np.concatenate([o4, d4], 1).astype(np.float32).tofile(rays)
    out = tmp_path / "rgb.raw"
    r = subprocess.run([EXAMPLE, str(out), "--rays", str(rays), str(mode), "8", "500", "13"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cpp = np.fromfile(out, np.float32).reshape(-1, 3)
    gpu_ctx.upload(scenes.box_scene(8, 8).flatten())
    py = gpu_ctx.integrate_rays(o4, d4, mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=13, key0=500)
    assert np.array_equal(cpp, py)
    assert py.sum() > 0
