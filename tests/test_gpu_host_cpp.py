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
