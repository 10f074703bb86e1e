"""Traversal counts of the device (rt_trace_count_device,
rt_trace_any_count_device) against the oracle's traversal of the same
exported BVH in the same visiting order (oc_flat_trace): interior nodes
fetched and primitive records tested must agree ray for ray, exactly.  The
algorithmic bytes of the traversal kernels' roofline are these counts times
128 B per node and 48 B per record, so this is what they rest on.  Both
split rules of the tree (median, surface-area heuristic) are checked."""
import numpy as np
import pytest

from oracle.oracle import flat_trace
from renderer_amd import BVH_MEDIAN, BVH_SAH, scenes
from renderer_amd.integrator import bvh_export
from tests.helpers import camera_rays, random_rays
from tests.test_oracle_flat_bvh import _sphere_cloud_scene

pytestmark = pytest.mark.gpu


def _device_counts(ctx, o4, d4, any_hit):
    import torch

    n = len(o4)
    to = torch.from_numpy(np.ascontiguousarray(o4)).cuda()
    td = torch.from_numpy(np.ascontiguousarray(d4)).cuda()
    nodes = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    prims = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    f = ctx.trace_any_count_device if any_hit else ctx.trace_count_device
    f(to.data_ptr(), td.data_ptr(), n, nodes.data_ptr(), prims.data_ptr())
    torch.cuda.synchronize()
    return nodes.cpu().numpy(), prims.cpu().numpy()


def _check(ctx, arr, o4, d4, label):
    for rule in (BVH_MEDIAN, BVH_SAH):
        arr.bvh_build = rule
        _check_tree(ctx, arr, o4, d4, f"{label}, {'SAH' if rule == BVH_SAH else 'median'} tree")
    arr.bvh_build = BVH_MEDIAN


def _check_tree(ctx, arr, o4, d4, label):
    ctx.upload(arr)
    bvh_nodes, recs, _ = bvh_export(arr)
    for any_hit in (False, True):
        print(f"{label}: {len(o4)} rays, any_hit={any_hit}", flush=True)
        dn, dp = _device_counts(ctx, o4, d4, any_hit)
        res = flat_trace(bvh_nodes, recs, o4, d4, any_hit=any_hit, threads=16)
        cn, cp = res[-2], res[-1]
        bad = np.nonzero((dn != cn) | (dp != cp))[0]
        print(f"  nodes/ray {dn.mean():.3f} (oracle {cn.mean():.3f}), records/ray {dp.mean():.3f} "
              f"(oracle {cp.mean():.3f}), mismatches {len(bad)}", flush=True)
        assert len(bad) == 0, [(int(i), int(dn[i]), int(cn[i]), int(dp[i]), int(cp[i])) for i in bad[:8]]
        assert dn.sum() > len(o4)


def test_counts_sphere_cloud(gpu_ctx):
    arr = _sphere_cloud_scene().flatten()
    o4, d4 = random_rays(100000, (-2.5, -1.2, -2.5), (2.5, 3.0, 2.5), seed=21)
    d4[::3, 3] = np.random.default_rng(22).uniform(0.05, 3.0, len(d4[::3]))
    _check(gpu_ctx, arr, o4, d4, "sphere cloud")


def test_counts_mesh_camera_rays(gpu_ctx):
    sc = scenes.mesh_scene(256, 256, n=128, seed=1)
    o4, d4 = camera_rays(sc, spp=1, seed=23)
    _check(gpu_ctx, sc.flatten(), o4, d4, "mesh 131072 triangles, camera rays")


def test_counts_box_scene(gpu_ctx):
    arr = scenes.box_scene(32, 32).flatten()
    o4, d4 = random_rays(50000, (0.02, 0.02, 0.02), (0.98, 0.98, 1.5), seed=24)
    _check(gpu_ctx, arr, o4, d4, "box scene")


@pytest.mark.parametrize("which", ["box", "mesh"])
def test_render_counts_visits_read_from_lds(gpu_ctx, which):
    # the traversal kernels read the tree's first 69 nodes from LDS and count
    # those visits apart (they carry no memory traffic in the roofline):
    # every visit of the box scene's 4-node tree is one, a 131K-triangle
    # mesh's deeper nodes are not
    from renderer_amd import MODE_UNBIASED

    sc = scenes.box_scene(32, 32) if which == "box" else scenes.mesh_scene(64, 64, n=128, seed=1)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.set_launch_config(count_traversal=1)
    try:
        gpu_ctx.render([(sc.camera, 2, 2)], mode=MODE_UNBIASED, max_depth=4, seed=5)
        s = gpu_ctx.stats()
    finally:
        gpu_ctx.set_launch_config(count_traversal=0)
    print(f"{which}: {s['bvh_nodes']} nodes; closest-hit visits {s['node_visits']} ({s['node_visits_lds']} from LDS), "
          f"any-hit {s['shadow_node_visits']} ({s['shadow_node_visits_lds']} from LDS)", flush=True)
    assert s["node_visits"] > 0 and s["shadow_node_visits"] > 0
    if which == "box":
        assert s["node_visits_lds"] == s["node_visits"]
        assert s["shadow_node_visits_lds"] == s["shadow_node_visits"]
    else:
        assert 0 < s["node_visits_lds"] < s["node_visits"]
        assert 0 < s["shadow_node_visits_lds"] < s["shadow_node_visits"]
