"""rt_trace_closest_device_noalloc, the allocation-free closest-hit entry
point (its traversal-stack overflow is a static device array), against
rt_trace_closest_device on the same rays: primitive IDs and t bit for bit,
on two calls in a row, the second with fewer rays, so that a stale or
shared scratch would show.  The scene is a stack of 262,144 parallel quads
crossed by rays along its axis: every box on the way is hit, so a ray
defers three children per 4-wide level and the stack grows past its LDS
and register entries into the overflow array."""
import ctypes

import numpy as np
import pytest

from renderer_amd import GpuContext, _abi
from renderer_amd.scene import DiffuseMaterial, IntersectableList, Mesh, Scene, Camera

pytestmark = pytest.mark.gpu

RT_STACK_LDS, RT_STACK_REG = 16, 4


def stacked_quads(n=262144):
    z = (np.arange(n, dtype=np.float64) + 0.5) / n
    corners = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64)
    pos = np.zeros((n, 4, 3), np.float32)
    pos[:, :, :2] = corners[None]
    pos[:, :, 2] = z[:, None]
    base = 4 * np.arange(n)[:, None]
    faces = np.concatenate([base + [0, 1, 2], base + [0, 2, 3]]).astype(np.int32)
    mesh = Mesh(pos.reshape(-1, 3), faces, DiffuseMaterial((0.5, 0.5, 0.5)))
    cam = Camera((0.5, 0.5, -2), (0.5, 0.5, 0), (0, 1, 0), 0.5, 8, 8)
    return Scene(cam, [], IntersectableList([mesh]))


def axis_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = np.zeros((n, 4), np.float32)
    d = np.zeros((n, 4), np.float32)
    o[:, :2] = rng.uniform(0.01, 0.99, (n, 2))
    up = rng.random(n) < 0.5
    o[:, 2] = np.where(up, -1.0, 2.0)
    o[:, 3] = 1e-3
    # mostly along the axis, some tilted (still inside the stack's footprint)
    d[:, :2] = rng.normal(0, 1e-3, (n, 2))
    d[:, 2] = np.where(up, 1.0, -1.0)
    d[:, 3] = np.inf
    # a quarter start inside the stack
    k = n // 4
    o[:k, 2] = rng.uniform(0, 1, k)
    return o, d


def test_noalloc_trace_matches_trace_closest_device():
    import torch

    lib = _abi.load()
    arrays = stacked_quads().flatten()
    with GpuContext(0) as ctx:
        ctx.upload(arrays)
        need = ctx.stats()["bvh_stack_need"]
        print(f"stacked quads: {ctx.stats()['bvh_prims']} records, stack need {need}", flush=True)
        assert need > RT_STACK_LDS + RT_STACK_REG
        for n, seed in ((1_000_000, 1), (50_000, 2)):
            o, d = axis_rays(n, seed)
            to = torch.from_numpy(o).cuda()
            td = torch.from_numpy(d).cuda()
            prim_a = torch.full((n,), -7, dtype=torch.int32, device="cuda")
            t_a = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
            prim_b = torch.full((n,), -9, dtype=torch.int32, device="cuda")
            t_b = torch.full((n,), -9.0, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            print(f"noalloc trace: {n} rays (seed {seed})", flush=True)
            st = lib.rt_trace_closest_device(ctx.h, to.data_ptr(), td.data_ptr(), n, prim_a.data_ptr(), t_a.data_ptr())
            assert st == 0, lib.rt_last_error(ctx.h)
            st = lib.rt_trace_closest_device_noalloc(ctx.h, to.data_ptr(), td.data_ptr(), n, prim_b.data_ptr(),
                                                     t_b.data_ptr())
            assert st == 0, lib.rt_last_error(ctx.h)
            pa, pb = prim_a.cpu().numpy(), prim_b.cpu().numpy()
            ta, tb = t_a.cpu().numpy(), t_b.cpu().numpy()
            assert np.array_equal(pa, pb)
            assert np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
            assert (pa >= 0).mean() > 0.99
        # arguments the entry point refuses
        assert lib.rt_trace_closest_device_noalloc(ctx.h, None, None, -1, None, None) == _abi.ERR_INVALID
        assert lib.rt_trace_closest_device_noalloc(ctx.h, to.data_ptr(), td.data_ptr(), (1 << 31) + 1,
                                                   prim_b.data_ptr(), t_b.data_ptr()) == _abi.ERR_INVALID
        assert lib.rt_trace_closest_device_noalloc(ctx.h, None, None, 0, None, None) == 0
