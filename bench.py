"""Benchmark: plain (unbiased) GPU path tracing of the 1,048,576-triangle mesh
scene at 1024 x 1024, 64 spp (8 x 8 strata), max depth 8, roulette after
depth 3, FP32, on one MI355X.

One step = one full render of that film (67.1M camera paths) with the scene
(BVH, triangles, materials, lights) and the film resident in HBM.  Prints one
JSON line with Mrays/s (camera + extension + shadow rays per second), the
roofline of the closest-hit traversal kernel and a CPU baseline (the oracle,
a restatement of the reference's PerPixelRenderer, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RAY_BYTES = 32 + 16      # ray (origin, tmin, direction, tmax) + hit (t, primitive, u, v)
NODE_BYTES = 64          # one interior node: both child boxes + child words
PRIM_BYTES = 48          # one primitive record


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=8, help="strata per axis (spp = value^2)")
    ap.add_argument("--subdiv", type=int, default=512, help="tetrahedron segments per edge (4 n^2 triangles)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy, the reference point for HBM counters")
    return ap.parse_args()


def cpu_baseline(arrays, scene, budget_s):
    """The oracle (the reference's algorithms: leaf-1 BVH visiting both
    children, per-sample critical section) rendering whole frames of the same
    workload at 1 spp each, new seed per frame, until the time budget is
    spent.  Returns Mrays/s over the frames rendered."""
    from oracle.oracle import Oracle

    threads = os.cpu_count() or 1
    t0 = time.time()
    orc = Oracle(arrays, accel=1)
    build_s = time.time() - t0
    cam = scene.camera
    rays = 0
    frames = 0
    t_render = 0.0
    while t_render < budget_s:
        t1 = time.time()
        _, r = orc.render((cam, 1, 1), mode=1, max_depth=8, rr_start=3, rr_prob=0.5, seed=frames, threads=threads)
        t_render += time.time() - t1
        rays += r
        frames += 1
    orc.close()
    return {"value": round(rays / t_render / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames} frames of {cam.width}x{cam.height} at 1 spp (seeds 0..{frames - 1}), {rays} rays "
                      f"in {t_render:.1f} s on {threads} threads; reference BVH build {build_s:.1f} s not timed"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from renderer_amd import MODE_UNBIASED, GpuContext, scenes

    log("building scene")
    t0 = time.time()
    scene = scenes.mesh_scene(args.res, args.res, n=args.subdiv, seed=1)
    arrays = scene.flatten()
    log(f"scene: {len(arrays.tri_vertices)} triangles in {time.time() - t0:.1f} s")
    ctx = GpuContext(0)
    t0 = time.time()
    ctx.upload(arrays)
    st = ctx.stats()
    log(f"upload + BVH build {time.time() - t0:.1f} s: {st['bvh_nodes']} nodes, depth {st['bvh_depth']}")
    film = torch.zeros((args.res, args.res, 3), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    views = [(scene.camera, args.spp, args.spp)]
    params = dict(mode=MODE_UNBIASED, max_depth=8, rr_start=3, rr_prob=0.5)

    for i in range(args.warmup):
        ctx.render_device(views, [film.data_ptr()], seed=1000 + i, **params)
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    rays = 0
    ext_ms = 0.0
    ext_launches = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render_device(views, [film.data_ptr()], seed=i, **params)
        s = ctx.stats()
        rays += s["camera_rays"] + s["extension_rays"] + s["shadow_rays"]
        ext_ms += s["extend_ms"]
        ext_launches += s["extend_launches"]
        log(f"step {i}: {s['render_ms']:.1f} ms, {(s['camera_rays'] + s['extension_rays'] + s['shadow_rays']) / 1e6:.1f} Mrays")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    last = ctx.stats()

    # algorithmic bytes of the closest-hit launches: counted in a separate,
    # untimed render of the same workload (same seeds as step 0)
    ctx.set_launch_config(count_traversal=1)
    ctx.render_device(views, [film.data_ptr()], seed=0, **params)
    cs = ctx.stats()
    ctx.set_launch_config(count_traversal=0)
    ext_rays = cs["camera_rays"] + cs["extension_rays"]
    alg_bytes = ext_rays * RAY_BYTES + cs["node_visits"] * NODE_BYTES + cs["prim_tests"] * PRIM_BYTES
    per_launch_bytes = alg_bytes / max(1, cs["extend_launches"])
    avg_launch_ms = ext_ms / max(1, ext_launches)
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "extend_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")  # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del a, b

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(arrays, scene, args.cpu_seconds)

    value = rays / dt / 1e6
    out = {
        "metric": "Mrays/s + equal-time relMSE, guided vs plain GPU path tracing, 1 MI355X",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced-sphere mesh scene)",
        "config": {
            "workload": f"plain unbiased PT, displaced sphere {len(arrays.tri_vertices)} tris + ground quad, "
                        f"area + point light, {args.res}x{args.res}, {args.spp * args.spp} spp, max depth 8, RR after 3",
            "triangles": int(len(arrays.tri_vertices)),
            "resolution": [args.res, args.res],
            "spp": args.spp * args.spp,
            "max_depth": 8,
            "paths_per_step": args.res * args.res * args.spp * args.spp,
            "rays_per_step": int(rays / args.steps),
            "path_slots": last["paths_per_pass"],
            "passes": last["passes"],
            "bvh_nodes": last["bvh_nodes"],
            "bvh_depth": last["bvh_depth"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_extend (closest-hit BVH traversal)",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            "avg_launch_ms": round(avg_launch_ms, 4),
            "alg_bytes_per_launch": int(per_launch_bytes),
            "nodes_per_ray": round(cs["node_visits"] / max(1, ext_rays), 2),
            "prims_per_ray": round(cs["prim_tests"] / max(1, ext_rays), 2),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
