"""Benchmark: plain (unbiased) GPU path tracing of the 1,048,576-triangle mesh
scene at 1024 x 1024, 64 spp (8 x 8 strata), max depth 8, roulette after
depth 3, FP32, on one MI355X.

One step = one full render of that film (67.1M camera paths) with the scene
(BVH, triangles, materials, lights) and the film resident in HBM.  Prints one
JSON line with Mrays/s (camera + extension + shadow rays per second), the
roofline of the closest-hit traversal kernel and a CPU baseline (the oracle,
a restatement of the reference's PerPixelRenderer, on a bounded sample).

Beside the timed steps (not part of `value`): the 32-view batched render
with autotuned launch configuration, and two equal-time comparisons of
guided vs plain PT (occluded-light room; mesh under a sky with a sun), each
with the guiding network's FP16 MFMA throughput.
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
SHADOW_RAY_BYTES = 32 + 16  # ray + its contribution record (rgb, path slot)
NODE_BYTES = 128         # one 4-wide node: four child boxes + child words
# node visits to the tree's first nodes read the traversal kernels' LDS copy
# of them: no memory traffic, so they count no bytes
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
    ap.add_argument("--no-guided", action="store_true", help="skip the equal-time guided vs plain comparisons")
    ap.add_argument("--no-env", action="store_true", help="skip the env-lit mesh comparison")
    ap.add_argument("--no-room", action="store_true", help="skip the occluded-light room comparison")
    ap.add_argument("--no-sah", action="store_true", help="skip the surface-area-heuristic tree run")
    ap.add_argument("--guide-bound", action="store_true",
                    help="also price the first guided bounce in the room against ideal proposals")
    ap.add_argument("--bound-points", type=int, default=64)
    ap.add_argument("--bound-grid", type=int, default=32, help="hemisphere bins per axis")
    ap.add_argument("--bound-rays", type=int, default=128, help="rays per bin")
    ap.add_argument("--no-batched", action="store_true", help="skip the 32-view batched render")
    ap.add_argument("--no-scene-file", action="store_true", help="skip the scene-file front-end timing")
    ap.add_argument("--no-film", action="store_true", help="skip the tonemapper and film-writer timing")
    ap.add_argument("--no-mlt", action="store_true", help="skip the Metropolis and ERPT renders")
    ap.add_argument("--room-res", type=str, default="1920x1080")
    ap.add_argument("--room-subdiv", type=int, default=256,
                    help="octahedron segments per edge of the 8 objects (256: 4,194,304 triangles)")
    ap.add_argument("--room-seconds", type=float, default=20.0, help="wall-time budget of each estimator")
    ap.add_argument("--room-ref-spp", type=int, default=4096)
    ap.add_argument("--room-pass-spp", type=int, default=16)
    ap.add_argument("--room-train-passes", type=int, default=16, help="guided passes that train the network")
    ap.add_argument("--room-epochs", type=int, default=4, help="training epochs over each pass's records")
    ap.add_argument("--room-lr", type=float, default=1e-3)
    ap.add_argument("--env-res", type=int, default=1024)
    ap.add_argument("--env-seconds", type=float, default=1.5)
    ap.add_argument("--env-ref-spp", type=int, default=4096)
    ap.add_argument("--env-pass-spp", type=int, default=16)
    ap.add_argument("--env-train-passes", type=int, default=8)
    ap.add_argument("--env-epochs", type=int, default=4)
    ap.add_argument("--env-lr", type=float, default=1e-3)
    return ap.parse_args()


def count_depth(stats, max_depth):
    """Depth of a counting render that traces the bounces the per-bounce
    closest-hit and any-hit launches of a timed render (its stats) traced:
    those before the bounce its passes moved to the finishing launch from
    (the counting render runs every bounce in per-bounce launches; paths are
    the same whatever the render's depth)."""
    fb = stats["finish_bounce"]
    return min(max_depth, fb) if fb > 0 else max_depth


def rel_mse(img, ref, eps=1e-2):
    """mean over pixels and channels of (I - R)^2 / (R^2 + eps)."""
    return float(((img - ref) ** 2 / (ref ** 2 + eps)).mean())


def batched_views(args, scene, arrays):
    """32 seeded views of the mesh scene (resolutions from {256^2, 512x384,
    1024x768, 1920x1080}, spp uniform in [8, 128], eyes on a sphere of radius
    3) traced together in one rt_render pipeline, with the launch
    configuration picked by the autotuner.  Returns Mrays/s and the choice."""
    import torch

    from renderer_amd import MODE_UNBIASED, GpuContext, scenes
    from renderer_amd.autotune import autotune

    views = scenes.batched_views(32, seed=7)
    ctx = GpuContext(0)
    ctx.upload(arrays)
    params = dict(mode=MODE_UNBIASED, max_depth=8, rr_start=3, rr_prob=0.5)
    t0 = time.time()
    tune = autotune(ctx, views, params, spp_divisor=16, repeats=1)
    tune_s = time.time() - t0
    films = [torch.zeros((c.height, c.width, 3), dtype=torch.float32, device="cuda:0") for c, _, _ in views]
    torch.cuda.synchronize()
    ctx.render_device(views, [f.data_ptr() for f in films], seed=1, **params)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.render_device(views, [f.data_ptr() for f in films], seed=2, **params)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = ctx.stats()
    # the same render with each pass replayed from one captured hipGraph
    eager = [f.clone() for f in films]
    for f in films:
        f.zero_()
    torch.cuda.synchronize()
    ctx.set_launch_config(path_slots=tune["path_slots"], blocks_per_cu=tune["blocks_per_cu"], use_graph=1)
    ctx.render_device(views, [f.data_ptr() for f in films], seed=1, **params)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.render_device(views, [f.data_ptr() for f in films], seed=2, **params)
    torch.cuda.synchronize()
    dt_graph = time.perf_counter() - t0
    same = all(bool(torch.equal(a, b)) for a, b in zip(eager, films))
    ctx.close()
    rays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
    return {
        "views": len(views),
        "pixels": int(sum(c.width * c.height for c, _, _ in views)),
        "paths": int(sum(c.width * c.height * sx * sy for c, sx, sy in views)),
        "seconds": round(dt, 3),
        "mrays_per_s": round(rays / dt / 1e6, 2),
        "rays": int(rays),
        "passes": st["passes"],
        "autotuned": {"path_slots": tune["path_slots"], "blocks_per_cu": tune["blocks_per_cu"],
                      "trials": len(tune["trials"]), "tuning_seconds": round(tune_s, 1)},
        "graph": {"seconds": round(dt_graph, 3), "mrays_per_s": round(rays / dt_graph / 1e6, 2),
                  "films_equal_eager": same},
    }


MLP_FWD_FLOP = 2 * (80 * 64 + 64 * 64 + 64 * 64 + 64 * 48)        # per sample: 32,768
MLP_TRAIN_FLOP = MLP_FWD_FLOP + 2 * (48 * 64 + 64 * 64 + 64 * 64) + 2 * (80 * 64 + 64 * 64 + 64 * 64 + 48 * 64)
MFMA_FP16_PEAK_TFLOPS = 2500.0


def equal_time(scene, desc, max_depth, budget_s, ref_spp, pass_spp, train_passes, epochs, lr):
    """Plain unbiased PT and guided PT of one scene each get the same
    wall-time budget; the guided budget includes the online training of its
    network, which trains during its first `train_passes` passes and then
    only samples.  Every pass counts towards the image.  Both are scored by
    relMSE against a high-sample plain render after every pass; scoring is
    not timed.  Also reports the network kernels' FP16 MFMA throughput from
    the library's own timers.  Returns the comparison as a dict."""
    import numpy as np
    import torch

    from renderer_amd import MODE_GUIDED, MODE_UNBIASED, GpuContext

    W, H = scene.camera.width, scene.camera.height
    arrays = scene.flatten()
    ctx = GpuContext(0)
    ctx.upload(arrays)
    ctx.set_launch_config(guide_epochs=epochs)
    common = dict(max_depth=max_depth, rr_start=3, rr_prob=0.5)
    dev = "cuda:0"
    net = dict(infer_ms=0.0, infer_samples=0, train_ms=0.0, train_samples=0, train_steps=0)

    def render(spp, seed, mode, train=0):
        f = torch.zeros((H, W, 3), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        sx = int(np.sqrt(spp))
        ctx.render_device([(scene.camera, sx, spp // sx)], [f.data_ptr()], seed=seed, mode=mode, guide_train=train,
                          guide_lr=lr, **common)
        if mode == MODE_GUIDED:
            s = ctx.stats()
            for k in net:
                net[k] += s["guide_" + k]
        return f

    # reference: plain PT at ref spp, rendered in K passes.  Its own noise
    # adds to every measured relMSE, E[(I - R)^2] = Var I + Var R, which
    # pulls the guided/plain ratio towards 1 when the compared images are not
    # far below the reference's spp; the spread of the K passes estimates
    # Var R, and the corrected relMSE subtracts its relMSE share.
    t0 = time.time()
    ref = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
    ref2 = torch.zeros_like(ref)
    per = min(256, ref_spp)
    K = max(2, ref_spp // per)
    for k in range(K):
        f = render(per, 100000 + k, MODE_UNBIASED).double()
        ref += f
        ref2 += f * f
    ref /= K
    var_ref = torch.clamp(ref2 - K * ref * ref, min=0.0) / (K * (K - 1))
    ref = ref.cpu().numpy()
    ref_noise = float((var_ref.cpu().numpy() / (ref ** 2 + 1e-2)).mean())
    log(f"reference {K * per} spp in {time.time() - t0:.1f} s (its relMSE share {ref_noise:.3g})")

    def progressive(mode, budget, training):
        acc = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
        n = 0
        curve = []
        # every pass is an unbiased estimate (the guided one for any network
        # state), so passes rendered while the network trains count too
        el = 0.0  # render (and training) time only
        while True:
            torch.cuda.synchronize()
            t_start = time.perf_counter()
            f = render(pass_spp, n, mode, train=1 if n < training else 0)
            acc += f.double()
            torch.cuda.synchronize()
            el += time.perf_counter() - t_start
            n += 1
            curve.append((el, n * pass_spp, rel_mse((acc / n).cpu().numpy(), ref) - ref_noise))
            if el >= budget:
                break
        return curve

    plain = progressive(MODE_UNBIASED, budget_s, 0)
    ctx.guide_init(7)
    guided = progressive(MODE_GUIDED, budget_s, train_passes)
    # the guided renderer with plain PT as fallback and error judge: after
    # training, six passes of each estimator decide which one renders the
    # rest of the budget (its own seeds; the probe is on the clock)
    from renderer_amd.progressive import GuidedProgressive
    ctx.guide_init(7)
    adaptive = []
    gp = GuidedProgressive(ctx, scene.camera, pass_spp=pass_spp, train_passes=train_passes, probe_passes=6, lr=lr,
                           **common)
    # network work of the fallback run too, so that kernel-trace totals of
    # the whole section can be divided by the vertices and records behind them
    net_fb = dict(infer_samples=0, train_samples=0, train_steps=0)

    def on_pass(t, spp, img):
        adaptive.append((t, spp, rel_mse(img.cpu().numpy(), ref) - ref_noise))
        s = ctx.stats()
        for k in net_fb:
            net_fb[k] += s["guide_" + k]

    _, ainfo = gp.run(budget_s, seed0=50000, on_pass=on_pass)
    ctx.close()
    rp = plain[-1][2]
    # time at which guided first reaches the plain estimator's final relMSE
    t_reach = next((t for t, _, e in guided if e <= rp), None)
    t_reach_a = next((t for t, _, e in adaptive if e <= rp), None)

    def tflops(flop, ms):
        return round(flop / (ms * 1e-3) / 1e12, 2) if ms > 0 else None

    inf_tf = tflops(net["infer_samples"] * MLP_FWD_FLOP, net["infer_ms"])
    trn_tf = tflops(net["train_samples"] * MLP_TRAIN_FLOP, net["train_ms"])
    return {
        "scene": desc,
        "budget_s": budget_s,
        "reference_spp": K * per,
        "reference_relmse_share": round(ref_noise, 6),
        "relmse": "against the reference, minus the reference's own share (estimated from its passes)",
        "pass_spp": pass_spp,
        "plain": {"relmse": round(rp, 6), "spp": plain[-1][1], "seconds": round(plain[-1][0], 3)},
        "guided": {"relmse": round(guided[-1][2], 6) if guided else None, "spp": guided[-1][1] if guided else 0,
                   "seconds": round(guided[-1][0], 3) if guided else None,
                   "training_passes": train_passes, "epochs": epochs, "lr": lr},
        "relmse_ratio_guided_over_plain": round(guided[-1][2] / rp, 4) if guided else None,
        "time_to_plain_relmse_s": round(t_reach, 3) if t_reach else None,
        "speedup_to_equal_relmse": round(plain[-1][0] / t_reach, 3) if t_reach else None,
        "guided_with_fallback": {
            "relmse": round(adaptive[-1][2], 6), "spp": adaptive[-1][1], "seconds": round(adaptive[-1][0], 3),
            "choice": ainfo["choice"],
            "probe": {k: round(v, 6) for k, v in ainfo.get("probe", {}).items()},
            "relmse_ratio_over_plain": round(adaptive[-1][2] / rp, 4),
            "speedup_to_equal_relmse": round(plain[-1][0] / t_reach_a, 3) if t_reach_a else None,
        },
        "network": {
            "infer_ms": round(net["infer_ms"], 2), "infer_samples": int(net["infer_samples"]),
            "infer_tflops": inf_tf,
            "train_ms": round(net["train_ms"], 2), "train_samples": int(net["train_samples"]),
            "train_steps": int(net["train_steps"]), "train_tflops": trn_tf,
            "flop_per_sample": {"forward": MLP_FWD_FLOP, "train": MLP_TRAIN_FLOP},
            "mfma_peak_tflops": MFMA_FP16_PEAK_TFLOPS,
            "infer_mfma_frac": round(inf_tf / MFMA_FP16_PEAK_TFLOPS, 4) if inf_tf else None,
            "train_mfma_frac": round(trn_tf / MFMA_FP16_PEAK_TFLOPS, 4) if trn_tf else None,
            "with_fallback_run": {k: int(net[k] + net_fb[k]) for k in net_fb},
        },
    }


def equal_time_room(args):
    """Occluded-light room: the only emitter outside, behind a door gap of 5%
    of the wall width; max depth 16."""
    from renderer_amd import scenes

    W, H = (int(x) for x in args.room_res.split("x"))
    t0 = time.time()
    scene = scenes.room_scene(W, H, n=args.room_subdiv, seed=3)
    ntri = 8 * 8 * args.room_subdiv ** 2 + 24
    log(f"room: {ntri} triangles, {W}x{H}, built in {time.time() - t0:.1f} s")
    desc = (f"occluded-light room, 12 quads + 8 displaced meshes ({ntri} triangles), {W}x{H}, max depth 16, "
            f"light outside behind a 5% door gap")
    return equal_time(scene, desc, 16, args.room_seconds, args.room_ref_spp, args.room_pass_spp,
                      args.room_train_passes, args.room_epochs, args.room_lr)


def sah_tree(ctx, arrays, views, params, film, args):
    """The headline workload (its median-split tree is the one the workload
    names) on a surface-area-heuristic tree of the same scene: same rays and
    image up to ties, less traversal work.  Untimed warm-up, then the same
    number of timed renders; nodes and records per ray from a counting
    render.  The context goes back to the median tree afterwards."""
    import torch

    from renderer_amd import BVH_MEDIAN, BVH_SAH

    t0 = time.time()
    arrays.bvh_build = BVH_SAH
    ctx.upload(arrays)
    st = ctx.stats()
    build_s = time.time() - t0
    for i in range(args.warmup):
        ctx.render_device(views, [film.data_ptr()], seed=1000 + i, **params)
    torch.cuda.synchronize()
    rays = ext_ms = sh_ms = launches = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render_device(views, [film.data_ptr()], seed=i, **params)
        s = ctx.stats()
        if i == 0:
            first = s
        rays += s["camera_rays"] + s["extension_rays"] + s["shadow_rays"]
        ext_ms += s["extend_ms"]
        sh_ms += s["shadow_ms"]
        launches += s["extend_launches"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_launch_config(count_traversal=1)
    ctx.render_device(views, [film.data_ptr()], seed=0, **dict(params, max_depth=count_depth(first, params["max_depth"])))
    cs = ctx.stats()
    ctx.set_launch_config(count_traversal=0)
    ext_rays = cs["camera_rays"] + cs["extension_rays"]
    out = {
        "mrays_per_s": round(rays / dt / 1e6, 2),
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "bvh_nodes": st["bvh_nodes"], "bvh_depth": st["bvh_depth"], "stack_need": st["bvh_stack_need"],
        "build_s": round(build_s, 2),
        "extend_avg_launch_ms": round(ext_ms / max(1, launches), 4),
        "shadow_avg_launch_ms": round(sh_ms / max(1, launches), 4),
        "nodes_per_ray": round(cs["node_visits"] / max(1, ext_rays), 2),
        "prims_per_ray": round(cs["prim_tests"] / max(1, ext_rays), 2),
        "shadow_nodes_per_ray": round(cs["shadow_node_visits"] / max(1, cs["shadow_rays"]), 2),
        "shadow_prims_per_ray": round(cs["shadow_prim_tests"] / max(1, cs["shadow_rays"]), 2),
    }
    arrays.bvh_build = BVH_MEDIAN
    ctx.upload(arrays)
    return out


def _vmf(mu, kappa, w):
    """vMF densities of directions w (..., 3) for lobes mu (L, 3), kappa (L):
    (..., L)."""
    import numpy as np

    k = np.maximum(kappa, 1e-6)
    c = k / (2 * np.pi * -np.expm1(-2 * k))
    return c * np.exp(k * (w @ mu.T - 1.0))


def _fit_vmf(dirs, q, lobes=8, iters=60, seed=0):
    """Weighted EM for a mixture of `lobes` vMF lobes on directions `dirs`
    with masses q (Banerjee et al. 2005, kappa by their approximation):
    the best an 8-lobe mixture does on a tabulated target."""
    import numpy as np

    rng = np.random.default_rng(seed)
    q = q / q.sum()
    mu = dirs[rng.choice(len(dirs), lobes, replace=False, p=q)]
    kappa = np.full(lobes, 10.0)
    w = np.full(lobes, 1.0 / lobes)
    for _ in range(iters):
        r = w * _vmf(mu, kappa, dirs) + 1e-300
        r /= r.sum(1, keepdims=True)
        rq = r * q[:, None]
        w = rq.sum(0) + 1e-12
        R = rq.T @ dirs / w[:, None]
        rl = np.clip(np.linalg.norm(R, axis=1), 1e-6, 1 - 1e-6)
        mu = R / rl[:, None]
        kappa = np.clip(rl * (3 - rl * rl) / (1 - rl * rl), 0.05, 2e4)
        w /= w.sum()
    return mu, kappa, w


def guide_bound(args):
    """How much better the first bounce could be sampled in the occluded-light
    room, at camera-visible points of its walls, floor and ceiling (diffuse,
    albedo 0.75): the incoming radiance over a 32 x 32 grid of
    cosine-weighted hemisphere bins, 128 jittered rays per bin (plain PT from
    there, rt_integrate_rays), gives the second moment of a one-sample
    estimate of the reflected radiance for any proposal that is piecewise
    cosine-shaped over the bins.  Reported as the gain in relMSE (variance /
    mean^2, summed over the points) over cosine sampling, for: the
    variance-optimal proposal (bin mass ~ sqrt of the second moment, fitted on
    half the rays, scored on the other half); it mixed with 5% cosine (the
    guided estimator's floor); the radiance-proportional proposal the
    training loss aims at, mixed the same way; the best 8-lobe vMF mixture
    of that target (weighted EM), mixed; and the trained network itself at
    the same points (its own mixing weight).  The network is trained as in
    the equal-time comparison.  The rest of the path stays plain PT, so this
    prices the first guided bounce only."""
    import numpy as np
    import torch

    from renderer_amd import LIGHT_AREA, MODE_GUIDED, MODE_UNBIASED, GpuContext, scenes

    W, H = (int(x) for x in args.room_res.split("x"))
    scene = scenes.room_scene(W, H, n=args.room_subdiv, seed=3)
    arrays = scene.flatten()
    ctx = GpuContext(0)
    ctx.upload(arrays)
    ctx.set_launch_config(guide_epochs=args.room_epochs)
    ctx.guide_init(7)
    common = dict(max_depth=16, rr_start=3, rr_prob=0.5)
    film = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    t0 = time.time()
    for k in range(args.room_train_passes):
        torch.cuda.synchronize()
        ctx.render_device([(scene.camera, 4, 4)], [film.data_ptr()], seed=k, mode=MODE_GUIDED, guide_train=1,
                          guide_lr=args.room_lr, **common)
    torch.cuda.synchronize()
    log(f"guide trained in {time.time() - t0:.1f} s")

    # camera rays of seeded pixels; keep hits on the 12 room quads
    cam = scene.camera
    rng = np.random.default_rng(11)
    w_ = cam.cop - cam.look_at
    w_ = w_ / np.linalg.norm(w_)
    u_ = np.cross(cam.up, w_)
    u_ = u_ / np.linalg.norm(u_)
    v_ = np.cross(w_, u_)
    th = np.tan(cam.fov / 2)
    m = 4096
    px, py = rng.random(m) * W, rng.random(m) * H
    dcam = ((-W / H * th + 2 * W / H * th * px / W)[:, None] * u_ + (th - 2 * th * py / H)[:, None] * v_ - w_)
    dcam /= np.linalg.norm(dcam, axis=1, keepdims=True)
    o4 = np.concatenate([np.tile(cam.cop, (m, 1)), np.full((m, 1), 1e-3)], 1).astype(np.float32)
    d4 = np.concatenate([dcam, np.full((m, 1), np.inf)], 1).astype(np.float32)
    prim, t = ctx.trace_closest(o4, d4)
    n_lights = sum(1 for l in arrays.lights if l.type == LIGHT_AREA)
    quad_ids = set(int(p) for p in np.unique(arrays.tri_prim)[:12])
    P = np.asarray(arrays.positions, np.float64).reshape(-1, 3)
    TV = np.asarray(arrays.tri_vertices).reshape(-1, 3)
    normal_of = {}
    for pid in quad_ids:
        tri = TV[np.nonzero(arrays.tri_prim == pid)[0][0]]
        nn = np.cross(P[tri[1]] - P[tri[0]], P[tri[2]] - P[tri[0]])
        normal_of[pid] = nn / np.linalg.norm(nn)
    pts = [i for i in range(m) if int(prim[i]) in quad_ids][: args.bound_points]
    del n_lights

    G, S = args.bound_grid, args.bound_rays  # bins per axis, rays per bin
    B = G * G
    gu, gv = np.meshgrid(np.arange(G), np.arange(G), indexing="ij")
    ju = (gu.reshape(-1, 1) + rng.random((B, S))) / G
    jv = (gv.reshape(-1, 1) + rng.random((B, S))) / G
    sin_t = np.sqrt(ju)                # cosine-weighted: equal cosine measure per bin
    cos_t = np.sqrt(1 - ju)
    phi = 2 * np.pi * jv
    local = np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], -1)  # (B, S, 3)
    rho = 0.75
    names = ["ideal", "ideal_mix5", "radiance_mix5", "vmf8_mix5", "network"]
    num = dict.fromkeys(["cosine"] + names, 0.0)
    per_point = []
    recs = np.zeros((len(pts), 16), np.float32)
    frames = []
    for j, i in enumerate(pts):
        x = o4[i, :3].astype(np.float64) + t[i] * dcam[i]
        nrm = normal_of[int(prim[i])]
        if nrm @ dcam[i] > 0:
            nrm = -nrm
        a = np.array([1.0, 0.0, 0.0]) if abs(nrm[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        tx = np.cross(a, nrm)
        tx /= np.linalg.norm(tx)
        ty = np.cross(nrm, tx)
        frames.append((x, nrm, tx, ty))
        recs[j, 0:3], recs[j, 3:6], recs[j, 6:9] = x, nrm, -dcam[i]
    head = ctx.guide_head(recs) if len(pts) else np.zeros((0, 48), np.float32)
    t0 = time.time()
    for j, (x, nrm, tx, ty) in enumerate(frames):
        dirs = local[..., 0:1] * tx + local[..., 1:2] * ty + local[..., 2:3] * nrm  # (B, S, 3)
        rays_o = np.concatenate([np.tile(x + 1e-4 * nrm, (B * S, 1)), np.full((B * S, 1), 1e-4)], 1)
        rays_d = np.concatenate([dirs.reshape(-1, 3), np.full((B * S, 1), np.inf)], 1)
        rgb = ctx.integrate_rays(rays_o.astype(np.float32), rays_d.astype(np.float32), key0=j * B * S,
                                 mode=MODE_UNBIASED, seed=99, max_depth=15, rr_start=2, rr_prob=0.5)
        L = (rgb @ np.array([0.299, 0.587, 0.114])).reshape(B, S).astype(np.float64)
        mean = L.mean(1)
        m2a, m2b = (L[:, 0::2] ** 2).mean(1), (L[:, 1::2] ** 2).mean(1)
        I = rho * mean.mean()
        if not I > 0:
            continue

        def rel_var(q):
            # one-sample second moment: sum_b q_b (rho L / (B q_b))^2
            return (rho * rho / (B * B) * np.sum(m2b / np.maximum(q, 1e-30)) - I * I) / (I * I)

        cosine = np.full(B, 1.0 / B)
        ideal = np.sqrt(m2a) + 1e-9 * np.sqrt(m2a).mean()
        ideal /= ideal.sum()
        mean_a = L[:, 0::2].mean(1)
        target = np.maximum(mean_a, 0) + 1e-9 * mean_a.mean()
        target /= target.sum()
        cdir = local.mean(1)
        cdir = cdir[:, 0:1] * tx + cdir[:, 1:2] * ty + cdir[:, 2:3] * nrm
        cdir /= np.linalg.norm(cdir, axis=1, keepdims=True)
        mu, kap, wt = _fit_vmf(cdir, target, seed=j)
        # bin mass of a density: sum over the bin's rays of pdf x solid angle
        omega = np.pi / (B * S) / np.maximum(local[..., 2], 1e-6)  # solid angle per ray
        vmf_q = (_vmf(mu, kap, dirs.reshape(-1, 3)) @ wt).reshape(B, S)
        vmf_q = (vmf_q * omega).sum(1)
        h = head[j]
        mu_n = h[0:40].reshape(8, 5)[:, 0:3].astype(np.float64)
        mu_n /= np.maximum(np.linalg.norm(mu_n, axis=1, keepdims=True), 1e-12)
        k_n = h[0:40].reshape(8, 5)[:, 3].astype(np.float64)
        w_n = h[0:40].reshape(8, 5)[:, 4].astype(np.float64)
        w_n = w_n / max(w_n.sum(), 1e-12)
        alpha = float(h[40])
        net_q = ((_vmf(mu_n, k_n, dirs.reshape(-1, 3)) @ w_n).reshape(B, S) * omega).sum(1)
        qs = {"cosine": cosine, "ideal": ideal, "ideal_mix5": 0.05 * cosine + 0.95 * ideal,
              "radiance_mix5": 0.05 * cosine + 0.95 * target, "vmf8_mix5": 0.05 * cosine + 0.95 * vmf_q,
              "network": (1 - alpha) * cosine + alpha * net_q}
        rv = {k: rel_var(q) for k, q in qs.items()}
        for k in num:
            num[k] += rv[k]
        per_point.append({"relvar_cosine": round(rv["cosine"], 4), "alpha": round(alpha, 3),
                          **{f"gain_{k}": round(rv["cosine"] / rv[k], 3) for k in names}})
    ctx.close()
    log(f"guide bound: {len(per_point)} points in {time.time() - t0:.1f} s")
    return {
        "scene": f"occluded-light room {W}x{H}, {8 * 8 * args.room_subdiv ** 2 + 24} triangles",
        "points": len(per_point),
        "bins": B, "rays_per_bin": S,
        "training": {"passes": args.room_train_passes, "spp_per_pass": 16, "epochs": args.room_epochs,
                     "lr": args.room_lr},
        "relmse_gain_over_cosine": {k: round(num["cosine"] / num[k], 3) for k in names},
        "median_point_gain": {k: round(float(np.median([pp[f"gain_{k}"] for pp in per_point])), 3) for k in names},
        "per_point": per_point[:8],
    }


def equal_time_env(args):
    """The 1,048,576-triangle mesh lit only by the procedural sky (512 x 256
    lat-long, sun disk of 0.5 degrees at radiance 5e4); max depth 8."""
    from renderer_amd import scenes

    t0 = time.time()
    scene = scenes.mesh_scene(args.env_res, args.env_res, n=args.subdiv, seed=1, env=True)
    log(f"env mesh scene built in {time.time() - t0:.1f} s")
    desc = (f"displaced sphere {4 * args.subdiv ** 2} tris + ground, lat-long sky 512x256 with a 0.5 deg sun at 5e4, "
            f"{args.env_res}x{args.env_res}, max depth 8")
    return equal_time(scene, desc, 8, args.env_seconds, args.env_ref_spp, args.env_pass_spp,
                      args.env_train_passes, args.env_epochs, args.env_lr)


def scene_file_load(args):
    """The scene-language front end on the benchmark mesh: the displaced
    sphere written as an OBJ file, read through obj(...) of a scene file
    (SceneGrammar::parse + resolveVars), evaluated into the scene description
    (buildScene), and dumped as --dump-bvh text.  Host work only."""
    import shutil
    import tempfile

    import numpy as np

    from renderer_amd import scenes
    from renderer_amd.scene_file import SceneFile

    P, F = scenes.displaced_sphere("tetra", args.subdiv, 0.1, 1)
    tmp = tempfile.mkdtemp(prefix="scene_file_")
    try:
        obj = os.path.join(tmp, "sphere.obj")
        with open(obj, "w") as f:
            f.write("".join(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in P.tolist()))
            f.write("".join(f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in F.tolist()))
        with open(os.path.join(tmp, "mesh.scene"), "w") as f:
            f.write("camera = camera([0 0.9 3.6], [0 -0.05 0], [0 1 0], 0.6981317008, 1024, 1024);\n"
                    "lights = lights{ arealight([-0.75 3 -0.75], [1.5 0 0], [0 0 1.5], [6 6 6]) "
                    "pointlight([2.5 2.5 2], [8 8 8]) };\n"
                    "intersectable = intersectables{ obj(\"sphere.obj\", phong([0.6 0.6 0.6], [0.3 0.3 0.3], 50)) "
                    "quad([-6 -1.15 -6], [-6 -1.15 6], [6 -1.15 6], [6 -1.15 -6], diffuse([0.5 0.5 0.5])) };\n")
        mb = os.path.getsize(obj) / 1e6
        t0 = time.perf_counter()
        sf = SceneFile.load(os.path.join(tmp, "mesh.scene"))
        t1 = time.perf_counter()
        a = sf.arrays()
        t2 = time.perf_counter()
        sf.dump_bvh(os.path.join(tmp, "mesh.bvh"))
        t3 = time.perf_counter()
        dump_mb = os.path.getsize(os.path.join(tmp, "mesh.bvh")) / 1e6
        same = bool(np.array_equal(a.positions.reshape(-1, 3, 3)[:len(F)], P[F]))
        sf.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {
        "workload": f"displaced sphere as OBJ ({len(F)} triangles, {mb:.1f} MB) + ground quad, 2 lights",
        "triangles": int(len(F)),
        "load_s": round(t1 - t0, 3),
        "evaluate_s": round(t2 - t1, 3),
        "obj_mb_per_s": round(mb / (t1 - t0), 1),
        "mtris_per_s": round(len(F) / (t2 - t0) / 1e6, 2),
        "dump_bvh_s": round(t3 - t2, 3),
        "dump_bvh_mb": round(dump_mb, 1),
        "geometry_equals_generator": same,
        "cores": 1,
    }


def film_outputs(ctx, film):
    """Film outputs: the tonemapper kernel on a large resident film (its HBM
    roofline: 12 B read + 3 B written per pixel), then the three writers on
    the benchmark film (host work)."""
    import shutil
    import tempfile

    import torch

    from renderer_amd.film import write_exr, write_image, write_mat

    side = 8192
    big = torch.rand((side, side, 3), dtype=torch.float32, device="cuda:0") * 1.5
    out = torch.empty((side, side, 3), dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    ms = []
    for _ in range(6):
        ctx.tonemap_device(big.data_ptr(), side, side, out.data_ptr(), 2.2)
        ms.append(ctx.stats()["tonemap_ms"])
    t = sorted(ms[1:])[len(ms[1:]) // 2]
    nbytes = side * side * 3 * 5
    gbs = nbytes / (t * 1e-3) / 1e9
    del big, out
    h, w = film.shape[0], film.shape[1]
    f = film.cpu().numpy()
    tmp = tempfile.mkdtemp(prefix="film_")
    try:
        t0 = time.perf_counter()
        rgb = ctx.tonemap(f, 2.2)
        t1 = time.perf_counter()
        write_image(os.path.join(tmp, "film.png"), rgb)
        t2 = time.perf_counter()
        write_exr(os.path.join(tmp, "film.exr"), f)
        t3 = time.perf_counter()
        write_mat(os.path.join(tmp, "film.m"), f)
        t4 = time.perf_counter()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {
        "tonemap": {"film": f"{side}x{side} float RGB resident in HBM, gamma 2.2", "kernel_ms": round(t, 4),
                    "bytes": nbytes, "achieved_gbs": round(gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                    "frac": round(gbs / HBM_PEAK_GBS, 4)},
        "writers": {"film": f"{w}x{h}", "tonemap_host_s": round(t1 - t0, 4), "png_s": round(t2 - t1, 4),
                    "exr_s": round(t3 - t2, 4), "mat_s": round(t4 - t3, 3)},
    }


def markov_chains(ctx, scene, arrays, cpu=True):
    """The Markov-chain renderers with the reference's default options on the
    benchmark scene (MAX_DEPTH 5): Metropolis with 16 mutations per pixel,
    ERPT with 100 mutations per chain from 2x2 initial samples per pixel, and
    the filtered Metropolis renderer (3 passes of two films, cross-filtered;
    its time includes the filtering).
    Throughput = proposals (path evaluations) per second; the oracle runs a
    bounded sample of the same Metropolis workload on the host cores."""
    import torch

    from renderer_amd import MLT_ERPT, MLT_FILTERED, MLT_METROPOLIS

    cam = scene.camera
    film = torch.zeros((cam.height, cam.width, 3), dtype=torch.float32, device="cuda:0")
    out = {}
    for name, kw in (("metropolis", dict(kind=MLT_METROPOLIS, mutations=16, bootstrap=10000, seed=1)),
                     ("erpt", dict(kind=MLT_ERPT, mutations=100, erpt_spp=(2, 2), seed=1)),
                     ("metropolisflt", dict(kind=MLT_FILTERED, mutations=16, bootstrap=10000, seed=1, flt_passes=3))):
        ctx.render_mlt_device(cam, film.data_ptr(), **kw)  # warm-up
        film.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = ctx.render_mlt_device(cam, film.data_ptr(), **kw)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        paths = r["mutations"] + r["seed_paths"]
        out[name] = {"mutations": r["mutations"], "chains": r["chains"], "seed_paths": r["seed_paths"],
                     "acceptance": round(r["accepted"] / max(1, r["mutations"]), 4), "seconds": round(wall, 3),
                     "mutations_per_s": round(r["mutations"] / wall, 1), "paths_per_s": round(paths / wall, 1),
                     "mrays_per_s": round(r["rays"] / wall / 1e6, 1)}
        log(f"{name}: {r['mutations']} mutations in {wall:.2f} s, {r['rays'] / wall / 1e6:.0f} Mrays/s")
    del film
    if cpu:
        from oracle.oracle import Oracle
        from renderer_amd import Camera
        threads = os.cpu_count() or 1
        orc = Oracle(arrays, accel=1)
        small = Camera(cam.cop, cam.look_at, cam.up, cam.fov, 128, 128)
        t0 = time.time()
        _, r = orc.render_mlt(small, threads=threads, kind=MLT_METROPOLIS, mutations=16, bootstrap=10000, seed=1)
        dt = time.time() - t0
        orc.close()
        out["metropolis"]["cpu_oracle"] = {"mutations_per_s": round(r["mutations"] / dt, 1), "cores": threads,
                                           "sample": f"128x128, 16 mutations per pixel ({r['mutations']} "
                                                     f"mutations, {r['seed_paths']} bootstrap paths) in {dt:.1f} s"}
    return out


def cpu_baseline(arrays, scene, budget_s):
    """The oracle (the reference's algorithms: leaf-1 BVH visiting both
    children, per-sample critical section) rendering whole frames of the same
    workload at 1 spp each, new seed per frame, until the time budget is
    spent.  Returns Mrays/s over the frames rendered."""
    from oracle.oracle import Oracle

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    cpu_model, cpu_mhz = host_cpu()
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
            "cpu": cpu_model, "cpu_mhz": cpu_mhz, "box_c1": box_c1(threads),
            "sample": f"{frames} frames of {cam.width}x{cam.height} at 1 spp (seeds 0..{frames - 1}), {rays} rays "
                      f"in {t_render:.1f} s on {threads} threads; reference BVH build {build_s:.1f} s not timed"}


def box_c1(threads):
    """The box scene at its configuration (256 x 256, 16 spp jittered, max
    depth 8, roulette after 3) rendered whole in both estimators, by the
    oracle on the host's cores and by the GPU: wall time and Mrays/s of each
    (one warm-up render on the GPU first)."""
    import torch

    from oracle.oracle import Oracle
    from renderer_amd import MODE_COMPAT, MODE_UNBIASED, GpuContext, scenes

    scene = scenes.box_scene(256, 256)
    arrays = scene.flatten()
    orc = Oracle(arrays, accel=1)
    ctx = GpuContext(0)
    ctx.upload(arrays)
    film = torch.zeros((256, 256, 3), dtype=torch.float32, device="cuda:0")
    out = {"scene": "box, 256x256, 16 spp, max depth 8, roulette after 3 (p 0.5)"}
    for name, mode in (("compat", MODE_COMPAT), ("unbiased", MODE_UNBIASED)):
        kw = dict(mode=mode, max_depth=8, rr_start=3, rr_prob=0.5)
        t0 = time.time()
        _, r = orc.render((scene.camera, 4, 4), seed=0, threads=threads, **kw)
        cpu_s = time.time() - t0
        ctx.render_device([(scene.camera, 4, 4)], [film.data_ptr()], seed=1, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.render_device([(scene.camera, 4, 4)], [film.data_ptr()], seed=0, **kw)
        torch.cuda.synchronize()
        gpu_s = time.perf_counter() - t0
        st = ctx.stats()
        grays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
        out[name] = {"cpu_s": round(cpu_s, 3), "cpu_mrays_per_s": round(r / cpu_s / 1e6, 3),
                     "gpu_s": round(gpu_s, 5), "gpu_mrays_per_s": round(grays / gpu_s / 1e6, 1)}
    orc.close()
    ctx.close()
    out["cores"] = threads
    return out


def host_cpu():
    """Model name and current clock (MHz, mean over the cores) from /proc/cpuinfo."""
    model, mhz = None, []
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model is None:
                    model = v.strip()
                elif k == "cpu MHz":
                    mhz.append(float(v))
    except OSError:
        pass
    return model, (round(sum(mhz) / len(mhz), 1) if mhz else None)


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
    sh_ms = 0.0
    ext_launches = 0
    fin_ms = 0.0
    fin_rays = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render_device(views, [film.data_ptr()], seed=i, **params)
        s = ctx.stats()
        rays += s["camera_rays"] + s["extension_rays"] + s["shadow_rays"]
        ext_ms += s["extend_ms"]
        sh_ms += s["shadow_ms"]
        ext_launches += s["extend_launches"]
        fin_ms += s["finish_ms"]
        fin_rays += s["finish_rays"] + s["finish_shadow_rays"]
        if i == 0:
            first = s
        log(f"step {i}: {s['render_ms']:.1f} ms, {(s['camera_rays'] + s['extension_rays'] + s['shadow_rays']) / 1e6:.1f} Mrays")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    last = ctx.stats()

    # algorithmic bytes of the closest-hit launches: counted in a separate,
    # untimed render of the same workload (same seeds as step 0), cut at the
    # bounce from which the timed render's passes went to the finishing
    # launch, so it counts exactly the bounces the per-bounce launches traced
    finish_depth = ctx.launch_config()["finish_depth"]
    cut = count_depth(first, params["max_depth"])
    ctx.set_launch_config(count_traversal=1)
    ctx.render_device(views, [film.data_ptr()], seed=0, **dict(params, max_depth=cut))
    cs = ctx.stats()
    ctx.set_launch_config(count_traversal=0)
    ext_rays = cs["camera_rays"] + cs["extension_rays"]
    timed_ext = first["camera_rays"] + first["extension_rays"] - first["finish_rays"]  # step 0: seed 0 too
    if ext_rays != timed_ext:
        log(f"warning: the counting render traced {ext_rays} closest-hit rays, the timed launches {timed_ext}")
    mem_nodes = cs["node_visits"] - cs["node_visits_lds"]
    alg_bytes = ext_rays * RAY_BYTES + mem_nodes * NODE_BYTES + cs["prim_tests"] * PRIM_BYTES
    per_launch_bytes = alg_bytes / max(1, cs["extend_launches"])
    avg_launch_ms = ext_ms / max(1, ext_launches)
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else None
    # the same for the any-hit launches (one per bounce, like the closest-hit ones)
    sh_mem_nodes = cs["shadow_node_visits"] - cs["shadow_node_visits_lds"]
    sh_bytes = (cs["shadow_rays"] * SHADOW_RAY_BYTES + sh_mem_nodes * NODE_BYTES +
                cs["shadow_prim_tests"] * PRIM_BYTES)
    sh_per_launch = sh_bytes / max(1, cs["extend_launches"])
    sh_avg_ms = sh_ms / max(1, ext_launches)
    sh_achieved = sh_per_launch / (sh_avg_ms * 1e-3) / 1e9 if sh_avg_ms > 0 else None
    # measured L2-to-fabric traffic and L2 hit rates of these kernels (rocprofv3
    # counters of the same workload, committed under profiles/)
    traffic = sh_traffic = l2 = sh_l2 = None
    tfile = os.path.join(ROOT, "profiles", "traversal_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        traffic = tr.get("k_extend", {}).get("hbm_bytes_per_launch")
        sh_traffic = tr.get("k_shadow", {}).get("hbm_bytes_per_launch")
        l2 = tr.get("k_extend", {}).get("l2_hit_rate")
        sh_l2 = tr.get("k_shadow", {}).get("l2_hit_rate")

    sah = None
    if not args.no_sah:
        log("the same workload on the surface-area-heuristic tree")
        sah = sah_tree(ctx, arrays, views, params, film, args)

    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")  # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del a, b

    mlt = None
    if not args.no_mlt:
        log("Metropolis and ERPT")
        mlt = markov_chains(ctx, scene, arrays, cpu=not args.no_cpu)

    fout = None
    if not args.no_film:
        log("film outputs")
        fout = film_outputs(ctx, film)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(arrays, scene, args.cpu_seconds)
    ctx.close()

    batched = None
    if not args.no_batched:
        log("batched views")
        batched = batched_views(args, scene, arrays)

    sfile = None
    if not args.no_scene_file:
        log("scene-file front end")
        sfile = scene_file_load(args)

    guided = env = bound = None
    if not args.no_guided:
        if not args.no_room:
            log("equal-time guided vs plain: occluded-light room")
            guided = equal_time_room(args)
        if args.guide_bound:
            log("first-bounce proposals in the room")
            bound = guide_bound(args)
        if not args.no_env:
            log("equal-time guided vs plain: env-lit mesh")
            env = equal_time_env(args)

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
            "lds_nodes_per_ray": round(cs["node_visits_lds"] / max(1, ext_rays), 2),
            "prims_per_ray": round(cs["prim_tests"] / max(1, ext_rays), 2),
            "l2_hit_rate": l2,
        },
        "roofline_shadow": {
            "bound": "hbm",
            "kernel": "k_shadow (any-hit BVH traversal of next-event rays)",
            "achieved": round(sh_achieved, 1) if sh_achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(sh_achieved / HBM_PEAK_GBS, 4) if sh_achieved else None,
            "traffic": sh_traffic,
            "avg_launch_ms": round(sh_avg_ms, 4),
            "alg_bytes_per_launch": int(sh_per_launch),
            "nodes_per_ray": round(cs["shadow_node_visits"] / max(1, cs["shadow_rays"]), 2),
            "lds_nodes_per_ray": round(cs["shadow_node_visits_lds"] / max(1, cs["shadow_rays"]), 2),
            "prims_per_ray": round(cs["shadow_prim_tests"] / max(1, cs["shadow_rays"]), 2),
            "l2_hit_rate": sh_l2,
        },
        "finish_launch": {
            "kernel": "k_finish (the bounces from finish_depth on, a lane per path)",
            "finish_depth": finish_depth,
            "from_bounce": first["finish_bounce"],
            "ms_per_step": round(fin_ms / args.steps, 3),
            "rays_share": round(fin_rays / max(1, rays), 4),
        },
        "cpu_baseline": cpu,
        "batched_views": batched,
        "scene_file": sfile,
        "film_outputs": fout,
        "markov_chains": mlt,
        "sah_tree": sah,
        "equal_time_guided_vs_plain": guided,
        "guide_first_bounce_bound": bound,
        "equal_time_guided_vs_plain_env": env,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
