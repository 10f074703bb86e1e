"""Launch-configuration autotuner for rt_render.

The wavefront pipeline has two free launch parameters that change only the
time, never the image (films are accumulated per pixel in sample order, so a
fixed seed gives bit-identical films for every choice):

  path_slots     paths in flight per pass (more slots: fewer passes and
                 fuller late-bounce launches, more HBM for path state)
  blocks_per_cu  persistent blocks per CU (the traversal kernels' LDS stacks
                 allow 8 resident blocks per CU; shading kernels fit fewer)

`autotune` times a reduced version of the caller's workload (the same views,
each sample count divided by `spp_divisor`) for every candidate pair and
keeps the fastest, optionally caching the choice in a JSON file keyed by a
signature of the workload.
"""
import json
import math
import os

DEFAULT_SLOTS = (1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27)
DEFAULT_BPC = (4, 8, 16)


def workload_signature(ctx, views, params):
    st = ctx.stats()
    pixels = sum(c.width * c.height for c, _, _ in views)
    samples = sum(c.width * c.height * sx * sy for c, sx, sy in views)
    return (f"prims{int(math.log2(max(1, st['bvh_prims'])))}_views{len(views)}_px{int(math.log2(max(1, pixels)))}"
            f"_s{int(math.log2(max(1, samples)))}_d{params.get('max_depth', 8)}_m{params.get('mode', 1)}")


def _reduced(views, divisor):
    out = []
    for cam, sx, sy in views:
        spp = max(1, (sx * sy) // divisor)
        nx = max(d for d in range(1, int(math.isqrt(spp)) + 1) if spp % d == 0)
        out.append((cam, nx, spp // nx))
    return out


def autotune(ctx, views, params, slots=DEFAULT_SLOTS, blocks_per_cu=DEFAULT_BPC, spp_divisor=8, repeats=2,
             cache_path=None, verbose=False):
    """Pick (path_slots, blocks_per_cu) for rendering `views` with `params`
    (keyword arguments of GpuContext.render) and apply it to ctx.  Returns a
    dict with the choice and every trial's time."""
    import numpy as np

    key = workload_signature(ctx, views, params)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            best = cache[key]
            ctx.set_launch_config(path_slots=best["path_slots"], blocks_per_cu=best["blocks_per_cu"])
            return dict(best, cached=True, key=key)
    trial_views = _reduced(views, spp_divisor)
    films = [np.zeros((c.height, c.width, 3), np.float32) for c, _, _ in trial_views]
    total = sum(c.width * c.height * sx * sy for c, sx, sy in trial_views)
    trials = []
    for s in slots:
        if s > 4 * total:  # more slots than the trial has samples: equivalent to a smaller choice
            continue
        for b in blocks_per_cu:
            ctx.set_launch_config(path_slots=int(s), blocks_per_cu=int(b))
            best_ms = float("inf")
            for r in range(repeats):
                ctx.render(trial_views, films=films, seed=1000 + r, **params)
                best_ms = min(best_ms, ctx.stats()["render_ms"])
            trials.append({"path_slots": int(s), "blocks_per_cu": int(b), "ms": round(best_ms, 3)})
            if verbose:
                print("autotune", trials[-1], flush=True)
    best = min(trials, key=lambda t: t["ms"])
    ctx.set_launch_config(path_slots=best["path_slots"], blocks_per_cu=best["blocks_per_cu"])
    result = {"path_slots": best["path_slots"], "blocks_per_cu": best["blocks_per_cu"], "trial_ms": best["ms"],
              "trials": trials, "spp_divisor": spp_divisor, "key": key, "cached": False}
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = {k: result[k] for k in ("path_slots", "blocks_per_cu", "trial_ms")}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return result
