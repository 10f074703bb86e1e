"""Shared test helpers: seeded ray batches, hit comparison, image metrics."""
import numpy as np

EPS = 0.001


def random_rays(n, lo, hi, seed, tmin=EPS, tmax=np.inf, normalize=True, axis_fraction=0.1):
    """Origins uniform in the box [lo, hi] (slightly enlarged), directions
    uniform on the sphere; a fraction of directions is snapped to have zero
    components (axis-parallel rays exercise the 0 * inf slab cases)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    ext = hi - lo
    o = rng.uniform(lo - 0.1 * ext, hi + 0.1 * ext, (n, 3))
    d = rng.normal(size=(n, 3))
    k = int(n * axis_fraction)
    if k:
        zero = rng.integers(0, 3, k)
        d[np.arange(k), zero] = 0.0
        two = rng.integers(0, 3, k // 2)
        d[np.arange(k // 2), two] = 0.0
        d[np.arange(k // 2), (two + 1) % 3] = 0.0
    d[np.all(d == 0, axis=1)] = (0.0, 1.0, 0.0)
    if normalize:
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    else:
        d *= rng.uniform(0.2, 3.0, (n, 1))
    o4 = np.concatenate([o, np.full((n, 1), tmin)], 1).astype(np.float32)
    d4 = np.concatenate([d, np.full((n, 1), tmax)], 1).astype(np.float32)
    return o4, d4


def camera_rays(scene, spp=1, seed=0):
    """Primary rays of a scene's camera through jittered pixel positions,
    with the reference's camera arithmetic (src/camera.cpp:34-58)."""
    cam = scene.camera
    rng = np.random.default_rng(seed)
    W, H = cam.width, cam.height
    w = cam.cop - cam.look_at
    w = w / np.linalg.norm(w)
    u = np.cross(cam.up, w)
    u = u / np.linalg.norm(u)
    v = np.cross(w, u)
    t = np.tan(cam.fov / 2)
    r = W / H * t
    py, px = np.mgrid[0:H, 0:W]
    px = np.repeat(px.reshape(-1), spp) + rng.random(H * W * spp)
    py = np.repeat(py.reshape(-1), spp) + rng.random(H * W * spp)
    x = -r + 2 * r * (px + 0.5) / W
    y = t - 2 * t * (py + 0.5) / H
    d = x[:, None] * u + y[:, None] * v - w
    n = len(d)
    o4 = np.concatenate([np.tile(cam.cop, (n, 1)), np.full((n, 1), EPS)], 1).astype(np.float32)
    d4 = np.concatenate([d, np.full((n, 1), np.inf)], 1).astype(np.float32)
    return o4, d4


def compare_hits(orc, o4, d4, prim_gpu, t_gpu, prim_ref, t_ref, rtol=1e-5):
    """Per-ray parity: same primitive and t within rtol, except
    (a) ties - the oracle intersects the GPU's primitive at the same t, and
    (b) edge grazing - the two primitives meet where the ray passes, so the
    oracle finds the GPU's primitive at the same t when its t-range is
    allowed a 1e-5 relative margin.  Returns (n_exact, n_tie, bad_indices)."""
    miss_g = prim_gpu < 0
    miss_r = prim_ref < 0
    same = (prim_gpu == prim_ref)
    t_ok = np.where(miss_g & miss_r, True,
                    np.abs(t_gpu - t_ref) <= rtol * np.maximum(np.abs(t_ref), 1e-3))
    exact = same & t_ok
    bad = []
    ties = 0
    for i in np.nonzero(~exact)[0]:
        if prim_gpu[i] >= 0 and prim_ref[i] >= 0:
            tg = orc.intersect_prim(prim_gpu[i], o4[i], d4[i])
            if np.isfinite(tg) and abs(tg - t_ref[i]) <= rtol * max(abs(t_ref[i]), 1e-3) and \
                    abs(t_gpu[i] - t_ref[i]) <= rtol * max(abs(t_ref[i]), 1e-3):
                ties += 1
                continue
        bad.append(i)
    return int(exact.sum()), ties, np.array(bad, np.int64)


def rel_mse(img, ref, eps=1e-2):
    """mean over pixels and channels of (I - R)^2 / (R^2 + eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))
