"""The Markov-chain renderers on the GPU against the oracle restatement:
the sample-driven integrator path by path, Metropolis and ERPT chain by
chain (both sides draw the same Philox numbers, so chains take the same
steps until a floating-point difference flips an acceptance test), exact
luminance conservation, bitwise reproducibility whatever the batch size,
and convergence to the path-traced image."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MLT_ERPT, MLT_METROPOLIS, MODE_COMPAT, scenes
from tests.test_oracle_mlt import _cornell, _lum

pytestmark = pytest.mark.gpu


def test_integrate_samples_matches_oracle(gpu_ctx):
    sc = scenes.box_scene(48, 48)
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    rng = np.random.default_rng(7)
    x = rng.random((20000, 11, 4), dtype=np.float32)
    print("integrate_samples: 20000 paths, box scene", flush=True)
    got = gpu_ctx.integrate_samples(sc.camera, x, max_depth=5, rr_start=3, termination_prob=0.5)
    want = Oracle(arr).integrate_samples(sc.camera, x, max_depth=5, rr_start=3, termination_prob=0.5)
    agree = np.all(np.abs(got - want) <= 1e-3 * np.abs(want) + 1e-5, axis=1).mean()
    print(f"agreement {agree:.5f}, nonzero {np.count_nonzero(want.sum(axis=1))}", flush=True)
    assert agree > 0.995
    assert np.count_nonzero(want.sum(axis=1)) > 2000


def _agree(a, b):
    d = np.abs(a - b)
    return np.all(d <= 1e-3 * np.abs(b) + 1e-6, axis=2).mean()


@pytest.mark.parametrize("kind", [MLT_METROPOLIS, MLT_ERPT])
def test_chains_match_oracle(gpu_ctx, kind):
    # off-centre: the e_d paths start at pixel corners, which a centred view
    # of the box lines up exactly with its edges, where the two sides break
    # ties between walls differently
    sc = _cornell(32, eye=(0.46, 0.53, 2.2), look_at=(0.52, 0.48, 0.0))
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    kw = dict(kind=kind, mutations=8, chains=2048, bootstrap=20000, seed=11) if kind == MLT_METROPOLIS else \
        dict(kind=kind, mutations=6, erpt_spp=(2, 2), seed=11)
    print(f"chains kind {kind}: 32x32", flush=True)
    img, res = gpu_ctx.render_mlt(sc.camera, **kw)
    ref, rref = Oracle(arr).render_mlt(sc.camera, **kw)
    print("gpu", res, "\noracle", rref, flush=True)
    for k in ("chains", "mutations", "seed_paths", "chain_seeds"):
        assert res[k] == rref[k], k
    assert res["b"] == pytest.approx(rref["b"], rel=1e-5)
    assert res["ed"] == pytest.approx(rref["ed"], rel=1e-5)
    assert abs(res["accepted"] - rref["accepted"]) <= 0.01 * rref["accepted"]
    agree = _agree(img, ref)
    print(f"pixel agreement {agree:.4f}, means {img.mean():.6f} {ref.mean():.6f}", flush=True)
    assert agree > 0.95
    assert img.mean() == pytest.approx(ref.mean(), rel=0.01)


def test_metropolis_conservation_and_reproducibility(gpu_ctx):
    sc = _cornell(40)
    gpu_ctx.upload(sc.flatten())
    kw = dict(kind=MLT_METROPOLIS, mutations=16, chains=6000, bootstrap=30000, seed=21)
    a, ra = gpu_ctx.render_mlt(sc.camera, **kw)
    assert _lum(a).sum() == pytest.approx(ra["b"] * 40 * 40, rel=1e-4)
    b, _ = gpu_ctx.render_mlt(sc.camera, **kw)
    assert np.array_equal(a, b)
    # batches of 1000 chains: the same chains, the same image bit for bit
    try:
        gpu_ctx.set_launch_config(path_slots=1000)
        c, rc = gpu_ctx.render_mlt(sc.camera, **kw)
    finally:
        gpu_ctx.set_launch_config(path_slots=1 << 24)
    assert rc["chains"] == ra["chains"] and rc["accepted"] == ra["accepted"]
    assert np.array_equal(a, c)


def test_erpt_batches_reproduce(gpu_ctx):
    sc = _cornell(24)
    gpu_ctx.upload(sc.flatten())
    kw = dict(kind=MLT_ERPT, mutations=10, erpt_spp=(2, 2), seed=5)
    a, ra = gpu_ctx.render_mlt(sc.camera, **kw)
    try:
        gpu_ctx.set_launch_config(path_slots=700)
        b, rb = gpu_ctx.render_mlt(sc.camera, **kw)
    finally:
        gpu_ctx.set_launch_config(path_slots=1 << 24)
    assert ra == {**rb, "render_ms": ra["render_ms"]}
    assert np.array_equal(a, b)


def test_metropolis_converges_to_path_tracing_on_gpu(gpu_ctx):
    sc = _cornell(64)
    gpu_ctx.upload(sc.flatten())
    pt = gpu_ctx.render([(sc.camera, 32, 32)], mode=MODE_COMPAT, max_depth=5, rr_start=3, rr_prob=0.0, seed=1)[0]
    mlt, res = gpu_ctx.render_mlt(sc.camera, kind=MLT_METROPOLIS, termination_prob=0.0, mutations=1024,
                                  chains=1 << 20, bootstrap=1 << 21, seed=2)
    blk = lambda im: im.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))
    rel = np.abs(blk(mlt) - blk(pt)) / blk(pt)
    print(f"MLT vs PT blocks: max {rel.max():.4f} mean {rel.mean():.4f}; {res['mutations']} mutations "
          f"in {res['render_ms']:.1f} ms, acceptance {res['accepted'] / res['mutations']:.3f}", flush=True)
    assert mlt.mean() == pytest.approx(pt.mean(), rel=0.005)
    assert rel.mean() < 0.01 and rel.max() < 0.04
