"""GPU guiding network against the FP64 oracle (forward, gradient, Adam),
learning a known lobe, and unbiasedness of guided rendering."""
import math

import numpy as np
import pytest

from oracle import guide_oracle as G
from oracle.oracle import Oracle
from renderer_amd import MODE_GUIDED, MODE_UNBIASED, scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu

W_SLICES = [(0, 5120), (5184, 9280), (9344, 13440), (13504, 16576)]
B_SLICES = [(5120, 5184), (9280, 9344), (13440, 13504), (16576, 16624)]


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _records(B, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((B, 3)).astype(np.float32)
    n = _unit(rng.normal(size=(B, 3)))
    wo = _unit(rng.normal(size=(B, 3)))
    wo = np.where((wo * n).sum(1, keepdims=True) < 0, -wo, wo)
    om = _unit(rng.normal(size=(B, 3)))
    pb = np.maximum((om * n).sum(1), 0) / math.pi + 0.02
    w = rng.uniform(0.1, 2.0, B)
    rec = np.zeros((B, 16), np.float32)
    rec[:, 0:3], rec[:, 3:6], rec[:, 6:9], rec[:, 9:12] = pos, n, wo, om
    rec[:, 12], rec[:, 13], rec[:, 14] = pb, w, 1.0
    return rec


def _fp16_weights(theta):
    th = theta.astype(np.float64).copy()
    for a, b in W_SLICES:
        th[a:b] = th[a:b].astype(np.float16).astype(np.float64)
    return th


def _oracle_inputs(rec):
    r = rec.astype(np.float32)
    x = G.features(r[:, 0:3].astype(np.float64), r[:, 3:6].astype(np.float64), r[:, 6:9].astype(np.float64))
    return x, r[:, 9:12].astype(np.float64), r[:, 12].astype(np.float64), r[:, 13].astype(np.float64)


@pytest.fixture()
def guide_ctx(gpu_ctx):
    gpu_ctx.upload(scenes.box_scene(8, 8).flatten())
    gpu_ctx.guide_init(0)
    gpu_ctx.guide_set_bounds((0, 0, 0), (1, 1, 1))
    return gpu_ctx


def _theta(seed):
    rng = np.random.default_rng(seed)
    th = G.init_params(seed).astype(np.float32)
    for a, b in B_SLICES:
        th[a:b] = rng.normal(0, 0.1, b - a)
    return th


def test_forward_matches_oracle(guide_ctx):
    th = _theta(1)
    guide_ctx.guide_set_params(th)
    rec = _records(1000, 2)
    y = guide_ctx.guide_forward(rec)
    x, _, _, _ = _oracle_inputs(rec)
    yr, _ = G.forward(_fp16_weights(th), x, fp16_acts=True)
    # outputs 41..47 feed nothing and are not computed
    assert np.all(y[:, 41:] == 0)
    y, yr = y[:, :41], yr[:, :41]
    err = np.abs(y - yr)
    scale = np.abs(yr).max()
    print(f"forward max err {err.max():.3e} scale {scale:.3f}", flush=True)
    assert err.max() <= 2e-2 * scale
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * scale


def test_head_matches_oracle(guide_ctx):
    """The values the guided shader samples from and evaluates: mean
    directions (renormalised), concentrations, mixture weights (divided by
    their sum) and the guide probability, against the oracle's head of its
    FP64 forward pass.  The outputs are stored as FP16."""
    th = _theta(7)
    guide_ctx.guide_set_params(th)
    rec = _records(3000, 8)
    out = guide_ctx.guide_head(rec).astype(np.float64)
    x, _, _, _ = _oracle_inputs(rec)
    yr, _ = G.forward(_fp16_weights(th), x, fp16_acts=True)
    hd = G.head(yr)
    lob = out[:, :40].reshape(-1, 8, 5)
    mu = lob[:, :, :3] / np.linalg.norm(lob[:, :, :3], axis=2, keepdims=True)
    ang = np.arccos(np.clip((mu * hd["mu"]).sum(2), -1, 1))
    wts = lob[:, :, 4] / lob[:, :, 4].sum(1, keepdims=True)
    print(f"head: direction err max {ang.max():.2e} rad, log-kappa err max "
          f"{np.abs(np.log(lob[:, :, 3]) - np.log(hd['kappa'])).max():.2e}, weight err max "
          f"{np.abs(wts - hd['pi']).max():.2e}, alpha err max {np.abs(out[:, 40] - hd['alpha']).max():.2e}", flush=True)
    assert np.all(out[:, 41:] == 0)
    assert np.percentile(ang, 99) < 0.02 and ang.max() < 0.1
    assert np.abs(np.log(lob[:, :, 3]) - np.log(hd["kappa"])).max() < 0.05
    assert np.abs(wts - hd["pi"]).max() < 0.01
    assert np.abs(out[:, 40] - hd["alpha"]).max() < 0.005


def test_gradient_matches_oracle(guide_ctx):
    th = _theta(3)
    guide_ctx.guide_set_params(th)
    rec = _records(4096, 4)
    loss, g = guide_ctx.guide_gradient(rec)
    x, om, pb, w = _oracle_inputs(rec)
    Lr, gr, _ = G.loss_and_grad(_fp16_weights(th), x, om, pb, w, fp16_acts=True)
    print(f"loss gpu {loss:.6f} oracle {Lr:.6f}", flush=True)
    assert loss == pytest.approx(Lr, rel=1e-2)
    for (a, b), (c, d) in zip(W_SLICES, B_SLICES):
        for lo, hi in ((a, b), (c, d)):
            ga, gb = g[lo:hi].astype(np.float64), gr[lo:hi]
            rel = np.linalg.norm(ga - gb) / max(np.linalg.norm(gb), 1e-12)
            cos = ga @ gb / max(np.linalg.norm(ga) * np.linalg.norm(gb), 1e-30)
            print(f"params {lo}:{hi} rel err {rel:.3e} cos {cos:.6f}", flush=True)
            assert rel < 5e-2 and cos > 0.998


def test_adam_first_step(guide_ctx):
    th = _theta(5)
    guide_ctx.guide_set_params(th)
    rec = _records(2048, 6)
    lr = 1e-3
    guide_ctx.guide_train(rec, steps=1, lr=lr)
    th1 = guide_ctx.guide_params()
    x, om, pb, w = _oracle_inputs(rec)
    _, gr, _ = G.loss_and_grad(_fp16_weights(th), x, om, pb, w, fp16_acts=True)
    big = np.abs(gr) > 1e-3 * np.abs(gr).max()
    step = (th1 - th).astype(np.float64)
    # first Adam step: -lr * g / (|g| + eps) = -lr * sign(g) where g is not tiny
    agree = np.mean(np.sign(step[big]) == -np.sign(gr[big]))
    assert agree > 0.99
    assert np.allclose(np.abs(step[big]), lr, rtol=0.02, atol=1e-6)


def test_learns_a_known_lobe(guide_ctx):
    """Records carry directions drawn uniformly with weights equal to a known
    vMF density over the uniform pdf; training must move the mixture towards
    that vMF (the loss decreases and the mixture mass concentrates on it)."""
    rng = np.random.default_rng(8)
    B = 16384
    rec = _records(B, 9)
    rec[:, 0:9] = rec[0, 0:9]  # one shading point: the network cannot memorise per-record positions
    om = _unit(rng.normal(size=(B, 3)))
    mu = np.array([0.6, 0.0, 0.8])
    kappa = 20.0
    target = kappa / (2 * math.pi * (1 - math.exp(-2 * kappa))) * np.exp(kappa * (om @ mu - 1))
    rec[:, 9:12] = om
    rec[:, 12] = 1 / (4 * math.pi)
    rec[:, 13] = target * 4 * math.pi
    guide_ctx.guide_set_params(G.init_params(10))
    l0, _ = guide_ctx.guide_gradient(rec)
    guide_ctx.guide_train(rec, steps=400, lr=5e-3)
    l1, _ = guide_ctx.guide_gradient(rec)
    th = guide_ctx.guide_params()
    x, omo, pbo, wo_ = _oracle_inputs(rec[:256])
    y, _ = G.forward(_fp16_weights(th), x, fp16_acts=True)
    pm, q, alpha = G.mixture_pdf(y, np.tile(mu, (256, 1)), np.full(256, 1 / (4 * math.pi)))
    # cross-entropy of the target with itself (its entropy) bounds the loss
    # from below; E[mu.w] = coth(k) - 1/k for a vMF
    h = -(math.log(kappa / (2 * math.pi * (1 - math.exp(-2 * kappa)))) + kappa * (1 / math.tanh(kappa) - 1 / kappa - 1))
    print(f"loss {l0:.4f} -> {l1:.4f} (target entropy {h:.4f}); guide density at the lobe mean {np.median(q):.2f} "
          f"(target {kappa / (2 * math.pi):.2f}); alpha {np.median(alpha):.3f}", flush=True)
    assert l1 < l0 - 1.0
    assert l1 > h - 0.2  # a proper density: no better than the target itself (up to sampling noise)
    assert np.median(q) > 0.4 * kappa / (2 * math.pi)
    assert np.median(alpha) > 0.7


def test_guided_render_is_unbiased(gpu_ctx):
    """Guided rendering (with online training) converges to the image of the
    unbiased estimator: over 8 seeds the mean images agree within 4
    standard errors, and the guided error against a high-sample oracle
    reference is of the same order."""
    sc = scenes.box_scene(32, 32)
    arr = sc.flatten()
    gpu_ctx.upload(arr)
    gpu_ctx.guide_init(1)
    orc = Oracle(arr)
    R, _ = orc.render((sc.camera, 32, 32), mode=MODE_UNBIASED, max_depth=8, seed=77)
    mg, mp, eg, ep = [], [], [], []
    for s in range(8):
        ig = gpu_ctx.render([(sc.camera, 8, 8)], mode=MODE_GUIDED, max_depth=8, seed=s, guide_train=1,
                            guide_batch=16384)[0]
        ip = gpu_ctx.render([(sc.camera, 8, 8)], mode=MODE_UNBIASED, max_depth=8, seed=100 + s)[0]
        assert np.isfinite(ig).all()
        mg.append(ig.mean(axis=(0, 1)))
        mp.append(ip.mean(axis=(0, 1)))
        eg.append(rel_mse(ig, R))
        ep.append(rel_mse(ip, R))
    mg, mp = np.array(mg), np.array(mp)
    se = np.sqrt(mg.var(0, ddof=1) / 8 + mp.var(0, ddof=1) / 8)
    print(f"means guided {mg.mean(0)} plain {mp.mean(0)} se {se}; relMSE {np.mean(eg):.4f} vs {np.mean(ep):.4f}",
          flush=True)
    assert np.all(np.abs(mg.mean(0) - mp.mean(0)) <= 4 * se + 1e-5)
    assert np.mean(eg) < 3 * np.mean(ep)


def test_guided_render_reports_network_work(gpu_ctx):
    """The library's own counters for the network: one inference per shaded
    vertex before the last bounce, GD_TRAIN_CHUNKS x epochs Adam steps per
    pass, timers running."""
    sc = scenes.box_scene(32, 32)
    gpu_ctx.upload(sc.flatten())
    gpu_ctx.guide_init(3)
    gpu_ctx.set_launch_config(guide_epochs=2)
    gpu_ctx.render([(sc.camera, 4, 4)], mode=MODE_GUIDED, max_depth=8, seed=1, guide_train=1, guide_batch=4096)
    st = gpu_ctx.stats()
    gpu_ctx.set_launch_config(guide_epochs=1)
    # every shaded vertex but those of the last bounce (max depth 8)
    assert st["camera_rays"] <= st["guide_infer_samples"] < st["camera_rays"] + st["extension_rays"]
    assert st["guide_train_steps"] == st["passes"] * 4 * 2
    assert 0 < st["guide_train_samples"] <= st["guide_train_steps"] * 4096
    assert st["guide_infer_ms"] > 0 and st["guide_train_ms"] > 0
    gpu_ctx.render([(sc.camera, 4, 4)], mode=MODE_UNBIASED, max_depth=8, seed=1)
    st = gpu_ctx.stats()
    assert st["guide_infer_samples"] == 0 and st["guide_train_steps"] == 0 and st["guide_infer_ms"] == 0


def test_trained_network_gains_over_plain_sampling(gpu_ctx):
    """What the learned proposal gains over the method without it: in the
    occluded-light room (the only light outside, behind a door gap of 5% of
    the wall width), after 24 passes of online training the guided estimator
    has a lower relMSE than plain BSDF sampling at equal samples per pixel
    (256), averaged over 4 seeds, against a 16,384-spp plain reference."""
    W, H = 320, 180
    sc = scenes.room_scene(W, H, n=32, seed=3)
    gpu_ctx.upload(sc.flatten())
    common = dict(max_depth=16, rr_start=3, rr_prob=0.5)
    ref = np.zeros((H, W, 3))
    for k in range(64):
        ref += gpu_ctx.render([(sc.camera, 16, 16)], mode=MODE_UNBIASED, seed=1000 + k, **common)[0]
    ref /= 64
    gpu_ctx.guide_init(7)
    gpu_ctx.set_launch_config(guide_epochs=4)
    try:
        for k in range(24):
            gpu_ctx.render([(sc.camera, 4, 4)], mode=MODE_GUIDED, seed=500 + k, guide_train=1, guide_lr=5e-3, **common)
    finally:
        gpu_ctx.set_launch_config(guide_epochs=1)
    eg, ep = [], []
    for s in range(4):
        ig = gpu_ctx.render([(sc.camera, 16, 16)], mode=MODE_GUIDED, seed=s, guide_train=0, **common)[0]
        ip = gpu_ctx.render([(sc.camera, 16, 16)], mode=MODE_UNBIASED, seed=s, **common)[0]
        assert np.isfinite(ig).all()
        eg.append(rel_mse(ig, ref))
        ep.append(rel_mse(ip, ref))
    gain = np.mean(ep) / np.mean(eg)
    print(f"relMSE plain {np.mean(ep):.4f} guided {np.mean(eg):.4f}: gain {gain:.3f} at 256 spp", flush=True)
    assert gain > 1.15
