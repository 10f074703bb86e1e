"""Progressive rendering of one view in passes with the learned proposal, and
plain path tracing as its fallback and error judge.

The guided estimator is unbiased whatever the network outputs, but it is
not always the better one for the time it takes: where next-event
estimation and BSDF sampling already do well (a sky with a sun handled by
environment importance sampling) the network's inference and the guided
shading cost more than its proposal gains.  GuidedProgressive therefore
renders in passes of `pass_spp` samples per pixel:

  1. `train_passes` guided passes that also train the network online;
  2. `probe_passes` guided and `probe_passes` plain passes, interleaved,
     without training: their spread gives each estimator's variance per
     pass (relative to a 7 x 7 box average of the mean of all passes so far,
     in which a rare bright sample weighs little; averaged over pixels),
     their clocks its time per pass;
  3. until the budget is spent, passes of the estimator with the smaller
     variance x time (the smaller error at equal time).

Every pass is an independent unbiased estimate of the image (passes use
distinct seeds), so the result, the mean of all passes, is unbiased too.
The reference renderer has no counterpart (it has no learned component);
the image it converges to is PerPixelRenderer's (src/perpixelrenderer.cpp).
"""
import time

from . import _abi


def relative_pass_variance(passes, eps=1e-2, mean=None):
    """Variance of one pass, estimated from K >= 2 independent passes (torch
    tensors of one shape), relative to mean^2 + eps and averaged over the
    pixels and channels (the relMSE convention of the comparisons).  `mean`:
    the image to normalise by (default: the passes' own mean; an estimate
    pooled over other unbiased passes as well keeps a rare bright sample
    from lowering its own score)."""
    import torch

    k = len(passes)
    if k < 2:
        raise ValueError("need at least two passes")
    stack = torch.stack([p.double() for p in passes])
    own = stack.mean(0)
    var = ((stack - own) ** 2).sum(0) / (k - 1)
    m = own if mean is None else mean.double()
    return float((var / (m * m + eps)).mean())


def smooth_image(img, k=7):
    """k x k box average of an (H, W, 3) image (edges over the pixels inside):
    a normalising image in which one bright sample weighs 1 / k^2."""
    import torch

    x = img.double().permute(2, 0, 1).unsqueeze(0)
    y = torch.nn.functional.avg_pool2d(x, k, stride=1, padding=k // 2, count_include_pad=False)
    return y.squeeze(0).permute(1, 2, 0)


class GuidedProgressive:
    """Passes of one view (camera, width x height) on a GpuContext holding the
    scene; see the module docstring for the schedule."""

    def __init__(self, ctx, camera, pass_spp=16, train_passes=16, probe_passes=6, lr=1e-3, max_depth=8, rr_start=3,
                 rr_prob=0.5):
        if pass_spp < 1 or train_passes < 0 or probe_passes < 2:
            raise ValueError("pass_spp >= 1, train_passes >= 0, probe_passes >= 2")
        self.ctx, self.camera = ctx, camera
        self.pass_spp, self.train_passes, self.probe_passes = pass_spp, train_passes, probe_passes
        self.lr = lr
        self.common = dict(max_depth=max_depth, rr_start=rr_start, rr_prob=rr_prob)
        sx = 1
        while (sx + 1) * (sx + 1) <= pass_spp:
            sx += 1
        self.spp_xy = (sx, pass_spp // sx)

    def _pass(self, seed, mode, train):
        import torch

        cam = self.camera
        f = torch.zeros((cam.height, cam.width, 3), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        self.ctx.render_device([(cam, self.spp_xy[0], self.spp_xy[1])], [f.data_ptr()], seed=seed, mode=mode,
                               guide_train=train, guide_lr=self.lr, **self.common)
        return f

    def run(self, budget_s, seed0=0, on_pass=None):
        """Render passes until `budget_s` seconds of rendering (training and the
        probe's variance estimate included; `on_pass` is called after every
        pass, outside the clock, with (seconds, samples per pixel, mean image
        float64)).  Returns (mean image, info)."""
        import torch

        cam = self.camera
        acc = torch.zeros((cam.height, cam.width, 3), dtype=torch.float64, device="cuda:0")
        n = 0
        el = 0.0
        probes = {_abi.MODE_GUIDED: [], _abi.MODE_UNBIASED: []}
        clock = {_abi.MODE_GUIDED: 0.0, _abi.MODE_UNBIASED: 0.0}
        choice = None
        info = {}

        def one(mode, train, keep):
            nonlocal n, el, acc
            torch.cuda.synchronize()
            t = time.perf_counter()
            f = self._pass(seed0 + n, mode, train)
            acc += f.double()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            el += dt
            n += 1
            if keep:
                probes[mode].append(f)
                clock[mode] += dt
            if on_pass is not None:
                on_pass(el, n * self.pass_spp, acc / n)

        for _ in range(self.train_passes):
            if el >= budget_s:
                break
            one(_abi.MODE_GUIDED, 1, False)
        for _ in range(self.probe_passes):
            if el >= budget_s:
                break
            one(_abi.MODE_GUIDED, 0, True)
            one(_abi.MODE_UNBIASED, 0, True)
        if len(probes[_abi.MODE_GUIDED]) >= 2 and len(probes[_abi.MODE_UNBIASED]) >= 2:
            torch.cuda.synchronize()
            t = time.perf_counter()
            pooled = smooth_image(acc / n)
            vg = relative_pass_variance(probes[_abi.MODE_GUIDED], mean=pooled)
            vp = relative_pass_variance(probes[_abi.MODE_UNBIASED], mean=pooled)
            torch.cuda.synchronize()
            el += time.perf_counter() - t
            tg = clock[_abi.MODE_GUIDED] / len(probes[_abi.MODE_GUIDED])
            tp = clock[_abi.MODE_UNBIASED] / len(probes[_abi.MODE_UNBIASED])
            choice = _abi.MODE_GUIDED if vg * tg <= vp * tp else _abi.MODE_UNBIASED
            info["probe"] = {"guided_var": vg, "plain_var": vp, "guided_pass_s": tg, "plain_pass_s": tp}
        else:
            choice = _abi.MODE_GUIDED
        probes.clear()
        info["choice"] = "guided" if choice == _abi.MODE_GUIDED else "plain"
        while el < budget_s:
            one(choice, 0, False)
        info.update(passes=n, spp=n * self.pass_spp, seconds=el)
        return acc / max(n, 1), info
