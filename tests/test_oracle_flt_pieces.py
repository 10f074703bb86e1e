"""Pins the pieces of the oracle's filtered Metropolis renderer
(src/metropolisfltrenderer.cpp, src/symmetricfilter.cpp) that no reference
vectors cover, each against something that does not share its code:

* cv::blur with BORDER_REFLECT (the symmetric filter's patch sums) against
  scipy.ndimage.uniform_filter in 'reflect' mode, which extends an image the
  same way (edge sample repeated), also where the window wraps past the
  image more than once;
* the [.1 .8 .1] sepFilter2D with BORDER_REFLECT_101 (the simple variance
  filter) against scipy.ndimage.correlate1d in 'mirror' mode;
* mathhelper var() of two films against its closed form ((a - b) / 2)^2;
* the importance map of a later pass (:231-239) at relative variances whose
  exponentials are exact by hand (e^-ln2 = 1/2, e^0 = 1, e^-inf = 0), its
  normalisation to mean 1 and its monotonicity;
* addSample's weighted running mean and sum of weighted squared deviations
  (West's update, :54-75) against the direct weighted formulas over the same
  deposits, and the rescale of a later pass as the linear reweighting of the
  samples so far."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle


@pytest.mark.parametrize("shape,p", [((5, 6), 7), ((17, 13), 3), ((17, 13), 7), ((3, 2), 9), ((64, 48), 15)])
def test_box_blur_is_the_reflect_box_filter(shape, p):
    img = np.random.default_rng(p * shape[0]).standard_normal(shape).astype(np.float32)
    got = oracle.box_blur(img, p)
    ref = ndimage.uniform_filter(img.astype(np.float64), size=p, mode="reflect")
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6 * p)


@pytest.mark.parametrize("shape", [(2, 2), (5, 7), (33, 17)])
def test_sep3_is_the_mirror_filter(shape):
    img = np.random.default_rng(shape[1]).random((*shape, 3)).astype(np.float32)
    got = oracle.sep3_reflect101(img)
    k = np.array([0.1, 0.8, 0.1])
    ref = ndimage.correlate1d(img.astype(np.float64), k, axis=1, mode="mirror")
    ref = ndimage.correlate1d(ref, k, axis=0, mode="mirror")
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)


def test_var2_closed_form():
    rng = np.random.default_rng(5)
    a = (rng.random(4096) * 10 - 2).astype(np.float32)
    b = (rng.random(4096) * 10 - 2).astype(np.float32)
    m, v = oracle.var2(a, b)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    np.testing.assert_allclose(m, (a64 + b64) / 2, rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(v, ((a64 - b64) / 2) ** 2, rtol=1e-4, atol=2e-5)
    assert (v >= 0).all()
    m, v = oracle.var2(a, a)
    assert (v == 0).all() and np.array_equal(m, a)


def _img(rows):
    """(H, W, 3) images of (variance, mean) pairs given per pixel, equal channels."""
    v = np.array([[[r[0]] * 3 for r in row] for row in rows], np.float32)
    m = np.array([[[r[1]] * 3 for r in row] for row in rows], np.float32)
    return v, m


def test_importance_closed_form():
    # relative variance vm / (mm^2 + 0.01): ln 2, 0 and a huge value give
    # raw importances 21 - 20 e^-rel = 11, 1 and 21; normalised by 3 / 33
    mm = 0.1
    v, m = _img([[(np.log(2.0) * (mm * mm + 1e-2), mm), (0.0, mm), (1e9, mm)]])
    imp = oracle.flt_importance(v, m)
    np.testing.assert_allclose(imp[0], [1.0, 1 / 11, 21 / 11], rtol=2e-5)
    # the 0.01 in the denominator: a zero mean, variance 0.01 ln 2 -> raw 11
    v, m = _img([[(0.01 * np.log(2.0), 0.0), (0.0, 5.0)]])
    np.testing.assert_allclose(oracle.flt_importance(v, m)[0], [11 / 6, 1 / 6], rtol=2e-5)


def test_importance_normalised_and_monotone():
    rng = np.random.default_rng(9)
    h, w = 24, 31
    var = (rng.random((h, w)) ** 4 * 3).astype(np.float32)
    mean = (rng.random((h, w)) + 0.05).astype(np.float32)
    v = np.repeat(var[..., None], 3, 2)
    m = np.repeat(mean[..., None], 3, 2)
    imp = oracle.flt_importance(v, m)
    assert abs(imp.mean() - 1.0) < 1e-5
    assert imp.max() / imp.min() <= 21.0 + 1e-3
    rel = var.astype(np.float64) / (mean.astype(np.float64) ** 2 + 1e-2)
    order = np.argsort(rel.ravel(), kind="stable")
    assert (np.diff(imp.ravel()[order]) >= -1e-6).all()  # more relative variance, more importance


def _state(h, w):
    return {"film": np.zeros((h, w, 3), np.float32), "mean": np.zeros((h, w, 3), np.float32),
            "m2": np.zeros((h, w, 3), np.float32), "sw": np.zeros((h, w), np.float32)}


def test_add_sample_is_the_weighted_mean_and_deviation():
    rng = np.random.default_rng(11)
    h, w, n = 4, 5, 600
    pos = rng.integers(0, h * w - 2, n)  # the last two pixels get no deposit
    wt = (rng.random(n) * 0.9 + 0.1).astype(np.float32)
    rgb = (rng.random((n, 3)) * 4).astype(np.float32)
    st = _state(h, w)
    oracle.flt_accumulate(st, pos, wt, rgb)
    for k in range(h * w):
        sel = pos == k
        y, x = divmod(k, w)
        if not sel.any():
            assert st["sw"][y, x] == 0 and (st["film"][y, x] == 0).all() and (st["m2"][y, x] == 0).all()
            continue
        ww = wt[sel].astype(np.float64)
        vv = rgb[sel].astype(np.float64)
        sw = ww.sum()
        mean = (ww[:, None] * vv).sum(0) / sw
        m2 = (ww[:, None] * (vv - mean) ** 2).sum(0)
        np.testing.assert_allclose(st["sw"][y, x], sw, rtol=1e-5)
        np.testing.assert_allclose(st["film"][y, x], (ww[:, None] * vv).sum(0), rtol=1e-5)
        np.testing.assert_allclose(st["mean"][y, x], mean, rtol=1e-4, atol=1e-5)
        np.testing.assert_allclose(st["m2"][y, x], m2, rtol=1e-3, atol=1e-4)


def test_rescale_reweights_the_samples_so_far():
    rng = np.random.default_rng(12)
    h, w, n = 3, 3, 200
    pos = rng.integers(0, h * w, n)
    wt = (rng.random(n) + 0.1).astype(np.float32)
    rgb = (rng.random((n, 3)) * 2).astype(np.float32)
    st = _state(h, w)
    oracle.flt_accumulate(st, pos, wt, rgb)
    before = {k: v.copy() for k, v in st.items()}
    sumimp = (rng.random((h, w)) + 0.5).astype(np.float32)
    newsum = sumimp + (rng.random((h, w)) + 0.5).astype(np.float32)
    r = (sumimp / newsum)[..., None].astype(np.float64)
    oracle.flt_accumulate(st, [], [], np.zeros((0, 3), np.float32), sumimp, newsum)
    # every weight so far scaled by r: the weighted sums scale by r, the
    # squared-deviation sums by r^2 (as the reference scales them), the
    # running mean and the weight sum are kept
    np.testing.assert_allclose(st["film"], r * before["film"], rtol=1e-6)
    np.testing.assert_allclose(st["m2"], r * r * before["m2"], rtol=1e-6)
    assert np.array_equal(st["mean"], before["mean"]) and np.array_equal(st["sw"], before["sw"])
    # deposits after the rescale add to the reweighted sums
    p2 = rng.integers(0, h * w, 50)
    w2 = (rng.random(50) + 0.1).astype(np.float32)
    c2 = rng.random((50, 3)).astype(np.float32)
    film = st["film"].astype(np.float64).copy()
    oracle.flt_accumulate(st, p2, w2, c2)
    for k in range(h * w):
        y, x = divmod(k, w)
        sel = p2 == k
        np.testing.assert_allclose(st["film"][y, x], film[y, x] + (w2[sel, None] * c2[sel]).sum(0), rtol=1e-5)
