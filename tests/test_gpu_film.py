"""The GPU tonemapper (rt_film_tonemap) against the oracle's
Tonemapper::map, byte for byte: random films over several decades, the
special values, values next to every grey-level boundary, sizes that are not
a multiple of the 4-value vector path, and a rendered film through the
device-pointer entry."""
import numpy as np
import pytest

from oracle.oracle import tonemap
from renderer_amd import MODE_UNBIASED, scenes

pytestmark = pytest.mark.gpu


def _values(seed):
    rng = np.random.default_rng(seed)
    parts = [rng.random(1 << 20) * 1.5,
             10.0 ** rng.uniform(-8, 3, 1 << 20),
             np.array([0, -0.0, 1, -1, np.nan, np.inf, -np.inf, 1e-45, 3.4e38, 255, 0.5])]
    k = np.arange(256, dtype=np.float64)
    for g in (2.2, 1.0, 1.8):
        b = ((k / 255.0) ** g).astype(np.float32)
        parts += [b, np.nextafter(b, np.float32(np.inf)), np.nextafter(b, np.float32(-np.inf))]
    v = np.concatenate([p.astype(np.float32) for p in parts])
    n = (len(v) // 3) * 3
    return v[:n].reshape(-1, 1, 3)


@pytest.mark.parametrize("gamma", [2.2, 1.0, 1.8, 0.5])
def test_tonemap_matches_oracle(gpu_ctx, gamma):
    film = _values(int(gamma * 10))
    print(f"tonemap gamma {gamma}: {film.size} values", flush=True)
    got = gpu_ctx.tonemap(film, gamma)
    ref = tonemap(film, gamma)
    bad = np.flatnonzero(got.reshape(-1) != ref.reshape(-1))
    assert len(bad) == 0, (film.reshape(-1)[bad[:8]], got.reshape(-1)[bad[:8]], ref.reshape(-1)[bad[:8]])


@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (3, 4), (33, 17)])
def test_tonemap_odd_sizes(gpu_ctx, shape):
    film = np.random.default_rng(3).random((*shape, 3)).astype(np.float32) * 1.3
    assert np.array_equal(gpu_ctx.tonemap(film, 2.2), tonemap(film, 2.2))


def test_tonemap_device_of_a_rendered_film(gpu_ctx):
    import torch

    sc = scenes.box_scene(96, 64)
    gpu_ctx.upload(sc.flatten())
    film = torch.zeros((64, 96, 3), dtype=torch.float32, device="cuda:0")
    out = torch.zeros((64, 96, 3), dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    gpu_ctx.render_device([(sc.camera, 2, 2)], [film.data_ptr()], mode=MODE_UNBIASED, max_depth=8, seed=4)
    gpu_ctx.tonemap_device(film.data_ptr(), 96, 64, out.data_ptr(), 2.2)
    f = film.cpu().numpy()
    assert f.mean() > 0
    assert np.array_equal(out.cpu().numpy(), tonemap(f, 2.2))
    assert gpu_ctx.stats()["tonemap_ms"] > 0


def test_tonemap_every_value_undecided(gpu_ctx):
    """A film of values next to grey-level boundaries only: the float
    estimate decides none of them, so each block's list of values for the
    double-precision path fills up and the rest are evaluated in place."""
    k = np.arange(1, 256, dtype=np.float64)
    b = ((k / 255.0) ** 2.2).astype(np.float32)
    v = np.concatenate([b, np.nextafter(b, np.float32(np.inf)), np.nextafter(b, np.float32(-np.inf))])
    film = np.tile(v, (1 << 21) // len(v) + 1)[: 3 * (1 << 19)].reshape(-1, 512, 3)
    got = gpu_ctx.tonemap(film, 2.2)
    assert np.array_equal(got, tonemap(film, 2.2))
