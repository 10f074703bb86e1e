"""Environment maps loaded the reference's way: Radiance .hdr files
(src/environmentmap.cpp:20-33) read by the library, and the angular light
probe lookup of EnvironmentMap::get (src/environmentmap.cpp:35-66).  Pins:
RGBE files written here from known values (flat and run-length scanlines),
probe directions whose image point is known in closed form (the probe
axis at the centre, directions perpendicular to it on the circle of radius
one half), white furnaces under a constant probe in both estimators, and
the sampling density built from the probe integrating to one."""
import math
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import _abi
from renderer_amd.scene import (Camera, DiffuseMaterial, EnvironmentMap, IntersectableList, Scene, Sphere)


def _rgbe_bytes(w, h, pixels, rle):
    """pixels: (h, w, 4) uint8 RGBE"""
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        row = pixels[y]
        if not rle:
            out += row.astype(np.uint8).tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for ch in range(4):
            vals = row[:, ch].tolist()
            x = 0
            while x < w:   # alternate a run of equal bytes and a literal stretch
                r = 1
                while x + r < w and vals[x + r] == vals[x] and r < 127:
                    r += 1
                if r > 2:
                    out += bytes([128 + r, vals[x]])
                    x += r
                else:
                    n = min(w - x, 3)
                    out += bytes([n] + vals[x:x + n])
                    x += n
    return bytes(out)


@pytest.mark.parametrize("rle", [False, True])
def test_rgbe_reader(tmp_path, rle):
    w, h = 12, 3
    px = np.zeros((h, w, 4), np.uint8)
    px[..., 0], px[..., 1], px[..., 2], px[..., 3] = 128, 64, 32, 129      # (1, .5, .25) at 2^(129-136)*128
    px[1, 2:9] = (200, 100, 50, 140)                                     # a run
    px[2, 5] = (7, 8, 9, 0)                                              # e = 0 -> black
    p = tmp_path / "t.hdr"
    p.write_bytes(_rgbe_bytes(w, h, px, rle))
    img = EnvironmentMap.load(p).image
    assert img.shape == (h, w, 3)
    assert np.array_equal(img[0, 0], [1.0, 0.5, 0.25])
    assert np.array_equal(img[1, 3], np.float32([200, 100, 50]) * 2.0 ** (140 - 136))
    assert np.array_equal(img[2, 5], [0, 0, 0])
    bad = tmp_path / "bad.hdr"
    bad.write_bytes(b"P6\n1 1\n255\n\x00\x00\x00")
    with pytest.raises(_abi.RtError) as e:
        EnvironmentMap.load(bad)
    assert e.value.status == _abi.ERR_IO


def _probe_oracle(img, coef=(1, 1, 1)):
    env = EnvironmentMap(img, coef, mapping="probe")
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.45, 8, 8), [env], IntersectableList([]))
    return Oracle(sc.flatten())


def test_probe_lookup_closed_form_points():
    W = H = 64
    img = np.zeros((H, W, 3), np.float32)
    img[..., 0] = np.arange(W)[None, :]       # column index in red
    img[..., 1] = np.arange(H)[:, None]       # row index in green
    orc = _probe_oracle(img)

    def texel(d):
        v = orc.env_value(np.asarray(d, np.float32) / np.linalg.norm(d))
        return int(v[1]), int(v[0])  # row, column
    # directions perpendicular to the probe axis (+x) land on radius 1/2:
    # column ((c + 1) / 2) * W - 1 for the image point c = (d.y, -d.z) / 2
    assert texel((0, 1, 0)) == (H // 2 - 1, int(0.75 * W) - 1)
    assert texel((0, -1, 0)) == (H // 2 - 1, int(0.25 * W) - 1)
    assert texel((0, 0, -1)) == (int(0.75 * H) - 1, W // 2 - 1)
    assert texel((0, 0, 1)) == (int(0.25 * H) - 1, W // 2 - 1)
    # the probe axis maps to the centre, the opposite direction to the rim
    assert texel((1, 1e-4, 0)) == (H // 2 - 1, W // 2 - 1)
    assert texel((-1, 1e-4, 0)) == (H // 2 - 1, W - 1)
    # 45 degrees from the axis: radius 1/4
    assert texel((1, 1, 0)) == (H // 2 - 1, int(0.625 * W) - 1)


@pytest.mark.parametrize("mode,expect", [(1, 0.6), (0, 2 * 0.6 / (3 * math.pi))])
def test_white_furnace_under_a_constant_probe(mode, expect):
    env = EnvironmentMap(np.ones((32, 32, 3), np.float32), mapping="probe")
    sc = Scene(Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.45, 24, 24), [env],
               IntersectableList([Sphere((0, 0, 0), 0.5, DiffuseMaterial((0.6, 0.6, 0.6)))]))
    img, _ = Oracle(sc.flatten()).render((sc.camera, 16, 16), mode=mode, max_depth=5 if mode == 0 else 4, seed=1)
    inside = np.zeros((24, 24), bool)
    inside[9:14, 9:14] = True
    assert img[inside].mean() == pytest.approx(expect, rel=0.01)


def test_probe_sampling_density_integrates_to_one():
    rng = np.random.default_rng(4)
    img = (rng.random((48, 48, 3)) * 3).astype(np.float32)
    img[20:24, 30:34] = 400.0  # a sun
    orc = _probe_oracle(img)
    for u0, u1 in rng.random((100, 2)):
        d, p = orc.env_sample(u0, u1)
        assert orc.env_pdf(d) == pytest.approx(p, rel=2e-3)
    v = rng.normal(size=(80000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    est = np.mean([orc.env_pdf(d) for d in v]) * 4 * math.pi
    assert est == pytest.approx(1.0, abs=0.05)


CAMPUS_PROBE = "/root/reference/environmentmaps/campus_probe.hdr"


def _decode_rgbe_rle(data):
    """Independent reader of Radiance files with new-style run-length
    scanlines (the format of the reference's probe): header lines up to an
    empty line, '-Y h +X w', then per scanline the marker (2, 2, w >> 8,
    w & 255) and the four channels each as runs (count > 128: count - 128
    copies of one byte) and literals (count bytes).  Values byte * 2^(e - 136)."""
    end = data.index(b"\n\n") + 2
    res_end = data.index(b"\n", end)
    ysign, h, xsign, w = data[end:res_end].split()
    assert ysign == b"-Y" and xsign == b"+X"
    h, w = int(h), int(w)
    pos = res_end + 1
    rgbe = np.zeros((h, w, 4), np.uint8)
    buf = np.frombuffer(data, np.uint8)
    for y in range(h):
        assert tuple(buf[pos:pos + 4]) == (2, 2, w >> 8, w & 255)
        pos += 4
        for ch in range(4):
            x = 0
            while x < w:
                c = int(buf[pos])
                pos += 1
                if c > 128:
                    rgbe[y, x:x + c - 128, ch] = buf[pos]
                    pos += 1
                    x += c - 128
                else:
                    rgbe[y, x:x + c, ch] = buf[pos:pos + c]
                    pos += c
                    x += c
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136))
    return (rgbe[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)


@pytest.mark.skipif(not os.path.exists(CAMPUS_PROBE), reason="the reference tree is not present")
def test_reference_campus_probe_decodes():
    """The reference's one data asset, the 640 x 640 angular light probe that
    src/environmentmap.cpp:20-33 loads: the library's reader agrees bit for
    bit with the independent decoder above, the values are finite and not
    negative, and the probe's disk holds almost all of the energy (an
    angular map uses only the inscribed circle; its corners are nearly
    black)."""
    env = EnvironmentMap.load(CAMPUS_PROBE)
    with open(CAMPUS_PROBE, "rb") as f:
        want = _decode_rgbe_rle(f.read())
    assert env.image.shape == (640, 640, 3)
    assert np.array_equal(env.image.view(np.uint32), want.view(np.uint32))
    assert np.all(np.isfinite(env.image)) and env.image.min() >= 0
    c = (np.arange(640) + 0.5) / 320.0 - 1.0
    r = np.hypot(c[None, :], c[:, None])
    assert env.image[r <= 1].sum() > 0.99 * env.image.sum()
    assert env.image[r < 0.98].mean() > 0
