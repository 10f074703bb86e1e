"""Film outputs on the CPU: the tonemapper oracle against closed forms, and
the writers (include/rt_film.h) of --save-img (read back by PIL, an
independent decoder), --save-exr (read back by a reader of the OpenEXR
layout written from the file-format description), --save-mat (a text
derived by hand from src/main.cpp:231-258) and the output-name rule."""
import os
import struct

import numpy as np
import pytest

from oracle.oracle import tonemap
from renderer_amd import _abi
from renderer_amd.film import unique_filename, write_exr, write_image, write_mat

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_tonemap_oracle_closed_forms():
    # 255 * c^(1/gamma), truncated, clamped to [0, 255]; NaN and negatives (NaN pow) -> 0
    x = np.array([0, 1, 0.5, 0.25, -1, np.nan, np.inf, -np.inf, 2, 1e-30, 0.9999999], np.float32)
    assert tonemap(x, 1.0).tolist() == [0, 255, 127, 63, 0, 0, 255, 0, 255, 0, 254]
    # gamma 0.5: the square, exact in double: 255 * 0.25 = 63.75, 255 * 0.0625 = 15.9375
    assert tonemap(np.float32([0.5, 0.25]), 0.5).tolist() == [63, 15]
    # gamma 2.2 at the grey levels k: c = ((k + 0.5) / 255)^2.2 maps back into [k, k + 1)
    k = np.arange(256)
    c = (((k + 0.5) / 255.0) ** 2.2).astype(np.float32)
    assert tonemap(c, 2.2).tolist() == [min(int(v), 255) for v in k]
    # monotone in the film value
    v = np.sort(np.random.default_rng(0).random(10000).astype(np.float32) * 1.2)
    t = tonemap(v, 2.2).astype(int)
    assert np.all(np.diff(t) >= 0)


def test_image_writers_read_back(tmp_path):
    from PIL import Image

    img = (np.random.default_rng(1).random((13, 17, 3)) * 256).astype(np.uint8)
    for ext in ("png", "ppm", "bmp", "PNG"):
        p = tmp_path / f"img.{ext}"
        write_image(p, img)
        back = np.asarray(Image.open(p).convert("RGB"))
        assert np.array_equal(back, img), ext
    with pytest.raises(_abi.RtError) as e:
        write_image(tmp_path / "img.jpg", img)
    assert e.value.status == _abi.ERR_UNSUPPORTED
    with pytest.raises(_abi.RtError) as e:
        write_image(tmp_path / "no" / "such" / "dir.png", img)
    assert e.value.status == _abi.ERR_IO


def read_exr(path):
    """Single-part scanline OpenEXR, uncompressed FLOAT channels."""
    b = open(path, "rb").read()
    assert struct.unpack_from("<I", b, 0)[0] == 20000630
    assert struct.unpack_from("<I", b, 4)[0] & 0xff == 2
    at, attrs = 8, {}
    while b[at] != 0:
        e = b.index(0, at)
        name = b[at:e].decode()
        e2 = b.index(0, e + 1)
        typ = b[e + 1:e2].decode()
        size = struct.unpack_from("<i", b, e2 + 1)[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + size])
        at = e2 + 5 + size
    at += 1
    chans, v = [], attrs["channels"][1]
    i = 0
    while v[i] != 0:
        e = v.index(0, i)
        name = v[i:e].decode()
        ptype, = struct.unpack_from("<i", v, e + 1)
        chans.append((name, ptype))
        i = e + 1 + 16
    xmin, ymin, xmax, ymax = struct.unpack("<4i", attrs["dataWindow"][1])
    assert attrs["compression"][1] == b"\x00" and attrs["lineOrder"][1] == b"\x00"
    assert attrs["displayWindow"][1] == attrs["dataWindow"][1]
    w, h = xmax - xmin + 1, ymax - ymin + 1
    offsets = struct.unpack_from(f"<{h}Q", b, at)
    planes = {name: np.zeros((h, w), np.float32) for name, _ in chans}
    for y in range(h):
        o = offsets[y]
        yy, n = struct.unpack_from("<ii", b, o)
        data = np.frombuffer(b, np.float32, count=n // 4, offset=o + 8).reshape(len(chans), w)
        for k, (name, ptype) in enumerate(chans):
            assert ptype == 2
            planes[name][yy - ymin] = data[k]
    return chans, planes


def test_exr_round_trip(tmp_path):
    film = np.random.default_rng(2).standard_normal((9, 11, 3)).astype(np.float32) * 100
    film[0, 0] = [np.inf, -0.0, 1e-40]
    write_exr(tmp_path / "f.exr", film)
    chans, planes = read_exr(tmp_path / "f.exr")
    assert [c for c, _ in chans] == ["B", "G", "R"]
    for name, ch in (("R", 0), ("G", 1), ("B", 2)):
        assert np.array_equal(planes[name].view(np.uint32), film[..., ch].view(np.uint32))


def test_mat_text_matches_the_hand_derived_file(tmp_path):
    film = np.array([[[0.5, 0.25, 1], [2, -0.125, 0]], [[1e-7, 3, 0.75], [10, 100.5, -0.0]]], np.float32)
    write_mat(tmp_path / "film_2x2.m", film)
    assert (tmp_path / "film_2x2.m").read_text() == open(os.path.join(GOLDEN, "film_2x2.m")).read()
    # the function name is the base name: up to the first '.'
    write_mat(tmp_path / "a.b.m", film)
    assert (tmp_path / "a.b.m").read_text().startswith("function im = a()\n")


def test_unique_filename(tmp_path):
    p = tmp_path / "out.png"
    assert unique_filename(p) == str(p)
    p.write_bytes(b"")
    assert unique_filename(p) == str(tmp_path / "out0.png")
    (tmp_path / "out0.png").write_bytes(b"")
    assert unique_filename(p) == str(tmp_path / "out1.png")
    q = tmp_path / "noext"
    q.write_bytes(b"")
    assert unique_filename(q) == str(tmp_path / "noext0")
