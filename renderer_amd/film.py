"""Film outputs (include/rt_film.h): the tonemapper on the GPU and the
writers of the reference's --save-exr, --save-img and --save-mat
(src/main.cpp:216-258).

    rgb8 = ctx.tonemap(film, gamma=2.2)            # Tonemapper::tonemap
    write_image(unique_filename("out.png"), rgb8)  # --save-img
    write_exr("out.exr", film)                     # --save-exr
    write_mat("out.m", film)                       # --save-mat

Films are (height, width, 3) float32 RGB arrays, as rt_render writes them.
"""
import ctypes

import numpy as np

from . import _abi


def _film(film):
    f = np.ascontiguousarray(film, np.float32)
    if f.ndim != 3 or f.shape[2] != 3:
        raise ValueError("film must be (height, width, 3)")
    return f


def _check(lib, st):
    if st != 0:
        raise _abi.RtError(int(st), lib.rt_film_last_error().decode(errors="replace"))


def write_exr(path, film):
    lib = _abi.load()
    f = _film(film)
    _check(lib, lib.rt_film_write_exr(str(path).encode(), f.ctypes.data, f.shape[1], f.shape[0]))


def write_mat(path, film):
    lib = _abi.load()
    f = _film(film)
    _check(lib, lib.rt_film_write_mat(str(path).encode(), f.ctypes.data, f.shape[1], f.shape[0]))


def write_image(path, rgb8):
    lib = _abi.load()
    a = np.ascontiguousarray(rgb8, np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("image must be (height, width, 3) uint8")
    _check(lib, lib.rt_film_write_image(str(path).encode(), a.ctypes.data, a.shape[1], a.shape[0]))


def unique_filename(path):
    lib = _abi.load()
    p = str(path).encode()
    buf = ctypes.create_string_buffer(len(p) + 32)
    n = lib.rt_unique_filename(p, buf, len(buf))
    _check(lib, n if n < 0 else 0)
    return buf.value.decode()
