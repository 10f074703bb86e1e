"""CPU-side checks of the C ABI library: it loads, exports every function the
headers declare, and reports errors without a GPU."""
import ctypes
import os
import re

import pytest

from renderer_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(rt_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_the_abi():
    names = _declared_functions()
    assert {"rt_create", "rt_destroy", "rt_scene_upload", "rt_trace_closest", "rt_trace_any", "rt_render",
            "rt_get_stats", "rt_set_fault"} <= names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    missing = [n for n in sorted(_declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes prototypes cover the whole header
    assert _declared_functions() <= set(_abi.PROTOTYPES)


def test_error_strings_and_invalid_arguments():
    lib = _abi.load()
    assert lib.rt_error_string(0) == b"ok"
    assert lib.rt_error_string(-6) == b"injected fault"
    assert lib.rt_version() >= 1
    assert lib.rt_destroy(None) == -1
    assert lib.rt_create(0, None) == -1
    assert lib.rt_set_fault(None, 1) == -1
    assert lib.rt_render(None, None, 0, None, None) == -1


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") == "", reason="no devices")
def test_create_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _abi.load()
    h = ctypes.c_void_p()
    assert lib.rt_create(0, ctypes.byref(h)) == -3
    assert not h.value
