"""CPU-side checks of the C-ABI library: it loads without a GPU and exports
every entry point include/ajisai_gpu.h declares (no compute call is made)."""
import ctypes as C
import os
import re

from ajisai_render_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "ajisai_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ajg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_entry_points():
    assert declared_entry_points() == capi.exported_symbols()


def test_library_loads_and_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    for name in declared_entry_points():
        assert hasattr(lib, name), name
    assert lib.ajg_abi_version() == 1


def test_camera_setup_host_only():
    from ajisai_render_amd.core import camera_setup
    from ajisai_render_amd import scenes
    cam = camera_setup(scenes.cornell().camera)
    assert cam.resolution[0] == 256
