"""CPU-side checks of the C-ABI library: it loads without a GPU and exports
every entry point include/ajisai_gpu.h declares (no compute call is made)."""
import ctypes as C
import os
import re

import pytest

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


@pytest.mark.parametrize("lang", ["c", "cpp"])
def test_header_compiles_as_c_and_cpp_and_links(tmp_path, lang):
    """include/ajisai_gpu.h compiles as strict C99 and C++17 (-Wall -Wextra
    -Werror -pedantic); a program linked against libajisai_gpu.so calls the
    entry points that need no GPU (ABI version, context creation's error
    path, null-handle errors, camera set-up) and exits 0."""
    import shutil
    import subprocess
    cc = shutil.which("gcc" if lang == "c" else "g++")
    if cc is None:
        pytest.skip("no host compiler")
    src = os.path.join(ROOT, "tests", "abi", "abi_c.c" if lang == "c" else "abi_cpp.cpp")
    std = "-std=c99" if lang == "c" else "-std=c++17"
    exe = str(tmp_path / f"abi_{lang}")
    libdir = os.path.dirname(capi.LIB_PATH)
    subprocess.run([cc, std, "-Wall", "-Wextra", "-Werror", "-pedantic",
                    "-I" + os.path.join(ROOT, "include"), src, "-L" + libdir, "-lajisai_gpu",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    assert "abi ok" in out.stdout
