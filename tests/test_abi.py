"""The C-ABI library loads, exports every declared symbol and validates its
arguments (CPU only: no compute call needs a device)."""
import ctypes
import os
import re

import pytest

import glasscamerapublic_amd as gcam
from glasscamerapublic_amd import _lib

HEADER_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gcam")


def declared_functions():
    names = set()
    for f in sorted(os.listdir(HEADER_DIR)):
        if f.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(HEADER_DIR, f)).read(), flags=re.S)
            names.update(re.findall(r"\b(gcam_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_what_the_binding_exports():
    assert declared_functions() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = gcam.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.gcam_abi_version() == 3


def test_status_strings():
    L = gcam.lib()
    assert L.gcam_status_string(0) == b"ok"
    assert L.gcam_status_string(1) == b"invalid argument"
    assert L.gcam_status_string(99) == b"unknown status"


@pytest.mark.parametrize("w,h,levels,radius,precision,msg", [
    (64, 64, 5, 10, 0, b"too small"),      # 4x4 coarsest level: no two 5x5 patches
    (1024, 1024, 0, 10, 0, b"levels"),
    (1024, 1024, 9, 10, 0, b"levels"),
    (1024, 1024, 5, 0, 0, b"search_radius"),
    (1024, 1024, 5, 11, 0, b"search_radius"),
    (0, 1024, 5, 4, 0, b"size"),
    (20000, 1024, 5, 4, 0, b"size"),
])
def test_create_rejects_bad_arguments(w, h, levels, radius, precision, msg):
    L = gcam.lib()
    cfg = _lib.ConfigC(radius, precision)
    handle = ctypes.c_void_p()
    st = L.gcam_pyramid_create(w, h, levels, ctypes.byref(cfg), ctypes.byref(handle))
    assert st == 1
    assert msg in L.gcam_last_error()
    assert not handle.value


def test_null_arguments_are_errors_not_crashes():
    L = gcam.lib()
    assert L.gcam_pyramid_create(64, 64, 2, None, None) == 1
    assert L.gcam_pyramid_denoise(None, None, None, None, None, 1.0, 0.0, 0, None, None) == 1
    assert L.gcam_downsample_image(None, None, None) == 1
    img = _lib.Image(0, 16, 16, 16, 0)
    assert L.gcam_basic_noise_statistics(ctypes.byref(img), ctypes.byref(img), None) == 1
    assert b"NULL" in L.gcam_last_error()
    L.gcam_pyramid_destroy(None)


def test_reserved_config_fields_must_be_zero():
    L = gcam.lib()
    cfg = _lib.ConfigC(4, 0)
    cfg.reserved[3] = 7
    handle = ctypes.c_void_p()
    assert L.gcam_pyramid_create(1024, 1024, 5, ctypes.byref(cfg), ctypes.byref(handle)) == 1
    assert b"reserved" in L.gcam_last_error()


def test_product_does_not_import_the_oracle():
    pkg = os.path.dirname(gcam.__file__)
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(root, f)).read()
                assert "import oracle" not in text and "gcam_oracle" not in text, f


def test_cpp_facade_compiles_and_links(tmp_path):
    """gcam::PyramidProcessor<5> (the reference-shaped C++ surface) compiles and
    links against the shared library with a plain host compiler."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "facade.cpp"
    src.write_text(
        '#include "gcam/pyramid_processor.hpp"\n'
        "int main() {\n"
        "  std::array<gcam::DenoiseParameters, 5> p; p.fill(gcam::defaultDenoiseParameters);\n"
        "  try { gcam::PyramidProcessor<5> proc(4032, 3024, 4); (void)proc; }\n"
        "  catch (const gcam::Error& e) { return e.status == GCAM_ERR_NO_DEVICE ? 0 : 1; }\n"
        "  return 0;\n"
        "}\n")
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = tmp_path / "facade"
    subprocess.run(["g++", "-std=c++17", f"-I{repo}/include", "-D__HIP_PLATFORM_AMD__", str(src), "-o", str(exe),
                    f"-L{libdir}", "-lgcam_denoise", f"-Wl,-rpath,{libdir}"], check=True)


def test_burst_create_rejects_bad_arguments_without_a_device():
    """Argument checks of gcam_burst_create that run before any device call."""
    import ctypes
    from glasscamerapublic_amd._lib import ConfigC, lib
    L = lib()
    h = ctypes.c_void_p()
    W = (ctypes.c_int32 * 2)(64, 64)
    H = (ctypes.c_int32 * 2)(64, 64)
    assert L.gcam_burst_create(0, W, H, 2, None, ctypes.byref(h)) == 1  # n_frames out of range
    assert L.gcam_burst_create(65, W, H, 2, None, ctypes.byref(h)) == 1
    assert L.gcam_burst_create(2, None, H, 2, None, ctypes.byref(h)) == 1
    cfg = ConfigC(4, 0, 0, 0)  # FP32: the burst runs the precision form
    assert L.gcam_burst_create(2, W, H, 2, ctypes.byref(cfg), ctypes.byref(h)) == 4
    cfg = ConfigC(4, 1, 0, 1)  # bilateral
    assert L.gcam_burst_create(2, W, H, 2, ctypes.byref(cfg), ctypes.byref(h)) == 4
    assert L.gcam_burst_create(2, W, H, 2, None, None) == 1
    assert not h.value
