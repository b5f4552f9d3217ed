"""The C ABI library loads and exports every entry point include/aurora_amd.h
declares; argument errors come back as status codes (no compute calls)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import pytest

from aurora_amd import native


def test_library_exports_every_header_symbol():
    lib = native.load()
    syms = native.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for s in native.header_symbols():
        assert s in exported, s  # unmangled: extern "C"


def test_abi_version():
    assert native.load().aur_abi_version() == native.ABI_VERSION


def test_null_context_is_an_error():
    lib = native.load()
    P = ctypes.c_void_p
    assert lib.aur_synchronize(None) == 2
    assert lib.aur_geometry_upload(None, 2, None, None, None, None) == 2
    assert lib.aur_rebuild_nmh(None, None) == 2
    assert lib.aur_frame_create(None, 4, 4, 1) == 2
    assert lib.aur_frame_read(None, None) == 2
    assert lib.aur_create(0, None) == 2
    assert lib.aur_last_error(None) == b"null context"
    assert lib.aur_stream(None) is None
    assert lib.aur_geometry_count(None) == 0
    lib.aur_destroy(None)  # no-op
    del P


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is present")
def test_create_without_device_reports_device_error():
    lib = native.load()
    h = ctypes.c_void_p()
    assert lib.aur_create(0, ctypes.byref(h)) == 3
    assert not h.value


def test_product_path_has_no_oracle_dependency():
    """The shipped library must not link or load the CPU oracle."""
    out = subprocess.run(["ldd", native.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    src_dir = os.path.join(os.path.dirname(native.LIB_PATH), "csrc")
    for name in os.listdir(src_dir):
        text = open(os.path.join(src_dir, name)).read()
        assert "oracle" not in text.lower(), name
    for mod in ("native.py", "renderer.py", "scenes.py"):
        text = open(os.path.join(os.path.dirname(native.LIB_PATH), mod)).read()
        assert "oracle" not in text, mod


ABI_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi")


def _gpu_present() -> bool:
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK)


def test_header_compiles_as_c_and_cpp_and_links():
    """tests/abi: the header as C99 (-pedantic -Werror) and C++17, record
    sizes and offsets against the reference's layouts (static asserts),
    every declared entry point linked from C, and the C++ Renderer class;
    aur_create reports AUR_ERR_DEVICE where no GPU is present."""
    subprocess.run(["make", "-s", "-B", "-C", ABI_DIR], check=True, capture_output=True)
    out = subprocess.run([os.path.join(ABI_DIR, "c_consumer")], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out[:2] == ["c-consumer", "ok"]
    assert int(out[2]) == len(native.header_symbols())
    assert int(out[4]) == native.ABI_VERSION
    if not _gpu_present():
        assert int(out[3]) == 3
        res = subprocess.run([os.path.join(ABI_DIR, "cpp_consumer")], capture_output=True, text=True)
        assert res.returncode == 0 and res.stdout.split() == ["cpp-consumer", "no-device", "3"]
