"""The C-ABI library loads and exports every function the public headers declare.

No device call is made here: these tests run on machines without a GPU.
"""
import ctypes as C
import glob
import os
import re

import pytest

import cgfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(cgfd_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "cgfd_run" in names and "cgfd_create" in names
    assert len(names) >= 20


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    lib = cgfd.load()
    assert hasattr(lib, name), f"{name} declared but not exported"


def test_binding_covers_exports():
    assert set(cgfd.EXPORTS) <= set(declared_functions())


def test_footprint_is_host_only_and_scales():
    lib = cgfd.load()
    d1 = cgfd.Desc(100, 100, 60, cgfd.STORE_F32, 1, 0)
    d2 = cgfd.Desc(100, 100, 60, cgfd.STORE_F32, 4, 0)
    d3 = cgfd.Desc(100, 100, 60, cgfd.STORE_F16, 1, 0)
    b1 = lib.cgfd_footprint_bytes(C.byref(d1), 20, 501)
    b2 = lib.cgfd_footprint_bytes(C.byref(d2), 20, 501)
    b3 = lib.cgfd_footprint_bytes(C.byref(d3), 20, 501)
    # 4 levels x 9 components of (106 x 106 x 66) padded points, at least
    assert b1 > 4 * 9 * 100 * 106 * 66 * 4
    assert b2 > 3 * b1 // 2 and b3 < b1
    bad = cgfd.Desc(0, 10, 10, 0, 1, 0)
    assert lib.cgfd_footprint_bytes(C.byref(bad), 0, 0) < 0


def test_server_arguments_checked_before_device_use():
    with pytest.raises(cgfd.CgfdError, match="max_batch"):
        cgfd.Server(max_batch=0)
    with pytest.raises(cgfd.CgfdError, match="chunk_steps"):
        cgfd.Server(chunk_steps=12)


def test_grid_over_32bit_component_offsets_is_refused_before_device_use():
    """Kernels address a component with 32-bit byte offsets: FP64 at 800^3 or FP32 at 1024^3 must be refused."""
    lib = cgfd.load()
    for d in (cgfd.Desc(800, 800, 800, cgfd.STORE_F64, 1, 0), cgfd.Desc(1024, 1024, 1024, cgfd.STORE_F32, 1, 0)):
        out = C.c_void_p()
        assert lib.cgfd_create(C.byref(d), C.byref(out)) == -1          # CGFD_E_ARG, not CGFD_E_NODEV
        assert b"4 GiB" in lib.cgfd_last_error()
        assert not out.value
    # the C3 precision grid stays inside the limit in every storage type
    for st in (cgfd.STORE_F32, cgfd.STORE_F16, cgfd.STORE_BF16, cgfd.STORE_F64):
        ok = cgfd.Desc(800, 800, 400, st, 1, 0)
        out = C.c_void_p()
        assert lib.cgfd_create(C.byref(ok), C.byref(out)) != -1 or b"4 GiB" not in lib.cgfd_last_error()


def test_footprint_counts_the_storage_padded_model():
    """Model arrays share the wavefield layout: 16-bit storage pads rows to 64 elements, not 32."""
    lib = cgfd.load()
    d = cgfd.Desc(400, 400, 200, cgfd.STORE_F16, 1, 0)
    b = lib.cgfd_footprint_bytes(C.byref(d), 0, 0)
    px = 512   # 64 (i=0 at 128 bytes) + 400 + 3 -> next multiple of 64 elements
    cs = px * 406 * 206
    assert b >= 4 * 9 * cs * 2 + 13 * cs * 4
