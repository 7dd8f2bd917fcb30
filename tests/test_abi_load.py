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
