"""The C-ABI library loads and exports every entry point include/d3fgpu.h declares.
No compute calls: this runs without a GPU."""
import ctypes
import os
import re

import d3fgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "d3fgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(d3fgpu.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_bindings_cover_the_header():
    assert sorted(d3fgpu.SIGNATURES) == declared()


def test_abi_version():
    assert d3fgpu.lib().fg_abi_version() == 1


def test_null_context_is_rejected():
    L = d3fgpu.lib()
    assert L.fg_destroy(None) == -1
    assert L.fg_sync(None) == -1
    assert L.fg_set_seed(None, 1) == -1


def test_force_descriptor_layout_matches_the_oracle():
    import oracle
    assert ctypes.sizeof(d3fgpu.ForceDesc) == ctypes.sizeof(oracle.ForceDesc)
    for (a, _), (b, _) in zip(d3fgpu.ForceDesc._fields_, oracle.ForceDesc._fields_):
        assert a == b
