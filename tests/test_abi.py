"""The C-ABI library loads without a GPU and exports every entry point that
include/muonx.h declares."""
import os
import re

import muon_amd
from muon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(muonx_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(muon_amd.lib, n), n


def test_bindings_cover_the_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared_functions() <= bound


def test_version_and_error_string():
    assert b"gfx950" in muon_amd.lib.muonx_version()
    assert isinstance(muon_amd.lib.muonx_last_error(), bytes)


def test_invalid_arguments_are_reported_not_thrown():
    rc = muon_amd.lib.muonx_host_scene_parse_file(b"/nonexistent/scene.muon", None, None)
    assert rc == -1
    import ctypes as C
    out = C.c_void_p()
    rc = muon_amd.lib.muonx_host_scene_parse_file(b"/nonexistent/scene.muon", None, C.byref(out))
    assert rc == -4 and b"cannot open" in muon_amd.lib.muonx_last_error()
