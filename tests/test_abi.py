"""The C-ABI library loads without a GPU and exports every entry point that
include/muonx.h declares; the Python ctypes mirrors have the header's struct
layouts."""
import ctypes as C
import os
import re

import pytest

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


MIRRORS = [("muonx_options", "Options"), ("muonx_scene_desc", "SceneDesc"), ("muonx_render_params", "RenderParams"),
           ("muonx_trace_stats", "TraceStatsC"), ("muonx_scene_info", "SceneInfo"), ("muonx_tuning", "Tuning"),
           ("muonx_autotune_report", "AutotuneReport"), ("muonx_query", "Query"),
           ("muonx_engine_config", "EngineConfig"), ("muonx_job", "Job"), ("muonx_job_result", "JobResult"),
           ("muonx_server_stats", "ServerStats")]


def test_python_struct_mirrors_match_the_header(tmp_path):
    """Every ctypes mirror in muon_amd/_lib.py has the size and field offsets
    of its C struct, as a C compiler lays the header out."""
    import shutil
    import subprocess
    from muon_amd import _lib
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "muonx.h"', "int main(void) {"]
    for cname, pyname in MIRRORS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in getattr(_lib, pyname)._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, pyname in MIRRORS:
        cls = getattr(_lib, pyname)
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_python_constants_mirror_the_header():
    """The integer constants the Python package restates carry the header's
    values: _lib's MUONX_* names, serving's SLICES_* orders (MUONX_SLICES_*),
    and the traversal mode DeviceScene.query passes as int(exact)."""
    import re
    from muon_amd import _lib, serving
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "include", "muonx.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define (MUONX_\w+) \(?(-?\d+)\)?", text)}
    mirrored = [(n, getattr(_lib, n)) for n in dir(_lib) if n.startswith("MUONX_")]
    mirrored += [("MUONX_" + n, getattr(serving, n)) for n in dir(serving) if n.startswith("SLICES_")]
    assert len(mirrored) >= 5
    for name, value in mirrored:
        assert defs[name] == value, name
    assert defs["MUONX_TRAVERSAL_FAST"] == int(False) and defs["MUONX_TRAVERSAL_EXACT"] == int(True)
