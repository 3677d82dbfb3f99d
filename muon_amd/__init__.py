"""muon_amd: MI355X (gfx950) wavefront path tracer behind muon's render surface.

The compute path is the HIP library muon_amd/lib/libmuonx.so (C ABI in
include/muonx.h). This package is a thin host-side mirror of the reference's
render interface (muon/renderer.h, muon/options.h) over that ABI; importing
it fails loudly if the native library has not been built.
"""
from ._lib import MuonxError, lib, library_path  # noqa: F401
from .renderer import Options, Renderer, Film, TraceStats  # noqa: F401
from .scene import HostScene, DeviceScene, Context  # noqa: F401

__all__ = ["MuonxError", "lib", "library_path", "Options", "Renderer", "Film", "TraceStats",
           "HostScene", "DeviceScene", "Context"]
