"""tspeed_amd -- MI355X (gfx950) engine for tspeed's leap-frog time-stepping path.

The engine is a C-ABI shared library (libtspeed_amd.so, include/tspeed_amd.h)
built in-tree by ``tspeed_amd.build``; the host C++ layer that mirrors the
reference's classes lives in tspeed_amd/csrc/host.  This package only loads and
drives the library; it has no compute path of its own.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
