"""MI355X (gfx950) implementation of HEXA TopOpt's inner solve loop.

``hexa``     -- ctypes bindings of libhexa (include/hexa.h) and mirrors of the
                reference's solver entry points;
``problems`` -- seeded structured-grid workloads;
``build``    -- in-tree hipcc build of libhexa.so.
"""
from . import problems  # noqa: F401

__all__ = ["problems", "hexa", "build"]
