"""nicp_amd: the NICP registration loop (yorsh87/nicp) on MI355X (gfx950).

The product is the C ABI library ``libnicp_amd.so`` (``include/nicp_amd.h``)
with the C++ drop-in classes of ``include/nicp/``. This Python package holds its
ctypes bindings (``capi``) and the seeded synthetic scene generator (``synth``).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
