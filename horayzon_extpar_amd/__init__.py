"""MI355X (gfx950) horizon and sky-view-factor library.

Drop-in for the HORAYZON_extpar call surface (horizon_svf.pyx):

    from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py

The computation runs in the HIP library ``libhorayzon.so`` (C-ABI declared in
``include/horayzon.h``); Python only validates arguments, allocates outputs
and calls it through ctypes.
"""

from .build import build_library, library_path  # noqa: F401

__all__ = ["build_library", "library_path"]
