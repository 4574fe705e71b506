"""The reference extension's module name (setup.py:40, horizon_svf.pyx:15):
``from horizon_svf import horizon_svf_comp_py`` works unchanged against
this library (same arguments, checks and outputs; see
horayzon_extpar_amd.horizon_svf)."""

from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py

__all__ = ["horizon_svf_comp_py"]
