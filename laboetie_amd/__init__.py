"""laboetie_amd: MI355X-native lattice-Boltzmann + moment-propagation hot path
of laboetie (maxlevesque/laboetie), behind the C-ABI declared in include/lbx.h.

Submodules:
  lbin     lb.in / geom.in readers with the reference grammar
  lbx      ctypes binding of liblbx.so (the HIP library)
"""
__all__ = ["lbin", "lbx"]
