"""MI355X-native pressure Poisson solver behind NGA2's ``linsol`` interface.

Modules:
  config   serial mirror of NGA2's grid/config (metrics, sync, integrate)
  incomp   caller-side operator assembly, divergence, shift_p, get_pgrad
  inputs   seeded synthetic workloads
  linsol   NgapStr: the hypre_str-compatible solver over libngapsolve.so
  _lib     ctypes binding of the C ABI (include/ngapsolve.h)
"""
__version__ = "0.1.0"
