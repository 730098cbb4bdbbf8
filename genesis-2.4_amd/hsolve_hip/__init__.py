"""MI355X backend for GENESIS's hsolve chanmode-4/5 step.

* :mod:`hsolve_hip.model`  -- setup compiler: synthetic populations as
  Hsolve-format arrays (libhsolve_model.so)
* :mod:`hsolve_hip.device` -- the device library libhsolve_hip.so behind the
  C ABI of include/hsolve_hip.h
"""
from . import model  # noqa: F401
