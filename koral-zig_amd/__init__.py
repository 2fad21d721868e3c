"""koral-zig_amd — MI355X-native per-step hot path of koral-zig.

Layout:
  csrc/        HIP kernels for gfx950 + the C-ABI (libkz_hip.so)
  host/        caller-side C++ (metric tables, PUFFY initializer, BC tables)
               that koral-zig keeps outside the hot path (libkz_host.so)
  abi_types.py ctypes mirrors of include/kz_types.h + PUFFY presets
  hip.py       Python binding of libkz_hip.so (the product entry points)
  host.py      Python binding of libkz_host.so
  build.py     compiles both libraries in-tree
"""
from . import abi_types  # noqa: F401
