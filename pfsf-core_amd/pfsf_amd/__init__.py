"""Host-side Python face of libpfsf (MI355X / gfx950).

``capi``      ctypes binding of the C ABI in include/pfsf/*.h
``device``    device-array calls (pfsf_hip.h) on torch CUDA tensors
``engine``    the reference tick-path API (pfsf.h) on host buffers
``workloads`` seeded synthetic islands
"""
from . import capi  # noqa: F401
