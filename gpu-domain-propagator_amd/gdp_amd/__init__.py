"""gdp_amd - host-side Python mirror of the MI355X domain propagator.

The compute lives in libGpuProp.so (HIP kernels for gfx950 behind the C ABI of
include/interface.h and include/gdp_ext.h); this package binds it with ctypes
and provides the synthetic workloads and an MPS reader.
"""
from .instances import Instance  # noqa: F401
from . import interface  # noqa: F401
from .mps import parse_mps, read_mps  # noqa: F401
from .interface import (  # noqa: F401
    CPU_LOOP, GPU_LOOP, MEGAKERNEL, GDP_OKAY, GDP_ERROR, GDP_NOTIMPLEMENTED, GDP_BINARY, GDP_INTEGER,
    GDP_CONTINUOUS, MAX_NUM_ROUNDS, GDP_INF, GDP_EPS, LIB_PATH, LEGACY_SYMBOLS, EXT_SYMBOLS, Problem, Outcome,
    load_library, device_count, device_name, DeviceArray, PinnedArray, Batch, Server, gpu_propagate, host_propagate, propagateGPUAtomic, propagateGPUReduction,
    propagateSequential, propagateFullOMP, propagateSequentialDisjoint,
)
