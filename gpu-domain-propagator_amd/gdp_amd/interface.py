"""ctypes binding of libGpuProp.so.

The functions ``propagateGPUAtomic``, ``propagateGPUReduction``,
``propagateSequential``, ``propagateFullOMP`` and ``propagateSequentialDisjoint``
mirror the reference's Python interface (fileReader/GPUDomPropInterface.py:35-145,
:195-247, :330-383, :385-429): same names, same argument order (note: n_vars
before n_cons), same return value ``(lbs, ubs)`` as lists.  The ``datatype``
argument selects the Double or Float entry point.

The numpy-level helpers (``gpu_propagate``, ``Problem`` ...) use the extension
API of include/gdp_ext.h and also return the round count and infeasibility.

Loading never falls back to anything: if the library is missing, importing
functions that need it raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_longlong, c_void_p
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GDP_LIBRARY", os.path.join(_PKG_DIR, "lib", "libGpuProp.so"))

GDP_OKAY = 1
GDP_ERROR = 0
GDP_NOTIMPLEMENTED = -18
GDP_BINARY, GDP_INTEGER, GDP_CONTINUOUS = 0, 1, 3
CPU_LOOP, GPU_LOOP, MEGAKERNEL = 0, 1, 2
MAX_NUM_ROUNDS = 100
GDP_INF = 1e20
GDP_EPS = 1e-9

#: every symbol include/interface.h and include/gdp_ext.h declare
LEGACY_SYMBOLS = (
    "propagateConstraintsGPUReductionDouble",
    "propagateConstraintsGPUReductionFloat",
    "propagateConstraintsGPUAtomicDouble",
    "propagateConstraintsGPUAtomicFloat",
    "propagateConstraintsSequentialDouble",
    "propagateConstraintsSequentialFloat",
    "propagateConstraintsFullOMPDouble",
    "propagateConstraintsFullOMPFloat",
    "propagateConstraintsSequentialDisjointDouble",
    "sequentialPropagateWithMeasureDouble",
    "atomicPropagateWithMeasureDouble",
)
EXT_SYMBOLS = (
    "gdpTuningDefaults",
    "gdpProblemCreate",
    "gdpProblemDestroy",
    "gdpProblemGetStats",
    "gdpPropagate",
    "gdpPropagateDevice",
    "gdpPropagateDeviceNoAlloc",
    "gdpProblemReserve",
    "gdpProblemWorkProfile",
    "gdpPropagateNodesDevice",
    "gdpAutotune",
    "gdpBatchAutotune",
    "gdpBatchCreate",
    "gdpBatchPropagate",
    "gdpBatchPropagateDevice",
    "gdpServerCreate",
    "gdpServerDestroy",
    "gdpServerSlots",
    "gdpServerSubmit",
    "gdpServerSubmitPinned",
    "gdpServerSetBase",
    "gdpServerSubmitDelta",
    "gdpServerStep",
    "gdpServerPending",
    "gdpServerCompleted",
    "gdpServerResultChanges",
    "gdpServerResult",
    "gdpServerStart",
    "gdpServerStop",
    "gdpSequentialPropagate",
    "gdpOMPPropagate",
    "gdpPropagateWithMeasure",
    "gdpWeakestBounds",
    "gdpParseMPS",
    "gdpMpsFree",
    "gdpDeviceCount",
    "gdpDeviceMalloc",
    "gdpDeviceFree",
    "gdpHostAlloc",
    "gdpHostFree",
    "gdpMemcpyHtoD",
    "gdpMemcpyDtoH",
    "gdpMemcpyDtoD",
    "gdpDeviceName",
    "gdpLastError",
    "gdpVersion",
)


class RunInfo(ctypes.Structure):
    _fields_ = [("rounds", c_int), ("infeasible", c_int), ("capped", c_int), ("kernel_ms", c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("item_nnz", c_int),
        ("mega_blocks_per_cu", c_int),
        ("mega_threads", c_int),
        ("graph_rounds", c_int),
        ("max_split_segs", c_int),
        ("compact_bounds", c_int),
        ("row_marking", c_int),
        ("mark_limit", c_int),
        ("group_log", c_int * 10),
    ]


class MpsModel(ctypes.Structure):
    _fields_ = [("n_cons", c_int), ("n_vars", c_int), ("nnz", c_int),
                ("row_ptrs", POINTER(c_int)), ("col_indices", POINTER(c_int)),
                ("vals", POINTER(c_double)), ("lhss", POINTER(c_double)), ("rhss", POINTER(c_double)),
                ("lbs", POINTER(c_double)), ("ubs", POINTER(c_double)), ("vartypes", POINTER(c_int)),
                ("name", ctypes.c_char * 128)]


class ProblemStats(ctypes.Structure):
    _fields_ = [
        ("n_cons", c_int),
        ("n_vars", c_int),
        ("nnz", c_int),
        ("n_cons_nonempty", c_int),
        ("n_items", c_int),
        ("n_long_rows", c_int),
        ("bin_rows", c_longlong * 11),
        ("bin_nnz", c_longlong * 11),
        ("n16", c_int),
        ("n32", c_int),
        ("nf", c_int),
        ("n_units", c_int),
        ("unit_state_bytes", c_longlong),
        ("mega_threads", c_int),
        ("mega_grid", c_int),
        ("run_state_bytes", c_longlong),
        ("n_binary", c_int),
        ("n_lds_binary", c_int),
    ]


_lib = None


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Loads libGpuProp.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libGpuProp.so not found at {p}; run __graft_entry__.build() first")
    lib = ctypes.CDLL(p)
    vp = c_void_p
    for name in ("propagateConstraintsGPUAtomicDouble", "propagateConstraintsGPUAtomicFloat"):
        getattr(lib, name).argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int]
        getattr(lib, name).restype = c_int
    for name in LEGACY_SYMBOLS:
        if "GPUAtomic" in name:
            continue
        getattr(lib, name).argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        getattr(lib, name).restype = c_int
    lib.gdpTuningDefaults.argtypes = [POINTER(Tuning)]
    lib.gdpTuningDefaults.restype = None
    lib.gdpProblemCreate.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, POINTER(Tuning),
                                     POINTER(vp)]
    lib.gdpProblemCreate.restype = c_int
    lib.gdpProblemDestroy.argtypes = [vp]
    lib.gdpProblemDestroy.restype = c_int
    lib.gdpProblemGetStats.argtypes = [vp, POINTER(ProblemStats)]
    lib.gdpProblemGetStats.restype = c_int
    lib.gdpPropagate.argtypes = [vp, vp, vp, c_int, c_int, POINTER(RunInfo)]
    lib.gdpPropagate.restype = c_int
    lib.gdpPropagateDevice.argtypes = [vp, vp, vp, c_int, c_int, POINTER(RunInfo)]
    lib.gdpPropagateDevice.restype = c_int
    lib.gdpPropagateDeviceNoAlloc.argtypes = [vp, vp, vp, c_int, c_int, POINTER(RunInfo)]
    lib.gdpPropagateDeviceNoAlloc.restype = c_int
    lib.gdpProblemWorkProfile.argtypes = [vp, vp, vp, c_int, vp, vp, vp, vp, POINTER(c_int), POINTER(RunInfo)]
    lib.gdpProblemWorkProfile.restype = c_int
    lib.gdpProblemReserve.argtypes = [vp, c_int]
    lib.gdpProblemReserve.restype = c_int
    lib.gdpPropagateNodesDevice.argtypes = [vp, c_int, vp, vp, vp, vp, c_int, c_int, POINTER(RunInfo)]
    lib.gdpPropagateNodesDevice.restype = c_int
    lib.gdpAutotune.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int, POINTER(Tuning),
                                POINTER(c_double)]
    lib.gdpAutotune.restype = c_int
    lib.gdpBatchAutotune.argtypes = [c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_int, POINTER(Tuning),
                                     POINTER(c_double)]
    lib.gdpBatchAutotune.restype = c_int
    lib.gdpBatchCreate.argtypes = [c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, POINTER(Tuning), POINTER(vp)]
    lib.gdpBatchCreate.restype = c_int
    lib.gdpBatchPropagate.argtypes = [vp, vp, vp, c_int, c_int, POINTER(RunInfo)]
    lib.gdpBatchPropagate.restype = c_int
    lib.gdpBatchPropagateDevice.argtypes = [vp, vp, vp, c_int, c_int, POINTER(RunInfo)]
    lib.gdpBatchPropagateDevice.restype = c_int
    lib.gdpServerCreate.argtypes = [vp, c_longlong, c_int, POINTER(vp)]
    lib.gdpServerCreate.restype = c_int
    lib.gdpServerDestroy.argtypes = [vp]
    lib.gdpServerDestroy.restype = c_int
    lib.gdpServerSlots.argtypes = [vp, POINTER(c_int)]
    lib.gdpServerSlots.restype = c_int
    lib.gdpServerSubmit.argtypes = [vp, vp, vp, vp, c_int, c_int, POINTER(c_longlong)]
    lib.gdpServerSubmit.restype = c_int
    lib.gdpServerSubmitPinned.argtypes = [vp, vp, vp, vp, vp, vp, c_int, c_int, POINTER(c_longlong)]
    lib.gdpServerSubmitPinned.restype = c_int
    lib.gdpServerSetBase.argtypes = [vp, vp, vp]
    lib.gdpServerSetBase.restype = c_int
    lib.gdpServerSubmitDelta.argtypes = [vp, c_int, vp, vp, vp, c_int, POINTER(c_longlong)]
    lib.gdpServerSubmitDelta.restype = c_int
    lib.gdpServerStep.argtypes = [vp, POINTER(c_int)]
    lib.gdpServerStep.restype = c_int
    lib.gdpServerPending.argtypes = [vp, POINTER(c_int)]
    lib.gdpServerPending.restype = c_int
    lib.gdpServerResultChanges.argtypes = [vp, c_longlong, vp, vp, vp, c_int, POINTER(c_int), POINTER(RunInfo),
                                           POINTER(c_int)]
    lib.gdpServerResultChanges.restype = c_int
    lib.gdpServerCompleted.argtypes = [vp, POINTER(c_longlong), c_int, POINTER(c_int)]
    lib.gdpServerCompleted.restype = c_int
    lib.gdpServerResult.argtypes = [vp, c_longlong, vp, vp, POINTER(RunInfo), POINTER(c_int)]
    lib.gdpServerResult.restype = c_int
    lib.gdpServerStart.argtypes = [vp]
    lib.gdpServerStart.restype = c_int
    lib.gdpServerStop.argtypes = [vp]
    lib.gdpServerStop.restype = c_int
    lib.gdpSequentialPropagate.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int,
                                           POINTER(RunInfo)]
    lib.gdpSequentialPropagate.restype = c_int
    lib.gdpOMPPropagate.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int, c_int,
                                    POINTER(RunInfo)]
    lib.gdpOMPPropagate.restype = c_int
    lib.gdpPropagateWithMeasure.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int, c_int,
                                            c_int, vp, vp, POINTER(c_int), POINTER(RunInfo)]
    lib.gdpPropagateWithMeasure.restype = c_int
    lib.gdpWeakestBounds.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.gdpWeakestBounds.restype = c_int
    lib.gdpParseMPS.argtypes = [ctypes.c_char_p, ctypes.c_size_t, POINTER(MpsModel)]
    lib.gdpParseMPS.restype = c_int
    lib.gdpMpsFree.argtypes = [POINTER(MpsModel)]
    lib.gdpMpsFree.restype = None
    lib.gdpDeviceCount.argtypes = [POINTER(c_int)]
    lib.gdpDeviceMalloc.argtypes = [POINTER(vp), ctypes.c_size_t]
    lib.gdpDeviceFree.argtypes = [vp]
    lib.gdpHostAlloc.argtypes = [POINTER(vp), ctypes.c_size_t]
    lib.gdpHostAlloc.restype = c_int
    lib.gdpHostFree.argtypes = [vp]
    lib.gdpHostFree.restype = c_int
    for f in ("gdpMemcpyHtoD", "gdpMemcpyDtoH", "gdpMemcpyDtoD"):
        getattr(lib, f).argtypes = [vp, vp, ctypes.c_size_t]
    lib.gdpDeviceName.argtypes = [ctypes.c_char_p, c_int]
    lib.gdpLastError.argtypes = []
    lib.gdpLastError.restype = ctypes.c_char_p
    lib.gdpVersion.argtypes = []
    lib.gdpVersion.restype = ctypes.c_char_p
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    return load_library().gdpLastError().decode()


def device_count() -> int:
    c = c_int(0)
    if load_library().gdpDeviceCount(ctypes.byref(c)) != GDP_OKAY:
        return 0
    return c.value


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    if load_library().gdpDeviceName(buf, 256) != GDP_OKAY:
        raise RuntimeError(last_error())
    return buf.value.decode()


class PinnedArray:
    """A float64 array in pinned host memory (gdpHostAlloc); `.array` is a
    numpy view of it.  For zero-copy server requests."""

    def __init__(self, size: int):
        self.lib = load_library()
        p = c_void_p()
        if self.lib.gdpHostAlloc(ctypes.byref(p), 8 * max(1, int(size))) != GDP_OKAY:
            raise RuntimeError(last_error())
        self.ptr = p.value
        self.array = np.ctypeslib.as_array(ctypes.cast(p, POINTER(c_double)), shape=(int(size),))

    def free(self):
        if self.ptr:
            self.array = None
            self.lib.gdpHostFree(c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A float64 array in device memory, allocated by the library's own HIP
    runtime (see gdpDeviceMalloc)."""

    def __init__(self, host: np.ndarray):
        self.lib = load_library()
        host = np.ascontiguousarray(host, np.float64)
        self.size = host.size
        self.nbytes = host.nbytes
        p = c_void_p()
        if self.lib.gdpDeviceMalloc(ctypes.byref(p), self.nbytes) != GDP_OKAY:
            raise RuntimeError(last_error())
        self.ptr = p.value
        self.upload(host)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, np.float64)
        assert host.nbytes == self.nbytes
        if self.lib.gdpMemcpyHtoD(c_void_p(self.ptr), c_void_p(host.ctypes.data), self.nbytes) != GDP_OKAY:
            raise RuntimeError(last_error())

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        if self.lib.gdpMemcpyDtoD(c_void_p(self.ptr), c_void_p(other.ptr), self.nbytes) != GDP_OKAY:
            raise RuntimeError(last_error())

    def numpy(self) -> np.ndarray:
        out = np.empty(self.size, np.float64)
        if self.lib.gdpMemcpyDtoH(c_void_p(out.ctypes.data), c_void_p(self.ptr), self.nbytes) != GDP_OKAY:
            raise RuntimeError(last_error())
        return out

    def free(self):
        if self.ptr:
            self.lib.gdpDeviceFree(c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# --------------------------------------------------------------------------
# reference-style list interface (fileReader/GPUDomPropInterface.py)
# --------------------------------------------------------------------------

def check_input(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes):
    """Same contract as GPUDomPropInterface.py:12-32."""
    assert len(csr_col_indices) == nnz
    assert len(csr_row_ptrs) == n_cons + 1
    assert len(csr_vals) == nnz
    assert len(lhss) == n_cons
    assert len(rhss) == n_cons
    assert len(lbs) == n_vars
    assert len(ubs) == n_vars
    assert len(vartypes) == n_vars


def _call_list(fname_double, fname_float, n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss,
               rhss, lbs, ubs, vartypes, datatype, extra=()):
    lib = load_library()
    check_input(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes)
    if datatype == c_double:
        fun = getattr(lib, fname_double)
    elif datatype == c_float and fname_float is not None:
        fun = getattr(lib, fname_float)
    else:
        raise Exception("unsupported datatype")
    c_col = (c_int * max(nnz, 1))(*csr_col_indices)
    c_rp = (c_int * (n_cons + 1))(*csr_row_ptrs)
    c_vals = (datatype * max(nnz, 1))(*csr_vals)
    c_lhss = (datatype * max(n_cons, 1))(*lhss)
    c_rhss = (datatype * max(n_cons, 1))(*rhss)
    c_lbs = (datatype * max(n_vars, 1))(*lbs)
    c_ubs = (datatype * max(n_vars, 1))(*ubs)
    c_vt = (c_int * max(n_vars, 1))(*vartypes)
    rc = fun(n_cons, n_vars, nnz, c_col, c_rp, c_vals, c_lhss, c_rhss, c_lbs, c_ubs, c_vt, *extra)
    if rc != GDP_OKAY:
        raise RuntimeError(f"{fun.__name__} returned {rc}: {last_error()}")
    return list(c_lbs)[:n_vars], list(c_ubs)[:n_vars]


def propagateGPUAtomic(n_vars: int, n_cons: int, nnz: int, csr_col_indices: List[int], csr_row_ptrs: List[int],
                       csr_vals: List[float], lhss: List[float], rhss: List[float], lbs: List[float],
                       ubs: List[float], vartypes: List[int], synctype: int = 0,
                       datatype=c_double) -> Tuple[List[float], List[float]]:
    return _call_list("propagateConstraintsGPUAtomicDouble", "propagateConstraintsGPUAtomicFloat", n_vars, n_cons,
                      nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, datatype,
                      extra=(c_int(synctype),))


def propagateGPUReduction(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs,
                          vartypes, datatype=c_double):
    return _call_list("propagateConstraintsGPUReductionDouble", "propagateConstraintsGPUReductionFloat", n_vars,
                      n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, datatype)


def propagateSequential(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs,
                        vartypes, datatype=c_double):
    return _call_list("propagateConstraintsSequentialDouble", "propagateConstraintsSequentialFloat", n_vars, n_cons,
                      nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, datatype)


def propagateFullOMP(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs,
                     vartypes, datatype=c_double):
    return _call_list("propagateConstraintsFullOMPDouble", "propagateConstraintsFullOMPFloat", n_vars, n_cons, nnz,
                      csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, datatype)


def propagateSequentialWithMeasure(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs,
                                   ubs, vartypes):
    """cpu_seq to the fixpoint, printing the progress measure of every round."""
    return _call_list("sequentialPropagateWithMeasureDouble", None, n_vars, n_cons, nnz, csr_col_indices,
                      csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, c_double)


def propagateGPUAtomicWithMeasure(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs,
                                  ubs, vartypes):
    """The GPU's rounds to the fixpoint, printing the progress measure of
    every round."""
    return _call_list("atomicPropagateWithMeasureDouble", None, n_vars, n_cons, nnz, csr_col_indices,
                      csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, c_double)


def propagateSequentialDisjoint(n_vars, n_cons, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs,
                                vartypes, datatype=c_double):
    return _call_list("propagateConstraintsSequentialDisjointDouble", None, n_vars, n_cons, nnz, csr_col_indices,
                      csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs, vartypes, datatype)


# --------------------------------------------------------------------------
# numpy interface over the extension API
# --------------------------------------------------------------------------

@dataclass
class Outcome:
    lbs: np.ndarray
    ubs: np.ndarray
    rounds: int
    infeasible: bool
    capped: bool
    kernel_ms: float


def _arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a: np.ndarray):
    return c_void_p(a.ctypes.data)


def make_tuning(tuning: Optional[dict]) -> Tuning:
    """gdp_tuning with the library defaults, overridden by `tuning`."""
    t = Tuning()
    load_library().gdpTuningDefaults(ctypes.byref(t))
    for k, v in (tuning or {}).items():
        if k == "group_log":
            for b, g in enumerate(v):
                t.group_log[b] = int(g)
        else:
            setattr(t, k, int(v))
    return t


def _outcome(lb, ub, info: RunInfo) -> "Outcome":
    return Outcome(lb, ub, info.rounds, bool(info.infeasible), bool(info.capped), info.kernel_ms)


class Problem:
    """A constraint matrix resident in device memory (gdpProblemCreate)."""

    def __init__(self, inst, tuning: Optional[dict] = None, envelope: bool = True):
        """envelope: use the instance's bounds as the bound envelope (every
        vector propagated later must lie inside them)."""
        self.lib = load_library()
        self.m, self.n = int(inst.m), int(inst.n)
        self.col = _arr(inst.col, np.int32)
        self.row_ptr = _arr(inst.row_ptr, np.int32)
        self.val = _arr(inst.val, np.float64)
        self.lhs = _arr(inst.lhs, np.float64)
        self.rhs = _arr(inst.rhs, np.float64)
        self.vartype = _arr(inst.vartype, np.int32)
        t = make_tuning(tuning)
        h = c_void_p()
        self.env_lb = _arr(inst.lb, np.float64) if envelope else None
        self.env_ub = _arr(inst.ub, np.float64) if envelope else None
        rc = self.lib.gdpProblemCreate(self.m, self.n, len(self.col), _ptr(self.col), _ptr(self.row_ptr),
                                       _ptr(self.val), _ptr(self.lhs), _ptr(self.rhs), _ptr(self.vartype),
                                       _ptr(self.env_lb) if envelope else None,
                                       _ptr(self.env_ub) if envelope else None, ctypes.byref(t), ctypes.byref(h))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpProblemCreate failed: {last_error()}")
        self.handle = h

    def stats(self) -> ProblemStats:
        s = ProblemStats()
        self.lib.gdpProblemGetStats(self.handle, ctypes.byref(s))
        return s

    def propagate(self, lbs, ubs, sync_type: int = CPU_LOOP, max_rounds: int = MAX_NUM_ROUNDS) -> Outcome:
        lb = np.array(lbs, dtype=np.float64, copy=True)
        ub = np.array(ubs, dtype=np.float64, copy=True)
        info = RunInfo()
        rc = self.lib.gdpPropagate(self.handle, _ptr(lb), _ptr(ub), sync_type, max_rounds, ctypes.byref(info))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpPropagate failed: {last_error()}")
        return Outcome(lb, ub, info.rounds, bool(info.infeasible), bool(info.capped), info.kernel_ms)

    def work_profile(self, lbs, ubs, max_rounds: int = MAX_NUM_ROUNDS):
        """Per-round work of one propagation (gdpProblemWorkProfile): returns
        (outcome, rows, nnz, cols, accepted), the last four int64 arrays with
        one entry per round run."""
        lb = np.array(lbs, dtype=np.float64, copy=True)
        ub = np.array(ubs, dtype=np.float64, copy=True)
        rows = np.zeros(max_rounds, np.int64)
        nnz = np.zeros(max_rounds, np.int64)
        cols = np.zeros(max_rounds, np.int64)
        acc = np.zeros(max_rounds, np.int64)
        n_rounds = c_int(0)
        info = RunInfo()
        rc = self.lib.gdpProblemWorkProfile(self.handle, _ptr(lb), _ptr(ub), max_rounds, _ptr(rows), _ptr(nnz),
                                            _ptr(cols), _ptr(acc), ctypes.byref(n_rounds), ctypes.byref(info))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpProblemWorkProfile failed: {last_error()}")
        k = n_rounds.value
        out = Outcome(lb, ub, info.rounds, bool(info.infeasible), bool(info.capped), info.kernel_ms)
        return out, rows[:k], nnz[:k], cols[:k], acc[:k]

    def propagate_device(self, d_lbs_ptr: int, d_ubs_ptr: int, sync_type: int = CPU_LOOP,
                         max_rounds: int = MAX_NUM_ROUNDS) -> RunInfo:
        """Bounds already in device memory (e.g. torch.Tensor.data_ptr())."""
        info = RunInfo()
        rc = self.lib.gdpPropagateDevice(self.handle, c_void_p(d_lbs_ptr), c_void_p(d_ubs_ptr), sync_type,
                                         max_rounds, ctypes.byref(info))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpPropagateDevice failed: {last_error()}")
        return info

    def propagate_device_no_alloc(self, d_lbs_ptr: int, d_ubs_ptr: int, sync_type: int = CPU_LOOP,
                                  max_rounds: int = MAX_NUM_ROUNDS) -> RunInfo:
        """propagate_device with the run state in the library's fixed device
        array: no device memory is allocated (gdpPropagateDeviceNoAlloc)."""
        info = RunInfo()
        rc = self.lib.gdpPropagateDeviceNoAlloc(self.handle, c_void_p(d_lbs_ptr), c_void_p(d_ubs_ptr), sync_type,
                                                max_rounds, ctypes.byref(info))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpPropagateDeviceNoAlloc failed: {last_error()}")
        return info

    def reserve(self, capacity: int):
        """Device state for `capacity` bound vectors (gdpProblemReserve)."""
        if self.lib.gdpProblemReserve(self.handle, capacity) != GDP_OKAY:
            raise RuntimeError(f"gdpProblemReserve failed: {last_error()}")

    def propagate_nodes_device(self, lb_ptrs: Sequence[int], ub_ptrs: Sequence[int],
                               hints: Optional[Sequence[Sequence[int]]] = None, sync_type: int = CPU_LOOP,
                               max_rounds: int = MAX_NUM_ROUNDS) -> List[RunInfo]:
        """Device-resident bound vectors (device pointers) in shared launches."""
        count = len(lb_ptrs)
        lbp = (c_void_p * count)(*lb_ptrs)
        ubp = (c_void_p * count)(*ub_ptrs)
        hp = hs = None
        if hints is not None:
            ptr = np.zeros(count + 1, np.int32)
            ptr[1:] = np.cumsum([len(h) for h in hints])
            flat = np.ascontiguousarray(np.concatenate([np.asarray(h, np.int32) for h in hints]) if ptr[-1] else
                                        np.zeros(1, np.int32), np.int32)
            hp, hs = _ptr(ptr), _ptr(flat)
            self._keep = (ptr, flat)
        infos = (RunInfo * count)()
        rc = self.lib.gdpPropagateNodesDevice(self.handle, count, lbp, ubp, hp, hs, sync_type, max_rounds, infos)
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpPropagateNodesDevice failed: {last_error()}")
        return list(infos)

    def propagate_nodes(self, lbs: Sequence[np.ndarray], ubs: Sequence[np.ndarray],
                        hints: Optional[Sequence[Sequence[int]]] = None, sync_type: int = CPU_LOOP,
                        max_rounds: int = MAX_NUM_ROUNDS) -> List["Outcome"]:
        """Host bound vectors: uploaded, propagated together, copied back."""
        dl = [DeviceArray(np.asarray(x, np.float64)) for x in lbs]
        du = [DeviceArray(np.asarray(x, np.float64)) for x in ubs]
        try:
            infos = self.propagate_nodes_device([d.ptr for d in dl], [d.ptr for d in du], hints, sync_type,
                                                max_rounds)
            return [_outcome(a.numpy(), b.numpy(), i) for a, b, i in zip(dl, du, infos)]
        finally:
            for d in dl + du:
                d.free()

    def close(self):
        if getattr(self, "handle", None):
            if self.lib.gdpProblemDestroy(self.handle) != GDP_OKAY:
                raise RuntimeError(f"gdpProblemDestroy failed: {last_error()}")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent instances propagated in shared launches (gdpBatchCreate).
    Each instance's bounds are its own; the envelope is each instance's
    bounds when `envelope`."""

    def __init__(self, instances, tuning: Optional[dict] = None, envelope: bool = True):
        self.lib = load_library()
        self.instances = list(instances)
        k = len(self.instances)
        self._keep = []

        def arrs(attr, dt):
            a = [_arr(getattr(x, attr), dt) for x in self.instances]
            self._keep.append(a)
            return (c_void_p * k)(*[x.ctypes.data for x in a])

        m = np.array([x.m for x in self.instances], np.int32)
        n = np.array([x.n for x in self.instances], np.int32)
        nnz = np.array([len(x.col) for x in self.instances], np.int32)
        self._keep += [m, n, nnz]
        self.n_offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
        h = c_void_p()
        rc = self.lib.gdpBatchCreate(k, _ptr(m), _ptr(n), _ptr(nnz), arrs("col", np.int32), arrs("row_ptr", np.int32),
                                     arrs("val", np.float64), arrs("lhs", np.float64), arrs("rhs", np.float64),
                                     arrs("vartype", np.int32), arrs("lb", np.float64) if envelope else None,
                                     arrs("ub", np.float64) if envelope else None, ctypes.byref(make_tuning(tuning)),
                                     ctypes.byref(h))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpBatchCreate failed: {last_error()}")
        self.handle = h

    def stats(self) -> ProblemStats:
        s = ProblemStats()
        self.lib.gdpProblemGetStats(self.handle, ctypes.byref(s))
        return s

    def concat_bounds(self, lbs=None, ubs=None):
        lbs = lbs if lbs is not None else [x.lb for x in self.instances]
        ubs = ubs if ubs is not None else [x.ub for x in self.instances]
        return (np.concatenate([np.asarray(x, np.float64) for x in lbs]),
                np.concatenate([np.asarray(x, np.float64) for x in ubs]))

    def propagate(self, lbs=None, ubs=None, sync_type: int = CPU_LOOP,
                  max_rounds: int = MAX_NUM_ROUNDS) -> List["Outcome"]:
        lb, ub = self.concat_bounds(lbs, ubs)
        infos = (RunInfo * len(self.instances))()
        rc = self.lib.gdpBatchPropagate(self.handle, _ptr(lb), _ptr(ub), sync_type, max_rounds, infos)
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpBatchPropagate failed: {last_error()}")
        o = self.n_offsets
        return [_outcome(lb[o[i]:o[i + 1]].copy(), ub[o[i]:o[i + 1]].copy(), infos[i])
                for i in range(len(self.instances))]

    def propagate_device(self, d_lb: int, d_ub: int, sync_type: int = CPU_LOOP,
                         max_rounds: int = MAX_NUM_ROUNDS) -> List[RunInfo]:
        infos = (RunInfo * len(self.instances))()
        rc = self.lib.gdpBatchPropagateDevice(self.handle, c_void_p(d_lb), c_void_p(d_ub), sync_type, max_rounds,
                                              infos)
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpBatchPropagateDevice failed: {last_error()}")
        return list(infos)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.gdpProblemDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Node requests against a device-resident matrix with continuous
    batching (gdpServerCreate)."""

    def __init__(self, problem: Problem, hbm_budget: int, max_slots: int = 1024):
        self.lib = load_library()
        self.problem = problem
        self.n = problem.n
        h = c_void_p()
        if self.lib.gdpServerCreate(problem.handle, int(hbm_budget), int(max_slots), ctypes.byref(h)) != GDP_OKAY:
            raise RuntimeError(f"gdpServerCreate failed: {last_error()}")
        self.handle = h
        c = c_int()
        self.lib.gdpServerSlots(self.handle, ctypes.byref(c))
        self.slots = c.value

    def submit(self, lb, ub, hints=None, max_rounds: int = MAX_NUM_ROUNDS) -> int:
        lb = _arr(lb, np.float64)
        ub = _arr(ub, np.float64)
        hs = _arr(hints, np.int32) if hints is not None and len(hints) else None
        t = c_longlong()
        rc = self.lib.gdpServerSubmit(self.handle, _ptr(lb), _ptr(ub), _ptr(hs) if hs is not None else None,
                                      0 if hs is None else len(hs), max_rounds, ctypes.byref(t))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpServerSubmit failed: {last_error()}")
        return t.value

    def submit_pinned(self, lb: "PinnedArray", ub: "PinnedArray", out_lb: "PinnedArray", out_ub: "PinnedArray",
                      hints=None, max_rounds: int = MAX_NUM_ROUNDS) -> int:
        """Zero-copy request (gdpServerSubmitPinned): the bounds are read from
        the pinned arrays lb / ub when the request is admitted and the result
        is written into out_lb / out_ub; keep all four alive and untouched
        until the ticket completes, then call result(ticket, copy=False)."""
        hs = _arr(hints, np.int32) if hints is not None and len(hints) else None
        t = c_longlong()
        rc = self.lib.gdpServerSubmitPinned(self.handle, c_void_p(lb.ptr), c_void_p(ub.ptr), c_void_p(out_lb.ptr),
                                            c_void_p(out_ub.ptr), _ptr(hs) if hs is not None else None,
                                            0 if hs is None else len(hs), max_rounds, ctypes.byref(t))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpServerSubmitPinned failed: {last_error()}")
        return t.value

    def set_base(self, lb, ub):
        """Base vector of delta requests (a propagated fixpoint)."""
        lb = _arr(lb, np.float64)
        ub = _arr(ub, np.float64)
        if self.lib.gdpServerSetBase(self.handle, _ptr(lb), _ptr(ub)) != GDP_OKAY:
            raise RuntimeError(f"gdpServerSetBase failed: {last_error()}")
        self.base = (lb.copy(), ub.copy())

    def submit_delta(self, cols, lb_vals, ub_vals, max_rounds: int = MAX_NUM_ROUNDS) -> int:
        """The base vector with columns `cols` set to (lb_vals, ub_vals)."""
        c = _arr(cols, np.int32)
        lv = _arr(lb_vals, np.float64)
        uv = _arr(ub_vals, np.float64)
        t = c_longlong()
        rc = self.lib.gdpServerSubmitDelta(self.handle, len(c), _ptr(c), _ptr(lv), _ptr(uv), max_rounds,
                                           ctypes.byref(t))
        if rc != GDP_OKAY:
            raise RuntimeError(f"gdpServerSubmitDelta failed: {last_error()}")
        return t.value

    def step(self) -> int:
        c = c_int()
        if self.lib.gdpServerStep(self.handle, ctypes.byref(c)) != GDP_OKAY:
            raise RuntimeError(f"gdpServerStep failed: {last_error()}")
        return c.value

    def start(self):
        """Run the server's own stepping thread (gdpServerStart): requests
        are stepped without calls to step(); completed() then returns the
        tickets finished since its last call."""
        if self.lib.gdpServerStart(self.handle) != GDP_OKAY:
            raise RuntimeError(f"gdpServerStart failed: {last_error()}")

    def stop(self):
        """End the stepping thread after its current step (gdpServerStop)."""
        if self.lib.gdpServerStop(self.handle) != GDP_OKAY:
            raise RuntimeError(f"gdpServerStop failed: {last_error()}")

    def pending(self) -> int:
        c = c_int()
        self.lib.gdpServerPending(self.handle, ctypes.byref(c))
        return c.value

    def completed(self) -> List[int]:
        """Tickets completed by the last step."""
        cap = max(1, self.slots)
        buf = (c_longlong * cap)()
        cnt = c_int()
        if self.lib.gdpServerCompleted(self.handle, buf, cap, ctypes.byref(cnt)) != GDP_OKAY:
            raise RuntimeError(f"gdpServerCompleted failed: {last_error()}")
        return [buf[i] for i in range(min(cnt.value, cap))]

    def result(self, ticket: int, copy: bool = True) -> Optional["Outcome"]:
        """Completed request (None if not complete).  copy=False: only the
        outcome (a zero-copy request's bounds are in its pinned arrays)."""
        lb = np.empty(self.n if copy else 0, np.float64)
        ub = np.empty(self.n if copy else 0, np.float64)
        info = RunInfo()
        ready = c_int()
        if self.lib.gdpServerResult(self.handle, ticket, _ptr(lb) if copy else None, _ptr(ub) if copy else None,
                                    ctypes.byref(info), ctypes.byref(ready)) != GDP_OKAY:
            raise RuntimeError(f"gdpServerResult failed: {last_error()}")
        return _outcome(lb, ub, info) if ready.value else None

    def result_changes(self, ticket: int):
        """Completed request as (columns, lbs, ubs, RunInfo): the columns
        whose bounds differ from the base vector; None if not complete."""
        cols = np.empty(self.n, np.int32)
        lb = np.empty(self.n, np.float64)
        ub = np.empty(self.n, np.float64)
        cnt = c_int()
        info = RunInfo()
        ready = c_int()
        if self.lib.gdpServerResultChanges(self.handle, ticket, _ptr(cols), _ptr(lb), _ptr(ub), self.n,
                                           ctypes.byref(cnt), ctypes.byref(info), ctypes.byref(ready)) != GDP_OKAY:
            raise RuntimeError(f"gdpServerResultChanges failed: {last_error()}")
        if not ready.value:
            return None
        k = cnt.value
        return cols[:k].copy(), lb[:k].copy(), ub[:k].copy(), info

    def close(self):
        if getattr(self, "handle", None):
            self.lib.gdpServerDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_dict(t: Tuning) -> dict:
    d = {f[0]: getattr(t, f[0]) for f in Tuning._fields_ if f[0] != "group_log"}
    d["group_log"] = list(t.group_log)
    return d


def autotune(inst, sync_type: int = MEGAKERNEL, start: Optional[dict] = None):
    """Fastest tuning for propagating the instance's bounds (gdpAutotune):
    returns (tuning dict, device ms)."""
    lib = load_library()
    a = [_arr(inst.col, np.int32), _arr(inst.row_ptr, np.int32), _arr(inst.val, np.float64),
         _arr(inst.lhs, np.float64), _arr(inst.rhs, np.float64), _arr(inst.vartype, np.int32),
         _arr(inst.lb, np.float64), _arr(inst.ub, np.float64)]
    t = make_tuning(start)
    ms = c_double()
    rc = lib.gdpAutotune(int(inst.m), int(inst.n), len(a[0]), *[_ptr(x) for x in a], sync_type, ctypes.byref(t),
                         ctypes.byref(ms))
    if rc != GDP_OKAY:
        raise RuntimeError(f"gdpAutotune failed: {last_error()}")
    return tuning_dict(t), ms.value


def batch_autotune(instances, sync_type: int = MEGAKERNEL, start: Optional[dict] = None):
    """Fastest tuning for a batch (gdpBatchAutotune): (tuning dict, device ms)."""
    lib = load_library()
    k = len(instances)
    keep = []

    def arrs(attr, dt):
        x = [_arr(getattr(i, attr), dt) for i in instances]
        keep.append(x)
        return (c_void_p * k)(*[y.ctypes.data for y in x])

    m = np.array([i.m for i in instances], np.int32)
    n = np.array([i.n for i in instances], np.int32)
    nnz = np.array([len(i.col) for i in instances], np.int32)
    lb = np.concatenate([np.asarray(i.lb, np.float64) for i in instances])
    ub = np.concatenate([np.asarray(i.ub, np.float64) for i in instances])
    t = make_tuning(start)
    ms = c_double()
    rc = lib.gdpBatchAutotune(k, _ptr(m), _ptr(n), _ptr(nnz), arrs("col", np.int32), arrs("row_ptr", np.int32),
                              arrs("val", np.float64), arrs("lhs", np.float64), arrs("rhs", np.float64),
                              arrs("vartype", np.int32), _ptr(lb), _ptr(ub), sync_type, ctypes.byref(t),
                              ctypes.byref(ms))
    if rc != GDP_OKAY:
        raise RuntimeError(f"gdpBatchAutotune failed: {last_error()}")
    return tuning_dict(t), ms.value


def gpu_propagate(inst, sync_type: int = CPU_LOOP, max_rounds: int = MAX_NUM_ROUNDS,
                  tuning: Optional[dict] = None) -> Outcome:
    p = Problem(inst, tuning)
    try:
        return p.propagate(inst.lb, inst.ub, sync_type, max_rounds)
    finally:
        p.close()


def host_propagate(inst, method: str = "seq", max_rounds: int = MAX_NUM_ROUNDS, threads: int = 0) -> Outcome:
    """The library's own host propagators (cpu_seq or cpu_omp) with run info."""
    lib = load_library()
    col = _arr(inst.col, np.int32)
    rp = _arr(inst.row_ptr, np.int32)
    val = _arr(inst.val, np.float64)
    lhs = _arr(inst.lhs, np.float64)
    rhs = _arr(inst.rhs, np.float64)
    vt = _arr(inst.vartype, np.int32)
    lb = np.array(inst.lb, dtype=np.float64, copy=True)
    ub = np.array(inst.ub, dtype=np.float64, copy=True)
    info = RunInfo()
    args = (int(inst.m), int(inst.n), len(col), _ptr(col), _ptr(rp), _ptr(val), _ptr(lhs), _ptr(rhs), _ptr(lb),
            _ptr(ub), _ptr(vt), max_rounds)
    if method == "seq":
        rc = lib.gdpSequentialPropagate(*args, ctypes.byref(info))
    elif method == "omp":
        rc = lib.gdpOMPPropagate(*args, threads, ctypes.byref(info))
    else:
        raise ValueError(method)
    if rc != GDP_OKAY:
        raise RuntimeError(f"host propagation failed: {last_error()}")
    return Outcome(lb, ub, info.rounds, bool(info.infeasible), bool(info.capped), info.kernel_ms)


MEASURE_CPU_SEQ = 0
MEASURE_GPU = 1


def _inst_arrays(inst):
    return (_arr(inst.col, np.int32), _arr(inst.row_ptr, np.int32), _arr(inst.val, np.float64),
            _arr(inst.lhs, np.float64), _arr(inst.rhs, np.float64), _arr(inst.vartype, np.int32))


def propagate_with_measure(inst, method: int = MEASURE_CPU_SEQ, max_rounds: int = MAX_NUM_ROUNDS,
                           print_measure: bool = False):
    """Propagation to the fixpoint with the progress measure of every round
    (gdpPropagateWithMeasure): returns (Outcome, p_fin, p_inf), the measures
    normalised to 0..100 as the reference prints them."""
    lib = load_library()
    col, rp, val, lhs, rhs, vt = _inst_arrays(inst)
    lb = np.array(inst.lb, dtype=np.float64, copy=True)
    ub = np.array(inst.ub, dtype=np.float64, copy=True)
    pf = np.zeros(max_rounds)
    pi = np.zeros(max_rounds)
    nr = c_int()
    info = RunInfo()
    rc = lib.gdpPropagateWithMeasure(int(inst.m), int(inst.n), len(col), _ptr(col), _ptr(rp), _ptr(val), _ptr(lhs),
                                     _ptr(rhs), _ptr(lb), _ptr(ub), _ptr(vt), method, max_rounds,
                                     int(print_measure), _ptr(pf), _ptr(pi), ctypes.byref(nr), ctypes.byref(info))
    if rc != GDP_OKAY:
        raise RuntimeError(f"gdpPropagateWithMeasure failed: {last_error()}")
    out = Outcome(lb, ub, info.rounds, bool(info.infeasible), bool(info.capped), info.kernel_ms)
    return out, pf[:nr.value].copy(), pi[:nr.value].copy()


def weakest_bounds(inst, lb=None, ub=None):
    """Weakest finite bounds of the bounds that start infinite
    (gdpWeakestBounds): returns (lb, ub)."""
    lib = load_library()
    col, rp, val, lhs, rhs, vt = _inst_arrays(inst)
    lb = np.array(inst.lb if lb is None else lb, dtype=np.float64, copy=True)
    ub = np.array(inst.ub if ub is None else ub, dtype=np.float64, copy=True)
    rc = lib.gdpWeakestBounds(int(inst.m), int(inst.n), len(col), _ptr(col), _ptr(rp), _ptr(val), _ptr(lhs),
                              _ptr(rhs), _ptr(lb), _ptr(ub), _ptr(vt))
    if rc != GDP_OKAY:
        raise RuntimeError(f"gdpWeakestBounds failed: {last_error()}")
    return lb, ub
