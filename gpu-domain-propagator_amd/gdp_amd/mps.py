"""MPS files as propagation instances.

The reference reads MIPLIB instances through python-mip and hands the
propagators n, m, nnz, the CSR matrix by constraint, lhs/rhs by constraint
sense and the variable bounds and types (fileReader/readerInterface.py:52-145).
Here the parsing is native (gdpParseMPS in libGpuProp.so); this module only
decompresses .mps.gz and converts the arrays to an Instance.
"""
from __future__ import annotations

import ctypes
import gzip
import os

import numpy as np

from .instances import Instance
from .interface import GDP_OKAY, MpsModel, last_error, load_library


def parse_mps(text: bytes, name: str = "") -> Instance:
    """Parses MPS text (bytes) into an Instance; raises ValueError on a parse error."""
    lib = load_library()
    model = MpsModel()
    if lib.gdpParseMPS(text, len(text), ctypes.byref(model)) != GDP_OKAY:
        raise ValueError(last_error() or "MPS parse error")
    try:
        m, n, nnz = model.n_cons, model.n_vars, model.nnz

        def arr(ptr, count, dt):
            return np.ctypeslib.as_array(ptr, shape=(max(count, 1),))[:count].astype(dt, copy=True)

        inst = Instance(
            m=m, n=n,
            col=arr(model.col_indices, nnz, np.int32),
            row_ptr=arr(model.row_ptrs, m + 1, np.int32),
            val=arr(model.vals, nnz, np.float64),
            lhs=arr(model.lhss, m, np.float64),
            rhs=arr(model.rhss, m, np.float64),
            lb=arr(model.lbs, n, np.float64),
            ub=arr(model.ubs, n, np.float64),
            vartype=arr(model.vartypes, n, np.int32),
            name=name or model.name.decode(errors="replace"),
        )
    finally:
        lib.gdpMpsFree(ctypes.byref(model))
    return inst


def read_mps(path: str) -> Instance:
    """Reads an .mps or .mps.gz file."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        text = f.read()
    base = os.path.basename(path)
    for ext in (".gz", ".mps"):
        if base.endswith(ext):
            base = base[: -len(ext)]
    return parse_mps(text, base)
