"""ctypes binding of the oracle (oracle/gdp_oracle.c).

Test infrastructure: the parity tests use it as the checker of the library.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "oracle", "lib", "libgdp_oracle.so")
SEG_LEN = 256  # canonical segment length shared with the kernels (kSegLen)


class Result(ctypes.Structure):
    _fields_ = [("rounds", c_int), ("infeasible", c_int), ("capped", c_int), ("loop_seconds", c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB}")
        L = ctypes.CDLL(LIB)
        vp = c_void_p
        L.oracle_seq_propagate.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int, c_int,
                                           POINTER(Result)]
        L.oracle_jacobi_propagate.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int, c_int,
                                              c_int, POINTER(Result)]
        L.oracle_seq_disjoint_propagate.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int,
                                                    POINTER(Result)]
        for f in ("oracle_adjust_upper", "oracle_adjust_lower"):
            getattr(L, f).argtypes = [c_int, c_double]
            getattr(L, f).restype = c_double
        for f in ("oracle_is_lb_better", "oracle_is_ub_better"):
            getattr(L, f).argtypes = [c_double, c_double, c_double]
            getattr(L, f).restype = c_int
        for f in ("oracle_tighten_var_upper", "oracle_tighten_var_lower"):
            getattr(L, f).argtypes = [c_double, c_double, c_double, c_int, c_double, c_double, c_int,
                                      POINTER(c_double)]
            getattr(L, f).restype = c_int
        L.oracle_can_cons_be_tightened.argtypes = [c_double, c_double, c_int, c_int, c_double, c_double, c_double]
        L.oracle_is_cons_infeasible.argtypes = [c_double, c_double, c_double, c_double]
        L.oracle_canonical_sum.argtypes = [vp, c_int, c_int]
        L.oracle_canonical_sum.restype = c_double
        L.oracle_row_activities_seq.argtypes = [c_int, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_row_activities_canon.argtypes = [c_int, vp, vp, vp, vp, vp, c_int, vp, vp]
        L.oracle_weakest_bounds.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_progress_max.argtypes = [c_int, vp, vp, vp, vp, vp, vp, POINTER(c_double), POINTER(c_int)]
        L.oracle_progress_round.argtypes = [c_int, vp, vp, vp, vp, vp, vp, vp, vp, POINTER(c_double),
                                            POINTER(c_int)]
        L.oracle_progress_normalize.argtypes = [c_double, c_double]
        L.oracle_progress_normalize.restype = c_double
        _lib = L
    return _lib


def _p(a):
    return c_void_p(a.ctypes.data)


def _arrays(inst):
    return (np.ascontiguousarray(inst.col, np.int32), np.ascontiguousarray(inst.row_ptr, np.int32),
            np.ascontiguousarray(inst.val, np.float64), np.ascontiguousarray(inst.lhs, np.float64),
            np.ascontiguousarray(inst.rhs, np.float64), np.ascontiguousarray(inst.vartype, np.int32))


def seq(inst, max_rounds=100, detect_infeasible=True, lb=None, ub=None):
    """cpu_seq fixpoint: returns (lb, ub, rounds, infeasible, capped)."""
    return seq_timed(inst, max_rounds, detect_infeasible, lb, ub)[0]


def seq_timed(inst, max_rounds=100, detect_infeasible=True, lb=None, ub=None):
    """seq() and the time of its round loop alone in seconds (no CSC build,
    no allocation: the reference's timing scope)."""
    col, rp, val, lhs, rhs, vt = _arrays(inst)
    lb = np.array(inst.lb if lb is None else lb, np.float64)
    ub = np.array(inst.ub if ub is None else ub, np.float64)
    r = Result()
    lib().oracle_seq_propagate(inst.m, inst.n, len(col), _p(col), _p(rp), _p(val), _p(lhs), _p(rhs), _p(lb),
                               _p(ub), _p(vt), max_rounds, int(detect_infeasible), ctypes.byref(r))
    return (lb, ub, r.rounds, bool(r.infeasible), bool(r.capped)), r.loop_seconds


def jacobi(inst, max_rounds=100, detect_infeasible=True, seg_len=SEG_LEN, lb=None, ub=None):
    """Round-synchronous fixpoint in the kernels' summation order."""
    col, rp, val, lhs, rhs, vt = _arrays(inst)
    lb = np.array(inst.lb if lb is None else lb, np.float64)
    ub = np.array(inst.ub if ub is None else ub, np.float64)
    r = Result()
    lib().oracle_jacobi_propagate(inst.m, inst.n, len(col), _p(col), _p(rp), _p(val), _p(lhs), _p(rhs), _p(lb),
                                  _p(ub), _p(vt), max_rounds, int(detect_infeasible), seg_len, ctypes.byref(r))
    return lb, ub, r.rounds, bool(r.infeasible), bool(r.capped)


def seq_disjoint(inst, max_rounds=100):
    col, rp, val, lhs, rhs, vt = _arrays(inst)
    lb = np.array(inst.lb, np.float64)
    ub = np.array(inst.ub, np.float64)
    r = Result()
    lib().oracle_seq_disjoint_propagate(inst.m, inst.n, len(col), _p(col), _p(rp), _p(val), _p(lhs), _p(rhs),
                                        _p(lb), _p(ub), _p(vt), max_rounds, ctypes.byref(r))
    return lb, ub, r.rounds, bool(r.infeasible), bool(r.capped)


def canonical_sum(x, seg_len=SEG_LEN):
    x = np.ascontiguousarray(x, np.float64)
    return lib().oracle_canonical_sum(_p(x), len(x), seg_len)


def row_activities(inst, row, canonical=False, lb=None, ub=None, seg_len=SEG_LEN):
    col, rp, val, _, _, _ = _arrays(inst)
    lb = np.ascontiguousarray(inst.lb if lb is None else lb, np.float64)
    ub = np.ascontiguousarray(inst.ub if ub is None else ub, np.float64)
    out = np.zeros(3)
    inf = np.zeros(2, np.int32)
    if canonical:
        lib().oracle_row_activities_canon(row, _p(col), _p(rp), _p(val), _p(lb), _p(ub), seg_len, _p(out), _p(inf))
    else:
        lib().oracle_row_activities_seq(row, _p(col), _p(rp), _p(val), _p(lb), _p(ub), _p(out), _p(inf))
    return out, inf


def tighten_upper(coeff, slack, surplus, num_inf, lb, ub, is_cont):
    nb = c_double()
    t = lib().oracle_tighten_var_upper(coeff, slack, surplus, num_inf, lb, ub, int(is_cont), ctypes.byref(nb))
    return bool(t), nb.value


def tighten_lower(coeff, slack, surplus, num_inf, lb, ub, is_cont):
    nb = c_double()
    t = lib().oracle_tighten_var_lower(coeff, slack, surplus, num_inf, lb, ub, int(is_cont), ctypes.byref(nb))
    return bool(t), nb.value


def weakest_bounds(inst, lb=None, ub=None):
    """Weakest finite bounds of the bounds that start infinite
    (computeWeakestBounds): returns (lb, ub)."""
    col, rp, val, lhs, rhs, vt = _arrays(inst)
    lb = np.array(inst.lb if lb is None else lb, np.float64)
    ub = np.array(inst.ub if ub is None else ub, np.float64)
    rc = lib().oracle_weakest_bounds(inst.m, inst.n, len(col), _p(col), _p(rp), _p(val), _p(lhs), _p(rhs), _p(lb),
                                     _p(ub), _p(vt))
    assert rc == 0
    return lb, ub


def progress_series(inst, driver="seq"):
    """The progress measure of every round of a propagation to the fixpoint,
    as the reference's measure entry points print it: the weakest bounds of
    the input, the cpu_seq fixpoint as the limit, and per round the
    normalised finite and infinite measures.  driver "seq" propagates with
    cpu_seq, "jacobi" round-synchronously.  Returns (rows, lb, ub) with rows
    a list of (round, p_fin, p_inf), or (None, j) if variable j has a finite
    limit upper bound but an infinite weakest one (the reference exits)."""
    L = lib()
    n = inst.n
    lb_in = np.array(inst.lb, np.float64)
    ub_in = np.array(inst.ub, np.float64)
    wl, wu = weakest_bounds(inst)
    ll, lu, _, _, _ = seq(inst)
    fmax, imax = c_double(), c_int()
    L.oracle_progress_max(n, _p(lb_in), _p(ub_in), _p(wl), _p(wu), _p(ll), _p(lu), ctypes.byref(fmax),
                          ctypes.byref(imax))
    run = seq if driver == "seq" else jacobi
    lb, ub, rounds, _, _ = run(inst)
    rows = []
    pl, pu = lb_in, ub_in
    inf_total = 0
    for r in range(1, rounds + 1):
        cl, cu, _, _, _ = run(inst, max_rounds=r)
        f, k = c_double(), c_int()
        rc = L.oracle_progress_round(n, _p(cl), _p(cu), _p(wl), _p(wu), _p(ll), _p(lu), _p(pl), _p(pu),
                                     ctypes.byref(f), ctypes.byref(k))
        if rc != 0:
            return None, -1 - rc
        rows.append((r, L.oracle_progress_normalize(f.value, fmax.value),
                     L.oracle_progress_normalize(float(k.value + inf_total), float(imax.value))))
        inf_total += k.value
        pl, pu = cl, cu
    return rows, lb, ub
