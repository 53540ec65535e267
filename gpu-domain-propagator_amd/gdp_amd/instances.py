"""MIP constraint systems in CSR form and the synthetic workloads of the
benchmarks.

Infinity is encoded as +-1e20 (GDP_INF), as the reference's C interface
expects (src/params.h:17); IEEE +-inf is accepted by the library as well.

Generators (all deterministic in their seed, numpy PCG64):
  small_mip        m=2000, n=1000, row lengths U{2..20}, integer
                   coefficients U{-10..10}\\{0}; 60% integer columns in [0,10],
                   40% continuous in [-100,100], 10% of bounds infinite; sides
                   from the activity at a feasible point widened by U[0,5],
                   30% of rows one-sided.  Parameters let the batched workload
                   draw other sizes of the same family.
  powerlaw_mip     m=1e6, n=5e5, Zipf row lengths in [2, 50000]; 70% integer
                   coefficients U{-20..20}\\{0}, 30% log-uniform in [1e-3,1e3]
                   with random sign; 50% binary, 20% integer [0,100], 30%
                   continuous [-1e4,1e4] with 5% infinite bounds; sides from a
                   feasible point plus slack U[0, 0.1*activity range].
  chain_mip        m=4e5, n=2e6, about 12.4M nonzeros, 2% of the variables
                   pre-fixed: small_mip's family at the cover/knapsack sizes,
                   49 rounds to the fixpoint (the many-round workload).
  cover_knapsack   m=4e5 half covering rows (0/1, length U{3..40}, >= 1), half
                   knapsack rows (weights U{1..1000}, length U{10..200},
                   capacity 0.3*sum of weights); n=2e6 binaries, 2% fixed
                   (a fraction one_frac of them to 1, the rest to 0).  With
                   the default seed the fixings to 1 overfill some knapsack
                   rows: round 1 proves the instance infeasible; with
                   one_frac=0 it is feasible and reaches its fixpoint in 2
                   rounds.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

INF = 1e20
CONT = 3


@dataclass
class Instance:
    m: int
    n: int
    col: np.ndarray      # int32 [nnz]
    row_ptr: np.ndarray  # int32 [m+1]
    val: np.ndarray      # float64 [nnz]
    lhs: np.ndarray      # float64 [m]
    rhs: np.ndarray      # float64 [m]
    lb: np.ndarray       # float64 [n]
    ub: np.ndarray       # float64 [n]
    vartype: np.ndarray  # int32 [n]
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def copy_bounds(self):
        return self.lb.copy(), self.ub.copy()

    def with_bounds(self, lb, ub) -> "Instance":
        return Instance(self.m, self.n, self.col, self.row_ptr, self.val, self.lhs, self.rhs,
                        np.asarray(lb, np.float64), np.asarray(ub, np.float64), self.vartype, self.name,
                        dict(self.meta))

    def validate(self):
        assert self.row_ptr.shape == (self.m + 1,) and self.row_ptr[0] == 0
        assert np.all(np.diff(self.row_ptr) >= 0)
        assert self.col.shape == (self.nnz,) and self.val.shape == (self.nnz,)
        assert self.nnz == 0 or (self.col.min() >= 0 and self.col.max() < self.n)
        assert self.lhs.shape == (self.m,) and self.rhs.shape == (self.m,)
        assert self.lb.shape == (self.n,) and self.ub.shape == (self.n,) and self.vartype.shape == (self.n,)


def from_rows(m, n, rows, vals, lhs, rhs, lb, ub, vartype, name="") -> Instance:
    """Builds an instance from per-row lists of column indices / values."""
    row_ptr = np.zeros(m + 1, np.int64)
    for i, r in enumerate(rows):
        row_ptr[i + 1] = row_ptr[i] + len(r)
    col = np.array([c for r in rows for c in r], np.int32)
    val = np.array([v for r in vals for v in r], np.float64)
    return Instance(m, n, col, row_ptr.astype(np.int32), val, np.asarray(lhs, np.float64),
                    np.asarray(rhs, np.float64), np.asarray(lb, np.float64), np.asarray(ub, np.float64),
                    np.asarray(vartype, np.int32), name)


def _csr_from_pairs(m: int, n: int, rows: np.ndarray, cols: np.ndarray):
    """Sorted, duplicate-free (row, col) pairs -> CSR structure."""
    key = rows.astype(np.int64) * n + cols.astype(np.int64)
    key = np.unique(key)
    r = (key // n).astype(np.int64)
    c = (key % n).astype(np.int32)
    counts = np.bincount(r, minlength=m)
    row_ptr = np.zeros(m + 1, np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    return row_ptr, r, c


def _row_ids(row_ptr: np.ndarray) -> np.ndarray:
    return np.repeat(np.arange(len(row_ptr) - 1, dtype=np.int64), np.diff(row_ptr))


def _feasible_point(rng, lb, ub, vartype, width=10.0):
    """A point inside the bounds (integral for integer columns); infinite
    bounds are replaced by the finite one minus/plus `width`."""
    lo = np.where(lb <= -INF, np.where(ub >= INF, -width, ub - width), lb)
    hi = np.where(ub >= INF, np.where(lb <= -INF, width, lb + width), ub)
    x = lo + (hi - lo) * rng.random(lb.shape[0])
    is_int = vartype != CONT
    x[is_int] = np.clip(np.round(x[is_int]), np.ceil(lo[is_int]), np.floor(hi[is_int]))
    return x


def _finalize(m, n, row_ptr, col, val, lhs, rhs, lb, ub, vartype, name, meta) -> Instance:
    inst = Instance(m, n, col.astype(np.int32), row_ptr.astype(np.int32), val.astype(np.float64),
                    lhs.astype(np.float64), rhs.astype(np.float64), lb.astype(np.float64), ub.astype(np.float64),
                    vartype.astype(np.int32), name, meta)
    inst.validate()
    return inst


def small_mip(seed: int = 0, m: int = 2000, n: int = 1000, len_lo: int = 2, len_hi: int = 20,
              int_frac: float = 0.6, fixed_frac: float = 0.0, name: str = "small_mip") -> Instance:
    rng = np.random.default_rng(seed)
    lengths = rng.integers(len_lo, len_hi + 1, size=m)
    lengths = np.minimum(lengths, n)
    rows = np.repeat(np.arange(m, dtype=np.int64), lengths)
    cols = rng.integers(0, n, size=rows.shape[0])
    row_ptr, r, c = _csr_from_pairs(m, n, rows, cols)
    nnz = int(row_ptr[-1])
    mag = rng.integers(1, 11, size=nnz)
    val = np.where(rng.random(nnz) < 0.5, -mag, mag).astype(np.float64)
    is_int = rng.random(n) < int_frac
    vartype = np.where(is_int, 1, CONT).astype(np.int32)
    lb = np.where(is_int, 0.0, -100.0)
    ub = np.where(is_int, 10.0, 100.0)
    lb[rng.random(n) < 0.1] = -INF
    ub[rng.random(n) < 0.1] = INF
    x0 = _feasible_point(rng, lb, ub, vartype)
    act = np.bincount(r, weights=val * x0[c], minlength=m)
    lhs = act - 5.0 * rng.random(m)
    rhs = act + 5.0 * rng.random(m)
    one_sided = rng.random(m) < 0.3
    drop_lhs = rng.random(m) < 0.5
    lhs[one_sided & drop_lhs] = -INF
    rhs[one_sided & ~drop_lhs] = INF
    meta = {"generator": "small_mip", "seed": seed}
    if fixed_frac > 0:
        # pre-fixed variables, at the feasible point the sides were built
        # from (drawn last: the rest of the instance is that of fixed_frac=0)
        fix = rng.random(n) < fixed_frac
        lb[fix] = x0[fix]
        ub[fix] = x0[fix]
        meta["fixed_frac"] = fixed_frac
    return _finalize(m, n, row_ptr, c, val, lhs, rhs, lb, ub, vartype, name, meta)


def zipf_lengths(rng, m: int, s: float, lo: int, hi: int) -> np.ndarray:
    k = np.arange(lo, hi + 1, dtype=np.float64)
    p = k ** (-s)
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    return (lo + np.searchsorted(cdf, rng.random(m), side="right")).clip(lo, hi)


def zipf_mean(s: float, lo: int, hi: int) -> float:
    k = np.arange(lo, hi + 1, dtype=np.float64)
    p = k ** (-s)
    return float((k * p).sum() / p.sum())


def powerlaw_mip(seed: int = 1, m: int = 1_000_000, n: int = 500_000, s: float = 1.95, len_lo: int = 2,
                 len_hi: int = 50_000, name: str = "powerlaw_mip") -> Instance:
    rng = np.random.default_rng(seed)
    lengths = zipf_lengths(rng, m, s, len_lo, min(len_hi, n))
    rows = np.repeat(np.arange(m, dtype=np.int64), lengths)
    cols = rng.integers(0, n, size=rows.shape[0])
    row_ptr, r, c = _csr_from_pairs(m, n, rows, cols)
    nnz = int(row_ptr[-1])
    is_real = rng.random(nnz) < 0.3
    ints = rng.integers(1, 21, size=nnz).astype(np.float64)
    reals = 10.0 ** rng.uniform(-3.0, 3.0, size=nnz)
    mag = np.where(is_real, reals, ints)
    val = np.where(rng.random(nnz) < 0.5, -mag, mag)
    u = rng.random(n)
    vartype = np.where(u < 0.5, 0, np.where(u < 0.7, 1, CONT)).astype(np.int32)
    lb = np.zeros(n)
    ub = np.where(vartype == 0, 1.0, 100.0)
    cont = vartype == CONT
    lb[cont] = -1e4
    ub[cont] = 1e4
    lb[cont & (rng.random(n) < 0.05)] = -INF
    ub[cont & (rng.random(n) < 0.05)] = INF
    x0 = _feasible_point(rng, lb, ub, vartype, width=1e4)
    act = np.bincount(r, weights=val * x0[c], minlength=m)
    lbf = np.where(lb <= -INF, x0 - 1e4, lb)
    ubf = np.where(ub >= INF, x0 + 1e4, ub)
    rng_w = np.bincount(r, weights=np.abs(val) * (ubf - lbf)[c], minlength=m)
    lhs = act - 0.1 * rng_w * rng.random(m)
    rhs = act + 0.1 * rng_w * rng.random(m)
    return _finalize(m, n, row_ptr, c, val, lhs, rhs, lb, ub, vartype, name,
                     {"generator": "powerlaw_mip", "seed": seed, "zipf_s": s, "len_hi": len_hi})


def cover_knapsack(seed: int = 2, m: int = 400_000, n: int = 2_000_000, fixed_frac: float = 0.02,
                   one_frac: float = 0.5, name: str = "cover_knapsack") -> Instance:
    rng = np.random.default_rng(seed)
    mc = m // 2
    mk = m - mc
    lengths = np.concatenate([rng.integers(3, 41, size=mc), rng.integers(10, 201, size=mk)])
    rows = np.repeat(np.arange(m, dtype=np.int64), lengths)
    cols = rng.integers(0, n, size=rows.shape[0])
    row_ptr, r, c = _csr_from_pairs(m, n, rows, cols)
    nnz = int(row_ptr[-1])
    is_cover = r < mc
    val = np.where(is_cover, 1.0, rng.integers(1, 1001, size=nnz).astype(np.float64))
    wsum = np.bincount(r, weights=val, minlength=m)
    lhs = np.where(np.arange(m) < mc, 1.0, -INF)
    rhs = np.where(np.arange(m) < mc, INF, 0.3 * wsum)
    vartype = np.zeros(n, np.int32)
    lb = np.zeros(n)
    ub = np.ones(n)
    fixed = rng.random(n) < fixed_frac
    to_one = rng.random(n) < one_frac
    lb[fixed & to_one] = 1.0
    ub[fixed & ~to_one] = 0.0
    return _finalize(m, n, row_ptr, c, val, lhs, rhs, lb, ub, vartype, name,
                     {"generator": "cover_knapsack", "seed": seed})


def chain_mip(seed: int = 1, m: int = 400_000, n: int = 2_000_000, len_hi: int = 60,
              fixed_frac: float = 0.02, name: str = "chain_mip") -> Instance:
    """The many-round workload: the structured binary instance's sizes
    (m = 400,000, n = 2,000,000, about 12.4M nonzeros) and its 2% of
    pre-fixed variables, built from the general family (small_mip: row
    lengths U{2..60}, 60% integer columns, continuous bounds converging over
    many rounds).  Seed 1 takes 49 round-synchronous rounds (cpu_seq 33).
    cover_knapsack itself cannot chain: a zero propagates only through a
    covering row whose other members are all zero, a one only through a
    knapsack row filled to 30% of its weight, and every fixing fraction that
    makes either likely makes some row infeasible in round 1."""
    return small_mip(seed=seed, m=m, n=n, len_hi=len_hi, fixed_frac=fixed_frac, name=name)


def batch_family(seed: int = 3, count: int = 256, m_lo: int = 1000, m_hi: int = 200_000,
                 mean_lo: float = 3.0, mean_hi: float = 30.0):
    """Sizes for the batched workload: m log-uniform, n = m/2, mean row length
    log-uniform; each instance is a small_mip with its own seed."""
    rng = np.random.default_rng(seed)
    ms = np.exp(rng.uniform(np.log(m_lo), np.log(m_hi), size=count)).astype(np.int64)
    means = np.exp(rng.uniform(np.log(mean_lo), np.log(mean_hi), size=count))
    specs = []
    for i in range(count):
        mean = float(means[i])
        hi = max(2, int(round(2 * mean - 2)))
        specs.append(dict(seed=int(seed * 100003 + i), m=int(ms[i]), n=max(2, int(ms[i]) // 2), len_lo=2,
                          len_hi=max(2, hi)))
    return specs


def tighten_node(inst: Instance, rng, k_lo: int = 1, k_hi: int = 50, base_lb=None, base_ub=None):
    """Bounds of a branch-and-bound node: k random integer columns of the base
    bounds (default: the root bounds; typically a propagated parent's)
    restricted to a random half of their domain.  Returns (lb, ub, columns)
    with the columns whose bounds changed."""
    lb = np.array(inst.lb if base_lb is None else base_lb, np.float64, copy=True)
    ub = np.array(inst.ub if base_ub is None else base_ub, np.float64, copy=True)
    ints = np.flatnonzero((inst.vartype != CONT) & (ub - lb >= 1) & (ub < INF) & (lb > -INF))
    if ints.size == 0:
        return lb, ub, np.zeros(0, np.int32)
    k = int(rng.integers(k_lo, k_hi + 1))
    pick = rng.choice(ints, size=min(k, ints.size), replace=False)
    for j in pick:
        mid = np.floor((lb[j] + ub[j]) / 2)
        if rng.random() < 0.5:
            ub[j] = mid
        else:
            lb[j] = mid + 1 if mid + 1 <= ub[j] else mid
    return lb, ub, np.sort(pick).astype(np.int32)
