"""Test instances: the reference's analytic fixtures (tests/golden), the
generalized facility-location closed form and small random systems."""
import json
import os

import numpy as np

from gdp_amd.instances import INF, Instance, from_rows, small_mip

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def analytic(name) -> Instance:
    d = load_json("analytic.json")[name]
    inst = Instance(d["m"], d["n"], np.array(d["col"], np.int32), np.array(d["row_ptr"], np.int32),
                    np.array(d["vals"], np.float64), np.array(d["lhs"], np.float64), np.array(d["rhs"], np.float64),
                    np.array(d["lb"], np.float64), np.array(d["ub"], np.float64), np.array(d["vartype"], np.int32),
                    name)
    inst.meta["lb_expected"] = np.array(d["lb_expected"], np.float64)
    inst.meta["ub_expected"] = np.array(d["ub_expected"], np.float64)
    return inst


ANALYTIC = ("achterberg", "savelsbergh1", "savelsbergh_cflp", "two_lines")


def generalized_cfp(k: int, capacity: int = 100) -> Instance:
    """Generalized capacitated facility location (the reference's
    GeneralizedCFP, test/test_setups.cuh:469-611, restated): k*k continuous
    flows x_ij >= 0 and k binary openings y_j; demand rows sum_j x_ij =
    (i+1) % capacity and capacity rows sum_i x_ij - capacity*y_j <= 0.
    Closed-form fixpoint: ub(x_ij) = (i+1) % capacity, y_j in [0,1]."""
    rows, vals, lhs, rhs = [], [], [], []
    for i in range(k):
        rows.append([i * k + t for t in range(k)])
        vals.append([1.0] * k)
        lhs.append(float((i + 1) % capacity))
        rhs.append(float((i + 1) % capacity))
    for i in range(k):
        rows.append([t * k + i for t in range(k)] + [k * k + i])
        vals.append([1.0] * k + [-float(capacity)])
        lhs.append(-INF)
        rhs.append(0.0)
    n = k * k + k
    lb = np.zeros(n)
    ub = np.array([INF] * (k * k) + [1.0] * k)
    vt = np.array([3] * (k * k) + [0] * k, np.int32)
    inst = from_rows(2 * k, n, rows, vals, lhs, rhs, lb, ub, vt, f"gcfp{k}")
    inst.meta["ub_expected"] = np.array([float((i + 1) % capacity) for i in range(k) for _ in range(k)] + [1.0] * k)
    inst.meta["lb_expected"] = np.zeros(n)
    return inst


def tiny_integer(seed: int, m: int = 4, n: int = 5, bound: int = 3) -> Instance:
    """Small all-integer system for brute-force soundness checks."""
    rng = np.random.default_rng(seed)
    rows, vals = [], []
    for _ in range(m):
        L = int(rng.integers(1, n + 1))
        cols = sorted(rng.choice(n, size=L, replace=False).tolist())
        rows.append(cols)
        vals.append([float(v) for v in rng.choice([-3, -2, -1, 1, 2, 3], size=L)])
    lb = rng.integers(-bound, 1, size=n).astype(np.float64)
    ub = lb + rng.integers(0, bound + 1, size=n)
    lhs, rhs = [], []
    for cols, v in zip(rows, vals):
        lo = sum(a * (lb[c] if a > 0 else ub[c]) for a, c in zip(v, cols))
        hi = sum(a * (ub[c] if a > 0 else lb[c]) for a, c in zip(v, cols))
        a, b = sorted(rng.uniform(lo - 1, hi + 1, size=2))
        lhs.append(float(np.floor(a)) if rng.random() < 0.7 else -INF)
        rhs.append(float(np.ceil(b)) if rng.random() < 0.7 else INF)
    return from_rows(m, n, rows, vals, lhs, rhs, lb, ub, np.ones(n, np.int32), f"tiny{seed}")


def varied(seed: int, m: int = 600, n: int = 400, long_rows: int = 3, long_len: int = 1500,
           real_frac: float = 0.3, inf_frac: float = 0.05) -> Instance:
    """Rows of every tree-width bin (lengths 1..512) plus a few rows longer
    than one segment, integer and real coefficients, some infinite bounds."""
    rng = np.random.default_rng(seed)
    lens = []
    for b in range(10):
        lo, hi = (1 << (b - 1)) + 1 if b > 0 else 1, 1 << b
        lens += rng.integers(lo, hi + 1, size=max(2, m // 10 // (1 + b // 3))).tolist()
    lens += [long_len + int(rng.integers(0, 700)) for _ in range(long_rows)]
    lens = [min(L, n) for L in lens]
    m = len(lens)
    rows, vals = [], []
    for L in lens:
        cols = np.sort(rng.choice(n, size=L, replace=False))
        ints = rng.integers(1, 11, size=L).astype(np.float64)
        reals = 10 ** rng.uniform(-2, 2, size=L)
        mag = np.where(rng.random(L) < real_frac, reals, ints)
        v = np.where(rng.random(L) < 0.5, -mag, mag)
        rows.append(cols.tolist())
        vals.append(v.tolist())
    is_int = rng.random(n) < 0.6
    vt = np.where(is_int, 1, 3).astype(np.int32)
    lb = np.where(is_int, 0.0, -50.0)
    ub = np.where(is_int, 20.0, 50.0)
    lb[rng.random(n) < inf_frac] = -INF
    ub[rng.random(n) < inf_frac] = INF
    x0 = np.where(lb <= -INF, np.minimum(ub, 50) - 5, np.where(ub >= INF, lb + 5, lb + (ub - lb) * rng.random(n)))
    x0[is_int] = np.round(x0[is_int])
    lhs, rhs = [], []
    for cols, v in zip(rows, vals):
        act = float(np.dot(v, x0[cols]))
        spread = 0.05 * float(np.sum(np.abs(v))) * 10
        lo = act - spread * rng.random()
        hi = act + spread * rng.random()
        r = rng.random()
        lhs.append(lo if r < 0.8 else -INF)
        rhs.append(hi if r > 0.2 else INF)
    return from_rows(m, n, rows, vals, lhs, rhs, lb, ub, vt, f"varied{seed}")


def infeasible_instance() -> Instance:
    """x + y <= -1 with x, y in [0, 5]: bounds cross in the first round."""
    return from_rows(2, 3, [[0, 1], [1, 2]], [[1.0, 1.0], [1.0, -1.0]], [-INF, -INF], [-1.0, 2.0],
                     [0, 0, 0], [5, 5, 5], [1, 1, 3], "infeasible")


def layout_edges(seed: int = 0, n: int = 90000) -> Instance:
    """Shapes at the edges of the matrix layout: empty rows first, in the
    middle and last; rows of 1, 3, 4, 5, 255, 256, 257, 512, 513 and 1024
    entries (padding and segment boundaries); one row of 70000 entries (274
    segments); most columns referenced by no row; binary, small-integer,
    wide-integer (beyond 16 bits) and continuous columns, some infinite
    bounds."""
    rng = np.random.default_rng(seed)
    lens = [0, 1, 3, 4, 5, 0, 255, 256, 257, 512, 513, 1024, 70000, 0, 2, 7, 0]
    lens += rng.integers(1, 40, size=300).tolist() + [0]
    used = rng.choice(n, size=75000, replace=False)  # columns any row may use
    rows, vals = [], []
    for L in lens:
        cols = np.sort(rng.choice(used, size=L, replace=False)) if L else np.zeros(0, np.int64)
        v = rng.integers(1, 9, size=L).astype(np.float64) * np.where(rng.random(L) < 0.5, -1.0, 1.0)
        rows.append(cols.tolist())
        vals.append(v.tolist())
    kind = rng.integers(0, 4, size=n)  # 0 binary, 1 small integer, 2 wide integer, 3 continuous
    lb = np.select([kind == 0, kind == 1, kind == 2], [0.0, -5.0, -100000.0], -50.0)
    ub = np.select([kind == 0, kind == 1, kind == 2], [1.0, 20.0, 200000.0], 50.0)
    cont = kind == 3
    lb[cont & (rng.random(n) < 0.1)] = -INF
    ub[cont & (rng.random(n) < 0.1)] = INF
    vartype = np.where(cont, 3, np.where(kind == 0, 0, 1)).astype(np.int32)
    x0 = np.where(kind == 3, rng.uniform(-10, 10, size=n), np.minimum(np.maximum(0.0, lb), ub))
    lhs, rhs = [], []
    for cols, v in zip(rows, vals):
        act = float(np.dot(v, x0[cols])) if cols else 0.0
        lhs.append(act - float(rng.uniform(0, 20)) if rng.random() < 0.7 else -INF)
        rhs.append(act + float(rng.uniform(0, 20)) if rng.random() < 0.7 else INF)
    return from_rows(len(lens), n, rows, vals, lhs, rhs, lb, ub, vartype, f"layout_edges{seed}")


def int32_overflow() -> Instance:
    """Integer columns with compact bound records (16-bit range: x, y; 32-bit
    range: v, t) receive candidates beyond the int32 range: x - w >= 3e9,
    y + w <= -3e9, v - w >= 5e9, t + w <= -5e9 with w in [0, 1].  The bounds
    cross in round 1 and the returned bounds are the exact candidates."""
    return from_rows(4, 5, [[0, 4], [1, 4], [2, 4], [3, 4]], [[1.0, -1.0], [1.0, 1.0], [1.0, -1.0], [1.0, 1.0]],
                     [3e9, -INF, 5e9, -INF], [INF, -3e9, INF, -5e9],
                     [0, 0, 0, -100000, 0], [5, 5, 100000, 0, 1], [1, 1, 1, 1, 3], "int32_overflow")


# Row patterns of gate_ties: offsets (in units of 1e-9) of the tie entries
# above R, in row order, and whether the sequential epsilon-maximum passes the
# row gate.  The running maximum starts at 0 and takes a value only when it
# exceeds the running one by more than 1e-9; the gate passes when that
# maximum exceeds R by more than 1e-9.
GATE_TIE_PATTERNS = (
    ((0.8, 1.5), False),        # 1.5 is within eps of 0.8: maximum stays at R+0.8e-9
    ((1.5, 0.8), True),
    ((0.3, 1.2), False),
    ((-2.0, -0.5, 0.9, 1.6), False),  # chain of replacements ends at R+0.9e-9
    ((-2.0, -0.5, 1.6, 0.9), True),
    ((1.2,), True),
    ((0.9,), False),
)


def gate_ties(lengths=(2, 7, 64, 200, 256, 300, 1000, 5000), R: float = 10.0) -> Instance:
    """Rows x_1 + ... + x_L <= R over continuous x in [0, u] whose row gate
    depends on the order of the entries: filler columns have u = 1 and the
    tie columns u = R + offset * 1e-9 (GATE_TIE_PATTERNS), spread over the
    row (for rows longer than 256 entries: over different 256-entry
    segments).  A row whose gate passes tightens its tie columns with
    u > R + 1e-9 to u = R; the expected bounds are in meta."""
    rows, vals, lhs, rhs = [], [], [], []
    ub = []
    ub_exp = []
    for L in lengths:
        for offsets, passes in GATE_TIE_PATTERNS:
            if len(offsets) > L:
                continue
            start = len(ub)
            pos = [int(round((t + 0.5) * L / len(offsets))) for t in range(len(offsets))]
            pos = [min(p, L - 1) for p in pos]
            u_row = [1.0] * L
            for p, off in zip(pos, offsets):
                u_row[p] = R + off * 1e-9
            rows.append(list(range(start, start + L)))
            vals.append([1.0] * L)
            lhs.append(-INF)
            rhs.append(R)
            ub.extend(u_row)
            for u in u_row:
                ub_exp.append(R if passes and u - R > 1e-9 else u)
    n = len(ub)
    inst = from_rows(len(rows), n, rows, vals, lhs, rhs, np.zeros(n), np.array(ub), np.full(n, 3, np.int32),
                     "gate_ties")
    inst.meta["ub_expected"] = np.array(ub_exp)
    inst.meta["lb_expected"] = np.zeros(n)
    return inst
