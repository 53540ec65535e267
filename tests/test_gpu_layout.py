"""The device-side row order and the host-side column order, checked through
the problem's layout summary (gdpProblemGetStats).

The rows of a problem are ordered on the device (a stable radix sort of the
global rows by tree-width bin, then a scan of the padded lengths), and the
host builds the sweep items from the (bin, instance) runs that it reads back.
The summary below is recomputed here from the instances' row lengths alone:
rows and nonzeros per bin, nonempty and long rows, the number of sweep items
(which depends on every run's length and so on the sort putting each row in
its run), the columns per bound storage class (with the 32-id padding of
each (class, instance) range in a batch), and the binary columns of a single
problem with the part of them a persistent launch reads from LDS.  The
propagation results on the same layouts are checked against the oracle by the
parity tests.
"""
import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import INF, batch_family, powerlaw_mip, small_mip

pytestmark = pytest.mark.gpu

SEG_LEN = 256
SEGS_PER_ITEM = 4
GROUP_LOG = (0, 0, 0, 1, 2, 3, 4, 5, 6, 6)


def _bin(length):
    """Tree-width bin of a nonempty row: 10 for rows longer than one segment,
    else ceil(log2(length))."""
    return np.where(length > SEG_LEN, 10, np.ceil(np.log2(np.maximum(length, 1))).astype(np.int64))


def _rows_per_item(b, item_nnz):
    logg = max(min(GROUP_LOG[b], b), b - 2)
    logg = min(logg, 6)
    per_pass = 64 >> logg
    rpi = max(1, item_nnz // (1 << b))
    return -(-rpi // per_pass) * per_pass


def _class(inst):
    """Bound storage class per column: 0 int16 records, 1 int32, 2 FP64."""
    lb, ub = inst.lb, inst.ub
    integral = (inst.vartype != 3) & np.isfinite(lb) & np.isfinite(ub) & (np.floor(lb) == lb) & (np.floor(ub) == ub)
    small = integral & (np.abs(lb) <= 32767) & (np.abs(ub) <= 32767)
    wide = integral & (np.abs(lb) <= 2147483647) & (np.abs(ub) <= 2147483647)
    return np.where(small, 0, np.where(wide, 1, 2))


# binary columns a persistent launch of 512-thread workgroups holds in LDS:
# 64 KB of 2-bit codes, in whole 16-byte words
LDS_BINARY_MAX = 1 << 18


def _n_binary(inst):
    """Integer columns whose bounds (the envelope) lie in [0, 1]."""
    return int(np.sum((_class(inst) == 0) & (inst.lb >= 0) & (inst.ub <= 1)))


def expected_summary(instances, batch):
    item_nnz = 1024
    bin_rows = np.zeros(11, np.int64)
    bin_nnz = np.zeros(11, np.int64)
    n_items = 0
    for x in instances:
        length = np.diff(x.row_ptr).astype(np.int64)
        length = length[length > 0]
        b = _bin(length)
        bin_rows += np.bincount(b, minlength=11)
        bin_nnz += np.bincount(b, weights=length, minlength=11).astype(np.int64)
        nseg = -(-length[b == 10] // SEG_LEN)
        n_items += int(np.sum(-(-nseg // SEGS_PER_ITEM)))
        for k in range(10):
            cnt = int(np.sum(b == k))
            n_items += -(-cnt // _rows_per_item(k, item_nnz))
    per_class = np.zeros(3, np.int64)
    for x in instances:
        cnt = np.bincount(_class(x), minlength=3)
        per_class += -(-cnt // 32) * 32 if batch else cnt
    n_bin = 0 if batch else _n_binary(instances[0])
    return dict(bin_rows=bin_rows, bin_nnz=bin_nnz, m_nonempty=int(bin_rows.sum()), n_long_rows=int(bin_rows[10]),
                n_items=n_items, classes=per_class, n_binary=n_bin,
                n_lds_binary=min(n_bin, LDS_BINARY_MAX) // 64 * 64)


def check_summary(stats, instances, batch):
    want = expected_summary(instances, batch)
    assert list(stats.bin_rows) == want["bin_rows"].tolist()
    assert list(stats.bin_nnz) == want["bin_nnz"].tolist()
    assert stats.n_cons_nonempty == want["m_nonempty"]
    assert stats.n_long_rows == want["n_long_rows"]
    assert stats.n_items == want["n_items"]
    assert [stats.n16, stats.n32, stats.nf] == want["classes"].tolist()
    assert stats.n_binary == want["n_binary"]
    assert stats.n_lds_binary == want["n_lds_binary"]
    assert stats.n_cons == sum(x.m for x in instances)
    assert stats.nnz == sum(x.nnz for x in instances)


@pytest.mark.parametrize("make", [
    lambda: fixtures.layout_edges(0),
    lambda: fixtures.varied(1),
    lambda: small_mip(seed=3),
    lambda: powerlaw_mip(seed=2, m=60000, n=30000, s=1.8, len_hi=5000),
], ids=["layout_edges", "varied", "small", "powerlaw"])
def test_single_instance_layout_summary(make):
    inst = make()
    p = I.Problem(inst)
    try:
        check_summary(p.stats(), [inst], batch=False)
    finally:
        p.close()


def test_batch_layout_summary():
    """Instances with empty rows, long rows and every bin, side by side: the
    runs of each (bin, instance) are cut apart."""
    insts = [fixtures.layout_edges(1), fixtures.varied(2), small_mip(seed=5), fixtures.layout_edges(3)]
    insts += [small_mip(seed=s, m=m, n=m // 2, len_hi=hi) for s, m, hi in [(6, 3000, 40), (7, 500, 3), (8, 9000, 300)]]
    b = I.Batch(insts)
    try:
        check_summary(b.stats(), insts, batch=True)
    finally:
        b.close()


def test_wide_batch_layout_summary():
    """Many instances (a binary search per row finds its instance)."""
    insts = [small_mip(**spec) for spec in batch_family(seed=9, count=64, m_hi=20000)]
    b = I.Batch(insts)
    try:
        check_summary(b.stats(), insts, batch=True)
    finally:
        b.close()


def singleton_rows(m=17_000_000, n=1_000_000, seed=0):
    """m rows of one entry each (a x_j <= b or >= b) over n columns: more
    rows than one pass of the scans' tile totals holds (4096 tiles of 4096
    entries), so the row order's scans run in several chunks."""
    from gdp_amd.instances import Instance
    rng = np.random.default_rng(seed)
    col = rng.integers(0, n, size=m).astype(np.int32)
    val = np.where(rng.random(m) < 0.5, -1.0, 1.0) * rng.integers(1, 5, size=m)
    is_int = rng.random(n) < 0.6
    lb = np.where(is_int, 0.0, -100.0)
    ub = np.where(is_int, 50.0, 100.0)
    # each row cuts its column's domain, keeping a point x0 feasible
    x0 = np.where(is_int, rng.integers(10, 40, size=n), rng.uniform(-50, 50, size=n))
    act = val * x0[col]
    slack = rng.uniform(0, 30, size=m)
    upper = rng.random(m) < 0.5
    lhs = np.where(upper, -INF, act - slack)
    rhs = np.where(upper, act + slack, INF)
    return Instance(m, n, col, np.arange(m + 1, dtype=np.int32), val, lhs, rhs, lb, ub,
                    np.where(is_int, 1, 3).astype(np.int32), "singletons")


def test_seventeen_million_singleton_rows(oracle):
    inst = singleton_rows()
    p = I.Problem(inst)
    try:
        check_summary(p.stats(), [inst], batch=False)
        got = p.propagate(inst.lb, inst.ub, I.MEGAKERNEL)
    finally:
        p.close()
    exp = oracle.jacobi(inst)
    np.testing.assert_array_equal(got.lbs, exp[0])
    np.testing.assert_array_equal(got.ubs, exp[1])
    assert got.rounds == exp[2] and got.infeasible == exp[3]


def test_binaries_split_between_lds_and_records(oracle):
    """More binary columns than LDS holds: the first 2^18 are read from LDS,
    the rest from their int16 records, in the same rows.  Bit-identical to the
    oracle, and to the step-kernel loop, which reads every record."""
    inst = powerlaw_mip(seed=4, m=200_000, n=700_000, s=2.2, len_hi=3000)
    p = I.Problem(inst)
    try:
        st = p.stats()
        check_summary(st, [inst], batch=False)
        assert st.n_lds_binary == LDS_BINARY_MAX < st.n_binary
        mega = p.propagate(inst.lb, inst.ub, I.MEGAKERNEL)
        loop = p.propagate(inst.lb, inst.ub, I.CPU_LOOP)
    finally:
        p.close()
    exp = oracle.jacobi(inst)
    assert exp[2] > 2
    for got in (mega, loop):
        np.testing.assert_array_equal(got.lbs, exp[0])
        np.testing.assert_array_equal(got.ubs, exp[1])
        assert got.rounds == exp[2] and got.infeasible == exp[3]
