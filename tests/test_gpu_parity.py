"""GPU parity: libGpuProp.so's HIP path against the oracle, through the C ABI.

Bar (per the propagation contract):
  * against the round-synchronous oracle (same arithmetic, same summation
    order): bounds identical bit for bit (compared as doubles, so -0.0 ==
    +0.0), identical round count, infeasibility flag and cap flag, for every
    sync type and for every round prefix (max_rounds = 1, 2, ...);
  * against cpu_seq (the reference's Gauss-Seidel fixpoint): integer bounds
    identical, continuous bounds within 1e-9 relative or 1e-9 absolute, same
    infeasibility verdict (round counts differ by design and are reported).
"""
import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import INF, cover_knapsack, powerlaw_mip, small_mip

pytestmark = pytest.mark.gpu

CONT_TOL = 1e-9


def _case(name):
    kind, _, arg = name.partition(":")
    if kind == "analytic":
        return fixtures.analytic(arg)
    if kind == "small":
        return small_mip(seed=int(arg))
    if kind == "small_int":
        return small_mip(seed=int(arg), int_frac=1.0)
    if kind == "varied":
        return fixtures.varied(int(arg))
    if kind == "gcfp":
        return fixtures.generalized_cfp(int(arg))
    if kind == "gate_ties":
        return fixtures.gate_ties()
    if kind == "infeasible":
        return fixtures.infeasible_instance()
    if kind == "int32_overflow":
        return fixtures.int32_overflow()
    if kind == "layout_edges":
        return fixtures.layout_edges(int(arg))
    if kind == "tiny":
        return fixtures.tiny_integer(int(arg))
    if kind == "powerlaw":
        return powerlaw_mip(seed=int(arg), m=20000, n=10000, len_hi=5000, s=1.8)
    raise KeyError(name)


CASES = ([f"analytic:{n}" for n in fixtures.ANALYTIC] + [f"small:{s}" for s in range(3)] +
         ["small_int:7", "varied:0", "varied:1", "varied:2", "gcfp:30", "infeasible", "int32_overflow", "tiny:3",
          "tiny:11", "powerlaw:5", "gate_ties", "layout_edges:0", "layout_edges:1"])


def assert_same(got, exp, label=""):
    lb, ub, rounds, infeasible, capped = exp
    np.testing.assert_array_equal(got.lbs, lb, err_msg=f"{label} lb")
    np.testing.assert_array_equal(got.ubs, ub, err_msg=f"{label} ub")
    assert (got.rounds, got.infeasible, got.capped) == (rounds, infeasible, capped), label


@pytest.mark.parametrize("sync", [I.CPU_LOOP, I.GPU_LOOP, I.MEGAKERNEL])
@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_round_synchronous_oracle(gpu, oracle, case, sync):
    inst = _case(case)
    exp = oracle.jacobi(inst)
    got = I.gpu_propagate(inst, sync)
    assert_same(got, exp, f"{case} sync={sync}")


@pytest.mark.parametrize("case", ["analytic:achterberg", "analytic:two_lines", "small:0", "varied:1", "powerlaw:5"])
def test_every_round_prefix_matches(gpu, oracle, case):
    inst = _case(case)
    full = oracle.jacobi(inst)
    p = I.Problem(inst)
    try:
        for k in sorted({1, 2, 3, 5, max(1, full[2] - 1), full[2]}):
            exp = oracle.jacobi(inst, max_rounds=k)
            for sync in (I.CPU_LOOP, I.MEGAKERNEL, I.GPU_LOOP):
                got = p.propagate(inst.lb, inst.ub, sync, max_rounds=k)
                assert_same(got, exp, f"{case} rounds<={k} sync={sync}")
    finally:
        p.close()


@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_cpu_seq_fixpoint(gpu, oracle, case):
    inst = _case(case)
    seq = oracle.seq(inst)
    got = I.gpu_propagate(inst, I.MEGAKERNEL)
    assert got.infeasible == seq[3]
    if seq[3] or seq[4] or got.capped:
        return
    is_int = inst.vartype != 3
    np.testing.assert_array_equal(got.lbs[is_int], seq[0][is_int])
    np.testing.assert_array_equal(got.ubs[is_int], seq[1][is_int])
    for g, s in ((got.lbs, seq[0]), (got.ubs, seq[1])):
        d = np.abs(g - s)[~is_int]
        bound = np.maximum(CONT_TOL, CONT_TOL * np.abs(s[~is_int]))
        assert np.all((d <= bound) | (g[~is_int] == s[~is_int])), f"{case}: max diff {d.max(initial=0)}"


def test_problem_reuse_many_bound_vectors(gpu, oracle):
    inst = small_mip(seed=11)
    rng = np.random.default_rng(0)
    p = I.Problem(inst)
    try:
        for t in range(6):
            lb, ub = inst.lb.copy(), inst.ub.copy()
            pick = rng.choice(inst.n, size=30, replace=False)
            for j in pick:
                if inst.vartype[j] != 3 and lb[j] > -INF and ub[j] < INF and ub[j] - lb[j] >= 1:
                    ub[j] = np.floor((lb[j] + ub[j]) / 2)
            exp = oracle.jacobi(inst.with_bounds(lb, ub))
            got = p.propagate(lb, ub, [I.CPU_LOOP, I.MEGAKERNEL, I.GPU_LOOP][t % 3])
            assert_same(got, exp, f"vector {t}")
    finally:
        p.close()


def test_legacy_entry_points(gpu, oracle):
    inst = fixtures.analytic("achterberg")
    args = (inst.n, inst.m, inst.nnz, inst.col.tolist(), inst.row_ptr.tolist(), inst.val.tolist(), inst.lhs.tolist(),
            inst.rhs.tolist(), inst.lb.tolist(), inst.ub.tolist(), inst.vartype.tolist())
    for st in (0, 1, 2):
        lbs, ubs = I.propagateGPUAtomic(*args, synctype=st)
        assert ubs == inst.meta["ub_expected"].tolist() and lbs == inst.meta["lb_expected"].tolist()
    lbs, ubs = I.propagateGPUReduction(*args)
    assert ubs == inst.meta["ub_expected"].tolist()
    from ctypes import c_float
    lbs, ubs = I.propagateGPUAtomic(*args, synctype=2, datatype=c_float)
    assert ubs == inst.meta["ub_expected"].tolist()
    tl = fixtures.analytic("two_lines")
    exp = oracle.jacobi(tl)
    targs = (tl.n, tl.m, tl.nnz, tl.col.tolist(), tl.row_ptr.tolist(), tl.val.tolist(), tl.lhs.tolist(),
             tl.rhs.tolist(), tl.lb.tolist(), tl.ub.tolist(), tl.vartype.tolist())
    lbs, ubs = I.propagateGPUAtomic(*targs, synctype=0, datatype=c_float)
    assert np.all(np.array(lbs, np.float64) <= exp[0]) and np.all(np.array(ubs, np.float64) >= exp[1])


def test_ieee_infinity(gpu, oracle):
    for name in ("savelsbergh1", "savelsbergh_cflp"):
        inst = fixtures.analytic(name)
        ieee = inst.with_bounds(np.where(inst.lb <= -INF, -np.inf, inst.lb), np.where(inst.ub >= INF, np.inf, inst.ub))
        ieee.lhs = np.where(inst.lhs <= -INF, -np.inf, inst.lhs)
        ieee.rhs = np.where(inst.rhs >= INF, np.inf, inst.rhs)
        exp = oracle.jacobi(ieee)
        for sync in (0, 2):
            got = I.gpu_propagate(ieee, sync)
            assert_same(got, exp, name)


def test_input_with_crossed_bounds(gpu):
    inst = fixtures.analytic("achterberg")
    lb = inst.lb.copy()
    lb[3] = 2.0  # above ub = 1
    got = I.Problem(inst).propagate(lb, inst.ub, I.MEGAKERNEL)
    assert got.infeasible and got.rounds == 0


TUNINGS = {
    "no_split": dict(max_split_segs=0),          # rows longer than 256 on one wave
    "split_upto_4": dict(max_split_segs=4),
    "fp64_bounds": dict(compact_bounds=0),       # no compact bound records
    "one_entry_per_lane": dict(group_log=[0, 1, 2, 3, 4, 5, 6, 6, 6, 6]),
    "small_items": dict(item_nnz=128),
    "marks_always": dict(mark_limit=1 << 30),     # row marks in every round, never a full sweep
    "mega_256": dict(mega_threads=256),          # persistent workgroups of 256 / 1024 threads
    "mega_1024": dict(mega_threads=1024),
}


@pytest.mark.parametrize("tuning", sorted(TUNINGS))
@pytest.mark.parametrize("case", ["powerlaw:5", "gate_ties", "varied:2", "small_int:7"])
def test_tunings_bit_exact(gpu, oracle, case, tuning):
    """Every layout option yields the same bits: the summation order is a
    function of the row length only."""
    inst = _case(case)
    exp = oracle.jacobi(inst)
    p = I.Problem(inst, TUNINGS[tuning])
    try:
        for sync in (I.CPU_LOOP, I.MEGAKERNEL):
            assert_same(p.propagate(inst.lb, inst.ub, sync), exp, f"{case} {tuning} sync={sync}")
    finally:
        p.close()


@pytest.mark.parametrize("one_frac", [0.5, 0.0])
def test_cover_knapsack_full_size(gpu, oracle, one_frac):
    """The structured binary instance at full size (m=400,000, n=2,000,000,
    25.3M nonzeros): as generated (fixings to 1 overfill knapsack rows, so
    round 1 proves infeasibility) and with fixings to 0 only (feasible)."""
    inst = cover_knapsack(one_frac=one_frac)
    exp = oracle.jacobi(inst)
    seq = oracle.seq(inst)
    assert exp[3] == seq[3] == (one_frac > 0)
    p = I.Problem(inst)
    try:
        for sync in (I.MEGAKERNEL, I.CPU_LOOP):
            assert_same(p.propagate(inst.lb, inst.ub, sync), exp, f"one_frac={one_frac} sync={sync}")
    finally:
        p.close()
    if not exp[3]:
        np.testing.assert_array_equal(exp[0], seq[0])
        np.testing.assert_array_equal(exp[1], seq[1])


def test_column_index_out_of_range_is_an_error(gpu):
    # checked on the device while the layout is built; no problem is created
    inst = fixtures.layout_edges(2)
    bad = inst.col.copy()
    bad[len(bad) // 2] = inst.n
    broken = type(inst)(inst.m, inst.n, bad, inst.row_ptr, inst.val, inst.lhs, inst.rhs, inst.lb, inst.ub,
                        inst.vartype, "bad_column")
    with pytest.raises(RuntimeError, match="column index out of range"):
        I.Problem(broken)
    bad[len(bad) // 2] = -1
    with pytest.raises(RuntimeError, match="column index out of range"):
        I.Batch([fixtures.analytic("achterberg"), broken])
    # the device is still usable afterwards
    exp = fixtures.analytic("achterberg")
    got = I.gpu_propagate(exp, I.MEGAKERNEL)
    np.testing.assert_allclose(got.lbs, exp.meta["lb_expected"], atol=1e-4)


def test_large_instance_through_staged_uploads(gpu, oracle):
    """9.4M nonzeros: the matrix, its coefficients and the bound vectors
    cross to the device through the library's pinned staging buffers in
    several chunks, and the drop-in entry point copies the result back."""
    import ctypes
    inst = powerlaw_mip(seed=3, m=500_000, n=250_000)
    assert inst.nnz * 8 > 64 << 20  # the coefficients take more than two 32 MiB chunks
    exp = oracle.jacobi(inst)
    p = I.Problem(inst)
    try:
        got = p.propagate(inst.lb, inst.ub, I.MEGAKERNEL)
    finally:
        p.close()
    np.testing.assert_array_equal(got.lbs, exp[0])
    np.testing.assert_array_equal(got.ubs, exp[1])
    assert got.rounds == exp[2]
    lib = I.load_library()
    lb, ub = inst.lb.copy(), inst.ub.copy()
    P = lambda a: ctypes.c_void_p(np.ascontiguousarray(a).ctypes.data)  # noqa: E731
    keep = [np.ascontiguousarray(a) for a in (inst.col, inst.row_ptr, inst.val, inst.lhs, inst.rhs, inst.vartype)]
    rc = lib.propagateConstraintsGPUAtomicDouble(inst.m, inst.n, inst.nnz, *(P(a) for a in keep[:5]), P(lb), P(ub),
                                                 P(keep[5]), I.MEGAKERNEL)
    assert rc == I.GDP_OKAY
    np.testing.assert_array_equal(lb, exp[0])
    np.testing.assert_array_equal(ub, exp[1])


@pytest.mark.parametrize("sync", [I.CPU_LOOP, I.GPU_LOOP, I.MEGAKERNEL])
@pytest.mark.parametrize("index", [56, 136, 150, 156])
def test_batch_tolerance_cases_alone(gpu, oracle, index, sync):
    """The batched workload's instances whose fixpoint lies more than 1e-9
    from cpu_seq's in one continuous bound (56, 136) and whose propagation
    reaches the round cap (150, 156), each propagated alone: bit for bit the
    oracle, the cap reported as such."""
    from gdp_amd.instances import batch_family
    inst = small_mip(**batch_family(count=256)[index])
    exp = oracle.jacobi(inst)
    assert exp[4] == (index in (150, 156))
    assert_same(I.gpu_propagate(inst, sync), exp, f"instance {index} sync={sync}")
