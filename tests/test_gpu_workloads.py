"""GPU parity at the full sizes of the benchmark workloads, and the
boundary's concurrency and ownership rules, through the C ABI.

* The 1M-row power-law MIP (the headline workload, Zipf(1.95), 19.3M
  nonzeros) and its literal Zipf(1.8) shape (41.6M nonzeros): bit for bit
  against the round-synchronous oracle.  Against cpu_seq: integer bounds
  identical, the same infeasibility verdict, continuous bounds within 1e-9
  relative or absolute except the bounds pinned per workload, and each
  fixpoint one of the other propagator up to the rounding of the two
  summation orders (a round of either propagator started from the other's
  fixpoint moves no integer bound and no continuous bound by more than
  2e-9).
* The 256 uneven MIPs of the batched workload (96M nonzeros in one launch),
  each against its own oracle runs; the two bounds beyond 1e-9 and the
  instance that reaches the round cap are pinned.
* The serving matrix (m = 200,000, 2M nonzeros) with 64 requests of every
  form through the server's stepping thread.
* Bound vectors of 5M columns through the staged copies in several chunks.
* Two threads sharing one problem; a server owning its problem's slots.
* The float entry points: exactly the FP64 result rounded outward.
"""
import threading

import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import CONT, batch_family, powerlaw_mip, small_mip, tighten_node

pytestmark = pytest.mark.gpu

CONT_TOL = 1e-9


def assert_same(got, exp, label=""):
    lb, ub, rounds, infeasible, capped = exp
    np.testing.assert_array_equal(got.lbs, lb, err_msg=f"{label} lb")
    np.testing.assert_array_equal(got.ubs, ub, err_msg=f"{label} ub")
    assert (got.rounds, bool(got.infeasible), bool(got.capped)) == (rounds, bool(infeasible), bool(capped)), label


def continuous_deviation(inst, lbs, ubs, ref_l, ref_u):
    """(integer mismatches, max continuous deviation relative or absolute,
    continuous bounds beyond CONT_TOL) of lbs/ubs against ref_l/ref_u."""
    is_int = inst.vartype != CONT
    bad_int = int(np.sum(lbs[is_int] != ref_l[is_int]) + np.sum(ubs[is_int] != ref_u[is_int]))
    worst, beyond = 0.0, 0
    for g, s in ((lbs, ref_l), (ubs, ref_u)):
        d = np.abs(g - s)[~is_int] / np.maximum(1.0, np.abs(s[~is_int]))
        d[g[~is_int] == s[~is_int]] = 0.0  # equal infinities
        worst = max(worst, float(d.max(initial=0.0)))
        beyond += int(np.sum(d > CONT_TOL))
    return bad_int, worst, beyond


def check_against_cpu_seq(inst, got, oracle, label, beyond_expected=0):
    """The engine's fixpoint against cpu_seq's (see the module docstring);
    beyond_expected: continuous bounds of this workload known to differ by
    more than 1e-9 (two epsilon-fixpoints reached along different paths)."""
    seq = oracle.seq(inst)
    assert bool(got.infeasible) == bool(seq[3]), label
    assert not seq[4], f"{label}: cpu_seq reached the round cap"
    bad_int, worst, beyond = continuous_deviation(inst, got.lbs, got.ubs, seq[0], seq[1])
    print(f"{label}: rounds {got.rounds} (cpu_seq {seq[2]}), integer mismatches {bad_int}, "
          f"max continuous deviation {worst:.3e}, beyond 1e-9: {beyond}", flush=True)
    assert bad_int == 0 and beyond == beyond_expected, (label, bad_int, worst, beyond)
    # each fixpoint is one of the other propagator up to the rounding of the
    # two summation orders: a round started from it moves no integer bound
    # and no continuous bound by more than 2e-9
    for name, run, start in (("cpu_seq from the engine's", oracle.seq, got),
                             ("round-synchronous from cpu_seq's", oracle.jacobi, seq)):
        sl, su = (start.lbs, start.ubs) if hasattr(start, "lbs") else (start[0], start[1])
        again = run(inst, lb=sl, ub=su)
        moved_int, moved, _ = continuous_deviation(inst, again[0], again[1], sl, su)
        assert moved_int == 0 and moved <= 2 * CONT_TOL, (label, name, moved_int, moved)
    return seq


# continuous bounds beyond 1e-9 of cpu_seq's per workload shape
POWER_LAW_BEYOND_TOL = {1.95: 0, 1.8: 2}


@pytest.mark.parametrize("zipf_s", [1.95, 1.8])
def test_power_law_million_rows(gpu, oracle, zipf_s):
    inst = powerlaw_mip(seed=1, s=zipf_s)
    print(f"power law s={zipf_s}: m={inst.m} n={inst.n} nnz={inst.nnz}", flush=True)
    exp = oracle.jacobi(inst)
    p = I.Problem(inst)
    try:
        syncs = (I.MEGAKERNEL, I.GPU_LOOP) if zipf_s == 1.95 else (I.MEGAKERNEL,)
        for sync in syncs:
            got = p.propagate(inst.lb, inst.ub, sync)
            assert_same(got, exp, f"s={zipf_s} sync={sync}")
        seq = check_against_cpu_seq(inst, got, oracle, f"s={zipf_s}", POWER_LAW_BEYOND_TOL[zipf_s])
        # the engine started from cpu_seq's fixpoint: the oracle's run, bit for bit
        assert_same(p.propagate(seq[0], seq[1], I.MEGAKERNEL), oracle.jacobi(inst, lb=seq[0], ub=seq[1]),
                    f"s={zipf_s} from cpu_seq's fixpoint")
    finally:
        p.close()


# ---------------------------------------------------------------------------
# the batched workload: 256 uneven MIPs in one launch
# ---------------------------------------------------------------------------

# Batch instances whose engine fixpoint and cpu_seq fixpoint differ by more
# than 1e-9 in one continuous bound each (two epsilon-fixpoints reached by
# different acceptance paths; the tolerance cases of the parity tests), and
# those whose round-synchronous propagation reaches the 100-round cap (150:
# cpu_seq converges in 84 rounds; 156: cpu_seq is capped too).
BEYOND_TOL = {56: 1, 136: 1}
CAPPED = {150, 156}


@pytest.fixture(scope="module")
def batch_run(gpu):
    insts = [small_mip(**sp) for sp in batch_family(count=256)]
    print(f"batch: {len(insts)} instances, {sum(x.nnz for x in insts)} nonzeros", flush=True)
    b = I.Batch(insts)
    try:
        outs = b.propagate(sync_type=I.MEGAKERNEL)
    finally:
        b.close()
    return insts, outs


@pytest.mark.parametrize("quarter", range(4))
def test_batch_of_256_each_instance(batch_run, oracle, quarter):
    insts, outs = batch_run
    for i in range(quarter * 64, (quarter + 1) * 64):
        inst, got = insts[i], outs[i]
        label = f"instance {i}"
        assert_same(got, oracle.jacobi(inst), label)
        assert bool(got.capped) == (i in CAPPED), label
        seq = oracle.seq(inst)
        assert bool(got.infeasible) == bool(seq[3]), label
        if got.capped or seq[3]:
            continue
        bad_int, worst, beyond = continuous_deviation(inst, got.lbs, got.ubs, seq[0], seq[1])
        assert bad_int == 0, label
        assert beyond == BEYOND_TOL.get(i, 0), (label, worst)
        # both are epsilon-fixpoints of each other's propagation (integer data:
        # the two summation orders give the same sums)
        again = oracle.seq(inst, lb=got.lbs, ub=got.ubs)
        assert again[2] == 1 and np.array_equal(again[0], got.lbs) and np.array_equal(again[1], got.ubs), label
        back = oracle.jacobi(inst, lb=seq[0], ub=seq[1])
        assert back[2] == 1 and np.array_equal(back[0], seq[0]) and np.array_equal(back[1], seq[1]), label


# ---------------------------------------------------------------------------
# serving: the 2M-nonzero matrix, requests of every form
# ---------------------------------------------------------------------------

def test_serving_matrix_requests_of_every_form(gpu, oracle):
    """64 node requests (full vectors copied, zero-copy with separate output
    arrays, zero-copy in place, changes to the propagated root) submitted
    from two threads while the server's stepping thread runs with 16 slots,
    so that requests queue and join running ones; every result bit for bit
    the oracle's propagation of that node alone."""
    import time

    inst = powerlaw_mip(seed=4, m=200_000, n=100_000, s=2.08, len_hi=5000)
    root = oracle.jacobi(inst)
    rng = np.random.default_rng(21)
    nodes = [tighten_node(inst, rng, 1, 50, root[0], root[1]) for _ in range(64)]
    expected = [oracle.jacobi(inst, lb=lb, ub=ub) for lb, ub, _ in nodes]
    p = I.Problem(inst)
    pins = []
    try:
        srv = I.Server(p, hbm_budget=8 << 30, max_slots=16)
        srv.set_base(root[0], root[1])
        io = {}
        for k in range(64):
            if k % 4 in (1, 2):
                a, b = I.PinnedArray(inst.n), I.PinnedArray(inst.n)
                a.array[:] = nodes[k][0]
                b.array[:] = nodes[k][1]
                outs = (I.PinnedArray(inst.n), I.PinnedArray(inst.n)) if k % 4 == 1 else (a, b)
                pins += [a, b] + ([outs[0], outs[1]] if k % 4 == 1 else [])
                io[k] = (a, b) + outs
        tickets, lock = {}, threading.Lock()

        def submit(ks):
            for k in ks:
                lb, ub, cols = nodes[k]
                if k % 4 == 0:
                    t = srv.submit(lb, ub, hints=cols)
                elif k % 4 in (1, 2):
                    a, b, ol, ou = io[k]
                    t = srv.submit_pinned(a, b, ol, ou, hints=cols)
                else:
                    t = srv.submit_delta(cols, lb[cols], ub[cols])
                with lock:
                    tickets[t] = k

        srv.start()
        th = [threading.Thread(target=submit, args=(list(range(h, 64, 2)),)) for h in (0, 1)]
        for t_ in th:
            t_.start()
        seen = []
        t0 = time.time()
        while len(seen) < 64 and time.time() - t0 < 60:
            seen += srv.completed()
            time.sleep(0.001)
        for t_ in th:
            t_.join()
        srv.stop()
        assert sorted(seen) == sorted(tickets), "every ticket completed exactly once"
        for t, k in tickets.items():
            exp = expected[k]
            if k % 4 in (1, 2):
                got = srv.result(t, copy=False)
                a, b, ol, ou = io[k]
                assert got is not None and got.rounds == exp[2], f"request {k}"
                np.testing.assert_array_equal(ol.array, exp[0], err_msg=f"request {k}")
                np.testing.assert_array_equal(ou.array, exp[1], err_msg=f"request {k}")
                if k % 4 == 1:  # separate inputs are left as they were
                    np.testing.assert_array_equal(a.array, nodes[k][0])
            elif k % 4 == 3:
                cols, cl, cu, info = srv.result_changes(t)
                lb, ub = root[0].copy(), root[1].copy()
                lb[cols] = cl
                ub[cols] = cu
                np.testing.assert_array_equal(lb, exp[0], err_msg=f"request {k}")
                np.testing.assert_array_equal(ub, exp[1], err_msg=f"request {k}")
                assert info.rounds == exp[2]
            else:
                assert_same(srv.result(t), exp, f"request {k}")
        srv.close()
    finally:
        p.close()
        for x in pins:
            x.free()


# ---------------------------------------------------------------------------
# the boundary
# ---------------------------------------------------------------------------

def test_bound_vectors_in_several_staging_chunks(gpu, oracle):
    """5M columns: each bound vector (40 MB) crosses PCIe through the pinned
    staging in two 32 MiB chunks each way, through gdpPropagate and the
    drop-in entry point."""
    import ctypes
    inst = small_mip(seed=4, m=40_000, n=5_000_000)
    assert inst.n * 8 > 32 << 20
    exp = oracle.jacobi(inst)
    p = I.Problem(inst)
    try:
        for sync in (I.MEGAKERNEL, I.CPU_LOOP):
            assert_same(p.propagate(inst.lb, inst.ub, sync), exp, f"sync={sync}")
    finally:
        p.close()
    lib = I.load_library()
    lb, ub = inst.lb.copy(), inst.ub.copy()
    keep = [np.ascontiguousarray(a) for a in (inst.col, inst.row_ptr, inst.val, inst.lhs, inst.rhs, inst.vartype)]
    P = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = lib.propagateConstraintsGPUAtomicDouble(inst.m, inst.n, inst.nnz, *(P(a) for a in keep[:5]), P(lb), P(ub),
                                                 P(keep[5]), I.GPU_LOOP)
    assert rc == I.GDP_OKAY
    np.testing.assert_array_equal(lb, exp[0])
    np.testing.assert_array_equal(ub, exp[1])


def test_two_threads_share_one_problem(gpu, oracle):
    """Calls on one problem from two threads are serialised by the problem,
    host-array copies included: with GPU_LOOP a call's copies must not land in
    another call's graph capture."""
    inst = powerlaw_mip(seed=6, m=20000, n=10000, len_hi=5000, s=1.8)
    root = oracle.jacobi(inst)
    rng = np.random.default_rng(3)
    nodes = [tighten_node(inst, rng, 1, 50, root[0], root[1]) for _ in range(16)]
    expected = [oracle.jacobi(inst, lb=lb, ub=ub) for lb, ub, _ in nodes]
    p = I.Problem(inst)
    results, errors = {}, []

    def run(ks):
        try:
            for k in ks:
                sync = I.GPU_LOOP if k % 3 else I.MEGAKERNEL
                results[k] = p.propagate(nodes[k][0], nodes[k][1], sync)
        except Exception as e:  # reported below
            errors.append(e)

    try:
        th = [threading.Thread(target=run, args=(list(range(h, 16, 2)),)) for h in (0, 1)]
        for t_ in th:
            t_.start()
        for t_ in th:
            t_.join()
        assert not errors, errors
        for k in range(16):
            assert_same(results[k], expected[k], f"vector {k}")
    finally:
        p.close()


def test_server_owns_its_problems_slots(gpu, oracle):
    """While a server is attached, the problem refuses work that would touch
    the server's unit slots, and its destruction; afterwards it is usable."""
    inst = small_mip(seed=3)
    exp = oracle.jacobi(inst)
    p = I.Problem(inst)
    srv = I.Server(p, hbm_budget=64 << 20, max_slots=4)
    try:
        with pytest.raises(RuntimeError, match="server is attached"):
            p.reserve(8)
        with pytest.raises(RuntimeError, match="server is attached"):
            p.propagate(inst.lb, inst.ub, I.MEGAKERNEL)
        dl, du = I.DeviceArray(inst.lb), I.DeviceArray(inst.ub)
        with pytest.raises(RuntimeError, match="server is attached"):
            p.propagate_device(dl.ptr, du.ptr, I.CPU_LOOP)
        with pytest.raises(RuntimeError, match="server is attached"):
            p.propagate_nodes_device([dl.ptr] * 6, [du.ptr] * 6, None, I.CPU_LOOP)
        with pytest.raises(RuntimeError, match="destroy the server first"):
            p.close()
        with pytest.raises(RuntimeError, match="server is attached"):
            I.Server(p, hbm_budget=64 << 20, max_slots=2)
        # the server itself still works
        t = srv.submit(inst.lb, inst.ub)
        while srv.pending():
            srv.step()
        assert_same(srv.result(t), exp, "served")
        dl.free()
        du.free()
    finally:
        srv.close()
    try:
        assert_same(p.propagate(inst.lb, inst.ub, I.MEGAKERNEL), exp, "after the server")
    finally:
        p.close()


def _round_outward(v, inp, lower):
    """Largest float32 not above v (lower) / smallest not below (upper) where
    the FP64 bound v moved from the float input inp; the input elsewhere."""
    v = np.asarray(v, np.float64)
    f = v.astype(np.float32)
    cands = np.stack([np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))])
    c64 = cands.astype(np.float64)
    if lower:
        pick = np.where(c64 <= v, c64, -np.inf).max(axis=0)
    else:
        pick = np.where(c64 >= v, c64, np.inf).min(axis=0)
    return np.where(v != inp.astype(np.float64), pick, inp.astype(np.float64)).astype(np.float32)


@pytest.mark.parametrize("entry,sync", [("propagateConstraintsGPUAtomicFloat", 0),
                                        ("propagateConstraintsGPUAtomicFloat", 1),
                                        ("propagateConstraintsGPUAtomicFloat", 2),
                                        ("propagateConstraintsGPUReductionFloat", None)])
def test_float_entry_points_round_the_fp64_result_outward(gpu, oracle, entry, sync):
    import ctypes
    for inst in (fixtures.analytic("two_lines"), small_mip(seed=3), fixtures.varied(0)):
        f32 = [np.ascontiguousarray(a, np.float32) for a in (inst.val, inst.lhs, inst.rhs, inst.lb, inst.ub)]
        as64 = inst.with_bounds(f32[3].astype(np.float64), f32[4].astype(np.float64))
        as64.val, as64.lhs, as64.rhs = (a.astype(np.float64) for a in f32[:3])
        exp = oracle.jacobi(as64)
        want_l = _round_outward(exp[0], f32[3], lower=True)
        want_u = _round_outward(exp[1], f32[4], lower=False)
        lb, ub = f32[3].copy(), f32[4].copy()
        col = np.ascontiguousarray(inst.col, np.int32)
        rp = np.ascontiguousarray(inst.row_ptr, np.int32)
        vt = np.ascontiguousarray(inst.vartype, np.int32)
        P = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        args = [inst.m, inst.n, inst.nnz, P(col), P(rp), P(f32[0]), P(f32[1]), P(f32[2]), P(lb), P(ub), P(vt)]
        if sync is not None:
            args.append(sync)
        assert getattr(I.load_library(), entry)(*args) == I.GDP_OKAY
        np.testing.assert_array_equal(lb, want_l, err_msg=f"{inst.name} {entry}")
        np.testing.assert_array_equal(ub, want_u, err_msg=f"{inst.name} {entry}")
