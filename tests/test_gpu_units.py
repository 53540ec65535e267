"""GPU parity of the shared-launch paths: a batch of independent instances,
many bound vectors against one matrix (with and without changed-column
hints), and the continuous-batching server.  Every unit must be
bit-identical to the oracle's round-synchronous propagation of that unit
alone (bounds, round count, infeasibility, cap)."""
import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import powerlaw_mip, small_mip, tighten_node

pytestmark = pytest.mark.gpu

SYNCS = [I.CPU_LOOP, I.GPU_LOOP, I.MEGAKERNEL]


def assert_same(got, exp, label=""):
    lb, ub, rounds, infeasible, capped = exp
    np.testing.assert_array_equal(got.lbs, lb, err_msg=f"{label} lb")
    np.testing.assert_array_equal(got.ubs, ub, err_msg=f"{label} ub")
    assert (got.rounds, bool(got.infeasible), bool(got.capped)) == (rounds, bool(infeasible), bool(capped)), label


def batch_instances():
    insts = [fixtures.analytic(n) for n in fixtures.ANALYTIC]
    insts += [small_mip(seed=s, m=200 + 150 * s, n=100 + 70 * s) for s in range(5)]
    insts += [small_mip(seed=7, int_frac=1.0), fixtures.generalized_cfp(12), fixtures.gate_ties(),
              fixtures.infeasible_instance(), fixtures.varied(1),
              powerlaw_mip(seed=5, m=6000, n=3000, len_hi=3000, s=1.8), fixtures.layout_edges(3)]
    return insts


@pytest.mark.parametrize("sync", SYNCS)
def test_batch_each_instance_matches_oracle(gpu, oracle, sync):
    insts = batch_instances()
    b = I.Batch(insts)
    try:
        outs = b.propagate(sync_type=sync)
    finally:
        b.close()
    for inst, got in zip(insts, outs):
        assert_same(got, oracle.jacobi(inst), f"{inst.name} sync={sync}")


def test_batch_matches_cpu_seq_fixpoint(gpu, oracle):
    """Each instance's fixpoint against its own cpu_seq: integer bounds
    exact, continuous within 1e-9 relative or absolute, same infeasibility."""
    insts = batch_instances()
    b = I.Batch(insts)
    try:
        outs = b.propagate(sync_type=I.MEGAKERNEL)
    finally:
        b.close()
    for inst, got in zip(insts, outs):
        seq = oracle.seq(inst)
        assert bool(got.infeasible) == bool(seq[3]), inst.name
        if seq[3] or seq[4] or got.capped:
            continue
        is_int = inst.vartype != 3
        np.testing.assert_array_equal(got.lbs[is_int], seq[0][is_int], err_msg=inst.name)
        np.testing.assert_array_equal(got.ubs[is_int], seq[1][is_int], err_msg=inst.name)
        for g, s in ((got.lbs, seq[0]), (got.ubs, seq[1])):
            d = np.abs(g - s)[~is_int]
            tol = np.maximum(1e-9, 1e-9 * np.abs(s[~is_int]))
            assert np.all(d <= tol), inst.name


def test_batch_round_prefixes(gpu, oracle):
    """A round cap applies to every instance of the batch."""
    insts = batch_instances()
    b = I.Batch(insts)
    try:
        for cap in (1, 2, 3):
            outs = b.propagate(sync_type=I.CPU_LOOP, max_rounds=cap)
            for inst, got in zip(insts, outs):
                assert_same(got, oracle.jacobi(inst, max_rounds=cap), f"{inst.name} cap={cap}")
    finally:
        b.close()


def test_batch_beyond_the_lds_unit_filter(gpu, oracle):
    """More instances than the persistent kernel notes in LDS per step
    (4096): its sweep then visits every item and each item checks its own
    instance's state.  Tiny instances that stop after different rounds."""
    insts = [small_mip(seed=s, m=8 + s % 13, n=6 + s % 7, len_lo=2, len_hi=5, name=f"tiny:{s}")
             for s in range(4100)]
    b = I.Batch(insts)
    try:
        outs = b.propagate(sync_type=I.MEGAKERNEL)
    finally:
        b.close()
    rounds = set()
    for inst, got in zip(insts, outs):
        assert_same(got, oracle.jacobi(inst), inst.name)
        rounds.add(got.rounds)
    assert len(rounds) > 2


def node_vectors(inst, oracle, count, seed, from_fixpoint=True):
    rng = np.random.default_rng(seed)
    base = oracle.jacobi(inst) if from_fixpoint else (inst.lb, inst.ub)
    nodes = [tighten_node(inst, rng, 1, 20, base[0], base[1]) for _ in range(count)]
    return nodes


@pytest.mark.parametrize("sync", SYNCS)
@pytest.mark.parametrize("case", ["small", "powerlaw"])
def test_nodes_match_oracle(gpu, oracle, case, sync):
    inst = small_mip(seed=3, int_frac=0.8) if case == "small" else powerlaw_mip(seed=5, m=20000, n=10000,
                                                                               len_hi=5000, s=1.8)
    nodes = node_vectors(inst, oracle, 24, seed=11)
    p = I.Problem(inst)
    try:
        outs = p.propagate_nodes([x[0] for x in nodes], [x[1] for x in nodes], sync_type=sync)
        hinted = p.propagate_nodes([x[0] for x in nodes], [x[1] for x in nodes], hints=[x[2] for x in nodes],
                                   sync_type=sync)
    finally:
        p.close()
    for k, (lb, ub, _) in enumerate(nodes):
        exp = oracle.jacobi(inst, lb=lb, ub=ub)
        assert_same(outs[k], exp, f"node {k}")
        assert_same(hinted[k], exp, f"hinted node {k}")


def test_single_hinted_nodes_with_binaries_in_lds(gpu, oracle):
    """One node vector per launch: the persistent kernel reads the binary
    columns from LDS, starting from a step that processes only the rows of
    the hinted columns, after launches of several vectors on the same units."""
    inst = powerlaw_mip(seed=5, m=20000, n=10000, len_hi=5000, s=1.8)
    nodes = node_vectors(inst, oracle, 6, seed=12)
    p = I.Problem(inst)
    try:
        assert p.stats().n_lds_binary > 0
        p.propagate_nodes([x[0] for x in nodes], [x[1] for x in nodes], sync_type=I.MEGAKERNEL)
        for k, (lb, ub, cols) in enumerate(nodes):
            got = p.propagate_nodes([lb], [ub], hints=[cols], sync_type=I.MEGAKERNEL)[0]
            assert_same(got, oracle.jacobi(inst, lb=lb, ub=ub), f"node {k}")
    finally:
        p.close()


def test_nodes_from_root_and_capacity_growth(gpu, oracle):
    """Vectors that are not near a fixpoint (no hints), more vectors than the
    reserved capacity, then fewer."""
    inst = small_mip(seed=4)
    nodes = node_vectors(inst, oracle, 9, seed=2, from_fixpoint=False)
    p = I.Problem(inst)
    try:
        p.reserve(2)
        for count in (9, 3):
            outs = p.propagate_nodes([x[0] for x in nodes[:count]], [x[1] for x in nodes[:count]],
                                     sync_type=I.MEGAKERNEL)
            for k in range(count):
                assert_same(outs[k], oracle.jacobi(inst, lb=nodes[k][0], ub=nodes[k][1]), f"node {k}")
    finally:
        p.close()


@pytest.mark.parametrize("case", ["small_int:7", "powerlaw:5", "gate_ties", "varied:2"])
def test_row_marking_off_same_bits(gpu, oracle, case):
    from test_gpu_parity import _case
    inst = _case(case)
    exp = oracle.jacobi(inst)
    for tuning in ({"row_marking": 0}, {"row_marking": 1}):
        p = I.Problem(inst, tuning)
        try:
            for sync in SYNCS:
                assert_same(p.propagate(inst.lb, inst.ub, sync), exp, f"{case} {tuning} sync={sync}")
        finally:
            p.close()


def _changed_row_sets(inst, oracle, rounds):
    """Per round k >= 2: the rows holding a column that round k-1 changed,
    and those columns (round-synchronous oracle, one round at a time)."""
    lens = np.diff(inst.row_ptr)
    row_of = np.repeat(np.arange(inst.m), lens)
    prev = (inst.lb, inst.ub)
    sets = []
    for k in range(1, rounds):
        cur = oracle.jacobi(inst, max_rounds=k)
        changed = (cur[0] != prev[0]) | (cur[1] != prev[1])
        rows = np.zeros(inst.m, bool)
        rows[row_of[changed[inst.col]]] = True
        sets.append((rows, changed))
        prev = cur[:2]
    return sets


@pytest.mark.parametrize("case", ["small:1", "powerlaw:5", "varied:2"])
def test_work_profile(gpu, oracle, case):
    """Round 1 processes every nonempty row; a later round processes every
    row (marking gave way to a full sweep) or exactly the rows holding a
    column the previous round changed.  A row longer than 256 entries is
    processed in 256-entry segments, and in a marked round only its segments
    holding a changed column (or not computed before in the run) are summed
    again.  The run itself is the propagation.  Runs of other bound vectors
    on the same problem come first: the marks they leave carry the same
    step numbers and must not make this run do more."""
    from test_gpu_parity import _case
    inst = _case(case)
    exp = oracle.jacobi(inst)
    lens = np.diff(inst.row_ptr)
    nonempty = lens > 0
    p = I.Problem(inst)
    try:
        for lb, ub, _ in node_vectors(inst, oracle, 3, seed=21, from_fixpoint=False):
            p.propagate(lb, ub, I.CPU_LOOP)
        out, rows, nnz, cols, acc = p.work_profile(inst.lb, inst.ub)
    finally:
        p.close()
    assert_same(out, exp, case)
    assert len(rows) == exp[2]
    all_cols = np.unique(inst.col).size
    assert (rows[0], nnz[0], cols[0]) == (nonempty.sum(), inst.nnz, all_cols)
    seg = 256
    computed = set()  # (row, segment) summed earlier in the run
    for r in np.flatnonzero(lens > seg):
        computed.update((r, g) for g in range((lens[r] + seg - 1) // seg))
    n_marked = 0
    for k, (marked, changed) in enumerate(_changed_row_sets(inst, oracle, exp[2]), start=1):
        if rows[k] == nonempty.sum() and nnz[k] == inst.nnz:
            for r in np.flatnonzero(lens > seg):
                computed.update((r, g) for g in range((lens[r] + seg - 1) // seg))
            continue
        n_marked += 1
        sel = marked & nonempty
        assert rows[k] == sel.sum(), f"round {k + 1}"
        want_nnz, want_cols = 0, set()
        for r in np.flatnonzero(sel):
            c = inst.col[inst.row_ptr[r]:inst.row_ptr[r + 1]]
            if lens[r] <= seg:
                want_nnz += lens[r]
                want_cols.update(c.tolist())
                continue
            for g in range((lens[r] + seg - 1) // seg):
                cg = c[g * seg:(g + 1) * seg]
                if (r, g) in computed and not changed[cg].any():
                    continue
                computed.add((r, g))
                want_nnz += cg.size
                want_cols.update(cg.tolist())
        assert nnz[k] == want_nnz, f"round {k + 1}"
        assert cols[k] == len(want_cols), f"round {k + 1}"
    assert n_marked > 0, "no round used row marks"
    p = I.Problem(inst, {"row_marking": 0})
    try:
        out, rows, nnz, cols, acc = p.work_profile(inst.lb, inst.ub)
    finally:
        p.close()
    assert_same(out, exp, case)
    assert np.all(rows == nonempty.sum()) and np.all(nnz == inst.nnz)


def test_server_results_alone_and_in_the_mix(gpu, oracle):
    """Requests submitted while others run join at the next step; every
    result equals the request propagated alone (oracle and the plain entry
    point)."""
    inst = powerlaw_mip(seed=6, m=20000, n=10000, len_hi=5000, s=1.8)
    nodes = node_vectors(inst, oracle, 40, seed=5)
    p = I.Problem(inst)
    try:
        alone = [p.propagate(lb, ub, I.CPU_LOOP) for lb, ub, _ in nodes]
        srv = I.Server(p, hbm_budget=64 << 20, max_slots=16)
        assert 1 <= srv.slots <= 16
        tickets = {}
        k = 0
        while k < len(nodes) or srv.pending():
            for _ in range(3):  # a few arrivals per step
                if k < len(nodes):
                    lb, ub, cols = nodes[k]
                    tickets[k] = srv.submit(lb, ub, hints=cols if k % 2 else None)
                    k += 1
            srv.step()
        for k, t in tickets.items():
            got = srv.result(t)
            assert got is not None, k
            exp = oracle.jacobi(inst, lb=nodes[k][0], ub=nodes[k][1])
            assert_same(got, exp, f"request {k}")
            np.testing.assert_array_equal(got.lbs, alone[k].lbs)
            np.testing.assert_array_equal(got.ubs, alone[k].ubs)
            assert got.rounds == alone[k].rounds
        assert srv.result(tickets[0]) is None  # results are handed out once
        srv.close()
    finally:
        p.close()


def test_server_budget_limits_slots(gpu):
    inst = small_mip(seed=1)
    p = I.Problem(inst)
    try:
        per = p.stats().unit_state_bytes + 64 * inst.n  # state, bound vectors, in-place inputs, staging, parking
        srv = I.Server(p, hbm_budget=3 * per + 1, max_slots=100)
        assert srv.slots == 3
        srv.close()
        with pytest.raises(RuntimeError):
            I.Server(p, hbm_budget=per // 2, max_slots=4)
    finally:
        p.close()


def test_server_zero_copy_requests(gpu, oracle):
    """Requests in caller-owned pinned memory (separate output arrays, or in
    place), mixed with copied and delta requests: results land in the
    caller's arrays, bit-identical to the oracle, and separate input arrays
    are left unchanged."""
    inst = powerlaw_mip(seed=6, m=20000, n=10000, len_hi=5000, s=1.8)
    base = oracle.jacobi(inst)
    nodes = node_vectors(inst, oracle, 24, seed=8)
    p = I.Problem(inst)
    pins = []
    try:
        srv = I.Server(p, hbm_budget=256 << 20, max_slots=6)
        srv.set_base(base[0], base[1])
        tickets, outs = {}, {}
        for k, (lb, ub, cols) in enumerate(nodes):
            if k % 6 == 0:  # separate output arrays
                a, b, ol, ou = (I.PinnedArray(inst.n) for _ in range(4))
                a.array[:] = lb
                b.array[:] = ub
                ol.array[:] = np.nan
                ou.array[:] = np.nan
                pins += [a, b, ol, ou]
                tickets[k] = srv.submit_pinned(a, b, ol, ou, hints=cols)
                outs[k] = (a, b, ol, ou)
            elif k % 6 == 3:  # in place: only the changed columns come back
                a, b = I.PinnedArray(inst.n), I.PinnedArray(inst.n)
                a.array[:] = lb
                b.array[:] = ub
                pins += [a, b]
                tickets[k] = srv.submit_pinned(a, b, a, b, hints=cols)
                outs[k] = (None, None, a, b)
            elif k % 3 == 1:
                tickets[k] = srv.submit(lb, ub, hints=cols)
            else:
                tickets[k] = srv.submit_delta(cols, lb[cols], ub[cols])
            if k % 5 == 4:
                srv.step()
        while srv.pending():
            srv.step()
        for k, t in tickets.items():
            exp = oracle.jacobi(inst, lb=nodes[k][0], ub=nodes[k][1])
            if k in outs:
                a, b, ol, ou = outs[k]
                got = srv.result(t, copy=False)
                assert got is not None and got.rounds == exp[2], f"request {k}"
                np.testing.assert_array_equal(ol.array, exp[0], err_msg=f"request {k}")
                np.testing.assert_array_equal(ou.array, exp[1], err_msg=f"request {k}")
                if a is not None:
                    np.testing.assert_array_equal(a.array, nodes[k][0])
                    np.testing.assert_array_equal(b.array, nodes[k][1])
            else:
                assert_same(srv.result(t), exp, f"request {k}")
        srv.close()
    finally:
        p.close()
        for x in pins:
            x.free()


def test_server_delta_requests(gpu, oracle):
    """Requests given as changes to a device-resident base (the propagated
    root) give the same bits as the full vectors and the oracle; the
    completed-ticket list names every request once."""
    inst = small_mip(seed=9, int_frac=0.8)
    base = oracle.jacobi(inst)
    nodes = node_vectors(inst, oracle, 30, seed=3)
    p = I.Problem(inst)
    try:
        srv = I.Server(p, hbm_budget=1 << 30, max_slots=8)
        srv.set_base(base[0], base[1])
        tickets = {}
        seen = []
        for k, (lb, ub, cols) in enumerate(nodes):
            if k % 3 == 0:
                tickets[k] = srv.submit(lb, ub, hints=cols)
            else:
                tickets[k] = srv.submit_delta(cols, lb[cols], ub[cols])
            if k % 4 == 3:
                srv.step()
                seen += srv.completed()
        while srv.pending():
            srv.step()
            seen += srv.completed()
        assert sorted(seen) == sorted(tickets.values())
        for k, t in tickets.items():
            exp = oracle.jacobi(inst, lb=nodes[k][0], ub=nodes[k][1])
            if k % 2:
                got = srv.result(t)
                assert_same(got, exp, f"request {k}")
            else:  # as changes against the base vector
                cols, cl, cu, info = srv.result_changes(t)
                lb, ub = base[0].copy(), base[1].copy()
                lb[cols] = cl
                ub[cols] = cu
                np.testing.assert_array_equal(lb, exp[0], err_msg=f"request {k}")
                np.testing.assert_array_equal(ub, exp[1], err_msg=f"request {k}")
                assert info.rounds == exp[2]
            assert srv.result(t) is None
        srv.close()
    finally:
        p.close()


def test_autotune_returns_a_valid_tuning(gpu, oracle):
    """The autotuner's pick gives the same bits (every candidate does)."""
    inst = powerlaw_mip(seed=5, m=20000, n=10000, len_hi=5000, s=1.8)
    tuning, ms = I.autotune(inst, I.MEGAKERNEL)
    assert ms > 0
    p = I.Problem(inst, tuning)
    try:
        assert_same(p.propagate(inst.lb, inst.ub, I.MEGAKERNEL), oracle.jacobi(inst), "autotuned")
    finally:
        p.close()
    insts = [small_mip(seed=s, m=300, n=150) for s in range(4)]
    bt, bms = I.batch_autotune(insts, I.CPU_LOOP)
    b = I.Batch(insts, bt)
    try:
        for x, got in zip(insts, b.propagate(sync_type=I.CPU_LOOP)):
            assert_same(got, oracle.jacobi(x), x.name)
    finally:
        b.close()


def test_server_results_waiting_for_collection(gpu, oracle):
    """A running server whose caller collects nothing until every request has
    completed: 12 copied requests with 4 slots, so 4 results go to the pinned
    outbox, 4 are parked in device memory and 4 go to host vectors; every
    result bit-identical to the oracle, each handed out once, and the entries
    reused by a second round of requests."""
    import time

    inst = powerlaw_mip(seed=7, m=20000, n=10000, len_hi=5000, s=1.8)
    nodes = node_vectors(inst, oracle, 24, seed=14)
    p = I.Problem(inst)
    try:
        srv = I.Server(p, hbm_budget=1 << 30, max_slots=4)
        assert srv.slots == 4
        srv.start()
        for rep in range(2):
            batch = nodes[12 * rep:12 * rep + 12]
            tickets = [srv.submit(lb, ub) for lb, ub, _ in batch]
            seen, t0 = [], time.time()
            while len(seen) < len(tickets) and time.time() - t0 < 60:
                seen += srv.completed()
                time.sleep(0.001)
            assert sorted(seen) == sorted(tickets), "every request completed before any result was taken"
            for t, (lb, ub, _) in zip(tickets, batch):
                assert_same(srv.result(t), oracle.jacobi(inst, lb=lb, ub=ub), f"rep {rep} ticket {t}")
                assert srv.result(t) is None
        srv.stop()
        srv.close()
    finally:
        p.close()


def test_server_stepping_thread(gpu, oracle):
    """The server's own stepping thread: requests of every form submitted
    from two caller threads while it steps, completions taken as they come;
    every result bit-identical to the oracle and every ticket completed
    once.  Stepping by hand is refused while the thread runs; the thread
    stops and restarts."""
    import threading
    import time

    inst = powerlaw_mip(seed=6, m=20000, n=10000, len_hi=5000, s=1.8)
    base = oracle.jacobi(inst)
    nodes = node_vectors(inst, oracle, 40, seed=12)
    p = I.Problem(inst)
    pins = []
    try:
        srv = I.Server(p, hbm_budget=256 << 20, max_slots=5)
        srv.set_base(base[0], base[1])
        for k in range(40):
            if k % 4 == 2:
                a, b = I.PinnedArray(inst.n), I.PinnedArray(inst.n)
                a.array[:] = nodes[k][0]
                b.array[:] = nodes[k][1]
                pins.append((k, a, b))
        pinned = {k: (a, b) for k, a, b in pins}
        tickets, lock = {}, threading.Lock()

        def submit(ks):
            for k in ks:
                lb, ub, cols = nodes[k]
                if k % 4 == 0:
                    t = srv.submit(lb, ub, hints=cols)
                elif k % 4 == 1:
                    t = srv.submit_delta(cols, lb[cols], ub[cols])
                elif k % 4 == 2:
                    a, b = pinned[k]
                    t = srv.submit_pinned(a, b, a, b, hints=cols)  # in place
                else:
                    t = srv.submit(lb, ub)
                with lock:
                    tickets[t] = k
                time.sleep(0.0005)

        for rep in range(2):
            for k, a, b in pins:
                a.array[:] = nodes[k][0]
                b.array[:] = nodes[k][1]
            tickets.clear()
            srv.start()
            with pytest.raises(RuntimeError, match="own thread"):
                srv.step()
            th = [threading.Thread(target=submit, args=(list(range(h, 40, 2)),)) for h in (0, 1)]
            for t_ in th:
                t_.start()
            seen = []
            t0 = time.time()
            while len(seen) < 40 and time.time() - t0 < 60:
                seen += srv.completed()
                time.sleep(0.001)
            for t_ in th:
                t_.join()
            srv.stop()
            assert sorted(seen) == sorted(tickets), "every ticket completed exactly once"
            assert srv.pending() == 0
            for t, k in tickets.items():
                exp = oracle.jacobi(inst, lb=nodes[k][0], ub=nodes[k][1])
                if k % 4 == 2:
                    got = srv.result(t, copy=False)
                    a, b = pinned[k]
                    assert got is not None and got.rounds == exp[2], f"rep {rep} request {k}"
                    np.testing.assert_array_equal(a.array, exp[0], err_msg=f"rep {rep} request {k}")
                    np.testing.assert_array_equal(b.array, exp[1], err_msg=f"rep {rep} request {k}")
                else:
                    assert_same(srv.result(t), exp, f"rep {rep} request {k}")
        srv.close()
    finally:
        p.close()
        for _, a, b in pins:
            a.free()
            b.free()


def test_problems_of_changing_sizes_reuse_device_memory(gpu, oracle):
    """Problems of different sizes created and destroyed in an interleaved
    order, some kept alive across others, each propagated through the
    persistent kernel and through host-driven steps: the device blocks that
    freed problems leave behind are handed to later ones, and every result
    stays bit-identical to the oracle."""
    rng = np.random.default_rng(17)
    alive = []
    for it in range(12):
        m = int(rng.choice([300, 2000, 9000, 30000]))
        inst = small_mip(seed=200 + it, m=m, n=max(2, m // 2), len_hi=int(rng.choice([5, 30, 400])))
        p = I.Problem(inst)
        exp = oracle.jacobi(inst)
        for sync in (I.MEGAKERNEL, I.CPU_LOOP):
            assert_same(p.propagate(inst.lb, inst.ub, sync), exp, f"iteration {it} sync {sync}")
        alive.append((p, inst, exp))
        if len(alive) > 3 or rng.random() < 0.4:
            q, qi, qe = alive.pop(int(rng.integers(0, len(alive))))
            assert_same(q.propagate(qi.lb, qi.ub, I.MEGAKERNEL), qe, f"iteration {it}, an older problem")
            q.close()
    for q, qi, qe in alive:
        assert_same(q.propagate(qi.lb, qi.ub, I.GPU_LOOP), qe, "the last problems")
        q.close()
