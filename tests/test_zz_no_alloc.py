"""The allocation-free entry point gdpPropagateDeviceNoAlloc against the entry
point it varies, gdpPropagateDevice, through the C ABI.

Its run state lives in one fixed device array per device, shared by every
call and keeping the last call's contents.  Each test therefore makes two
calls in a row, the second on a smaller problem, and checks every result
bit for bit against gdpPropagateDevice on the same input (and against the
round-synchronous oracle).  A problem run only through the allocation-free
entry point never allocates its own run state (its stats show no units).
"""
import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import from_rows, powerlaw_mip, small_mip

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _propagate(inst, sync, no_alloc):
    """One propagation of the root bounds on a fresh problem."""
    p = I.Problem(inst)
    try:
        dl, du = I.DeviceArray(inst.lb), I.DeviceArray(inst.ub)
        if no_alloc:
            info = p.propagate_device_no_alloc(dl.ptr, du.ptr, sync)
            assert p.stats().n_units == 0, "the allocation-free call allocated the problem's run state"
        else:
            info = p.propagate_device(dl.ptr, du.ptr, sync)
        out = (dl.numpy(), du.numpy(), info.rounds, bool(info.infeasible), bool(info.capped))
        dl.free()
        du.free()
        return out
    finally:
        p.close()


def _check(inst, sync, oracle, label):
    print(f"no-alloc case {label} sync={sync} m={inst.m} n={inst.n} nnz={inst.nnz}", flush=True)
    got = _propagate(inst, sync, no_alloc=True)
    ref = _propagate(inst, sync, no_alloc=False)
    np.testing.assert_array_equal(got[0], ref[0], err_msg=f"{label} lb")
    np.testing.assert_array_equal(got[1], ref[1], err_msg=f"{label} ub")
    assert got[2:] == ref[2:], (label, got[2:], ref[2:])
    exp = oracle.jacobi(inst)
    np.testing.assert_array_equal(got[0], exp[0], err_msg=f"{label} lb vs oracle")
    np.testing.assert_array_equal(got[1], exp[1], err_msg=f"{label} ub vs oracle")
    assert got[2:] == (exp[2], exp[3], exp[4]), (label, got[2:], exp[2:])


@pytest.mark.parametrize("sync", [I.CPU_LOOP, I.GPU_LOOP, I.MEGAKERNEL])
def test_two_calls_second_smaller(gpu, oracle, sync):
    # a problem with split rows (a 70,000-entry row) first, then a smaller one
    _check(fixtures.layout_edges(0), sync, oracle, "layout_edges:0")
    _check(small_mip(seed=1), sync, oracle, "small_mip:1")


def test_power_law_then_small(gpu, oracle):
    big = powerlaw_mip(seed=5, m=20000, n=10000, len_hi=5000, s=1.8)
    _check(big, I.MEGAKERNEL, oracle, "powerlaw:5")
    _check(fixtures.gate_ties(), I.MEGAKERNEL, oracle, "gate_ties")


def test_owned_state_kept_across_calls(gpu, oracle):
    """The same problem alternating between its own run state and the array:
    neither disturbs the other."""
    inst = small_mip(seed=2)
    exp = oracle.jacobi(inst)
    p = I.Problem(inst)
    try:
        for k, no_alloc in enumerate([False, True, False, True]):
            dl, du = I.DeviceArray(inst.lb), I.DeviceArray(inst.ub)
            fn = p.propagate_device_no_alloc if no_alloc else p.propagate_device
            info = fn(dl.ptr, du.ptr, I.GPU_LOOP)
            np.testing.assert_array_equal(dl.numpy(), exp[0], err_msg=f"call {k}")
            np.testing.assert_array_equal(du.numpy(), exp[1], err_msg=f"call {k}")
            assert (info.rounds, bool(info.infeasible)) == (exp[2], exp[3])
            dl.free()
            du.free()
    finally:
        p.close()


def test_run_state_beyond_the_array_is_an_error(gpu):
    """3M continuous columns need more than the array holds: refused on the
    host, before any launch."""
    rows, n = 3000, 3_000_000
    per = n // rows
    inst = from_rows(rows, n, [list(range(i * per, (i + 1) * per)) for i in range(rows)],
                     [[1.0] * per for _ in range(rows)], [-1e20] * rows, [1e6] * rows, [0.0] * n, [1.0] * n,
                     [3] * n)
    p = I.Problem(inst)
    try:
        assert p.stats().run_state_bytes > SCRATCH_BYTES
        dl, du = I.DeviceArray(inst.lb), I.DeviceArray(inst.ub)
        with pytest.raises(RuntimeError, match="bytes of run state"):
            p.propagate_device_no_alloc(dl.ptr, du.ptr, I.MEGAKERNEL)
        dl.free()
        du.free()
    finally:
        p.close()
