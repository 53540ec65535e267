"""The C ABI of libGpuProp.so without a GPU: exported symbols, argument
checking, and the library's own host propagators against the oracle.

The host cpu_seq of the library and the oracle's cpu_seq are independent
restatements of the reference's sequential propagator, so they must agree
bit for bit (bounds and rounds)."""
import ctypes
import subprocess

import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import INF, from_rows, small_mip


def _declared_functions(header):
    import re
    import os
    text = open(os.path.join(fixtures.GOLDEN, "..", "..", "include", header)).read()
    return set(re.findall(r"^(?:GDP_RETCODE|void|const char \*)\s*([A-Za-z_]\w*)\s*\(", text, re.M))


def test_exports_every_declared_symbol(gdp):
    declared = _declared_functions("interface.h") | _declared_functions("gdp_ext.h")
    assert set(I.LEGACY_SYMBOLS) == _declared_functions("interface.h")
    assert set(I.EXT_SYMBOLS) == _declared_functions("gdp_ext.h")
    out = subprocess.run(["nm", "-D", "--defined-only", I.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    assert declared <= exported, declared - exported
    lib = gdp.load_library()
    for name in declared:
        assert hasattr(lib, name)
    # nothing else with C linkage: internal helpers would collide with a
    # client's symbols of the same name in a drop-in library
    unmangled = {s for s in exported if not s.startswith("_Z")}
    stray = {s for s in unmangled if s not in I.LEGACY_SYMBOLS and not s.startswith("gdp")}
    assert not stray, stray
    assert unmangled == declared, unmangled ^ declared


def test_library_has_gfx950_code(gdp):
    # the embedded device code object names its target
    data = open(I.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def _args(inst, dtype=np.float64):
    return (np.ascontiguousarray(inst.col, np.int32), np.ascontiguousarray(inst.row_ptr, np.int32),
            np.ascontiguousarray(inst.val, dtype), np.ascontiguousarray(inst.lhs, dtype),
            np.ascontiguousarray(inst.rhs, dtype), np.array(inst.lb, dtype), np.array(inst.ub, dtype),
            np.ascontiguousarray(inst.vartype, np.int32))


def _call(gdp, name, inst, *extra, dtype=np.float64):
    lib = gdp.load_library()
    col, rp, val, lhs, rhs, lb, ub, vt = _args(inst, dtype)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = getattr(lib, name)(inst.m, inst.n, len(col), P(col), P(rp), P(val), P(lhs), P(rhs), P(lb), P(ub), P(vt),
                            *extra)
    return rc, lb, ub


def test_zero_size_problem_is_ok(gdp):
    inst = from_rows(0, 0, [], [], [], [], [], [], [])
    for name in ("propagateConstraintsSequentialDouble", "propagateConstraintsFullOMPDouble",
                 "propagateConstraintsGPUReductionDouble", "propagateConstraintsSequentialDisjointDouble"):
        rc, _, _ = _call(gdp, name, inst)
        assert rc == gdp.GDP_OKAY
    rc, _, _ = _call(gdp, "propagateConstraintsGPUAtomicDouble", inst, 0)
    assert rc == gdp.GDP_OKAY


def test_unknown_sync_type_is_an_error(gdp):
    inst = fixtures.analytic("achterberg")
    for st in (3, -1, 99):
        rc, lb, ub = _call(gdp, "propagateConstraintsGPUAtomicDouble", inst, st)
        assert rc == gdp.GDP_ERROR
        np.testing.assert_array_equal(lb, inst.lb)  # untouched
    rc, _, _ = _call(gdp, "propagateConstraintsGPUAtomicFloat", inst, 7, dtype=np.float32)
    assert rc == gdp.GDP_ERROR


@pytest.mark.parametrize("name", fixtures.ANALYTIC)
def test_sequential_measure_entry_reaches_the_analytic_fixpoint(gdp, name):
    inst = fixtures.analytic(name)
    rc, lb, ub = _call(gdp, "sequentialPropagateWithMeasureDouble", inst)
    assert rc == gdp.GDP_OKAY
    np.testing.assert_allclose(lb, inst.meta["lb_expected"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(ub, inst.meta["ub_expected"], rtol=0, atol=1e-4)


def _instances():
    out = [fixtures.analytic(n) for n in fixtures.ANALYTIC]
    out += [small_mip(seed=s) for s in range(4)]
    out += [fixtures.varied(s) for s in range(3)]
    out += [fixtures.generalized_cfp(25), fixtures.infeasible_instance()]
    return out


@pytest.mark.parametrize("idx", range(12))
def test_host_seq_matches_oracle_bitwise(gdp, oracle, idx):
    inst = _instances()[idx]
    exp = oracle.seq(inst)
    got = I.host_propagate(inst, "seq")
    np.testing.assert_array_equal(got.lbs, exp[0])
    np.testing.assert_array_equal(got.ubs, exp[1])
    assert (got.rounds, got.infeasible, got.capped) == (exp[2], exp[3], exp[4])
    # the legacy entry point gives the same bounds
    rc, lb, ub = _call(gdp, "propagateConstraintsSequentialDouble", inst)
    assert rc == gdp.GDP_OKAY
    np.testing.assert_array_equal(lb, exp[0])
    np.testing.assert_array_equal(ub, exp[1])


@pytest.mark.parametrize("idx", range(12))
def test_host_seq_disjoint_matches_oracle_bitwise(gdp, oracle, idx):
    inst = _instances()[idx]
    if inst.name == "infeasible":
        return
    exp = oracle.seq_disjoint(inst)
    rc, lb, ub = _call(gdp, "propagateConstraintsSequentialDisjointDouble", inst)
    assert rc == gdp.GDP_OKAY
    np.testing.assert_array_equal(lb, exp[0])
    np.testing.assert_array_equal(ub, exp[1])


@pytest.mark.parametrize("threads", [1, 4])
def test_host_omp_fixpoint(gdp, oracle, threads):
    for name in fixtures.ANALYTIC:
        inst = fixtures.analytic(name)
        got = I.host_propagate(inst, "omp", threads=threads)
        np.testing.assert_allclose(got.lbs, inst.meta["lb_expected"], atol=1e-4, rtol=0)
        np.testing.assert_allclose(got.ubs, inst.meta["ub_expected"], atol=1e-4, rtol=0)
    inst = small_mip(seed=5, int_frac=1.0)
    exp = oracle.seq(inst)
    got = I.host_propagate(inst, "omp", threads=threads)
    np.testing.assert_array_equal(got.lbs, exp[0])
    np.testing.assert_array_equal(got.ubs, exp[1])
    rc, lb, ub = _call(gdp, "propagateConstraintsFullOMPDouble", inst)
    assert rc == gdp.GDP_OKAY
    np.testing.assert_array_equal(lb, exp[0])


def test_float_entry_points_round_outward(gdp, oracle):
    inst = fixtures.analytic("two_lines")
    exp = oracle.seq(inst)
    for name in ("propagateConstraintsSequentialFloat", "propagateConstraintsFullOMPFloat"):
        rc, lb, ub = _call(gdp, name, inst, dtype=np.float32)
        assert rc == gdp.GDP_OKAY
        assert np.all(lb.astype(np.float64) <= exp[0]) and np.all(ub.astype(np.float64) >= exp[1])
        np.testing.assert_allclose(lb, exp[0], rtol=1e-6)
        np.testing.assert_allclose(ub, exp[1], rtol=1e-6)


def test_reference_style_list_interface(gdp, oracle):
    inst = fixtures.analytic("achterberg")
    args = (inst.n, inst.m, inst.nnz, inst.col.tolist(), inst.row_ptr.tolist(), inst.val.tolist(),
            inst.lhs.tolist(), inst.rhs.tolist(), inst.lb.tolist(), inst.ub.tolist(), inst.vartype.tolist())
    lbs, ubs = I.propagateSequential(*args)
    assert ubs == inst.meta["ub_expected"].tolist()
    lbs, ubs = I.propagateFullOMP(*args)
    assert ubs == inst.meta["ub_expected"].tolist()
    lbs, ubs = I.propagateSequentialDisjoint(*args)
    assert ubs == inst.meta["ub_expected"].tolist()
    with pytest.raises(AssertionError):
        I.propagateSequential(inst.n + 1, *args[1:])


def test_ieee_infinity_same_as_1e20(gdp, oracle):
    for name in ("savelsbergh1", "savelsbergh_cflp"):
        inst = fixtures.analytic(name)
        a = I.host_propagate(inst, "seq")
        ieee = inst.with_bounds(np.where(inst.lb <= -INF, -np.inf, inst.lb), np.where(inst.ub >= INF, np.inf, inst.ub))
        ieee.lhs = np.where(inst.lhs <= -INF, -np.inf, inst.lhs)
        ieee.rhs = np.where(inst.rhs >= INF, np.inf, inst.rhs)
        b = I.host_propagate(ieee, "seq")
        fin = a.ubs < INF
        np.testing.assert_array_equal(a.ubs[fin], b.ubs[fin])
        np.testing.assert_array_equal(a.lbs, b.lbs)
        assert np.all(np.isinf(b.ubs[~fin]))
        assert a.rounds == b.rounds
