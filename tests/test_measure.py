"""Progress measure (src/progressMeasureMethods/) without a GPU: the
oracle against hand-derived known answers, the library's weakest-bounds
search and cpu_seq measure against the oracle, the printed format of the
sequential measure entry point, and its error path.

Known answers (tests/golden/progress_measure.json, each derived by hand):
the reference's own weakest-bounds example, the same example without its
first row (where the search reaches the reference test's analytic value),
a chain whose upper bound moves twice after becoming finite (a fractional
P_fin), and an input on which the reference's search loses a weakest bound
(the reference exits there)."""
import ctypes
import json
import os

import numpy as np
import pytest

import fixtures
from gdp_amd import interface as I
from gdp_amd.instances import INF, Instance, small_mip

MEASURES = json.load(open(os.path.join(fixtures.GOLDEN, "progress_measure.json")))
KNOWN = ("example1", "example1_without_first_row", "chain")


def measure_instance(name) -> Instance:
    d = MEASURES[name]
    return Instance(d["m"], d["n"], np.array(d["col"], np.int32), np.array(d["row_ptr"], np.int32),
                    np.array(d["vals"], np.float64), np.array(d["lhs"], np.float64), np.array(d["rhs"], np.float64),
                    np.array(d["lb"], np.float64), np.array(d["ub"], np.float64), np.array(d["vartype"], np.int32),
                    name)


def leq_mip(seed, m=300, n=150):
    """small_mip with every row one-sided (<=): inputs on which the
    reference's weakest-bounds search keeps every finite limit bound finite."""
    inst = small_mip(seed, m=m, n=n)
    inst.lhs[:] = -INF
    return inst


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def _check_series(rows, expected, fin_tol=0.0):
    assert len(rows) == len(expected)
    for (r, f, k), (er, ef, ek) in zip(rows, expected):
        assert r == er
        assert abs(f - ef) <= fin_tol
        assert abs(k - ek) <= 0.005  # P_inf as printed with two decimals


@pytest.mark.parametrize("name", KNOWN)
def test_oracle_weakest_bounds_known_answers(oracle, name):
    d = MEASURES[name]
    wl, wu = oracle.weakest_bounds(measure_instance(name))
    np.testing.assert_array_equal(wl, d["weakest_lb"])
    np.testing.assert_array_equal(wu, d["weakest_ub"])


def test_reference_example_and_its_analytic_value(oracle):
    # The reference's test expects ub0 = 5 for its example; its search
    # expands every matrix entry once, so the path to 5 is cut (derivation in
    # the fixture).  Without the example's first row the same chain is the
    # only source of ub0 and the search reaches 5.
    d = MEASURES["example1"]
    _, wu = oracle.weakest_bounds(measure_instance("example1"))
    assert wu[0] == 1.0 and d["reference_test_ub"][0] == 5.0
    _, wu = oracle.weakest_bounds(measure_instance("example1_without_first_row"))
    assert wu[0] == d["reference_test_ub"][0]


@pytest.mark.parametrize("name", KNOWN)
@pytest.mark.parametrize("driver", ["seq", "jacobi"])
def test_oracle_measure_known_answers(oracle, name, driver):
    rows, lb, ub = oracle.progress_series(measure_instance(name), driver)
    _check_series(rows, MEASURES[name]["series"], fin_tol=1e-12)


def test_oracle_measure_defect(oracle):
    rows, var = oracle.progress_series(measure_instance("defect"), "seq")
    assert rows is None and var == 0


@pytest.mark.parametrize("name", KNOWN)
def test_library_known_answers(gdp, name):
    d = MEASURES[name]
    inst = measure_instance(name)
    wl, wu = I.weakest_bounds(inst)
    np.testing.assert_array_equal(wl, d["weakest_lb"])
    np.testing.assert_array_equal(wu, d["weakest_ub"])
    out, pf, pi = I.propagate_with_measure(inst, I.MEASURE_CPU_SEQ)
    _check_series(list(zip(range(1, len(pf) + 1), pf, pi)), d["series"], fin_tol=1e-12)


def _legacy(name, inst):
    lib = I.load_library()
    col, rp, val = (np.ascontiguousarray(inst.col, np.int32), np.ascontiguousarray(inst.row_ptr, np.int32),
                    np.ascontiguousarray(inst.val, np.float64))
    lhs, rhs = np.ascontiguousarray(inst.lhs, np.float64), np.ascontiguousarray(inst.rhs, np.float64)
    lb, ub = np.array(inst.lb, np.float64), np.array(inst.ub, np.float64)
    vt = np.ascontiguousarray(inst.vartype, np.int32)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = getattr(lib, name)(inst.m, inst.n, len(col), P(col), P(rp), P(val), P(lhs), P(rhs), P(lb), P(ub), P(vt))
    return rc, lb, ub


def test_sequential_entry_prints_the_measure(gdp, oracle, capfd):
    d = MEASURES["chain"]
    inst = measure_instance("chain")
    capfd.readouterr()
    rc, lb, ub = _legacy("sequentialPropagateWithMeasureDouble", inst)
    out = capfd.readouterr().out
    assert rc == I.GDP_OKAY
    assert [ln for ln in out.splitlines() if ln.startswith("round=")] == d["lines"]
    assert "=== cpu_seq execution with measure of progress ===" in out
    el, eu, _, _, _ = oracle.seq(inst)
    assert np.array_equal(_bits(lb), _bits(el)) and np.array_equal(_bits(ub), _bits(eu))
    np.testing.assert_array_equal(ub, d["limit_ub"])


def test_defect_returns_error(gdp, capfd):
    inst = measure_instance("defect")
    capfd.readouterr()
    for name in ("sequentialPropagateWithMeasureDouble",):
        rc, lb, ub = _legacy(name, inst)
        assert rc == I.GDP_ERROR
        assert "varidx: 0" in capfd.readouterr().out
        np.testing.assert_array_equal(lb, inst.lb)
        np.testing.assert_array_equal(ub, inst.ub)
    with pytest.raises(RuntimeError, match="infinite weakest upper bound"):
        I.propagate_with_measure(inst, I.MEASURE_CPU_SEQ)


def test_empty_problem(gdp):
    inst = Instance(0, 0, np.zeros(0, np.int32), np.zeros(1, np.int32), np.zeros(0), np.zeros(0), np.zeros(0),
                    np.zeros(0), np.zeros(0), np.zeros(0, np.int32), "empty")
    rc, _, _ = _legacy("sequentialPropagateWithMeasureDouble", inst)
    assert rc == I.GDP_OKAY


def test_bad_arguments(gdp):
    inst = measure_instance("chain")
    with pytest.raises(RuntimeError):
        I.propagate_with_measure(inst, method=7)
    with pytest.raises(RuntimeError):
        I.propagate_with_measure(inst, max_rounds=0)


@pytest.mark.parametrize("seed", range(12))
def test_weakest_bounds_match_oracle(gdp, oracle, seed):
    # both-sided rows: many candidates, including the ones that reset a
    # weakest bound (the defect above)
    inst = small_mip(seed, m=200, n=100)
    wl, wu = I.weakest_bounds(inst)
    ol, ou = oracle.weakest_bounds(inst)
    assert np.array_equal(_bits(wl), _bits(ol)) and np.array_equal(_bits(wu), _bits(ou))
    # finite input bounds are kept (checkWeakestBoundsResult, weakest_bounds.h:62-96)
    fl, fu = inst.lb > -INF + 1e-9, inst.ub < INF - 1e-9
    np.testing.assert_array_equal(wl[fl], inst.lb[fl])
    np.testing.assert_array_equal(wu[fu], inst.ub[fu])


@pytest.mark.parametrize("seed", range(12))
def test_seq_measure_matches_oracle(gdp, oracle, seed):
    inst = leq_mip(seed)
    rows, el, eu = oracle.progress_series(inst, "seq")
    assert rows is not None
    out, pf, pi = I.propagate_with_measure(inst, I.MEASURE_CPU_SEQ)
    # same sums in the same order: identical values
    np.testing.assert_array_equal(pf, [r[1] for r in rows])
    np.testing.assert_array_equal(pi, [r[2] for r in rows])
    assert np.array_equal(_bits(out.lbs), _bits(el)) and np.array_equal(_bits(out.ubs), _bits(eu))
    # properties the measure has by construction: bounds only tighten, so
    # both measures never decrease; at the fixpoint every bound is its limit
    assert np.all(np.diff(pf) >= 0) and np.all(np.diff(pi) >= 0)
    assert np.all((pf >= 0) & (pf <= 100)) and np.all((pi >= 0) & (pi <= 100))
    assert pf[-1] == 100.0 and pi[-1] == 100.0
    # the weakest bounds never lie inside the limit box
    wl, wu = I.weakest_bounds(inst)
    ll, lu, _, _, _ = oracle.seq(inst)
    fl, fu = ll > -INF + 1e-9, lu < INF - 1e-9
    assert np.all(wl[fl] <= ll[fl] + 1e-9) and np.all(wu[fu] >= lu[fu] - 1e-9)


@pytest.mark.parametrize("name", ["b1c1s1", "gen-ip036", "drayage-25-27", "p0201"])
def test_fixtures_match_oracle(gdp, oracle, name):
    inst = gdp.read_mps(os.path.join(fixtures.GOLDEN, "mps", name + ".mps.gz"))
    wl, wu = I.weakest_bounds(inst)
    ol, ou = oracle.weakest_bounds(inst)
    assert np.array_equal(_bits(wl), _bits(ol)) and np.array_equal(_bits(wu), _bits(ou))
    res = oracle.progress_series(inst, "seq")
    if res[0] is None:
        # b1c1s1: a finite limit upper bound whose weakest bound is lost
        with pytest.raises(RuntimeError, match="infinite weakest upper bound"):
            I.propagate_with_measure(inst, I.MEASURE_CPU_SEQ)
        return
    rows = res[0]
    _, pf, pi = I.propagate_with_measure(inst, I.MEASURE_CPU_SEQ)
    np.testing.assert_array_equal(pf, [r[1] for r in rows])
    np.testing.assert_array_equal(pi, [r[2] for r in rows])
