"""Grid sequencing's stop logic and input checks on the CPU (the library's
host code: gust_loose_stop, gust_check_sequencing_options), on synthetic
histories with the iterations test_Sequencing.cpp pins (92 and 137 for the
residual branch, 88 and 50 and 29 for the monitor branch, the reference
window, the released hold), and the checks' messages
(Tests/CMakeLists.txt:242-280 abort_sequencing_*)."""
import math

import pytest

from ablsteadyrans_amd import sequencing as SQ
from ablsteadyrans_amd.steady import MonitorSpec, SteadyCriterion, SteadyParams, SteadyStatus


def params(**kw):
    p = SteadyParams(residual_tol=1.0e-12, residual_drop_orders=4, residual_ref_iterations=5, monitor_rel_tol=1.0e-3,
                     monitor_window=50)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


NOISY = [lambda n: 1.1 if n % 2 == 0 else 0.9]
ONE_MON = [MonitorSpec("speed", 1.0)]


def first_loose(c, o, res, mon, n_max, hold=None):
    if hold is None:
        hold = c.params.residual_ref_iterations
    for n in range(n_max):
        c.add([f(n) for f in res], [f(n) for f in mon])
        stop, why = SQ.loose_stop(c, o, hold)
        if stop:
            return n, why
    return -1, ""


def test_loose_residual_branch_stops_at_the_pinned_iteration():
    p = params()
    o = SQ.SequencingOptions(enabled=True)
    res = [lambda n: 2.0 * 0.9 ** n, lambda n: 1.0 + n if n < 2 else 3.0 * 0.95 ** (n - 2), lambda n: 1.0e-13]
    c = SteadyCriterion(p, ["u", "k", "p"], ONE_MON)
    n, why = first_loose(c, o, res, NOISY, 1000)
    assert n == 92 and "orders below its reference" in why
    assert c.references()[1] == 3.0
    assert c.verdict().status == SteadyStatus.NotSteady
    o.drop_orders = 3
    c3 = SteadyCriterion(p, ["u", "k", "p"], ONE_MON)
    assert first_loose(c3, o, res, NOISY, 1000)[0] == 137


def test_loose_residual_branch_waits_for_the_reference():
    c = SteadyCriterion(params(), ["u"], ONE_MON)
    o = SQ.SequencingOptions(enabled=True)
    assert first_loose(c, o, [lambda n: 1.0 if n == 0 else 1.0e-6], NOISY, 100)[0] == 4


def test_loose_monitor_branch_stops_at_the_pinned_iteration():
    p = params()
    o = SQ.SequencingOptions(enabled=True)
    specs = [MonitorSpec("speedup", 1.0), MonitorSpec("ustar", 0.1)]
    res = [lambda n: 1.0]
    mon = [lambda n: 1 + 0.5 * 0.9 ** n, lambda n: 0.002 * 0.95 ** n]
    assert first_loose(SteadyCriterion(p, ["u"], specs), o, res, mon, 2000)[0] == -1   # off by default
    o.monitor_tol = 0.01
    n, why = first_loose(SteadyCriterion(p, ["u"], specs), o, res, mon, 2000)
    assert n == 88 and "every monitor within" in why
    flat = [lambda n: 1.0, lambda n: 0.0]
    assert first_loose(SteadyCriterion(p, ["u"], specs), o, res, flat, 200)[0] == 50
    q = params(monitor_window=10, residual_ref_iterations=30)
    assert first_loose(SteadyCriterion(q, ["u"], specs), o, res, flat, 200)[0] == 29


def test_non_finite_history_never_stops():
    p = params()
    o = SQ.SequencingOptions(enabled=True)
    res = [lambda n: 1.0 if n < 5 else 0.0, lambda n: math.inf if n == 7 else 0.0]
    mon = [lambda n: 1.0]
    assert first_loose(SteadyCriterion(p, ["u", "v"], ONE_MON), o, res, mon, 6)[0] == 5
    c2 = SteadyCriterion(p, ["u", "v"], ONE_MON)
    for n in range(200):
        c2.add([res[0](n), res[1](n)], [1.0])
        if n >= 7:
            assert not SQ.loose_stop(c2, o, p.residual_ref_iterations)[0], n
    assert c2.verdict().status == SteadyStatus.Diverged


def test_seeded_reference_is_the_drop_origin():
    p = params()
    o = SQ.SequencingOptions(enabled=True)
    res = [lambda n: 1.0e-2]
    own = SteadyCriterion(p, ["u"], ONE_MON)
    seeded = SteadyCriterion(p, ["u"], ONE_MON)
    seeded.seed_reference([1.0])
    assert first_loose(own, o, res, NOISY, 100)[0] == -1
    assert first_loose(seeded, o, res, NOISY, 100)[0] == 4
    assert seeded.references() == [1.0]
    raised = SteadyCriterion(p, ["u"], ONE_MON)
    raised.seed_reference([1.0])
    raised.add([5.0], [1.0])
    assert raised.references() == [5.0]


def test_released_hold_stops_from_the_first_iterations():
    p = params()
    o = SQ.SequencingOptions(enabled=True)
    res = [lambda n: 1.0e-2]

    def seeded(pp=p):
        c = SteadyCriterion(pp, ["u"], ONE_MON)
        c.seed_reference([1.0])
        return c

    assert first_loose(seeded(), o, res, NOISY, 100, hold=5)[0] == 4
    assert first_loose(seeded(), o, res, NOISY, 100, hold=0)[0] == 2
    assert first_loose(seeded(params(residual_ref_iterations=1)), o, res, NOISY, 100, hold=0)[0] == 0
    falling = [lambda n: 3.0e-2 if n < 3 else 1.0e-2]
    assert first_loose(seeded(), o, falling, NOISY, 100, hold=0)[0] == 3
    rising = [lambda n: 5.0 if n == 0 else (0.1 if n <= 2 else 4.0e-2)]
    assert first_loose(seeded(), o, rising, NOISY, 100, hold=0)[0] == 3
    om = SQ.SequencingOptions(enabled=True, monitor_tol=0.01)
    for hold in (0, 5):
        n, why = first_loose(seeded(), om, [lambda n: 1.0], [lambda n: 2.0], 100, hold=hold)
        assert n == 50 and "every monitor" in why, hold
    assert first_loose(seeded(params(monitor_window=2)), om, [lambda n: 1.0], [lambda n: 2.0], 100, hold=0)[0] == 4


def _has(err, s):
    return any(s in e for e in err)


def test_input_checks_name_their_keys():
    p = params()
    ok = SQ.SequencingOptions(enabled=True)
    assert SQ.check_sequencing_options(ok, p, 3, (32, 16)) == []
    assert SQ.check_sequencing_options(ok, p, 4, (32, 16)) == []
    assert _has(SQ.check_sequencing_options(ok, p, 5, (32, 16)), "at least 2 cells in x and in y on the coarsest level")
    for d in (0.0, -1.0, 4.5, math.nan):
        o = SQ.SequencingOptions(enabled=True, drop_orders=d)
        assert _has(SQ.check_sequencing_options(o, p, 3, (32, 16)), "steady.sequencing_drop_orders"), d
    assert SQ.check_sequencing_options(SQ.SequencingOptions(enabled=True, drop_orders=4), p, 3, (32, 16)) == []
    assert SQ.check_sequencing_options(SQ.SequencingOptions(enabled=True, monitor_tol=0), p, 3, (32, 16)) == []
    for t in (1.0, -0.1, math.nan, math.inf):
        o = SQ.SequencingOptions(enabled=True, monitor_tol=t)
        assert _has(SQ.check_sequencing_options(o, p, 3, (32, 16)), "steady.sequencing_monitor_tol"), t
    assert _has(SQ.check_sequencing_options(ok, p, 1, (32, 16)), "needs mesh.hierarchy_levels >= 2")
    for strict_orders in (1.0, 0.0):
        assert SQ.check_sequencing_options(SQ.SequencingOptions(), params(residual_drop_orders=strict_orders), 1,
                                           (32, 16)) == []
    off = SQ.SequencingOptions(given=tuple(SQ.GIVEN_BITS))
    err = SQ.check_sequencing_options(off, p, 1, (32, 16))
    for k in ("sequencing_drop_orders", "sequencing_monitor_tol", "sequencing_hold", "sequencing_cfl_start"):
        assert _has(err, f"steady.{k} has no effect unless steady.sequencing = 1"), k
    for hold in (0, 1):
        o = SQ.SequencingOptions(enabled=True, hold=hold)
        assert SQ.check_sequencing_options(o, p, 3, (32, 16)) == []
        o.cfl_previous = True
        got = _has(SQ.check_sequencing_options(o, p, 3, (32, 16)),
                   "steady.sequencing_cfl_start = previous needs steady.sequencing_hold = 0")
        assert got == (hold == 1)


def test_the_reference_abort_messages():
    """The texts Tests/CMakeLists.txt:245-258 expects (abort_sequencing_range,
    abort_sequencing_coarsest)."""
    p = params(residual_drop_orders=8)
    o = SQ.SequencingOptions(enabled=True, drop_orders=9, monitor_tol=1, hold=1, cfl_previous=True)
    err = SQ.check_sequencing_options(o, p, 1, (64, 64))
    assert _has(err, "steady.sequencing_drop_orders = 9.000e+00: must be > 0 and <= steady.residual_drop_orders = "
                     "8.000e+00")
    assert _has(err, "steady.sequencing_monitor_tol = 1.000e+00: must be 0 (off) or > 0 and < 1")
    assert _has(err, "steady.sequencing = 1 needs mesh.hierarchy_levels >= 2 (it is 1)")
    assert _has(err, "steady.sequencing_cfl_start = previous needs steady.sequencing_hold = 0")
    err = SQ.check_sequencing_options(SQ.SequencingOptions(enabled=True), params(), 5, (64, 16))
    assert _has(err, "mesh.n_cell x and y / 2^(mesh.hierarchy_levels - 1) = 4 1")


def test_the_library_refuses_a_null_criterion():
    import ctypes as C
    from ablsteadyrans_amd import _lib
    o = SQ.SequencingOptions(enabled=True).c()
    stop = C.c_int32(0)
    with pytest.raises(_lib.GustError, match="null"):
        _lib.check(_lib.lib.gust_loose_stop(None, C.byref(o), 5, C.byref(stop), None, 0))
