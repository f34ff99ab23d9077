"""The steady criterion (host C++ in libgust_amd.so) on the synthetic
histories of Tests/Unit_Tests/test_SteadyCriterion.cpp: the iterations the
reference pins there (136, 38, 49, 50, 4602; from an independent
implementation of the definition), the message texts, divergence, the window
boundary in exact binary arithmetic, parameter checks, restart, and the
driver's r_n and CFL steps (Source/Driver.cpp:325-334, 421-437). Host only:
no device call is made."""
from __future__ import annotations

import math

import pytest

from ablsteadyrans_amd import _lib
from ablsteadyrans_amd.steady import MonitorSpec, SteadyCriterion, SteadyParams, SteadyStatus, validate


def mentions(v, s):
    return any(s in r for r in v.reasons)


def first_steady(c, res, mon, n_max):
    for n in range(n_max):
        c.add([f(n) for f in res], [f(n) for f in mon])
        if c.verdict().status == SteadyStatus.Steady:
            return n
    return -1


CONV_RES = [lambda n: 2.0 * 0.8 ** n, lambda n: 0.5 * 0.85 ** n]
CONV_MON = [lambda n: 0.3 + 0.05 * 0.9 ** n, lambda n: 0.4 * (1 - 0.5 * 0.93 ** n)]
CONV_SPEC = [MonitorSpec("speedup", 1.0), MonitorSpec("ustar", 1.0e-6)]


def test_converging_history_is_steady_at_the_pinned_iteration():
    # test_SteadyCriterion.cpp:54-60
    c = SteadyCriterion(SteadyParams(), ["u", "p"], CONV_SPEC)
    assert first_steady(c, CONV_RES, CONV_MON, 1000) == 136


def test_stalling_history_is_never_steady():
    # test_SteadyCriterion.cpp:62-74
    c = SteadyCriterion(SteadyParams(), ["u"], [("speedup", 1.0)])
    assert first_steady(c, [lambda n: max(1.0e-3, 0.8 ** n)], [lambda n: 0.3], 2000) == -1
    v = c.verdict()
    assert v.status == SteadyStatus.NotSteady
    assert mentions(v, "residual 'u' = 0.001 is not below steady.residual_tol")
    assert mentions(v, "dropped 3 orders")


def test_residual_at_the_tolerance_is_not_below_it():
    # test_SteadyCriterion.cpp:76-89
    p = SteadyParams()
    at = SteadyCriterion(p, ["u"], [])
    at.add([1.0], [])
    for _ in range(10):
        at.add([p.residual_tol], [])
    assert at.verdict().status == SteadyStatus.NotSteady
    below = SteadyCriterion(p, ["u"], [])
    below.add([1.0], [])
    below.add([0.98 * p.residual_tol], [])
    assert below.verdict().status == SteadyStatus.Steady


def test_oscillating_monitor_is_never_steady():
    # test_SteadyCriterion.cpp:91-101
    c = SteadyCriterion(SteadyParams(), ["u"], [("speedup", 1.0)])
    mon = [lambda n: 1.0 + 0.005 * math.sin(2.0 * math.pi * n / 37.0)]
    assert first_steady(c, [lambda n: 0.8 ** n], mon, 2000) == -1
    v = c.verdict()
    assert len(v.reasons) == 1
    assert mentions(v, "monitor 'speedup' changed by")


def test_oscillating_residual_is_never_steady():
    c = SteadyCriterion(SteadyParams(), ["u"], [])
    assert first_steady(c, [lambda n: 1.0e-3 * (1.5 + math.sin(0.3 * n))], [], 2000) == -1


def test_non_finite_values_mean_diverged():
    # test_SteadyCriterion.cpp:112-130
    c = SteadyCriterion(SteadyParams(), ["u", "k"], [("speedup", 1.0)])
    c.add([1.0, 1.0], [0.3])
    c.add([0.5, math.nan], [0.3])
    v = c.verdict()
    assert v.status == SteadyStatus.Diverged
    assert mentions(v, "residual 'k' became NaN or infinite")
    for _ in range(200):
        c.add([0.0, 0.0], [0.3])
    assert c.verdict().status == SteadyStatus.Diverged
    d = SteadyCriterion(SteadyParams(), ["u"], [("ustar", 1.0)])
    d.add([1.0], [math.inf])
    assert d.verdict().status == SteadyStatus.Diverged
    assert mentions(d.verdict(), "monitor 'ustar'")


def test_near_zero_monitor_uses_its_scale():
    # test_SteadyCriterion.cpp:132-142
    res = [lambda n: 0.0]
    mon = [lambda n: (1.0 if n % 2 else -1.0) * 1.0e-17]
    assert first_steady(SteadyCriterion(SteadyParams(), ["u"], [("speedup", 1.0)]), res, mon, 200) == 50
    assert first_steady(SteadyCriterion(SteadyParams(), ["u"], [("speedup", 1.0e-20)]), res, mon, 200) == -1


def test_window_needs_all_its_values_and_the_exact_edge():
    # test_SteadyCriterion.cpp:144-171 (binary-exact values: tolerance 2^-10, scale 1)
    c = SteadyCriterion(SteadyParams(), ["u"], [("m", 1.0)])
    for _ in range(50):
        c.add([0.0], [1.0])
    v = c.verdict()
    assert v.iteration == 49 and v.status == SteadyStatus.NotSteady
    assert mentions(v, "50 of the 51 values")
    c.add([0.0], [1.0])
    assert c.verdict().status == SteadyStatus.Steady
    p = SteadyParams(monitor_window=2, monitor_rel_tol=1.0 / 1024.0)
    at = 1.0 + 1.0 / 1024.0
    above = at + 1.0 / (1024.0 * 1024.0)
    edge = SteadyCriterion(p, ["u"], [("m", 1.0)])
    for x in (at, 1.0, 1.0):
        edge.add([0.0], [x])
    assert edge.verdict().status == SteadyStatus.Steady
    over = SteadyCriterion(p, ["u"], [("m", 1.0)])
    for x in (above, 1.0, 1.0):
        over.add([0.0], [x])
    assert over.verdict().status == SteadyStatus.NotSteady


def test_drop_is_measured_from_the_early_peak():
    # test_SteadyCriterion.cpp:173-186
    res = [lambda n: 1.0e-2 * 10.0 ** n if n < 3 else 10.0 * 0.7 ** (n - 3)]
    assert first_steady(SteadyCriterion(SteadyParams(), ["u"], []), res, [], 500) == 38
    assert first_steady(SteadyCriterion(SteadyParams(residual_ref_iterations=1), ["u"], []), res, [], 500) == 49


@pytest.mark.parametrize("key,change", [
    ("steady.residual_tol", dict(residual_tol=0.0)),
    ("steady.residual_drop_orders", dict(residual_drop_orders=-1.0)),
    ("steady.residual_ref_iterations", dict(residual_ref_iterations=0)),
    ("steady.monitor_rel_tol", dict(monitor_rel_tol=math.nan)),
    ("steady.monitor_window", dict(monitor_window=1)),
    ("steady.residual_drop_orders", dict(residual_drop_orders=308 + 1.0)),
])
def test_parameter_checks_name_their_keys(key, change):
    # test_SteadyCriterion.cpp:188-219
    assert validate(SteadyParams()) == []
    err = validate(SteadyParams(**change))
    assert len(err) == 1 and key in err[0]


def test_constructor_and_add_refusals():
    with pytest.raises(_lib.GustError, match="scale must be finite and > 0"):
        SteadyCriterion(SteadyParams(), ["u"], [("m", 0.0)])
    with pytest.raises(_lib.GustError, match="at least one residual"):
        SteadyCriterion(SteadyParams(), [], [])
    with pytest.raises(_lib.GustError, match="invalid steady criterion"):
        SteadyCriterion(SteadyParams(monitor_window=1), ["u"], [])
    c = SteadyCriterion(SteadyParams(), ["u", "p"], [])
    with pytest.raises(_lib.GustError, match="got 1 residuals and 0 monitors, expected 2 and 0"):
        c.add([1.0], [])


def test_slow_monitor_stops_near_its_limit():
    # test_SteadyCriterion.cpp:221-232
    A, r = 0.1, 0.999
    c = SteadyCriterion(SteadyParams(), ["u"], [("ustar", 1.0)])
    n = first_steady(c, [lambda k: 0.5 ** k], [lambda k: 1.0 + A * r ** k], 10000)
    assert n == 4602
    assert A * r ** n <= 1.0e-3 * (1.0 + A * r ** n)


@pytest.mark.parametrize("period", [37.0, 100.0, 200.0])
def test_oscillations_up_to_four_windows_are_never_steady(period):
    c = SteadyCriterion(SteadyParams(), ["u"], [("speedup", 1.0)])
    mon = [lambda k: 1.0 + 0.01 * math.sin(2.0 * math.pi * k / period)]
    assert first_steady(c, [lambda k: 0.5 ** k], mon, 5000) == -1


def test_zero_residual_reference_does_not_block():
    c = SteadyCriterion(SteadyParams(), ["theta"], [("m", 1.0)])
    assert first_steady(c, [lambda k: 0.0 if k < 5 else 1.0e-16], [lambda k: 1.0], 500) == 50


def test_restored_history_gives_the_same_verdicts():
    # test_SteadyCriterion.cpp:258-291
    whole = SteadyCriterion(SteadyParams(), ["u", "p"], CONV_SPEC)
    before = SteadyCriterion(SteadyParams(), ["u", "p"], CONV_SPEC)

    def feed(c, n):
        c.add([f(n) for f in CONV_RES], [f(n) for f in CONV_MON])

    for n in range(70):
        feed(whole, n)
        feed(before, n)
    saved = before.history()
    after = SteadyCriterion(SteadyParams(), ["u", "p"], CONV_SPEC)
    after.restore(saved)
    assert after.iterations() == 70
    for n in range(70, 300):
        feed(whole, n)
        feed(after, n)
        a, b = whole.verdict(), after.verdict()
        assert (a.status, a.iteration, a.reasons) == (b.status, b.iteration, b.reasons), n
    other = SteadyCriterion(SteadyParams(), ["u"], CONV_SPEC)
    with pytest.raises(_lib.GustError, match="does not match"):
        other.restore(saved)
    with pytest.raises(_lib.GustError, match="does not match"):
        after.restore(saved[:-3])


def test_slowly_growing_monitor_is_not_steady():
    c = SteadyCriterion(SteadyParams(), ["u"], [("k", 1.0)])
    assert first_steady(c, [lambda k: 0.0], [lambda k: 1.0 + 1.0e-6 * 1.05 ** k], 120) == -1
    assert mentions(c.verdict(), "is not converging over the window")


def test_thresholds_and_seeded_reference():
    p = SteadyParams(residual_drop_orders=6)
    c = SteadyCriterion(p, ["u", "p"], [])
    c.seed_reference([3.0, 0.0])
    c.add([1.0, 2.0e-6], [])
    # min(tol, max(ref, tol) 10^-6): ref 3 -> 3e-6; ref 2e-6 floored at tol 5e-5 -> 5e-11
    t = c.residual_thresholds()
    assert t[0] == min(5e-5, 3.0 * 1e-6) and t[1] == min(5e-5, 5e-5 * 1e-6)
    with pytest.raises(_lib.GustError, match="before the first iteration"):
        c.seed_reference([1.0, 1.0])
    d = SteadyCriterion(p, ["u"], [])
    with pytest.raises(_lib.GustError, match="one finite reference >= 0"):
        d.seed_reference([-1.0])


def test_driver_relative_residual_and_cfl():
    """Source/Driver.cpp:325-334 and 421-437: r_n = 1 during the hold, then
    max |R_q| / max(ref_q, tol); cfl0 during the hold, then the SER law."""
    c = SteadyCriterion(SteadyParams(), ["u", "p"], [])
    hold = 5
    res = [[2.0, 1.0], [4.0, 0.5], [1.0, 3.0], [0.5, 0.2], [0.25, 0.1], [0.1, 0.3]]
    for n, r in enumerate(res):
        c.add(r, [])
        rn = c.relative_residual(r, hold)
        if n + 1 <= hold:
            assert rn == 1.0
            assert c.next_cfl(hold, 2.0, rn, 1.0, 10.0, 1.0, 2.0) == 1.0
        else:
            # references: max of the first five values, 4 and 3
            assert rn == max(0.1 / 4.0, 0.3 / 3.0)
            # SER: min(cfl_max, growth cfl, cfl0 r^-p) = min(10, 4, 1/0.1) = 4
            assert c.next_cfl(hold, 2.0, rn, 1.0, 10.0, 1.0, 2.0) == 4.0
    c.add([math.nan, 1.0], [])
    assert c.relative_residual([math.nan, 1.0], hold, previous=0.125) == 0.125
