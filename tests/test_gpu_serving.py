"""The serving loop on the GPU: each request's result is the same alone and
in the mix (bitwise: every reduction on the path has a fixed order and the
tiles share no device memory), and served requests report latencies."""
import numpy as np
import pytest

from ablsteadyrans_amd import serving as S

pytestmark = pytest.mark.gpu


def _requests():
    specs = [S.TileSpec(24, 20, 16, 50.0, hills=[(600.0, 500.0, 60.0, 300.0)], direction=30.0, dz0=0.0),
             S.TileSpec(16, 28, 20, 50.0, hills=[(400.0, 700.0, 40.0, 250.0), (200.0, 300.0, 30.0, 200.0)],
                        direction=200.0, dz0=0.0),
             S.TileSpec(20, 20, 24, 50.0, hills=[(500.0, 500.0, 80.0, 400.0)], direction=310.0, dz0=0.0)]
    reqs = [S.Request(0, "steady", specs[0], max_iterations=25),
            S.Request(1, "init", specs[1]),
            S.Request(2, "steady", specs[2], max_iterations=25, precision=0),
            S.Request(3, "steady", specs[1], max_iterations=25)]
    return reqs


def test_results_are_the_same_alone_and_in_the_mix(cuda):
    reqs = _requests()
    alone = [S.run_request(r, keep_fields=True) for r in reqs]
    mixed, wall = S.Server(slots=4).run(reqs, keep_fields=True)
    for a, m in zip(alone, mixed):
        assert a.error is None and m.error is None, (a.error, m.error)
        assert a.iterations == m.iterations
        for name in a.fields:
            assert np.array_equal(a.fields[name], m.fields[name]), (a.rid, name)
    s = S.summary(mixed, wall)
    assert s["errors"] == 0 and s["all"]["n"] == 4
    assert all(r.start >= r.arrival and r.finish > r.start for r in mixed)


def test_arrivals_are_waited_for_and_the_budget_serialises(cuda):
    reqs = _requests()[:2]
    reqs[1].arrival = 0.3
    res, wall = S.Server(slots=2).run(reqs)
    assert res[1].start >= res[1].arrival - 1e-3
    assert wall >= 0.3
    # a budget that fits one tile at a time runs them one after the other
    budget = max(r.spec.cells for r in reqs) * S.BYTES_PER_CELL
    reqs[1].arrival = 0.0
    res, _ = S.Server(slots=2, hbm_budget=budget).run(reqs)
    first, second = sorted(res, key=lambda r: r.start)
    assert second.start >= first.finish - 1e-3
