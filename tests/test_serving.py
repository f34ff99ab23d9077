"""Host logic of the serving loop: the tile distribution, latency summary and
HBM admission refusal (no device work)."""
import numpy as np
import pytest

from ablsteadyrans_amd import serving as S


def test_random_tiles_follow_the_batched_distribution():
    """nx, ny in 32..128, nz in 32..80, dx 20 m, levels from 1 m by 1.08;
    1-5 hills, H in 20-300 m (scaled down together to half the top when
    their sum would exceed it), L in 100-600 m; direction in 0-360."""
    rng = np.random.default_rng(1)
    scaled = 0
    for _ in range(200):
        t = S.random_tile(rng)
        assert 32 <= t.nx <= 128 and 32 <= t.ny <= 128 and 32 <= t.nz <= 80
        assert t.dx == 20.0 and t.dz0 == 1.0 and t.stretch == 1.08
        zl = t.zlevels()
        assert zl[1] == 1.0 and abs(zl[-1] - (1.08 ** t.nz - 1) / 0.08) <= 1e-9 * zl[-1]
        assert 1 <= len(t.hills) <= 5
        total = sum(h[2] for h in t.hills)
        assert total <= 0.5 * zl[-1] * (1 + 1e-12)
        if total < 0.5 * zl[-1] * (1 - 1e-12):   # not scaled: the drawn heights
            assert all(h[2] >= 20.0 for h in t.hills), t.hills
        else:
            scaled += 1
        for (xc, yc, H, L) in t.hills:
            assert H <= 300.0 and H <= S.MAX_ASPECT * L * (1 + 1e-12) and 100.0 <= L <= 600.0
            assert 0 <= xc <= t.nx * t.dx and 0 <= yc <= t.ny * t.dx
        assert 0.0 <= t.direction < 360.0
    assert 0 < scaled < 200


def test_poisson_arrivals_have_the_rate():
    rng = np.random.default_rng(2)
    a = S.poisson_arrivals(20000, 20.0, rng)
    assert np.all(np.diff(a) > 0)
    assert abs(a[-1] / 20000 - 1 / 20.0) < 0.002


def test_summary_percentiles_and_counts():
    res = []
    for i in range(100):
        r = S.Result(i, "init" if i % 4 else "steady", 1000, 0.0, 0.0, 0.01 * (i + 1), iterations=10, status=1)
        res.append(r)
    res.append(S.Result(100, "init", 10, 0.0, 0.0, 5.0, error="boom"))
    s = S.summary(res, 2.0)
    assert s["requests"] == 101 and s["errors"] == 1
    lat = [0.01 * (i + 1) for i in range(100)]
    assert s["all"]["p50_s"] == pytest.approx(np.percentile(lat, 50), abs=1e-4)
    assert s["all"]["p99_s"] == pytest.approx(np.percentile(lat, 99), abs=1e-4)
    assert s["steady"]["n"] == 25 and s["steady"]["steady"] == 25
    assert s["init"]["n"] == 75


def test_server_refuses_a_request_larger_than_the_budget():
    req = S.Request(0, "init", S.TileSpec(128, 128, 80))
    with pytest.raises(ValueError, match="whole HBM budget"):
        S.Server(slots=2, hbm_budget=1e6).run([req])
    with pytest.raises(ValueError, match="slots"):
        S.Server(slots=0)
