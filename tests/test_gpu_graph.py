"""One SIMPLEC iteration captured as a HIP graph (gust_simplec_capture /
gust_simplec_replay, gust_steady_solve_graphed; the pressure solve in its
fixed-step device form):

* replays equal eager iterations of the same form bit for bit, iteration
  by iteration (every field and the stage norms), with eager work and waits
  between the replays (a replay's slots start at their initial values every
  time), on the column deck, the 3-D hill deck and an open-channel hill tile;
* the graphed driver reaches the deck's verdict in the same iterations and
  to the same bits as the eager driver with the same fixed pressure steps;
  with the steps learned from the first iteration, in the eager driver's
  iterations (within 5 %);
* ms per iteration, graphed and eager, on the column and 3-D hill decks;
* too few fixed pressure steps are an error, with the solver's message.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "p", "theta", "tke")


def _fields(deck):
    return {n: deck.state[n].numpy()[0, 2:-2, 2:-2, 2:-2].copy() for n in NAMES}


def _deck(name):
    from ablsteadyrans_amd import decks as D
    from ablsteadyrans_amd import mesh as M
    if name == "column":
        return D.prepare(D.neutral_column())
    if name == "hill3d":
        return D.prepare(D.neutral_hill_3d())
    return D.prepare(D.open_abl("hill_40x32x24", (40, 32, 24), (800.0, 640.0), 2.0,
                                M.Terrain(kind="hills", hills=[(400.0, 320.0, 40.0, 120.0)]), sweeps=4, cfl_max=4.0,
                                rotation_relax=1.0))


def _dtau(deck, dt, cfl):
    from ablsteadyrans_amd import pseudotime as PT
    from ablsteadyrans_amd.fab import Fab
    nu = Fab.at_loc(deck.tile.geom, 0, 1, 1)
    nu.data.zero_()
    PT.local_dtau(deck.mesh, deck.state["u"], deck.state["v"], deck.state["w"], cfl, dt, nu=nu)


@pytest.mark.parametrize("name,steps", [("column", 4), ("hill3d", 8), ("open", 8)])
def test_replays_equal_eager_iterations_bitwise(cuda, name, steps):
    import torch
    from ablsteadyrans_amd.fab import Fab
    n_it = 6
    print(f"graph vs eager: {name}, {steps} pressure steps, {n_it} iterations", flush=True)
    runs = {}
    side = torch.cuda.Stream()   # the null stream cannot be captured
    for mode in ("eager", "graph"):
        with torch.cuda.stream(side):
            deck = _deck(name)
            st = deck.simplec
            st.set_pressure_steps(steps)
            dt = Fab.at_loc(deck.tile.geom, 0, 1, 0)
            st.init(deck.state)
            hist = []
            for it in range(n_it):
                _dtau(deck, dt, 1.0 + 0.5 * it)
                if mode == "eager" or it == 0:
                    st.iterate(deck.state, dt)
                    if mode == "graph" and it == 0:
                        st.capture(deck.state, dt)
                else:
                    # eager work and a wait between replays
                    junk = torch.randn(1000, device="cuda").sum()
                    torch.cuda.synchronize()
                    del junk
                    st.replay()
                hist.append((_fields(deck), st.norms()))
            runs[mode] = hist
    for it in range(n_it):
        fe, ne = runs["eager"][it]
        fg, ng = runs["graph"][it]
        for n in NAMES:
            assert np.array_equal(fe[n], fg[n]), (name, it, n)
        assert ne == ng, (name, it)


@pytest.mark.parametrize("name,steps", [("column", 4), ("hill3d", 8)])
def test_graphed_driver_equals_the_eager_driver(cuda, name, steps):
    import torch
    out = {}
    for mode in ("eager", "graph"):
        deck = _deck(name)
        deck.simplec.set_pressure_steps(steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if mode == "eager":
            r = deck.simplec.solve(deck.state, deck.driver)
        else:
            r = deck.simplec.solve_graphed(deck.state, deck.driver, pressure_steps=steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[mode] = (r.iterations, r.status, _fields(deck), dt)
        print(f"  {name} {mode}: {r.iterations} iterations, status {r.status}, {dt:.3f} s = "
              f"{dt * 1e3 / r.iterations:.3f} ms per iteration", flush=True)
    assert out["eager"][1] == out["graph"][1] == 1
    assert out["eager"][0] == out["graph"][0]
    for n in NAMES:
        assert np.array_equal(out["eager"][2][n], out["graph"][2][n]), n


def test_too_few_pressure_steps_are_an_error(cuda):
    from ablsteadyrans_amd import _lib
    from ablsteadyrans_amd.fab import Fab
    deck = _deck("hill3d")
    st = deck.simplec
    st.set_pressure_steps(1)
    dt = Fab.at_loc(deck.tile.geom, 0, 1, 0)
    st.init(deck.state)
    _dtau(deck, dt, 1.0)
    with pytest.raises(_lib.GustError, match="did not converge in its 1 fixed GMRES steps"):
        st.iterate(deck.state, dt)


@pytest.mark.parametrize("name", ["column", "hill3d"])
def test_graphed_driver_with_learned_steps(cuda, name):
    """pressure_steps 0: the first iteration's GMRES count + 1 fixes the
    steps (grown by 2 after a miss); the verdict and the iteration count stay
    those of the eager driver with GMRES to the tolerance, within 5 %."""
    import torch
    out = {}
    for mode in ("eager", "graph"):
        deck = _deck(name)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = deck.simplec.solve(deck.state, deck.driver) if mode == "eager" else \
            deck.simplec.solve_graphed(deck.state, deck.driver)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[mode] = (r.iterations, r.status, dt)
        print(f"  {name} {mode} (steps learned): {r.iterations} iterations, status {r.status}, {dt:.3f} s = "
              f"{dt * 1e3 / r.iterations:.3f} ms per iteration", flush=True)
    assert out["eager"][1] == out["graph"][1] == 1
    assert abs(out["graph"][0] - out["eager"][0]) <= 0.05 * out["eager"][0]
