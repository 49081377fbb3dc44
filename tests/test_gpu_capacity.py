"""Neighbour lists that outgrow their reserved width (main.cpp:1487-1495, 1598-1602: the
reference sizes NeighborInd at every rebuild).

The deck is a sparse cloud collapsing on its centre (decks.collapsing_cloud_3d): its lists
more than double over 200 steps, past the width the first build reserves.  Every step path
(eager, graph replay, timed steps) must widen the lists inside the step that overflowed,
finish it and go on, with the oracle's results.  The surrogate's steps widen them too.
"""
import numpy as np
import pytest

from mphexplicit_amd import decks
from oracle.oracle import Oracle, by_id
from tests.util import normwise, oracle_lists_by_id

pytestmark = pytest.mark.gpu

STEPS = 200


@pytest.fixture(scope="module")
def solver_cls(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    from mphexplicit_amd.solver import Solver
    return Solver


@pytest.fixture(scope="module")
def oracle_state():
    d = decks.collapsing_cloud_3d()
    orc = Oracle(d, nthreads=16)
    rebuilds = orc.step(STEPS)
    ptr, ind = orc.neighbors()
    f = orc.fields()
    return d, by_id(f), oracle_lists_by_id(f, ptr, ind), rebuilds


def _rows(s):
    ptr, ind = s.neighbors()
    return [tuple(ind[ptr[i]:ptr[i + 1]]) for i in range(len(ptr) - 1)]


@pytest.mark.parametrize("mode", ["eager", "graph", "timed"])
def test_lists_grow_and_match_oracle(solver_cls, oracle_state, mode, monkeypatch):
    monkeypatch.setenv("MPH_TUNE", "4")
    d, of, rows_o, rb_o = oracle_state
    print("case collapsing_cloud3d", mode, flush=True)
    s = solver_cls.from_deck(d)
    w0 = s.stats()["list_width"]
    for _ in range(STEPS // 50):
        if mode == "eager":
            s.step(50)
        elif mode == "graph":
            s.step_graph(50)
        else:
            s.step_timed(50)
    st = s.stats()
    print("width %d -> %d, longest list %d, rebuilds gpu %d oracle %d" % (
        w0, st["list_width"], st["max_list_len"], st["rebuilds"] - 1, rb_o), flush=True)
    assert st["list_width"] > w0, "the deck did not outgrow the first width"
    assert st["steps"] == STEPS
    assert s.time() == pytest.approx(d.data.dt * STEPS, rel=1e-12)
    gf = s.fields()
    for a, b in (("position", "x"), ("velocity", "v"), ("force", "force"), ("density_a", "dens_a"),
                 ("vol_strain_p", "vol_strain"), ("pressure_p", "press_p")):
        err = normwise(gf[a], of[b])
        print("%s %s %.2e" % (mode, a, err), flush=True)
        # rounding differences of 200 steps, as the multi-step parity bound allows
        assert err <= 1e-9, (a, err)
    assert _rows(s) == rows_o
    assert np.array_equal(gf["neighbor_count"], of["ncount"])


def test_paths_bitwise_after_growth(solver_cls, monkeypatch):
    monkeypatch.setenv("MPH_TUNE", "4")
    d = decks.collapsing_cloud_3d()
    out = {}
    for mode in ("eager", "graph", "timed"):
        s = solver_cls.from_deck(d)
        getattr(s, {"eager": "step", "graph": "step_graph", "timed": "step_timed"}[mode])(STEPS)
        out[mode] = (s.fields(), s.stats())
        s.close()
    for mode in ("graph", "timed"):
        for k in ("position", "velocity", "force", "pressure_p", "neighbor_count"):
            assert np.array_equal(out[mode][0][k], out["eager"][0][k]), (mode, k)
        assert out[mode][1]["rebuilds"] == out["eager"][1]["rebuilds"]
        assert out[mode][1]["list_width"] == out["eager"][1]["list_width"]


def test_rejected_block_that_outgrew_the_lists(solver_cls, monkeypatch):
    # a judged block whose surrogate steps outgrow the list width is rejected and replayed
    # with plain steps: the call succeeds and leaves the plain path's results (the replay
    # starts with a rebuild, so sums are ordered differently from an uninterrupted run)
    monkeypatch.setenv("MPH_TUNE", "4")
    from mphexplicit_amd.solver import Surrogate
    d = decks.collapsing_cloud_3d()
    a = solver_cls.from_deck(d)
    b = solver_cls.from_deck(d)
    w0 = b.stats()["list_width"]
    sur = Surrogate(b, seed=7)
    sur.set_weights(np.zeros_like(sur.weights()))
    stats = sur.advance(1, k=STEPS // 2, tol=1e-6)
    print("rejected block:", stats, "width %d -> %d" % (w0, b.stats()["list_width"]), flush=True)
    assert stats["blocks"] == 1 and stats["rollbacks"] == 1
    assert b.stats()["list_width"] > w0
    a.step(STEPS)
    assert b.time() == pytest.approx(a.time(), rel=1e-12)
    fa, fb = a.fields(), b.fields()
    for k in ("position", "velocity", "pressure_p"):
        err = normwise(fb[k], fa[k])
        print("replayed vs plain %s %.2e" % (k, err), flush=True)
        assert err <= 1e-9, (k, err)


def test_surrogate_steps_widen_lists(solver_cls, monkeypatch):
    # a network with zero weights predicts no force: its steps carry the cloud in ballistically,
    # faster than the pressure lets the plain path, and the lists outgrow the width inside the
    # surrogate's own steps, whose rebuilds are checked on the host and widened
    monkeypatch.setenv("MPH_TUNE", "4")
    from mphexplicit_amd.solver import Surrogate
    d = decks.collapsing_cloud_3d()
    s = solver_cls.from_deck(d)
    w0 = s.stats()["list_width"]
    sur = Surrogate(s, seed=7)
    sur.set_weights(np.zeros_like(sur.weights()))
    sur.step(STEPS // 2)
    st = s.stats()
    print("surrogate: width %d -> %d" % (w0, st["list_width"]), flush=True)
    assert st["list_width"] > w0
    f = s.fields(names=("position", "velocity"))
    assert np.isfinite(f["position"]).all() and np.isfinite(f["velocity"]).all()
    # training reads the pairs of the widened lists (its pair table is widened with them)
    losses = sur.train(1, 1)
    assert np.isfinite(losses).all()
    # judged blocks (kept or replayed with plain steps) go on from the widened state
    stats = sur.advance(2, k=5, tol=0.01)
    print("judged:", stats, flush=True)
    assert stats["blocks"] == 2
