"""The measured workloads at their full sizes, on the GPU, against the oracle and against
fixtures written by the reference's own source (tests/golden/make_config_golden.py).

* 2D square droplet, 80x80 fluid particles (decks.droplet_2d): first-step parity, and the
  2000-step history of fluid momentum, sum m|v|, kinetic energy and centre of mass against
  the reference's, with the rebuild counts of every 100 steps;
* 3D dam break, 64x32x64 fluid particles in a 192x32x96 walled box (decks.dam_break_3d):
  first-step parity and the 1000-step history against the reference's;
* 3D two-fluid dam break with 1,048,576 fluid particles (decks.dam_break_1m, 1,272,064 with
  the walls), FP64: initial and step-1 neighbour sets identical by id, every field after
  one step within 1e-12 normwise of the oracle; the mixed path against FP64 after 20 steps
  with the per-field bounds of test_gpu_mixed;
* the 64-case uneven batch over 1000 replayed graph steps: sampled cases bit-identical to
  their single-case runs, every case finite with its particles;
* the surrogate on the 1M deck: a short training and two judged blocks (a property test:
  finite state, the judge's simulated time, errors within the bound on kept blocks).

Neighbour sets are compared exactly as sorted int64 keys id_i * n + id_j (tests.util).
"""
import os

import numpy as np
import pytest

from mphexplicit_amd import decks
from oracle.oracle import Oracle, by_id
from tests.util import compare_fields, csr_keys, fields_for, pair_keys_by_id, rest_scales

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def solver_mod(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    from mphexplicit_amd import solver
    return solver


def _same_lists(s, orc, label):
    gp, gi = s.neighbors()
    op, oi = orc.neighbors()
    kg = csr_keys(gp, gi)
    ko = pair_keys_by_id(orc.fields()["index"], op, oi)
    print("%s: %d list entries (gpu %d)" % (label, len(ko), len(kg)), flush=True)
    assert len(kg) == len(ko), label
    assert np.array_equal(kg, ko), label


def _first_step(solver_mod, deck, label, pressure_vs_exact=False):
    orc = Oracle(deck, nthreads=16)
    s = solver_mod.Solver.from_deck(deck)
    _same_lists(s, orc, label + " init")
    of = by_id(orc.fields())
    gf = s.fields()
    assert np.array_equal(gf["position"], of["x"])
    assert np.array_equal(gf["neighbor_count"], of["ncount"])
    compare_fields(gf, of, label + " init", names=("dens_a", "vol_strain"))
    orc.step(1)
    s.step(1)
    of = by_id(orc.fields())
    gf = s.fields()
    assert np.array_equal(gf["position"], of["x"])
    _same_lists(s, orc, label + " step 1")
    assert np.array_equal(gf["neighbor_count"], of["ncount"])
    # the 64x32x64 dam break is a lattice at rest: its first-step pressure is rounding noise
    # (compared at the scales of tests.util.rest_scales); the jittered 1M deck's is not
    p_scale, f_scale, v_scale = rest_scales(deck)
    names = fields_for(deck)
    if pressure_vs_exact:
        names = tuple(k for k in names if k != "press_p")
    compare_fields(gf, of, label + " step 1", names=names, p_scale=p_scale, f_scale=f_scale, v_scale=v_scale)
    if pressure_vs_exact:
        _pressure_vs_exact(deck, orc, of, gf, p_scale, label)
    return s


def _pressure_vs_exact(deck, orc, of, gf, p_scale, label):
    """P = -lambda D + kappa e^P (main.cpp:1905) with e^P = sum w_P - N0p, a difference of
    sums near N0p (main.cpp:1850-1856): its FP64 rounding depends on the summation order,
    which differs between the two codes (the reference's bitonic cell order against our
    clusters).  Each code's e^P is measured against the exact sum of the same terms (long
    double, oracle.vol_strain_exact): the GPU must be as close to it as the reference's own
    FP64 sum (which the oracle reproduces bit for bit), and P must agree with the
    reference to 1e-12 of the pressure of a unit strain, K N0p."""
    exact = by_id({"index": orc.fields()["index"], "e": orc.vol_strain_exact()})["e"]
    e_ref, e_gpu = of["vol_strain"], gf["vol_strain_p"]
    err_ref = np.abs(e_ref - exact).max()
    err_gpu = np.abs(e_gpu - exact).max()
    kappa = np.where(exact < 0.0, 0.0, np.asarray(deck.data.bulk_modulus)[of["type"]])
    lam = np.asarray(deck.data.bulk_viscosity)[of["type"]]
    p_exact = -lam * of["divergence"] + kappa * exact
    pmax = np.abs(of["press_p"]).max()
    dp_ref = np.abs(of["press_p"] - p_exact).max()
    dp_gpu = np.abs(gf["pressure_p"] - p_exact).max()
    dp = np.abs(gf["pressure_p"] - of["press_p"]).max()
    print("%s: e^P from the exact sum: reference %.2e, gpu %.2e; P from exact: reference %.2e, gpu %.2e of "
          "max|P|; P gpu vs reference %.2e of max|P|, %.2e of K N0p" % (
              label, err_ref, err_gpu, dp_ref / pmax, dp_gpu / pmax, dp / pmax, dp / p_scale), flush=True)
    assert err_gpu <= 2.0 * err_ref + 1e-18
    assert dp_gpu <= 2.0 * dp_ref
    assert dp / p_scale <= 1e-12


def _history_row(deck, f):
    t = f["type"]
    fl = t < 2
    m = np.asarray(deck.data.density)[t[fl]] * deck.spacing ** deck.dim
    v = f["velocity"][fl]
    x = f["position"][fl]
    return np.concatenate([(m[:, None] * v).sum(0), (m[:, None] * np.abs(v)).sum(0),
                           [0.5 * float((m * (v * v).sum(1)).sum())], (m[:, None] * x).sum(0) / m.sum()])


def _history(solver_mod, deck, steps, every):
    s = solver_mod.Solver.from_deck(deck)
    rows = [_history_row(deck, s.fields(names=("type", "position", "velocity")))]
    rebuilds = []
    last = s.stats()["rebuilds"]
    for k in range(steps // every):
        s.step(every)
        rows.append(_history_row(deck, s.fields(names=("type", "position", "velocity"))))
        r = s.stats()["rebuilds"]
        rebuilds.append(r - last)
        last = r
    return np.array(rows), np.array(rebuilds), s


def _check_history(name, h, rb, g, tol):
    ref, rref = g["history"], g["rebuilds"]
    scale = np.abs(ref[:, 3:6]).max()                       # sum m |v|
    dp = np.abs(h[:, 0:3] - ref[:, 0:3]).max() / scale
    da = np.abs(h[:, 3:6] - ref[:, 3:6]).max() / scale
    dk = np.abs(h[:, 6] - ref[:, 6]).max() / np.abs(ref[:, 6]).max()
    dc = np.abs(h[:, 7:10] - ref[:, 7:10]).max()
    print("%s: momentum %.2e, sum m|v| %.2e, kinetic energy %.2e (relative), centre of mass %.2e m" % (
        name, dp, da, dk, dc), flush=True)
    print("%s rebuilds gpu %s\n%s rebuilds ref %s" % (name, rb.tolist(), name, rref.tolist()), flush=True)
    assert dp <= tol and da <= tol and dk <= tol, (dp, da, dk)
    assert dc <= tol * 1e-2
    # a rebuild is decided by comparing a displacement with (0.05 l0)^2 (main.cpp:1102):
    # rounding can move a borderline trigger by a step, never change the rate
    assert np.abs(rb - rref).max() <= 1 and abs(int(rb.sum()) - int(rref.sum())) <= 3


def test_droplet80_first_step(solver_mod):
    _first_step(solver_mod, decks.droplet_2d(), "droplet2d 80x80")


def test_droplet80_2000_steps(solver_mod):
    g = np.load(os.path.join(GOLDEN, "config_droplet2d_80.npz"), allow_pickle=False)
    d = decks.droplet_2d()
    assert int(g["n"]) == d.n
    h, rb, s = _history(solver_mod, d, int(g["steps"]), int(g["every"]))
    # the wall-free periodic droplet conserves momentum (antisymmetric pair forces,
    # main.cpp:1930, 1767, 1975): both codes keep it at rounding level
    scale = np.abs(g["history"][:, 3:6]).max()
    assert np.abs(h[:, 0:3]).max() <= 1e-11 * scale
    assert np.abs(g["history"][:, 0:3]).max() <= 1e-11 * scale
    _check_history("droplet2d 80x80, 2000 steps", h, rb, g, 1e-8)
    assert s.time() == float(g["time"])


def test_dambreak64_first_step(solver_mod):
    _first_step(solver_mod, decks.dam_break_3d(), "dambreak3d 64x32x64")


def test_dambreak64_1000_steps(solver_mod):
    g = np.load(os.path.join(GOLDEN, "config_dambreak3d_64.npz"), allow_pickle=False)
    d = decks.dam_break_3d()
    assert int(g["n"]) == d.n
    h, rb, s = _history(solver_mod, d, int(g["steps"]), int(g["every"]))
    _check_history("dambreak3d 64x32x64, 1000 steps", h, rb, g, 1e-9)
    assert s.time() == float(g["time"])


def test_dambreak1m_fp64_first_step(solver_mod, monkeypatch):
    # the headline deck: 1,272,064 particles, FP64, neighbour sets by id and every field
    # within 1e-12 of the oracle after one step
    monkeypatch.setenv("MPH_TUNE", "4")
    d = decks.dam_break_1m()
    assert d.n == 1272064
    s = _first_step(solver_mod, d, "dambreak3d 1M", pressure_vs_exact=True)
    s.close()


def test_dambreak1m_mixed_vs_fp64(solver_mod, monkeypatch):
    from tests.test_gpu_mixed import TOL_STEP20, _errors
    d = decks.dam_break_1m()
    monkeypatch.setenv("MPH_TUNE", "4")
    s64 = solver_mod.Solver.from_deck(d, precision=solver_mod.MPH_FP64)
    s64.step(20)
    f64 = s64.fields(names=tuple(TOL_STEP20))
    s64.close()
    monkeypatch.setenv("MPH_TUNE", "8")
    s32 = solver_mod.Solver.from_deck(d, precision=solver_mod.MPH_MIXED)
    s32.step(20)
    f32 = s32.fields(names=tuple(TOL_STEP20))
    err = _errors(f32, f64, d, list(TOL_STEP20))
    print("dambreak3d 1M mixed vs FP64, 20 steps:", " ".join("%s=%.2e" % kv for kv in err.items()), flush=True)
    bad = {k: e for k, e in err.items() if not e <= TOL_STEP20[k]}
    assert not bad, bad


def test_batch64_1000_graph_steps(solver_mod, monkeypatch):
    monkeypatch.setenv("MPH_TUNE", "8")
    cases = decks.batch_cases(count=64)
    b = solver_mod.Solver.from_decks(cases)
    for _ in range(10):
        b.step_graph(100)
    b.sync()
    assert b.stats()["steps"] == 1000
    n = [d.n for d in cases]
    for k in (int(np.argmin(n)), int(np.argmax(n)), 40):
        d = cases[k]
        s = solver_mod.Solver.from_deck(d)
        s.step(1000)
        names = ("position", "velocity", "pressure_p", "neighbor_count")
        fs, fb = s.fields(names=names), b.fields(case=k, names=names)
        for nm in names:
            assert np.array_equal(fs[nm], fb[nm]), (d.name, nm)
        assert s.time() == b.time(case=k)
        s.close()
    for k, d in enumerate(cases):
        f = b.fields(case=k, names=("type", "position", "velocity"))
        assert np.array_equal(f["type"], d.type), d.name
        assert np.isfinite(f["position"]).all() and np.isfinite(f["velocity"]).all(), d.name


def test_surrogate_1m_judged(solver_mod, monkeypatch):
    monkeypatch.setenv("MPH_TUNE", "4")
    d = decks.dam_break_1m()
    s = solver_mod.Solver.from_deck(d)
    s.step(100)
    sur = solver_mod.Surrogate(s, train_stride=16, lr=1e-3, seed=5, u_scale=0.1, a_scale=10.0, dens_scale=50.0)
    losses = sur.train(24, 4)
    print("surrogate 1M: loss %.3g -> %.3g" % (losses[0], losses[-1]), flush=True)
    assert np.isfinite(losses).all() and losses[-1] < losses[0]
    t0 = s.time()
    k, blocks, tol = 10, 2, 0.01
    st = sur.advance(blocks, k=k, tol=tol)
    print("surrogate 1M judged:", st, flush=True)
    assert st["blocks"] == blocks
    kept = blocks - st["rollbacks"]
    assert st["sur_steps"] == k * kept
    # a kept block advances k surrogate steps and the judge's two plain steps; a rejected
    # one is replayed with 2k plain steps
    dt = d.data.dt
    assert st["sim_time"] == pytest.approx(blocks * 2 * k * dt + kept * 2 * dt, rel=1e-9)
    assert s.time() - t0 == pytest.approx(st["sim_time"], rel=1e-9)
    if st["rollbacks"] == 0:
        assert st["max_err_v"] <= tol and st["max_err_n"] <= tol
    f = s.fields(names=("position", "velocity"))
    assert np.isfinite(f["position"]).all() and np.isfinite(f["velocity"]).all()
