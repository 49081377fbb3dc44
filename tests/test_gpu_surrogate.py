"""The pairwise-message surrogate on the GPU (mph_sur_* through the C ABI):

* forward (BF16 matrix cores, FP32 accumulation) against the CPU oracle with the same
  bfloat16 roundings (oracle/surrogate.py), on a 3D dam-break state;
* loss and gradient of the project's backward kernels against the oracle's analytic
  gradient (the oracle's gradient is pinned by finite differences in
  test_surrogate_oracle.py);
* training on the solver's own trajectory lowers the loss;
* the judge rolls back an untrained network and keeps a trained one;
* what the trained network gains: 2 dt surrogate steps against plain steps at 2 dt and the
  untrained network, all measured against the plain FP64 run at dt.
"""
import copy

import numpy as np
import pytest

from mphexplicit_amd import decks
from oracle import surrogate as S
from oracle.oracle import Oracle
from tests.util import normwise

pytestmark = pytest.mark.gpu

SCALES = dict(u_scale=0.1, a_scale=10.0, dens_scale=50.0)


@pytest.fixture(scope="module")
def mods(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    from mphexplicit_amd import solver
    return solver


def _oracle_inputs(deck, s):
    f = s.fields(names=("position", "velocity", "vol_strain_p", "type"))
    ptr, ind = s.neighbors()
    sc = Oracle(deck).scalars()
    n0p, maxr = sc[1], sc[3]
    lo, hi = deck.domain_min, deck.domain_max
    pi, pj = S.pairs_within(f, ptr, ind, lo, hi, maxr)
    X, basis = S.features(f, pi, pj, lo, hi, maxr, n0p, SCALES["u_scale"], SCALES["dens_scale"])
    return f, pi, X, basis, S.mass_ratio(f, deck.data.density)


def _state(mods, steps=6):
    d = decks.small_dam_break_3d()
    s = mods.Solver.from_deck(d)
    s.step(steps)
    return d, s


def test_forward_matches_oracle(mods):
    d, s = _state(mods)
    sur = mods.Surrogate(s, seed=7, **SCALES)
    w = sur.weights()
    a_gpu = sur.predict()
    f, pi, X, basis, ratio = _oracle_inputs(d, s)
    a_orc = S.predict(w, X, basis, pi, ratio, SCALES["a_scale"], emulate=True)
    fl = f["type"] < 2
    err = normwise(a_gpu[fl], a_orc[fl])
    print("surrogate forward normwise vs oracle %.2e, pairs %d" % (err, len(pi)), flush=True)
    # same bf16 roundings; differences come from FP32 vs FP64 arithmetic before a rounding
    assert err <= 2e-3
    assert np.all(a_gpu[~fl] == 0.0)


def test_messages_conserve_momentum(mods):
    # wall-free periodic droplet, one fluid: m_ji = -m_ij and symmetric pair sets make the
    # predicted accelerations sum to zero up to FP32 rounding, for any weights
    d = decks.small_droplet_2d()
    s = mods.Solver.from_deck(d)
    s.step(5)
    sur = mods.Surrogate(s, seed=3, **SCALES)
    rng = np.random.default_rng(0)
    sur.set_weights(sur.weights() * 20.0 + rng.normal(size=sur.weights().shape).astype(np.float32) * 0.05)
    a = sur.predict()
    total, scale = np.abs(a.sum(0)).max(), np.abs(a).sum(0).max()
    print("sum a %.3e of sum |a| %.3e" % (total, scale), flush=True)
    assert total <= 1e-5 * scale


def test_gradient_matches_oracle(mods):
    d, s = _state(mods)
    sur = mods.Surrogate(s, seed=11, **SCALES)
    w = sur.weights()
    rng = np.random.default_rng(3)
    target = rng.normal(size=(d.n, 3)) * 5.0
    loss_g, g_gpu = sur.gradient(target)
    f, pi, X, basis, ratio = _oracle_inputs(d, s)
    train = np.nonzero(f["type"] < 2)[0]
    loss_o, g_orc = S.loss_and_grad(w, X, basis, pi, ratio, train, target, SCALES["a_scale"], emulate=True)
    print("loss gpu %.8e oracle %.8e" % (loss_g, loss_o), flush=True)
    assert abs(loss_g - loss_o) <= 1e-4 * loss_o
    o = 0
    for name, shp in S.SHAPES:
        m = int(np.prod(shp))
        err = normwise(g_gpu[o:o + m], g_orc[o:o + m])
        print("grad %s normwise %.2e" % (name, err), flush=True)
        # bf16 inputs of the weight-gradient products (emulated by the oracle), FP32 sums
        assert err <= 1e-2, name
        o += m


def _trained_weights(mods, deck, samples=300, iters=8, lr=1e-3, warm=0):
    s = mods.Solver.from_deck(deck)
    s.step(warm)
    sur = mods.Surrogate(s, seed=5, lr=lr, **SCALES)
    losses = sur.train(samples, iters)
    w = sur.weights()
    sur.close()
    s.close()
    return w, losses


def test_training_lowers_loss(mods):
    w, losses = _trained_weights(mods, decks.small_dam_break_3d(), samples=20, iters=20, lr=2e-3)
    first, last = losses[:5].mean(), losses[-20:].mean()
    print("training loss first %.4e last %.4e" % (first, last), flush=True)
    assert np.all(np.isfinite(losses))
    assert last < 0.3 * first


def test_training_across_decks(mods):
    # one network trained on trajectories of several decks in turn (a 2D droplet, then a 3D
    # dam break), as the surrogate of the 1M-particle deck is: the weights carry from one
    # context to the next, and on the dam break training goes on from them
    w2d, l2d = _trained_weights(mods, decks.small_droplet_2d(), samples=20, iters=10, lr=2e-3)
    s = mods.Solver.from_deck(decks.small_dam_break_3d())
    s.step(6)
    sur = mods.Surrogate(s, seed=9, lr=2e-3, **SCALES)
    sur.set_weights(w2d)
    assert np.array_equal(sur.weights(), w2d)
    losses = sur.train(20, 10)
    w = sur.weights()
    print("2D droplet loss %.3e -> %.3e; dam break %.3e -> %.3e" % (l2d[0], l2d[-1], losses[0], losses[-1]),
          flush=True)
    assert np.all(np.isfinite(l2d)) and np.all(np.isfinite(losses))
    assert l2d[-10:].mean() < l2d[:5].mean()
    assert losses[-10:].mean() < losses[:5].mean()
    assert not np.array_equal(w, w2d)
    sur.close()
    s.close()


def _errors(a, b):
    fl = b["type"] < 2
    ev = np.linalg.norm(a["velocity"][fl] - b["velocity"][fl]) / np.linalg.norm(b["velocity"][fl])
    en = np.linalg.norm(a["density_a"][fl] - b["density_a"][fl]) / np.linalg.norm(b["density_a"][fl])
    return float(ev), float(en)


def test_network_against_plain_2dt(mods):
    """What the trained network gains over the method without it.  A surrogate step
    advances 2 dt; without the network the same 2 dt steps are plain MPH steps at 2 dt
    (mph_set_time_step).  Both start from the same developed state and are compared with
    plain steps at dt after 2M dt.  Measured on this deck: the trained network is far
    closer than the untrained one and its judged blocks are kept at the 1% bound, but plain
    steps at 2 dt are more accurate still (and cheaper: the bench's surrogate leg), so the
    network gains nothing over them here; the assertions state that measured ordering."""
    deck = decks.small_dam_break_3d()
    # trained on the window that is judged (from step 100, like the bench): its per-step
    # error measures 0.8-1.0% there, trained from rest 0.95-1.16%, at the judge's 1% bound
    w, losses = _trained_weights(mods, deck, warm=100)
    print("training loss first %.4e last %.4e" % (losses[:5].mean(), losses[-20:].mean()), flush=True)
    names = ("velocity", "density_a", "type")
    W = 100                                 # plain steps before the comparison (flow developed)
    M = 40                                  # surrogate steps = 2 M plain steps of dt

    def start(d=deck, steps=W):
        s = mods.Solver.from_deck(d)
        s.step(steps)
        return s

    # plain FP64 at dt: the reference trajectory
    ref = start()
    ref.step(2 * M)
    fr = ref.fields(names=names)
    # without the network: plain steps at 2 dt from the same state
    p2 = start()
    p2.set_dt(2.0 * deck.data.dt)
    p2.step(M)
    results = {"plain_2dt": _errors(p2.fields(names=names), fr)}
    for label, weights in (("untrained", None), ("trained", w)):
        s = start()
        sur = mods.Surrogate(s, seed=5, **SCALES)
        if weights is not None:
            sur.set_weights(weights)
        sur.step(M)
        # an untrained network can compress the fluid until the lists widen: the run goes on
        ev, en = _errors(s.fields(names=names), fr)
        results[label] = (float(np.nan_to_num(ev, nan=np.inf)), float(np.nan_to_num(en, nan=np.inf)))
        assert abs(s.time() - ref.time()) <= 1e-12 * ref.time()
        s2 = start()
        sur2 = mods.Surrogate(s2, seed=5, **SCALES)
        if weights is not None:
            sur2.set_weights(weights)
        st = sur2.advance(blocks=3, k=10, tol=0.01)
        results[label + "_judged"] = st
        print(label, "judged", st, flush=True)
        f2 = s2.fields(names=names)
        assert np.all(np.isfinite(f2["velocity"]))       # rollbacks keep the state sane
    print("relative L2 (velocity, density) against plain dt after %d dt from step %d: %s" % (
        2 * M, W, {k: v for k, v in results.items() if not k.endswith("judged")}), flush=True)
    tr, un, p2e = results["trained"], results["untrained"], results["plain_2dt"]
    # the trained network is closer to the plain run than the untrained one
    assert tr[0] < un[0] and tr[1] < un[1]
    # the judge rejects every block of the untrained network and keeps blocks of the trained one
    assert results["untrained_judged"]["rollbacks"] == 3
    assert results["trained_judged"]["sur_steps"] > 0
    # the method without the network, plain steps at 2 dt, is the more accurate one here
    assert p2e[0] < tr[0] and p2e[1] < tr[1]
    assert p2e[0] < 0.01 and p2e[1] < 0.01


def test_judge_rolls_back_nan_weights(mods):
    # corrupted weights: every block is rolled back and the state stays the plain path's
    deck = decks.small_dam_break_3d()
    s = mods.Solver.from_deck(deck)
    s.step(20)
    ref = mods.Solver.from_deck(deck)
    ref.step(20 + 2 * 10 * 2)
    sur = mods.Surrogate(s, seed=5, **SCALES)
    w = sur.weights()
    w[100:200] = np.nan
    sur.set_weights(w)
    st = sur.advance(blocks=2, k=10, tol=0.01)
    assert st["rollbacks"] == 2 and st["sur_steps"] == 0
    a, b = s.fields(names=("velocity", "position")), ref.fields(names=("velocity", "position"))
    # the replayed blocks are plain steps: the result is the plain run's (same operations)
    assert np.all(np.isfinite(a["velocity"]))
    assert normwise(a["velocity"], b["velocity"]) <= 1e-9
