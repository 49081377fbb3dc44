"""Mass and momentum over 1000 steps: the HIP path against the oracle (which is bit-identical
to the reference's own build over the same 1000 steps, see test_oracle.py).

Particle mass is rho_t V0 per type (main.cpp:1636), so total mass is conserved exactly when
no particle is lost or retyped.  Momentum:
* 2D droplet (periodic, no walls, g = 0, one fluid): the pair forces are antisymmetric
  (main.cpp:1930, 1767, 1975), so sum m v stays at rounding level in both codes; the drift
  is bounded relative to sum m |v| and the two trajectories agree to that level;
* 3D dam break (gravity, walls): sum m v of the fluid follows gravity and the wall forces;
  the GPU trajectory matches the oracle's every 100 steps.
Sampled every 100 steps; the bounds are stated with each assertion.
"""
import numpy as np
import pytest

from mphexplicit_amd import decks
from oracle.oracle import Oracle, by_id

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver_cls(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    from mphexplicit_amd.solver import Solver
    return Solver


def _mass(deck, types):
    rho = np.array(deck.data.density)
    return rho[types] * deck.spacing ** deck.dim


def _trajectories(solver_cls, deck, steps=1000, every=100):
    orc = Oracle(deck, nthreads=16)
    s = solver_cls.from_deck(deck)
    out = {"gpu": [], "oracle": [], "abs_gpu": [], "abs_oracle": []}
    orc_rebuilds = 0
    for k in range(steps // every):
        s.step(every)
        orc_rebuilds += orc.step(every)
        g = s.fields(names=("type", "velocity"))
        o = by_id(orc.fields())
        for tag, t, v in (("gpu", g["type"], g["velocity"]), ("oracle", o["type"], o["v"])):
            fl = t < 2
            m = _mass(deck, t)[fl]
            out[tag].append((m[:, None] * v[fl]).sum(0))
            out["abs_" + tag].append((m[:, None] * np.abs(v[fl])).sum(0))
        assert np.array_equal(g["type"], deck.type)          # no particle lost or retyped
        print("%s step %d P_gpu %s P_oracle %s" % (deck.name, (k + 1) * every, out["gpu"][-1],
                                                   out["oracle"][-1]), flush=True)
    tr = {k: np.array(v) for k, v in out.items()}
    tr["rebuilds_oracle"] = orc_rebuilds
    tr["rebuilds_gpu"] = s.stats()["rebuilds"] - 1         # the count includes the initial build
    return tr, s, orc


def test_mass_exact(solver_cls):
    d = decks.drift_dam_break_3d()
    s = solver_cls.from_deck(d)
    m0 = _mass(d, d.type).sum()
    s.step(1000)
    f = s.fields(names=("type",))
    assert _mass(d, f["type"]).sum() == m0


def test_momentum_drift_droplet(solver_cls):
    d = decks.small_droplet_2d()
    tr, s, orc = _trajectories(solver_cls, d)
    scale = tr["abs_oracle"].max()
    # conservation: both codes keep |sum m v| below 1e-11 of sum m |v| over 1000 steps
    assert np.abs(tr["oracle"]).max() <= 1e-11 * scale
    assert np.abs(tr["gpu"]).max() <= 1e-11 * scale
    # and their drifts agree at that level
    assert np.abs(tr["gpu"] - tr["oracle"]).max() <= 1e-11 * scale


def test_momentum_trajectory_dam_break(solver_cls):
    d = decks.drift_dam_break_3d()
    tr, s, orc = _trajectories(solver_cls, d)
    scale = np.abs(tr["oracle"]).max()
    rel = np.abs(tr["gpu"] - tr["oracle"]).max() / scale
    print("dam break momentum trajectory: max deviation / max |P| = %.2e" % rel, flush=True)
    # FP64 rounding differences (summation order) stay below 1e-11 of the momentum scale
    # over 1000 steps that include about 226 list rebuilds (measured: 7e-15)
    assert rel <= 1e-11
    # rebuild decisions compare a displacement with (0.05 l0)^2 (main.cpp:1102): rounding
    # can move a borderline trigger by a step, not change the rebuild rate
    print("rebuilds gpu %d oracle %d" % (tr["rebuilds_gpu"], tr["rebuilds_oracle"]), flush=True)
    assert abs(tr["rebuilds_gpu"] - tr["rebuilds_oracle"]) <= 2
