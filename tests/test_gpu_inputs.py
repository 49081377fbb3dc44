"""Edge cases of the C ABI on the GPU.

* Inputs that mph_create must reject with an error instead of running: a type outside 0..3,
  a particle outside the periodic domain (or not finite), a case without particles, a
  non-positive time step, a dimension other than 2 or 3.  The library checks them on the host
  before any kernel runs (the reference exits or reads past its arrays instead:
  source/errorfunc.cpp:8-31, main.cpp:549, 652).
* A negative step count is an error; zero steps change nothing.
* The smallest cases against the oracle: one particle (a list holding only itself, all sums
  empty) and one interacting pair.
"""
import dataclasses

import numpy as np
import pytest

from mphexplicit_amd import decks
from oracle.oracle import Oracle, by_id
from tests.util import FP64_TOL, normwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mod(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    from mphexplicit_amd import solver
    return solver


def _base():
    return decks.small_droplet_2d()


def _with(d, **kw):
    return dataclasses.replace(d, **kw)


def _bad_type(d):
    t = d.type.copy()
    t[7] = 5
    return _with(d, type=t)


def _outside(d):
    x = d.x.copy()
    x[3, 0] = d.domain_max[0] + 2.0 * (d.domain_max[0] - d.domain_min[0])
    return _with(d, x=x)


def _not_finite(d):
    x = d.x.copy()
    x[11, 1] = np.nan
    return _with(d, x=x)


BAD_DECKS = {
    "type": (_bad_type, "outside 0..3"),
    "outside": (_outside, "outside the periodic domain"),
    "nan": (_not_finite, "outside the periodic domain"),
}


@pytest.mark.parametrize("name", list(BAD_DECKS))
def test_create_rejects_bad_particles(mod, name):
    make, msg = BAD_DECKS[name]
    d = make(_base())
    with pytest.raises(mod.MphError) as e:
        mod.Solver.from_deck(d)
    print(name, "->", e.value, flush=True)
    assert msg in str(e.value)


def test_create_rejects_empty_case(mod):
    d = _base()
    empty = _with(d, type=d.type[:0].copy(), x=d.x[:0].copy(), v=d.v[:0].copy())
    with pytest.raises(mod.MphError) as e:
        mod.Solver.from_decks([d, empty])
    assert "no particles" in str(e.value)


@pytest.mark.parametrize("field,value,msg", [("dt", 0.0, "dt"), ("dt", float("inf"), "dt"), ("dim", 4, "dim")])
def test_create_rejects_bad_parameters(mod, field, value, msg):
    d = _base()
    p = mod.params_from_deck(d)
    setattr(p, field, value)
    with pytest.raises(mod.MphError) as e:
        mod.Solver([p], [d.type], [d.x], [d.v])
    assert msg in str(e.value)


def test_step_counts(mod):
    d = _base()
    s = mod.Solver.from_deck(d)
    with pytest.raises(mod.MphError):
        s.step(-1)
    before = s.fields(names=("position", "velocity"))
    s.step(0)
    after = s.fields(names=("position", "velocity"))
    assert np.array_equal(before["position"], after["position"])
    assert np.array_equal(before["velocity"], after["velocity"])
    assert s.time() == 0.0
    s.close()


def _tiny(n):
    d = _base()
    tiny = _with(d, type=d.type[:n].copy(), x=d.x[:n].copy(), v=d.v[:n].copy(), name="tiny%d" % n)
    if n == 2:
        # one pair two spacings apart: inside every kernel radius (3.5 spacings)
        tiny.x[1] = tiny.x[0] + np.array([2.0 * d.spacing, 0.0, 0.0])
    return tiny


@pytest.mark.parametrize("n", [1, 2])
def test_smallest_cases_match_oracle(mod, n):
    d = _tiny(n)
    orc = Oracle(d, nthreads=1)
    s = mod.Solver.from_deck(d)
    orc.step(5)
    s.step(5)
    of = by_id(orc.fields())
    gf = s.fields()
    assert np.array_equal(gf["neighbor_count"], of["ncount"])
    for k, ok in (("position", "x"), ("velocity", "v"), ("density_a", "dens_a"), ("pressure_p", "press_p")):
        ref = of[ok]
        scale = np.max(np.abs(ref))
        err = normwise(gf[k], ref) if scale > 0 else float(np.max(np.abs(gf[k])))
        print(n, k, "%.2e" % err, flush=True)
        assert err <= FP64_TOL, k
    s.close()


def test_walls_only_deck_matches_oracle(mod):
    # no fluid at all: no wall particle has a list, every sum is zero and static walls stay
    from mphexplicit_amd import decks
    d = decks.walls_only_3d()
    orc = Oracle(d, nthreads=1)
    s = mod.Solver.from_deck(d)
    orc.step(5)
    s.step(5)
    of = by_id(orc.fields())
    gf = s.fields()
    assert np.array_equal(gf["neighbor_count"], of["ncount"])
    assert not gf["neighbor_count"].any()
    assert np.array_equal(gf["position"], of["x"])
    for k in ("velocity", "force", "density_a", "pressure_p"):
        assert not np.any(gf[k]), k
    ptr, ind = s.neighbors()
    assert len(ind) == 0 and not np.any(ptr)
    s.close()
