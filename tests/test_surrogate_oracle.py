"""The surrogate's CPU oracle (oracle/surrogate.py), pinned without the GPU:

* its analytic gradient against central finite differences of its own exact FP64 loss;
* a planted mistake in the backward pass is caught by the same check;
* features of hand-built pairs against their closed forms (periodic wrap included);
* the bfloat16 rounding against known bit patterns.
"""
import numpy as np
import pytest

from oracle import surrogate as S


def _problem(seed=0, n=12, npairs=160):
    rng = np.random.default_rng(seed)
    w = np.zeros(S.NPARAMS)
    p = S.unflatten(w)
    p["W1"] = rng.uniform(-0.5, 0.5, (17, 64))
    p["W2"] = rng.uniform(-0.2, 0.2, (64, 64))
    p["b2"] = rng.uniform(-0.1, 0.1, 64)
    p["W3"] = rng.uniform(-0.2, 0.2, (64, 64))
    p["b3"] = rng.uniform(-0.1, 0.1, 64)
    p["W4"] = rng.uniform(-0.2, 0.2, (64, 3))
    p["b4"] = rng.uniform(-0.1, 0.1, 3)
    w = S.flatten(p)
    X = rng.normal(size=(npairs, 17))
    X[:, 16] = 1.0
    e = rng.normal(size=(npairs, 3))
    e /= np.linalg.norm(e, axis=1)[:, None]
    u = rng.normal(size=(npairs, 3))
    basis = dict(e=e, u=u, ue=(u * e).sum(1))
    pi = rng.integers(0, n, npairs)
    ratio = rng.uniform(0.8, 1.25, n)
    target = rng.normal(size=(n, 3)) * 5.0
    train = np.arange(0, n, 2)
    return w, X, basis, pi, ratio, train, target


def _loss(w, X, basis, pi, ratio, train, target, a_scale):
    return S.loss_and_grad(w, X, basis, pi, ratio, train, target, a_scale, emulate=False)[0]


def test_gradient_matches_finite_differences():
    w, X, basis, pi, ratio, train, target = _problem()
    a_scale = 3.0
    loss, g = S.loss_and_grad(w, X, basis, pi, ratio, train, target, a_scale, emulate=False)
    rng = np.random.default_rng(1)
    # coordinates from every tensor, biases included
    idx, o = [], 0
    for _, shp in S.SHAPES:
        m = int(np.prod(shp))
        idx += list(o + rng.choice(m, size=min(m, 12), replace=False))
        o += m
    worst = 0.0
    for k in idx:
        h = 1e-6 * max(1.0, abs(w[k]))
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (_loss(wp, X, basis, pi, ratio, train, target, a_scale) -
              _loss(wm, X, basis, pi, ratio, train, target, a_scale)) / (2 * h)
        worst = max(worst, abs(fd - g[k]) / max(1e-3 * np.abs(g).max(), abs(g[k])))
    assert worst <= 1e-6, worst


def test_planted_backward_error_is_caught(monkeypatch):
    w, X, basis, pi, ratio, train, target = _problem(seed=3)
    good = S.dsilu
    monkeypatch.setattr(S, "dsilu", lambda z: good(z) * (1.0 + 0.01 * np.sign(z)))   # a 1% slip
    loss, g = S.loss_and_grad(w, X, basis, pi, ratio, train, target, 2.0, emulate=False)
    k = S.NPARAMS - 100       # a W4 weight: its gradient does not involve dsilu
    j = 100                   # a W1 weight: its gradient goes through all three dsilu
    h = 1e-6
    for kk, should_match in ((k, True), (j, False)):
        wp, wm = w.copy(), w.copy()
        wp[kk] += h
        wm[kk] -= h
        fd = (_loss(wp, X, basis, pi, ratio, train, target, 2.0) -
              _loss(wm, X, basis, pi, ratio, train, target, 2.0)) / (2 * h)
        rel = abs(fd - g[kk]) / max(abs(fd), 1e-12)
        assert (rel <= 1e-6) == should_match, (kk, rel)


def test_features_closed_form():
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([10.0, 10.0, 10.0])
    f = {"position": np.array([[1.0, 5.0, 5.0], [3.0, 5.0, 5.0], [9.5, 5.0, 5.0], [5.0, 5.0, 5.0]]),
         "velocity": np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0], [0.0, 0.1, 0.0], [0.0, 0.0, 0.0]]),
         "vol_strain_p": np.array([0.0, 0.1, -0.1, 0.0]), "type": np.array([0, 1, 2, 0])}
    pi, pj = np.array([0, 0, 1]), np.array([1, 2, 0])
    X, B = S.features(f, pi, pj, lo, hi, maxr=3.5, n0p=1.0, u_scale=0.1, dens_scale=10.0)
    # pair (0,1): along +x at r = 2, relative velocity (2, 0, 0) in units of u_scale
    np.testing.assert_allclose(X[0, :4], [2.0 / 3.5, (2.0 / 3.5) ** 2, 2.0, 4.0])
    np.testing.assert_array_equal(X[0, 4:10], [0, 1, 0, 0, 0, 0])        # classes {0, 1}
    np.testing.assert_allclose(X[0, 10:17], [1.0, 1.0, 0.0, 0, 0, 0, 1.0], atol=1e-12)
    np.testing.assert_allclose(B["e"][0], [1.0, 0.0, 0.0])
    # pair (0,2): across the periodic face, x_j - x_i = -1.5; classes {0, wall}
    np.testing.assert_allclose(X[1, :4], [1.5 / 3.5, (1.5 / 3.5) ** 2, 0.0, 1.0], atol=1e-12)
    np.testing.assert_array_equal(X[1, 4:10], [0, 0, 1, 0, 0, 0])
    np.testing.assert_allclose(B["e"][1], [-1.0, 0.0, 0.0])
    # (1,0) is (0,1) seen from the other side: same features, opposite basis and message
    np.testing.assert_allclose(X[2], X[0], atol=1e-12)
    np.testing.assert_allclose(B["e"][2], -B["e"][0])
    np.testing.assert_allclose(B["u"][2], -B["u"][0])
    C = np.array([[0.3, -0.7, 0.2]] * 3)
    m = S.messages(C, B)
    np.testing.assert_allclose(m[2], -m[0], atol=1e-15)


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties go to even (1.0);
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties go to even (1 + 2^-6)
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -12, -2.5])
    got = S.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2.0 ** -6
    assert got[3] == 1.0 + 2.0 ** -7 and got[4] == -2.5
    # every bf16 value keeps 8 significant bits: the low 16 bits of its FP32 pattern are zero
    r = np.random.default_rng(0).normal(size=1000) * 100
    assert np.all(S.bf16(r).astype(np.float32).view(np.uint32) & 0xFFFF == 0)
    assert np.max(np.abs(S.bf16(r) - r) / np.abs(r)) <= 2.0 ** -8


def test_forward_emulation_close_to_exact():
    w, X, basis, pi, ratio, train, target = _problem(seed=5)
    a = S.predict(w, X, basis, pi, ratio, 1.0, emulate=False)
    b = S.predict(w, X, basis, pi, ratio, 1.0, emulate=True)
    # bfloat16 weights and activations: about 2^-8 relative per rounding
    assert np.max(np.abs(a - b)) / np.max(np.abs(a)) < 0.05


def _silu(z):
    import math
    return z / (1.0 + math.exp(-z))


def test_forward_closed_form_one_pair():
    # a hand-built network on one pair whose output follows from the weights by hand: unit 0
    # of layer 1 reads r/h (feature 0), unit 1 reads u.e (feature 2) plus the constant feature
    # 16 (layer 1's bias), layers 2 and 3 pass units 0 and 1 on with weight 1 and biases
    # b2, b3, and the outputs are c0 = 2 h3[0] + b4[0], c1 = -h3[1], c2 = b4[2]:
    #   h1 = (silu(x0), silu(x2 + 0.3)),  h2 = silu(h1 + b2),  h3 = silu(h2 + b3),
    # a_i = a_scale rho_0/rho_i (c0 e + c1 u + c2 (u.e) e) for the pair's single message
    import math
    w = np.zeros(S.NPARAMS)
    p = S.unflatten(w)
    p["W1"][0, 0] = 1.0
    p["W1"][2, 1] = 1.0
    p["W1"][16, 1] = 0.3
    p["W2"][0, 0] = p["W2"][1, 1] = 1.0
    p["b2"][:2] = (0.1, -0.2)
    p["W3"][0, 0] = p["W3"][1, 1] = 1.0
    p["b3"][:2] = (-0.05, 0.25)
    p["W4"][0, 0] = 2.0
    p["W4"][1, 1] = -1.0
    p["b4"][:] = (0.07, 0.0, -0.4)
    w = S.flatten(p)
    # two particles 2 l0 apart in a periodic box, across the face; relative velocity
    lo, hi = np.zeros(3), np.full(3, 0.02)
    x = np.array([[0.0005, 0.01, 0.01], [0.0195, 0.0102, 0.0101]])   # j is 1e-3 to the -x side of i
    v = np.array([[0.01, 0.0, -0.02], [-0.03, 0.015, 0.0]])
    fields = {"position": x, "velocity": v, "vol_strain_p": np.array([0.01, -0.02]), "type": np.array([0, 0])}
    maxr, n0p, u_scale, dens_scale, a_scale = 3.5e-3, 0.95, 0.1, 50.0, 10.0
    pi, pj = np.array([0]), np.array([1])
    X, basis = S.features(fields, pi, pj, lo, hi, maxr, n0p, u_scale, dens_scale)
    a = S.predict(w, X, basis, pi, np.array([1.0, 1.0]), a_scale, emulate=False)
    # by hand
    d = np.array([-1.0e-3, 0.0002, 0.0001])                            # minimum image x_j - x_i
    r = math.sqrt((d * d).sum())
    e = d / r
    u = (v[1] - v[0]) / u_scale
    ue = float(u @ e)
    h1 = (_silu(r / maxr), _silu(ue + 0.3))
    h2 = (_silu(h1[0] + 0.1), _silu(h1[1] - 0.2))
    h3 = (_silu(h2[0] - 0.05), _silu(h2[1] + 0.25))
    c0, c1, c2 = 2.0 * h3[0] + 0.07, -h3[1], -0.4
    want = a_scale * (c0 * e + c1 * u + c2 * ue * e)
    assert np.allclose(X[0, [0, 1, 2, 16]], [r / maxr, (r / maxr) ** 2, ue, 1.0], rtol=1e-12)
    assert np.allclose(a[0], want, rtol=1e-12, atol=1e-15)
    assert np.array_equal(a[1], np.zeros(3))
