"""CPU FP64 oracle of the pairwise-message surrogate -- test infrastructure only.

Only tests/ may import this module; the product path (mphexplicit_amd) never does.

It restates the network of the HIP kernels (mphexplicit_amd/csrc/mph_surrogate.h) in numpy
FP64: the 17 pair features (symmetric under i <-> j), three hidden SiLU layers of 64, three
coefficients of the antisymmetric message m = c0 e + c1 u + c2 (u.e) e, the sum over the
pairs of a particle, the loss and its analytic gradient.  With ``emulate=True`` it
rounds where the GPU rounds (weights of W1..W4 and every matrix-core input to bfloat16,
round-to-nearest-even); with ``emulate=False`` it is the exact FP64 function, which the
central finite-difference check pins.

The network has no counterpart in the reference; the features use the reference's
periodic difference (source/main.cpp:70, 1680), its neighbour set (main.cpp:1408: r <=
MaxRadius) and the volumetric strain VolStrainP = sum w_P - N0p (main.cpp:1850-1856, 948).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 17, 64, 3
SHAPES = [("W1", (17, 64)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (64, 3)), ("b4", (3,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAMS == 9603


def unflatten(w):
    out, o = {}, 0
    for k, s in SHAPES:
        m = int(np.prod(s))
        out[k] = np.asarray(w[o:o + m], np.float64).reshape(s)
        o += m
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[k], np.float64).ravel() for k, _ in SHAPES])


def bf16(x):
    """FP64 -> FP32 -> bfloat16 (round to nearest even), returned as FP64"""
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(x))


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def periodic_diff(xi, xj, lo, hi):
    """x_j - x_i per axis as main.cpp:1680: Mod(x_j - x_i + W/2, W) - W/2"""
    w = hi - lo
    hw = 0.5 * w
    a = xj - xi + hw
    return (a - w * np.floor(a / w)) - hw


def pairs_within(fields, ptr, ind, lo, hi, maxr):
    """(i, j) for fluid i and j != i in i's neighbour list with r <= MaxRadius (original ids)"""
    x, t = fields["position"], fields["type"]
    ii, jj = [], []
    for i in range(len(ptr) - 1):
        if t[i] >= 2:
            continue
        js = ind[ptr[i]:ptr[i + 1]]
        js = js[js != i]
        d = periodic_diff(x[i][None, :], x[js], lo, hi)
        q2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        keep = js[q2 <= maxr * maxr]
        ii.append(np.full(len(keep), i))
        jj.append(keep)
    return np.concatenate(ii).astype(np.int64), np.concatenate(jj).astype(np.int64)


def features(fields, pi, pj, lo, hi, maxr, n0p, u_scale, dens_scale):
    """features of each pair (i, j), all symmetric under i <-> j, and the message basis.

    X [P][17]: r/h, (r/h)^2, u.e, |u|^2, one-hot of the unordered class pair (00, 01, 02,
    11, 12, 22; class = type 0, type 1, wall), e_i + e_j, (e_i - e_j)^2, e_i e_j (e = dens_scale
    VolStrainP / N0p), three zeros and the constant 1.  Basis: e = x_ij / r, u = (v_j - v_i) /
    u_scale and u.e, all antisymmetric except u.e, so m_ij = c0 e + c1 u + c2 (u.e) e is
    antisymmetric."""
    x, v, n, t = fields["position"], fields["velocity"], fields["vol_strain_p"], fields["type"]
    d = periodic_diff(x[pi], x[pj], lo, hi)
    r = np.sqrt((d * d).sum(1))
    e = d / r[:, None]
    u = (v[pj] - v[pi]) / u_scale
    ue = (u * e).sum(1)
    cls = np.minimum(t, 2)
    a, b = np.minimum(cls[pi], cls[pj]), np.maximum(cls[pi], cls[pj])
    cp = np.where(a == 0, b, np.where(a == 1, 2 + b, 5))
    ei, ej = n[pi] / n0p * dens_scale, n[pj] / n0p * dens_scale
    X = np.zeros((len(pi), N_IN))
    X[:, 0] = r / maxr
    X[:, 1] = (r / maxr) ** 2
    X[:, 2] = ue
    X[:, 3] = (u * u).sum(1)
    X[np.arange(len(pi)), 4 + cp] = 1.0
    X[:, 10] = ei + ej
    X[:, 11] = (ei - ej) ** 2
    X[:, 12] = ei * ej
    X[:, 16] = 1.0
    return X, dict(e=e, u=u, ue=ue)


def mass_ratio(fields, density):
    """rho_0 / rho_i of every particle (the messages are forces per unit volume)"""
    rho = np.asarray(density, np.float64)
    return rho[0] / rho[fields["type"]]


def messages(C, basis):
    """m = c0 e + c1 u + c2 (u.e) e from the network's outputs C [P][3]"""
    return ((C[:, 0] + C[:, 2] * basis["ue"])[:, None] * basis["e"] + C[:, 1][:, None] * basis["u"])


def forward(w, X, emulate=True):
    """per-pair messages and the stored activations"""
    p = unflatten(w)
    rw = bf16 if emulate else (lambda a: a)
    W1, W2, W3, W4 = rw(p["W1"]), rw(p["W2"]), rw(p["W3"]), rw(p["W4"])
    Xr = rw(X)
    Z1 = Xr @ W1
    H1 = rw(silu(Z1))
    Z2 = H1 @ W2 + p["b2"]
    H2 = rw(silu(Z2))
    Z3 = H2 @ W3 + p["b3"]
    H3 = rw(silu(Z3))
    M = H3 @ W4 + p["b4"]
    return M, dict(X=Xr, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3, H3=H3, W1=W1, W2=W2, W3=W3, W4=W4)


def predict(w, X, basis, pi, ratio, a_scale, emulate=True):
    """a_i = a_scale rho_0/rho_i sum_j m_ij, [n][3] (zero for particles without pairs)"""
    C, _ = forward(w, X, emulate)
    a = np.zeros((len(ratio), 3))
    np.add.at(a, pi, messages(C, basis))
    return a_scale * ratio[:, None] * a


def loss_and_grad(w, X, basis, pi, ratio, train, target, a_scale, emulate=True):
    """L = mean over the training particles of |a - target|^2 / a_scale^2 and dL/dw.

    train: original ids of the training particles; target [n][3]; ratio = rho_0/rho_i.
    With emulate, the inputs of the weight-gradient products are rounded to bfloat16 as the
    matrix-core reduction does (the FP32 data path is kept in FP64)."""
    n = target.shape[0]
    C, c = forward(w, X, emulate)
    s = a_scale * ratio
    a = np.zeros((n, 3))
    np.add.at(a, pi, messages(C, basis))
    a *= s[:, None]
    mask = np.zeros(n, bool)
    mask[train] = True
    N = len(train)
    d = (a - target) / a_scale
    loss = float((d[mask] ** 2).sum() / N)
    dA = np.where(mask[:, None], 2.0 * d / (N * a_scale), 0.0)      # dL/da
    g = (dA * s[:, None])[pi]                                        # dL/dm per pair
    ge = (g * basis["e"]).sum(1)
    dM = np.stack([ge, (g * basis["u"]).sum(1), ge * basis["ue"]], 1)  # dL/dc
    rw = bf16 if emulate else (lambda q: q)
    dz3 = (dM @ c["W4"].T) * dsilu(c["Z3"])
    dz2 = (dz3 @ c["W3"].T) * dsilu(c["Z2"])
    dz1 = (dz2 @ c["W2"].T) * dsilu(c["Z1"])
    grad = {
        "W4": rw(c["H3"]).T @ rw(dM), "b4": rw(dM).sum(0),
        "W3": rw(c["H2"]).T @ rw(dz3), "b3": rw(dz3).sum(0),
        "W2": rw(c["H1"]).T @ rw(dz2), "b2": rw(dz2).sum(0),
        "W1": c["X"].T @ rw(dz1),
    }
    return loss, flatten(grad)
