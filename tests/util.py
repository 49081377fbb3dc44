"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np

# Field names of the oracle / reference (oracle.oracle.FIELDS) -> mph_fields names
ORACLE_TO_ABI = {
    "index": "particle_index", "type": "type", "x": "position", "v": "velocity",
    "force": "force", "dens_a": "density_a", "vol_strain": "vol_strain_p",
    "divergence": "divergence_p", "press_p": "pressure_p", "press_a": "pressure_a",
    "gcenter": "gravity_center", "ncount": "neighbor_count",
}

# relative normwise tolerance of the FP64 GPU path against the oracle, per field
FP64_TOL = 1.0e-12


def normwise(a, b) -> float:
    """max_i |a_i - b_i| / max_i |b_i|  (0 if both are identically zero)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    if den == 0.0:
        return 0.0 if num == 0.0 else np.inf
    return float(num / den)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def csr_rows(ptr, ind):
    return [tuple(ind[ptr[i]:ptr[i + 1]]) for i in range(len(ptr) - 1)]


def oracle_lists_by_id(fields, ptr, ind):
    """Reference-style lists (current order) -> rows by original id, each sorted ascending."""
    idx = np.asarray(fields["index"])
    n = len(idx)
    rows = [None] * n
    for i in range(n):
        rows[idx[i]] = tuple(sorted(int(idx[j]) for j in ind[ptr[i]:ptr[i + 1]]))
    return rows


COMPARED = ("x", "v", "force", "dens_a", "vol_strain", "divergence", "press_p", "press_a", "gcenter")


def pressure_scale(deck):
    """K_max * N0p: the pressure of a unit volumetric strain.  P = -lambda D + kappa e^P
    (main.cpp:1905) is a multiple of the strain sum, whose rounding error is a multiple of
    N0p, so this is the scale that FP64 differences in P are measured against when the
    pressure field itself is at rounding level (a lattice deck at its first step)."""
    from oracle.oracle import Oracle
    return max(deck.data.bulk_modulus) * Oracle(deck).scalars()[1]


def fields_for(deck, names=COMPARED):
    """GravityCenter is only computed when some surface tension is non-zero: without surface
    tension it enters no force (CofA = 0) and is not an output field."""
    if not any(t != 0.0 for t in deck.data.surface_tension):
        names = tuple(k for k in names if k != "gcenter")
    return names


def rest_scales(deck):
    """Scales of a deck whose pressure is at rounding level (a lattice at rest, step 1):
    the pressure of a unit volumetric strain, K*N0p (see pressure_scale), the force that
    pressure exerts across a particle's face, K*N0p*l0^(d-1), and the velocity that force
    gives the lightest particle in one step.  The pressure forces of such a state are
    rounding noise of P, so the force and velocity differences are measured against these
    (the largest force is then gravity alone, a far smaller number)."""
    p = pressure_scale(deck)
    f = p * deck.spacing ** (deck.dim - 1)
    m = min(deck.data.density) * deck.spacing ** deck.dim
    return p, f, f / m * deck.data.dt


def compare_fields(gf, of, label, names=COMPARED, tol=FP64_TOL, p_scale=0.0, f_scale=0.0, v_scale=0.0):
    """normwise error of every named field (oracle names) of the GPU fields gf against the
    oracle's of; asserts each is within tol.  Where the oracle's pressure is at rounding level
    (below 1e-9 of p_scale), pressure, force and velocity are measured at the scales of
    rest_scales instead (those given as non-zero)."""
    errs = {}
    at_rest = "press_p" in of and p_scale > 0.0 and np.max(np.abs(of["press_p"])) < 1e-9 * p_scale
    scale = {"press_p": p_scale, "force": f_scale, "v": v_scale}
    for k in names:
        a, b = gf[ORACLE_TO_ABI[k]], of[k]
        if at_rest and scale.get(k, 0.0) > 0.0:
            errs[k] = float(np.max(np.abs(a - b)) / scale[k])
        else:
            errs[k] = normwise(a, b)
    print(label, " ".join("%s=%.2e" % kv for kv in errs.items()), flush=True)
    bad = {k: e for k, e in errs.items() if not e <= tol}
    assert not bad, "%s: fields beyond %g: %s" % (label, tol, bad)
    return errs


def pair_keys_by_id(index, ptr, ind):
    """sorted int64 keys id_i * n + id_j of every list entry (rows in any order, entries by
    current index) -- an exact, order-free form of a set of neighbour lists"""
    index = np.asarray(index, np.int64)
    n = len(index)
    counts = np.diff(np.asarray(ptr, np.int64))
    keys = np.repeat(index, counts) * n + index[np.asarray(ind, np.int64)]
    keys.sort()
    return keys


def csr_keys(ptr, ind):
    """the same keys for CSR lists that are already keyed by original id"""
    n = len(ptr) - 1
    counts = np.diff(np.asarray(ptr, np.int64))
    return np.repeat(np.arange(n, dtype=np.int64), counts) * n + np.asarray(ind, np.int64)
