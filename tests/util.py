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
