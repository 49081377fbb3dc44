"""Write the golden fixtures of tests/golden from the reference's own source.

The reference (source/main.cpp) is built under oracle/_ref by oracle/Makefile; this script
runs it through oracle.oracle.Reference on the small decks of mphexplicit_amd.decks and
stores inputs-derived outputs as compressed .npz files (fields keyed by original particle
id, neighbour lists by id, momentum histories).  Run from the repo root:

    make -C oracle && python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mphexplicit_amd import decks  # noqa: E402
from oracle.oracle import Reference, by_id  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("x", "v", "force", "dens_a", "vol_strain", "divergence", "press_p", "press_a",
          "gcenter", "ncount", "type")

# name -> (deck factory, store every particle?, momentum steps)
CASES = {
    "droplet2d": (decks.small_droplet_2d, True, 100),
    "movingwall2d": (decks.small_moving_wall_2d, True, 50),
    "bridge2d": (decks.sample_bridge, False, 20),
    "dambreak3d": (decks.small_dam_break_3d, False, 10),
    "twofluid3d": (decks.small_two_fluid_3d, False, 10),
    "rotatingwall3d": (decks.small_rotating_wall_3d, True, 50),
}
SAMPLE = 200


def lists_by_id(fields, ptr, ind):
    idx = fields["index"]
    n = len(idx)
    rows = [None] * n
    for i in range(n):
        rows[idx[i]] = np.sort(idx[ind[ptr[i]:ptr[i + 1]]])
    lens = np.array([len(r) for r in rows], np.int64)
    p = np.concatenate([[0], np.cumsum(lens)])
    return p, np.concatenate(rows).astype(np.int32)


def momentum(fields):
    m = fields["mass"][:, None]
    return (m * fields["v"]).sum(axis=0)


def main(names):
    for name in names or list(CASES):
        factory, full, msteps = CASES[name]
        d = factory()
        ref = Reference(d, nthreads=os.cpu_count() or 1)
        rng = np.random.default_rng(2024)
        ids = np.arange(d.n) if full else np.sort(rng.choice(d.n, size=min(SAMPLE, d.n), replace=False))
        out = {"sample_ids": ids.astype(np.int32), "n": np.int64(d.n)}
        f0 = ref.fields()
        fb = by_id(f0)
        for k in ("x", "dens_a", "vol_strain", "gcenter", "ncount", "type"):
            out["init_" + k] = fb[k][ids]
        ptr, ind = ref.neighbors()
        if full:
            p, q = lists_by_id(f0, ptr, ind)
            out["init_nbr_ptr"], out["init_nbr_ind"] = p, q
        out["init_nbr_total"] = np.int64(len(ind))
        mom = [momentum(f0)]
        snaps = {1: None, msteps: None}
        for s in range(1, msteps + 1):
            ref.step(1)
            f = ref.fields()
            mom.append(momentum(f))
            if s in snaps:
                fb = by_id(f)
                for k in FIELDS:
                    out["step%d_%s" % (s, k)] = fb[k][ids]
                out["step%d_time" % s] = np.float64(ref.time)
        out["momentum"] = np.array(mom)
        out["mass_total"] = np.float64(f0["mass"].sum())
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **out)
        print(name, d.n, os.path.getsize(path), "bytes", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
