"""Write the long-run fixtures of the measured workloads from the reference's own source.

The reference (source/main.cpp) is built under oracle/_ref by oracle/Makefile.  This script
runs it through oracle.oracle.Reference on two workloads of mphexplicit_amd.decks at their
full size and length and stores, every `every` steps, data only:

* ``config_droplet2d_80.npz``  the 2D square droplet of 80x80 particles over 2000 steps
  (``decks.droplet_2d()``): fluid momentum sum m v, sum m |v|, kinetic energy, centre of
  mass, and the list rebuilds of each interval;
* ``config_dambreak3d_64.npz`` the 3D dam break of 64x32x64 fluid particles in a
  192x32x96 walled box over 1000 steps (``decks.dam_break_3d()``): the same quantities.

Run from the repository root in the container that has the reference (minutes):
    make -C oracle && python tests/golden/make_config_golden.py [droplet2d_80|dambreak3d_64]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mphexplicit_amd import decks  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# name -> (deck factory, steps, sampling interval)
CONFIGS = {
    "droplet2d_80": (decks.droplet_2d, 2000, 100),
    "dambreak3d_64": (decks.dam_break_3d, 1000, 100),
}


def history_row(deck, fields):
    """[sum m v (3), sum m |v| (3), kinetic energy, centre of mass (3)] of the fluid"""
    t = np.asarray(fields["type"])
    fl = t < 2
    m = np.asarray(deck.data.density)[t[fl]] * deck.spacing ** deck.dim
    v = np.asarray(fields["v"])[fl]
    x = np.asarray(fields["x"])[fl]
    mom = (m[:, None] * v).sum(0)
    mabs = (m[:, None] * np.abs(v)).sum(0)
    ke = 0.5 * float((m * (v * v).sum(1)).sum())
    com = (m[:, None] * x).sum(0) / m.sum()
    return np.concatenate([mom, mabs, [ke], com])


def run(name):
    from oracle.oracle import Reference
    factory, steps, every = CONFIGS[name]
    d = factory()
    t0 = time.time()
    ref = Reference(d, nthreads=os.cpu_count() or 1)
    rows = [history_row(d, ref.fields())]
    rebuilds = []
    for k in range(steps // every):
        rebuilds.append(ref.step(every))
        rows.append(history_row(d, ref.fields()))
        print("%s step %d rebuilds %d (%.0f s)" % (name, (k + 1) * every, rebuilds[-1], time.time() - t0),
              flush=True)
    out = {"n": np.int64(d.n), "steps": np.int64(steps), "every": np.int64(every),
           "history": np.array(rows), "rebuilds": np.array(rebuilds, np.int64),
           "time": np.float64(ref.time),
           "columns": np.array(["Px", "Py", "Pz", "Ax", "Ay", "Az", "KE", "Cx", "Cy", "Cz"])}
    path = os.path.join(OUT, "config_%s.npz" % name)
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes", flush=True)


if __name__ == "__main__":
    for nm in (sys.argv[1:] or list(CONFIGS)):
        run(nm)
