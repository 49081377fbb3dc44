"""Write the fixture of the command-line driver test from the reference's own main().

Runs the reference build (oracle/_ref, compiled from source/main.cpp by oracle/Makefile) on
a small 2D droplet deck with .prof output every 5 steps and .vtk output every 10 steps over
20 steps, in a temporary directory, and stores in driver_droplet2d.npz (data only): the
names, times and particle rows of its .prof files, and the names and every array of its
.vtk files (points, label, DensityA, VolStrainP, VirialPressureAtParticle, neighbor,
velocity, force: the layout of main.cpp:698-834).

Run from the repository root in the container that has the reference:
    python tests/golden/make_driver_golden.py
"""
import ctypes
import glob
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mphexplicit_amd import decks  # noqa: E402


def driver_deck():
    d = decks.small_droplet_2d()
    dt = d.data.dt
    d.data.end_time = 20 * dt
    d.data.output_interval = 5 * dt
    d.data.vtk_output_interval = 10 * dt
    return d


def read_prof(path):
    with open(path) as f:
        t = float(f.readline())
        f.readline()
        rows = np.loadtxt(f, ndmin=2)
    return t, rows


VTK_SCALARS = ("label", "DensityA", "VolStrainP", "VirialPressureAtParticle", "neighbor")
VTK_VECTORS = ("velocity", "force")


def read_vtk(path):
    """Arrays of a legacy ASCII .vtk file in the layout of main.cpp:698-834:
    {"points": [n,3], "<scalar>": [n], "<vector>": [n,3]}"""
    with open(path) as f:
        tok = f.read().split("\n")
    out, i = {}, 0
    while i < len(tok):
        line = tok[i].split()
        if line[:1] == ["POINTS"]:
            n = int(line[1])
            out["points"] = np.array([[float(a) for a in tok[i + 1 + k].split()] for k in range(n)])
            i += 1 + n
            continue
        if line[:1] == ["SCALARS"]:
            n = len(out["points"])
            out[line[1]] = np.array([float(tok[i + 2 + k]) for k in range(n)])
            i += 2 + n
            continue
        if line[:1] == ["VECTORS"]:
            n = len(out["points"])
            out[line[1]] = np.array([[float(a) for a in tok[i + 1 + k].split()] for k in range(n)])
            i += 1 + n
            continue
        i += 1
    return out


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libmphref2d.so"))
    d = driver_deck()
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        dpath, gpath = d.write(tmp, "deck")
        args = [b"mph", dpath.encode(), gpath.encode(), b"prof%03d.prof", b"vtk%03d.vtk", b"log.txt", b"4"]
        argv = (ctypes.c_char_p * len(args))(*args)
        cwd = os.getcwd()
        os.chdir(tmp)
        try:
            rc = lib.mphref_main(len(args), argv)
        finally:
            os.chdir(cwd)
        assert rc == 0, rc
        profs = sorted(glob.glob(os.path.join(tmp, "prof*.prof")))
        vtks = sorted(os.path.basename(p) for p in glob.glob(os.path.join(tmp, "*.vtk")))
        out["prof_names"] = np.array([os.path.basename(p) for p in profs])
        out["vtk_names"] = np.array(vtks)
        times = []
        for k, p in enumerate(profs):
            t, rows = read_prof(p)
            times.append(t)
            out["prof_%d" % k] = rows
        out["prof_times"] = np.array(times)
        for k, name in enumerate(vtks):
            for field, arr in read_vtk(os.path.join(tmp, name)).items():
                out["vtk_%d_%s" % (k, field)] = arr
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "driver_droplet2d.npz"), **out)
    print("wrote", len(out["prof_names"]), "prof and", len(out["vtk_names"]), "vtk names")


if __name__ == "__main__":
    main()
