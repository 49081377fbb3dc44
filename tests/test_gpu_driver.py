"""The command-line driver (mphexplicit_amd/mph_explicit) against the reference's own main().

The fixture tests/golden/driver_droplet2d.npz holds what the reference build wrote for a 2D
droplet with .prof output every 5 steps and .vtk output every 10 steps over 20 steps
(written by tests/golden/make_driver_golden.py).  The driver, given the same command line,
must write the same files at the same times, with the same particles: rows are matched by
position (both codes print %e), and positions and velocities must agree to the printed
precision (7 significant digits).  Every .vtk file (main.cpp:698-834) is compared array by
array: points, label and neighbor count exactly after matching the particles by position,
DensityA, VolStrainP, VirialPressureAtParticle, velocity and force to the printed float
precision.  The graph-replay mode writes identical files.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.golden.make_driver_golden import VTK_SCALARS, VTK_VECTORS, driver_deck, read_prof, read_vtk

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mphexplicit_amd", "mph_explicit")
GOLDEN = os.path.join(ROOT, "tests", "golden", "driver_droplet2d.npz")


def _run(tmp, extra=()):
    d = driver_deck()
    dpath, gpath = d.write(str(tmp), "deck")
    cmd = [DRIVER, dpath, gpath, "prof%03d.prof", "vtk%03d.vtk", "log.txt", "4", "-2d", *extra]
    print(" ".join(cmd), flush=True)
    # the eager and graph runs are compared bit for bit: pin the cluster size
    env = dict(os.environ, MPH_TUNE="4")
    r = subprocess.run(cmd, cwd=str(tmp), capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    return d


@pytest.fixture(scope="module")
def golden(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    if not os.path.exists(DRIVER):
        pytest.fail("driver not built: python -m mphexplicit_amd.build")
    return np.load(GOLDEN, allow_pickle=False)


def _match(rows_g, rows_r, scale_x, scale_v):
    """largest position and velocity difference after matching every row of the driver's
    file to the reference row nearest in position (same type)"""
    assert rows_g.shape == rows_r.shape
    assert np.array_equal(np.sort(rows_g[:, 0]), np.sort(rows_r[:, 0]))
    dx = dv = 0.0
    used = np.zeros(len(rows_r), bool)
    for row in rows_g:
        d = np.abs(rows_r[:, 1:4] - row[1:4]).max(1)
        d[rows_r[:, 0] != row[0]] = np.inf
        k = int(np.argmin(d))
        assert not used[k], "two rows matched the same reference particle"
        used[k] = True
        dx = max(dx, d[k])
        dv = max(dv, np.abs(rows_r[k, 4:7] - row[4:7]).max())
    return dx / scale_x, dv / scale_v


def test_driver_matches_reference_main(golden, tmp_path):
    _run(tmp_path)
    names = sorted(p for p in os.listdir(tmp_path) if p.endswith(".prof"))
    assert names == list(golden["prof_names"])
    vtks = sorted(p for p in os.listdir(tmp_path) if p.endswith(".vtk"))
    assert vtks == list(golden["vtk_names"])
    for k, name in enumerate(names):
        t, rows = read_prof(os.path.join(tmp_path, name))
        assert t == float(golden["prof_times"][k]), name
        ref = golden["prof_%d" % k]
        sx = np.abs(ref[:, 1:4]).max()
        sv = max(np.abs(ref[:, 4:7]).max(), 1e-30)
        ex, ev = _match(rows, ref, sx, sv)
        print(name, "position %.1e velocity %.1e (relative to the field maximum)" % (ex, ev), flush=True)
        # %e keeps 7 significant digits: one unit in the last place is 1e-6 relative
        assert ex <= 2e-6, name
        assert ev <= 2e-6 or k == 0, name
    with open(os.path.join(tmp_path, "vtk010.vtk")) as f:
        head = [next(f) for _ in range(5)]
    assert head[0].startswith("# vtk DataFile") and head[4].startswith("POINTS 400")


def test_driver_vtk_arrays_match_reference(golden, tmp_path):
    _run(tmp_path)
    names = list(golden["vtk_names"])
    for k, name in enumerate(names):
        g = read_vtk(os.path.join(tmp_path, name))
        r = {f: golden["vtk_%d_%s" % (k, f)] for f in ("points",) + VTK_SCALARS + VTK_VECTORS}
        n = len(r["points"])
        assert len(g["points"]) == n, name
        # match every particle of the driver's file to the reference's particle at the same
        # printed position (both print (float) x with %e; the lattice positions are distinct)
        key = lambda p: [tuple(row) for row in p]
        at = {t: i for i, t in enumerate(key(r["points"]))}
        assert len(at) == n, "reference positions are not distinct in " + name
        perm = np.array([at.get(t, -1) for t in key(g["points"])])
        assert (perm >= 0).all() and len(set(perm.tolist())) == n, name + ": particles do not match by position"
        assert np.array_equal(g["label"], r["label"][perm]), name
        assert np.array_equal(g["neighbor"], r["neighbor"][perm]), name
        errs = {}
        for f in ("DensityA", "VolStrainP", "VirialPressureAtParticle", "velocity", "force"):
            if name == "output.vtk" and f == "VirialPressureAtParticle":
                continue        # main.cpp:365 prints the virial before it is ever computed
            a, b = g[f], r[f][perm]
            den = np.abs(b).max()
            errs[f] = float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a).max())
        print(name, " ".join("%s=%.1e" % kv for kv in errs.items()), flush=True)
        # both print (float) of doubles that agree to about 1e-13: one float rounding and the
        # 7 printed digits, 2e-6 of the field's largest value
        bad = {f: e for f, e in errs.items() if not e <= 2e-6}
        assert not bad, (name, bad)


def test_driver_graph_mode_writes_identical_files(golden, tmp_path):
    a, b = tmp_path / "eager", tmp_path / "graph"
    a.mkdir()
    b.mkdir()
    _run(a)
    _run(b, ("-graph",))
    for name in sorted(p for p in os.listdir(a) if p.endswith(".prof") or p.endswith(".vtk")):
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
