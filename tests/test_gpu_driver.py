"""The command-line driver (mphexplicit_amd/mph_explicit) against the reference's own main().

The fixture tests/golden/driver_droplet2d.npz holds what the reference build wrote for a 2D
droplet with .prof output every 5 steps and .vtk output every 10 steps over 20 steps
(written by tests/golden/make_driver_golden.py).  The driver, given the same command line,
must write the same files at the same times, with the same particles: rows are matched by
position (both codes print %e), and positions and velocities must agree to the printed
precision (7 significant digits).  The graph-replay mode writes identical files.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.golden.make_driver_golden import driver_deck, read_prof

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


def test_driver_graph_mode_writes_identical_files(golden, tmp_path):
    a, b = tmp_path / "eager", tmp_path / "graph"
    a.mkdir()
    b.mkdir()
    _run(a)
    _run(b, ("-graph",))
    for name in sorted(p for p in os.listdir(a) if p.endswith(".prof") or p.endswith(".vtk")):
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
