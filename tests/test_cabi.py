"""The C-ABI library without a GPU: symbols, host formats, error behaviour."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mphexplicit_amd import decks
from mphexplicit_amd.abi import MphFields, MphParams, MphStats, params_from_deck
from mphexplicit_amd.solver import LIB_PATH, MphError, Solver, load_library, read_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mph_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 20
    for nm in names:
        assert hasattr(lib, nm), nm


def test_struct_sizes_match_header_layout():
    # mph_params: 2 ints + 63 doubles (no padding after the two ints on x86-64)
    assert C.sizeof(MphParams) == 8 + 8 * (1 + 3 + 3 + 5 + 3 + 5 * 4 + 16 + 3 + 12 * 3)
    assert C.sizeof(MphFields) == 13 * 8
    # mph_stats: 3 long long, 4 int, eval_slots, list_grid and 4 bytes of tail padding
    assert C.sizeof(MphStats) == 3 * 8 + 4 * 4 + 8 + 4 + 4


def test_params_default_matches_reference_statics():
    lib = load_library()
    p = MphParams()
    lib.mph_params_default(C.byref(p))
    assert p.dt == 1.0e100 and p.dim == 3 and p.particle_spacing == 0.0


@pytest.mark.parametrize("factory", [decks.small_two_fluid_3d, decks.small_moving_wall_2d])
def test_data_and_grid_round_trip(tmp_path, factory):
    d = factory()
    dp, gp = d.write(str(tmp_path))
    p, t, x, v = read_case(dp, gp, dim=d.dim)
    ref = params_from_deck(d)
    assert bytes(p) == bytes(ref)
    assert np.array_equal(t, d.type) and np.array_equal(x, d.x) and np.array_equal(v, d.v)


def test_data_file_crlf_comments_and_unknown_lines(tmp_path):
    # the sample decks of the reference are CRLF with blank lines; '#' starts a comment
    d = decks.sample_bridge()
    dp, gp = d.write(str(tmp_path))
    text = open(dp).read().replace("\n", "\r\n")
    text = "#######\r\n" + text + "\r\n\r\nNotAKeyword 3\r\n   Dt 7.5e-6\r\n"
    crlf = os.path.join(str(tmp_path), "crlf.data")
    with open(crlf, "w", newline="") as f:
        f.write(text)
    p, _, _, _ = read_case(crlf, gp, dim=2)
    assert p.dt == 7.5e-6                       # later lines override, leading blanks allowed
    assert p.radius_ratio_p == 1.75 and list(p.gravity) == [0.0, -9.8, 0.0]
    assert p.surface_tension[0] == 0.072


def test_grid_errors(tmp_path):
    lib = load_library()
    p = MphParams()
    n = C.c_int(0)
    assert lib.mph_read_grid_header(b"/nonexistent/x.grid", C.byref(p), C.byref(n)) == -2
    assert b"cannot open" in lib.mph_last_error()
    bad = tmp_path / "short.grid"
    bad.write_text("0.0\n3 1e-3 0 1 0 1 0 1\n0 0.1 0.1 0 0 0 0\n")
    assert lib.mph_read_grid_header(str(bad).encode(), C.byref(p), C.byref(n)) == 0 and n.value == 3
    t = np.zeros(3, np.int32)
    x = np.zeros((3, 3))
    v = np.zeros((3, 3))
    rc = lib.mph_read_grid_particles(str(bad).encode(), 3, t.ctypes.data, x.ctypes.data, v.ctypes.data)
    assert rc == -3 and b"particle line 1" in lib.mph_last_error()
    assert lib.mph_read_data_file(b"/nonexistent/x.data", C.byref(p)) == -2


def test_prof_file_is_a_grid_file(tmp_path):
    # .prof has the .grid layout (main.cpp:672-696 vs 549-639), so it restarts a run
    lib = load_library()
    d = decks.small_droplet_2d()
    p = params_from_deck(d)
    path = str(tmp_path / "out.prof").encode()
    rc = lib.mph_write_prof_file(path, C.byref(p), 0.25, d.n, d.type.ctypes.data,
                                 np.ascontiguousarray(d.x).ctypes.data, np.ascontiguousarray(d.v).ctypes.data)
    assert rc == 0
    q, t, x, v = read_case(os.devnull, path.decode(), dim=2)
    assert q.time == 0.25
    assert np.array_equal(t, d.type)
    assert np.allclose(x, d.x, rtol=1e-6, atol=0)          # %e keeps 7 significant digits


def test_vtk_layout(tmp_path):
    lib = load_library()
    n = 5
    t = np.array([0, 1, 2, 3, 0], np.int32)
    x = np.arange(15, dtype=np.float64).reshape(5, 3) * 1e-3
    z = np.zeros(n)
    nc = np.arange(n, dtype=np.int32)
    f3 = np.ones((n, 3))
    path = str(tmp_path / "a.vtk")
    rc = lib.mph_write_vtk_file(path.encode(), n, t.ctypes.data, x.ctypes.data, f3.ctypes.data, z.ctypes.data,
                                z.ctypes.data, z.ctypes.data, nc.ctypes.data, f3.ctypes.data)
    assert rc == 0
    lines = open(path).read().split("\n")
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[4] == "POINTS 5 float"
    heads = [l for l in lines if l.startswith(("SCALARS", "VECTORS"))]
    assert heads == ["SCALARS label float 1", "SCALARS DensityA float 1", "SCALARS VolStrainP float 1",
                     "SCALARS VirialPressureAtParticle float 1", "SCALARS neighbor float 1",
                     "VECTORS velocity float", "VECTORS force float"]
    assert lines[5] == "%e %e %e" % (0.0, np.float32(1e-3), np.float32(2e-3))


def test_no_gpu_fails_loudly():
    lib = load_library()
    if lib.mph_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(MphError) as e:
        Solver.from_deck(decks.small_droplet_2d())
    assert "GPU" in str(e.value)
