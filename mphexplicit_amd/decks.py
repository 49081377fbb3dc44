"""Seeded input decks in the reference's file formats.

The reference reads a ``.data`` keyword file (source/main.cpp:489-539) and a ``.grid``
particle file (source/main.cpp:541-670).  This module writes both, for

* the workloads the solver is measured and tested on (2D square droplet, 3D dam break,
  3D two-fluid dam break at ~1M fluid particles, batches of droplets/bridges), and
* small parity decks, including the two sample cases shipped with the reference
  (results/Bridge_l00100_deg000_re35_mu01_g98 and results/SquareDropSkip_l00400_rv35_rp17),
  whose cuboid lists are reproduced here as data and expanded into particles with the
  lattice rule of generator/generator.cpp:170-193.

Positions are written with ``%.17g`` so that the reference's ``sscanf("%lf")`` and our
readers obtain identical doubles.  Randomness (lattice jitter, batch sampling) comes from
``numpy.random.default_rng(seed)`` (PCG64), so every deck is a pure function of its
arguments.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

FLUID_TYPES = (0, 1)
WALL_TYPES = (2, 3)


@dataclass
class DataParams:
    """Keywords of the reference's ``.data`` file (source/main.cpp:502-521)."""
    dt: float = 1.0e-4
    output_interval: float = 1.0e30
    vtk_output_interval: float = 1.0e30
    end_time: float = 1.0
    radius_ratio_a: float = 3.5
    radius_ratio_p: float = 3.5
    radius_ratio_v: float = 3.5
    density: Sequence[float] = (1000.0,) * 4
    bulk_modulus: Sequence[float] = (1.0e4,) * 4
    bulk_viscosity: Sequence[float] = (0.1,) * 4
    shear_viscosity: Sequence[float] = (0.1,) * 4
    surface_tension: Sequence[float] = (0.0,) * 4
    interaction_ratio: Sequence[Sequence[float]] = ((1.0,) * 4,) * 4
    gravity: Sequence[float] = (0.0, 0.0, 0.0)
    wall_center: Sequence[Sequence[float]] = ((0.0,) * 3,) * 4
    wall_velocity: Sequence[Sequence[float]] = ((0.0,) * 3,) * 4
    wall_omega: Sequence[Sequence[float]] = ((0.0,) * 3,) * 4


@dataclass
class Deck:
    """One case: parameters, particles and the simulated dimension."""
    name: str
    dim: int
    data: DataParams
    spacing: float
    domain_min: np.ndarray
    domain_max: np.ndarray
    type: np.ndarray            # int32 [n]
    x: np.ndarray               # float64 [n,3]
    v: np.ndarray               # float64 [n,3]
    time: float = 0.0
    steps: int = 1
    notes: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    def write(self, directory: str, stem: str | None = None) -> tuple[str, str]:
        stem = stem or self.name
        os.makedirs(directory, exist_ok=True)
        dpath = os.path.join(directory, stem + ".data")
        gpath = os.path.join(directory, stem + ".grid")
        write_data(dpath, self.data)
        write_grid(gpath, self)
        return dpath, gpath


def _g(x: float) -> str:
    return "%.17g" % float(x)


def write_data(path: str, p: DataParams) -> None:
    """Write a ``.data`` file in the keyword layout that source/main.cpp:502-521 parses."""
    four = lambda k, a: "%s %s\n" % (k, " ".join(_g(t) for t in a))
    lines = [
        "# written by mphexplicit_amd.decks\n",
        "Dt %s\n" % _g(p.dt),
        "OutputInterval %s\n" % _g(p.output_interval),
        "VtkOutputInterval %s\n" % _g(p.vtk_output_interval),
        "EndTime %s\n" % _g(p.end_time),
        "RadiusRatioA %s\n" % _g(p.radius_ratio_a),
        "RadiusRatioP %s\n" % _g(p.radius_ratio_p),
        "RadiusRatioV %s\n" % _g(p.radius_ratio_v),
        four("Density", p.density),
        four("BulkModulus", p.bulk_modulus),
        four("BulkViscosity", p.bulk_viscosity),
        four("ShearViscosity", p.shear_viscosity),
        four("SurfaceTension", p.surface_tension),
    ]
    for t in range(4):
        lines.append(four("InteractionRatio(Type%d)" % t, p.interaction_ratio[t]))
    lines.append("Gravity %s\n" % " ".join(_g(g) for g in p.gravity))
    for w in (2, 3):
        lines.append("Wall%d  Center %s Velocity %s Omega %s\n" % (
            w, " ".join(_g(c) for c in p.wall_center[w]),
            " ".join(_g(c) for c in p.wall_velocity[w]),
            " ".join(_g(c) for c in p.wall_omega[w])))
    with open(path, "w") as f:
        f.writelines(lines)


def write_grid(path: str, deck: Deck) -> None:
    """Write a ``.grid`` file (layout of source/main.cpp:549-639) with exact doubles."""
    n = deck.n
    lo, hi = deck.domain_min, deck.domain_max
    with open(path, "w") as f:
        f.write("%s\n" % _g(deck.time))
        f.write("%d %s %s %s %s %s %s %s\n" % (n, _g(deck.spacing), _g(lo[0]), _g(hi[0]),
                                               _g(lo[1]), _g(hi[1]), _g(lo[2]), _g(hi[2])))
        rows = ["%d %s %s %s %s %s %s\n" % (int(t), _g(a), _g(b), _g(c), _g(d), _g(e), _g(h))
                for t, (a, b, c), (d, e, h) in zip(deck.type, deck.x, deck.v)]
        f.writelines(rows)


# --------------------------------------------------------------------------------------
# lattice helpers


def _lattice(counts, spacing, origin):
    """Particle centres ``origin + (i+0.5)*spacing`` on an ``counts`` lattice (x slowest)."""
    axes = [origin[d] + (np.arange(counts[d]) + 0.5) * spacing for d in range(3)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([a.reshape(-1) for a in g], axis=1)


def _jitter(rng, x, amp, dim):
    j = rng.uniform(-amp, amp, size=x.shape)
    if dim == 2:
        j[:, 2] = 0.0
    return x + j


def _wall_shell(interior_counts, layers, spacing, axes_walled, dim):
    """Wall lattice: ``layers`` particle layers outside the interior box on the walled axes.

    The interior occupies cells [0, n_d) on each axis; walls fill [-layers, n_d+layers) on
    walled axes, minus the interior.  Un-walled axes keep [0, n_d) (periodic).
    """
    lo = [(-layers if d in axes_walled else 0) for d in range(3)]
    hi = [(interior_counts[d] + layers if d in axes_walled else interior_counts[d]) for d in range(3)]
    if dim == 2:
        lo[2], hi[2] = 0, 1
    idx = np.stack(np.meshgrid(*[np.arange(lo[d], hi[d]) for d in range(3)], indexing="ij"), -1).reshape(-1, 3)
    inside = np.ones(len(idx), dtype=bool)
    for d in axes_walled:
        inside &= (idx[:, d] >= 0) & (idx[:, d] < interior_counts[d])
    idx = idx[~inside]
    x = (idx + 0.5) * spacing
    if dim == 2:
        x[:, 2] = 0.0
    return x


# --------------------------------------------------------------------------------------
# workloads


def sound_speed_params(rho, c, l0, cfl=0.2):
    """Bulk modulus K = rho c^2 and dt = cfl * l0 / c for a chosen sound speed c."""
    return rho * c * c, cfl * l0 / c


def droplet_2d(n_side=80, l0=1.0e-3, jitter=0.01, seed=1, mu=0.1, sigma=0.07, rho=1000.0,
               ratio=3.5, steps=2000, v_max=0.1, name=None) -> Deck:
    """2D square droplet of ``n_side``^2 type-0 particles, no walls, no gravity.

    Sound speed is 10x ``v_max``; dt from CFL 0.2 (then checked against the viscous limit
    0.125 rho l0^2 / mu).
    """
    rng = np.random.default_rng(seed)
    c = 10.0 * v_max
    K, dt = sound_speed_params(rho, c, l0)
    dt = min(dt, 0.125 * rho * l0 * l0 / mu)
    half = 0.5 * n_side * l0
    x = _lattice((n_side, n_side, 1), l0, (-half, -half, 0.0))
    x[:, 2] = 0.0
    x = _jitter(rng, x, jitter * l0, 2)
    margin = max(0.02, 2 * 3.6 * l0 + 0.5 * half)
    L = half + margin
    data = DataParams(dt=dt, end_time=dt * steps, radius_ratio_a=ratio, radius_ratio_p=ratio,
                      radius_ratio_v=ratio, density=(rho,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(mu,) * 4, shear_viscosity=(mu,) * 4,
                      surface_tension=(sigma,) * 4)
    n = x.shape[0]
    return Deck(name or "droplet2d_%d" % n_side, 2, data, l0, np.array([-L, -L, -5e-4]),
                np.array([L, L, 5e-4]), np.zeros(n, np.int32), x, np.zeros((n, 3)), steps=steps)


def dam_break_3d(fluid=(64, 32, 64), box=(192, 32, 96), layers=3, l0=1.0e-3, jitter=0.0, seed=2,
                 two_fluid=False, mu=(1.0e-3, 5.0e-3), rho=(1000.0, 800.0), bulk_visc=2.0,
                 steps=1000, sigma=0.0, ratio=3.5, pad=4, name=None) -> Deck:
    """3D dam break: a fluid block in the (x,z) corner of a walled box, gravity -9.8 on z.

    Walls are ``layers`` layers of type-2 particles on the x and z faces; y is periodic
    with width box[1]*l0 (the fluid spans it).  Sound speed is 10*sqrt(2 g H) with H the
    fluid height; K = rho c^2 per fluid type, dt = 0.2 l0 / c.  With ``two_fluid`` the
    upper half (z) is type 1 (density rho[1], viscosity mu[1]).
    """
    rng = np.random.default_rng(seed)
    H = fluid[2] * l0
    c = 10.0 * math.sqrt(2.0 * 9.8 * H)
    dt = 0.2 * l0 / c
    xf = _lattice(fluid, l0, (0.0, 0.0, 0.0))
    tf = np.zeros(len(xf), np.int32)
    if two_fluid:
        tf[xf[:, 2] >= 0.5 * H] = 1
    if jitter > 0:
        xf = _jitter(rng, xf, jitter * l0, 3)
    xw = _wall_shell(box, layers, l0, (0, 2), 3)
    x = np.concatenate([xf, xw])
    t = np.concatenate([tf, np.full(len(xw), 2, np.int32)])
    r0, r1 = rho
    dens = (r0, r1, r0, r0)
    K = tuple(d * c * c for d in dens)
    visc = (mu[0], mu[1], mu[0], mu[0])
    data = DataParams(dt=dt, end_time=dt * steps, radius_ratio_a=ratio, radius_ratio_p=ratio,
                      radius_ratio_v=ratio, density=dens, bulk_modulus=K,
                      bulk_viscosity=(bulk_visc,) * 4, shear_viscosity=visc,
                      surface_tension=(sigma,) * 4, gravity=(0.0, 0.0, -9.8))
    lo = np.array([-(layers + pad) * l0, 0.0, -(layers + pad) * l0])
    hi = np.array([(box[0] + layers + pad) * l0, box[1] * l0, (box[2] + layers + pad) * l0])
    nm = name or ("dambreak3d_%dx%dx%d" % fluid)
    return Deck(nm, 3, data, l0, lo, hi, t, x, np.zeros((len(x), 3)), steps=steps,
                notes={"fluid": int(len(xf)), "wall": int(len(xw)), "sound_speed": c})


def dam_break_1m(**kw) -> Deck:
    """~1M-fluid-particle two-fluid 3D dam break: 128x64x128 fluid, 384x64x192 box."""
    args = dict(fluid=(128, 64, 128), box=(384, 64, 192), jitter=0.01, seed=3, two_fluid=True,
                steps=500, name="dambreak3d_1m")
    args.update(kw)
    return dam_break_3d(**args)


# --------------------------------------------------------------------------------------
# .boid -> .grid, as the reference's generator does it (generator/generator.cpp)


@dataclass
class Boid:
    """Contents of a ``.boid`` file (generator/generator.cpp:49-165): the domain, the particle
    distance written into the .grid header, and cuboids (spacing, type, lower, upper,
    velocity) filled with particles."""
    particle_distance: float
    lower: tuple
    upper: tuple
    cuboids: list


def parse_boid(text: str) -> Boid:
    """Parse ``.boid`` text the way generator/generator.cpp:49-165 reads it: top-level lines
    whose first word is ParticleDistance, LowerDomain, UpperDomain or StartCuboid ('#' lines
    skipped, other lines ignored); after StartCuboid, whitespace-separated tokens (across
    lines) Spacing, Type, Lower, Upper, Velocity until EndCuboid, whose line's remainder is
    read as the next line.  Raises ValueError where the generator reports an error."""
    pd, lo, hi, cubs = -1.0, [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], []
    lines = text.splitlines(keepends=True)
    k = 0
    pending = None                     # remainder of the line that ended a cuboid
    while pending is not None or k < len(lines):
        if pending is not None:
            buf, pending = pending, None
        else:
            buf, k = lines[k], k + 1
        if buf.startswith("#"):
            continue
        w = buf.split()
        if not w:
            continue
        if w[0] == "ParticleDistance":
            if len(w) < 2:
                raise ValueError("ParticleDistance count not 1")
            pd = float(w[1])
        elif w[0] in ("LowerDomain", "UpperDomain"):
            if len(w) < 4:
                raise ValueError("%s count not 3" % w[0])
            (lo if w[0] == "LowerDomain" else hi)[:] = [float(a) for a in w[1:4]]
        elif w[0] == "StartCuboid":
            # token stream from the next line on (fscanf after fgets)
            rest = "".join(lines[k:])
            toks = rest.split()
            cub = {"Spacing": None, "Type": None, "Lower": None, "Upper": None, "Velocity": None}
            t = 0
            while True:
                if t >= len(toks):
                    raise ValueError("cuboid without EndCuboid")
                tok = toks[t]
                t += 1
                if tok == "EndCuboid":
                    break
                if tok == "Spacing":
                    cub["Spacing"] = float(toks[t]); t += 1
                elif tok == "Type":
                    cub["Type"] = int(toks[t]); t += 1
                elif tok in ("Lower", "Upper", "Velocity"):
                    cub[tok] = tuple(float(a) for a in toks[t:t + 3]); t += 3
                else:
                    raise ValueError("no such indication: %s" % tok)
            missing = [key for key, v in cub.items() if v is None]
            if missing:
                raise ValueError("cuboid without %s" % ", ".join(missing))
            cubs.append((cub["Spacing"], cub["Type"], cub["Lower"], cub["Upper"], cub["Velocity"]))
            # continue with the rest of the line that holds EndCuboid
            consumed, line_k, pos = 0, k, 0
            for line_k in range(k, len(lines)):
                n_here = len(lines[line_k].split())
                if consumed + n_here >= t:
                    break
                consumed += n_here
            words = lines[line_k].split()
            after = words[t - consumed:]
            k = line_k + 1
            if after:
                pending = " ".join(after) + "\n"
    return Boid(pd, tuple(lo), tuple(hi), cubs)


def _boid_particles(cuboids):
    """types, positions and velocities of the cuboids' particles (generator.cpp:170-193):
    per axis count = round(width / spacing) and spacing = width / count, centres from
    lower + spacing/2 while below upper - 0.49 spacing, accumulated by repeated addition"""
    types, pos, vel = [], [], []
    for space, ctype, lower, upper, velocity in cuboids:
        lower = np.array(lower, float)
        upper = np.array(upper, float)
        width = upper - lower
        # C round() rounds halves away from zero
        count = [int(math.copysign(math.floor(abs(width[d] / space) + 0.5), width[d] / space)) for d in range(3)]
        spacing = [width[d] / count[d] for d in range(3)]
        px = lower[0] + 0.5 * spacing[0]
        while px < upper[0] - 0.49 * spacing[0]:
            py = lower[1] + 0.5 * spacing[1]
            while py < upper[1] - 0.49 * spacing[1]:
                pz = lower[2] + 0.5 * spacing[2]
                while pz < upper[2] - 0.49 * spacing[2]:
                    types.append(ctype)
                    pos.append((px, py, pz))
                    vel.append(tuple(velocity))
                    pz += spacing[2]
                py += spacing[1]
            px += spacing[0]
    return types, pos, vel


def grid_text_from_boid(boid: Boid) -> str:
    """The .grid file the reference's generator writes for a .boid (generator.cpp:197-217)"""
    types, pos, vel = _boid_particles(boid.cuboids)
    lo, hi = boid.lower, boid.upper
    out = ["%f\n" % 0.0,
           "%d %e  %e %e %e  %e %e %e\n" % (len(types), boid.particle_distance, lo[0], hi[0], lo[1], hi[1], lo[2],
                                           hi[2])]
    out += ["%d   %e %e %e    %e %e %e\n" % (t, p[0], p[1], p[2], v[0], v[1], v[2])
            for t, p, v in zip(types, pos, vel)]
    return "".join(out)


def boid_to_grid(case: str) -> str:
    """``GeneratorForMph <case>`` (generator.cpp:37-47): read <case>.boid, write <case>.grid"""
    with open(case + ".boid") as f:
        boid = parse_boid(f.read())
    path = case + ".grid"
    with open(path, "w") as f:
        f.write(grid_text_from_boid(boid))
    return path


# --------------------------------------------------------------------------------------
# the reference's two sample decks (results/*/box.boid, box.data), expanded like
# generator/generator.cpp:170-193 and written with its %e precision (:197-217)


def _boid_expand(cuboids):
    types, pos, vel = _boid_particles(cuboids)
    t = np.array(types, np.int32)
    # the generator prints %e (7 significant digits); the solver reads those decimals back
    x = np.array([[float("%e" % c) for c in p] for p in pos])
    v = np.array([[float("%e" % c) for c in q] for q in vel])
    return t, x, v


# the cuboids of the two sample .boid files (results/*/box.boid), as data
SAMPLE_BRIDGE_CUBOIDS = [
    (0.0001, 1, (-0.002, 0.000, -0.00005), (0.002, 0.003, 0.00005), (0.0, 0.0, 0.0)),
    (0.0001, 2, (-0.005, -0.001, -0.00005), (0.005, 0.000, 0.00005), (0.0, 0.0, 0.0)),
    (0.0001, 2, (-0.005, 0.003, -0.00005), (0.005, 0.004, 0.00005), (0.0, 0.0, 0.0)),
]
SAMPLE_SQUARE_DROP_CUBOIDS = [
    (0.04, 1, (0.2, 0.68, -0.02), (1.2, 1.68, 0.02), (0.0, -0.8, 0.0)),
    (0.04, 2, (0.0, 0.0, -0.02), (0.4, 0.6, 0.02), (0.0, 0.0, 0.0)),
    (0.04, 2, (0.0, -0.20, -0.02), (3.0, 0.00, 0.02), (0.0, 0.0, 0.0)),
]


def sample_bridge() -> Deck:
    """results/Bridge_l00100_deg000_re35_mu01_g98 (2D liquid bridge between plates)."""
    t, x, v = _boid_expand(SAMPLE_BRIDGE_CUBOIDS)
    data = DataParams(dt=6.25e-6, output_interval=1.0, vtk_output_interval=2.0e-3, end_time=2.0e-2,
                      radius_ratio_a=3.5, radius_ratio_p=1.75, radius_ratio_v=3.5,
                      density=(1000.0,) * 4, bulk_modulus=(1.0e4,) * 4, bulk_viscosity=(0.1,) * 4,
                      shear_viscosity=(0.1,) * 4, surface_tension=(0.072,) * 4,
                      gravity=(0.0, -9.8, 0.0))
    lo = np.array([float("%e" % a) for a in (-0.005, -0.001, -0.001)])
    hi = np.array([float("%e" % a) for a in (0.005, 0.005, 0.001)])
    return Deck("sample_bridge", 2, data, 1.0e-4, lo, hi, t, x, v, steps=3201)


def sample_square_drop() -> Deck:
    """results/SquareDropSkip_l00400_rv35_rp17 (2D square drop falling onto a step)."""
    t, x, v = _boid_expand(SAMPLE_SQUARE_DROP_CUBOIDS)
    data = DataParams(dt=5.0e-4, output_interval=1.0, vtk_output_interval=5.0e-2, end_time=1.0,
                      radius_ratio_a=3.5, radius_ratio_p=1.75, radius_ratio_v=3.5,
                      density=(1.0e3,) * 4, bulk_modulus=(1.0e6,) * 4, bulk_viscosity=(2.0e3,) * 4,
                      shear_viscosity=(1.0e3,) * 4, surface_tension=(0.0,) * 4,
                      gravity=(0.0, -1.0, 0.0))
    lo = np.array([0.0, -0.2, -0.1])
    hi = np.array([3.0, 3.0, 0.1])
    return Deck("sample_square_drop", 2, data, 0.04, lo, hi, t, x, v, steps=2001)


# --------------------------------------------------------------------------------------
# small parity decks


def small_droplet_2d(n_side=20, seed=11) -> Deck:
    return droplet_2d(n_side=n_side, seed=seed, steps=200, name="small_droplet2d")


def small_dam_break_3d(seed=12) -> Deck:
    return dam_break_3d(fluid=(12, 12, 12), box=(24, 12, 18), jitter=0.01, seed=seed, steps=100,
                        name="small_dambreak3d")


def wide_dam_break_3d(seed=15) -> Deck:
    """Small 3D dam break whose every width exceeds eight list cut-offs (as the 1M deck's do):
    10x30x10 fluid in a 30x30x30 walled box, y periodic."""
    return dam_break_3d(fluid=(10, 30, 10), box=(30, 30, 30), jitter=0.01, seed=seed, steps=100,
                        name="wide_dambreak3d")


def periodic_flow_3d(n_side=20, l0=1.0e-3, drift=(0.02, 0.05, 0.0), seed=16, steps=40, name=None) -> Deck:
    """Fluid filling a periodic cube (no walls, no gravity) that drifts as a whole with a
    random perturbation: particles, and the members of a cluster, cross the periodic faces
    between rebuilds (every axis wide enough for the sweeps' image path)."""
    rng = np.random.default_rng(seed)
    rho, mu = 1000.0, 1.0e-3
    vmax = float(np.linalg.norm(drift))
    K, dt = sound_speed_params(rho, 10.0 * vmax, l0)
    x = _jitter(rng, _lattice((n_side,) * 3, l0, (0.0, 0.0, 0.0)), 0.01 * l0, 3)
    v = np.asarray(drift, np.float64)[None, :] + rng.uniform(-0.2, 0.2, size=x.shape) * vmax
    data = DataParams(dt=dt, end_time=dt * steps, density=(rho,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(mu,) * 4, shear_viscosity=(mu,) * 4, surface_tension=(0.0,) * 4)
    return Deck(name or "periodic_flow3d", 3, data, l0, np.zeros(3), np.full(3, n_side * l0),
                np.zeros(len(x), np.int32), x, v, steps=steps)


def walls_only_3d(l0=1.0e-3, steps=5, name=None) -> Deck:
    """Wall particles only (types 2 and 3, static), no fluid: every wall particle has no fluid
    partner, so none has a list (main.cpp:1416) and every sum is zero."""
    x = _lattice((10, 10, 6), l0, (0.0, 0.0, 0.0))
    t = np.where(x[:, 2] < 3 * l0, 2, 3).astype(np.int32)
    K, dt = sound_speed_params(1000.0, 1.0, l0)
    data = DataParams(dt=dt, end_time=dt * steps, bulk_modulus=(K,) * 4, gravity=(0.0, 0.0, -9.8))
    return Deck(name or "walls_only3d", 3, data, l0, np.zeros(3), np.array([20.0, 20.0, 12.0]) * l0,
                t, x, np.zeros_like(x), steps=steps)


def narrow_periodic_3d(shape=(16, 5, 8), l0=1.0e-3, seed=20, steps=20, name=None) -> Deck:
    """Fluid filling a periodic box with axes narrower than three cells (here one cell on y and
    two on z): the reference's stencil then visits cells more than once (main.cpp:1365-1391),
    and every list holds such repeats."""
    rng = np.random.default_rng(seed)
    rho, mu = 1000.0, 1.0e-3
    K, dt = sound_speed_params(rho, 1.0, l0)
    x = _jitter(rng, _lattice(shape, l0, (0.0, 0.0, 0.0)), 0.02 * l0, 3)
    v = rng.uniform(-0.02, 0.02, size=x.shape)
    data = DataParams(dt=dt, end_time=dt * steps, density=(rho,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(mu,) * 4, shear_viscosity=(mu,) * 4, surface_tension=(0.0,) * 4)
    return Deck(name or "narrow_periodic3d", 3, data, l0, np.zeros(3), np.asarray(shape, np.float64) * l0,
                np.zeros(len(x), np.int32), x, v, steps=steps)


def sparse_random_3d(n=1500, side=24, l0=1.0e-3, seed=17, steps=10, name=None) -> Deck:
    """Particles at uniformly random positions and velocities in a periodic cube, about 0.11
    per l0^3 (a tenth of the lattice density), no gravity: cells hold a few particles far apart,
    so some members of a cluster have no other member within the cut-off, and some pairs are
    much closer than l0."""
    rng = np.random.default_rng(seed)
    rho, mu = 1000.0, 1.0e-3
    K, dt = sound_speed_params(rho, 1.0, l0)
    L = side * l0
    x = rng.uniform(0.0, L, size=(n, 3))
    data = DataParams(dt=dt, end_time=dt * steps, density=(rho,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(mu,) * 4, shear_viscosity=(mu,) * 4, surface_tension=(0.0,) * 4)
    v = rng.uniform(-0.05, 0.05, size=(n, 3))
    return Deck(name or "sparse_random3d", 3, data, l0, np.zeros(3), np.full(3, L), np.zeros(n, np.int32), x,
                v, steps=steps)


def collapsing_cloud_3d(n_side=14, spread=1.45, l0=1.0e-3, vmax=0.2, seed=18, steps=200, name=None) -> Deck:
    """A sparse cubic cloud (lattice spacing ``spread``*l0, about a third of the fluid
    density) falling in on its centre, v = -alpha x with the corners at ``vmax``, in a
    periodic box without gravity.  Sound speed 10 vmax, dt from CFL 0.2.  After ``steps``
    = 200 steps the cloud has shrunk by about a quarter and its neighbour lists have more
    than doubled, past the width reserved at the first build."""
    rng = np.random.default_rng(seed)
    rho, mu = 1000.0, 1.0e-3
    half = 0.5 * n_side * spread * l0
    x = _jitter(rng, _lattice((n_side,) * 3, spread * l0, (-half, -half, -half)), 0.01 * l0, 3)
    alpha = vmax / float(np.max(np.linalg.norm(x, axis=1)))
    v = -alpha * x
    K, dt = sound_speed_params(rho, 10.0 * vmax, l0)
    L = half + 6.0 * l0
    data = DataParams(dt=dt, end_time=dt * steps, density=(rho,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(mu,) * 4, shear_viscosity=(mu,) * 4, surface_tension=(0.0,) * 4)
    return Deck(name or "collapsing_cloud3d", 3, data, l0, np.full(3, -L), np.full(3, L),
                np.zeros(len(x), np.int32), x, v, steps=steps)


def drift_dam_break_3d(seed=14) -> Deck:
    """Small 3D dam break for 1000-step runs: 8x11x8 fluid in a 14x11x12 walled box."""
    return dam_break_3d(fluid=(8, 11, 8), box=(14, 11, 12), jitter=0.01, seed=seed, steps=1000,
                        name="drift_dambreak3d")


def small_two_fluid_3d(seed=13) -> Deck:
    """3D two-fluid block with walls, surface tension and non-unit interaction ratios."""
    d = dam_break_3d(fluid=(10, 12, 10), box=(16, 12, 14), jitter=0.02, seed=seed, steps=50,
                     two_fluid=True, sigma=0.07, name="small_twofluid3d")
    st = (0.07, 0.05, 0.06, 0.06)
    ir = ((1.0, 0.5, 0.8, 0.8), (0.5, 1.0, 0.7, 0.7), (1.0, 1.0, 1.0, 1.0), (1.0, 1.0, 1.0, 1.0))
    d.data.surface_tension = st
    d.data.interaction_ratio = ir
    return d


def small_two_fluid_3d_no_surface_tension(seed=13) -> Deck:
    """The 3D two-fluid block with non-unit interaction ratios and no surface tension: n^A
    (DensityA) then weighs pairs by InteractionRatio but enters no force, so it is summed
    only when the fields are read."""
    d = small_two_fluid_3d(seed=seed)
    d.data.surface_tension = (0.0,) * 4
    d.name = "small_twofluid3d_nosurf"
    return d


def small_moving_wall_2d(seed=14) -> Deck:
    """2D fluid patch on a translating+rotating wall plate (types 2 and 3)."""
    rng = np.random.default_rng(seed)
    l0 = 1.0e-3
    xf = _lattice((16, 10, 1), l0, (-8 * l0, 0.0, 0.0))
    xf[:, 2] = 0.0
    xf = _jitter(rng, xf, 0.01 * l0, 2)
    xw2 = _lattice((24, 3, 1), l0, (-12 * l0, -3 * l0, 0.0)); xw2[:, 2] = 0.0
    xw3 = _lattice((3, 8, 1), l0, (-12 * l0, 0.0, 0.0)); xw3[:, 2] = 0.0
    x = np.concatenate([xf, xw2, xw3])
    t = np.concatenate([np.zeros(len(xf), np.int32), np.full(len(xw2), 2, np.int32),
                        np.full(len(xw3), 3, np.int32)])
    K, dt = sound_speed_params(1000.0, 2.0, l0)
    wc = ((0, 0, 0), (0, 0, 0), (0.0, -1.5e-3, 0.0), (-10.5e-3, 4e-3, 0.0))
    wv = ((0, 0, 0), (0, 0, 0), (0.05, 0.0, 0.0), (0.0, 0.0, 0.0))
    wo = ((0, 0, 0), (0, 0, 0), (0.0, 0.0, 0.5), (0.0, 0.0, -0.8))
    data = DataParams(dt=dt, end_time=dt * 50, density=(1000.0,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(0.05,) * 4, shear_viscosity=(0.05,) * 4,
                      surface_tension=(0.05,) * 4, gravity=(0.0, -9.8, 0.0),
                      wall_center=wc, wall_velocity=wv, wall_omega=wo)
    L = 0.04
    return Deck("small_movingwall2d", 2, data, l0, np.array([-L, -L, -5e-4]), np.array([L, L, 5e-4]),
                t, x, np.zeros((len(x), 3)), steps=50)


def small_rotating_wall_3d(seed=19) -> Deck:
    """3D fluid block on a translating plate (type 2) beside a wall (type 3); both walls
    rotate about axes off the coordinate axes, so every component of the in-place rotation
    r <- R r (main.cpp:2033-2035) reads an already rotated one."""
    rng = np.random.default_rng(seed)
    l0 = 1.0e-3
    xf = _jitter(rng, _lattice((8, 8, 8), l0, (-4 * l0, 0.0, -4 * l0)), 0.01 * l0, 3)
    xw2 = _lattice((14, 3, 14), l0, (-7 * l0, -3 * l0, -7 * l0))
    xw3 = _lattice((3, 8, 14), l0, (-7 * l0, 0.0, -7 * l0))
    x = np.concatenate([xf, xw2, xw3])
    t = np.concatenate([np.zeros(len(xf), np.int32), np.full(len(xw2), 2, np.int32),
                        np.full(len(xw3), 3, np.int32)])
    K, dt = sound_speed_params(1000.0, 2.0, l0)
    wc = ((0, 0, 0), (0, 0, 0), (0.0, -1.5e-3, 0.0), (-6e-3, 4e-3, 0.0))
    wv = ((0, 0, 0), (0, 0, 0), (0.04, 0.0, -0.02), (0.0, 0.01, 0.0))
    wo = ((0, 0, 0), (0, 0, 0), (0.3, -0.4, 0.5), (0.6, 0.2, -0.7))
    data = DataParams(dt=dt, end_time=dt * 50, density=(1000.0,) * 4, bulk_modulus=(K,) * 4,
                      bulk_viscosity=(0.05,) * 4, shear_viscosity=(0.05,) * 4,
                      surface_tension=(0.05,) * 4, gravity=(0.0, -9.8, 0.0),
                      wall_center=wc, wall_velocity=wv, wall_omega=wo)
    L = 0.03
    return Deck("small_rotatingwall3d", 3, data, l0, np.array([-L, -L, -L]), np.array([L, L, L]),
                t, x, np.zeros((len(x), 3)), steps=50)


# --------------------------------------------------------------------------------------
# batched cases


def batch_cases(count=64, seed=4, n_min=2.0e3, n_max=6.0e4, l0=1.0e-3) -> List[Deck]:
    """Independent droplets and liquid bridges, 2D and 3D, with log-uniform particle counts.

    Viscosity ~ U[0.01, 1] Pa s, surface tension ~ U[0.01, 0.1] N/m, density 1000.
    Sound speed = 10 x (capillary velocity scale), floored at 0.5 m/s; dt from CFL 0.2 and
    the viscous limit.
    """
    rng = np.random.default_rng(seed)
    decks = []
    for k in range(count):
        target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        dim = 2 if rng.uniform() < 0.5 else 3
        bridge = rng.uniform() < 0.5
        mu = float(rng.uniform(0.01, 1.0))
        sigma = float(rng.uniform(0.01, 0.1))
        rho = 1000.0
        sub = np.random.default_rng(int(rng.integers(1 << 31)))
        if dim == 2:
            side = max(8, int(round(math.sqrt(target))))
            counts = (side, side, 1)
        else:
            side = max(6, int(round(target ** (1.0 / 3.0))))
            counts = (side, side, side)
        R = 0.5 * side * l0
        v_cap = min(sigma / mu, math.sqrt(sigma / (rho * R)))
        c = max(0.5, 10.0 * 2.0 * v_cap)
        K, dt = sound_speed_params(rho, c, l0)
        dt = min(dt, 0.1 * rho * l0 * l0 / mu)
        xf = _lattice(counts, l0, (-R, -R, -R if dim == 3 else 0.0))
        if dim == 2:
            xf[:, 2] = 0.0
        xf = _jitter(sub, xf, 0.01 * l0, dim)
        parts = [xf]
        types = [np.zeros(len(xf), np.int32)]
        if bridge:
            # plates of 3 layers below and above the fluid along y, wider than the fluid
            wc = list(counts)
            wc[0] = side + 8
            wc[1] = 3
            o_lo = (-R - 4 * l0, -R - 3 * l0, (-R if dim == 3 else 0.0))
            o_hi = (-R - 4 * l0, R, (-R if dim == 3 else 0.0))
            for o in (o_lo, o_hi):
                w = _lattice(tuple(wc), l0, o)
                if dim == 2:
                    w[:, 2] = 0.0
                parts.append(w)
                types.append(np.full(len(w), 2, np.int32))
        x = np.concatenate(parts)
        t = np.concatenate(types)
        L = R + 8 * l0 + 4 * 3.6 * l0
        lo = np.array([-L, -L, -L if dim == 3 else -5e-4])
        hi = np.array([L, L, L if dim == 3 else 5e-4])
        data = DataParams(dt=dt, end_time=dt * 1000, density=(rho,) * 4, bulk_modulus=(K,) * 4,
                          bulk_viscosity=(mu,) * 4, shear_viscosity=(mu,) * 4,
                          surface_tension=(sigma,) * 4)
        decks.append(Deck("batch%02d_%dd_%s" % (k, dim, "bridge" if bridge else "drop"), dim, data,
                          l0, lo, hi, t, x, np.zeros((len(x), 3)), steps=1000))
    return decks


if __name__ == "__main__":
    # python -m mphexplicit_amd.decks <case>: the reference's `GeneratorForMph <case>`
    import sys
    if len(sys.argv) != 2:
        sys.exit("usage: python -m mphexplicit_amd.decks <case>   (reads <case>.boid, writes <case>.grid)")
    print(boid_to_grid(sys.argv[1]))
