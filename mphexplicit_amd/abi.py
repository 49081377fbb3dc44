"""ctypes mirrors of the structs declared in include/mph.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

_D3 = C.c_double * 3
_D4 = C.c_double * 4


class MphParams(C.Structure):
    """``mph_params``: the .data keywords + .grid header (source/main.cpp:502-521, 549-557)."""
    _fields_ = [
        ("dim", C.c_int),
        ("precision", C.c_int),
        ("particle_spacing", C.c_double),
        ("domain_min", _D3),
        ("domain_max", _D3),
        ("time", C.c_double),
        ("dt", C.c_double),
        ("output_interval", C.c_double),
        ("vtk_output_interval", C.c_double),
        ("end_time", C.c_double),
        ("radius_ratio_a", C.c_double),
        ("radius_ratio_p", C.c_double),
        ("radius_ratio_v", C.c_double),
        ("density", _D4),
        ("bulk_modulus", _D4),
        ("bulk_viscosity", _D4),
        ("shear_viscosity", _D4),
        ("surface_tension", _D4),
        ("interaction_ratio", _D4 * 4),
        ("gravity", _D3),
        ("wall_center", _D3 * 4),
        ("wall_velocity", _D3 * 4),
        ("wall_omega", _D3 * 4),
    ]


class MphFields(C.Structure):
    _fields_ = [
        ("particle_index", C.POINTER(C.c_int)),
        ("type", C.POINTER(C.c_int)),
        ("position", C.POINTER(C.c_double)),
        ("velocity", C.POINTER(C.c_double)),
        ("force", C.POINTER(C.c_double)),
        ("density_a", C.POINTER(C.c_double)),
        ("gravity_center", C.POINTER(C.c_double)),
        ("pressure_a", C.POINTER(C.c_double)),
        ("vol_strain_p", C.POINTER(C.c_double)),
        ("divergence_p", C.POINTER(C.c_double)),
        ("pressure_p", C.POINTER(C.c_double)),
        ("virial_pressure", C.POINTER(C.c_double)),
        ("neighbor_count", C.POINTER(C.c_int)),
    ]


class MphStats(C.Structure):
    _fields_ = [
        ("steps", C.c_longlong),
        ("rebuilds", C.c_longlong),
        ("list_entries", C.c_longlong),
        ("list_width", C.c_int),
        ("max_list_len", C.c_int),
        ("cluster_size", C.c_int),
        ("clusters", C.c_int),
        ("eval_slots", C.c_longlong),
        ("list_grid", C.c_int),
    ]


class MphTiming(C.Structure):
    _fields_ = [
        ("move_ms", C.c_double),
        ("rebuild_ms", C.c_double),
        ("sweep_a_ms", C.c_double),
        ("sweep_b_ms", C.c_double),
        ("pairs", C.c_longlong),
        ("rebuilds", C.c_longlong),
    ]


class MphSurParams(C.Structure):
    """mph_sur_params of include/mph.h"""
    _fields_ = [
        ("u_scale", C.c_double), ("a_scale", C.c_double), ("dens_scale", C.c_double),
        ("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double),
        ("train_stride", C.c_int), ("pad0", C.c_int), ("seed", C.c_ulonglong),
        ("radius_ratio", C.c_double),
    ]


class MphSurStats(C.Structure):
    """mph_sur_stats of include/mph.h"""
    _fields_ = [
        ("blocks", C.c_longlong), ("rollbacks", C.c_longlong), ("sur_steps", C.c_longlong),
        ("plain_steps", C.c_longlong), ("max_err_v", C.c_double), ("max_err_n", C.c_double),
        ("last_err_v", C.c_double), ("last_err_n", C.c_double), ("sim_time", C.c_double),
    ]


SUR_NPARAMS = 9603

FIELD_SHAPES = {
    "particle_index": (np.int32, 1), "type": (np.int32, 1), "position": (np.float64, 3),
    "velocity": (np.float64, 3), "force": (np.float64, 3), "density_a": (np.float64, 1),
    "gravity_center": (np.float64, 3), "pressure_a": (np.float64, 1),
    "vol_strain_p": (np.float64, 1), "divergence_p": (np.float64, 1),
    "pressure_p": (np.float64, 1), "virial_pressure": (np.float64, 1),
    "neighbor_count": (np.int32, 1),
}


def params_from_deck(deck, precision: int = 0) -> MphParams:
    """Fill ``mph_params`` from a :class:`mphexplicit_amd.decks.Deck`."""
    p = MphParams()
    d = deck.data
    p.dim = deck.dim
    p.precision = precision
    p.particle_spacing = deck.spacing
    for k in range(3):
        p.domain_min[k] = float(deck.domain_min[k])
        p.domain_max[k] = float(deck.domain_max[k])
        p.gravity[k] = float(d.gravity[k])
    p.time = deck.time
    p.dt = d.dt
    p.output_interval = d.output_interval
    p.vtk_output_interval = d.vtk_output_interval
    p.end_time = d.end_time
    p.radius_ratio_a = d.radius_ratio_a
    p.radius_ratio_p = d.radius_ratio_p
    p.radius_ratio_v = d.radius_ratio_v
    for t in range(4):
        p.density[t] = d.density[t]
        p.bulk_modulus[t] = d.bulk_modulus[t]
        p.bulk_viscosity[t] = d.bulk_viscosity[t]
        p.shear_viscosity[t] = d.shear_viscosity[t]
        p.surface_tension[t] = d.surface_tension[t]
        for u in range(4):
            p.interaction_ratio[t][u] = d.interaction_ratio[t][u]
        for k in range(3):
            p.wall_center[t][k] = d.wall_center[t][k]
            p.wall_velocity[t][k] = d.wall_velocity[t][k]
            p.wall_omega[t][k] = d.wall_omega[t][k]
    return p


def as_ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
