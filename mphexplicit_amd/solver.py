"""Python front end of the C ABI (include/mph.h), loaded with ctypes.

``Solver`` drives the reference's per-step loop (source/main.cpp:374-452) on the GPU:
create (main.cpp:324-364), step (main.cpp:386-431), fields for output (main.cpp:698-834).
It never computes anything itself; if the HIP library is missing it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Iterable, List, Sequence

import numpy as np

from .abi import (FIELD_SHAPES, SUR_NPARAMS, MphFields, MphParams, MphStats, MphSurParams, MphSurStats,
                  MphTiming, params_from_deck)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmph_hip.so")

MPH_FP64, MPH_MIXED = 0, 1
ORDER_CURRENT, ORDER_ORIGINAL = 0, 1

_lib = None


class MphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("mph error %d: %s" % (code, msg))
        self.code = code


def load_library() -> C.CDLL:
    """Load libmph_hip.so (built by ``python -m mphexplicit_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmph_hip.so not built; run `python -m mphexplicit_amd.build`")
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    sig = {
        "mph_params_default": (None, [C.POINTER(MphParams)]),
        "mph_read_data_file": (C.c_int, [C.c_char_p, C.POINTER(MphParams)]),
        "mph_read_grid_header": (C.c_int, [C.c_char_p, C.POINTER(MphParams), ip]),
        "mph_read_grid_particles": (C.c_int, [C.c_char_p, C.c_int, vp, vp, vp]),
        "mph_write_prof_file": (C.c_int, [C.c_char_p, C.POINTER(MphParams), C.c_double, C.c_int, vp, vp, vp]),
        "mph_write_vtk_file": (C.c_int, [C.c_char_p, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]),
        "mph_device_count": (C.c_int, []),
        "mph_create": (vp, [C.POINTER(MphParams), C.c_int, vp, vp, vp]),
        "mph_create_batch": (vp, [C.c_int, vp, vp, vp, vp, vp]),
        "mph_step": (C.c_int, [vp, C.c_int]),
        "mph_step_graph": (C.c_int, [vp, C.c_int]),
        "mph_step_timed": (C.c_int, [vp, C.c_int, C.POINTER(MphTiming)]),
        "mph_sync": (C.c_int, [vp]),
        "mph_particle_count": (C.c_int, [vp, C.c_int]),
        "mph_get_time": (C.c_int, [vp, C.c_int, dp]),
        "mph_get_fields": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(MphFields)]),
        "mph_compute_virial": (C.c_int, [vp]),
        "mph_get_neighbors": (C.c_int, [vp, C.c_int, vp, vp]),
        "mph_get_neighbors_noalloc": (C.c_int, [vp, C.c_int, vp, vp, C.c_longlong, vp]),
        "mph_get_stats": (C.c_int, [vp, C.POINTER(MphStats)]),
        "mph_destroy": (None, [vp]),
        "mph_last_error": (C.c_char_p, []),
        "mph_set_time_step": (C.c_int, [vp, C.c_int, C.c_double]),
        "mph_sur_params_default": (None, [C.POINTER(MphSurParams)]),
        "mph_sur_create": (vp, [vp, C.POINTER(MphSurParams)]),
        "mph_sur_get_weights": (C.c_int, [vp, vp]),
        "mph_sur_set_weights": (C.c_int, [vp, vp]),
        "mph_sur_predict": (C.c_int, [vp, vp]),
        "mph_sur_gradient": (C.c_int, [vp, vp, dp, vp]),
        "mph_sur_train": (C.c_int, [vp, C.c_int, C.c_int, vp]),
        "mph_sur_step": (C.c_int, [vp, C.c_int]),
        "mph_sur_advance": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.POINTER(MphSurStats)]),
        "mph_sur_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise MphError(rc, load_library().mph_last_error().decode())


def read_case(data_path: str, grid_path: str, dim: int = 3, precision: int = MPH_FP64):
    """Read a .data/.grid pair with the library's readers -> (params, type, x, v)."""
    lib = load_library()
    p = MphParams()
    lib.mph_params_default(C.byref(p))
    _check(lib.mph_read_data_file(data_path.encode(), C.byref(p)))
    n = C.c_int(0)
    _check(lib.mph_read_grid_header(grid_path.encode(), C.byref(p), C.byref(n)))
    t = np.zeros(n.value, np.int32)
    x = np.zeros((n.value, 3))
    v = np.zeros((n.value, 3))
    _check(lib.mph_read_grid_particles(grid_path.encode(), n.value, t.ctypes.data, x.ctypes.data, v.ctypes.data))
    p.dim = dim
    p.precision = precision
    return p, t, x, v


class Solver:
    """One context: one case, or several independent cases stepped together."""

    def __init__(self, params: Sequence[MphParams], types: Sequence[np.ndarray],
                 positions: Sequence[np.ndarray], velocities: Sequence[np.ndarray]):
        lib = load_library()
        self.lib = lib
        k = len(params)
        arr = (MphParams * k)(*params)
        counts = np.array([len(t) for t in types], np.int32)
        t = np.ascontiguousarray(np.concatenate(types), np.int32)
        x = np.ascontiguousarray(np.concatenate(positions), np.float64)
        v = np.ascontiguousarray(np.concatenate(velocities), np.float64)
        self.h = lib.mph_create_batch(k, C.cast(arr, C.c_void_p), counts.ctypes.data, t.ctypes.data,
                                      x.ctypes.data, v.ctypes.data)
        if not self.h:
            raise MphError(-1, lib.mph_last_error().decode())
        self.ncases = k
        self.counts = counts

    @classmethod
    def from_deck(cls, deck, precision: int = MPH_FP64) -> "Solver":
        return cls([params_from_deck(deck, precision)], [deck.type], [deck.x], [deck.v])

    @classmethod
    def from_decks(cls, decks: Iterable, precision: int = MPH_FP64) -> "Solver":
        decks = list(decks)
        return cls([params_from_deck(d, precision) for d in decks], [d.type for d in decks],
                   [d.x for d in decks], [d.v for d in decks])

    def step(self, n: int = 1) -> None:
        _check(self.lib.mph_step(self.h, n))

    def step_graph(self, n: int = 1) -> None:
        _check(self.lib.mph_step_graph(self.h, n))

    def step_timed(self, n: int = 1) -> dict:
        t = MphTiming()
        _check(self.lib.mph_step_timed(self.h, n, C.byref(t)))
        return {k: getattr(t, k) for k, _ in MphTiming._fields_}

    def set_dt(self, dt: float, case: int = 0) -> None:
        _check(self.lib.mph_set_time_step(self.h, case, dt))

    def sync(self) -> None:
        _check(self.lib.mph_sync(self.h))

    def time(self, case: int = 0) -> float:
        t = C.c_double(0.0)
        _check(self.lib.mph_get_time(self.h, case, C.byref(t)))
        return t.value

    def fields(self, case: int = 0, order: int = ORDER_ORIGINAL,
               names: Sequence[str] | None = None) -> Dict[str, np.ndarray]:
        n = int(self.counts[case])
        names = list(names or FIELD_SHAPES.keys())
        out, f = {}, MphFields()
        for nm in names:
            dt, comp = FIELD_SHAPES[nm]
            a = np.zeros((n, comp) if comp > 1 else n, dt)
            out[nm] = a
            ct = C.c_int if dt == np.int32 else C.c_double
            setattr(f, nm, a.ctypes.data_as(C.POINTER(ct)))
        _check(self.lib.mph_get_fields(self.h, case, order, C.byref(f)))
        return out

    def compute_virial(self) -> None:
        _check(self.lib.mph_compute_virial(self.h))

    def neighbors(self, case: int = 0):
        """CSR (ptr, ind) by original id, rows ascending, self included (reference lists)."""
        n = int(self.counts[case])
        ptr = np.zeros(n + 1, np.int64)
        _check(self.lib.mph_get_neighbors(self.h, case, ptr.ctypes.data, None))
        ind = np.zeros(max(int(ptr[-1]), 1), np.int32)
        _check(self.lib.mph_get_neighbors(self.h, case, ptr.ctypes.data, ind.ctypes.data))
        return ptr, ind[: int(ptr[-1])]

    def stats(self) -> dict:
        s = MphStats()
        _check(self.lib.mph_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in MphStats._fields_}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mph_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sur_params(**kw) -> MphSurParams:
    """mph_sur_params with the library defaults, overridden by keyword"""
    p = MphSurParams()
    load_library().mph_sur_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Surrogate:
    """The pairwise-message surrogate bound to a single-case Solver (mph_sur_* of mph.h)."""

    def __init__(self, solver: "Solver", **params):
        self.lib = load_library()
        self.solver = solver
        self.params = sur_params(**params)
        h = self.lib.mph_sur_create(solver.h, C.byref(self.params))
        if not h:
            raise MphError(-1, self.lib.mph_last_error().decode())
        self.h = h

    def weights(self) -> np.ndarray:
        w = np.zeros(SUR_NPARAMS, np.float32)
        _check(self.lib.mph_sur_get_weights(self.h, w.ctypes.data))
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (SUR_NPARAMS,)
        _check(self.lib.mph_sur_set_weights(self.h, w.ctypes.data))

    def predict(self) -> np.ndarray:
        a = np.zeros((int(self.solver.counts[0]), 3))
        _check(self.lib.mph_sur_predict(self.h, a.ctypes.data))
        return a

    def gradient(self, target: np.ndarray):
        t = np.ascontiguousarray(target, np.float64)
        loss = C.c_double(0.0)
        g = np.zeros(SUR_NPARAMS, np.float32)
        _check(self.lib.mph_sur_gradient(self.h, t.ctypes.data, C.byref(loss), g.ctypes.data))
        return loss.value, g

    def train(self, samples: int, iters: int) -> np.ndarray:
        losses = np.zeros(max(samples * iters, 1))
        _check(self.lib.mph_sur_train(self.h, samples, iters, losses.ctypes.data))
        return losses[: samples * iters]

    def step(self, n: int = 1) -> None:
        _check(self.lib.mph_sur_step(self.h, n))

    def advance(self, blocks: int, k: int = 10, tol: float = 0.01) -> dict:
        st = MphSurStats()
        _check(self.lib.mph_sur_advance(self.h, blocks, k, tol, C.byref(st)))
        return {f: getattr(st, f) for f, _ in MphSurStats._fields_}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mph_sur_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
