"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Two CPU implementations of the reference's step sit behind one interface:

* ``Oracle``      -- our plain-C restatement (oracle/mph_oracle.c, built as liboracle_mph.so);
* ``Reference``   -- the reference's own source/main.cpp compiled under oracle/_ref by
                     oracle/Makefile (present only where that build was made).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, and
only as the checker.  Fields come back in the solver's current order together with
``particle_index``; :func:`by_id` reorders them to original ids.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

from mphexplicit_amd.abi import params_from_deck

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_mph.so")
REF_DIR = os.path.join(HERE, "_ref")

FIELDS = ("index", "type", "x", "v", "force", "dens_a", "vol_strain", "divergence",
          "press_p", "press_a", "gcenter", "ncount", "mass")
_VEC = {"x", "v", "force", "gcenter"}
_INT = {"index", "type", "ncount"}


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_mph.so"], check=True)


def reference_available(dim: int) -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libmphref%dd.so" % dim))


def _alloc(n):
    out = {}
    for k in FIELDS:
        if k in _INT:
            out[k] = np.zeros(n, np.int32)
        elif k in _VEC:
            out[k] = np.zeros((n, 3), np.float64)
        else:
            out[k] = np.zeros(n, np.float64)
    return out


def _ptrs(out):
    return [out[k].ctypes.data_as(C.c_void_p) for k in FIELDS]


def by_id(fields: dict) -> dict:
    """Reorder every per-particle array by original particle id."""
    order = np.argsort(fields["index"], kind="stable")
    return {k: (v[order] if isinstance(v, np.ndarray) and v.shape[:1] == order.shape else v)
            for k, v in fields.items()}


def neighbor_sets_by_id(index, ptr, ind):
    """CSR lists in current order -> {original id: sorted tuple of original neighbour ids}."""
    index = np.asarray(index)
    out = {}
    for i in range(len(index)):
        row = index[np.asarray(ind[ptr[i]:ptr[i + 1]])]
        out[int(index[i])] = tuple(sorted(int(j) for j in row))
    return out


class Oracle:
    """Our C restatement of the reference step (oracle/mph_oracle.c)."""

    def __init__(self, deck, nthreads: int = 1):
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        lib = C.CDLL(ORACLE_SO)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        lib.orc_step.argtypes = [C.c_void_p, C.c_int]
        for f in ("orc_last_rebuild", "orc_count"):
            getattr(lib, f).argtypes = [C.c_void_p]
        lib.orc_time.restype = C.c_double
        lib.orc_time.argtypes = [C.c_void_p]
        lib.orc_neighbor_total.restype = C.c_long
        lib.orc_neighbor_total.argtypes = [C.c_void_p]
        lib.orc_get.argtypes = [C.c_void_p] + [C.c_void_p] * len(FIELDS)
        lib.orc_neighbors.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_virial.argtypes = [C.c_void_p, C.c_void_p]
        lib.orc_vol_strain_exact.argtypes = [C.c_void_p, C.c_void_p]
        lib.orc_scalars.argtypes = [C.c_void_p, C.c_void_p]
        lib.orc_destroy.argtypes = [C.c_void_p]
        self.lib = lib
        self.params = params_from_deck(deck)
        t = np.ascontiguousarray(deck.type, np.int32)
        x = np.ascontiguousarray(deck.x, np.float64)
        v = np.ascontiguousarray(deck.v, np.float64)
        self.h = lib.orc_create(C.byref(self.params), deck.n, t.ctypes.data, x.ctypes.data,
                                v.ctypes.data, nthreads)
        if not self.h:
            raise ValueError("oracle rejected the deck")
        self.n = deck.n

    def step(self, n: int = 1) -> int:
        return self.lib.orc_step(self.h, n)

    def last_rebuild(self) -> bool:
        return bool(self.lib.orc_last_rebuild(self.h))

    @property
    def time(self) -> float:
        return self.lib.orc_time(self.h)

    def fields(self) -> dict:
        out = _alloc(self.n)
        self.lib.orc_get(self.h, *_ptrs(out))
        return out

    def neighbors(self):
        tot = self.lib.orc_neighbor_total(self.h)
        ptr = np.zeros(self.n + 1, np.int64)
        ind = np.zeros(max(tot, 1), np.int32)
        self.lib.orc_neighbors(self.h, ptr.ctypes.data, ind.ctypes.data)
        return ptr, ind[:tot]

    def virial(self) -> np.ndarray:
        vp = np.zeros(self.n)
        self.lib.orc_virial(self.h, vp.ctypes.data)
        return vp

    def scalars(self) -> np.ndarray:
        o = np.zeros(16)
        self.lib.orc_scalars(self.h, o.ctypes.data)
        return o

    def vol_strain_exact(self) -> np.ndarray:
        """VolStrainP (current order) with long double sums: the exact value for the same
        inputs, against which both codes' FP64 summation rounding is measured"""
        out = np.zeros(self.n)
        self.lib.orc_vol_strain_exact(self.h, out.ctypes.data)
        return out

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ref_libs = {}


class Reference:
    """The reference's own source/main.cpp built under oracle/_ref (one state per process
    and per dimension, since the reference keeps its state in file-scope statics)."""

    def __init__(self, deck, nthreads: int = 1, workdir: str | None = None):
        path = os.path.join(REF_DIR, "libmphref%dd.so" % deck.dim)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        if path not in _ref_libs:
            lib = C.CDLL(path)
            lib.mphref_load.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
            lib.mphref_step.argtypes = [C.c_int]
            lib.mphref_time.restype = C.c_double
            lib.mphref_neighbor_total.restype = C.c_long
            lib.mphref_get.argtypes = [C.c_void_p] * len(FIELDS)
            lib.mphref_neighbors.argtypes = [C.c_void_p, C.c_void_p]
            lib.mphref_virial.argtypes = [C.c_void_p]
            lib.mphref_scalars.argtypes = [C.c_void_p]
            _ref_libs[path] = lib
        self.lib = _ref_libs[path]
        tmp = workdir or tempfile.mkdtemp(prefix="mphref_")
        dpath, gpath = deck.write(tmp, "deck")
        self.n = self.lib.mphref_load(dpath.encode(), gpath.encode(), b"/dev/null", nthreads)

    def step(self, n: int = 1) -> int:
        return self.lib.mphref_step(n)

    def last_rebuild(self) -> bool:
        return bool(self.lib.mphref_last_rebuild())

    @property
    def time(self) -> float:
        return self.lib.mphref_time()

    def fields(self) -> dict:
        out = _alloc(self.n)
        self.lib.mphref_get(*_ptrs(out))
        return out

    def neighbors(self):
        tot = self.lib.mphref_neighbor_total()
        ptr = np.zeros(self.n + 1, np.int64)
        ind = np.zeros(max(tot, 1), np.int32)
        self.lib.mphref_neighbors(ptr.ctypes.data, ind.ctypes.data)
        return ptr, ind[:tot]

    def virial(self) -> np.ndarray:
        vp = np.zeros(self.n)
        self.lib.mphref_virial(vp.ctypes.data)
        return vp

    def scalars(self) -> np.ndarray:
        o = np.zeros(16)
        self.lib.mphref_scalars(o.ctypes.data)
        return o
