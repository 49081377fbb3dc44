"""mph_get_neighbors_noalloc: the neighbour lists in CSR form written into the caller's
device buffers, with the library's scratch in a static device array (no allocation).
Checked against mph_get_neighbors (host form, itself checked against the oracle by the
parity tests) bit for bit, on two calls in a row, the second on a smaller case, and once
more on the first context afterwards; bad arguments are rejected before any kernel."""
import ctypes as C
import os

import numpy as np
import pytest

from mphexplicit_amd import decks

pytestmark = pytest.mark.gpu

MPH_ERR_ARG = -1


@pytest.fixture(scope="module")
def hip(gpu_available):
    if not gpu_available:
        pytest.skip("no GPU")
    lib = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0
        # garbage, so that every output value must be written by the call
        assert hip.hipMemset(self.p, 0xA5, max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0     # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _check(hip, solver, label):
    from mphexplicit_amd.solver import load_library
    lib = load_library()
    ptr_h, ind_h = solver.neighbors()
    n = len(ptr_h) - 1
    total_h = int(ptr_h[-1])
    dptr, dind = DevArray(hip, 8 * (n + 1)), DevArray(hip, 4 * total_h)
    total = C.c_longlong(-1)
    rc = lib.mph_get_neighbors_noalloc(solver.h, 0, dptr.p, dind.p, total_h, C.byref(total))
    assert rc == 0, lib.mph_last_error().decode()
    assert total.value == total_h
    assert np.array_equal(dptr.get(np.int64, n + 1), ptr_h), label
    assert np.array_equal(dind.get(np.int32, total_h), ind_h), label
    print("%s: %d particles, %d list entries, device rows identical" % (label, n, total_h), flush=True)
    # too small an ind buffer, and host memory, are rejected
    assert lib.mph_get_neighbors_noalloc(solver.h, 0, dptr.p, dind.p, total_h - 1, C.byref(total)) == MPH_ERR_ARG
    assert total.value == total_h
    host = np.zeros(n + 1, np.int64)
    assert lib.mph_get_neighbors_noalloc(solver.h, 0, host.ctypes.data, None, 0, None) == MPH_ERR_ARG
    dptr.free()
    dind.free()


def test_noalloc_neighbors_match_host_form(hip):
    from mphexplicit_amd.solver import Solver
    big = Solver.from_deck(decks.small_dam_break_3d())
    big.step(3)
    _check(hip, big, "dambreak3d")
    small = Solver.from_deck(decks.small_droplet_2d())          # a smaller case: the second call
    small.step(3)
    _check(hip, small, "droplet2d")
    big.step(2)                                                 # the first context again, after a rebuild
    _check(hip, big, "dambreak3d again")
