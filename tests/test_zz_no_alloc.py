"""calculateNForcesNoAlloc (include/lj_hip.h): the compatibility force call
with its device scratch in a file-scope array instead of allocations.

Through the C-ABI, against calculateNForces on the same inputs: forces,
pressure and RDF bins bit-identical (the same kernels on the same plan), on
two calls in a row, the second with a smaller input, so that the second call
finds the first one's contents in the shared array; periodic and open boxes.
Sizes beyond the array are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import ljamd
from helpers import perturbed_state
from ljamd._lib import LJ_ERR_ARG, lib, ptr
from oracle_lib import ORC

pytestmark = pytest.mark.gpu


def _no_alloc(arr, pos4, L, periodic, dr2):
    """copyArrayToDevice + calculateNForcesNoAlloc + copyArrayFromDevice on the
    caller's device arrays (allocated once, outside the call)."""
    N = arr.N
    pos4 = np.ascontiguousarray(pos4, np.float32)
    lib().copyArrayToDevice(arr.d_pos, ptr(pos4), N)
    pressure = C.c_float(0.0)
    rdf = np.zeros(256, np.int32)
    rc = lib().calculateNForcesNoAlloc(arr.d_pos, arr.d_force, C.byref(pressure), N, float(L),
                                       int(periodic), ptr(rdf), float(dr2), 256, 1)
    assert rc == 0, lib().lj_last_error()
    frc = np.zeros(4 * N, np.float32)
    lib().copyArrayFromDevice(ptr(frc), arr.d_force, 0, N)
    return frc, pressure.value, rdf


@pytest.mark.parametrize("sizes,periodic", [((16384, 2048), 1), ((5000, 700), 0),
                                            ((65536, 257), 1)])
def test_no_alloc_matches_calculateNForces(sizes, periodic):
    for N in sizes:  # the second call is smaller and reuses the array
        print(f"\n[case] no-alloc N={N} periodic={periodic}", flush=True)
        rho = 0.85 if N > 1000 else 0.3
        L, pos, _ = perturbed_state(N, rho, 1.0, seed=N)
        dr2 = ORC.orc_rdf_dr2(N)
        arr = ljamd.CompatArrays(N)
        try:
            exp = arr.calculate(pos, np.float32(L), periodic, dr2)
            got = _no_alloc(arr, pos, np.float32(L), periodic, dr2)
        finally:
            arr.close()
        assert np.array_equal(got[0], exp[0])
        assert got[1] == exp[1]
        assert np.array_equal(got[2], exp[2])
        # a real histogram (the clamped last bin of the int counts wraps at
        # N = 65536, as the reference's int histogram would)
        assert np.count_nonzero(got[2]) > 10


def test_no_alloc_rejects_oversize():
    d = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = lib().calculateNForcesNoAlloc(d, d, None, 65537, 10.0, 1, None, 0.1, 256, 1)
    assert rc == LJ_ERR_ARG
    assert b"65536" in lib().lj_last_error()
    rc = lib().calculateNForcesNoAlloc(None, None, None, 100, 10.0, 1, None, 0.1, 256, 1)
    assert rc == LJ_ERR_ARG
