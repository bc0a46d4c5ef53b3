"""The drop-in at link level, on the CPU: the reference's own MDSystem class,
compiled from its sources with USE_CUDA_TOOLKIT (its GPU branch,
MDSystem.cpp:9-25,114-140,237-251) and linked against libljhip.so instead of
MDSystem.cu (oracle/Makefile, oracle/_ref/libref_hip.so).

Every symbol the reference's translation units leave undefined and that the
reference's CUDA file would have provided must resolve in libljhip.so; the
library loads with immediate binding; without a device its GPU branch runs
through the compatibility set and each call reports the missing device
instead of exiting.  The compiled reference never leaves this machine, so
these checks run only where it was built (skipped elsewhere)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import ljamd
from ljamd import _lib as L
from oracle_lib import ORACLE_DIR, ref_lib

REF_HIP = os.path.join(ORACLE_DIR, "_ref", "libref_hip.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF_HIP),
                               reason="reference class not built here (oracle/_ref)")

# the C-ABI the reference's MDSystem.cpp declares under USE_CUDA_TOOLKIT (MDSystem.cpp:10-25)
COMPAT = {"allocateNBodyArrays", "allocateArray", "deleteNBodyArrays", "deleteArray",
          "calculateNForces", "copyArrayFromDevice", "copyArrayToDevice",
          "registerGLBufferObject", "unregisterGLBufferObject", "threadSync", "threadExit"}


def _nm(path, flag):
    out = subprocess.run(["nm", "-D", flag, path], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@needs_ref
def test_reference_class_links_against_libljhip():
    undefined = _nm(REF_HIP, "--undefined-only")
    used = COMPAT & undefined
    # the reference class calls these through its GPU branch (MDSystem.cpp:114-140,237-251)
    assert {"allocateArray", "deleteArray", "calculateNForces", "copyArrayToDevice",
            "copyArrayFromDevice"} <= used, used
    exported = _nm(L.LIB_PATH, "--defined-only")
    assert used <= exported, used - exported
    # the dynamic section names libljhip.so as a needed library
    dyn = subprocess.run(["readelf", "-d", REF_HIP], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    assert "libljhip.so" in dyn
    # immediate binding: every undefined symbol resolves at load time
    C.CDLL(REF_HIP, mode=os.RTLD_NOW | os.RTLD_LOCAL)


@needs_ref
@pytest.mark.skipif(ljamd.lib().lj_device_count() > 0, reason="a HIP device is present")
def test_reference_class_without_device_reports_errors(capfd):
    """The reference's GPU branch (useCUDA = true) on a machine without a
    device: construction, CalculateForces and a step run through the
    compatibility set, whose calls report the missing device (stderr and
    lj_last_error()) instead of exiting, as the reference's checkCudaErrors
    would."""
    lib = ref_lib("hip")
    h = lib.refh_create(64, 1.0, 0.5, 0, 0, 1, 256)
    try:
        assert h
        frc = np.zeros(4 * 64, np.float32)
        out = np.zeros(3, np.float64)
        rdf = np.zeros(256, np.int32)
        lib.refh_forces(h, frc, out, rdf)
        assert ljamd.lib().lj_last_error()  # the last compat call reported its failure
        lib.refh_integrate(h, 0.005, 1, None)
    finally:
        lib.refh_destroy(h)
    err = capfd.readouterr().err
    assert "ljhip: " in err and "no HIP device" in err, err[:400]
    assert "calculateNForces" in err
