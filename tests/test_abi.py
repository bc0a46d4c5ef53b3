"""The C-ABI library loads and exports every symbol include/lj_hip.h declares;
argument errors are reported without a device; no compute calls (CPU suite)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import ljamd
from ljamd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_builds_and_loads():
    assert os.path.exists(L.LIB_PATH)
    assert ljamd.lib().lj_version().startswith(b"ljhip")


def test_every_declared_symbol_is_exported():
    decl = declared_functions()
    assert {"calculateNForces", "copyArrayToDevice", "lj_create", "lj_integrate"} <= decl
    missing = decl - exported_symbols()
    assert not missing, missing


def test_python_signatures_cover_header():
    decl = declared_functions()
    sig = {n for n, _, _ in L.SIGNATURES}
    assert decl <= sig, decl - sig


def test_reference_compat_signatures():
    """Same names/arity as the reference's declarations (MDSystem.cpp:13-24)."""
    hdr = open(os.path.join(ROOT, "include", "lj_hip.h")).read()
    for decl in [
        "void allocateNBodyArrays(float* vel[2], int numBodies);",
        "void allocateArray(float** dest, int number);",
        "void deleteNBodyArrays(float* vel[2]);",
        "void deleteArray(float* arr);",
        "void copyArrayFromDevice(float* host, const float* device, unsigned int pbo, int numBodies);",
        "void copyArrayToDevice(float* device, const float* host, int numBodies);",
        "void registerGLBufferObject(unsigned int pbo);",
        "void unregisterGLBufferObject(unsigned int pbo);",
        "void threadSync(void);",
        "void threadExit(void);",
    ]:
        assert decl in hdr, decl
    assert re.search(r"void calculateNForces\(float\* Pos, float\* Force, float\* host_pressure, "
                     r"int numBodies,\s+float host_L, int Lperiodic, int\* host_RDF, "
                     r"float host_dr2,\s+int blockSize, int q\);", hdr)


def test_config_defaults_match_reference():
    cfg = L.LJConfig()
    ljamd.lib().lj_config_init(C.byref(cfg))
    # MDSystemConfiguration() defaults, MDSystem.h:32-39
    assert (cfg.N, cfg.T0, cfg.rho, cfg.canonical, cfg.boundary) == (128, 1.5, 0.2, 0, 0)
    assert (cfg.eps, cfg.sigma) == (1.0, 1.0)


def test_dr2_rule():
    from oracle_lib import ORC
    for N in (16, 100, 400, 2048, 16384, 131072):
        assert ljamd.lib().lj_rdf_dr2_default(N) == ORC.orc_rdf_dr2(N)


def test_argument_errors_without_compute():
    cfg = L.LJConfig()
    ljamd.lib().lj_config_init(C.byref(cfg))
    h = C.c_void_p()
    for field, bad in (("N", 0), ("rho", -1.0), ("boundary", 5), ("canonical", 2), ("eps", 0.0)):
        c2 = L.LJConfig.from_buffer_copy(cfg)
        setattr(c2, field, bad)
        rc = ljamd.lib().lj_create(C.byref(c2), C.byref(h))
        assert rc == L.LJ_ERR_ARG, (field, rc)
        assert ljamd.lib().lj_last_error()
    assert ljamd.lib().lj_create(None, C.byref(h)) == L.LJ_ERR_ARG
    for fn in ("lj_destroy",):
        assert getattr(ljamd.lib(), fn)(None) == 0
    assert ljamd.lib().lj_synchronize(None) == L.LJ_ERR_ARG
    assert ljamd.lib().lj_integrate(None, 0.005, 1, None) == L.LJ_ERR_ARG


@pytest.mark.skipif(ljamd.lib().lj_device_count() > 0, reason="a HIP device is present")
def test_fails_loudly_without_device():
    cfg = L.LJConfig()
    ljamd.lib().lj_config_init(C.byref(cfg))
    h = C.c_void_p()
    rc = ljamd.lib().lj_create(C.byref(cfg), C.byref(h))
    assert rc == L.LJ_ERR_DEVICE
    assert b"no HIP device" in ljamd.lib().lj_last_error()
    with pytest.raises(ljamd.LJError):
        ljamd.LJSystem(100, 0.5)


def test_compat_null_arguments_reported():
    ljamd.lib().allocateArray(None, 4)
    assert b"allocateArray" in ljamd.lib().lj_last_error()
    ljamd.lib().copyArrayToDevice(None, None, 4)
    assert b"copyArrayToDevice" in ljamd.lib().lj_last_error()
    ljamd.lib().registerGLBufferObject(3)  # no-op as in the reference


def test_mdsystem_class_header_compiles():
    """The C++ class over the C-ABI builds with the reference's C++ standard
    (C++11) and with C++17, warnings as errors."""
    src = os.path.join(ROOT, "tests", "cpp", "mdsystem_driver.cpp")
    for std in ("c++11", "c++17"):
        r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            "-I" + os.path.join(ROOT, "include"), src],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message, no work done."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
