"""The task programs' host parts on the CPU (lennard-jones-cuda_amd/tasks):

  * TimeAverage of task_common.hpp against the oracle's restatement of the
    reference's TimeAverage / NumberStatistics (oracle/lj_oracle_flucs.c, itself
    pinned bit-exactly to the reference compiled here), bit for bit, directly
    and through the device-statistics layout the counters hand over;
  * the parameter files of the reference (input/*, data the reference's own
    tasks read) are read as the reference's ReadParametersFromFile reads them:
    `name value` pairs, '#' comments to the end of the line (the expected
    pairs are written out below);
  * without a device the programs stop with an error instead of running."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import tavg_eval

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TASKS = os.path.join(ROOT, "lennard-jones-cuda_amd", "tasks")
LIB = os.path.join(ROOT, "lennard-jones-cuda_amd", "lib")


@pytest.fixture(scope="module")
def tavg_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tavg") / "tavg_check")
    r = subprocess.run(["g++", "-std=c++11", "-O2", "-ffp-contract=off", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), "-I" + TASKS,
                        os.path.join(ROOT, "tests", "cpp", "tavg_check.cpp"), "-o", exe],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return exe


@pytest.mark.parametrize("seq", ["counts", "reals", "ar1", "two_valued"])
def test_time_average_matches_oracle(tavg_exe, seq):
    rng = np.random.default_rng(11)
    if seq == "counts":
        v = rng.binomial(400, 0.3, 2500).astype(np.float64)
    elif seq == "reals":
        v = 0.05 + 0.01 * rng.standard_normal(1500)
    elif seq == "ar1":
        v = np.zeros(3000)
        for t in range(1, v.size):
            v[t] = 0.8 * v[t - 1] + rng.standard_normal()
        v += 20.0
    else:
        v = (rng.random(1000) < 0.3).astype(np.float64)
    r = subprocess.run([tavg_exe], input="\n".join(repr(float(x)) for x in v), text=True,
                       stdout=subprocess.PIPE, check=True)
    rows = [np.array([float(x) for x in line.split()]) for line in r.stdout.splitlines()]
    exp = tavg_eval(v)
    for got in rows:
        assert np.array_equal(np.isnan(got), np.isnan(exp))
        ok = ~np.isnan(exp)
        assert np.array_equal(got[ok], exp[ok]), (got, exp)


# the reference's input files and what its reader makes of them
INPUTS = {
    "N400.Tst1.4.isotherm": [("N", "400"), ("T*", "1.4"), ("rho*_min", "0.01"),
                             ("rho*_max", "0.71"), ("drho*", "0.01"), ("teq", "50"),
                             ("tfin", "5000"), ("dt*", "0.004"), ("useCUDA", "1")],
    "N400.Tst1.4.rhost0.05": [("N", "400"), ("T*", "1.4"), ("rho*", "0.05"), ("teq", "50"),
                              ("tfin", "50000"), ("dt*", "0.004"), ("canonical", "1"),
                              ("subvolume_spacing", "0.05"), ("useCUDA", "1")],
    "N400.ust1.708.rhost0.05": [("N", "400"), ("u*", "1.708"), ("rho*", "0.05"), ("teq", "50"),
                                ("tfin", "50000"), ("dt*", "0.004"), ("canonical", "0"),
                                ("subvolume_spacing", "0.05"), ("useCUDA", "1")],
}
# the same files written out here (the reference's input directory is not on
# every machine that runs the tests): comments, blank lines, trailing text
INPUT_TEXT = {
    "N400.Tst1.4.isotherm": "# isotherm\n# Format: <parameter-name> <parameter-value>\n\nN 400\n"
                            "\n# The temperature\nT* 1.4\nrho*_min 0.01\nrho*_max 0.71\n"
                            "drho* 0.01\nteq 50.\ntfin 5000.\ndt* 0.004\nuseCUDA 1",
    "N400.Tst1.4.rhost0.05": "# canonical\nN 400\nT* 1.4\n# u* 1.708\nrho* 0.05\nteq 50.\n"
                             "tfin 50000.\ndt* 0.004\ncanonical 1\nsubvolume_spacing 0.05\n"
                             "useCUDA 1",
    "N400.ust1.708.rhost0.05": "N 400\n# T* 1.4\nu* 1.708\nrho* 0.05\nteq 50.\ntfin 50000.\n"
                               "dt* 0.004\ncanonical 0\nsubvolume_spacing 0.05\nuseCUDA 1\n",
}


@pytest.mark.parametrize("name", sorted(INPUTS))
def test_parameter_files_read_as_the_reference(tmp_path, name):
    prog = "run_isotherm" if name.endswith("isotherm") else "run_fluctuations"
    exe = os.path.join(LIB, prog)
    if not os.path.exists(exe):
        pytest.skip("task programs not built")
    path = tmp_path / name
    ref = os.path.join("/root/reference/input", name)
    path.write_text(open(ref).read() if os.path.exists(ref) else INPUT_TEXT[name])
    r = subprocess.run([exe, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, cwd=tmp_path, timeout=60)
    got = [tuple(line[len("Reading input parameter "):].split(" = "))
           for line in r.stdout.splitlines() if line.startswith("Reading input parameter ")]
    assert got == INPUTS[name]
    import ljamd
    if ljamd.lib().lj_device_count() == 0:
        assert r.returncode == 1 and "no HIP device" in r.stderr
