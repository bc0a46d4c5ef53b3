"""The header-only C++ MDSystem class (include/lj_mdsystem.hpp) driven by a
C++ program the way the reference's task programs drive MDSystem
(configure, Integrate in a loop, read the public fields), against the
reference CPU path restated in the oracle from the same seeded state:
per-step U within |dU|/N <= 5e-6 over 200 EVN steps (N = 400, rho* = 0.05,
T* = 1.4, periodic, dt = 0.005) and NdNdr2 bit-exact with the reference GPU
binning on the final positions; the speed distribution (initvelo /
updatevelo) bit-exact against the oracle's counts."""
import os
import subprocess

import numpy as np
import pytest

from helpers import shear_fp64
from oracle_lib import ORC, OracleSystem, lattice_state, rdf, velocity_histogram

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "lennard-jones-cuda_amd", "lib", "mdsystem_driver")


def _run(*args):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is not built (run __graft_entry__.build())")
    r = subprocess.run([DRIVER, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {"step": [], "iteration_lines": []}
    for line in r.stdout.splitlines():
        if line.startswith("Iteration:"):  # MDSystem::Fluctuations' progress line
            out["iteration_lines"].append(line)
            continue
        key, *vals = line.split()
        if key == "step":
            out["step"].append([float(v) for v in vals])
        else:
            out[key] = [float(v) for v in vals]
    return out


def test_mdsystem_class_vs_reference_cpu():
    N, rho, T0, dt, steps, seed = 400, 0.05, 1.4, 0.005, 200, 3
    print(f"\n[case] C++ MDSystem N={N} EVN periodic {steps} steps", flush=True)
    out = _run(N, rho, T0, 0, 0, dt, steps, seed)
    L, pos, vel = lattice_state(N, rho, T0, seed)
    o = OracleSystem(N, T0, rho, 0, 0, nthreads=8)
    o.set_state(pos, vel)
    o.forces()
    o.params()
    U0, T0g, K0, V0, P0 = out["init"]
    assert abs(U0 - o.s.U) <= 1e-6 * abs(o.s.U) and abs(V0 - o.s.V) <= 1e-5 * abs(o.s.V)
    # Pshear: 0 as on the reference's GPU branch; UpdateShearPressure gives the
    # CPU path's value (MDSystem.cpp:299,309,335,353)
    ps0, ps = out["pshear"]
    _, scale = shear_fp64(pos, vel, L, True)
    print(f"  Pshear {ps:.10g} vs oracle {o.s.Pshear:.10g} (scale {scale:.3g})")
    assert ps0 == 0.0 and abs(ps - o.s.Pshear) <= 1e-7 * scale
    ho = o.integrate(dt, steps)
    st = np.array(out["step"])
    assert np.array_equal(st[:, 0], np.arange(steps))
    assert np.allclose(st[:, 1], dt * (np.arange(steps) + 1), rtol=0, atol=1e-12)
    dU = np.abs(st[:, 2] - ho[:, 0]) / N
    print(f"  max |dU|/N {dU.max():.2e}  final U {st[-1, 2]:.9g} vs {ho[-1, 0]:.9g}")
    assert dU.max() <= 5e-6  # about 10x the measured 5.4e-7
    assert np.max(np.abs(st[:, 3] - ho[:, 1])) <= 2e-4
    # NdNdr2 of the final positions: the reference GPU rule (MDSystem.cu:39), bit-exact
    p = np.array(out["pos"], np.float32)
    counts = rdf(p, L, True, 1, 256, dr2=o.s.rdf_dr2)
    assert np.array_equal(np.array(out["rdf"], np.int64), counts)
    assert out["cpu_mode_rejected"] == [1.0]
    # NdNdr2 after the loop was binned inside the last step, before the
    # periodic wrap: against the wrapped positions at most a few pairs move
    rs = np.array(out["rdfstep"], np.int64)
    re = rdf(np.array(out["posstep"], np.float32), L, True, 1, 256, dr2=o.s.rdf_dr2)
    assert rs.sum() == re.sum() and np.abs(rs - re).sum() <= 8
    _flucs_check(out, L, N)
    # speed distribution: initvelo(3.0, 0.05), then updatevelo after 10 steps,
    # against the oracle's counts with the reference's normalisation
    # (MDSystem.cpp:651-694), bit for bit
    nb = int(3.0 / 0.05) + 1
    c0 = velocity_histogram(np.array(out["vel0"], np.float32), 0.05, nb)
    v0 = np.array(out["velo0"]).reshape(-1, 2)
    assert v0.shape[0] == nb
    assert np.array_equal(v0[:, 0], 0.05 * (0.5 + np.arange(nb)))
    y0 = c0 / 0.05 / N
    assert np.array_equal(v0[:, 1], y0)
    shag = v0[1, 0] - v0[0, 0]
    c1 = velocity_histogram(np.array(out["vel1"], np.float32), shag, nb)
    v1 = np.array(out["velo1"]).reshape(-1, 2)
    assert np.array_equal(v1[:, 1], (y0 * 1 + c1 / shag / N) / 2)
    # Maxwell(v) (MDSystem.cpp:696-699): unit normalisation and mean speed sqrt(8T/pi)
    T, m = out["maxwell"][0], np.array(out["maxwell"][1:])
    v = 0.02 * np.arange(m.size)
    assert abs(np.trapezoid(m, v) - 1.0) <= 1e-4
    assert abs(np.trapezoid(v * m, v) - np.sqrt(8 * T / np.pi)) <= 1e-3


def _flucs_check(out, L, N):
    """Fluctuations(0.3) on posstep and on posstep2: the closed-cube counts of
    the oracle (MDSystem.cpp:708-719) in the reference's moment sums."""
    n1 = ORC.orc_fl_closed_cube(N, np.array(out["posstep"], np.float32), L, 0.3)
    n2 = ORC.orc_fl_closed_cube(N, np.array(out["posstep2"], np.float32), L, 0.3)
    assert out["flucs"] == [float(n1 + n2), float(n1 * n1 + n2 * n2), float(n1 ** 3 + n2 ** 3),
                            float(n1 ** 4 + n2 ** 4), 2.0]
    lines = out["iteration_lines"]
    assert len(lines) == 2 and lines[0].startswith(f"Iteration: 1\t<N> = {n1}\t")
    assert lines[1].endswith("Binom: 0.7")


def test_mdsystem_class_rdf_every_step_and_fluctuations():
    """Hard walls leave positions in place, so NdNdr2 after Integrate (binned
    inside the last step, as the reference's CalculateForces does) equals the
    GPU-rule binning of the positions the step left, bit for bit; Fluctuations
    keeps the reference's moment sums of the closed-cube count."""
    N, rho, T0, dt, steps, seed = 400, 0.05, 1.4, 0.005, 30, 5
    print(f"\n[case] C++ MDSystem N={N} TVN hard wall {steps} steps", flush=True)
    out = _run(N, rho, T0, 1, 1, dt, steps, seed)
    L = ORC.orc_box_length(N, rho)
    p = np.array(out["posstep"], np.float32)
    exp = rdf(p, L, False, 1, 256, dr2=ORC.orc_rdf_dr2(N))
    assert np.array_equal(np.array(out["rdfstep"], np.int64), exp)
    _flucs_check(out, L, N)
