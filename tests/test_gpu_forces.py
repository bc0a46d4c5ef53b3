"""Force / energy / virial parity of the HIP force pass (context API).

Reference: MDSystem.cpp:237-311 (CPU path), MDSystem.cu:21-138 (GPU kernel).
Checks, on identical seeded states:
  * against the FP64 oracle: per-particle forces, E_pot, virial W;
  * against the reference-faithful CPU path (oracle restatement, bit-exact to
    the compiled reference): forces, V and the virial pressure term;
at N = 400 (rho* = 0.05), 2048 and 16384 (rho* = 0.85) with periodic and hard-wall
boxes, plus the exact lattice states of these workloads.

Tolerances (FP32 pair arithmetic, FP64 cross-slice sums):
  forces vs FP64:  global ||dF||/||F|| <= 2e-6,
                   per particle |dF_i| / max(|F_i|, F_rms) <= 1e-5;
  E_pot vs FP64: relative <= 1e-6;  virial W: relative <= 1e-5;
  forces vs reference CPU path: per particle <= 1e-5 + the reference's own
                   deviation from FP64 at that particle (it accumulates in float).
"""
import os

import numpy as np
import pytest

import ljamd
from helpers import force_errors, perturbed_state, rel, shear_fp64, xyz
from oracle_lib import ORC, OracleSystem, forces_fp64, forces_ref, lattice_state

pytestmark = pytest.mark.gpu

CASES = [
    (400, 0.05, 1.4, ljamd.BC_PERIODIC),
    (400, 0.05, 1.4, ljamd.BC_WALL),
    (2048, 0.85, 1.0, ljamd.BC_PERIODIC),
    (2048, 0.85, 1.0, ljamd.BC_WALL),
    (2048, 0.3, 1.0, ljamd.BC_NONE),
    (5000, 0.6, 1.0, ljamd.BC_PERIODIC),
    (16384, 0.85, 1.0, ljamd.BC_PERIODIC),
    (16384, 0.85, 1.0, ljamd.BC_WALL),
]


def _gpu_forces(N, rho, bc, pos, precision=ljamd.PREC_FP32, jslices=0):
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc, precision=precision, jslices=jslices) as s:
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        return s.force, obs


@pytest.mark.parametrize("N,rho,T0,bc", CASES)
def test_forces_vs_fp64(N, rho, T0, bc):
    L, pos, _ = perturbed_state(N, rho, T0, seed=11)
    print(f"\n[case] forces N={N} rho={rho} bc={bc}", flush=True)
    f4, obs = _gpu_forces(N, rho, bc, pos)
    F64, ew, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    err = force_errors(xyz(f4), F64)
    eE, eW = rel(obs["V"], E), rel(obs["W"], W)
    ew_err = np.max(np.abs(f4.reshape(N, 4)[:, 3] - ew[:, 0]) / np.maximum(np.abs(ew[:, 0]), 1.0))
    print(f"  vs FP64: {err}  E rel {eE:.2e}  W rel {eW:.2e}  Force.w max {ew_err:.2e}")
    assert err["global_rel"] <= 2e-6
    assert err["max_rel_floor"] <= 1e-5
    assert eE <= 1e-6 and eW <= 1e-5  # W has cancellation between repulsive/attractive parts
    assert ew_err <= 1e-5


@pytest.mark.parametrize("N,rho,T0,bc", [c for c in CASES if c[0] <= 5000])
def test_forces_vs_reference_cpu(N, rho, T0, bc):
    L, pos, _ = perturbed_state(N, rho, T0, seed=23)
    f4, obs = _gpu_forces(N, rho, bc, pos)
    fr, V, P, Ps, _ = forces_ref(pos, L, bc, ORC.orc_rdf_dr2(N), nthreads=8)
    F64, _, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    Fg, Fr = xyz(f4), xyz(fr)
    n = np.linalg.norm(F64, axis=1)
    floor = np.sqrt(np.mean(n ** 2))
    own = np.linalg.norm(Fr - F64, axis=1)  # the reference's own deviation
    d = np.linalg.norm(Fg - Fr, axis=1)
    excess = np.max((d - own) / np.maximum(n, floor))
    print(f"\n  vs ref CPU: global {force_errors(Fg, Fr)['global_rel']:.2e} excess {excess:.2e} "
          f"ref-own {np.max(own / np.maximum(n, floor)):.2e}  V {rel(obs['V'], V):.2e} "
          f"Pvir {rel(obs['W'] / 3, P):.2e}")
    assert excess <= 1e-5
    assert rel(obs["V"], V) <= 1e-5
    assert rel(obs["W"] / 3.0, P) <= 1e-5  # reference P before CalculateParameters = W/3


@pytest.mark.parametrize("N,rho", [(400, 0.05), (16384, 0.85), (2048, 0.85)])
def test_config_lattice_state(N, rho):
    """The workloads' own initial state (lattice, no displacement)."""
    L, pos, vel = lattice_state(N, rho, 1.0, seed=1)
    f4, obs = _gpu_forces(N, rho, ljamd.BC_PERIODIC, pos)
    F64, _, E, W = forces_fp64(pos, L, True)
    fr = forces_ref(pos, L, ljamd.BC_PERIODIC, ORC.orc_rdf_dr2(N), nthreads=8)[0]
    # A lattice is nearly force-free: each |F_i| is a sum of pair forces of
    # magnitude up to ~100 that cancel, so FP32 pair math leaves an absolute
    # error far above the tiny net force (the reference's own FP32 path is off
    # by 2.2e-5 relative at N = 2048). The bar is the reference's own error.
    err = force_errors(xyz(f4), F64)
    own = force_errors(xyz(fr), F64)
    print(f"\n  lattice N={N}: {err} ref-own {own} E {rel(obs['V'], E):.2e} "
          f"W {rel(obs['W'], W):.2e}")
    assert err["global_rel"] <= max(2e-6, own["global_rel"])
    assert err["max_rel_floor"] <= max(1e-5, own["max_rel_floor"])
    assert rel(obs["V"], E) <= 1e-6 and rel(obs["W"], W) <= 1e-5


@pytest.mark.parametrize("prec", [ljamd.PREC_FP32, ljamd.PREC_COMPENSATED])
@pytest.mark.parametrize("jslices", [1, 3, 64])
def test_slicing_and_precision_modes(prec, jslices):
    N, rho = 3000, 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=5)
    f4, obs = _gpu_forces(N, rho, ljamd.BC_PERIODIC, pos, precision=prec, jslices=jslices)
    F64, _, E, W = forces_fp64(pos, L, True)
    err = force_errors(xyz(f4), F64)
    print(f"\n  prec={prec} S={jslices}: {err}")
    assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5
    assert rel(obs["V"], E) <= 1e-6


def test_eps_sigma():
    N, rho = 1500, 0.5
    eps, sigma = 1.7, 1.15
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=3)
    pos = pos.reshape(N, 4)
    pos[:, :3] *= sigma  # same reduced configuration, scaled box
    pos = pos.reshape(-1).copy()
    with ljamd.LJSystem(N, rho / sigma ** 3, 1.0, eps=eps, sigma=sigma) as s:
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        f4 = s.force
        Ls = s.L
    F64, _, E, W = forces_fp64(pos, Ls, True, eps=eps, sigma=sigma)
    err = force_errors(xyz(f4), F64)
    assert err["global_rel"] <= 2e-6, err
    assert rel(obs["V"], E) <= 1e-6 and rel(obs["W"], W) <= 1e-6


def test_deterministic_repeat():
    N, rho = 4096, 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=9)
    a, oa = _gpu_forces(N, rho, ljamd.BC_PERIODIC, pos)
    b, ob = _gpu_forces(N, rho, ljamd.BC_PERIODIC, pos)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert oa["V"] == ob["V"] and oa["W"] == ob["W"]


def test_translation_invariance_periodic():
    """Shifting every particle by a vector (mod L) leaves forces unchanged."""
    N, rho = 2000, 0.7
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=4)
    p = pos.reshape(N, 4).copy()
    p2 = p.copy()
    p2[:, :3] = np.mod(p2[:, :3] + np.float32(0.37 * L), np.float32(L)).astype(np.float32)
    a, oa = _gpu_forces(N, rho, ljamd.BC_PERIODIC, p.reshape(-1))
    b, ob = _gpu_forces(N, rho, ljamd.BC_PERIODIC, p2.reshape(-1))
    err = force_errors(xyz(b), xyz(a))
    assert err["global_rel"] <= 1e-5 and rel(ob["V"], oa["V"]) <= 1e-6


def test_newton_third_law():
    N, rho = 3000, 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=8)
    f4, _ = _gpu_forces(N, rho, ljamd.BC_PERIODIC, pos)
    F = xyz(f4)
    tot = np.abs(F.sum(axis=0)).max()
    assert tot <= 1e-5 * np.abs(F).sum() / N * np.sqrt(N)


@pytest.mark.parametrize("N,bc", [(12800, ljamd.BC_PERIODIC), (13001, ljamd.BC_PERIODIC),
                                  (13001, ljamd.BC_WALL), (16384, ljamd.BC_NONE)])
def test_pair_once_kernel(N, bc):
    """The pair-once kernel (FP32 mode, N >= 3072, groups of 256; 13001 ends in
    a partial group) against FP64 and against the sliced kernel (jslices = 32)
    on the same state; two evaluations are bit-identical."""
    rho = 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=13)
    print(f"\n[case] pair-once N={N} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc) as s:
        pair_once, nslots = s.force_plan()
        assert pair_once and nslots in ((N + 511) // 512, (N + 255) // 256)
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        f4 = s.force
        obs2 = s.CalculateForces()
        assert np.array_equal(s.force, f4) and obs2["V"] == obs["V"] and obs2["W"] == obs["W"]
    f_sl, obs_sl = _gpu_forces(N, rho, bc, pos, jslices=32)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc, jslices=32) as s:
        assert s.force_plan() == (False, 32)
    F64, ew, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    err = force_errors(xyz(f4), F64)
    err_sl = force_errors(xyz(f_sl), F64)
    ew_err = np.max(np.abs(f4.reshape(N, 4)[:, 3] - ew[:, 0]) / np.maximum(np.abs(ew[:, 0]), 1.0))
    print(f"  pair-once {err} sliced {err_sl} E {rel(obs['V'], E):.2e} W {rel(obs['W'], W):.2e} "
          f"Force.w {ew_err:.2e}")
    assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5
    assert rel(obs["V"], E) <= 1e-6 and rel(obs["W"], W) <= 1e-5
    assert ew_err <= 1e-5
    assert force_errors(xyz(f4), xyz(f_sl))["global_rel"] <= 2e-6


@pytest.mark.parametrize("N,bc", [(16384, ljamd.BC_PERIODIC), (13001, ljamd.BC_WALL),
                                  (8192, ljamd.BC_NONE)])
def test_compensated_pair_once(N, bc):
    """The compensated mode's pair-once pass (N >= 3072; blocks of 32 pairs in
    FP32 added by TwoSum on both sides of every pair; 13001 ends in a partial
    group) against FP64, the FP32 pair-once pass and the compensated sliced
    kernel (LJ_PAIR_ONCE_COMP=0) on one state; two evaluations are
    bit-identical."""
    rho = 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=17)
    print(f"\n[case] compensated pair-once N={N} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc, precision=ljamd.PREC_COMPENSATED) as s:
        assert s.force_plan() == (True, (N + 255) // 256)
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        fc = s.force.copy()
        obs2 = s.CalculateForces()
        assert np.array_equal(s.force, fc) and obs2["V"] == obs["V"] and obs2["W"] == obs["W"]
    os.environ["LJ_PAIR_ONCE_COMP"] = "0"
    try:
        with ljamd.LJSystem(N, rho, 1.0, boundary=bc, precision=ljamd.PREC_COMPENSATED) as s:
            assert not s.force_plan()[0]
            s.set_state(pos, np.zeros_like(pos))
            s.CalculateForces()
            fsl = s.force.copy()
    finally:
        os.environ.pop("LJ_PAIR_ONCE_COMP", None)
    f32, _ = _gpu_forces(N, rho, bc, pos)
    F64, ew, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    ec, esl, e32 = (force_errors(xyz(f), F64) for f in (fc, fsl, f32))
    ew_err = np.max(np.abs(fc.reshape(N, 4)[:, 3] - ew[:, 0]) / np.maximum(np.abs(ew[:, 0]), 1.0))
    print(f"  compensated pair-once {ec}\n  compensated sliced {esl}\n  fp32 pair-once {e32}\n"
          f"  E {rel(obs['V'], E):.2e} W {rel(obs['W'], W):.2e} Force.w {ew_err:.2e}")
    assert ec["global_rel"] <= 2e-6 and ec["max_rel_floor"] <= 1e-5
    assert ec["global_rel"] <= e32["global_rel"]
    assert rel(obs["V"], E) <= 1e-6 and rel(obs["W"], W) <= 1e-5 and ew_err <= 1e-5


@pytest.mark.parametrize("N,rho,bc", [(2048, 0.85, ljamd.BC_PERIODIC), (2000, 0.3, ljamd.BC_WALL),
                                      (1500, 0.2, ljamd.BC_NONE), (16384, 0.85, ljamd.BC_PERIODIC)])
def test_shear_pressure_vs_reference_cpu(N, rho, bc):
    """Pshear as the reference CPU path leaves it after CalculateForces and
    CalculateParameters (MDSystem.cpp:299,309,335,353), on demand: against the
    oracle's restatement (bit-exact to the reference) and an FP64 sum, within
    1e-7 of the sum of the terms' magnitudes (measured on MI355X: at most
    8.4e-9 from the oracle; the oracle 1.1e-9 from the FP64 sum).  The
    observables keep Pshear = 0, as the reference's GPU branch does."""
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=29)
    print(f"\n[case] shear pressure N={N} rho={rho} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc) as s:
        s.set_state(pos, vel)
        ps = s.shear_pressure()
        obs = s.CalculateParameters()
    o = OracleSystem(N, 1.0, rho, ljamd.EVN, bc, nthreads=16)
    o.set_state(pos, vel)
    o.forces()
    o.params()
    ref = o.s.Pshear
    p64, scale = shear_fp64(pos, vel, L, bc == ljamd.BC_PERIODIC)
    print(f"  GPU {ps:.10g}  oracle {ref:.10g}  FP64 {p64:.10g}  scale {scale:.4g}  "
          f"|GPU - oracle| / scale {abs(ps - ref) / scale:.2e}")
    assert abs(ps - ref) <= 1e-7 * scale and abs(ps - p64) <= 1e-7 * scale
    assert abs(ref - p64) <= 1e-7 * scale
    assert obs["Pshear"] == 0.0

