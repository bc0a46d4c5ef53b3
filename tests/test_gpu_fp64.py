"""The FP64 precision mode (LJ_PREC_FP64) against the FP64 CPU oracle.

Forces/E/W of double positions vs orc_forces_fp64d (relative 1e-12), and EVN /
TVN trajectories vs orc_fp64_step, the reference's integrator formulas in
double (positions 1e-9 absolute, energies 1e-10 relative after 40 steps).
The FP32 and compensated modes are then measured against it on one state.
"""
import numpy as np
import pytest

import ljamd
from helpers import force_errors, perturbed_state, xyz
from oracle_lib import fp64_steps, forces_fp64d

pytestmark = pytest.mark.gpu


def _x4(x3):
    a = np.zeros((x3.shape[0], 4))
    a[:, :3] = x3
    return a.ravel()


@pytest.mark.parametrize("bc", [ljamd.BC_PERIODIC, ljamd.BC_WALL, ljamd.BC_NONE])
def test_forces_fp64(bc):
    N, rho = 2048, 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=41)
    rng = np.random.default_rng(2)
    x = xyz(pos) + rng.uniform(-1e-3, 1e-3, (N, 3))  # not float-representable
    print(f"\n[case] fp64 forces N={N} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc, eps=1.3, sigma=0.9,
                        precision=ljamd.PREC_FP64) as s:
        s.set_array_f64(ljamd.ARRAY_POS, _x4(x))
        assert np.array_equal(s.get_array_f64(ljamd.ARRAY_POS).reshape(N, 4)[:, :3], x)
        obs = s.CalculateForces()
        f = s.get_array_f64(ljamd.ARRAY_FORCE).reshape(N, 4)
        f32 = s.force.reshape(N, 4)
    F, ew, E, W = forces_fp64d(x, L, bc == ljamd.BC_PERIODIC, 1.3, 0.9)
    d = np.linalg.norm(f[:, :3] - F) / np.linalg.norm(F)
    print(f"  ||dF||/||F|| {d:.2e}  E {obs['V']:.12g} vs {E:.12g}  W {obs['W']:.12g} vs {W:.12g}")
    assert d <= 1e-12
    assert abs(obs["V"] - E) <= 1e-12 * abs(E) and abs(obs["W"] - W) <= 1e-11 * abs(W)
    assert np.max(np.abs(f[:, 3] - ew[:, 0])) <= 1e-12 * np.max(np.abs(ew[:, 0]))
    # the float mirror is the rounded double force
    assert np.array_equal(f32[:, :3], f[:, :3].astype(np.float32))


@pytest.mark.parametrize("canon,bc", [(ljamd.EVN, ljamd.BC_WALL), (ljamd.EVN, ljamd.BC_PERIODIC),
                                      (ljamd.TVN, ljamd.BC_PERIODIC), (ljamd.TVN, ljamd.BC_WALL)])
def test_trajectory_fp64(canon, bc):
    N, rho, T0, dt, n = 1500, 0.3, 1.0, 0.002, 40
    L, pos, vel = perturbed_state(N, rho, T0, seed=43)
    print(f"\n[case] fp64 trajectory N={N} canon={canon} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, T0, canonical=canon, boundary=bc,
                        precision=ljamd.PREC_FP64) as s:
        s.set_state(pos, vel)
        s.CalculateForces()
        hist = s.Integrate(dt, n)
        x = s.get_array_f64(ljamd.ARRAY_POS).reshape(N, 4)[:, :3]
        v = s.get_array_f64(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3]
    xo, vo, _, obs = fp64_steps(xyz(pos), xyz(vel), L, bc, canon, T0, dt, n)
    dx, dv = np.max(np.abs(x - xo)), np.max(np.abs(v - vo))
    dV = np.max(np.abs(hist[:, 3] - obs[:, 0])) / np.max(np.abs(obs[:, 0]))
    dK = np.max(np.abs(hist[:, 2] - obs[:, 1])) / np.max(np.abs(obs[:, 1]))
    print(f"  max|dx| {dx:.2e} max|dv| {dv:.2e} V rel {dV:.2e} K rel {dK:.2e}")
    assert dx <= 1e-9 and dv <= 1e-8
    assert dV <= 1e-10 and dK <= 1e-10


def test_modes_against_fp64():
    """One state of the precision workload's density: FP32 and compensated
    pair sums against the FP64 mode on the same (float) positions."""
    N, rho = 16384, 0.3
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=44)
    res = {}
    for prec in (ljamd.PREC_FP64, ljamd.PREC_FP32, ljamd.PREC_COMPENSATED):
        with ljamd.LJSystem(N, rho, 1.0, boundary=ljamd.BC_WALL, precision=prec) as s:
            s.set_state(pos)
            obs = s.CalculateForces()
            res[prec] = (s.get_array_f64(ljamd.ARRAY_FORCE).reshape(N, 4)[:, :3], obs["V"],
                         obs["W"])
    F64, E64, W64 = res[ljamd.PREC_FP64]
    for prec in (ljamd.PREC_FP32, ljamd.PREC_COMPENSATED):
        F, E, W = res[prec]
        err = force_errors(F, F64)
        print(f"\n  prec {prec}: {err} E rel {abs(E - E64) / abs(E64):.2e} "
              f"W rel {abs(W - W64) / abs(W64):.2e}")
        assert err["global_rel"] <= 2e-6 and abs(E - E64) <= 1e-6 * abs(E64)
    e32 = force_errors(res[ljamd.PREC_FP32][0], F64)["global_rel"]
    ecp = force_errors(res[ljamd.PREC_COMPENSATED][0], F64)["global_rel"]
    assert ecp <= e32


def test_setup_operations_fp64():
    N, rho, T0 = 1000, 0.5, 1.5
    with ljamd.LJSystem(N, rho, T0, boundary=ljamd.BC_PERIODIC, precision=ljamd.PREC_FP64) as s:
        s.sample_initial_state(5)
        v = s.get_array_f64(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3]
        s.set_array_f64(ljamd.ARRAY_VEL, _x4(1.1 * v))
        s.RenormalizeVelocities(True)
        v2 = s.get_array_f64(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3]
        assert abs(np.sum(v2 * v2) / (3 * N) - T0) <= 1e-13 * T0
        obs = s.CalculateParameters()
        assert abs(obs["T"] - T0) <= 1e-13 * T0


def test_precision_workload_first_steps_vs_oracle():
    """The precision study's workload at full size (N = 131072, rho* = 0.3,
    hard wall, EVN, dt = 0.002): the FP64 mode against the FP64 CPU oracle
    over its first 2 steps."""
    from oracle_lib import lattice_state
    N, rho, T0, dt, n = 131072, 0.3, 1.0, 0.002, 2
    L, pos, vel = lattice_state(N, rho, T0, 5)
    print(f"\n[case] precision workload N={N} fp64 vs oracle, {n} steps", flush=True)
    with ljamd.LJSystem(N, rho, T0, boundary=ljamd.BC_WALL, precision=ljamd.PREC_FP64) as s:
        s.set_state(pos, vel)
        s.CalculateForces()
        hist = s.Integrate(dt, n)
        x = s.get_array_f64(ljamd.ARRAY_POS).reshape(N, 4)[:, :3]
        f = s.get_array_f64(ljamd.ARRAY_FORCE).reshape(N, 4)[:, :3]
    xo, vo, fo, obs = fp64_steps(xyz(pos), xyz(vel), L, ljamd.BC_WALL, ljamd.EVN, T0, dt, n,
                                 nthreads=16)
    dF = np.linalg.norm(f - fo) / np.linalg.norm(fo)
    print(f"  max|dx| {np.max(np.abs(x - xo)):.2e}  ||dF||/||F|| {dF:.2e}  "
          f"V {hist[-1, 3]:.12g} vs {obs[-1, 0]:.12g}")
    assert np.max(np.abs(x - xo)) <= 1e-11 and dF <= 1e-12
    assert np.all(np.abs(hist[:, 3] - obs[:, 0]) <= 1e-12 * np.abs(obs[:, 0]))


def test_precision_workload_modes_full_size():
    """FP32 and compensated against the FP64 mode on the precision study's
    workload at full size (N = 131072, rho* = 0.3, hard wall, EVN,
    dt = 0.002), over its first 100 steps from one seeded state: the first
    state's forces, E_pot and P over every step.  Measured on MI355X (the
    compensated mode on the pair-once pass): forces 7.7e-7 / 5.7e-7 (global),
    E_pot 3.8e-7 / 2.3e-7, P 1.1e-5 / 8.3e-6; the tolerances are about 3x
    those."""
    import types

    from ljamd import precision as pr

    a = types.SimpleNamespace(N=131072, rho=0.3, T0=1.0, dt=0.002, steps=100, chunk=100, seed=5)
    print(f"\n[case] precision workload N={a.N}, 3 modes x {a.steps} steps", flush=True)
    runs = {name: pr.run_mode(name, prec, a, lambda msg: None) for name, prec in pr.MODES}
    ref = runs["fp64"]
    errs = {}
    for name in ("fp32", "compensated"):
        r = runs[name]
        fe = force_errors(r["F0"], ref["F0"])["global_rel"]
        dE = np.max(np.abs(r["hist"][:, 3] - ref["hist"][:, 3]) / np.abs(ref["hist"][:, 3]))
        dP = np.max(np.abs(r["hist"][:, 4] - ref["hist"][:, 4]) / np.abs(ref["hist"][:, 4]))
        errs[name] = fe
        print(f"  {name}: forces {fe:.2e}  E_pot {dE:.2e}  P {dP:.2e}  "
              f"force pass {r['force_us']:.0f} us (fp64 {ref['force_us']:.0f} us)")
        assert fe <= 2.5e-6 and dE <= 1.2e-6 and dP <= 3.5e-5
    # the compensated per-particle sums are the more accurate ones
    assert errs["compensated"] <= errs["fp32"]
