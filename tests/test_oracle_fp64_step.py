"""Pins of the FP64 integrator oracle (orc_fp64_step, the reference step of the
precision mode): the reference-faithful float step from the same state, time
reversibility, the O(dt^2) energy error and the TVN half-step invariant."""
import numpy as np
import pytest

from helpers import perturbed_state, xyz
from oracle_lib import OracleSystem, fp64_steps


@pytest.mark.parametrize("canon", [0, 1])
@pytest.mark.parametrize("bc", [0, 1])
def test_one_step_matches_reference_float_step(canon, bc):
    # the reference CPU step (float state, double expressions) from the same
    # state differs from the FP64 step only by float rounding
    N, rho, T0, dt = 300, 0.4, 1.2, 0.005
    L, pos, vel = perturbed_state(N, rho, T0, seed=31)
    o = OracleSystem(N, T0, rho, canon, bc, nthreads=4)
    o.set_state(pos, vel)
    o.forces()
    o.integrate(dt, 1)
    x, v, f, obs = fp64_steps(xyz(pos), xyz(vel), L, bc, canon, T0, dt, 1)
    assert np.max(np.abs(x - xyz(o.pos))) <= 2e-5 * L
    assert np.max(np.abs(v - xyz(o.vel))) <= 1e-4 * np.max(np.abs(v))
    assert np.max(np.abs(f - xyz(o.frc))) <= 1e-4 * np.max(np.abs(f))
    assert abs(obs[0, 0] - o.s.V) <= 1e-5 * abs(o.s.V)
    assert abs(obs[0, 1] - o.s.K) <= 1e-5 * abs(o.s.K)


def test_time_reversible_open_box():
    N, rho, dt, n = 200, 0.3, 0.004, 40
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=8)
    x0, v0 = xyz(pos), xyz(vel)
    x, v, _, _ = fp64_steps(x0, v0, L, 2, 0, 1.0, dt, n)
    x, v, _, _ = fp64_steps(x, -v, L, 2, 0, 1.0, dt, n)
    assert np.max(np.abs(x - x0)) <= 1e-9
    assert np.max(np.abs(v + v0)) <= 1e-8


def test_energy_error_scales_dt2():
    N, rho, T = 216, 0.5, 1.0
    L, pos, vel = perturbed_state(N, rho, T, seed=4)
    errs = []
    for dt in (0.004, 0.002):
        n = int(round(0.2 / dt))
        x, v, f, obs = fp64_steps(xyz(pos), xyz(vel), L, 0, 0, T, dt, n)
        _, _, _, o0 = fp64_steps(xyz(pos), xyz(vel), L, 0, 0, T, 0.0, 1)
        U0 = o0[0, 0] + o0[0, 1]
        U = obs[:, 0] + obs[:, 1]
        errs.append(np.max(np.abs(U - U0)))
    assert 3.0 <= errs[0] / errs[1] <= 5.0, errs


def test_tvn_half_step_invariant():
    # v_new = 2 chi v' - v_old, so sum(((v_new + v_old) / 2)^2) / 3N = T0
    N, rho, T0, dt = 250, 0.6, 1.3, 0.005
    L, pos, vel = perturbed_state(N, rho, 0.7, seed=2)
    x, v, f = xyz(pos), xyz(vel), None
    for _ in range(5):
        x1, v1, f, _ = fp64_steps(x, v, L, 2, 1, T0, dt, 1)
        tv = 0.5 * (v1 + v)
        assert abs(np.sum(tv * tv) / (3 * N) - T0) <= 1e-12 * T0
        x, v = x1, v1
