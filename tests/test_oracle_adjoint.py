"""Pins of the FP64 adjoint oracle (CPU): every derivative it returns against
central finite differences of the FP64 forward quantities, and the full
trajectory gradient (reverse sweep over velocity Verlet) against central
differences of the FP64 trajectory loss."""
import numpy as np
import pytest

from helpers import perturbed_state, xyz
from oracle_lib import adjoint_pass_fp64, forces_fp64d, traj_grad_fp64, traj_loss_fp64


def _state(N=64, rho=0.5, seed=3):
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=seed, amp=0.08)
    return L, xyz(pos), xyz(vel)


@pytest.mark.parametrize("periodic", [0, 1])
def test_hvp_and_grad_w_vs_fd(periodic):
    L, x, _ = _state()
    rng = np.random.default_rng(1)
    z = rng.normal(size=x.shape)
    eps, s = 1.3, 1.05
    hz, f, gw, sc = adjoint_pass_fp64(x, z, L, periodic, eps, s)
    h = 1e-6
    fp = forces_fp64d(x + h * z, L, periodic, eps, s)[0]
    fm = forces_fp64d(x - h * z, L, periodic, eps, s)[0]
    fd = (fp - fm) / (2 * h)
    assert np.max(np.abs(hz - fd)) <= 1e-6 * np.max(np.abs(fd))
    # F and E, W agree with the force oracle
    F, _, E, W = forces_fp64d(x, L, periodic, eps, s)
    assert np.max(np.abs(f - F)) <= 1e-12 * np.max(np.abs(F))
    assert abs(sc[0] - E) <= 1e-12 * abs(E) and abs(sc[1] - W) <= 1e-12 * abs(W)
    assert abs(sc[5] - np.sum(F * z)) <= 1e-10 * np.sum(np.abs(F * z))
    # dW/dx along a random direction
    d = rng.normal(size=x.shape)
    Wp = forces_fp64d(x + h * d, L, periodic, eps, s)[3]
    Wm = forces_fp64d(x - h * d, L, periodic, eps, s)[3]
    assert abs((Wp - Wm) / (2 * h) - np.sum(gw * d)) <= 1e-6 * np.sum(np.abs(gw * d))


def test_sigma_derivatives_vs_fd():
    L, x, _ = _state(seed=5)
    z = np.random.default_rng(2).normal(size=x.shape)
    eps, s, h = 0.9, 1.1, 1e-6
    _, _, _, sc = adjoint_pass_fp64(x, z, L, 1, eps, s)
    Fp, _, Ep, Wp = forces_fp64d(x, L, 1, eps, s + h)
    Fm, _, Em, Wm = forces_fp64d(x, L, 1, eps, s - h)
    assert abs((Ep - Em) / (2 * h) - sc[2]) <= 1e-6 * abs(sc[2])
    assert abs((Wp - Wm) / (2 * h) - sc[3]) <= 1e-6 * abs(sc[3])
    fsz = np.sum((Fp - Fm) / (2 * h) * z)
    assert abs(fsz - sc[4]) <= 1e-6 * max(abs(sc[4]), 1.0)


@pytest.mark.parametrize("periodic", [0, 1])
def test_trajectory_gradient_vs_fd(periodic):
    L, x, v = _state(N=64, rho=0.4, seed=7)
    eps, s, dt, M = 1.0, 1.0, 0.005, 30
    out, g2, dx0 = traj_grad_fp64(x, v, L, periodic, eps, s, dt, M, pt=1.0, wp=1.0)
    assert np.allclose(out, traj_loss_fp64(x, v, L, periodic, eps, s, dt, M), rtol=1e-13, atol=0)
    h = 1e-5
    lp = traj_loss_fp64(x, v, L, periodic, eps + h, s, dt, M)[0]
    lm = traj_loss_fp64(x, v, L, periodic, eps - h, s, dt, M)[0]
    assert abs((lp - lm) / (2 * h) - g2[0]) <= 1e-6 * max(abs(g2[0]), 1.0)
    lp = traj_loss_fp64(x, v, L, periodic, eps, s + h, dt, M)[0]
    lm = traj_loss_fp64(x, v, L, periodic, eps, s - h, dt, M)[0]
    assert abs((lp - lm) / (2 * h) - g2[1]) <= 1e-6 * max(abs(g2[1]), 1.0)
    scale = np.sqrt(np.mean(dx0 ** 2))
    for (i, c) in [(0, 0), (5, 1), (17, 2), (40, 0), (63, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[i, c] += h
        xm[i, c] -= h
        fd = (traj_loss_fp64(xp, v, L, periodic, eps, s, dt, M)[0]
              - traj_loss_fp64(xm, v, L, periodic, eps, s, dt, M)[0]) / (2 * h)
        assert abs(fd - dx0[i, c]) <= 1e-5 * max(abs(dx0[i, c]), scale), (i, c, fd, dx0[i, c])
