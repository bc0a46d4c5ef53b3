"""Differentiable LJ path (HIP adjoint kernel + reverse sweep) against the FP64
CPU oracle and central finite differences.

Workload (the differentiable configuration): N = 2048, rho* = 0.85, T* = 1.0,
periodic, 200 velocity-Verlet EVN steps of dt = 0.005 in FP32, checkpoints
every 20 steps, loss = <E_pot>_t + (<P>_t - 1)^2, gradients with respect to
eps, sigma and the initial positions.
Tolerances: adjoint pass vs FP64 pass 1e-5 (relative, FP32 pair math);
trajectory gradient vs FP64 CPU gradient 1e-4 relative (x0: relative to the
rms component); vs central differences of the FP64 CPU loss (h = 1e-4): 1e-4
in an open box, and in the periodic box 1e-4 for eps but only 3e-3 (sigma)
and 1e-3 (x0), since the minimum-image force jumps at |dx| = L/2.
"""
import numpy as np
import pytest

import ljamd
from helpers import perturbed_state, xyz
from oracle_lib import adjoint_pass_fp64, lattice_state, traj_grad_fp64, traj_loss_fp64

pytestmark = pytest.mark.gpu

N_LIQ, RHO_LIQ, T_LIQ, DT4, M4, K4 = 2048, 0.85, 1.0, 0.005, 200, 20


@pytest.mark.parametrize("periodic", [True, False])
def test_adjoint_pass_vs_fp64(periodic):
    N, rho = 1500, 0.6
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=12)
    z = np.random.default_rng(3).normal(size=(N, 4)).astype(np.float32)
    z[:, 3] = 0
    bc = ljamd.BC_PERIODIC if periodic else ljamd.BC_NONE
    eps, sigma = 1.2, 0.95
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc, eps=eps, sigma=sigma) as s:
        hz, f, gw, sc = s.adjoint_pass(pos, z)
        Ls = s.L
    hz0, f0, gw0, sc0 = adjoint_pass_fp64(xyz(pos), z[:, :3].astype(np.float64), Ls, periodic,
                                         eps, sigma)
    for got, exp in ((hz[:, :3], hz0), (f[:, :3], f0), (gw[:, :3], gw0)):
        assert np.linalg.norm(got - exp) <= 1e-5 * np.linalg.norm(exp)
    for k in range(6):
        assert abs(sc[k] - sc0[k]) <= 1e-5 * max(abs(sc0[k]), 1e-3 * np.abs(sc0[:6]).max()), k


def _liquid_n2048_state(seed=4):
    return lattice_state(N_LIQ, RHO_LIQ, T_LIQ, seed)


def test_trajectory_gradient_liquid_n2048():
    L, pos, vel = _liquid_n2048_state()
    print("\n[case] adjoint liquid N=2048 M=200 k=20", flush=True)
    with ljamd.LJSystem(N_LIQ, RHO_LIQ, T_LIQ) as s:
        s.set_state(pos, vel)
        loss_g = s.trajectory_loss(DT4, M4)
        out, g2, dx0, times = s.trajectory_grad(DT4, M4, K4)
        out1, g21, dx01, times1 = s.trajectory_grad(DT4, M4, 1)    # store every state
        out7, g27, dx07, _ = s.trajectory_grad(DT4, M4, 7)
        # state untouched by the gradient calls
        assert np.array_equal(s.pos, pos)
    # checkpoint spacing does not change the result (deterministic recompute)
    assert np.array_equal(g2, g21) and np.array_equal(g2, g27)
    assert np.array_equal(dx0, dx01) and np.array_equal(dx0, dx07)
    assert np.array_equal(out, loss_g)
    x, v = xyz(pos), xyz(vel)
    ref_out, ref_g2, ref_dx0 = traj_grad_fp64(x, v, L, 1, 1.0, 1.0, DT4, M4)
    rel_loss = abs(out[0] - ref_out[0]) / abs(ref_out[0])
    rel_g = np.abs(g2 - ref_g2) / np.abs(ref_g2)
    d = dx0[:, :3] - ref_dx0
    rel_x = np.linalg.norm(d) / np.linalg.norm(ref_dx0)
    rms = np.sqrt(np.mean(ref_dx0 ** 2))
    max_x = np.max(np.abs(d)) / rms
    fwd, rec, rev = times
    print(f"  loss {out[0]:.9g} (FP64 {ref_out[0]:.9g}, rel {rel_loss:.2e}); dL/deps {g2[0]:.9g} vs "
          f"{ref_g2[0]:.9g}; dL/dsigma {g2[1]:.9g} vs {ref_g2[1]:.9g}; rel {rel_g}")
    print(f"  dL/dx0: ||d||/||g|| {rel_x:.2e}, max|d|/rms {max_x:.2e}")
    print(f"  cost: forward {fwd:.2f} ms, recompute {rec:.2f} ms, reverse {rev:.2f} ms; "
          f"gradient/forward {(rec + rev) / fwd:.2f}x (k=20), {times1[2] / times1[0]:.2f}x (k=1)")
    assert rel_loss <= 1e-5
    assert np.all(rel_g <= 1e-4)
    assert rel_x <= 1e-4 and max_x <= 1e-3
    # Central differences of the FP64 loss, h = 1e-4. With the minimum image and
    # no cutoff the pair force jumps where a pair separation crosses L/2, so the
    # loss is only piecewise smooth in sigma and x0: differences include the
    # shifts of those crossing times, which the pathwise gradient leaves out
    # (FP64 oracle, this case: sigma 1.1e-3, x0 1.9e-4, eps 3e-5). The smooth
    # open-box case below holds sigma and x0 to 1e-4.
    h = 1e-4
    fd_e = (traj_loss_fp64(x, v, L, 1, 1 + h, 1.0, DT4, M4)[0]
            - traj_loss_fp64(x, v, L, 1, 1 - h, 1.0, DT4, M4)[0]) / (2 * h)
    fd_s = (traj_loss_fp64(x, v, L, 1, 1.0, 1 + h, DT4, M4)[0]
            - traj_loss_fp64(x, v, L, 1, 1.0, 1 - h, DT4, M4)[0]) / (2 * h)
    print(f"  FD: dL/deps {fd_e:.9g}, dL/dsigma {fd_s:.9g}")
    assert abs(fd_e - g2[0]) <= 1e-4 * abs(fd_e)
    assert abs(fd_s - g2[1]) <= 3e-3 * abs(fd_s)
    for (i, c) in [(3, 0), (1000, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[i, c] += h
        xm[i, c] -= h
        fd = (traj_loss_fp64(xp, v, L, 1, 1.0, 1.0, DT4, M4)[0]
              - traj_loss_fp64(xm, v, L, 1, 1.0, 1.0, DT4, M4)[0]) / (2 * h)
        print(f"  FD dL/dx[{i},{c}] {fd:.6g} vs {dx0[i, c]:.6g}")
        assert abs(fd - dx0[i, c]) <= 1e-3 * max(abs(fd), rms)


def test_trajectory_gradient_open_box_fd():
    """Same workload in an open box (smooth loss): the GPU gradient against
    central differences of the FP64 loss for eps, sigma and x0."""
    L, pos, vel = _liquid_n2048_state()
    print("\n[case] adjoint open box N=2048 M=200 k=20", flush=True)
    with ljamd.LJSystem(N_LIQ, RHO_LIQ, T_LIQ, boundary=ljamd.BC_NONE) as s:
        s.set_state(pos, vel)
        out, g2, dx0, _ = s.trajectory_grad(DT4, M4, K4)
    x, v = xyz(pos), xyz(vel)
    ref_out, ref_g2, ref_dx0 = traj_grad_fp64(x, v, L, 0, 1.0, 1.0, DT4, M4)
    rel_g = np.abs(g2 - ref_g2) / np.abs(ref_g2)
    rel_x = np.linalg.norm(dx0[:, :3] - ref_dx0) / np.linalg.norm(ref_dx0)
    print(f"  vs FP64 adjoint: loss rel {abs(out[0] - ref_out[0]) / abs(ref_out[0]):.2e}, "
          f"eps/sigma rel {rel_g}, x0 {rel_x:.2e}")
    h = 1e-4
    fd_e = (traj_loss_fp64(x, v, L, 0, 1 + h, 1.0, DT4, M4)[0]
            - traj_loss_fp64(x, v, L, 0, 1 - h, 1.0, DT4, M4)[0]) / (2 * h)
    fd_s = (traj_loss_fp64(x, v, L, 0, 1.0, 1 + h, DT4, M4)[0]
            - traj_loss_fp64(x, v, L, 0, 1.0, 1 - h, DT4, M4)[0]) / (2 * h)
    print(f"  FD: dL/deps {fd_e:.9g} vs {g2[0]:.9g}, dL/dsigma {fd_s:.9g} vs {g2[1]:.9g}")
    assert abs(fd_e - g2[0]) <= 1e-4 * abs(fd_e)
    assert abs(fd_s - g2[1]) <= 1e-4 * abs(fd_s)
    rms = np.sqrt(np.mean(ref_dx0 ** 2))
    for (i, c) in [(3, 0), (1000, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[i, c] += h
        xm[i, c] -= h
        fd = (traj_loss_fp64(xp, v, L, 0, 1.0, 1.0, DT4, M4)[0]
              - traj_loss_fp64(xm, v, L, 0, 1.0, 1.0, DT4, M4)[0]) / (2 * h)
        print(f"  FD dL/dx[{i},{c}] {fd:.6g} vs {dx0[i, c]:.6g}")
        assert abs(fd - dx0[i, c]) <= 1e-4 * max(abs(fd), rms)


def test_trajectory_gradient_open_box_small():
    N, rho = 300, 0.3
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=21)
    with ljamd.LJSystem(N, rho, 1.0, boundary=ljamd.BC_NONE, eps=1.1, sigma=1.05) as s:
        s.set_state(pos, vel)
        out, g2, dx0, _ = s.trajectory_grad(0.004, 50, 10, p_target=0.2, w_p=3.0)
    ref_out, ref_g2, ref_dx0 = traj_grad_fp64(xyz(pos), xyz(vel), L, 0, 1.1, 1.05, 0.004, 50,
                                              pt=0.2, wp=3.0)
    assert np.all(np.abs(g2 - ref_g2) <= 1e-4 * np.abs(ref_g2))
    assert np.linalg.norm(dx0[:, :3] - ref_dx0) <= 1e-4 * np.linalg.norm(ref_dx0)


def test_trajectory_gradient_pair_once_forces():
    """The forward and recompute force passes of the gradient on the pair-once
    kernel (split items, forced on at N = 300: two groups of 256, 4 slots per
    particle) against the FP64 gradient, as on the sliced kernel above."""
    import os

    N, rho = 300, 0.3
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=21)
    old = os.environ.get("LJ_PAIR_ONCE")
    os.environ["LJ_PAIR_ONCE"] = "1"
    try:
        with ljamd.LJSystem(N, rho, 1.0, boundary=ljamd.BC_NONE, eps=1.1, sigma=1.05) as s:
            assert s.force_plan() == (True, 4)
            s.set_state(pos, vel)
            out, g2, dx0, _ = s.trajectory_grad(0.004, 50, 10, p_target=0.2, w_p=3.0)
    finally:
        if old is None:
            os.environ.pop("LJ_PAIR_ONCE", None)
        else:
            os.environ["LJ_PAIR_ONCE"] = old
    ref_out, ref_g2, ref_dx0 = traj_grad_fp64(xyz(pos), xyz(vel), L, 0, 1.1, 1.05, 0.004, 50,
                                              pt=0.2, wp=3.0)
    assert np.all(np.abs(g2 - ref_g2) <= 1e-4 * np.abs(ref_g2))
    assert np.linalg.norm(dx0[:, :3] - ref_dx0) <= 1e-4 * np.linalg.norm(ref_dx0)


def test_trajectory_errors():
    with ljamd.LJSystem(200, 0.5, 1.0, canonical=ljamd.TVN) as s:
        with pytest.raises(ljamd.LJError):
            s.trajectory_grad(0.005, 10)
    with ljamd.LJSystem(200, 0.5, 1.0, boundary=ljamd.BC_WALL) as s:
        with pytest.raises(ljamd.LJError):
            s.trajectory_loss(0.005, 10)


def test_torch_autograd_binding():
    torch = pytest.importorskip("torch")
    from ljamd.autograd import trajectory_loss

    N, rho, M = 512, 0.7, 40
    L, pos, vel = lattice_state(N, rho, 1.0, 9)
    dev = torch.device("cuda:0")
    x0 = torch.tensor(pos.reshape(N, 4), device=dev, requires_grad=True)
    v0 = torch.tensor(vel.reshape(N, 4), device=dev)
    eps = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    with ljamd.LJSystem(N, rho, 1.0) as s:
        loss = trajectory_loss(x0, eps, sig, s, v0, dt=0.005, nsteps=M, checkpoint_every=10)
        loss.backward()
        s.set_state(pos, vel)
        out, g2, dx0, _ = s.trajectory_grad(0.005, M, 10)
    assert abs(loss.item() - out[0]) <= 1e-12 * abs(out[0])
    assert abs(eps.grad.item() - g2[0]) <= 1e-12 * abs(g2[0])
    assert abs(sig.grad.item() - g2[1]) <= 1e-12 * abs(g2[1])
    assert np.array_equal(x0.grad.cpu().numpy()[:, :3], dx0[:, :3])
