"""Shared test utilities: seeded states and error metrics."""
import numpy as np

from oracle_lib import ORC, lattice_state


def perturbed_state(N, rho, T0, seed, amp=0.05):
    """Config lattice + seeded Gaussian velocities, positions displaced by
    N(0, amp*dL) per component (seeded) so that forces are generic."""
    L, pos, vel = lattice_state(N, rho, T0, seed)
    n = ORC.orc_lattice_n(N)
    rng = np.random.default_rng(seed + 17)
    p = pos.reshape(N, 4).copy()
    p[:, :3] += (rng.normal(0.0, amp * L / n, (N, 3))).astype(np.float32)
    return L, p.reshape(-1).copy(), vel


def xyz(a4):
    return np.asarray(a4, np.float64).reshape(-1, 4)[:, :3]


def force_errors(F, Fref):
    """Global and per-particle relative errors of forces [N,3].
    per-particle: |dF_i| / max(|F_i|, F_rms) (floor = rms force)."""
    F = np.asarray(F, np.float64)
    Fref = np.asarray(Fref, np.float64)
    d = np.linalg.norm(F - Fref, axis=1)
    n = np.linalg.norm(Fref, axis=1)
    floor = float(np.sqrt(np.mean(n ** 2)))
    return {
        "global_rel": float(np.linalg.norm(F - Fref) / max(np.linalg.norm(Fref), 1e-300)),
        "max_rel_floor": float(np.max(d / np.maximum(n, floor))),
        "floor": floor,
    }


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def shear_fp64(pos, vel, L, periodic, rows=256):
    """Pshear of the state in FP64 (eps = sigma = 1): (sum over i < j of
    -r_x F_y + sum_i -v_x v_y) / L^3, and the same with every term's
    magnitude (the scale of the tolerance: the stress averages to zero)."""
    x = xyz(pos)
    v = xyz(vel)
    N = x.shape[0]
    tot, mag = 0.0, 0.0
    for i0 in range(0, N, rows):
        d = x[i0:i0 + rows, None, :] - x[None, :, :]
        if periodic:
            d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=2)
        keep = np.arange(N)[None, :] > np.arange(i0, min(N, i0 + rows))[:, None]
        r2 = np.where(keep, r2, 1.0)
        ir2 = 1.0 / r2
        ir6 = ir2 ** 3
        fac = np.where(keep, 4.0 * ir2 * (12.0 * ir6 * ir6 - 6.0 * ir6), 0.0)
        t = -d[:, :, 0] * fac * d[:, :, 1]
        tot += float(np.sum(t))
        mag += float(np.sum(np.abs(t)))
    kin = -v[:, 0] * v[:, 1]
    V = L ** 3
    return (tot + float(np.sum(kin))) / V, (mag + float(np.sum(np.abs(kin)))) / V
