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
