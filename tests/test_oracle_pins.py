"""Pins of the oracle (CPU suite).

1. Bit-exact against the reference's own CPU path compiled from its sources
   (oracle/_ref/libref_cpu.so): forces, V, P, Pshear, RDF bins and full
   EVN / TVN trajectories for all three boundary conditions.
2. What the mathematics fixes: LJ dimer closed forms, F = -grad V by central
   differences, Newton's third law, the virial identity, the force-free
   periodic simple-cubic lattice and its lattice-sum energy, RDF totals and
   rule relations, time reversibility, the TVN half-step invariant, O(dt^2)
   energy error of velocity Verlet, ideal-gas pressure, and the initial state.
"""
import math

import numpy as np
import pytest

from helpers import perturbed_state
from oracle_lib import (ORC, RDF_RBINS, RDF_REF_CPU, RDF_REF_GPU, OracleSystem, RefSystem,
                        forces_fp64, forces_fp64d, forces_ref, lattice_state, rdf, rdf_edges2,
                        ref_lib, velocity_histogram)

REF = ref_lib("cpu")
needs_ref = pytest.mark.skipif(REF is None, reason="reference not built (oracle/_ref)")


# ------------------------------------------------------------------ 1
@needs_ref
@pytest.mark.parametrize("N,rho,T0,bc,canon", [
    (400, 0.05, 1.4, 0, 0), (343, 0.85, 1.0, 0, 1), (300, 0.6, 1.0, 1, 1),
    (200, 0.3, 1.0, 2, 0), (150, 0.9, 2.0, 1, 0)])
def test_oracle_bitexact_vs_reference(N, rho, T0, bc, canon):
    L, pos, vel = perturbed_state(N, rho, T0, seed=99)
    R = RefSystem(REF, N, T0, rho, canon, bc)
    try:
        assert R.L == L and R.dr2 == ORC.orc_rdf_dr2(N)
        R.set_state(pos, vel)
        fr, V, P, Ps, rdfr = R.forces()
        fo, Vo, Po, Pso, rdfo = forces_ref(pos, L, bc, ORC.orc_rdf_dr2(N))
        m = np.arange(4 * N) % 4 < 3
        assert np.array_equal(fr[m].view(np.int32), fo[m].view(np.int32))
        assert (V, P, Ps) == (Vo, Po, Pso)
        assert np.array_equal(rdfr, rdfo)
        fo8, Vo8, Po8, Pso8, rdfo8 = forces_ref(pos, L, bc, ORC.orc_rdf_dr2(N), nthreads=4)
        assert np.array_equal(fo8[m], fo[m]) and np.array_equal(rdfo8, rdfo)
        assert abs(Vo8 - Vo) <= 1e-12 * abs(Vo) + 1e-12
        # trajectory
        O = OracleSystem(N, T0, rho, canon, bc)
        O.set_state(pos, vel)
        O.forces()
        O.params()
        R.set_state(pos, vel)
        R.forces()
        R.params()
        ho = O.integrate(0.005, 40)
        hr = R.integrate(0.005, 40)
        p1, v1, f1 = R.state()
        assert np.array_equal(ho, hr)
        assert np.array_equal(p1[m], O.pos[m]) and np.array_equal(v1[m], O.vel[m])
    finally:
        R.close()


@needs_ref
def test_rdf_normalisation_matches_reference():
    """RDF() of the reference (MDSystem.cpp:633-649) from the oracle's counts."""
    N, rho = 300, 0.4
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=1)
    R = RefSystem(REF, N, 1.0, rho, 0, 0)
    try:
        R.set_state(pos, vel)
        R.forces()
        x = np.zeros(256)
        y = np.zeros(256)
        REF.refh_rdf_normalised(R.h, x, y)
        counts = rdf(pos, L, 1, RDF_REF_CPU, 256, dr2=R.dr2)
        dr2 = float(np.float32(R.dr2))
        r = np.sqrt((np.arange(256) + 0.5) * dr2)
        n0 = N / L / L / L  # MDSystem.cpp:637 evaluation order
        # int / float is a float division in the reference (MDSystem.cpp:643)
        q = (counts.astype(np.float32) / np.float32(R.dr2)).astype(np.float64)
        g = q / 2.0 / math.pi / r / n0 / N
        assert np.array_equal(x, r) and np.array_equal(y, g)
    finally:
        R.close()


def test_velocity_distribution_matches_reference():
    """initvelo / updatevelo of the reference (MDSystem.cpp:651-694) from the
    oracle's speed counts, bit for bit."""
    N, rho = 500, 0.3
    L, pos, vel = perturbed_state(N, rho, 1.3, seed=6)
    R = RefSystem(REF, N, 1.3, rho, 0, 0)
    try:
        R.set_state(pos, vel)
        x, y = R.velo(False, vmax=4.0, step=0.07)
        nb = int(4.0 / 0.07) + 1
        c = velocity_histogram(vel, 0.07, nb)
        assert x.size == nb and c.sum() > 0.9 * N
        assert np.array_equal(x, 0.07 * (0.5 + np.arange(nb)))
        assert np.array_equal(y, c / 0.07 / N)
        vel2 = (vel * np.float32(1.1)).astype(np.float32)
        R.set_state(pos, vel2)
        x2, y2 = R.velo(True)
        shag = x[1] - x[0]
        c2 = velocity_histogram(vel2, shag, nb)
        assert np.array_equal(x2, x) and np.array_equal(y2, (y * 1 + c2 / shag / N) / 2)
    finally:
        R.close()


# ------------------------------------------------------------------ 2
def _dimer(r, L=50.0):
    pos = np.zeros(8, np.float32)
    pos[0:3] = [10.0, 10.0, 10.0]
    pos[4:7] = [10.0 + r, 10.0, 10.0]
    return pos, L


def test_dimer_closed_forms():
    rmin = 2.0 ** (1.0 / 6.0)
    pos, L = _dimer(rmin)
    F, ew, E, W = forces_fp64(pos, L, 1)
    r32 = float(np.float32(10.0 + rmin)) - 10.0
    assert abs(E - 4 * (r32 ** -12 - r32 ** -6)) < 1e-12 and abs(E + 1.0) < 1e-6
    assert np.abs(F).max() < 1e-4
    pos, L = _dimer(1.0)
    F, ew, E, W = forces_fp64(pos, L, 1)
    assert abs(E) < 1e-12 and abs(abs(F[0, 0]) - 24.0) < 1e-9
    assert F[0, 0] < 0 < F[1, 0]  # repulsive: particle 0 pushed to -x
    assert abs(W - 24.0) < 1e-9  # r . f = 24 at r = 1
    # reference-faithful path: same values in its convention
    fr, V, P, Ps, _ = forces_ref(pos, L, 0, 0.1)
    assert abs(fr[0] + 24.0) < 1e-5 and abs(V) < 1e-6 and abs(P - 24.0 / 3) < 1e-5


def test_dimer_scaling():
    """V(r) = 4 (r^-12 - r^-6) at several r (FP64 oracle, exact float positions)."""
    for r in (0.95, 1.3, 2.0, 3.7):
        pos, L = _dimer(r)
        rr = float(np.float32(10.0 + r)) - 10.0
        _, _, E, W = forces_fp64(pos, L, 0)
        assert abs(E - 4 * (rr ** -12 - rr ** -6)) <= 1e-12 * max(1, abs(E))
        assert abs(W - 24 * (2 * rr ** -12 - rr ** -6)) <= 1e-11 * max(1, abs(W))


@pytest.mark.parametrize("periodic", [0, 1])
def test_force_is_minus_gradient(periodic):
    N, rho = 60, 0.4
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=3, amp=0.1)
    x = pos.reshape(N, 4)[:, :3].astype(np.float64).copy()
    F, _, E, _ = forces_fp64d(x, L, periodic)
    h = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(12):
        i, c = rng.integers(N), rng.integers(3)
        xp, xm = x.copy(), x.copy()
        xp[i, c] += h
        xm[i, c] -= h
        Ep = forces_fp64d(xp, L, periodic)[2]
        Em = forces_fp64d(xm, L, periodic)[2]
        fd = -(Ep - Em) / (2 * h)
        assert abs(fd - F[i, c]) <= 1e-5 * max(1.0, abs(F[i, c])), (i, c, fd, F[i, c])


@pytest.mark.parametrize("periodic", [0, 1])
def test_newton_and_virial_identity(periodic):
    N, rho = 40, 0.5
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=4, amp=0.1)
    F, ew, E, W = forces_fp64(pos, L, periodic)
    assert np.abs(F.sum(axis=0)).max() <= 1e-9 * np.abs(F).sum()
    # virial identity by an explicit pair loop (i<j) with minimum image
    x = pos.reshape(N, 4)[:, :3].astype(np.float64)
    Wd, Ed = 0.0, 0.0
    for i in range(N):
        for j in range(i + 1, N):
            d = x[i] - x[j]
            if periodic:
                d -= L * np.rint(d / L)
            r2 = d @ d
            Wd += 24 * (2 * r2 ** -6 - r2 ** -3)
            Ed += 4 * (r2 ** -6 - r2 ** -3)
    assert abs(Wd - W) <= 1e-12 * abs(W) and abs(Ed - E) <= 1e-12 * abs(E)
    assert abs(ew[:, 1].sum() / 2 - W) <= 1e-12 * abs(W)


def test_periodic_lattice_force_free_and_lattice_sum():
    """n^3 = N simple-cubic lattice in a periodic box: F_i = 0 by symmetry; the
    energy equals an independent lattice sum over integer offsets."""
    n, a = 7, 1.2  # odd n: no image sits exactly at L/2
    N = n ** 3
    L = n * a
    rho = N / L ** 3
    Lc, pos, _ = lattice_state(N, rho, 1.0, 1)
    assert abs(Lc - L) < 1e-12
    k = np.arange(N)
    x = np.stack([(k % n + 0.5) * a, ((k // n) % n + 0.5) * a, (k // (n * n) + 0.5) * a], 1)
    F, _, E, _ = forces_fp64d(x, L, 1)
    assert np.abs(F).max() < 1e-9
    # the float-rounded lattice of the workloads is force-free up to rounding
    Ff, _, Ef, _ = forces_fp64(pos, Lc, 1)
    assert np.abs(Ff).max() < 1e-4 and abs(Ef - E) < 1e-6 * abs(E)
    # lattice sum: each particle sees offsets k in [-(n-1)/2, (n-1)/2]^3 \ 0 (minimum image)
    ks = np.arange(-(n // 2), n // 2 + 1)
    K = np.stack(np.meshgrid(ks, ks, ks, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    K = K[np.any(K != 0, axis=1)]
    r2 = np.sum((K * a) ** 2, axis=1)
    e_site = np.sum(r2 ** -6 - r2 ** -3)
    E_lat = 4 * N * e_site / 2
    assert abs(E - E_lat) <= 1e-9 * abs(E_lat)


def test_rdf_rules_relations():
    N, rho = 500, 0.5
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=8)
    dr2 = ORC.orc_rdf_dr2(N)
    cpu = rdf(pos, L, 1, RDF_REF_CPU, 256, dr2=dr2)
    gpu = rdf(pos, float(np.float32(L)), 1, RDF_REF_GPU, 256, dr2=dr2)
    assert gpu.sum() == N * (N - 1)  # clamped: every ordered pair lands in a bin
    assert np.all(cpu % 2 == 0)
    assert cpu.sum() <= N * (N - 1)
    # below the clamp bin the rules differ only by rounding at bin edges
    assert np.abs(gpu[:255] - cpu[:255]).sum() <= 4
    assert gpu[255] >= cpu[255]
    # uniform-r bins: total = #pairs with r < rmax (FP64 brute force, edge pairs aside)
    rmax = L / 2
    rb = rdf(pos, L, 1, RDF_RBINS, 150, rmax=rmax)
    x = pos.reshape(N, 4)[:, :3].astype(np.float64)
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.rint(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    assert abs(rb.sum() - np.sum(r < rmax)) <= 2
    hist64, _ = np.histogram(r[np.isfinite(r)], bins=150, range=(0, rmax))
    assert np.abs(rb - hist64).sum() <= 4
    # permutation invariance
    perm = np.random.default_rng(0).permutation(N)
    pp = pos.reshape(N, 4)[perm].reshape(-1).copy()
    assert np.array_equal(rdf(pp, L, 1, RDF_RBINS, 150, rmax=rmax), rb)
    e2 = rdf_edges2(150, rmax)
    assert e2[0] == 0 and np.all(np.diff(e2) > 0)


def test_time_reversibility_oracle():
    N, rho, T0 = 300, 0.3, 1.0
    L, pos, vel = perturbed_state(N, rho, T0, seed=5, amp=0.02)
    O = OracleSystem(N, T0, rho, 0, 2, nthreads=4)
    O.set_state(pos, vel)
    O.forces()
    O.integrate(0.002, 30)
    v = O.vel.copy()
    v.reshape(N, 4)[:, :3] *= -1
    O.vel[:] = v
    O.integrate(0.002, 30)
    err = np.abs(O.pos.reshape(N, 4)[:, :3] - pos.reshape(N, 4)[:, :3]).max()
    assert err < 1e-3


def test_tvn_half_step_invariant():
    """TVN (MDSystem.cpp:465-510): v_new = 2 chi v' - v_old with
    sum (chi v')^2 / (3N) = T0, i.e. mean((v_new + v_old)/2)^2 over 3N = T0."""
    N, rho, T0 = 400, 0.6, 1.3
    L, pos, vel = perturbed_state(N, rho, T0, seed=6)
    O = OracleSystem(N, T0, rho, 1, 0)
    O.set_state(pos, 1.5 * vel)
    O.forces()
    for _ in range(3):
        v_old = O.vel.reshape(N, 4)[:, :3].astype(np.float64).copy()
        O.integrate(0.005, 1)
        v_new = O.vel.reshape(N, 4)[:, :3].astype(np.float64)
        half = 0.5 * (v_new + v_old)
        assert abs(np.sum(half ** 2) / (3 * N) - T0) <= 1e-5 * T0


def test_verlet_energy_error_scales_dt2():
    """Dimer oscillation: max |dE| over a period scales as dt^2."""
    errs = []
    for dt in (0.008, 0.004):
        pos = np.zeros(8, np.float32)  # near the origin: float rounding stays below dt^2
        pos[0:3] = [0.5, 0.5, 0.5]
        pos[4:7] = [1.55, 0.5, 0.5]
        vel = np.zeros(8, np.float32)
        O = OracleSystem(2, 1.0, 2 / 40.0 ** 3, 0, 2)
        O.set_state(pos, vel)
        O.forces()
        O.params()
        E0 = O.s.U
        h = O.integrate(dt, int(round(2.0 / dt)))
        errs.append(np.abs(h[:, 0] - E0).max())
    ratio = errs[0] / errs[1]
    assert 3.5 < ratio < 4.5, (errs, ratio)


def test_ideal_gas_pressure():
    """No interaction energy within reach (huge box, 2 particles): P = rho T."""
    pos = np.zeros(8, np.float32)
    pos[0:3] = [1.0, 1.0, 1.0]
    pos[4:7] = [400.0, 400.0, 400.0]
    vel = np.array([1, 0.5, -0.3, 0, -1, -0.5, 0.3, 0], np.float32)
    rho = 2 / 1000.0 ** 3
    O = OracleSystem(2, 1.0, rho, 0, 2)
    O.set_state(pos, vel)
    O.forces()
    O.params()
    assert abs(O.s.P - rho * O.s.T) <= 1e-12 * rho * O.s.T + 1e-25


def test_initial_state_invariants():
    N, rho, T0 = 1000, 0.85, 1.0
    L, pos, vel = lattice_state(N, rho, T0, seed=3)
    n = ORC.orc_lattice_n(N)
    assert n == 10
    p = pos.reshape(N, 4)
    k = np.arange(N)
    dL = L / n
    exp = np.stack([(k % n + 0.5) * dL, ((k // n) % n + 0.5) * dL, (k // (n * n) + 0.5) * dL], 1)
    assert np.array_equal(p[:, :3], exp.astype(np.float32))
    v = vel.reshape(N, 4)[:, :3].astype(np.float64)
    assert np.abs(v.sum(axis=0)).max() <= 1e-4
    assert abs(ORC.orc_kinetic_temperature(N, vel) - T0) <= 1e-6
    # cube roots: the reference's ceil(pow(N, 1/3)) (MDSystem.cpp:150)
    for N_, nn in ((400, 8), (16384, 26), (2048, 13), (131072, 51), (512, 8), (1000, 10)):
        assert ORC.orc_lattice_n(N_) == nn


def test_splitmix_known_values():
    """SplitMix64 from state 0: published first outputs."""
    import ctypes as C
    st = C.c_uint64(0)
    u = [ORC.orc_splitmix_uniform(C.byref(st)) for _ in range(3)]
    z = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert u == [(x >> 11) * 2.0 ** -53 for x in z]


def test_gaussian_moments():
    N = 20000
    vel = np.zeros(4 * N, np.float32)
    ORC.orc_gaussian_velocities(N, 2.0, 12345, vel)
    v = vel.reshape(N, 4)[:, :3].ravel().astype(np.float64)
    assert abs(np.mean(v ** 2) - 2.0) < 1e-5
    k = np.mean(v ** 4) / np.mean(v ** 2) ** 2
    assert abs(k - 3.0) < 0.1  # Gaussian kurtosis
