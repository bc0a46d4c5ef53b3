"""Device-resident time step (MDSystem::Integrate, MDSystem.cpp:438-583) against
the reference CPU path restated in the oracle (bit-exact to the compiled
reference).  Forces differ from the reference's at the rounding level, so
trajectories are compared step by step with the divergence reported; the
integrator arithmetic itself follows the reference's double/float pattern."""
import numpy as np
import pytest

import ljamd
from helpers import perturbed_state, xyz
from oracle_lib import ORC, OracleSystem, lattice_state

pytestmark = pytest.mark.gpu


def _pair(N, rho, T0, canon, bc, pos, vel, precision=ljamd.PREC_FP32):
    g = ljamd.LJSystem(N, rho, T0, canonical=canon, boundary=bc, precision=precision)
    g.set_state(pos, vel)
    g.CalculateForces()
    g.CalculateParameters()
    o = OracleSystem(N, T0, rho, canon, bc, nthreads=16)
    o.set_state(pos, vel)
    o.forces()
    o.params()
    return g, o


def test_initial_state_matches_oracle():
    for N, rho, T0 in [(400, 0.05, 1.4), (16384, 0.85, 1.0), (2048, 0.85, 1.0)]:
        L, pos, vel = lattice_state(N, rho, T0, seed=2024)
        with ljamd.LJSystem(N, rho, T0, seed=2024) as s:
            assert abs(s.L - L) == 0.0
            assert np.array_equal(s.pos, pos)
            v = s.vel.reshape(N, 4)
            assert np.array_equal(v[:, :3], vel.reshape(N, 4)[:, :3])
            obs = s.observables()
            assert abs(obs["T"] - T0) <= 1e-6 * T0  # rescaled to exactly T0 (float state)
            assert obs["t"] == 0.0 and obs["av_iters"] == 0
        vv = vel.reshape(N, 4)[:, :3].astype(np.float64)
        assert np.abs(vv.sum(axis=0)).max() <= 1e-4 * np.sqrt(N)


@pytest.mark.parametrize("N,rho,T0,canon,bc,nsteps,tolU", [
    # tolerances about 10x the measured divergence (MI355X): 7e-7, 7e-7, 1.4e-6,
    # 4.4e-7, 1.4e-7, 3.5e-7, 8.5e-7 at the sampled steps
    (400, 0.05, 1.4, ljamd.EVN, ljamd.BC_PERIODIC, 400, 1e-5),   # the N = 400 workload, a slice
    (2048, 0.85, 1.0, ljamd.TVN, ljamd.BC_PERIODIC, 60, 1e-5),   # the liquid at N = 2048 (TVN)
    (2048, 0.85, 1.0, ljamd.TVN, ljamd.BC_WALL, 60, 2e-5),       # the liquid at N = 2048 (hard wall)
    (1000, 0.3, 1.0, ljamd.EVN, ljamd.BC_WALL, 100, 5e-6),
    (600, 0.2, 1.0, ljamd.EVN, ljamd.BC_NONE, 60, 2e-6),
    # the pair-once pass with a partial last group of 256 (16 and 201 particles)
    (10000, 0.85, 1.0, ljamd.TVN, ljamd.BC_PERIODIC, 10, 2e-6),
    (13001, 0.85, 1.0, ljamd.EVN, ljamd.BC_WALL, 10, 2e-6),
])
def test_trajectory_vs_reference(N, rho, T0, canon, bc, nsteps, tolU):
    L, pos, vel = lattice_state(N, rho, T0, seed=77)
    print(f"\n[case] traj N={N} canon={canon} bc={bc}", flush=True)
    g, o = _pair(N, rho, T0, canon, bc, pos, vel)
    assert g.force_plan()[0] == (N >= 3072)
    try:
        hg = g.Integrate(0.005, nsteps)
        ho = o.integrate(0.005, nsteps)
        dU = np.abs(hg[:, 0] - ho[:, 0]) / N
        dT = np.abs(hg[:, 1] - ho[:, 1])
        dx = np.max(np.abs(xyz(g.pos) - xyz(o.pos)))
        k = [0, nsteps // 4, nsteps // 2, nsteps - 1]
        print(f"  |dU|/N at steps {k}: {dU[k]}  |dT|: {dT[k]}  final max|dx| {dx:.2e}; "
              f"max |dU|/N {dU.max():.2e}, max |dT| {dT.max():.2e}")
        assert dU.max() <= tolU
        assert dT.max() <= 10 * tolU
        obs = g.observables()
        assert obs["step"] == nsteps and abs(obs["t"] - 0.005 * nsteps) < 1e-12
    finally:
        g.close()


def test_evn_energy_conservation():
    N, rho, T0 = 2048, 0.85, 1.0
    L, pos, vel = lattice_state(N, rho, T0, seed=5)
    with ljamd.LJSystem(N, rho, T0) as s:
        s.set_state(pos, vel)
        s.CalculateForces()
        s.CalculateParameters()
        h = s.Integrate(0.002, 1000)
    U = h[:, 0] / N
    drift = abs(U[-1] - U[200])
    fluct = np.std(U[200:])
    print(f"\n  EVN dt=0.002: u drift {drift:.2e} fluct {fluct:.2e}")
    assert drift < 5e-3 and fluct < 5e-3


def test_time_reversibility():
    N, rho, T0 = 1000, 0.5, 1.0
    L, pos, vel = perturbed_state(N, rho, T0, seed=6, amp=0.02)
    with ljamd.LJSystem(N, rho, T0, boundary=ljamd.BC_NONE) as s:
        s.set_state(pos, vel)
        s.CalculateForces()
        s.Integrate(0.002, 40)
        v = s.vel
        v.reshape(N, 4)[:, :3] *= -1
        s.set_array(ljamd.ARRAY_VEL, v)
        s.Integrate(0.002, 40)
        back = s.pos
    err = np.max(np.abs(xyz(back) - xyz(pos)))
    assert err < 1e-3, err


@pytest.mark.parametrize("N,canon,bc", [
    (3000, ljamd.TVN, ljamd.BC_PERIODIC), (3000, ljamd.EVN, ljamd.BC_WALL),
    (10000, ljamd.TVN, ljamd.BC_PERIODIC), (10000, ljamd.EVN, ljamd.BC_PERIODIC),
    (10000, ljamd.TVN, ljamd.BC_WALL), (10000, ljamd.EVN, ljamd.BC_NONE)])
def test_graph_replay_deterministic(N, canon, bc):
    # 41 steps run as 32 + 8 + 1 (graphs of 32 and 8 steps, where each step's
    # kick kernel does the next drift and, with the pair-once pass, the next
    # step's first integrator kernel finalises the step; then single steps), as
    # (8 + 1) + 32 and as 7 x 1 + (8 + 5 x 1) + (8 + 8 + 5 x 1): the same
    # arithmetic, the same bits, history rows included; N = 10000 takes the
    # pair-once force pass
    rho, T0 = 0.85, 1.0
    L, pos, vel = lattice_state(N, rho, T0, seed=9)
    print(f"\n[case] graph replay N={N} canon={canon} bc={bc}", flush=True)
    outs = []
    for chunks in ([41], [9, 32], [7, 13, 21]):
        with ljamd.LJSystem(N, rho, T0, canonical=canon, boundary=bc) as s:
            assert s.force_plan()[0] == (N >= 3072)
            s.set_state(pos, vel)
            s.CalculateForces()
            hs = [s.Integrate(0.005, c) for c in chunks]
            obs = s.observables()
            outs.append((s.pos, s.vel, np.concatenate(hs), s.force,
                         [obs[k] for k in ("U", "T", "K", "V", "P", "t", "step", "av_iters")]))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
    assert outs[0][4][6] == 41


def test_renormalize_and_modes():
    N, rho, T0 = 500, 0.3, 1.3
    L, pos, vel = lattice_state(N, rho, T0, seed=4)
    with ljamd.LJSystem(N, rho, T0) as s:
        s.set_state(pos, 2.0 * vel)
        s.CalculateForces()
        s.CalculateParameters()
        s.RenormalizeVelocities(True)
        obs = s.CalculateParameters()
        assert abs(obs["T"] - T0) <= 1e-6
        s.RenormalizeVelocitiesToEnergy(1.0)
        obs = s.CalculateParameters()
        assert abs(obs["U"] / N - 1.0) <= 1e-5
        s.set_mode(ljamd.TVN, ljamd.BC_WALL)
        s.Integrate(0.005, 5)
        with pytest.raises(ljamd.LJError):
            s.set_mode(3, 0)


@pytest.mark.parametrize("canon,prec", [(ljamd.EVN, ljamd.PREC_FP32), (ljamd.TVN, ljamd.PREC_FP32),
                                        (ljamd.TVN, ljamd.PREC_COMPENSATED)])
def test_step_totals_of_pair_once_pass(canon, prec):
    """At N >= 3072 the step's force pass sums E and W per workgroup (no
    per-particle energies), in the FP32 and the compensated mode.  Its V and
    W must equal a full force evaluation at the same positions; Force.w reads
    0 after the step, as documented."""
    N, rho, T0, dt = 16384, 0.85, 1.0, 0.005
    L, pos, vel = lattice_state(N, rho, T0, seed=5)
    with ljamd.LJSystem(N, rho, T0, canonical=canon, precision=prec) as s:
        assert s.force_plan()[0]
        s.set_state(pos, vel)
        s.CalculateForces()
        hist = s.Integrate(dt, 2)
        o = s.observables()
        f_step = s.force.reshape(N, 4).copy()
        o2 = s.CalculateForces()  # full evaluation (per-particle energies) at the same positions
        f_full = s.force.reshape(N, 4)
    print(f"\n  V {hist[-1, 3]:.10g} vs {o2['V']:.10g}; W {o['W']:.10g} vs {o2['W']:.10g}")
    assert abs(hist[-1, 3] - o2["V"]) <= 1e-6 * abs(o2["V"])
    assert abs(o["W"] - o2["W"]) <= 1e-6 * abs(o2["W"])
    assert np.all(f_step[:, 3] == 0.0) and np.any(f_full[:, 3] != 0.0)
    d = np.linalg.norm(f_step[:, :3] - f_full[:, :3]) / np.linalg.norm(f_full[:, :3])
    assert d <= 1e-6


@pytest.mark.parametrize("N,canon,bc,nsteps,tolU", [
    (10000, ljamd.TVN, ljamd.BC_PERIODIC, 10, 2e-6),
    (13001, ljamd.EVN, ljamd.BC_WALL, 10, 2e-6),
])
def test_compensated_trajectory_vs_reference(N, canon, bc, nsteps, tolU):
    """The compensated mode's time step (pair-once pass with compensated sums,
    partial last groups) against the reference CPU path, with the
    tolerances of the FP32 mode's trajectory test at these sizes."""
    rho, T0 = 0.85, 1.0
    L, pos, vel = lattice_state(N, rho, T0, seed=77)
    print(f"\n[case] compensated traj N={N} canon={canon} bc={bc}", flush=True)
    g, o = _pair(N, rho, T0, canon, bc, pos, vel, precision=ljamd.PREC_COMPENSATED)
    assert g.force_plan()[0]
    try:
        hg = g.Integrate(0.005, nsteps)
        ho = o.integrate(0.005, nsteps)
        dU = np.abs(hg[:, 0] - ho[:, 0]) / N
        dT = np.abs(hg[:, 1] - ho[:, 1])
        print(f"  max |dU|/N {dU.max():.2e}, max |dT| {dT.max():.2e}")
        assert dU.max() <= tolU and dT.max() <= 10 * tolU
    finally:
        g.close()

