"""Fluctuation observables on the device (lj_fl_*, lj_batch_fl_*) and the
sampling inside the time step (lj_set_step_sampling), against the oracle's
restatement of the reference's counters (oracle/lj_oracle_flucs.c, pinned to
run-fluctuations-aux.h and MDSystem::Fluctuations compiled from source):

  * counts of every ladder type, bit-exact, on lattice, perturbed and
    out-of-box (hard-wall) states;
  * statistics accumulated over a run that samples after every step, bit-exact
    against the oracle's TimeAverage sums over the same states (the run's
    states are the single steps' states, which are read back one by one);
  * the same per box of a batch;
  * the RDF of every step (LJ_SAMPLE_RDF_LAST / SUM) against the oracle's
    binning of the positions each step's force pass saw."""
import ctypes as C

import numpy as np
import pytest

import ljamd
from helpers import perturbed_state
from oracle_lib import (ORC, RDF_REF_CPU, RDF_REF_GPU, OrcTavg, fl_nsub_batch, fl_vz_batch,
                        lattice_state, rdf)

pytestmark = pytest.mark.gpu

STEP = 0.05


def _ladders(T0):
    vmax = 3.0 * np.sqrt(T0)  # MomentumFlucsAverage's vcut_max (run-fluctuations-aux.h:432,462)
    return [
        (ljamd.FL_SLAB_X, 0, STEP, 0.0), (ljamd.FL_SLAB_Y, 0, STEP, 0.0),
        (ljamd.FL_SLAB_Z, 0, STEP, 0.0), (ljamd.FL_CUBE, 0, STEP, 0.0),
        (ljamd.FL_VEL_Z, 0, STEP, vmax), (ljamd.FL_VEL_X, 0, 0.1, vmax),
        (ljamd.FL_CUBE, 1, 0.5, 0.0), (ljamd.FL_SLAB_Z, 1, 0.5, 0.0),
        (ljamd.FL_VEL_Y, 1, 0.0, 1.0), (ljamd.FL_CUBE_CLOSED, 1, 0.3, 0.0),
    ]


def _oracle_counts(pos, vel, L, T0):
    """The counts of _ladders(T0) by the oracle, concatenated in ladder order."""
    N = pos.size // 4
    vmax = 3.0 * np.sqrt(T0)
    out = [fl_nsub_batch(pos, L, STEP, t) for t in (0, 1, 2, 3)]
    out.append(fl_vz_batch(vel, vmax, STEP, 2))
    out.append(fl_vz_batch(vel, vmax, 0.1, 0))
    out.append([ORC.orc_fl_nsub(N, pos, L, 0.5, 3), ORC.orc_fl_nsub(N, pos, L, 0.5, 2),
                ORC.orc_fl_vz(N, vel, 1.0, 1), ORC.orc_fl_closed_cube(N, pos, L, 0.3)])
    return np.concatenate([np.asarray(o, np.int64) for o in out])


def _states():
    out = []
    for N, rho, T0, seed in [(400, 0.05, 1.4, 1), (2048, 0.85, 1.0, 2), (5000, 0.3, 1.2, 3)]:
        out.append((N, rho, T0) + lattice_state(N, rho, T0, seed))
        out.append((N, rho, T0) + perturbed_state(N, rho, T0, seed, amp=0.3))
    N, rho, T0, L, pos, vel = out[1]
    p = pos.reshape(N, 4).copy()
    p[:9, :3] = np.float32(-1e-3)  # a hard wall leaves particles just outside
    p[9:20, :3] = np.float32(L + 2e-3)
    out.append((N, rho, T0, L, p.reshape(-1), vel))
    return out


STATES = _states()


@pytest.mark.parametrize("k", range(len(STATES)))
def test_counts_bitexact_vs_oracle(k):
    N, rho, T0, L, pos, vel = STATES[k]
    print(f"\n[case] fluctuation counts N={N} state {k}", flush=True)
    with ljamd.LJSystem(N, rho, T0) as s:
        s.set_state(pos, vel)
        s.fl_configure(_ladders(T0))
        exp = _oracle_counts(pos, vel, L, T0)
        assert s.fl_num_entries() == exp.size
        s.fl_sample()
        got = s.fl_counts()
        assert np.array_equal(got, exp)
        # the values of the entries: the accumulated ladder steps and the fractions
        vals = s.fl_values()
        assert vals[0] == STEP and vals[1] == STEP + STEP
        assert vals[-1] == 0.3 and vals[-2] == 1.0
        st = s.fl_stats()
        assert np.array_equal(st[:, 1], exp.astype(np.float64))
        assert np.all(st[:, 7] == 1.0) and np.all(st[:, 5] == 0.0)


def _tavg_sums(series):
    """Oracle TimeAverage sums of each entry's count series (samples x entries):
    the statistics layout of lj_fl_stats."""
    T, E = series.shape
    out = np.zeros((E, ljamd.FL_NSTAT))
    for e in range(E):
        a = OrcTavg()
        ORC.orc_tavg_init(C.byref(a))
        for t in range(T):
            ORC.orc_tavg_add(C.byref(a), float(series[t, e]))
        out[e, :5] = list(a.sums)
        out[e, 5] = a.corr
        out[e, 6] = a.last
        out[e, 7] = a.iters
    return out


@pytest.mark.parametrize("N,rho,T0,canon,bc", [(400, 0.05, 1.4, ljamd.TVN, ljamd.BC_PERIODIC),
                                               (4096, 0.6, 1.0, ljamd.EVN, ljamd.BC_WALL)])
def test_statistics_of_sampled_run_bitexact(N, rho, T0, canon, bc):
    """Sampling after every step inside the step graphs (40 steps in one call)
    against single steps read back one by one and counted by the oracle; the
    two runs end in the same bits (sampling steps are single steps)."""
    steps, dt = 40, 0.004
    L, pos, vel = lattice_state(N, rho, T0, seed=5)
    print(f"\n[case] sampled run N={N} canon={canon} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, T0, canonical=canon, boundary=bc) as a:
        a.set_state(pos, vel)
        a.CalculateForces()
        a.CalculateParameters()
        a.fl_configure(_ladders(T0))
        a.set_step_sampling(fluctuations=True)
        ha = a.Integrate(dt, steps)
        st = a.fl_stats()
        last = a.fl_counts()
        pa, va = a.pos, a.vel
    series = []
    with ljamd.LJSystem(N, rho, T0, canonical=canon, boundary=bc) as b:
        b.set_state(pos, vel)
        b.CalculateForces()
        b.CalculateParameters()
        hb = []
        for _ in range(steps):
            hb.append(b.Integrate(dt, 1)[0])
            series.append(_oracle_counts(b.pos, b.vel, L, T0))
        pb, vb = b.pos, b.vel
    series = np.array(series)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert np.array_equal(ha, np.array(hb))
    assert np.array_equal(last, series[-1])
    exp = _tavg_sums(series)
    assert np.array_equal(st, exp)


def test_batch_counts_and_sampled_statistics():
    """Every box of a batch against the oracle with its own L; sampling after
    every step of lj_batch_integrate against lj_batch_fl_sample after single
    steps."""
    N, rho, T0 = ljamd.ensemble_boxes(6, seed=3, n_range=(300, 2500))
    T = float(np.max(T0))
    print("\n[case] batch fluctuation counts", flush=True)
    with ljamd.LJBatch(N, rho, T0, seed=9) as b:
        b.fl_configure(_ladders(T))
        b.fl_sample()
        got = b.fl_counts()
        for k in range(len(N)):
            L = ORC.orc_box_length(int(N[k]), float(rho[k]))
            pos = b.get_array(ljamd.ARRAY_POS, box=k)
            vel = b.get_array(ljamd.ARRAY_VEL, box=k)
            assert np.array_equal(got[k], _oracle_counts(pos, vel, L, T)), k
        b.fl_reset()
        b.set_step_sampling(True)
        b.integrate(0.005, 12)
        st_a = b.fl_stats()
        pa = b.get_array(ljamd.ARRAY_POS)
    with ljamd.LJBatch(N, rho, T0, seed=9) as c:
        c.fl_configure(_ladders(T))
        for _ in range(12):
            c.integrate(0.005, 1)
            c.fl_sample()
        st_b = c.fl_stats()
        pb = c.get_array(ljamd.ARRAY_POS)
    assert np.array_equal(pa, pb)
    assert np.array_equal(st_a, st_b)
    assert np.all(st_a[:, :, 7] == 12.0)


@pytest.mark.parametrize("bc,periodic", [(ljamd.BC_WALL, False), (ljamd.BC_NONE, False)])
def test_step_rdf_bitexact(bc, periodic):
    """With hard walls or an open box the boundary conditions do not move
    positions, so the positions a step's force pass sees are those it leaves:
    the step RDF equals the oracle's binning of them, bit for bit (both
    reference rules); the summed mode adds the steps up."""
    N, rho, T0 = 2048, 0.6, 1.0
    L, pos, vel = perturbed_state(N, rho, T0, seed=13)
    print(f"\n[case] step RDF bc={bc}", flush=True)
    for rule in (RDF_REF_GPU, RDF_REF_CPU):
        with ljamd.LJSystem(N, rho, T0, boundary=bc) as s:
            s.set_state(pos, vel)
            s.CalculateForces()
            s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_LAST, rdf_rule=rule, rdf_nbins=256)
            s.Integrate(0.005, 7)
            got, n = s.step_rdf_counts()
            exp = rdf(s.pos, L, periodic, rule, 256, dr2=ORC.orc_rdf_dr2(N))
            assert n == 1 and np.array_equal(got.astype(np.int64), exp)
            s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_SUM, rdf_rule=rule, rdf_nbins=256)
            total = np.zeros(256, np.int64)
            for _ in range(5):
                s.Integrate(0.005, 1)
                total += rdf(s.pos, L, periodic, rule, 256, dr2=ORC.orc_rdf_dr2(N))
            got, n = s.step_rdf_counts(clear=True)
            assert n == 5 and np.array_equal(got.astype(np.int64), total)
            got, n = s.step_rdf_counts()
            assert n == 0 and not got.any()


def test_step_rdf_periodic_and_trajectory_unchanged():
    """Periodic box: the step histogram bins the drifted positions before the
    wrap, so it can differ from a binning of the wrapped final positions only
    by pairs whose images round differently; and sampling leaves the trajectory
    bit-identical to a run without it."""
    N, rho, T0 = 3000, 0.85, 1.0
    L, pos, vel = lattice_state(N, rho, T0, seed=3)
    runs = []
    for sample in (False, True):
        with ljamd.LJSystem(N, rho, T0, canonical=ljamd.TVN) as s:
            s.set_state(pos, vel)
            s.CalculateForces()
            if sample:
                s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_LAST, rdf_rule=RDF_REF_CPU)
            h = s.Integrate(0.005, 20)
            runs.append((h, s.pos, s.vel))
            if sample:
                got, _ = s.step_rdf_counts()
                exp = rdf(s.pos, L, True, RDF_REF_CPU, 256, dr2=ORC.orc_rdf_dr2(N))
                d = np.abs(got.astype(np.int64) - exp).sum()
                print(f"\n  periodic step RDF vs wrapped final positions: {d} of {exp.sum()}")
                assert d <= 1e-4 * exp.sum()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_fl_errors():
    with ljamd.LJSystem(300, 0.3) as s:
        with pytest.raises(ljamd.LJError):
            s.fl_sample()
        with pytest.raises(ljamd.LJError):
            s.set_step_sampling(fluctuations=True)
        with pytest.raises(ljamd.LJError):
            s.fl_configure([(9, 0, 0.05, 0.0)])
        with pytest.raises(ljamd.LJError):
            s.fl_configure([(ljamd.FL_VEL_X, 0, 0.05, 0.0)])  # vmax missing
        with pytest.raises(ljamd.LJError):
            s.fl_configure([(ljamd.FL_CUBE, 0, 1e-5, 0.0)])  # over 4096 entries
        s.fl_configure([(ljamd.FL_CUBE, 0, 0.05, 0.0)])
        s.fl_configure([])
        with pytest.raises(ljamd.LJError):
            s.fl_counts()
