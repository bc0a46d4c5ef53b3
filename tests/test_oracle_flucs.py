"""Pins of the oracle's fluctuation counters and time averages
(oracle/lj_oracle_flucs.c), CPU only.

Against the reference itself, compiled here from its sources
(oracle/_ref/libref_cpu.so with oracle/ref_flucs_harness.cpp), bit for bit:
GetNSubsystemBatch (all four types), GetNsubVzBatch, GetNSubsystem, GetNsubVz
and MDSystem::Fluctuations on lattice, perturbed and out-of-box states, and
TimeAverage's statistics on integer and real sequences.

Without the reference: cumulative counts against direct counts of the same
inequalities (the cube ladder against GetNSubsystem's direct comparison at
the same alpha, slabs and velocity windows against brute-force counts),
totals, and the statistics against numpy's mean and variance and the
closed-form scaled-variance error of a two-valued sequence."""
import numpy as np
import pytest

from helpers import perturbed_state
from oracle_lib import (ORC, RefSystem, fl_ladder, fl_nsub_batch, fl_vz_batch, lattice_state,
                        ref_lib, tavg_eval)

REF = ref_lib("cpu")
needs_ref = pytest.mark.skipif(REF is None or not hasattr(REF, "refh_fl_nsub_batch"),
                               reason="reference not built (oracle/_ref)")


def _states():
    out = []
    for N, rho, T0, seed in [(400, 0.05, 1.4, 1), (2048, 0.85, 1.0, 2), (1000, 0.3, 1.2, 3)]:
        L, pos, vel = lattice_state(N, rho, T0, seed)
        out.append(("lattice", N, rho, T0, L, pos, vel))
        L, pos, vel = perturbed_state(N, rho, T0, seed, amp=0.3)
        out.append(("perturbed", N, rho, T0, L, pos, vel))
        # hard-wall states leave positions slightly outside [0, L] (MDSystem.cpp:425-435)
        p = pos.reshape(N, 4).copy()
        p[:7, :3] = np.float32(-1e-3)
        p[7:14, :3] = np.float32(L + 2e-3)
        out.append(("outside", N, rho, T0, L, p.reshape(-1), vel))
    return out


STATES = _states()
STEPS = [0.05, 0.1, 0.03, 0.01, 0.125]


def _ref_sys(N, rho, T0, pos, vel):
    r = RefSystem(REF, N, T0, rho, 0, 0)
    r.set_state(pos, vel)
    return r


@needs_ref
@pytest.mark.parametrize("k", range(len(STATES)))
def test_counters_bitexact_vs_reference(k):
    name, N, rho, T0, L, pos, vel = STATES[k]
    r = _ref_sys(N, rho, T0, pos, vel)
    assert abs(r.L - L) == 0.0
    buf = np.zeros(4096, np.int32)
    try:
        for step in STEPS:
            for typ in range(4):
                m = REF.refh_fl_nsub_batch(r.h, step, typ, buf, buf.size)
                assert np.array_equal(fl_nsub_batch(pos, L, step, typ), buf[:m]), (name, step, typ)
            vmax = np.sqrt(T0) * 3.0
            for typ in range(3):
                m = REF.refh_fl_vz_batch(r.h, vmax, step, typ, buf, buf.size)
                assert np.array_equal(fl_vz_batch(vel, vmax, step, typ), buf[:m]), (name, step, typ)
        for frac in (0.5, 0.3, 0.123, 0.95):
            for typ in range(4):
                assert ORC.orc_fl_nsub(N, pos, L, frac, typ) == REF.refh_fl_nsub(r.h, frac, typ)
            for typ in range(3):
                assert ORC.orc_fl_vz(N, vel, frac * 2.0, typ) == REF.refh_fl_vz(r.h, frac * 2.0, typ)
        # MDSystem::Fluctuations: the closed-cube count, through the moment sums
        mom = np.zeros(5)
        prev = 0.0
        for frac in (0.5, 0.2):
            REF.refh_fl_fluctuations(r.h, frac, mom)
            n = ORC.orc_fl_closed_cube(N, pos, L, frac)
            assert mom[0] - prev == n
            prev = mom[0]
    finally:
        r.close()


@needs_ref
@pytest.mark.parametrize("seq", ["counts", "reals", "ar1", "constant_plus_one"])
def test_time_average_bitexact_vs_reference(seq):
    rng = np.random.default_rng(5)
    if seq == "counts":
        v = rng.binomial(400, 0.3, 3000).astype(np.float64)
    elif seq == "reals":
        v = -3.2 + 0.01 * rng.standard_normal(2000)
    elif seq == "ar1":
        v = np.zeros(5000)
        for t in range(1, v.size):
            v[t] = 0.9 * v[t - 1] + rng.standard_normal()
        v += 50.0
    else:
        v = np.full(100, 7.0)
        v[50] = 8.0
    exp = np.zeros(7)
    REF.refh_tavg(np.ascontiguousarray(v), v.size, exp)
    got = tavg_eval(v)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    ok = ~np.isnan(exp)
    assert np.array_equal(got[ok], exp[ok]), (got, exp)


@pytest.mark.parametrize("k", range(len(STATES)))
def test_counters_against_direct_counts(k):
    name, N, rho, T0, L, pos, vel = STATES[k]
    p = pos.reshape(N, 4).astype(np.float64)
    v = vel.reshape(N, 4).astype(np.float64)
    for step in (0.05, 0.1):
        alphas = fl_ladder(step, 1.0 - 1e-9)
        # the cube ladder entry k equals the direct open-cube count at alpha_k
        cube = fl_nsub_batch(pos, L, step, 3)
        for a, c in zip(alphas, cube):
            tL = L * a ** (1.0 / 3.0)
            direct = int(np.sum(np.all(np.abs(p[:, :3] - 0.5 * L) < 0.5 * tL, axis=1)))
            assert c == direct == ORC.orc_fl_nsub(N, pos, L, float(a), 3)
        # slabs: bin index (int)(x / L / step) counted cumulatively
        for typ in range(3):
            got = fl_nsub_batch(pos, L, step, typ)
            t = np.trunc(p[:, typ] / L / step)
            # a negative index (left of the box) is skipped, as the unsigned
            # size comparison of run-fluctuations-aux.h:232 does
            exp = np.array([np.sum((t >= 0) & (t <= i)) for i in range(alphas.size)])
            assert np.array_equal(got, exp)
            assert np.all(np.diff(got) >= 0) and got[-1] <= N
        vmax = 3.0 * np.sqrt(T0)
        va = fl_ladder(step, 1.0 + 1e-9)
        for typ in range(3):
            got = fl_vz_batch(vel, vmax, step, typ)
            t = np.trunc(np.abs(v[:, typ] / vmax) / step)
            exp = np.array([np.sum(t <= i) for i in range(va.size)])
            assert np.array_equal(got, exp)
    # closed cube of fraction 1: every particle inside [0, L]^3
    inside = int(np.sum(np.all((p[:, :3] >= 0) & (p[:, :3] <= L), axis=1)))
    assert ORC.orc_fl_closed_cube(N, pos, L, 1.0) == inside


def test_ladder_is_accumulated():
    a = fl_ladder(0.1, 1.0 - 1e-9)
    acc, exp = 0.1, []
    while acc < 1.0 - 1e-9:
        exp.append(acc)
        acc += 0.1
    assert np.array_equal(a, np.array(exp))
    # 0.1 accumulated nine times is 0.8999999999999999, ten times 0.9999999999999999,
    # which the 1 - 1e-9 limit excludes
    assert a.size == 9 and a[-1] != 0.9
    assert fl_ladder(0.05, 1.0 - 1e-9).size == 19 and fl_ladder(0.05, 1.0 + 1e-9).size == 20


def test_time_average_closed_forms():
    rng = np.random.default_rng(9)
    v = rng.poisson(40.0, 4000).astype(np.float64)
    out = tavg_eval(v)
    assert out[0] == pytest.approx(v.mean(), rel=1e-13)
    assert out[1] == pytest.approx(v.var(), rel=1e-9)
    assert out[4] == pytest.approx(v.var() / v.mean(), rel=1e-9)
    assert out[6] == pytest.approx(np.sqrt(v.var() / v.size), rel=1e-9)
    # uncorrelated samples: s near 1 is not guaranteed, but s > 0 and the
    # time-average error is the plain error times sqrt(s)
    assert out[2] > 0 and out[3] == pytest.approx(out[6] * np.sqrt(out[2]), rel=1e-15)
    # two-valued sequence (fraction q of ones): variance q(1-q), scaled variance
    # (1-q), and its error from the central moments mu2 = q(1-q),
    # mu3 = q(1-q)(1-2q), mu4 = q(1-q)(1-3q+3q^2)
    q, n = 0.25, 4000
    b = np.zeros(n)
    b[: int(q * n)] = 1.0
    out = tavg_eval(b)
    mu2, mu3, mu4 = q * (1 - q), q * (1 - q) * (1 - 2 * q), q * (1 - q) * (1 - 3 * q + 3 * q * q)
    c1dev = (mu4 - mu2 * mu2) / n
    c2dev = mu2 / n
    cov = mu3 / n
    err = abs(mu2 / q) * np.sqrt(c1dev / mu2 ** 2 + c2dev / q ** 2 - 2 * cov / mu2 / q)
    assert out[4] == pytest.approx(1 - q, rel=1e-12)
    assert out[5] == pytest.approx(err, rel=1e-9)
