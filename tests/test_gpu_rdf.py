"""RDF histogram parity: bit-exact bin counts against the oracle on identical
positions, for the three binning rules (MDSystem.cpp:273-285 CPU rule,
MDSystem.cu:26-42 GPU rule, uniform-r bins of the workloads)."""
import numpy as np
import pytest

import ljamd
from helpers import perturbed_state
from oracle_lib import ORC, RDF_RBINS, RDF_REF_CPU, RDF_REF_GPU, forces_ref, rdf

pytestmark = pytest.mark.gpu

RULES = {RDF_REF_CPU: ljamd.RDF_REF_CPU, RDF_REF_GPU: ljamd.RDF_REF_GPU,
         RDF_RBINS: ljamd.RDF_RBINS}


@pytest.mark.parametrize("N,rho,bc,rule,nbins", [
    (400, 0.05, ljamd.BC_PERIODIC, RDF_REF_CPU, 256),
    (400, 0.05, ljamd.BC_PERIODIC, RDF_REF_GPU, 256),
    (400, 0.05, ljamd.BC_PERIODIC, RDF_RBINS, 200),   # the N = 400 workload: 200 bins to L/2
    (2048, 0.85, ljamd.BC_WALL, RDF_REF_CPU, 256),
    (2048, 0.85, ljamd.BC_WALL, RDF_RBINS, 300),
    (3001, 0.3, ljamd.BC_PERIODIC, RDF_REF_GPU, 256),
    (16384, 0.85, ljamd.BC_PERIODIC, RDF_RBINS, 500),  # the N = 16384 workload: 500 bins
    (16384, 0.85, ljamd.BC_PERIODIC, RDF_REF_CPU, 256),
])
def test_rdf_bitexact(N, rho, bc, rule, nbins):
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=31)
    print(f"\n[case] rdf N={N} bc={bc} rule={rule} nbins={nbins}", flush=True)
    dr2 = ORC.orc_rdf_dr2(N)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc) as s:
        s.set_state(pos)
        got = s.rdf(RULES[rule], nbins, dr2=dr2, rmax=L / 2)
    exp = rdf(pos, L, bc == ljamd.BC_PERIODIC, rule, nbins, dr2=dr2, rmax=L / 2)
    assert np.array_equal(got.astype(np.int64), exp), np.nonzero(got.astype(np.int64) - exp)
    if rule != RDF_REF_GPU:
        assert np.all(exp % 2 == 0)  # ordered pairs: every bin even
    if rule == RDF_RBINS:
        assert exp.sum() <= N * (N - 1)


def test_rdf_matches_reference_cpu_counts():
    """The CPU rule equals the reference CPU path's NdNdr2 (MDSystem.cpp:254,281-285)."""
    N, rho = 1200, 0.4
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=2)
    dr2 = ORC.orc_rdf_dr2(N)
    _, _, _, _, rdf_ref = forces_ref(pos, L, 0, dr2, nthreads=8)
    with ljamd.LJSystem(N, rho, 1.0) as s:
        s.set_state(pos)
        got = s.rdf(ljamd.RDF_REF_CPU, 256, dr2=dr2)
    assert np.array_equal(got.astype(np.int64), rdf_ref.astype(np.int64))


def test_rdf_accumulate():
    N, rho = 800, 0.5
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=12)
    with ljamd.LJSystem(N, rho, 1.0) as s:
        s.set_state(pos)
        a = s.rdf(ljamd.RDF_RBINS, 100, rmax=L / 2)
        b = s.rdf(ljamd.RDF_RBINS, 100, rmax=L / 2, accumulate=True)
        c = s.rdf(ljamd.RDF_RBINS, 100, rmax=L / 2, accumulate=False)
    assert np.array_equal(b, 2 * a) and np.array_equal(c, a)


def test_rdf_timing_leaves_rdf_exact():
    """lj_time_rdf runs reps + 1 passes into the device histogram; the next
    lj_rdf (even asked to accumulate) starts from zero and stays exact."""
    N, rho = 800, 0.5
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=12)
    with ljamd.LJSystem(N, rho, 1.0) as s:
        s.set_state(pos)
        a = s.rdf(ljamd.RDF_RBINS, 100, rmax=L / 2)
        us = s.time_rdf(ljamd.RDF_RBINS, 100, rmax=L / 2, reps=3)
        b = s.rdf(ljamd.RDF_RBINS, 100, rmax=L / 2, accumulate=True)
        with pytest.raises(ljamd.LJError):
            s.time_rdf(ljamd.RDF_RBINS, 100, reps=0)
    assert us > 0 and np.array_equal(a, b)


def test_rdf_bad_args():
    with ljamd.LJSystem(100, 0.5, 1.0) as s:
        with pytest.raises(ljamd.LJError):
            s.rdf(7, 10)
        with pytest.raises(ljamd.LJError):
            s.rdf(ljamd.RDF_RBINS, 0)
        with pytest.raises(ljamd.LJError):
            s.rdf(ljamd.RDF_RBINS, 100000)
