"""The compatibility set (reference MDSystem.cu:155-299) and the drop-in check:
the reference's own MDSystem class (compiled from its sources with
USE_CUDA_TOOLKIT and linked against libljhip.so, oracle/_ref/libref_hip.so)
drives the HIP path through its unchanged CPU/GPU switch; its results are
compared with the same class's CPU path (libref_cpu.so)."""
import numpy as np
import pytest

import ljamd
from helpers import force_errors, perturbed_state, rel, xyz
from oracle_lib import ORC, RDF_REF_GPU, RefSystem, forces_fp64, rdf, ref_lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,rho,periodic", [(400, 0.05, 1), (2048, 0.85, 1), (2048, 0.85, 0),
                                            (16384, 0.85, 1), (257, 0.3, 1)])
def test_calculateNForces_semantics(N, rho, periodic):
    """Force = (4 sum r s, sum (r^-12 - r^-6)), pressure = sum (4/6) r.f, RDF GPU rule."""
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=41)
    print(f"\n[case] compat N={N} periodic={periodic}", flush=True)
    dr2 = ORC.orc_rdf_dr2(N)
    arr = ljamd.CompatArrays(N)
    try:
        frc, pressure, rdf_got = arr.calculate(pos, np.float32(L), periodic, dr2)
    finally:
        arr.close()
    F64, ew, E, W = forces_fp64(pos, float(np.float32(L)), periodic, skip_r2=1e-6)
    err = force_errors(xyz(frc), F64)
    assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5, err
    w = frc.reshape(N, 4)[:, 3]
    assert np.max(np.abs(w - ew[:, 0]) / np.maximum(np.abs(ew[:, 0]), 1.0)) <= 1e-5
    assert rel(pressure, W / 3.0) <= 1e-6
    exp = rdf(pos, float(np.float32(L)), periodic, RDF_REF_GPU, 256, dr2=dr2)
    assert np.array_equal(rdf_got.astype(np.int64), exp)


def test_coincident_particles_skipped():
    """r^2 <= 1e-6 pairs are skipped (MDSystem.cu:37): a duplicated particle
    contributes nothing."""
    N, rho = 300, 0.3
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=7)
    p = pos.reshape(N, 4).copy()
    p[5, :3] = p[4, :3]
    arr = ljamd.CompatArrays(N)
    try:
        frc, pressure, rdf_got = arr.calculate(p.reshape(-1), np.float32(L), 1, 0.1)
    finally:
        arr.close()
    assert np.all(np.isfinite(frc))
    F64, _, _, W = forces_fp64(p.reshape(-1), float(np.float32(L)), 1, skip_r2=1e-6)
    assert force_errors(xyz(frc), F64)["global_rel"] <= 2e-6


def test_compat_errors_do_not_crash():
    # null device arrays: reported, not a crash
    ljamd.lib().calculateNForces(None, None, None, 10, 1.0, 1, None, 0.1, 256, 1)
    assert b"null" in ljamd.lib().lj_last_error()


@pytest.mark.parametrize("N,rho,bc,canon", [(400, 0.05, 0, 0), (1000, 0.85, 0, 1),
                                            (1000, 0.6, 1, 1)])
def test_reference_class_on_hip_path(N, rho, bc, canon):
    """Reference MDSystem, useCUDA=true (our C-ABI) vs useCUDA=false (its CPU)."""
    hip, cpu = ref_lib("hip"), ref_lib("cpu")
    if hip is None or cpu is None:
        pytest.skip("reference not built into oracle/_ref")
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=3)
    rg = RefSystem(hip, N, 1.0, rho, canon, bc, use_gpu=True)
    rc = RefSystem(cpu, N, 1.0, rho, canon, bc, use_gpu=False)
    try:
        rg.set_state(pos, vel)
        rc.set_state(pos, vel)
        fg, _, Pg, _, rdfg = rg.forces()
        fc, Vc, Pc, _, rdfc = rc.forces()
        F64, _, E, W = forces_fp64(pos, L, bc == 0, skip_r2=1e-6)
        assert force_errors(xyz(fg), F64)["max_rel_floor"] <= 1e-5
        assert rel(Pg, W / 3) <= 1e-6 and rel(Pc, W / 3) <= 1e-5
        # GPU-rule RDF of the reference's GPU branch == oracle GPU rule
        exp = rdf(pos, float(np.float32(L)), bc == 0, RDF_REF_GPU, 256, dr2=rg.dr2)
        assert np.array_equal(rdfg.astype(np.int64), exp)
        # short trajectory through the reference integrator on both paths
        rg.params()
        rc.params()
        hg = rg.integrate(0.005, 50)
        hc = rc.integrate(0.005, 50)
        dU = np.abs(hg[:, 0] - hc[:, 0]) / N
        print(f"\n  ref-class traj N={N}: max |dU|/N over 50 steps {dU.max():.2e}")
        assert dU.max() <= 1e-4
    finally:
        rg.close()
        rc.close()
