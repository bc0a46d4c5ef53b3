"""The compatibility set (reference MDSystem.cu:155-299) on the device: the
forces, pressure and RDF that the reference class's GPU branch
(MDSystem.cpp:240-251) reads from calculateNForces, against the FP64 oracle and
the oracle's GPU-rule RDF.  The reference class itself, compiled with
USE_CUDA_TOOLKIT and linked against libljhip.so, is checked on the CPU
(tests/test_ref_class_link.py): the compiled reference stays on the machine
that has its sources."""
import numpy as np
import pytest

import ljamd
from helpers import force_errors, perturbed_state, rel, xyz
from oracle_lib import ORC, RDF_REF_GPU, forces_fp64, rdf

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
