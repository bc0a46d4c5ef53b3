"""Small and ragged systems through every device path: the LJ dimer's closed
forms, N = 1 and 2, sizes just off a wave (64) or a pair-once group (256),
and the pair-once kernel forced on at small N (one group: a diagonal item
only; a partial last group of one particle), against the FP64 oracle, the
reference's drift arithmetic and the exact RDF counts of the oracle."""
import os

import numpy as np
import pytest

import ljamd
from helpers import force_errors, perturbed_state, rel, xyz
from oracle_lib import ORC, RDF_RBINS, RDF_REF_CPU, RDF_REF_GPU, forces_fp64, rdf

pytestmark = pytest.mark.gpu

RULES = {RDF_REF_CPU: ljamd.RDF_REF_CPU, RDF_REF_GPU: ljamd.RDF_REF_GPU,
         RDF_RBINS: ljamd.RDF_RBINS}


@pytest.mark.parametrize("r,V,F", [(2.0 ** (1.0 / 6.0), -1.0, 0.0), (1.0, 0.0, 24.0),
                                   (1.5, 4 * (1.5 ** -12 - 1.5 ** -6), None)])
def test_dimer_closed_form(r, V, F):
    """V = 4 (r^-12 - r^-6), |F| = 24 |2 r^-13 - r^-7| (sigma = eps = 1)."""
    F = abs(24 * (2 * r ** -13 - r ** -7)) if F is None else F
    pos = np.array([[1.0, 1.0, 1.0, 0.0], [1.0 + r, 1.0, 1.0, 0.0]], np.float32).reshape(-1)
    with ljamd.LJSystem(2, 0.001, 1.0, boundary=ljamd.BC_NONE) as s:
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        f = s.force.reshape(2, 4)
    assert abs(obs["V"] - V) <= 1e-6 * max(1.0, abs(V))
    assert abs(abs(f[0, 0]) - F) <= 1e-5 * max(1.0, F) and abs(f[0, 1]) + abs(f[0, 2]) == 0.0
    assert np.array_equal(f[0, :3], -f[1, :3])  # Newton's third law, exactly for two bodies


def test_single_particle():
    """No pairs: zero force, energy and virial; a free particle drifts at
    constant velocity (EVN); every RDF bin is zero."""
    pos = np.array([2.0, 3.0, 4.0, 0.0], np.float32)
    vel = np.array([0.5, -0.25, 0.125, 0.0], np.float32)
    with ljamd.LJSystem(1, 0.01, 1.0, boundary=ljamd.BC_NONE) as s:
        s.set_state(pos, vel)
        obs = s.CalculateForces()
        assert np.all(s.force == 0.0) and obs["V"] == 0.0 and obs["W"] == 0.0
        s.Integrate(0.01, 10)
        p = s.pos
        for rule in (ljamd.RDF_REF_CPU, ljamd.RDF_REF_GPU, ljamd.RDF_RBINS):
            assert not np.any(s.rdf(rule, 50, rmax=2.0))
    # the reference's drift x = (float)(x + (dt v + dt^2 f / 2)) in double
    # (MDSystem.cpp:447-449) with f = 0, ten times: the same bits
    x = pos[:3].copy()
    for _ in range(10):
        x = (x.astype(np.float64) + 0.01 * vel[:3].astype(np.float64)).astype(np.float32)
    assert np.array_equal(p[:3], x)


@pytest.mark.parametrize("N,bc", [(2, ljamd.BC_PERIODIC), (3, ljamd.BC_PERIODIC), (63, ljamd.BC_WALL),
                                  (65, ljamd.BC_PERIODIC), (257, ljamd.BC_PERIODIC),
                                  (257, ljamd.BC_NONE)])
def test_ragged_forces_and_rdf(N, bc):
    rho = 0.5
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=41)
    print(f"\n[case] ragged N={N} bc={bc}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc) as s:
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        f4 = s.force
        got = {rule: s.rdf(RULES[rule], nb, dr2=ORC.orc_rdf_dr2(N), rmax=L / 2)
               for rule, nb in ((RDF_REF_CPU, 256), (RDF_REF_GPU, 256), (RDF_RBINS, 40))}
    F64, _, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    err = force_errors(xyz(f4), F64)
    assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5, err
    # 3e-6: in these few-particle boxes (L = 1.6 to 5 sigma) e = (b - a) / 2
    # cancels more of the per-particle float sums (measured 1.23e-6 at N = 2, 65)
    assert rel(obs["V"], E) <= 3e-6 and rel(obs["W"], W) <= 1e-5
    for rule, nb in ((RDF_REF_CPU, 256), (RDF_REF_GPU, 256), (RDF_RBINS, 40)):
        exp = rdf(pos, L, bc == ljamd.BC_PERIODIC, rule, nb, dr2=ORC.orc_rdf_dr2(N), rmax=L / 2)
        assert np.array_equal(got[rule].astype(np.int64), exp), rule


@pytest.fixture
def pair_once_env():
    old = os.environ.get("LJ_PAIR_ONCE")
    yield
    if old is None:
        os.environ.pop("LJ_PAIR_ONCE", None)
    else:
        os.environ["LJ_PAIR_ONCE"] = old


@pytest.mark.parametrize("prec", [ljamd.PREC_FP32, ljamd.PREC_COMPENSATED])
@pytest.mark.parametrize("N,bc", [(256, ljamd.BC_PERIODIC), (257, ljamd.BC_WALL),
                                  (700, ljamd.BC_PERIODIC)])
def test_pair_once_forced_small(N, bc, prec, pair_once_env):
    """LJ_PAIR_ONCE=1 at small N, in the FP32 and the compensated mode: one
    group (the diagonal item alone), a last group of a single particle, three
    groups; forces against FP64, and a 20-step TVN run (the step variant)
    against the sliced kernel's (LJ_PAIR_ONCE=0) from the same state."""
    rho = 0.6
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=43)
    print(f"\n[case] pair-once forced N={N} bc={bc} prec={prec}", flush=True)
    runs = {}
    for flag in ("1", "0"):
        os.environ["LJ_PAIR_ONCE"] = flag
        with ljamd.LJSystem(N, rho, 1.0, canonical=ljamd.TVN, boundary=bc, precision=prec) as s:
            assert s.force_plan()[0] == (flag == "1")
            s.set_state(pos, vel)
            obs = s.CalculateForces()
            f4 = s.force
            h = s.Integrate(0.005, 20)
            runs[flag] = (f4, obs, h, s.pos)
    F64, _, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    for flag in ("1", "0"):
        err = force_errors(xyz(runs[flag][0]), F64)
        assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5, (flag, err)
        assert rel(runs[flag][1]["V"], E) <= 1e-6
    dU = np.max(np.abs(runs["1"][2][:, 0] - runs["0"][2][:, 0])) / N
    dx = np.max(np.abs(xyz(runs["1"][3]) - xyz(runs["0"][3])))
    assert dU <= 1e-5 and dx <= 1e-4, (dU, dx)


@pytest.fixture
def waves_env():
    old = {k: os.environ.get(k) for k in ("LJ_PAIR_ONCE", "LJ_PAIR_ONCE_WAVES")}
    yield
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("N,bc", [(1300, ljamd.BC_PERIODIC), (1100, ljamd.BC_WALL)])
def test_pair_once_sixteen_wave_groups(N, bc, waves_env):
    """LJ_PAIR_ONCE_WAVES=4 (groups of 512, 16-wave workgroups, single diagonal
    items also in the time step): forces against FP64 and a 20-step TVN run
    against the default groups of 256 from the same state."""
    rho = 0.6
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=47)
    print(f"\n[case] pair-once 16-wave groups N={N} bc={bc}", flush=True)
    os.environ["LJ_PAIR_ONCE"] = "1"
    runs = {}
    for waves in ("4", "2"):
        os.environ["LJ_PAIR_ONCE_WAVES"] = waves
        with ljamd.LJSystem(N, rho, 1.0, canonical=ljamd.TVN, boundary=bc) as s:
            ng = (N + 128 * int(waves) - 1) // (128 * int(waves))
            # groups of 256 this few run every item as two workgroups (2 ng slots)
            assert s.force_plan() == (True, 2 * ng if waves == "2" else ng)
            s.set_state(pos, vel)
            obs = s.CalculateForces()
            f4 = s.force
            h = s.Integrate(0.005, 20)
            runs[waves] = (f4, obs, h, s.pos)
    F64, _, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    for waves in ("4", "2"):
        err = force_errors(xyz(runs[waves][0]), F64)
        assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5, (waves, err)
        assert rel(runs[waves][1]["V"], E) <= 1e-6 and rel(runs[waves][1]["W"], W) <= 1e-5
    dU = np.max(np.abs(runs["4"][2][:, 0] - runs["2"][2][:, 0])) / N
    dx = np.max(np.abs(xyz(runs["4"][3]) - xyz(runs["2"][3])))
    assert dU <= 1e-5 and dx <= 1e-4, (dU, dx)


@pytest.fixture
def split_env():
    old = {k: os.environ.get(k) for k in ("LJ_PAIR_ONCE", "LJ_PAIR_ONCE_SPLIT")}
    yield
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("N,bc", [(8192, ljamd.BC_PERIODIC), (7001, ljamd.BC_WALL),
                                  (6144, ljamd.BC_NONE), (300, ljamd.BC_PERIODIC)])
def test_pair_once_split_items(N, bc, split_env):
    """Split items (each group pair run by two workgroups, one 64-particle j
    sub-tile of every wave each; the default while twice the items fit in one
    round of resident workgroups) against the unsplit pass (LJ_PAIR_ONCE_SPLIT
    = 0, the diagonal items paired in the time step) on the same state: slot
    counts, forces against FP64 for both, the two passes' forces against each
    other, V and W of the step variant against a full evaluation, and a 20-step
    TVN run.  7001 ends in a partial group of 89; N = 300 forces the kernel on
    at two groups."""
    rho = 0.85
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=53)
    print(f"\n[case] pair-once split N={N} bc={bc}", flush=True)
    if N < 3072:
        os.environ["LJ_PAIR_ONCE"] = "1"
    ng = (N + 255) // 256
    runs = {}
    for flag in ("1", "0"):
        os.environ["LJ_PAIR_ONCE_SPLIT"] = flag
        with ljamd.LJSystem(N, rho, 1.0, canonical=ljamd.TVN, boundary=bc) as s:
            assert s.force_plan() == (True, 2 * ng if flag == "1" else ng)
            s.set_state(pos, vel)
            obs = s.CalculateForces()
            f4 = s.force
            h = s.Integrate(0.005, 20)
            o_step = s.observables()
            o_full = s.CalculateForces()  # full evaluation at the positions of step 20
            runs[flag] = (f4, obs, h, s.pos, o_step, o_full)
    F64, _, E, W = forces_fp64(pos, L, bc == ljamd.BC_PERIODIC)
    for flag in ("1", "0"):
        f4, obs, h, _, o_step, o_full = runs[flag]
        err = force_errors(xyz(f4), F64)
        print(f"  split={flag}: {err} V {rel(obs['V'], E):.2e} W {rel(obs['W'], W):.2e}")
        assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5, (flag, err)
        # 3e-6 (as for the ragged boxes above): e = (b - a) / 2 cancels part of
        # the per-particle float sums; measured 1.23e-6 unsplit at N = 7001 with
        # hard walls, 2.27e-6 split at N = 300 (V = -280 over 300 particles)
        assert rel(obs["V"], E) <= 3e-6 and rel(obs["W"], W) <= 1e-5
        assert abs(h[-1, 3] - o_full["V"]) <= 1e-6 * abs(o_full["V"])
        assert abs(o_step["W"] - o_full["W"]) <= 1e-6 * abs(o_full["W"])
    assert force_errors(xyz(runs["1"][0]), xyz(runs["0"][0]))["global_rel"] <= 2e-6
    dU = np.max(np.abs(runs["1"][2][:, 0] - runs["0"][2][:, 0])) / N
    dx = np.max(np.abs(xyz(runs["1"][3]) - xyz(runs["0"][3])))
    assert dU <= 1e-5 and dx <= 1e-4, (dU, dx)


# the kernel choice at the sizes the measurements cover (rho* = 0.85, FP32):
# (N, pair-once, partial slots per particle); split items give 2 per group
PLAN_CASES = [(2048, False, 32), (3071, False, 48), (3072, True, 24), (4096, True, 32),
              (6144, True, 48), (8192, True, 32), (9216, True, 72), (10000, True, 80),
              (11264, True, 44), (12288, True, 48), (13001, True, 51), (16384, True, 64)]


@pytest.mark.parametrize("N,pair_once,slots", PLAN_CASES)
def test_force_plan_choice(N, pair_once, slots, split_env):
    """Sliced below N = 3072; pair-once above, with split items (2 slots per
    group of 256) where the occupancy model of a 256-CU device predicts a 10%
    faster pass: 12-31, 33-38 and 40-41 groups (N = 3072-7936, 8193-9728,
    9985-10496), not 32 groups (N = 8192) nor from 42 on."""
    for k in ("LJ_PAIR_ONCE", "LJ_PAIR_ONCE_SPLIT"):
        os.environ.pop(k, None)
    with ljamd.LJSystem(N, 0.85, 1.0, canonical=ljamd.TVN) as s:
        assert s.force_plan() == (pair_once, slots)


def test_maxwell_initial_state():
    """The reference's own initial-velocity sampler (MDSystem.cpp:36-53,
    147-181), seeded: zero momentum and T = T0 exactly (float rounding), speeds
    Maxwell-distributed at T0 (Kolmogorov-Smirnov against the chi distribution
    with 3 degrees of freedom, scale sqrt(T0)), isotropic directions, the
    lattice positions unchanged, reproducible per seed."""
    from scipy import stats

    N, rho, T0 = 16384, 0.85, 1.3
    with ljamd.LJSystem(N, rho, T0) as s:
        s.sample_initial_state(5, ljamd.INIT_MAXWELL)
        pos = s.get_array(ljamd.ARRAY_POS).reshape(N, 4)
        v = s.get_array(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3].astype(np.float64)
        s.sample_initial_state(5, ljamd.INIT_MAXWELL)
        v_again = s.get_array(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3]
        s.sample_initial_state(6, ljamd.INIT_MAXWELL)
        v_other = s.get_array(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3]
        s.sample_initial_state(5)
        pos_g = s.get_array(ljamd.ARRAY_POS).reshape(N, 4)
        v_g = s.get_array(ljamd.ARRAY_VEL).reshape(N, 4)[:, :3]
        with pytest.raises(ljamd.LJError):
            s.sample_initial_state(5, 7)
    assert np.array_equal(v, v_again) and not np.array_equal(v, v_other)
    assert np.array_equal(pos, pos_g) and not np.array_equal(v[:, :3].astype(np.float32), v_g)
    assert np.all(np.abs(v.sum(axis=0)) <= 1e-3)  # float momentum of 16384 particles
    assert abs(np.sum(v * v) / (3 * N) - T0) <= 1e-6 * T0
    speed = np.linalg.norm(v, axis=1)
    ks = stats.kstest(speed, stats.chi(df=3, scale=np.sqrt(T0)).cdf)
    print(f"\n  KS statistic {ks.statistic:.4f} (p = {ks.pvalue:.3f})")
    assert ks.pvalue > 1e-3
    cth = v[:, 2] / speed
    assert abs(np.mean(cth)) < 0.03 and abs(np.mean(cth ** 2) - 1 / 3) < 0.02
    # the Gaussian draw, for contrast, is Maxwell-distributed too (same law)
    assert stats.kstest(np.linalg.norm(v_g.astype(np.float64), axis=1),
                        stats.chi(df=3, scale=np.sqrt(T0)).cdf).pvalue > 1e-3


@pytest.mark.parametrize("jslices", [0, 32])
def test_stamped_force_timing(jslices):
    """lj_time_force_kernel_stamped: each launch's own duration (hipExtLaunchKernel
    start / stop events) is positive and no longer than the back-to-back time
    per launch, for the pair-once (jslices 0 at N = 16384) and the sliced
    kernel; the stamped launches leave the forces as the plain ones do;
    out-of-range repetition counts are refused."""
    N = 16384
    with ljamd.LJSystem(N, 0.85, 1.0, seed=3, jslices=jslices) as s:
        s.CalculateForces()
        f0 = s.force.copy()
        k, span = s.time_force_kernel_stamped(50)
        assert 0.0 < k <= span * 1.02
        s.CalculateForces()
        assert np.array_equal(s.force, f0)
        for bad in (0, 5000):
            with pytest.raises(ljamd.LJError):
                s.time_force_kernel_stamped(bad)
