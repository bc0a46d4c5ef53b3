"""One GPU's share of a force pass sharded by i-particles (lj_forces_shard),
the single-GPU part of the multi-GPU design: whole 64-particle tiles of rows,
about N / G of them, against all N.  With the slicing of a full sliced evaluation the shares
reassemble it bit for bit and their E_pot / W partials add up to its totals;
with the default slicing each share fills the device, and at N = 16384 over
8 shares one share's kernel is timed against the whole pass."""
import os

import numpy as np
import pytest

import ljamd
from helpers import force_errors, perturbed_state, xyz
from oracle_lib import forces_fp64, lattice_state

pytestmark = pytest.mark.gpu


# slice counts whose j chunks are multiples of 8 (kept as given) and whose
# full pass runs the plain sliced grid (at least four waves per SIMD)
@pytest.mark.parametrize("N,bc,G,S,prec", [(4096, ljamd.BC_PERIODIC, 4, 64, ljamd.PREC_FP32),
                                           (6000, ljamd.BC_WALL, 3, 75, ljamd.PREC_FP32),
                                           (16384, ljamd.BC_PERIODIC, 8, 32, ljamd.PREC_FP32),
                                           (3001, ljamd.BC_NONE, 7, 94, ljamd.PREC_FP32),
                                           (6000, ljamd.BC_PERIODIC, 5, 75, ljamd.PREC_COMPENSATED)])
def test_shards_reassemble_full_pass(N, bc, G, S, prec):
    """Shares with the full pass's j slices: bit-identical force rows
    (Force.w included), E_pot and W partials summing to the totals; shares
    of uneven size (6000 / 3: 1984 + 1984 + 2032 rows; 3001 / 7); the
    compensated mode's sliced kernel (TwoSum blocks) alike."""
    rho = 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=31)
    print(f"\n[case] shards N={N} bc={bc} G={G} S={S} prec={prec}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, boundary=bc, jslices=S, precision=prec) as s:
        assert s.force_plan() == (False, S)
        s.set_state(pos, np.zeros_like(pos))
        obs = s.CalculateForces()
        full = s.force.copy()
        s.set_array(ljamd.ARRAY_FORCE, np.zeros_like(full))
        E = W = 0.0
        for g in range(G):
            e, w = s.forces_shard(g, G, jslices=S)
            E += e
            W += w
        sharded = s.force.copy()
    print(f"  E {E:.15g} vs {obs['V']:.15g}  W {W:.15g} vs {obs['W']:.15g}")
    assert np.array_equal(sharded, full)
    assert abs(E - obs["V"]) <= 1e-12 * abs(obs["V"])
    assert abs(W - obs["W"]) <= 1e-12 * abs(obs["W"])


def test_shard_default_slicing_n16384():
    """N = 16384 over 8 shares with the default slicing (each share fills the
    device): forces against FP64; one share's kernel against the whole pass
    on the same grid (one share)."""
    N, rho, G = 16384, 0.85, 8
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=31)
    print(f"\n[case] shard timing N={N} G={G}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0) as s:
        s.set_state(pos, np.zeros_like(pos))
        for g in range(G):
            s.forces_shard(g, G)
        f = s.force.copy()
        t8 = s.time_forces_shard(0, G, reps=200)
        t1 = s.time_forces_shard(0, 1, reps=50)
    F64, _, E, W = forces_fp64(pos, L, True)
    err = force_errors(xyz(f), F64)
    print(f"  vs FP64 {err}; one share of {G}: {t8:.2f} us, the whole pass: {t1:.2f} us "
          f"(ratio {t1 / t8:.2f})")
    assert err["global_rel"] <= 2e-6 and err["max_rel_floor"] <= 1e-5
    assert t8 < t1


def test_shard_bad_args():
    with ljamd.LJSystem(512, 0.5, 1.0) as s:
        for shard, nshards, js in [(0, 0, 0), (3, 3, 0), (-1, 2, 0), (0, 2, -1), (0, 513, 0)]:
            with pytest.raises(ljamd.LJError):
                s.forces_shard(shard, nshards, jslices=js)
    with ljamd.LJSystem(512, 0.5, 1.0, precision=ljamd.PREC_FP64) as s:
        with pytest.raises(ljamd.LJError):
            s.forces_shard(0, 2)


def test_pair_once_items_share_n16384():
    """The other sharding: the time step's pair-once pass with its items dealt
    out over 8 GPUs; one share's launch against the whole pass (timing only;
    the partial entries of the items are disjoint by construction)."""
    N, rho = 16384, 0.85
    print(f"\n[case] pair-once item share N={N}", flush=True)
    with ljamd.LJSystem(N, rho, 1.0, canonical=ljamd.TVN) as s:
        assert s.force_plan() == (True, 64)
        t8 = s.time_force_items_share(0, 8)
        t1 = s.time_force_items_share(0, 1, reps=50)
        with pytest.raises(ljamd.LJError):
            s.time_force_items_share(8, 8)
    print(f"  one share of 8: {t8:.2f} us, the whole pass: {t1:.2f} us (ratio {t1 / t8:.2f})")
    assert t8 < t1
    with ljamd.LJSystem(4096, rho, 1.0) as s:  # split items: no paired diagonals
        with pytest.raises(ljamd.LJError):
            s.time_force_items_share(0, 2)


@pytest.mark.parametrize("canon,bc", [(ljamd.TVN, ljamd.BC_PERIODIC), (ljamd.EVN, ljamd.BC_WALL)])
def test_emulated_item_shards_step_bitwise(canon, bc):
    """The time step with its pair-once pass run as 8 consecutive launches on
    contiguous item ranges (LJ_EMULATE_SHARDS=8, as 8 GPUs would run their
    shares): every item writes its own partial entries and totals, so the
    trajectory is bit-identical to the single launch's."""
    N, rho, T0 = 16384, 0.85, 1.0
    L, pos, vel = lattice_state(N, rho, T0, seed=5)
    print(f"\n[case] emulated item shards canon={canon} bc={bc}", flush=True)
    res = {}
    for g in ("1", "8"):
        os.environ["LJ_EMULATE_SHARDS"] = g
        try:
            with ljamd.LJSystem(N, rho, T0, canonical=canon, boundary=bc) as s:
                assert s.force_plan() == (True, 64)
                s.set_state(pos, vel)
                s.CalculateForces()
                h = s.Integrate(0.005, 20)
                res[g] = (h, s.pos.copy(), s.vel.copy())
        finally:
            os.environ.pop("LJ_EMULATE_SHARDS", None)
    for a, b in zip(res["1"], res["8"]):
        assert np.array_equal(a, b)

