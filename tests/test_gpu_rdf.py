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
    # the cell-list pass (periodic; cells per axis at reach 1 / reach 2, the
    # layout taken by default being reach 2 in all four):
    (16384, 0.85, ljamd.BC_PERIODIC, RDF_REF_GPU, 256),  # 5 / 10
    (3000, 0.5, ljamd.BC_PERIODIC, RDF_REF_GPU, 256),     # L = 18.2: 3 / 7
    (6000, 0.85, ljamd.BC_PERIODIC, RDF_REF_CPU, 256),    # L = 19.2: 3 / 7
    (9000, 0.3, ljamd.BC_PERIODIC, RDF_REF_GPU, 128),     # L = 31.1, range r^2 < 12.8: 8 / 17
    # 512 bins of dr2 = 0.1 (range r^2 < 51.2): 3 / 7 cells per axis, 48
    # particles per reach-2 cell, about 3000 staged j per cell (two stages)
    (16384, 0.85, ljamd.BC_PERIODIC, RDF_REF_CPU, 512),
    (16384, 0.85, ljamd.BC_PERIODIC, RDF_REF_GPU, 512),
    # hard walls: the cell list without wrap (10 cells per axis at N = 16384;
    # the perturbed lattice puts some particles just outside [0, L])
    (16384, 0.85, ljamd.BC_WALL, RDF_REF_CPU, 256),
    (16384, 0.85, ljamd.BC_WALL, RDF_REF_GPU, 256),
    (9000, 0.3, ljamd.BC_WALL, RDF_REF_GPU, 128),
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


def test_rdf_cells_out_of_span_falls_back():
    """Coordinates beyond [-L, 2 L) void the cell-list pass's bound: lj_rdf
    then counts with the all-pairs pass, still bit-exact (one particle moved
    by +3 L, a periodic image of itself)."""
    N, rho = 6000, 0.85
    L, pos, _ = perturbed_state(N, rho, 1.0, seed=8)
    p = pos.reshape(-1, 4)
    p[17, 0] += np.float32(3.0 * L)
    pos = p.reshape(-1).copy()
    dr2 = ORC.orc_rdf_dr2(N)
    for rule in (RDF_REF_GPU, RDF_REF_CPU):
        with ljamd.LJSystem(N, rho, 1.0) as s:
            s.set_state(pos)
            got = s.rdf(RULES[rule], 256, dr2=dr2)
            again = s.rdf(RULES[rule], 256, dr2=dr2)
        exp = rdf(pos, L, True, rule, 256, dr2=dr2)
        assert np.array_equal(got.astype(np.int64), exp)
        assert np.array_equal(again, got)


def _cells_on_and_off(code, tmp_path, timeout=300):
    """Run `code` (argv: package path, output .npz) in three child processes:
    LJ_RDF_CELLS=1 (the cell-list pass with reach 1: cells of side >= the far
    bound's root), =2 (reach 2: half that side) and =0 (the all-pairs pass).
    Return the three lists of saved arrays."""
    import os
    import subprocess
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    py = os.path.join(os.path.dirname(here), "lennard-jones-cuda_amd", "python")
    res = {}
    for cells in ("1", "2", "0"):
        outf = str(tmp_path / f"cells{cells}.npz")
        env = dict(os.environ, LJ_RDF_CELLS=cells, PYTHONPATH=here + os.pathsep + py)
        r = subprocess.run([sys.executable, "-c", code, py, outf], env=env, cwd=here,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=timeout)
        assert r.returncode == 0, r.stderr[-2000:]
        with np.load(outf) as z:
            res[cells] = [z[k] for k in sorted(z.files, key=lambda t: int(t.split("_")[1]))]
    return res["1"], res["2"], res["0"]


def test_step_rdf_cells_equal_all_pairs_pass(tmp_path):
    """The step RDF of a periodic box at N = 6000 (3 cells per axis at reach
    1, 7 at reach 2): summed over 10 captured steps with each cell-list
    layout, and in a child process with LJ_RDF_CELLS=0 (the all-pairs pass)
    from the same state: the same counts, bit for bit, for both reference
    rules; the trajectory too."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import ljamd
from oracle_lib import lattice_state
N, rho, T0 = 6000, 0.85, 1.0
L, pos, vel = lattice_state(N, rho, T0, 21)
out = []
for rule in (ljamd.RDF_REF_GPU, ljamd.RDF_REF_CPU):
    with ljamd.LJSystem(N, rho, T0, canonical=ljamd.TVN) as s:
        s.set_state(pos, vel)
        s.CalculateForces()
        s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_SUM, rdf_rule=rule, rdf_nbins=256)
        s.Integrate(0.005, 10)
        got, n = s.step_rdf_counts()
        out.append(got.astype(np.int64))
        out.append(s.pos.astype(np.float64))
np.savez(sys.argv[2], *out)
"""
    on1, on2, off = _cells_on_and_off(code, tmp_path)
    for a, b, c in zip(on1, on2, off):
        assert np.array_equal(a, c)
        assert np.array_equal(b, c)
    assert on1[0].sum() > 0 and on1[2].sum() > 0


@pytest.mark.parametrize("N", [70000, 131072])
def test_rdf_cells_large_boxes_equal_all_pairs_pass(N, tmp_path):
    """Above 65536 particles the cell sort runs over many workgroups: per-block
    counts in LDS (N = 70000: 8^3 cells at reach 1) or global counters (N =
    131072: 10^3 cells at reach 1; reach 2 has 17^3 and 21^3 cells, global
    counters).  Both reference rules, perturbed lattice and after 20 TVN steps:
    the same counts as the all-pairs pass, bit for bit, and every pair counted
    (the GPU rule's total is N (N - 1): each ordered pair once)."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import ljamd
from helpers import perturbed_state
from oracle_lib import ORC
N = int(sys.argv[3])
L, pos, vel = perturbed_state(N, 0.85, 1.0, seed=5)
dr2 = ORC.orc_rdf_dr2(N)
out = []
with ljamd.LJSystem(N, 0.85, 1.0, canonical=ljamd.TVN) as s:
    s.set_state(pos, vel)
    for rule in (ljamd.RDF_REF_GPU, ljamd.RDF_REF_CPU):
        out.append(s.rdf(rule, 256, dr2=dr2).astype(np.int64))
    s.CalculateForces()
    s.Integrate(0.005, 20)
    for rule in (ljamd.RDF_REF_GPU, ljamd.RDF_REF_CPU):
        out.append(s.rdf(rule, 256, dr2=dr2).astype(np.int64))
np.savez(sys.argv[2], *out)
""".replace("int(sys.argv[3])", str(N))
    on1, on2, off = _cells_on_and_off(code, tmp_path)
    for a, b, c in zip(on1, on2, off):
        assert np.array_equal(a, c)
        assert np.array_equal(b, c)
    assert on1[0].sum() == N * (N - 1) and on1[2].sum() == N * (N - 1)
    assert 0 < on1[1].sum() < N * (N - 1)


def test_step_rdf_out_of_span_reported_after_lj_rdf():
    """A coordinate outside [-L, 2 L) seen by a sampled step's RDF pass makes
    lj_step_rdf_counts fail (the cell-list pass counted nothing for that
    step), also when an lj_rdf call (whose own flag is separate) comes in
    between; the sampling then resumes from a state inside the box."""
    N, rho = 6000, 0.85
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=8)
    p = pos.reshape(-1, 4).copy()
    p[17, 0] += np.float32(3.0 * L)
    with ljamd.LJSystem(N, rho, 1.0, canonical=ljamd.TVN) as s:
        s.set_state(p.reshape(-1), vel)
        s.CalculateForces()
        s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_SUM, rdf_rule=ljamd.RDF_REF_GPU, rdf_nbins=256)
        s.Integrate(0.005, 1)  # the force pass sees x + 3 L; the step wraps it
        got = s.rdf(ljamd.RDF_REF_GPU, 256, dr2=ORC.orc_rdf_dr2(N))
        assert got.sum() == N * (N - 1)
        with pytest.raises(ljamd.LJError):
            s.step_rdf_counts()
        # the periodic wrap moves a coordinate by one L per step: start again
        # from the state inside the box
        s.set_state(pos, vel)
        s.step_rdf_counts(clear=True)  # the flag was reported once
        s.Integrate(0.005, 2)
        counts, n = s.step_rdf_counts()
        assert n == 2 and counts.sum() == 2 * N * (N - 1)


def test_rdf_cells_crowded_cells_equal_all_pairs_pass(tmp_path):
    """Cells holding more particles than a workgroup has lanes: N = 70000,
    1024 bins of dr2 = 0.18 (range r^2 < 184), 3 cells per axis at reach 1
    (2600 particles each) and 6 at reach 2 (324 each: two i rounds, 14 j
    stages).  Both rules, both layouts: the all-pairs pass's counts."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import ljamd
from helpers import perturbed_state
N = 70000
L, pos, vel = perturbed_state(N, 0.85, 1.0, seed=6)
out = []
with ljamd.LJSystem(N, 0.85, 1.0) as s:
    s.set_state(pos, vel)
    for rule in (ljamd.RDF_REF_GPU, ljamd.RDF_REF_CPU):
        out.append(s.rdf(rule, 1024, dr2=0.18).astype(np.int64))
np.savez(sys.argv[2], *out)
"""
    on1, on2, off = _cells_on_and_off(code, tmp_path)
    for a, b, c in zip(on1, on2, off):
        assert np.array_equal(a, c)
        assert np.array_equal(b, c)
    N = 70000
    assert on2[0].sum() == N * (N - 1) and 0 < on2[1].sum() < N * (N - 1)


def test_rdf_wall_cells_equal_all_pairs_pass(tmp_path):
    """Hard walls at N = 16384: the captured per-step sums of 30 TVN steps
    and the trajectory, then both reference rules on that state with the
    particles near each wall moved 0.4 outside [0, L] (the clamped edge
    cells), counted over the wall cell list, equal the all-pairs pass
    (LJ_RDF_CELLS=0), bit for bit."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import ljamd
from helpers import perturbed_state
N = 16384
L, pos, vel = perturbed_state(N, 0.85, 1.0, seed=12)
out = []
with ljamd.LJSystem(N, 0.85, 1.0, canonical=ljamd.TVN, boundary=ljamd.BC_WALL) as s:
    s.set_state(pos, vel)
    s.CalculateForces()
    s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_SUM, rdf_rule=ljamd.RDF_REF_GPU, rdf_nbins=256)
    s.Integrate(0.005, 30)
    got, n = s.step_rdf_counts()
    out.append(got.astype(np.int64))
    out.append(s.pos.astype(np.float64))
    p = s.pos.reshape(-1, 4).copy()
    xyz = p[:, :3]
    xyz[xyz < 1.0] -= np.float32(0.4)
    xyz[xyz > L - 1.0] += np.float32(0.4)
    p[:, :3] = xyz
    out.append(np.array([np.sum((xyz < 0) | (xyz > L))]))
    s.set_state(p.reshape(-1), s.vel)
    for rule in (ljamd.RDF_REF_GPU, ljamd.RDF_REF_CPU):
        out.append(s.rdf(rule, 256, dr2=0.1).astype(np.int64))
np.savez(sys.argv[2], *out)
"""
    on1, on2, off = _cells_on_and_off(code, tmp_path)
    for a, b, c in zip(on1, on2, off):
        assert np.array_equal(a, c)
        assert np.array_equal(b, c)
    N = 16384
    assert on1[0].sum() == 30 * N * (N - 1)
    assert on1[2][0] > 1000  # coordinates outside [0, L]: the clamped edge cells
    assert on1[3].sum() == N * (N - 1) and 0 < on1[4].sum() < N * (N - 1)


def test_step_rdf_after_lj_rdf_grows_the_cell_workspace():
    """lj_rdf with more bins than the per-step sampling reallocates the
    cell-list workspace the captured steps use: the steps are captured again,
    and the sampled sums and the trajectory equal a run without that call."""
    N, rho, T0 = 16384, 0.85, 1.0
    L, pos, vel = perturbed_state(N, rho, T0, seed=14)
    runs = []
    for grow in (False, True):
        with ljamd.LJSystem(N, rho, T0, canonical=ljamd.TVN) as s:
            s.set_state(pos, vel)
            s.CalculateForces()
            s.set_step_sampling(rdf=ljamd.SAMPLE_RDF_SUM, rdf_rule=ljamd.RDF_REF_GPU, rdf_nbins=256)
            s.Integrate(0.005, 5)
            if grow:  # 1024 bins over the same range: a larger workspace
                got = s.rdf(ljamd.RDF_REF_CPU, 1024, dr2=0.025)
                assert 0 < got.sum() < N * (N - 1)
            s.Integrate(0.005, 5)
            counts, n = s.step_rdf_counts()
            runs.append((counts.astype(np.int64), n, s.pos.copy()))
    assert runs[0][1] == runs[1][1] == 10
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][2], runs[1][2])
