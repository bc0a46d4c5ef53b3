"""The oracle against committed outputs of the reference's own CPU path
(tests/golden/*.npz, written by tests/golden/make_golden.py from MDSystem
compiled from its sources): one force evaluation (forces, V, P, Pshear, the
256 RDF bins) and a 40-step EVN / TVN trajectory (history rows, final
positions and velocities), bit-exact, for periodic, hard-wall and open boxes.
Unlike the live comparison in test_oracle_pins.py this needs no build of the
reference."""
import glob
import os

import numpy as np
import pytest

from oracle_lib import ORC, OracleSystem, forces_ref

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_cpu_*.npz")))


def case(path):
    name = os.path.basename(path)[len("ref_cpu_"):-len(".npz")]
    f = name.split("_")
    return (int(f[0][1:]), float(f[1][3:]), float(f[2][1:]), int(f[3][2:]), int(f[4][1:]))


def test_golden_fixtures_present():
    assert len(FILES) == 5


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_oracle_matches_reference_outputs(path):
    N, rho, T0, bc, canon = case(path)
    g = np.load(path)  # data only (no pickles)
    pos, vel, L = g["pos"], g["vel"], float(g["L"])
    m = np.arange(4 * N) % 4 < 3
    fo, Vo, Po, Pso, rdfo = forces_ref(pos, L, bc, ORC.orc_rdf_dr2(N))
    assert np.array_equal(fo[m].view(np.int32), g["force"][m].view(np.int32))
    assert (Vo, Po, Pso) == (float(g["V"]), float(g["P"]), float(g["Pshear"]))
    assert np.array_equal(rdfo, g["rdf"])
    O = OracleSystem(N, T0, rho, canon, bc)
    O.set_state(pos, vel)
    O.forces()
    O.params()
    h = O.integrate(float(g["dt"]), int(g["steps"]))
    assert np.array_equal(h, g["hist"])
    assert np.array_equal(O.pos[m], g["pos_end"][m]) and np.array_equal(O.vel[m], g["vel_end"][m])
