"""Writes the reference-run fixtures of tests/test_oracle_golden.py: the
reference's own CPU path (MDSystem compiled from its sources by
oracle/Makefile into oracle/_ref/libref_cpu.so) run on seeded states, its
outputs saved as data.  Run from the repository root after building the
test-oracle libraries (`python -c "import __graft_entry__ as g; g.build()"`):

    python tests/golden/make_golden.py

Each case stores the input state (positions, velocities), the reference's
forces, V, P, Pshear and 256 RDF bins of one force evaluation, and the history
rows and final positions / velocities of 40 steps (dt = 0.005)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import perturbed_state  # noqa: E402
from oracle_lib import RefSystem, ref_lib  # noqa: E402

CASES = [  # N, rho, T0, boundary (0 periodic, 1 wall, 2 open), canonical (0 EVN, 1 TVN)
    (400, 0.05, 1.4, 0, 0), (343, 0.85, 1.0, 0, 1), (300, 0.6, 1.0, 1, 1),
    (200, 0.3, 1.0, 2, 0), (150, 0.9, 2.0, 1, 0)]
STEPS, DT, SEED = 40, 0.005, 99


def case_name(N, rho, T0, bc, canon):
    return f"ref_cpu_n{N}_rho{rho}_t{T0}_bc{bc}_c{canon}.npz"


def main():
    lib = ref_lib("cpu")
    if lib is None:
        sys.exit("oracle/_ref/libref_cpu.so is not built (it needs the reference sources)")
    for N, rho, T0, bc, canon in CASES:
        L, pos, vel = perturbed_state(N, rho, T0, seed=SEED)
        R = RefSystem(lib, N, T0, rho, canon, bc)
        try:
            R.set_state(pos, vel)
            f, V, P, Ps, rdf = R.forces()
            R.set_state(pos, vel)
            R.forces()
            R.params()
            hist = R.integrate(DT, STEPS)
            p1, v1, _ = R.state()
            np.savez_compressed(os.path.join(HERE, case_name(N, rho, T0, bc, canon)),
                                pos=pos, vel=vel, force=f, V=V, P=P, Pshear=Ps, rdf=rdf,
                                hist=hist, pos_end=p1, vel_end=v1, L=L, steps=STEPS, dt=DT)
        finally:
            R.close()
        print("wrote", case_name(N, rho, T0, bc, canon))


if __name__ == "__main__":
    main()
