"""Speed histogram on the device (lj_velocity_histogram) against the oracle's
restatement of MDSystem::initvelo / updatevelo binning (MDSystem.cpp:651-694),
bit-exact, including speeds past the last bin and bin edges."""
import numpy as np
import pytest

import ljamd
from helpers import perturbed_state
from oracle_lib import velocity_histogram

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,T0,step,vmax", [(400, 1.4, 0.12, 12.0), (5000, 1.0, 0.05, 3.0),
                                           (16384, 2.0, 0.013, 2.0)])
def test_velocity_histogram_matches_oracle(N, T0, step, vmax):
    rho = 0.5
    L, pos, vel = perturbed_state(N, rho, T0, seed=17)
    v = vel.reshape(N, 4).copy()
    # a few speeds exactly on bin edges (float |v|^2 of an axis-aligned vector)
    v[:8, :3] = 0
    v[:8, 0] = (np.arange(8) * step).astype(np.float32)
    v = v.reshape(-1)
    nb = int(vmax / step) + 1
    with ljamd.LJSystem(N, rho, T0) as s:
        s.set_state(pos, v)
        got = s.velocity_histogram(step, nb)
    exp = velocity_histogram(v, step, nb)
    assert np.array_equal(got.astype(np.int64), exp)
    assert exp.sum() <= N


def test_velocity_histogram_errors():
    with ljamd.LJSystem(100, 0.3, 1.0) as s:
        for step, nb in ((0.0, 10), (-1.0, 10), (0.1, 0), (0.1, 20000)):
            with pytest.raises(ljamd.LJError):
                s.velocity_histogram(step, nb)
