"""Context behaviour around the time step: contexts used from two host threads
at once, mode changes between differentiable trajectories, and device-to-device
state transfer into an FP64 context.  Each compares with a reference run of the
same library on the same state (single-threaded, a fresh context, a host-side
transfer); the kernels are deterministic, so the results are compared bit for
bit."""
import threading

import numpy as np
import pytest

import ljamd
from helpers import perturbed_state
from ljamd._hip import DeviceBuffer
from ljamd._lib import check, lib
from oracle_lib import lattice_state

pytestmark = pytest.mark.gpu


def _run(N, rho, canon, bc, seed, nsteps):
    L, pos, vel = lattice_state(N, rho, 1.0, seed)
    with ljamd.LJSystem(N, rho, 1.0, canonical=canon, boundary=bc) as s:
        s.set_state(pos, vel)
        s.CalculateForces()
        s.CalculateParameters()
        h = s.Integrate(0.005, nsteps)
        return h, s.pos, s.vel, s.rdf(ljamd.RDF_REF_CPU, 256)


def test_two_contexts_from_two_host_threads():
    """Two contexts (the sliced pass at N = 2048 TVN periodic, the pair-once
    pass at N = 6144 EVN hard wall) stepped concurrently from two host threads
    (ctypes releases the interpreter lock during each call) give the bits of
    the same runs made one after the other on one thread."""
    cases = [(2048, 0.85, ljamd.TVN, ljamd.BC_PERIODIC, 11, 300),
             (6144, 0.6, ljamd.EVN, ljamd.BC_WALL, 12, 200)]
    print("\n[case] two host threads, N = 2048 and 6144", flush=True)
    serial = [_run(*c) for c in cases]
    results = [None, None]
    errors = []

    def worker(k):
        try:
            results[k] = _run(*cases[k])
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for got, exp in zip(results, serial):
        for a, b in zip(got, exp):
            assert np.array_equal(a, b)


def test_trajectory_loss_after_mode_change():
    """A periodic context that computed a trajectory loss, switched to an open
    box by set_mode, gives the loss of a fresh open-box context on the same
    state: the captured FP64 step of the differentiable path is rebuilt for the
    new boundary."""
    N, rho, dt, M = 512, 0.3, 0.005, 30
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=8)
    with ljamd.LJSystem(N, rho, 1.0) as s:
        s.set_state(pos, vel)
        periodic = s.trajectory_loss(dt, M)
        s.set_mode(ljamd.EVN, ljamd.BC_NONE)
        switched = s.trajectory_loss(dt, M)
        g_switched = s.trajectory_grad(dt, M, 10)
    with ljamd.LJSystem(N, rho, 1.0, boundary=ljamd.BC_NONE) as f:
        f.set_state(pos, vel)
        fresh = f.trajectory_loss(dt, M)
        g_fresh = f.trajectory_grad(dt, M, 10)
    print(f"\n  loss periodic {periodic[0]:.12g}, after set_mode {switched[0]:.12g}, "
          f"fresh open box {fresh[0]:.12g}")
    assert periodic[0] != fresh[0]
    assert np.array_equal(switched, fresh)
    assert np.array_equal(g_switched[1], g_fresh[1])
    assert np.array_equal(g_switched[2], g_fresh[2])


def test_fp64_context_state_from_device_pointer():
    """lj_set_array_device into an LJ_PREC_FP64 context reaches the double
    state that its steps use: 10 steps after a device-side transfer equal 10
    steps after the same values set from the host."""
    N, rho = 1500, 0.6
    L, pos, vel = perturbed_state(N, rho, 1.0, seed=21)
    dp = DeviceBuffer.from_array(pos.reshape(N, 4))
    dv = DeviceBuffer.from_array(vel.reshape(N, 4))
    outs = []
    for via_device in (False, True):
        with ljamd.LJSystem(N, rho, 1.0, canonical=ljamd.TVN, precision=ljamd.PREC_FP64) as s:
            if via_device:
                check(lib().lj_set_array_device(s.h, ljamd.ARRAY_POS, dp.ptr, None))
                check(lib().lj_set_array_device(s.h, ljamd.ARRAY_VEL, dv.ptr, None))
            else:
                s.set_array(ljamd.ARRAY_POS, pos)
                s.set_array(ljamd.ARRAY_VEL, vel)
            s.CalculateForces()
            s.CalculateParameters()
            h = s.Integrate(0.005, 10)
            outs.append((h, s.get_array_f64(ljamd.ARRAY_POS)))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    # and the state did move from the transferred values
    assert np.max(np.abs(outs[1][1].reshape(N, 4)[:, :3] - pos.reshape(N, 4)[:, :3])) > 1e-4
