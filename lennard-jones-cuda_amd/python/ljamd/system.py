"""Python view of one lj_ctx: the device-resident box of the C-ABI.

Method names follow the reference class MDSystem (src/library/MDSystem.h) so
that test code reads like the reference's own call sites.
"""
import ctypes as C

import numpy as np

from . import _lib as _L
from ._lib import check, lib, ptr


class LJSystem:
    """One box on the GPU: MDSystem's state and time step behind lj_hip.h."""

    def __init__(self, N, rho, T0=1.0, canonical=_L.EVN, boundary=_L.BC_PERIODIC, eps=1.0,
                 sigma=1.0, precision=_L.PREC_FP32, jslices=0, device=0, seed=None):
        cfg = _L.LJConfig()
        lib().lj_config_init(C.byref(cfg))
        cfg.N, cfg.rho, cfg.T0 = int(N), float(rho), float(T0)
        cfg.canonical, cfg.boundary = int(canonical), int(boundary)
        cfg.eps, cfg.sigma = float(eps), float(sigma)
        cfg.precision, cfg.jslices, cfg.device = int(precision), int(jslices), int(device)
        h = C.c_void_p()
        check(lib().lj_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.cfg = cfg
        self.N = int(N)
        if seed is not None:
            self.sample_initial_state(seed)

    # lifetime
    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().lj_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # state
    @property
    def L(self):
        return lib().lj_box_length(self.h)

    def sample_initial_state(self, seed, velocities=None):
        """Lattice positions and seeded velocities: the Gaussian draw of the
        benchmark configurations (default) or INIT_MAXWELL, the reference's
        own Maxwell-speed sampler."""
        if velocities is None:
            check(lib().lj_sample_initial_state(self.h, int(seed)))
        else:
            check(lib().lj_sample_initial_state_ex(self.h, int(seed), int(velocities)))

    def set_array(self, kind, a4):
        a4 = np.ascontiguousarray(a4, np.float32).reshape(-1)
        if a4.size != 4 * self.N:
            raise ValueError("expected float32[4N]")
        check(lib().lj_set_array(self.h, int(kind), ptr(a4)))

    def get_array(self, kind):
        out = np.empty(4 * self.N, np.float32)
        check(lib().lj_get_array(self.h, int(kind), ptr(out)))
        return out

    def set_array_f64(self, kind, a4):
        a4 = np.ascontiguousarray(a4, np.float64).reshape(-1)
        if a4.size != 4 * self.N:
            raise ValueError("expected float64[4N]")
        check(lib().lj_set_array_f64(self.h, int(kind), ptr(a4)))

    def get_array_f64(self, kind):
        out = np.empty(4 * self.N, np.float64)
        check(lib().lj_get_array_f64(self.h, int(kind), ptr(out)))
        return out

    def set_state(self, pos4, vel4=None, force4=None):
        self.set_array(_L.ARRAY_POS, pos4)
        if vel4 is not None:
            self.set_array(_L.ARRAY_VEL, vel4)
        if force4 is not None:
            self.set_array(_L.ARRAY_FORCE, force4)

    @property
    def pos(self):
        return self.get_array(_L.ARRAY_POS)

    @property
    def vel(self):
        return self.get_array(_L.ARRAY_VEL)

    @property
    def force(self):
        return self.get_array(_L.ARRAY_FORCE)

    # MDSystem methods
    def CalculateForces(self):
        o = _L.LJObs()
        check(lib().lj_compute_forces(self.h, C.byref(o)))
        return o.as_dict()

    def CalculateParameters(self):
        o = _L.LJObs()
        check(lib().lj_compute_parameters(self.h, C.byref(o)))
        return o.as_dict()

    def observables(self):
        o = _L.LJObs()
        check(lib().lj_get_observables(self.h, C.byref(o)))
        return o.as_dict()

    def Integrate(self, dt, nsteps=1):
        """nsteps of MDSystem::Integrate; returns per-step {U,T,K,V,P} [nsteps, 5]."""
        hist = np.zeros((max(nsteps, 1), 5), np.float64)
        check(lib().lj_integrate(self.h, float(dt), int(nsteps), ptr(hist)))
        return hist[:nsteps]

    def integrate_async(self, dt, nsteps=1):
        check(lib().lj_integrate_async(self.h, float(dt), int(nsteps)))

    def rdf(self, rule=_L.RDF_REF_CPU, nbins=256, dr2=0.0, rmax=0.0, accumulate=False):
        counts = np.zeros(nbins, np.uint64)
        check(lib().lj_rdf(self.h, int(rule), int(nbins), float(dr2), float(rmax), ptr(counts),
                           int(bool(accumulate))))
        return counts

    def RenormalizeVelocities(self, recompute_tkin=False):
        check(lib().lj_renormalize_velocities(self.h, int(bool(recompute_tkin))))

    def RenormalizeVelocitiesToEnergy(self, ust):
        check(lib().lj_renormalize_to_energy(self.h, float(ust)))

    def resetAveraging(self):
        check(lib().lj_reset_averaging(self.h))

    def ApplyBoundaryConditions(self):
        check(lib().lj_apply_boundary_conditions(self.h))

    def set_mode(self, canonical, boundary):
        check(lib().lj_set_mode(self.h, int(canonical), int(boundary)))

    def synchronize(self):
        check(lib().lj_synchronize(self.h))

    def device_arrays(self):
        p, v, f = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().lj_device_arrays(self.h, C.byref(p), C.byref(v), C.byref(f)))
        return p.value, v.value, f.value

    def time_rdf(self, rule=_L.RDF_REF_GPU, nbins=256, dr2=0.0, rmax=0.0, reps=20):
        """Average microseconds of one RDF pass (lj_time_rdf)."""
        us = C.c_double()
        check(lib().lj_time_rdf(self.h, int(rule), int(nbins), float(dr2), float(rmax), int(reps),
                                C.byref(us)))
        return us.value

    def time_force_kernel(self, reps=20):
        us = C.c_double()
        check(lib().lj_time_force_kernel(self.h, int(reps), C.byref(us)))
        return us.value

    def time_force_kernel_stamped(self, reps=200):
        """(mean of the launches' own durations, back-to-back time per launch),
        microseconds, over `reps` force-pass launches (FP32 mode)."""
        k, s = C.c_double(), C.c_double()
        check(lib().lj_time_force_kernel_stamped(self.h, int(reps), C.byref(k), C.byref(s)))
        return k.value, s.value

    def velocity_histogram(self, step, nbins):
        """Speed histogram with the binning of MDSystem::initvelo (uint64[nbins])."""
        out = np.zeros(int(nbins), np.uint64)
        check(lib().lj_velocity_histogram(self.h, float(step), int(nbins), ptr(out)))
        return out

    def forces_shard(self, shard, nshards, jslices=0):
        """One GPU's share of a force pass sharded by i over nshards GPUs (at
        most ceil(N/64)): the forces on its rows, whole 64-particle tiles
        [64*(shard*T//nshards), 64*((shard+1)*T//nshards)) with T = ceil(N/64),
        go to the context's force array; returns that share's (E_pot, W)
        partial sums."""
        ev = np.zeros(2, np.float64)
        check(lib().lj_forces_shard(self.h, int(shard), int(nshards), int(jslices), None, ptr(ev)))
        return float(ev[0]), float(ev[1])

    def time_forces_shard(self, shard, nshards, jslices=0, reps=50):
        """Microseconds per launch of one share's force kernel."""
        us = C.c_double()
        check(lib().lj_time_forces_shard(self.h, int(shard), int(nshards), int(jslices), int(reps),
                                         C.byref(us)))
        return us.value

    def time_force_items_share(self, shard, nshards, reps=200):
        """Microseconds per launch of one share of the pair-once step pass with
        its items dealt out over nshards GPUs."""
        us = C.c_double()
        check(lib().lj_time_force_items_share(self.h, int(shard), int(nshards), int(reps),
                                              C.byref(us)))
        return us.value

    def shear_pressure(self):
        """Pshear of the current state as the reference CPU path computes it
        (MDSystem.cpp:299,309,335,353); one all-pairs pass."""
        out = C.c_double()
        check(lib().lj_shear_pressure(self.h, C.byref(out)))
        return out.value

    def force_plan(self):
        """(pair_once, nslots) of the force kernel this context runs."""
        p, n = C.c_int(), C.c_int()
        check(lib().lj_force_plan_info(self.h, C.byref(p), C.byref(n)))
        return bool(p.value), n.value

    def forces_device(self, d_pos, d_force, d_ev=None, stream=None):
        check(lib().lj_forces_device(self.h, C.c_void_p(d_pos), C.c_void_p(d_force),
                                     C.c_void_p(d_ev) if d_ev else None,
                                     C.c_void_p(stream) if stream else None))


    # ---- fluctuation counters (lj_fl_*) and sampling inside the step ----

    def fl_configure(self, ladders):
        """ladders: (type, single, alpha, vmax) tuples (LJ_FL_* types)."""
        arr = _fl_ladders(ladders)
        check(lib().lj_fl_configure(self.h, arr, len(ladders)))

    def fl_num_entries(self):
        e = C.c_int()
        check(lib().lj_fl_num_entries(self.h, C.byref(e)))
        return e.value

    def fl_values(self):
        out = np.zeros(self.fl_num_entries(), np.float64)
        check(lib().lj_fl_values(self.h, ptr(out)))
        return out

    def fl_sample(self):
        check(lib().lj_fl_sample(self.h))

    def fl_counts(self):
        out = np.zeros(self.fl_num_entries(), np.int64)
        check(lib().lj_fl_counts(self.h, ptr(out)))
        return out

    def fl_stats(self):
        out = np.zeros((self.fl_num_entries(), _L.FL_NSTAT), np.float64)
        check(lib().lj_fl_stats(self.h, ptr(out)))
        return out

    def fl_reset(self):
        check(lib().lj_fl_reset(self.h))

    def set_step_sampling(self, rdf=_L.SAMPLE_RDF_OFF, rdf_rule=_L.RDF_REF_CPU, rdf_nbins=256,
                          rdf_dr2=0.0, rdf_rmax=0.0, fluctuations=False):
        smp = _L.LJStepSampling(int(rdf), int(rdf_rule), int(rdf_nbins), float(rdf_dr2),
                                float(rdf_rmax), int(bool(fluctuations)))
        self._rdf_nbins = int(rdf_nbins)
        check(lib().lj_set_step_sampling(self.h, C.byref(smp)))

    def step_rdf_counts(self, clear=False):
        """(uint64[nbins] step-sampled RDF counts, steps they hold)."""
        out = np.zeros(getattr(self, "_rdf_nbins", 256), np.uint64)
        n = C.c_int64()
        check(lib().lj_step_rdf_counts(self.h, ptr(out), C.byref(n), int(bool(clear))))
        return out, n.value


def _fl_ladders(ladders):
    arr = (_L.LJFlLadder * max(len(ladders), 1))()
    for k, (typ, single, alpha, vmax) in enumerate(ladders):
        arr[k] = _L.LJFlLadder(int(typ), int(single), float(alpha), float(vmax))
    return arr


# ---- the compatibility set (reference MDSystem.cu:155-299) ----

class CompatArrays:
    """Device arrays obtained through allocateArray / copyArrayToDevice."""

    def __init__(self, N):
        self.N = N
        self.d_pos = C.c_void_p()
        self.d_force = C.c_void_p()
        lib().allocateArray(C.byref(self.d_pos), N)
        lib().allocateArray(C.byref(self.d_force), N)
        if not self.d_pos.value or not self.d_force.value:
            raise _L.LJError(_L.LJ_ERR_ALLOC, lib().lj_last_error().decode())

    def close(self):
        if self.d_pos.value:
            lib().deleteArray(self.d_pos)
        if self.d_force.value:
            lib().deleteArray(self.d_force)
        self.d_pos = C.c_void_p()
        self.d_force = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def calculate(self, pos4, L, periodic, dr2, block=256):
        """copyArrayToDevice + calculateNForces + copyArrayFromDevice, as
        MDSystem::CalculateForces' GPU branch (MDSystem.cpp:242-250)."""
        pos4 = np.ascontiguousarray(pos4, np.float32)
        lib().copyArrayToDevice(self.d_pos, ptr(pos4), self.N)
        pressure = C.c_float(0.0)
        rdf = np.zeros(256, np.int32)
        lib().calculateNForces(self.d_pos, self.d_force, C.byref(pressure), self.N, float(L),
                               int(periodic), ptr(rdf), float(dr2), int(block), 1)
        frc = np.zeros(4 * self.N, np.float32)
        lib().copyArrayFromDevice(ptr(frc), self.d_force, 0, self.N)
        return frc, pressure.value, rdf


# ---- differentiable path (lj_trajectory_*, lj_adjoint_pass_device) ----

def _set_lj_params(self, eps, sigma):
    check(lib().lj_set_lj_params(self.h, float(eps), float(sigma)))


def _trajectory_loss(self, dt, nsteps, p_target=1.0, w_p=1.0):
    """(loss, mean E_pot, mean P) of nsteps EVN steps from the current state."""
    out = np.zeros(3)
    check(lib().lj_trajectory_loss(self.h, float(dt), int(nsteps), float(p_target), float(w_p),
                                   ptr(out)))
    return out


def _trajectory_grad(self, dt, nsteps, checkpoint_every=20, p_target=1.0, w_p=1.0, d_dx0=None):
    """loss3, (dL/deps, dL/dsigma), dL/dx0 float32[N,4], times {fwd, recompute, reverse} ms."""
    out, g2, times = np.zeros(3), np.zeros(2), np.zeros(3)
    dx0 = np.zeros(4 * self.N, np.float32)
    check(lib().lj_trajectory_grad(self.h, float(dt), int(nsteps), int(checkpoint_every),
                                   float(p_target), float(w_p), ptr(out), ptr(g2), ptr(dx0),
                                   C.c_void_p(d_dx0) if d_dx0 else None, ptr(times)))
    return out, g2, dx0.reshape(self.N, 4), times


def _adjoint_pass(self, pos4, z4):
    """(H z, F, dW/dx) as float64 [N,4] and the 8 scalars, at host positions pos4."""
    from ._hip import DeviceBuffer
    N = self.N
    dp = DeviceBuffer.from_array(np.ascontiguousarray(pos4, np.float32).reshape(N, 4))
    dz = DeviceBuffer.from_array(np.ascontiguousarray(z4, np.float32).reshape(N, 4))
    bufs = [DeviceBuffer(32 * N) for _ in range(3)]
    sc = DeviceBuffer(64)
    check(lib().lj_adjoint_pass_device(self.h, dp.ptr, dz.ptr, bufs[0].ptr, bufs[1].ptr,
                                       bufs[2].ptr, sc.ptr, None))
    hz, f, gw = (b.download(np.float64, (N, 4)) for b in bufs)
    return hz, f, gw, sc.download(np.float64, (8,))


LJSystem.set_lj_params = _set_lj_params
LJSystem.trajectory_loss = _trajectory_loss
LJSystem.trajectory_grad = _trajectory_grad
LJSystem.adjoint_pass = _adjoint_pass
