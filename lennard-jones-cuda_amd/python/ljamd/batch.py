"""Batched ensemble of independent boxes (lj_batch_* of include/lj_hip.h)."""
import ctypes as C

import numpy as np

from . import _lib as _L
from ._lib import check, lib, ptr


def ensemble_boxes(nboxes=512, seed=7, n_range=(200, 4000), rho_range=(0.05, 0.9),
                   t_range=(0.8, 2.0)):
    """The batched-ensemble workload: N ~ U{200..4000}, rho* ~ U[0.05, 0.9],
    T* ~ U[0.8, 2.0], seeded (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    N = rng.integers(n_range[0], n_range[1] + 1, nboxes).astype(np.int32)
    rho = rng.uniform(rho_range[0], rho_range[1], nboxes)
    T0 = rng.uniform(t_range[0], t_range[1], nboxes)
    return N, rho, T0


class LJBatch:
    def __init__(self, N, rho, T0, canonical=_L.EVN, boundary=_L.BC_PERIODIC, device=0,
                 seed=None):
        self.N = np.ascontiguousarray(N, np.int32)
        self.rho = np.ascontiguousarray(rho, np.float64)
        self.T0 = np.ascontiguousarray(T0, np.float64)
        self.nboxes = int(self.N.size)
        self.offsets = np.concatenate([[0], np.cumsum(self.N)[:-1]]).astype(np.int64)
        h = C.c_void_p()
        check(lib().lj_batch_create(self.nboxes, ptr(self.N), ptr(self.rho), ptr(self.T0),
                                    int(canonical), int(boundary), int(device), C.byref(h)))
        self.h = h
        if seed is not None:
            self.sample_initial_state(seed)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().lj_batch_destroy(self.h)
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

    @property
    def Ntot(self):
        return int(self.N.sum())

    def sample_initial_state(self, seed):
        check(lib().lj_batch_sample_initial_state(self.h, int(seed)))

    def set_array(self, kind, a4, box=-1):
        a4 = np.ascontiguousarray(a4, np.float32).reshape(-1)
        n = self.Ntot if box < 0 else int(self.N[box])
        if a4.size != 4 * n:
            raise ValueError("size mismatch")
        check(lib().lj_batch_set_array(self.h, int(box), int(kind), ptr(a4)))

    def get_array(self, kind, box=-1):
        n = self.Ntot if box < 0 else int(self.N[box])
        out = np.empty(4 * n, np.float32)
        check(lib().lj_batch_get_array(self.h, int(box), int(kind), ptr(out)))
        return out

    def compute_forces(self):
        check(lib().lj_batch_compute_forces(self.h))

    def integrate(self, dt, nsteps):
        check(lib().lj_batch_integrate(self.h, float(dt), int(nsteps)))

    def observables(self):
        arr = (_L.LJObs * self.nboxes)()
        check(lib().lj_batch_get_observables(self.h, C.cast(arr, C.c_void_p)))
        return [o.as_dict() for o in arr]

    def set_subvolume(self, fraction):
        check(lib().lj_batch_set_subvolume(self.h, float(fraction)))

    def stats(self):
        out = np.zeros((self.nboxes, _L.BATCH_NSTAT), np.float64)
        check(lib().lj_batch_get_stats(self.h, ptr(out)))
        return out

    def reset_stats(self):
        check(lib().lj_batch_reset_stats(self.h))

    def set_tuning(self, ipl, jchunk, blocks_per_cu, pair_once=0):
        t = _L.LJBatchTuning(int(ipl), int(jchunk), int(blocks_per_cu), int(pair_once))
        check(lib().lj_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        """(ipl, jchunk, blocks_per_cu, pair_once)"""
        t = _L.LJBatchTuning()
        check(lib().lj_batch_get_tuning(self.h, C.byref(t)))
        return (t.ipl, t.jchunk, t.blocks_per_cu, t.pair_once)

    def time_force(self, reps=5):
        us = C.c_double()
        check(lib().lj_batch_time_force(self.h, int(reps), C.byref(us)))
        return us.value

    def autotune(self, reps=3):
        t = _L.LJBatchTuning()
        us = C.c_double()
        check(lib().lj_batch_autotune(self.h, int(reps), C.byref(t), C.byref(us)))
        return (t.ipl, t.jchunk, t.blocks_per_cu, t.pair_once), us.value

    def integrate_hist(self, dt, nsteps):
        """nsteps batched steps; returns {U, T, K, V, P} [nsteps, nboxes, 5]."""
        h = np.zeros((max(int(nsteps), 1), self.nboxes, 5), np.float64)
        check(lib().lj_batch_integrate_hist(self.h, float(dt), int(nsteps), ptr(h)))
        return h[:nsteps]

    # ---- fluctuation counters of every box (lj_batch_fl_*) ----

    def fl_configure(self, ladders):
        from .system import _fl_ladders
        check(lib().lj_batch_fl_configure(self.h, _fl_ladders(ladders), len(ladders)))

    def fl_num_entries(self):
        e = C.c_int()
        check(lib().lj_batch_fl_num_entries(self.h, C.byref(e)))
        return e.value

    def fl_sample(self):
        check(lib().lj_batch_fl_sample(self.h))

    def fl_counts(self):
        out = np.zeros((self.nboxes, self.fl_num_entries()), np.int64)
        check(lib().lj_batch_fl_counts(self.h, ptr(out)))
        return out

    def fl_stats(self):
        out = np.zeros((self.nboxes, self.fl_num_entries(), _L.FL_NSTAT), np.float64)
        check(lib().lj_batch_fl_stats(self.h, ptr(out)))
        return out

    def fl_reset(self):
        check(lib().lj_batch_fl_reset(self.h))

    def set_step_sampling(self, fluctuations):
        check(lib().lj_batch_set_step_sampling(self.h, int(bool(fluctuations))))
