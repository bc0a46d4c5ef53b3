"""ctypes bindings to the test-oracle libraries under oracle/ (test infrastructure).

liboracle.so is the C restatement of the reference CPU path; oracle/_ref holds
the reference's own MDSystem class compiled from its sources (when it was
buildable).  Nothing in the product imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")

RDF_REF_CPU, RDF_REF_GPU, RDF_RBINS = 0, 1, 2


class OrcSys(C.Structure):
    _fields_ = [("N", C.c_int), ("T0", C.c_double), ("rho", C.c_double),
                ("canonical", C.c_int), ("bc", C.c_int), ("L", C.c_double),
                ("rdf_dr2", C.c_float), ("pos", C.POINTER(C.c_float)),
                ("vel", C.POINTER(C.c_float)), ("frc", C.POINTER(C.c_float)),
                ("rdf", C.c_int * 256), ("U", C.c_double), ("T", C.c_double),
                ("K", C.c_double), ("V", C.c_double), ("P", C.c_double),
                ("Pshear", C.c_double), ("t", C.c_double),
                ("av_U_tot", C.c_double), ("av_T_tot", C.c_double),
                ("av_p_tot", C.c_double), ("av_iters", C.c_int),
                ("nthreads", C.c_int)]


def _load_oracle():
    lib = C.CDLL(os.path.join(ORACLE_DIR, "liboracle.so"))
    lib.orc_lattice_n.restype = C.c_int
    lib.orc_lattice_n.argtypes = [C.c_int]
    lib.orc_box_length.restype = C.c_double
    lib.orc_box_length.argtypes = [C.c_int, C.c_double]
    lib.orc_rdf_dr2.restype = C.c_float
    lib.orc_rdf_dr2.argtypes = [C.c_int]
    lib.orc_lattice_positions.argtypes = [C.c_int, C.c_double, _f32p]
    lib.orc_gaussian_velocities.argtypes = [C.c_int, C.c_double, C.c_uint64, _f32p]
    lib.orc_splitmix_uniform.restype = C.c_double
    lib.orc_splitmix_uniform.argtypes = [C.POINTER(C.c_uint64)]
    lib.orc_correct_momentum.argtypes = [C.c_int, _f32p]
    lib.orc_kinetic_temperature.restype = C.c_double
    lib.orc_kinetic_temperature.argtypes = [C.c_int, _f32p]
    lib.orc_fast_round.restype = C.c_int
    lib.orc_fast_round.argtypes = [C.c_float]
    lib.orc_forces_ref.argtypes = [C.c_int, _f32p, C.c_double, C.c_int, C.c_float, _f32p,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), _i32p]
    lib.orc_forces_ref_omp.argtypes = lib.orc_forces_ref.argtypes + [C.c_int]
    lib.orc_forces_fp64.argtypes = [C.c_int, _f32p, C.c_double, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    lib.orc_forces_fp64d.argtypes = [C.c_int, _f64p, C.c_double, C.c_int, C.c_double,
                                     C.c_double, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    lib.orc_rdf.argtypes = [C.c_int, _f32p, C.c_double, C.c_int, C.c_int, C.c_int,
                            C.c_float, C.c_double, _i64p, C.c_int]
    lib.orc_rdf_edges2.argtypes = [C.c_int, C.c_double, _f32p]
    lib.orc_velocity_histogram.argtypes = [C.c_int, _f32p, C.c_double, C.c_int, _i64p]
    lib.orc_sys_init.restype = C.c_int
    lib.orc_sys_init.argtypes = [C.POINTER(OrcSys), C.c_int, C.c_double, C.c_double,
                                 C.c_int, C.c_int]
    lib.orc_sys_free.argtypes = [C.POINTER(OrcSys)]
    for n in ("orc_sys_forces", "orc_sys_params", "orc_sys_apply_bc"):
        getattr(lib, n).argtypes = [C.POINTER(OrcSys)]
    lib.orc_sys_integrate.argtypes = [C.POINTER(OrcSys), C.c_double]
    lib.orc_fp64_step.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_int, C.c_int,
                                  C.c_double, C.c_double, C.c_double, C.c_double, _f64p, C.c_int]
    lib.orc_adjoint_pass_fp64.argtypes = [C.c_int, _f64p, _f64p, C.c_double, C.c_int, C.c_double,
                                          C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_int]
    lib.orc_traj_loss_fp64.argtypes = [C.c_int, _f64p, _f64p, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                       _f64p, C.c_int]
    lib.orc_traj_grad_fp64.argtypes = [C.c_int, _f64p, _f64p, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                       _f64p, _f64p, C.c_void_p, C.c_int]
    _i32p_ = _i32p
    lib.orc_fl_ladder.restype = C.c_int
    lib.orc_fl_ladder.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_int]
    lib.orc_fl_nsub_batch.restype = C.c_int
    lib.orc_fl_nsub_batch.argtypes = [C.c_int, _f32p, C.c_double, C.c_double, C.c_int, _i32p_,
                                      C.c_int]
    lib.orc_fl_vz_batch.restype = C.c_int
    lib.orc_fl_vz_batch.argtypes = [C.c_int, _f32p, C.c_double, C.c_double, C.c_int, _i32p_,
                                    C.c_int]
    lib.orc_fl_nsub.restype = C.c_int
    lib.orc_fl_nsub.argtypes = [C.c_int, _f32p, C.c_double, C.c_double, C.c_int]
    lib.orc_fl_vz.restype = C.c_int
    lib.orc_fl_vz.argtypes = [C.c_int, _f32p, C.c_double, C.c_int]
    lib.orc_fl_closed_cube.restype = C.c_longlong
    lib.orc_fl_closed_cube.argtypes = [C.c_int, _f32p, C.c_double, C.c_double]
    lib.orc_tavg_init.argtypes = [C.c_void_p]
    lib.orc_tavg_add.argtypes = [C.c_void_p, C.c_double]
    lib.orc_tavg_eval.argtypes = [C.c_void_p, _f64p]
    return lib


ORC = _load_oracle()


class OrcTavg(C.Structure):
    _fields_ = [("sums", C.c_double * 5), ("nobs", C.c_int64), ("last", C.c_double),
                ("corr", C.c_double), ("iters", C.c_int)]


def fl_ladder(step, limit):
    """Alpha ladder of the reference's batch counters (accumulated steps)."""
    n = ORC.orc_fl_ladder(step, limit, None, 0)
    a = np.zeros(n, np.float64)
    ORC.orc_fl_ladder(step, limit, a.ctypes.data, n)
    return a


def fl_nsub_batch(pos, L, step, typ):
    cnt = np.zeros(4096, np.int32)
    m = ORC.orc_fl_nsub_batch(pos.size // 4, np.ascontiguousarray(pos, np.float32), L, step, typ,
                              cnt, cnt.size)
    return cnt[:m].copy()


def fl_vz_batch(vel, vmax, step, typ):
    cnt = np.zeros(4096, np.int32)
    m = ORC.orc_fl_vz_batch(vel.size // 4, np.ascontiguousarray(vel, np.float32), vmax, step, typ,
                            cnt, cnt.size)
    return cnt[:m].copy()


def tavg_eval(values):
    """TimeAverage statistics of a sequence (mean, variance, s, mean error,
    scaled variance, scaled variance error, plain mean error)."""
    a = OrcTavg()
    ORC.orc_tavg_init(C.byref(a))
    for v in values:
        ORC.orc_tavg_add(C.byref(a), float(v))
    out = np.zeros(7, np.float64)
    ORC.orc_tavg_eval(C.byref(a), out)
    return out


def lattice_state(N, rho, T0, seed):
    """Config initial state: sc lattice + seeded Gaussian velocities (float4 AoS)."""
    L = ORC.orc_box_length(N, rho)
    pos = np.zeros(4 * N, np.float32)
    vel = np.zeros(4 * N, np.float32)
    ORC.orc_lattice_positions(N, L, pos)
    ORC.orc_gaussian_velocities(N, T0, seed, vel)
    return L, pos, vel


def forces_ref(pos, L, bc, dr2, nthreads=1):
    N = pos.size // 4
    frc = np.zeros(4 * N, np.float32)
    V, P, Ps = C.c_double(), C.c_double(), C.c_double()
    rdf = np.zeros(256, np.int32)
    if nthreads > 1:
        ORC.orc_forces_ref_omp(N, pos, L, bc, dr2, frc, C.byref(V), C.byref(P), C.byref(Ps),
                               rdf, nthreads)
    else:
        ORC.orc_forces_ref(N, pos, L, bc, dr2, frc, C.byref(V), C.byref(P), C.byref(Ps), rdf)
    return frc, V.value, P.value, Ps.value, rdf


def forces_fp64(pos, L, periodic, eps=1.0, sigma=1.0, skip_r2=0.0, nthreads=0):
    """FP64 oracle: (f[N,3], ew[N,2] = (Force.w convention, per-i virial), E, W)."""
    N = pos.size // 4
    f = np.zeros(3 * N, np.float64)
    ew = np.zeros(2 * N, np.float64)
    E, W = C.c_double(), C.c_double()
    ORC.orc_forces_fp64(N, pos, L, int(periodic), eps, sigma, skip_r2,
                        f.ctypes.data, ew.ctypes.data, C.byref(E), C.byref(W), nthreads)
    return f.reshape(N, 3), ew.reshape(N, 2), E.value, W.value


def forces_fp64d(x3, L, periodic, eps=1.0, sigma=1.0, nthreads=0):
    N = x3.size // 3
    f = np.zeros(3 * N, np.float64)
    ew = np.zeros(2 * N, np.float64)
    E, W = C.c_double(), C.c_double()
    ORC.orc_forces_fp64d(N, np.ascontiguousarray(x3, np.float64).ravel(), L, int(periodic),
                         eps, sigma, f.ctypes.data, ew.ctypes.data, C.byref(E), C.byref(W),
                         nthreads)
    return f.reshape(N, 3), ew.reshape(N, 2), E.value, W.value


def rdf(pos, L, periodic, rule, nbins, dr2=0.1, rmax=0.0, nthreads=0):
    counts = np.zeros(nbins, np.int64)
    ORC.orc_rdf(pos.size // 4, pos, L, int(periodic), rule, nbins, dr2, rmax, counts, nthreads)
    return counts


def velocity_histogram(vel, step, nbins):
    counts = np.zeros(nbins, np.int64)
    ORC.orc_velocity_histogram(vel.size // 4, np.ascontiguousarray(vel, np.float32), step, nbins,
                               counts)
    return counts


def rdf_edges2(nbins, rmax):
    e = np.zeros(nbins + 1, np.float32)
    ORC.orc_rdf_edges2(nbins, rmax, e)
    return e


class OracleSystem:
    """orc_sys: the reference CPU MDSystem restated (MDSystem.cpp:66-583)."""

    def __init__(self, N, T0, rho, canonical, bc, nthreads=1):
        self.s = OrcSys()
        rc = ORC.orc_sys_init(C.byref(self.s), N, T0, rho, int(canonical), bc)
        if rc:
            raise ValueError(f"orc_sys_init failed ({rc})")
        self.s.nthreads = nthreads
        self.N = N

    def __del__(self):
        try:
            ORC.orc_sys_free(C.byref(self.s))
        except Exception:
            pass

    def _arr(self, p):
        return np.ctypeslib.as_array(p, shape=(4 * self.N,))

    @property
    def pos(self):
        return self._arr(self.s.pos)

    @property
    def vel(self):
        return self._arr(self.s.vel)

    @property
    def frc(self):
        return self._arr(self.s.frc)

    def set_state(self, pos, vel):
        self.pos[:] = pos
        self.vel[:] = vel
        self.s.t = 0.0

    def forces(self):
        ORC.orc_sys_forces(C.byref(self.s))

    def params(self):
        ORC.orc_sys_params(C.byref(self.s))

    def integrate(self, dt, n=1):
        hist = []
        for _ in range(n):
            ORC.orc_sys_integrate(C.byref(self.s), dt)
            hist.append((self.s.U, self.s.T, self.s.K, self.s.V, self.s.P))
        return np.array(hist)


def ref_lib(kind="cpu"):
    """The reference's own MDSystem compiled from source (oracle/_ref), or None."""
    path = os.path.join(ORACLE_DIR, "_ref", f"libref_{kind}.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.refh_create.restype = C.c_void_p
    lib.refh_create.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                C.c_int]
    lib.refh_destroy.argtypes = [C.c_void_p]
    lib.refh_set_state.argtypes = [C.c_void_p, _f32p, _f32p]
    lib.refh_set_force.argtypes = [C.c_void_p, _f32p]
    lib.refh_forces.argtypes = [C.c_void_p, _f32p, _f64p, _i32p]
    lib.refh_params.argtypes = [C.c_void_p]
    lib.refh_observables.argtypes = [C.c_void_p, _f64p]
    lib.refh_integrate.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_void_p]
    lib.refh_get_state.argtypes = [C.c_void_p, _f32p, _f32p, _f32p]
    lib.refh_L.restype = C.c_double
    lib.refh_L.argtypes = [C.c_void_p]
    lib.refh_dr2.restype = C.c_float
    lib.refh_dr2.argtypes = [C.c_void_p]
    if hasattr(lib, "refh_fl_nsub_batch"):
        for n in ("refh_fl_nsub_batch", "refh_fl_vz_batch"):
            getattr(lib, n).restype = C.c_int
        lib.refh_fl_nsub_batch.argtypes = [C.c_void_p, C.c_double, C.c_int, _i32p, C.c_int]
        lib.refh_fl_vz_batch.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _i32p,
                                         C.c_int]
        lib.refh_fl_nsub.restype = C.c_int
        lib.refh_fl_nsub.argtypes = [C.c_void_p, C.c_double, C.c_int]
        lib.refh_fl_vz.restype = C.c_int
        lib.refh_fl_vz.argtypes = [C.c_void_p, C.c_double, C.c_int]
        lib.refh_fl_fluctuations.argtypes = [C.c_void_p, C.c_double, _f64p]
        lib.refh_tavg.argtypes = [_f64p, C.c_int, _f64p]
    lib.refh_fast_round.restype = C.c_int
    lib.refh_fast_round.argtypes = [C.c_void_p, C.c_float]
    lib.refh_rdf_normalised.argtypes = [C.c_void_p, _f64p, _f64p]
    lib.refh_velo.restype = C.c_int
    lib.refh_velo.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _f64p, _f64p, C.c_int]
    lib.refh_maxwell.restype = C.c_double
    lib.refh_maxwell.argtypes = [C.c_void_p, C.c_double]
    return lib


class RefSystem:
    """The reference MDSystem object behind the harness (oracle/ref_harness.cpp)."""

    def __init__(self, lib, N, T0, rho, canonical, bc, use_gpu=False, block=256):
        self.lib = lib
        self.N = N
        self.h = lib.refh_create(N, T0, rho, int(canonical), bc, int(use_gpu), block)

    def close(self):
        if self.h:
            self.lib.refh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def L(self):
        return self.lib.refh_L(self.h)

    @property
    def dr2(self):
        return self.lib.refh_dr2(self.h)

    def set_state(self, pos, vel, frc=None):
        self.lib.refh_set_state(self.h, pos, vel)
        if frc is not None:
            self.lib.refh_set_force(self.h, frc)

    def forces(self):
        frc = np.zeros(4 * self.N, np.float32)
        out = np.zeros(3, np.float64)
        rdf = np.zeros(256, np.int32)
        self.lib.refh_forces(self.h, frc, out, rdf)
        return frc, out[0], out[1], out[2], rdf

    def params(self):
        self.lib.refh_params(self.h)

    def observables(self):
        out = np.zeros(8, np.float64)
        self.lib.refh_observables(self.h, out)
        return dict(zip(("U", "T", "K", "V", "P", "Pshear", "t", "L"), out))

    def integrate(self, dt, n):
        hist = np.zeros((n, 5), np.float64)
        self.lib.refh_integrate(self.h, dt, n, hist.ctypes.data)
        return hist

    def velo(self, update=False, vmax=12.0, step=0.12):
        """initvelo(vmax, step) or updatevelo(): the curve (x, y) (MDSystem.cpp:651-694)."""
        x = np.zeros(100000, np.float64)
        y = np.zeros(100000, np.float64)
        n = self.lib.refh_velo(self.h, int(update), vmax, step, x, y, x.size)
        return x[:n].copy(), y[:n].copy()

    def state(self):
        p = np.zeros(4 * self.N, np.float32)
        v = np.zeros(4 * self.N, np.float32)
        f = np.zeros(4 * self.N, np.float32)
        self.lib.refh_get_state(self.h, p, v, f)
        return p, v, f


def adjoint_pass_fp64(x3, z3, L, periodic, eps=1.0, sigma=1.0, nthreads=0):
    """FP64 (H z, F, dW/dx, scalars{E, W, dE/ds, dW/ds, (dF/ds)^T z, F^T z})."""
    x3 = np.ascontiguousarray(x3, np.float64).ravel()
    z3 = np.ascontiguousarray(z3, np.float64).ravel()
    N = x3.size // 3
    hz, f, gw, sc = (np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N), np.zeros(8))
    ORC.orc_adjoint_pass_fp64(N, x3, z3, L, int(periodic), eps, sigma, hz.ctypes.data,
                              f.ctypes.data, gw.ctypes.data, sc.ctypes.data, nthreads)
    return hz.reshape(N, 3), f.reshape(N, 3), gw.reshape(N, 3), sc


def traj_loss_fp64(x0, v0, L, periodic, eps, sigma, dt, M, pt=1.0, wp=1.0, nthreads=0):
    x0 = np.ascontiguousarray(x0, np.float64).ravel()
    v0 = np.ascontiguousarray(v0, np.float64).ravel()
    out = np.zeros(3)
    ORC.orc_traj_loss_fp64(x0.size // 3, x0, v0, L, int(periodic), eps, sigma, dt, M, pt, wp,
                           out, nthreads)
    return out


def traj_grad_fp64(x0, v0, L, periodic, eps, sigma, dt, M, pt=1.0, wp=1.0, nthreads=0):
    x0 = np.ascontiguousarray(x0, np.float64).ravel()
    v0 = np.ascontiguousarray(v0, np.float64).ravel()
    N = x0.size // 3
    out, g2, dx0 = np.zeros(3), np.zeros(2), np.zeros(3 * N)
    ORC.orc_traj_grad_fp64(N, x0, v0, L, int(periodic), eps, sigma, dt, M, pt, wp, out, g2,
                           dx0.ctypes.data, nthreads)
    return out, g2, dx0.reshape(N, 3)


def fp64_steps(x3, v3, L, bc, canonical, T0, dt, nsteps, eps=1.0, sigma=1.0, nthreads=0):
    """nsteps FP64 Integrate(dt) steps from double state (x3, v3) [N,3].
    Returns (x, v, f, obs[nsteps, 3] = (E_pot, E_kin, W) after each step)."""
    x = np.ascontiguousarray(x3, np.float64).copy()
    v = np.ascontiguousarray(v3, np.float64).copy()
    N = x.shape[0]
    f = forces_fp64d(x, L, bc == 0, eps, sigma, nthreads)[0].copy()
    obs = np.zeros((nsteps, 3))
    o = np.zeros(3)
    for k in range(nsteps):
        ORC.orc_fp64_step(N, x.ravel(), v.ravel(), f.ravel(), L, int(bc), int(canonical), T0, eps,
                          sigma, dt, o, nthreads)
        obs[k] = o
    return x, v, f, obs
