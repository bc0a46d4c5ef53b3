"""ctypes declarations of libljhip.so (include/lj_hip.h).

The library is loaded from the package tree (lennard-jones-cuda_amd/lib).  A
missing library raises ImportError: there is no CPU fallback on this path.
"""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_PKG, "lib", "libljhip.so")

LJ_OK = 0
LJ_ERR_ARG, LJ_ERR_DEVICE, LJ_ERR_ALLOC, LJ_ERR_STATE, LJ_ERR_UNSUPPORTED = -1, -2, -3, -4, -5
BC_PERIODIC, BC_WALL, BC_NONE = 0, 1, 2
EVN, TVN = 0, 1
INIT_GAUSSIAN, INIT_MAXWELL = 0, 1
RDF_REF_CPU, RDF_REF_GPU, RDF_RBINS = 0, 1, 2
PREC_FP32, PREC_COMPENSATED, PREC_FP64 = 0, 1, 2
ARRAY_POS, ARRAY_VEL, ARRAY_FORCE = 0, 1, 2


class LJConfig(C.Structure):
    _fields_ = [("N", C.c_int), ("T0", C.c_double), ("rho", C.c_double),
                ("canonical", C.c_int), ("boundary", C.c_int), ("eps", C.c_double),
                ("sigma", C.c_double), ("precision", C.c_int), ("jslices", C.c_int),
                ("device", C.c_int)]


class LJObs(C.Structure):
    _fields_ = [("U", C.c_double), ("T", C.c_double), ("K", C.c_double), ("V", C.c_double),
                ("P", C.c_double), ("Pshear", C.c_double), ("t", C.c_double), ("W", C.c_double),
                ("av_U_tot", C.c_double), ("av_T_tot", C.c_double), ("av_p_tot", C.c_double),
                ("av_iters", C.c_int64), ("step", C.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class LJBatchTuning(C.Structure):
    _fields_ = [("ipl", C.c_int), ("jchunk", C.c_int), ("blocks_per_cu", C.c_int),
                ("pair_once", C.c_int)]


FL_SLAB_X, FL_SLAB_Y, FL_SLAB_Z, FL_CUBE, FL_VEL_X, FL_VEL_Y, FL_VEL_Z, FL_CUBE_CLOSED = range(8)
FL_NSTAT = 8
SAMPLE_RDF_OFF, SAMPLE_RDF_LAST, SAMPLE_RDF_SUM = 0, 1, 2


class LJFlLadder(C.Structure):
    _fields_ = [("type", C.c_int), ("single", C.c_int), ("alpha", C.c_double),
                ("vmax", C.c_double)]


class LJStepSampling(C.Structure):
    _fields_ = [("rdf", C.c_int), ("rdf_rule", C.c_int), ("rdf_nbins", C.c_int),
                ("rdf_dr2", C.c_float), ("rdf_rmax", C.c_double), ("fluctuations", C.c_int)]


BATCH_NSTAT = 13
BATCH_STAT_NAMES = ("samples", "U", "U2", "V", "V2", "K", "K2", "P", "P2", "n1", "n2", "n3", "n4")


class LJError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ljhip error {code}: {msg}")
        self.code = code


_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_f = C.c_float

# (name, restype, argtypes) of every C-ABI entry point
SIGNATURES = [
    ("lj_last_error", C.c_char_p, []),
    ("lj_device_count", _i, []),
    ("lj_version", C.c_char_p, []),
    ("allocateNBodyArrays", None, [_vp, _i]),
    ("allocateArray", None, [C.POINTER(_vp), _i]),
    ("deleteNBodyArrays", None, [_vp]),
    ("deleteArray", None, [_vp]),
    ("allocateIntArray", None, [C.POINTER(_vp), _i]),
    ("deleteIntArray", None, [_vp]),
    ("copyArrayFromDevice", None, [_vp, _vp, C.c_uint, _i]),
    ("copyArrayToDevice", None, [_vp, _vp, _i]),
    ("registerGLBufferObject", None, [C.c_uint]),
    ("unregisterGLBufferObject", None, [C.c_uint]),
    ("threadSync", None, []),
    ("calculateNForces", None, [_vp, _vp, _vp, _i, _f, _i, _vp, _f, _i, _i]),
    ("calculateNForcesNoAlloc", _i, [_vp, _vp, _vp, _i, _f, _i, _vp, _f, _i, _i]),
    ("threadExit", None, []),
    ("lj_config_init", None, [C.POINTER(LJConfig)]),
    ("lj_create", _i, [C.POINTER(LJConfig), C.POINTER(_vp)]),
    ("lj_destroy", _i, [_vp]),
    ("lj_box_length", _d, [_vp]),
    ("lj_rdf_dr2_default", _f, [_i]),
    ("lj_sample_initial_state", _i, [_vp, C.c_uint64]),
    ("lj_sample_initial_state_ex", _i, [_vp, C.c_uint64, _i]),
    ("lj_set_array", _i, [_vp, _i, _vp]),
    ("lj_get_array", _i, [_vp, _i, _vp]),
    ("lj_set_array_f64", _i, [_vp, _i, _vp]),
    ("lj_force_plan_info", _i, [_vp, _vp, _vp]),
    ("lj_velocity_histogram", _i, [_vp, C.c_double, _i, _vp]),
    ("lj_shear_pressure", _i, [_vp, C.POINTER(_d)]),
    ("lj_forces_shard", _i, [_vp, _i, _i, _i, _vp, _vp]),
    ("lj_time_forces_shard", _i, [_vp, _i, _i, _i, _i, C.POINTER(_d)]),
    ("lj_time_force_items_share", _i, [_vp, _i, _i, _i, C.POINTER(_d)]),
    ("lj_get_array_f64", _i, [_vp, _i, _vp]),
    ("lj_compute_forces", _i, [_vp, C.POINTER(LJObs)]),
    ("lj_compute_parameters", _i, [_vp, C.POINTER(LJObs)]),
    ("lj_get_observables", _i, [_vp, C.POINTER(LJObs)]),
    ("lj_integrate", _i, [_vp, _d, _i, _vp]),
    ("lj_integrate_async", _i, [_vp, _d, _i]),
    ("lj_rdf", _i, [_vp, _i, _i, _f, _d, _vp, _i]),
    ("lj_renormalize_velocities", _i, [_vp, _i]),
    ("lj_renormalize_to_energy", _i, [_vp, _d]),
    ("lj_reset_averaging", _i, [_vp]),
    ("lj_apply_boundary_conditions", _i, [_vp]),
    ("lj_set_mode", _i, [_vp, _i, _i]),
    ("lj_synchronize", _i, [_vp]),
    ("lj_device_arrays", _i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    ("lj_time_force_kernel", _i, [_vp, _i, C.POINTER(_d)]),
    ("lj_time_force_kernel_stamped", _i, [_vp, _i, C.POINTER(_d), C.POINTER(_d)]),
    ("lj_time_rdf", _i, [_vp, _i, _i, _f, _d, _i, C.POINTER(_d)]),
    ("lj_forces_device", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("lj_set_array_device", _i, [_vp, _i, _vp, _vp]),
    ("lj_get_array_device", _i, [_vp, _i, _vp, _vp]),
    ("lj_set_lj_params", _i, [_vp, _d, _d]),
    ("lj_adjoint_pass_device", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("lj_trajectory_loss", _i, [_vp, _d, _i, _d, _d, _vp]),
    ("lj_trajectory_grad", _i, [_vp, _d, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp]),
    ("lj_batch_create", _i, [_i, _vp, _vp, _vp, _i, _i, _i, C.POINTER(_vp)]),
    ("lj_batch_destroy", _i, [_vp]),
    ("lj_batch_num_particles", _i, [_vp]),
    ("lj_batch_sample_initial_state", _i, [_vp, C.c_uint64]),
    ("lj_batch_set_array", _i, [_vp, _i, _i, _vp]),
    ("lj_batch_get_array", _i, [_vp, _i, _i, _vp]),
    ("lj_batch_compute_forces", _i, [_vp]),
    ("lj_batch_integrate", _i, [_vp, _d, _i]),
    ("lj_batch_integrate_hist", _i, [_vp, _d, _i, _vp]),
    ("lj_batch_get_observables", _i, [_vp, _vp]),
    ("lj_batch_set_subvolume", _i, [_vp, _d]),
    ("lj_batch_get_stats", _i, [_vp, _vp]),
    ("lj_batch_reset_stats", _i, [_vp]),
    ("lj_batch_set_tuning", _i, [_vp, C.POINTER(LJBatchTuning)]),
    ("lj_batch_get_tuning", _i, [_vp, C.POINTER(LJBatchTuning)]),
    ("lj_batch_time_force", _i, [_vp, _i, C.POINTER(_d)]),
    ("lj_batch_autotune", _i, [_vp, _i, C.POINTER(LJBatchTuning), C.POINTER(_d)]),
    ("lj_fl_configure", _i, [_vp, _vp, _i]),
    ("lj_fl_num_entries", _i, [_vp, C.POINTER(_i)]),
    ("lj_fl_values", _i, [_vp, _vp]),
    ("lj_fl_sample", _i, [_vp]),
    ("lj_fl_counts", _i, [_vp, _vp]),
    ("lj_fl_stats", _i, [_vp, _vp]),
    ("lj_fl_reset", _i, [_vp]),
    ("lj_set_step_sampling", _i, [_vp, C.POINTER(LJStepSampling)]),
    ("lj_step_rdf_counts", _i, [_vp, _vp, C.POINTER(C.c_int64), _i]),
    ("lj_batch_fl_configure", _i, [_vp, _vp, _i]),
    ("lj_batch_fl_num_entries", _i, [_vp, C.POINTER(_i)]),
    ("lj_batch_fl_sample", _i, [_vp]),
    ("lj_batch_fl_counts", _i, [_vp, _vp]),
    ("lj_batch_fl_stats", _i, [_vp, _vp]),
    ("lj_batch_fl_reset", _i, [_vp]),
    ("lj_batch_set_step_sampling", _i, [_vp, _i]),
]

_lib = None


def lib():
    """The loaded libljhip.so (ImportError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libljhip.so not built at {LIB_PATH}: run "
                              "`python lennard-jones-cuda_amd/build.py` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != LJ_OK:
        raise LJError(rc, lib().lj_last_error().decode())
    return rc


def ptr(a):
    """ctypes void* of a numpy array or an int address."""
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    return C.c_void_p(a.ctypes.data)
