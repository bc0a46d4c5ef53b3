"""ljamd -- Python access to the MI355X Lennard-Jones time step (libljhip.so).

The compute path is the HIP library; this package only declares its C-ABI
(include/lj_hip.h) for ctypes and wraps a context in a small class.
"""
from ._lib import (ARRAY_FORCE, ARRAY_POS, ARRAY_VEL, BC_NONE, BC_PERIODIC, BC_WALL, EVN,
                   FL_CUBE, FL_CUBE_CLOSED, INIT_GAUSSIAN, INIT_MAXWELL, FL_NSTAT, FL_SLAB_X, FL_SLAB_Y, FL_SLAB_Z, FL_VEL_X,
                   FL_VEL_Y, FL_VEL_Z, LIB_PATH, PREC_COMPENSATED, PREC_FP32, PREC_FP64, RDF_RBINS,
                   RDF_REF_CPU, RDF_REF_GPU, SAMPLE_RDF_LAST, SAMPLE_RDF_OFF, SAMPLE_RDF_SUM, TVN,
                   LJConfig, LJError, LJObs, SIGNATURES, check, lib)
from .batch import LJBatch, ensemble_boxes
from .system import CompatArrays, LJSystem

__all__ = [
    "ARRAY_FORCE", "ARRAY_POS", "ARRAY_VEL", "BC_NONE", "BC_PERIODIC", "BC_WALL", "EVN", "TVN",
    "LIB_PATH", "PREC_COMPENSATED", "PREC_FP32", "PREC_FP64", "RDF_RBINS", "RDF_REF_CPU", "RDF_REF_GPU",
    "LJConfig", "LJError", "LJObs", "SIGNATURES", "check", "lib", "CompatArrays", "LJSystem",
    "LJBatch", "ensemble_boxes", "FL_CUBE", "FL_CUBE_CLOSED", "FL_NSTAT", "FL_SLAB_X", "FL_SLAB_Y",
    "FL_SLAB_Z", "FL_VEL_X", "FL_VEL_Y", "FL_VEL_Z", "SAMPLE_RDF_LAST", "SAMPLE_RDF_OFF",
    "SAMPLE_RDF_SUM", "INIT_GAUSSIAN", "INIT_MAXWELL",
]
