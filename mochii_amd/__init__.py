"""MI355X-native photon-packet transport and cell balance for MoCHII models.

The compute path is the HIP library libmochii_gpu.so (C-ABI in
include/mochii_gpu.h); this package binds it with ctypes and mirrors the
reference's Fortran driver API (module procedures of ionizing_field_mod,
ion_balance_mod, thermal_mod, gas_opacity_mod) at pass granularity.

Importing the package does not need a GPU; creating a Model does, and fails
loudly (MochiiError) when the library or the device is missing -- there is no
CPU fallback.
"""
from .lib import (MC_FP64, MC_MIXED, MC_GRID_AMR, MC_GRID_CAR, MC_EXT_SPH, MC_EXT_REC,
                  MC_DIFFUSE_THRESHOLD, MC_DIFFUSE_EXPONENTIAL, MC_DIFFUSE_MILNE,
                  MochiiError, Params, library, library_path)
from .model import Model, run_model
from . import atomic

__all__ = ["MC_FP64", "MC_MIXED", "MC_GRID_AMR", "MC_GRID_CAR", "MC_EXT_SPH", "MC_EXT_REC",
           "MC_DIFFUSE_THRESHOLD", "MC_DIFFUSE_EXPONENTIAL", "MC_DIFFUSE_MILNE",
           "MochiiError", "Params", "Model", "run_model", "library", "library_path", "atomic"]
