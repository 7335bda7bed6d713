"""fjmodels_amd -- MI355X (gfx950) implementation of the fJmodels SCF iteration.

The compute path is libfjmodels.so (HIP kernels behind the C ABI in
include/fjmodels.h); the C++ facade libfjfacade.so and the `fjmodels` driver
mirror the reference's Potential / computeNewPhi / setDF / tabulateDelta
surface.  This Python module is a thin ctypes binding used by tests and the
benchmark; importing it does not touch the GPU.
"""
from .capi import (Batch, Context, DFParams, FJError, df_params, FJ_QUAD_LOG22, FJ_QUAD_TANHSINH22,
                   FJ_POISSON_QUADPACK, FJ_POISSON_BLOCKSCAN, FJ_PREC_FP64, FJ_PREC_MIXED, FJ_PREC_FAST, LIB_PATH,
                   NOALLOC_MAX_POINTS)
from .scf import log_grid, Model
from .sweep import Sweep, draw_models

__all__ = ["Batch", "Context", "DFParams", "FJError", "df_params", "log_grid", "Model", "Sweep", "draw_models",
           "LIB_PATH",
           "FJ_QUAD_LOG22", "FJ_QUAD_TANHSINH22", "FJ_POISSON_QUADPACK", "FJ_POISSON_BLOCKSCAN",
           "FJ_PREC_FP64", "FJ_PREC_MIXED", "FJ_PREC_FAST", "NOALLOC_MAX_POINTS"]
