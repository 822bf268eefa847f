"""transfer_amd -- MI355X-native drop-in for uranix/Transfer's CudaContext +
kernels.cu path: the even-parity spherical-harmonic radiative-transfer
operator on P1 tetrahedral elements and its preconditioned CG, as
hand-written gfx950 HIP behind a C ABI (include/transfer.h)."""
from .api import (Angular, Batch, Config, HipContext, MLP, mlp_features, Mesh, mesh_view_of, abs_normal, Stencil, TransferError, check,
                  lib, rng_uniform, LIB_PATH, TR_SOLVER_REFERENCE, TR_SOLVER_PCG_FP64, TR_SOLVER_MIXED,
                  TR_CTX_GENERAL_NORMALS, TR_OK, TR_ERR_ARG, TR_ERR_HIP, TR_ERR_STATE, TR_ERR_NOCONV,
                  TET_DTYPE, FACE_DTYPE, expn, step_analytic, slab_analytic_U)
from . import problems

__all__ = ["Angular", "Batch", "Config", "MLP", "mlp_features", "HipContext", "Mesh", "mesh_view_of", "abs_normal", "Stencil", "TransferError", "check",
           "lib", "rng_uniform", "LIB_PATH", "problems", "TR_SOLVER_REFERENCE",
           "TR_SOLVER_PCG_FP64", "TR_SOLVER_MIXED", "TR_CTX_GENERAL_NORMALS", "TR_OK", "TR_ERR_ARG", "TR_ERR_HIP",
           "TR_ERR_STATE", "TR_ERR_NOCONV", "TET_DTYPE", "FACE_DTYPE", "expn", "step_analytic",
           "slab_analytic_U"]
