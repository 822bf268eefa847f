"""transfer_amd -- MI355X-native drop-in for uranix/Transfer's CudaContext +
kernels.cu path: the even-parity spherical-harmonic radiative-transfer
operator on P1 tetrahedral elements and its preconditioned CG, as
hand-written gfx950 HIP behind a C ABI (include/transfer.h)."""
from .api import (Angular, Batch, Config, HipContext, MLP, mlp_features, Mesh, Stencil, TransferError, check,
                  lib, rng_uniform, LIB_PATH, TR_SOLVER_REFERENCE, TR_SOLVER_PCG_FP64, TR_SOLVER_MIXED,
                  TR_CTX_GENERAL_NORMALS,
                  TET_DTYPE, FACE_DTYPE, expn, step_analytic, slab_analytic_U)
from . import problems

__all__ = ["Angular", "Batch", "Config", "MLP", "mlp_features", "HipContext", "Mesh", "Stencil", "TransferError", "check",
           "lib", "rng_uniform", "LIB_PATH", "problems", "TR_SOLVER_REFERENCE",
           "TR_SOLVER_PCG_FP64", "TR_SOLVER_MIXED", "TR_CTX_GENERAL_NORMALS", "TET_DTYPE", "FACE_DTYPE", "expn", "step_analytic",
           "slab_analytic_U"]
