"""sparsebackends_amd: the block-sparse contraction backend of
Ritvik1sharma/SparseBackends rebuilt for MI355X (gfx950).

Storage types and contract!/permutedims mirror the reference's Julia API
(SparseBackends/src/SparseBackends.jl:4-8); all work runs in
libsparsebackends_amd.so (HIP), reached through its C ABI.
"""
from ._abi import SBError, load, LIB_PATH
from .storage import (NewBlockSparseSorted, COOTensor, DenseTensor, AliasedBlockSparse, blocksparse_from_dense, coo_from_dense,
                      to_dense, rekey, retypeval, KEY_TYPES, VAL_TYPES, context, bind_context, memory, trim, synchronize, profile, profile_reset, profile_read, set_svd_sweeps)
from .tensoralgebra import (contract_, contract_bs_bs_to_dense_, Plan, HeffPlan, PlanBatch, PlanGraph, set_graphs, labels_array, EXACT_JOIN, SKIP_ZERO_SLICES, AUTOTUNE,
                            PREC_FP64, PREC_FP32, PREC_BF16X3)
from .krylov import eigsolve, eigsolve_refined, eigsolve_batch, eigsolve_multi, ConvergenceInfo
from .factorize import blocksparse_svd, blocksparse_left_qr, left_factorize, qr_sectors, Spectrum

__all__ = ["SBError", "load", "LIB_PATH", "NewBlockSparseSorted", "COOTensor", "DenseTensor", "AliasedBlockSparse",
           "blocksparse_from_dense", "coo_from_dense", "to_dense", "rekey", "retypeval", "KEY_TYPES", "VAL_TYPES", "context", "bind_context", "memory", "trim", "synchronize", "profile", "profile_reset", "profile_read", "set_svd_sweeps", "contract_",
           "contract_bs_bs_to_dense_", "Plan", "HeffPlan", "PlanBatch", "PlanGraph", "set_graphs", "labels_array", "EXACT_JOIN", "SKIP_ZERO_SLICES", "AUTOTUNE", "PREC_FP64",
           "PREC_FP32", "PREC_BF16X3", "eigsolve", "eigsolve_refined", "eigsolve_batch", "eigsolve_multi", "ConvergenceInfo",
           "blocksparse_svd", "Spectrum", "blocksparse_left_qr", "left_factorize", "qr_sectors"]
