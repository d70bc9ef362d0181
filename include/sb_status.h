/* Status codes, tensor kinds and contraction flags shared by the
 * sparsebackends_amd C ABI (include/sparsebackends_amd.h).
 *
 * The reference raises Julia exceptions; each exception class it raises on
 * the contraction path maps to one status code here, so a caller can tell
 * the classes apart exactly as a Julia `@test_throws` would:
 *   AssertionError  (@assert, e.g. SparseBackends/src/blocksparse/storage.jl:126-132)
 *   ErrorException  (error(), e.g. tensoralgebra/contract_bs_bs.jl:603,621)
 *   ArgumentError   (e.g. blocksparse/conversions.jl:232-234, coo/storage.jl:141-142)
 *   KeyError        (Dict lookup of a missing label, e.g. contract.jl:100 mapB[rlab])
 *   MethodError     (no kernel method for the operand kinds, contract.jl:100,102)
 *   BoundsError     (storage.jl:28, coo/storage.jl:26)
 */
#ifndef SB_STATUS_H
#define SB_STATUS_H

#define SB_OK            0
#define SB_E_ASSERT      1  /* AssertionError */
#define SB_E_ERROR       2  /* ErrorException */
#define SB_E_ARGUMENT    3  /* ArgumentError */
#define SB_E_KEY         4  /* KeyError */
#define SB_E_METHOD      5  /* MethodError: no kernel for these operand kinds */
#define SB_E_BOUNDS      6  /* BoundsError */
#define SB_E_DEVICE      7  /* HIP runtime / launch failure */
#define SB_E_NOMEM       8  /* device or host allocation failed */
#define SB_E_HANDLE      9  /* null handle or handle of the wrong kind */
#define SB_E_LIMIT      10  /* a size beyond what the library supports */
#define SB_E_NOCONV     11  /* LAPACKException: an SVD did not converge (the sweep cap of the
                               one-sided Jacobi, sb_ctx_set_svd_sweeps, was reached) */

/* tensor kinds (the reference's storage types) */
#define SB_KIND_BLOCKSPARSE 1  /* NewBlockSparseSorted  (blocksparse/storage.jl:3-9) */
#define SB_KIND_COO         2  /* COOTensor             (coo/storage.jl:1-6) */
#define SB_KIND_DENSE       3  /* plain Array (column-major) */
#define SB_KIND_ALIASED     4  /* AliasedBlockSparse    (aliased/storage.jl:46-59) */

/* declared key integer types (rekey) and payload element types (retypeval):
 * keys are stored as int32 and payloads as float64 on the device whatever
 * the declared type; sb_tensor_types reports the declaration */
#define SB_KEY_INT64   0  /* the reference's default K = Int */
#define SB_KEY_INT32   1
#define SB_KEY_INT16   2
#define SB_KEY_INT8    3
#define SB_KEY_UINT64  4
#define SB_KEY_UINT32  5
#define SB_KEY_UINT16  6
#define SB_KEY_UINT8   7
#define SB_VAL_FLOAT64 0  /* the default T = Float64 */
#define SB_VAL_FLOAT32 1
#define SB_VAL_FLOAT16 2
#define SB_VAL_INT64   3
#define SB_VAL_INT32   4
#define SB_VAL_INT16   5
#define SB_VAL_INT8    6

/* contraction flags.  With flags == 0 every kernel follows the reference
 * code, including its documented divergences from the dense mathematics. */
/* multi-label BS x BS merge-join: join on the order the keys are sorted in,
 * instead of the reference's first-axis-first comparison
 * (tensoralgebra/contract_bs_bs.jl:430-442,687-699), which can skip
 * matches when two or more shared labels sit in the key prefix. */
#define SB_FLAG_EXACT_JOIN        (1u << 0)
/* COO x Dense: create no output block whose contribution is all zero
 * (the reference creates one for every entry, contract_coo_dense.jl:113-122
 * in the included file; its test coo_dense.jl:164-189 expects none). */
#define SB_FLAG_SKIP_ZERO_SLICES  (1u << 1)
/* Box products (BS x BS contract_shared!, strided route): choose the tile
 * shape and grid density by timing every candidate on the plan's own work at
 * plan creation, instead of the padded-work cost model; the choice is cached
 * per block-size histogram (log2 bins of the output blocks' row and column
 * extents), so later plans of the same histogram reuse it without timing.
 * Speed only: every choice gives the same result. */
#define SB_FLAG_AUTOTUNE          (1u << 2)

/* precision form of the block GEMMs (bits 8..11).  Storage stays float64. */
#define SB_PREC_MASK    (0xFu << 8)
#define SB_PREC_FP64    (0u << 8)   /* FP64 MFMA, the drop-in */
#define SB_PREC_FP32    (1u << 8)   /* operands rounded to FP32, FP32 MFMA, FP32 accumulate */
#define SB_PREC_BF16X3  (2u << 8)   /* operands split in three bf16 terms, FP32 accumulate */

/* tile shape of the BS x BS block products (bits 16..19): 0 chooses from the
 * operands' block boxes at plan time; 1 + c forces shape c of
 * workgroup tiles 16x128, 128x16, 32x32, 64x32, 32x64, 64x64, 16x64, 64x16 (c = 0..7) or
 * wave tiles 16x16, 32x16, 16x32, 32x32, 64x16, 16x64 (c = 8..13); speed only */
#define SB_TILE_SHIFT   16
#define SB_TILE_MASK    (0xFu << SB_TILE_SHIFT)

#endif
