/* sparsebackends_amd: block-sparse tensor contraction for MI355X (gfx950).
 *
 * C ABI of libsparsebackends_amd.so.  It replaces the storage types and the
 * contract/permute entry points of the reference backend
 * (Ritvik1sharma/SparseBackends, Julia), so a Julia `ccall` (or any FFI) can
 * drive it from the reference's ITensors boundary (wrapped_contract,
 * SparseBackends/src/tensor_wrappers.jl:614-676).
 *
 * Conventions (those of the reference, kept so results compare bit for bit):
 *  - tensors are device-resident; handles own their device memory;
 *  - keys and ids are 1-based; a block-sparse key is the tuple of the first
 *    P = N - N2 coordinates, passed row-wise (nblocks x P int64);
 *  - keys of a NewBlockSparseSorted are sorted by the column-major linear
 *    index of the prefix (axis 1 fastest); ids are payload slots;
 *  - payloads and dense arrays are float64, column-major; block `id` lives at
 *    data[(id-1)*blksize .. id*blksize);
 *  - a label is (id, plev) as two uint64 (tensor_wrappers.jl:577-580),
 *    so label arrays hold 2*rank uint64;
 *  - every call returns SB_OK or one status of sb_status.h, and
 *    sb_last_error() describes the last failure of the calling thread.
 * All work is ordered on the context's stream; a call that returns host data
 * synchronises that stream.  One context per host thread; calls on one
 * context must not overlap.
 */
#ifndef SPARSEBACKENDS_AMD_H
#define SPARSEBACKENDS_AMD_H

#include <stdint.h>

#include "sb_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sb_ctx_s* sb_ctx_t;      /* device + stream + device-memory cache */
typedef struct sb_tensor_s* sb_tensor_t; /* a device-resident tensor of one storage kind */
typedef struct sb_plan_s* sb_plan_t;     /* a contraction with its block structure fixed */
typedef struct sb_batch_s* sb_batch_t;   /* independent plan chains executed together */
typedef struct sb_graph_s* sb_graph_t;   /* a plan chain replayed as a HIP graph */

/* ------------------------------------------------------------ runtime */
const char* sb_last_error(void);
int sb_version(void);
/* New: the reference runs in one Julia process on the CPU; a context binds a
 * HIP device and a stream (per-thread scratch of tensor_wrappers.jl:582-598). */
int sb_ctx_create(int device, sb_ctx_t* out);
int sb_ctx_destroy(sb_ctx_t ctx);
/* run on a caller's hipStream_t (NULL: the context's own stream) */
int sb_ctx_set_stream(sb_ctx_t ctx, void* hip_stream);
int sb_ctx_synchronize(sb_ctx_t ctx);
/* device bytes held by live tensors / kept in the cache for reuse */
int sb_ctx_memory(sb_ctx_t ctx, int64_t* in_use, int64_t* cached);
int sb_ctx_trim(sb_ctx_t ctx);
/* New (the reference's only tracing is the SB_TRACE prints, e.g.
 * contract_bs_dense.jl:112-114): time the numeric steps of plan executions
 * with HIP events on the context's stream, aggregated per kernel name.
 * enable: 0 off, 1 every step, 2 only the block products (grouped_gemm), whose
 * timing then disturbs the stream least.  Plan chains run eagerly (no graph
 * replays) while profiling is on. */
int sb_ctx_profile(sb_ctx_t ctx, int enable);
/* the sweep cap of the one-sided Jacobi SVDs (blocksparse_svd, the sector SVD;
 * default 40): a factorisation still rotating after it returns SB_E_NOCONV,
 * as the reference's LAPACK SVD raises LAPACKException when it does not converge */
int sb_ctx_set_svd_sweeps(sb_ctx_t ctx, int sweeps);
int sb_ctx_profile_reset(sb_ctx_t ctx);
/* synchronises; writes lines "name count total_ms min_ms max_ms" into buf
 * (NUL-terminated, at most cap bytes); *needed = bytes the full text takes */
int sb_ctx_profile_read(sb_ctx_t ctx, char* buf, int64_t cap, int64_t* needed);

int sb_tensor_free(sb_tensor_t t);
int sb_tensor_kind(sb_tensor_t t, int* kind);

/* ------------------------------------------------------------ NewBlockSparseSorted */
/* NewBlockSparseSorted{T,N,N2}(dims) (blocksparse/storage.jl:86-93): no blocks */
int sb_bs_new(sb_ctx_t ctx, int N, int N2, const int64_t* dims, sb_tensor_t* out);
/* NewBlockSparseSorted{T,N,N2,P}(dims, blksize, keys, ids, data) (storage.jl:3-13):
 * host arrays uploaded as given (keys nblocks x P, ids nblocks, data nblocks*blksize) */
int sb_bs_from_host(sb_ctx_t ctx, int N, int N2, const int64_t* dims, int64_t nblocks, const int64_t* keys,
                    const int64_t* ids, const double* data, sb_tensor_t* out);
int sb_bs_info(sb_tensor_t t, int* N, int* N2, int64_t* dims, int64_t* nblocks, int64_t* blksize);
/* the fields keys/ids/data copied to host arrays */
int sb_bs_to_host(sb_ctx_t ctx, sb_tensor_t t, int64_t* keys, int64_t* ids, double* data);
/* device pointers of the fields (int32 keys/ids, float64 payload), valid until the tensor changes.
 * Asking for the payload pointer marks the payload as written from outside:
 * later contractions rescan it for zero padding on every use. */
int sb_bs_device_ptrs(sb_tensor_t t, const int32_t** keys, const int32_t** ids, double** data);
/* assignment to the data field (A.data .= x on the struct of storage.jl:3-9):
 * replace the payload (nblocks*blksize values, host or device memory) keeping
 * keys and ids */
int sb_bs_set_payload(sb_ctx_t ctx, sb_tensor_t t, const double* data, int from_device);
/* getindex / setindex! (storage.jl:100-124); I holds N 1-based coordinates */
int sb_bs_getindex(sb_ctx_t ctx, sb_tensor_t t, const int64_t* I, double* v);
int sb_bs_setindex(sb_ctx_t ctx, sb_tensor_t t, double v, const int64_t* I);
/* sort!(::NewBlockSparseSorted) (storage.jl:403-410): lexicographic keys, ids follow */
int sb_bs_sort_lex(sb_ctx_t ctx, sb_tensor_t t);
/* is_dense (storage.jl:413-433) */
int sb_bs_is_dense(sb_tensor_t t, int* out);
/* blocksparse_from_dense(A, Val(N2); atol, rtol, dropzeros) (blocksparse/conversions.jl:21-75) */
int sb_bs_from_dense(sb_ctx_t ctx, sb_tensor_t dense, int N2, double atol, double rtol, int dropzeros,
                     sb_tensor_t* out);
/* to_dense(B; init, merged_axes) (conversions.jl:83-187); groups: ngroups pairs of 1-based axes */
int sb_bs_to_dense(sb_ctx_t ctx, sb_tensor_t t, double init, int ngroups, const int64_t* groups, sb_tensor_t* out);
/* fuse_two_axes! (conversions.jl:299-314): may permute t in place (fuse_dense_axes!) */
int sb_bs_fuse_two_axes(sb_ctx_t ctx, sb_tensor_t t, int64_t ax1, int64_t ax2, sb_tensor_t* out);

/* ------------------------------------------------------------ COOTensor */
/* COOTensor{T,N}(dims) (coo/storage.jl:14-15) */
int sb_coo_new(sb_ctx_t ctx, int N, const int64_t* dims, sb_tensor_t* out);
/* COOTensor{T,N}(dims, keys, vals, dirty) (coo/storage.jl:10-12); keys nnz x N */
int sb_coo_from_host(sb_ctx_t ctx, int N, const int64_t* dims, int64_t nnz, const int64_t* keys, const double* vals,
                     int dirty, sb_tensor_t* out);
int sb_coo_info(sb_tensor_t t, int* N, int64_t* dims, int64_t* nnz, int* dirty);
int sb_coo_to_host(sb_ctx_t ctx, sb_tensor_t t, int64_t* keys, double* vals);
int sb_coo_device_ptrs(sb_tensor_t t, const int32_t** keys, double** vals);
/* getindex / setindex!(...; dirty) (coo/storage.jl:49-60, 84-134) */
int sb_coo_getindex(sb_ctx_t ctx, sb_tensor_t t, const int64_t* I, double* v);
int sb_coo_setindex(sb_ctx_t ctx, sb_tensor_t t, double v, const int64_t* I, int dirty);
/* sort!(::COOTensor) = sort_coo! (coo/storage.jl:182-196) */
int sb_coo_sort(sb_ctx_t ctx, sb_tensor_t t);
/* coo_from_dense(A; atol, rtol) (coo/conversions.jl:13-26) and to_dense (:4-10) */
int sb_coo_from_dense(sb_ctx_t ctx, sb_tensor_t dense, double atol, double rtol, sb_tensor_t* out);
int sb_coo_to_dense(sb_ctx_t ctx, sb_tensor_t t, sb_tensor_t* out);
/* fuse_axes_coo (coo/conversions.jl:174-223) */
int sb_coo_fuse_axes(sb_ctx_t ctx, sb_tensor_t t, int64_t ax1, int64_t ax2, sb_tensor_t* out);

/* ------------------------------------------------------------ dense arrays */
int sb_dense_from_host(sb_ctx_t ctx, int N, const int64_t* dims, const double* data, sb_tensor_t* out);
/* wrap (borrow != 0, caller keeps the memory alive) or copy caller device memory */
int sb_dense_from_device(sb_ctx_t ctx, int N, const int64_t* dims, double* dptr, int borrow, sb_tensor_t* out);
int sb_dense_info(sb_tensor_t t, int* N, int64_t* dims);
int sb_dense_to_host(sb_ctx_t ctx, sb_tensor_t t, double* data);
int sb_dense_device_ptr(sb_tensor_t t, double** data);

/* ------------------------------------------------------------ AliasedBlockSparse */
/* AliasedBlockSparse{T,N,N2}(dims) (aliased/storage.jl:122-132): no blocks, no templates */
int sb_al_new(sb_ctx_t ctx, int N, int N2, const int64_t* dims, sb_tensor_t* out);
/* the struct of aliased/storage.jl:46-59 from host arrays: templates (ntemplates*blksize),
 * keys (nblocks x P, sorted by the column-major prefix lin), alias_ids (1-based), scalars */
int sb_al_from_host(sb_ctx_t ctx, int N, int N2, const int64_t* dims, int64_t ntemplates, const double* templates,
                    int64_t nblocks, const int64_t* keys, const int64_t* alias_ids, const double* scalars,
                    sb_tensor_t* out);
int sb_al_info(sb_tensor_t t, int* N, int* N2, int64_t* dims, int64_t* nblocks, int64_t* blksize,
               int64_t* ntemplates);
int sb_al_to_host(sb_ctx_t ctx, sb_tensor_t t, int64_t* keys, int64_t* alias_ids, double* scalars, double* templates);
/* getindex (aliased/storage.jl:138-147) */
int sb_al_getindex(sb_ctx_t ctx, sb_tensor_t t, const int64_t* I, double* v);
/* compression_ratio (aliased/storage.jl:163-164) */
int sb_al_compression_ratio(sb_tensor_t t, double* out);
/* to_blocksparse (aliased/conversions.jl:18-32) */
int sb_al_to_blocksparse(sb_ctx_t ctx, sb_tensor_t t, sb_tensor_t* out);
/* to_dense(A; init) (aliased/conversions.jl:45-64) */
int sb_al_to_dense(sb_ctx_t ctx, sb_tensor_t t, double init, sb_tensor_t* out);

/* ------------------------------------------------------------ permutes */
/* Base.permutedims: NewBlockSparseSorted (storage.jl:188-264), COOTensor
 * (coo/storage.jl:153-166, keys only, marks dirty), AliasedBlockSparse
 * (aliased/storage.jl:172-239, templates only), dense arrays; perm 1-based */
int sb_permutedims(sb_ctx_t ctx, sb_tensor_t t, const int64_t* perm, sb_tensor_t* out);
/* permutedims! (storage.jl:271-392; coo/storage.jl:168-180) */
int sb_permutedims_inplace(sb_ctx_t ctx, sb_tensor_t t, const int64_t* perm);

/* ------------------------------------------------------------ contraction */
/* contract!(C, labelsC, A, labelsA, B, labelsB) (tensoralgebra/contract.jl:37-104).
 * C was made by sb_bs_new / sb_coo_new with its dims and N2 and is overwritten.
 * A and B are not modified.  Operand kinds follow the reference's methods:
 * BS x BS, BS x Dense, COO x COO, COO x BS, COO x Dense (sparse operand first,
 * as wrapped_contract arranges, tensor_wrappers.jl:561-573), and into an
 * AliasedBlockSparse C: Aliased x Dense, Aliased x Aliased (contract_aliased.jl,
 * contract_aliased_shared.jl), COO x Dense (contract_aliased_coo_dense.jl).
 * flags: sb_status.h. */
int sb_contract(sb_ctx_t ctx, sb_tensor_t C, const uint64_t* labelsC, int nC, sb_tensor_t A, const uint64_t* labelsA,
                int nA, sb_tensor_t B, const uint64_t* labelsB, int nB, uint32_t flags);
/* contract_bs_bs_to_dense! (contract_bs_bs.jl:736-754): P_C = 0 sink into a dense C */
int sb_contract_bs_bs_to_dense(sb_ctx_t ctx, sb_tensor_t Cdense, const uint64_t* labelsC, int nC, sb_tensor_t A,
                               const uint64_t* labelsA, int nA, sb_tensor_t B, const uint64_t* labelsB, int nB,
                               uint32_t flags);

/* New: the reference recomputes the block structure on every call.  A plan
 * runs sb_contract once (C gets its keys/ids and values) and keeps the numeric
 * phase; sb_plan_execute recomputes C's values from the operands' current
 * values, which must keep their block structure (same keys, ids and payload
 * buffers).  For the families whose output blocks depend on values (rank-1 over a
 * dense-tail index, COO products, the zero-slice flag) the block list of the
 * first call is kept; in the rank-1 families (contract_dense_bb!,
 * contract_dense_bd!) the r slices that contribute follow each replay's values,
 * in the COO families the contributing (entry, block) pairs are the first
 * call's.  Plans of one context may be executed back to back and
 * captured into a hipGraph by the caller. */
int sb_plan_create(sb_ctx_t ctx, sb_tensor_t C, const uint64_t* labelsC, int nC, sb_tensor_t A,
                   const uint64_t* labelsA, int nA, sb_tensor_t B, const uint64_t* labelsB, int nB, uint32_t flags,
                   sb_plan_t* out);
int sb_plan_execute(sb_ctx_t ctx, sb_plan_t plan);
/* New (no counterpart in the reference, which applies the two-site effective
 * Hamiltonian of DMRG as four contract! calls psi*L*W1*W2*R, ITensorMPS's
 * product(ProjMPO, psi) as test_product_dmrg_state.jl contracts it): the
 * whole H_eff*psi as one plan whose execution is one fused kernel, with no
 * intermediate tensor.  Layouts (those of the four-call chain):
 *   psi [Ql, Qr | s1, s2, l, r]  L [Ql', Ql | wl, l', l]  R [Qr', Qr | wr, r', r]
 *   W1 [ | wl, s1', s1, wm]  W2 [ | wm, s2', s2, wr]  (one dense block each)
 * Hpsi, made by sb_bs_new with psi's dims and N2 = 4, gets the block list the
 * chain's last contract! gives it (keys and ids bit for bit) and at every
 * execution the value of the chain for psi's current values (summation order
 * of its own: FP64 to rounding; a NaN or Inf in psi reaches the outputs its
 * row feeds, which may be fewer than the chain's dense MPO steps spread it to).  The sector extents are the environments'
 * nonzero reach at creation: entries of psi beyond it are not read (the chain
 * multiplies them by zero environment entries); L, R, W1 and W2 must not change
 * afterwards (an execution after a library call changed them fails).  flags:
 * SB_PREC_FP64 or SB_PREC_FP32 (products in FP32, storage FP64).  SB_E_LIMIT
 * when an output block reaches beyond 4096 entries or the LDS tiles: then the
 * four contract! plans do the work. */
int sb_heff_plan_create(sb_ctx_t ctx, sb_tensor_t Hpsi, sb_tensor_t psi, sb_tensor_t L, sb_tensor_t W1,
                        sb_tensor_t W2, sb_tensor_t R, uint32_t flags, sb_plan_t* out);
/* flops: sum of 2*M*K*N over contributing block pairs (actual extents);
 * gemm_bytes / perm_bytes: algorithmic bytes of the block products / permutes */
int sb_plan_stats(sb_plan_t plan, double* flops, double* gemm_bytes, double* perm_bytes, int64_t* npairs,
                  int64_t* nblocks_out);
/* distinct payload bytes of the block-product operands (A, B) and result C */
int sb_plan_operand_bytes(sb_plan_t plan, double* a_bytes, double* b_bytes, double* c_bytes);
/* New (no counterpart in the reference): the box product's launch shape —
 * tile configuration (-1: the plan is not one box product), workgroups per
 * CU (0: as many as are resident), and for a plan made with SB_FLAG_AUTOTUNE
 * the number of candidates timed (0: the choice came from the histogram
 * cache or the cost model) and the best time measured in ms. */
int sb_plan_box_config(sb_plan_t plan, int* tile, int* wg_per_cu, int* candidates, double* best_ms);
/* forget the autotuner's cached choices (all devices) */
int sb_autotune_reset(void);
/* New (no counterpart in the reference): the matrix-core FLOPs one
 * execution of the plan issues, tile padding included (the fused H_eff plan
 * counts them; 0 for plans that do not) */
int sb_plan_mfma_flops(sb_plan_t plan, double* out);
int sb_plan_free(sb_plan_t plan);

/* New (no counterpart in the reference, which re-derives every contraction on
 * every call): a chain of plans executed in order as one HIP graph, e.g. the
 * four contractions of one H_eff*psi.  sb_graph_launch runs the chain: the
 * first launch eagerly, the second captured into a graph (kernel nodes only),
 * later ones as one graph launch while the chain's buffers and its host-side
 * decisions (which operands need a box scan) match the capture; results are
 * those of executing the plans one by one.  Chains with a plan that is not a
 * box product, and every launch while the context profiles, run eagerly.
 * sb_ctx_graphs sets whether eigensolves replay their operator this way:
 * 1 always, 0 never, -1 (default) when krylovdim * maxiter >= 8.  A capture
 * takes in whatever is launched on the context's stream meanwhile: a context
 * used by several threads at once must have graphs off (mode 0) and must not
 * launch chain graphs. */
int sb_graph_create(sb_ctx_t ctx, const sb_plan_t* plans, int nplans, sb_graph_t* out);
/* the same for one execution of every chain of a batch (sb_batch_execute with
 * all chains): its fused stages replayed as one graph launch; the batched
 * eigensolve uses it while all its chains are running */
int sb_batch_graph_create(sb_ctx_t ctx, sb_batch_t batch, sb_graph_t* out);
int sb_graph_launch(sb_ctx_t ctx, sb_graph_t graph);
int sb_graph_stats(sb_graph_t graph, int64_t* eager, int64_t* captures, int64_t* replays, int64_t* nodes);
int sb_graph_free(sb_graph_t graph);
int sb_ctx_graphs(sb_ctx_t ctx, int mode);

/* New (no counterpart in the reference, which contracts one problem per call;
 * it is the batched schedule the many-small-problems workloads need): a batch
 * of nchains independent plan chains of nstages plans each, plans[c*nstages+s]
 * being stage s of chain c (e.g. the four contractions of one H_eff*psi per
 * problem).  sb_batch_execute runs stage by stage; where every plan of a stage
 * is one box product of one precision form (the strided BS x BS route), the
 * stage is ONE device launch over the union of the chains' work lists, with
 * bit-identical results to executing each plan alone.  Other stages run their
 * plans one after the other.  active (ascending chain numbers; NULL or
 * nactive = 0: all) selects the chains to execute.  The batch refers to the
 * plans, which must outlive it. */
int sb_batch_create(sb_ctx_t ctx, const sb_plan_t* plans, int nchains, int nstages, sb_batch_t* out);
int sb_batch_execute(sb_ctx_t ctx, sb_batch_t batch, const int* active, int nactive);
/* flops of one execution of every chain; number of fused stages */
int sb_batch_stats(sb_batch_t batch, double* flops, int* fused_stages);
/* New: time every workgroup tile shape for each fused stage's joint launch
 * over all chains and keep the fastest (also done at creation when a stage's
 * plans were made with SB_FLAG_AUTOTUNE); the stage's shape (-1: not fused)
 * and its measured time */
int sb_batch_autotune(sb_ctx_t ctx, sb_batch_t batch);
int sb_batch_stage_tile(sb_batch_t batch, int stage, int* tile, double* tuned_ms);
int sb_batch_free(sb_batch_t batch);

/* ------------------------------------------------------------ key and value type conversions
 * The device stores keys as int32 and payloads as float64 whatever the
 * declared types; these check the reference's conditions on the device and
 * return a copy declaring the new type (read it back with sb_tensor_types).
 * A failed check returns SB_E_ERROR with the reference's message. */
/* rekey(A, K2; check) (blocksparse/conversions.jl:327-342 for NewBlockSparseSorted,
 * coo/conversions.jl:41-53 for COOTensor, aliased/conversions.jl:76-93 for
 * AliasedBlockSparse): key_type one of SB_KEY_*; with check, every prefix
 * dimension (block-sparse, aliased) and key coordinate must fit K2 */
int sb_rekey(sb_ctx_t ctx, sb_tensor_t A, int key_type, int check, sb_tensor_t* out);
/* retypeval(A::COOTensor, T2) (coo/conversions.jl:65-82): val_type one of
 * SB_VAL_*; every value must survive the round trip Float64(T2(v)) == v
 * (integer T2: v integral and inside T2's range) */
int sb_coo_retypeval(sb_ctx_t ctx, sb_tensor_t A, int val_type, sb_tensor_t* out);
/* the declared key and payload types of a tensor (SB_KEY_*, SB_VAL_*) */
int sb_tensor_types(sb_tensor_t t, int* key_type, int* val_type);

/* ------------------------------------------------------------ Krylov vector operations
 * The storage hooks a Krylov solver reaches through ITensors' ExternalStorage
 * (SparseBackends/src/tensor_wrappers.jl:792-923), on block-sparse tensors. */
/* norm(w::WrappedBlockSparse) = norm(data) (tensor_wrappers.jl:793) */
int sb_bs_norm(sb_ctx_t ctx, sb_tensor_t A, double* out);
/* copy(::WrappedBlockSparse) (tensor_wrappers.jl:817-823): a new tensor */
int sb_bs_copy(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t* out);
/* _external_similar -> _zero_similar (tensor_wrappers.jl:898-904,915-923): same keys and ids, zero data */
int sb_bs_similar(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t* out);
/* alpha * wbs (tensor_wrappers.jl:801-807): a new tensor with data alpha .* data */
int sb_bs_scale(sb_ctx_t ctx, sb_tensor_t A, double alpha, sb_tensor_t* out);
/* _external_fill! (tensor_wrappers.jl:850-853): fill!(data, v) */
int sb_bs_fill(sb_ctx_t ctx, sb_tensor_t A, double v);
/* _external_map_storage! -> _apply_elementwise! (tensor_wrappers.jl:856-862,877-886)
 * with f(r, a) = alpha*a + beta*r, position by position over the payloads;
 * SB_E_ERROR when the payload lengths differ (the reference's only check).
 * Covers scale! (A == R), axpy!, axpby! and copy! (alpha 1, beta 0). */
int sb_bs_axpby(sb_ctx_t ctx, double alpha, sb_tensor_t A, double beta, sb_tensor_t R);
/* wA + wB (tensor_wrappers.jl:826-840) and ITensors._add (:842-847): copy of A,
 * every block of B added by key, missing keys created by _ensure_block! */
int sb_bs_add(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t B, sb_tensor_t* out);
/* inner product over blocks of equal key (the scalar ITensors' inner gets
 * from the P_C = 0 contraction, tensor_wrappers.jl:646-652) */
int sb_bs_dot(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t B, double* out);
/* New (a helper for Krylov loops outside the library, no counterpart in the
 * reference): a new tensor with A's keys and ids holding B's block of each of
 * those keys (zero where B has none), the projection of an operator's image
 * onto the vector's block space that sb_eigsolve_lanczos applies */
int sb_bs_restrict(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t B, sb_tensor_t* out);
/* New, with no counterpart in the reference: sb_bs_norm for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Writes norm(A.data) into
 * out_device[0] (device memory) on the context's stream, without allocating
 * and without synchronising: the reduction's 1024 partial sums live in a
 * fixed 8 KiB scratch array per GPU, so every payload length fits.  The
 * scratch is shared by all calls on one GPU: calls that overlap on one GPU
 * (several streams or threads) race on it.  Same kernels as sb_bs_norm, so
 * the same value bit for bit. */
int sb_bs_norm_noalloc(sb_ctx_t ctx, sb_tensor_t A, double* out_device);

/* ------------------------------------------------------------ eigensolver
 * Replaces KrylovKit's eigsolve(A, x0, 1, :SR; ishermitian=true, krylovdim,
 * maxiter, tol) as two-site DMRG calls it on H_eff (call sites
 * test_product_dmrg_state.jl:221-226, test_product.jl:235-238; KrylovKit 0.10.2
 * is not part of the reference's sources).  FP64 Lanczos with full
 * reorthogonalisation (classical Gram-Schmidt twice) and thick restarts.
 * The operator is the chain of plans: plans[0] reads x, plans[nplans-1] writes
 * y, y has x's dims and its blocks are matched to x's by key: blocks of x
 * absent from y count as zero, blocks of y absent from x are left out (the
 * solve works in x's block space; for a charge-conserving H_eff those blocks
 * are structural zeros of the contraction).  The
 * plans' precision flags select the GEMM form; every vector operation is FP64.
 * krylovdim in 2..64 vectors per factorisation, maxiter factorisations at
 * most, converged when the residual norm of the lowest Ritz pair is <= tol.
 * On return x holds the normalised Ritz vector; ritz (capacity krylovdim, may
 * be NULL) the Ritz values of the last factorisation, ascending, nritz many. */
int sb_eigsolve_lanczos(sb_ctx_t ctx, const sb_plan_t* plans, int nplans, sb_tensor_t x, sb_tensor_t y,
                        int krylovdim, int maxiter, double tol, double* value, double* ritz, int* nritz,
                        double* normres, int64_t* numops, int* numiter, int* converged);

/* New (no counterpart in the reference, whose eigsolve takes one operator):
 * the same solve for every chain c of a batch, from xs[c] (first plan of
 * chain c reads it) to ys[c] (its last plan writes it), in lock step: each
 * iteration applies the chains still running as ONE sb_batch_execute and
 * synchronises once for all of them.  Every chain's result and Ritz vector
 * are bit for bit what sb_eigsolve_lanczos gives for it alone.  Output
 * arrays (may be NULL) hold nchains entries. */
int sb_eigsolve_lanczos_batch(sb_ctx_t ctx, sb_batch_t batch, const sb_tensor_t* xs, const sb_tensor_t* ys,
                              int nchains, int krylovdim, int maxiter, double tol, double* values, double* normres,
                              int64_t* numops, int* numiter, int* converged);

/* New (no counterpart in the reference; for serving loops that co-batch the
 * eigensolves of independent problems held by different host threads): the
 * lock-step solve of sb_eigsolve_lanczos_batch over chains whose plans and
 * vectors may belong to other contexts of the same device, one context per
 * thread.  plans[c*nstages+s] is stage s of chain c.  For the duration of the
 * call every such context is synchronised and then launches on ctx's stream;
 * at the end ctx's stream is synchronised and they get their own streams
 * back.  The other contexts must not be used by other threads during the
 * call.  Results are bit for bit those of sb_eigsolve_lanczos per chain. */
int sb_eigsolve_lanczos_multi(sb_ctx_t ctx, const sb_plan_t* plans, int nchains, int nstages, const sb_tensor_t* xs,
                              const sb_tensor_t* ys, int krylovdim, int maxiter, double tol, double* values,
                              double* normres, int64_t* numops, int* numiter, int* converged);

/* ------------------------------------------------------------ two-site SVD
 * blocksparse_svd (SparseBackends/src/ops_factorize.jl:50-148; ITensor-level
 * wrapper itensor_blocksparse_svd, tensor_wrappers.jl:975-1005) of a
 * NewBlockSparseSorted{T,4,2,2} A with prefix (L, R) and dense tail (s1, s2):
 * one SVD per L value of its d_s1 x (n_R * d_s2) block row, the singular
 * values sorted descending across all L, the kept count min(maxdim, all)
 * lowered while above mindim and below cutoff * max (maxdim < 0: no limit).
 * U: dims (n_L, k, d_s1), keys (L, nb); SV: dims (k, n_R, d_s2), keys (nb, R);
 * the singular value goes into SV (ortho_right = 0, "left") or U (1, "right").
 * svs receives min(svs_cap, k) kept singular values (may be NULL); truncerr =
 * sum of the discarded singular values squared.  d_s1 at most 32.  k = 0
 * (nothing kept) fails like the reference's constructor (dims .>= 1). */
int sb_bs_svd(sb_ctx_t ctx, sb_tensor_t A, int ortho_right, int64_t maxdim, int64_t mindim, double cutoff,
              sb_tensor_t* U, sb_tensor_t* SV, double* svs, int64_t svs_cap, int64_t* newdim, double* truncerr);

/* New (the sector-format counterpart of blocksparse_svd; the reference holds
 * psi only element-sparse): block-diagonal SVD of a two-site tensor
 * psi [Ql, Qr | s1, s2, l, r] (N = 6, N2 = 4; l, r padded in-sector axes) by
 * fused charge sector.  rowsec[il + nl*s1] (colsec[ir + nr*s2]) is the sector,
 * 0..nsec-1 or -1, of left sector il with site state s1 (right sector ir with
 * s2); dl[nl] and dr[nr] the actual extents of the bond sectors.  Sector m's
 * matrix has rows (Ql, s1, l) and columns (Qr, s2, r), both ascending; its
 * SVD runs in one workgroup (one-sided Jacobi in LDS; a sector matrix must
 * fit 160 KB).  Truncation as sb_bs_svd.  New bond: the sectors with kept
 * values in ascending order, padded to the largest kept count Dm.
 * A [Ql, Qm | s1, l, m] and B [Qm, Qr | s2, m, r] (ids in key order); the
 * singular values go into B (ortho_right = 0) or A (1).  kept[nsec]: kept
 * count per sector; svs: min(svs_cap, total) kept values, sector by sector. */
int sb_bs_svd_sectors(sb_ctx_t ctx, sb_tensor_t psi, int64_t nsec, const int64_t* rowsec, const int64_t* colsec,
                      const int64_t* dl, const int64_t* dr, int ortho_right, int64_t maxdim, int64_t mindim,
                      double cutoff, sb_tensor_t* A, sb_tensor_t* B, int64_t* kept, double* svs, int64_t svs_cap,
                      int64_t* newdim, double* truncerr);

/* ------------------------------------------------------------ QR
 * blocksparse_left_qr (SparseBackends/src/ops_factorize.jl:8-43) of a
 * NewBlockSparseSorted{T,3,1,2} A with prefix (l, r) and dense s: the dense
 * (n_l d_s) x n_r matrix of its nonzero l and r values, thin Householder QR
 * with LAPACK's sign conventions.  Q {T,3,2} (dims (A.dims[1], d_s, k)) has
 * one block per l value (keys ascending, ids 1..n_l); R is a dense k x n_r
 * array (columns in ascending r value).  An A without blocks fails with
 * ArgumentError. */
int sb_bs_left_qr(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t* Q, sb_tensor_t* R);
/* left_factorize (tensor_wrappers.jl:954-970) below its ITensor layer: Q as
 * above and R scattered into a dense k x d_r array (column r of the r value r,
 * zero where A has no block with that r) */
int sb_bs_left_factorize(sb_ctx_t ctx, sb_tensor_t A, sb_tensor_t* Q, sb_tensor_t* R);
/* New (the sector-format counterpart of blocksparse_left_qr, as
 * sb_bs_svd_sectors is of blocksparse_svd): QR of a site tensor
 * A [Ql, Qr | s, l, r] sector by sector.  For right sector ir the matrix has
 * rows (Ql, s, l) over A's blocks (il, ir) with rowsec[il + nl*s] == ir
 * (ascending il, s, then l < dl[il]) and columns r < dr[ir]; one thin QR per
 * sector.  Q has A's block list and the k = min(rows, dr) columns of every
 * sector's factor (kept[ir] = k); R [Qr, Qr | m, r] holds the k x dr factors
 * on its diagonal keys. */
int sb_bs_qr_sectors(sb_ctx_t ctx, sb_tensor_t A, const int64_t* rowsec, const int64_t* dl, const int64_t* dr,
                     sb_tensor_t* Q, sb_tensor_t* R, int64_t* kept);

#ifdef __cplusplus
}
#endif

#endif
