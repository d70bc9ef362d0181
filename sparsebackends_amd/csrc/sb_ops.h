// Storage operations on device tensors (used by the C ABI and by contract).
#pragma once

#include <memory>

#include "sb_kernels.h"

namespace sb {

using TP = std::unique_ptr<Tensor>;

TP bs_new(Ctx*, const std::vector<i64>& dims, int N2);   // public constructor (0 < N2)
TP bs_make(Ctx*, const std::vector<i64>& dims, int N2);  // inner constructor (0 <= N2)
TP dense_new(Ctx*, const std::vector<i64>& dims);
TP al_new(Ctx*, const std::vector<i64>& dims, int N2);    // AliasedBlockSparse (0 < N2)
double al_getindex(Ctx*, const Tensor& A, const i64* I);
TP al_to_blocksparse(Ctx*, const Tensor& A);
// _check_no_cross_perm (storage.jl:126-132)
void check_no_cross_perm(const std::vector<i64>& perm, int N, int P);
TP coo_new(Ctx*, const std::vector<i64>& dims);
// permutedims(::NewBlockSparseSorted, perm) (blocksparse/storage.jl:188-264); perm 1-based
TP bs_permute(Ctx*, const Tensor& A, const std::vector<i64>& perm);
// permutedims!(...) (storage.jl:271-392): requires packed ids
void bs_permute_inplace(Ctx*, Tensor& A, const std::vector<i64>& perm);
// Base.permutedims of a dense array; perm 1-based
TP dense_permute(Ctx*, const Tensor& B, const std::vector<i64>& perm);
// permutedims(::COOTensor, perm) (coo/storage.jl:153-166): key remap, marks dirty
TP coo_permute(Ctx*, const Tensor& A, const std::vector<i64>& perm);
// sort_coo! (coo/storage.jl:182-189): stable column-major sort when dirty
void coo_sort(Ctx*, Tensor& A);
// permutedims + sort! in one: returns the permuted, column-major sorted keys
// (values allocated, not filled) and idx with out entry i = in entry idx[i]
TP coo_permute_sorted(Ctx*, const Tensor& A, const std::vector<i64>& perm, DBuf<i32>& idx);
// sort!(::NewBlockSparseSorted) (storage.jl:403-410): lexicographic keys, ids follow
void bs_sort_lex(Ctx*, Tensor& A);
bool bs_is_dense(const Tensor& A);
TP bs_from_dense(Ctx*, const Tensor& D, int N2, double atol, double rtol, bool dropzeros);
TP bs_to_dense(Ctx*, const Tensor& B, double init, const std::vector<std::pair<i64, i64>>& groups);
TP coo_from_dense(Ctx*, const Tensor& D, double atol, double rtol);
TP coo_to_dense(Ctx*, const Tensor& A);
TP bs_fuse_two_axes(Ctx*, Tensor& A, i64 ax1, i64 ax2);
TP coo_fuse_axes(Ctx*, const Tensor& A, i64 ax1, i64 ax2);
double bs_getindex(Ctx*, const Tensor& A, const i64* I);
void bs_setindex(Ctx*, Tensor& A, double v, const i64* I);
double coo_getindex(Ctx*, const Tensor& A, const i64* I);
void coo_setindex(Ctx*, Tensor& A, double v, const i64* I, bool dirty_kw);
// block offsets (elements) of all blocks: (ids[i]-1)*blksize
DBuf<i64> block_offsets(Ctx*, const Tensor& A);

}  // namespace sb
