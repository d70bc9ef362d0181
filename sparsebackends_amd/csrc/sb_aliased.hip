// AliasedBlockSparse on the device (aliased/storage.jl:46-59): block i is
// scalars[i] * templates[alias_ids[i]].  Kernels of the aliased contraction
// families (contract_aliased.jl, contract_aliased_shared.jl,
// contract_aliased_coo_dense.jl) that are not plain block products:
// contribution bookkeeping, the alias / accumulate resolution of every output
// key (_aliased_contribute!, contract_aliased.jl:75-114) and the exact-order
// accumulation of demoted blocks.  Compiled without floating-point contraction:
// the accumulated blocks and summed scalars repeat the reference's IEEE
// operations one for one.
#include "sb_kernels.h"

namespace sb {

static inline unsigned nb256a(i64 n) { return (unsigned)((n + 255) / 256); }
#define LAUNCH_A(kern, n, ...)                                                                   \
  do {                                                                                           \
    if ((n) > 0) {                                                                               \
      hipLaunchKernelGGL(kern, dim3(nb256a(n)), dim3(256), 0, ctx->stream, __VA_ARGS__);        \
      HIP_CHECK(hipGetLastError());                                                              \
    }                                                                                            \
  } while (0)

// out[i*bs + j] = scalars[i] * templates[(ids[i]-1)*bs + j]  (to_blocksparse, conversions.jl:18-32)
__global__ void al_expand_kernel(i64 n, i64 bs, const i32* __restrict__ ids, const double* __restrict__ sc,
                                 const double* __restrict__ tm, double* __restrict__ out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * bs) return;
  i64 i = t / bs, j = t - i * bs;
  out[t] = sc[i] * tm[(i64)(ids[i] - 1) * bs + j];
}
void k_al_expand(Ctx* ctx, i64 n, i64 bs, const i32* ids, const double* sc, const double* tm, double* out) {
  LAUNCH_A(al_expand_kernel, n * bs, n, bs, ids, sc, tm, out);
}

// combination key of contribution p: x = idsA[pa[p]]; y = (mode 0) keysA[pa[p]*wA + col],
// (mode 1) idsB[pb[p]], (mode 2) 1 + sum_t (keysA[pa[p]*wA + cols[t]] - 1) * str[t]
__global__ void al_xy_kernel(int mode, const i32* __restrict__ pa, const i32* __restrict__ pb, i64 np,
                             const i32* __restrict__ idsA, const i32* __restrict__ keysA, int wA, int col,
                             const i32* __restrict__ idsB, SArr cols, IArr str, int ncols, i32* __restrict__ xy) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const i32 a = pa[p];
  i32 y;
  if (mode == 0)
    y = keysA[(i64)a * wA + col];
  else if (mode == 1)
    y = idsB[pb[p]];
  else {
    i64 l = 1;
    for (int t = 0; t < ncols; ++t) l += (i64)(keysA[(i64)a * wA + cols.v[t]] - 1) * str.v[t];
    y = (i32)l;
  }
  xy[2 * p] = idsA[a];
  xy[2 * p + 1] = y;
}
void k_al_xy(Ctx* ctx, int mode, const i32* pa, const i32* pb, i64 np, const i32* idsA, const i32* keysA, int wA,
             int col, const i32* idsB, const int* cols, const i64* str, int ncols, i32* xy) {
  SArr c{};
  IArr s{};
  for (int t = 0; t < ncols; ++t) {
    c.v[t] = cols[t];
    s.v[t] = str[t];
  }
  LAUNCH_A(al_xy_kernel, np, mode, pa, pb, np, idsA, keysA, wA, col, idsB, c, s, ncols, xy);
}

// lin of a 2-column key (x, y) over (nx, ny), 1-based entries
__global__ void al_xy_lin_kernel(const i32* __restrict__ xy, i64 np, i64 nx, u64* __restrict__ lin) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  lin[p] = (u64)(xy[2 * p] - 1) + (u64)nx * (u64)(xy[2 * p + 1] - 1);
}
void k_al_xy_lin(Ctx* ctx, const i32* xy, i64 np, i64 nx, u64* lin) { LAUNCH_A(al_xy_lin_kernel, np, xy, np, nx, lin); }

// per unique key s (ids[s] = its first-seen rank): every member p of segment s
// gets val[order[q]] = ids[s]; the unique key's columns go to slot ids[s]-1
__global__ void al_seg_ids_kernel(i64 ns, const i32* __restrict__ seg, const i32* __restrict__ order,
                                  const i32* __restrict__ ids, const i32* __restrict__ ukeys, int w,
                                  i32* __restrict__ val, i32* __restrict__ byid) {
  i64 s = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const i32 id = ids[s];
  for (i32 q = seg[s]; q < seg[s + 1]; ++q) val[order[q]] = id;
  if (byid)
    for (int c = 0; c < w; ++c) byid[(i64)(id - 1) * w + c] = ukeys[s * w + c];
}
void k_al_seg_ids(Ctx* ctx, i64 ns, const i32* seg, const i32* order, const i32* ids, const i32* ukeys, int w,
                  i32* val, i32* byid) {
  LAUNCH_A(al_seg_ids_kernel, ns, ns, seg, order, ids, ukeys, w, val, byid);
}

// off[t] = (u[t*w + c] - 1) * scale + add
__global__ void al_offsets_kernel(const i32* __restrict__ u, int w, int c, i64 n, i64 scale, i64 add,
                                  i64* __restrict__ off) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  off[t] = (i64)(u[t * w + c] - 1) * scale + add;
}
void k_al_offsets(Ctx* ctx, const i32* u, int w, int c, i64 n, i64 scale, i64 add, i64* off) {
  LAUNCH_A(al_offsets_kernel, n, u, w, c, n, scale, add, off);
}

// Resolution of every output key g (its contributions order[seg[g] .. seg[g+1]) in
// traversal order): aliased while all share one combined template (demote_any:
// any second contribution demotes, contract_aliased_coo_dense.jl:96-118);
// first_diff = position of the contribution that demotes it.
__global__ void al_groups_kernel(i64 ng, const i32* __restrict__ seg, const i32* __restrict__ order,
                                 const i32* __restrict__ ctid, int demote_any, i32* __restrict__ is_acc,
                                 i32* __restrict__ first_diff, i32* __restrict__ aid) {
  i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const i32 q0 = seg[g], q1 = seg[g + 1];
  const i32 t0 = ctid ? ctid[order[q0]] : 0;
  i32 fd = 0;
  for (i32 q = q0 + 1; q < q1; ++q)
    if (demote_any || ctid[order[q]] != t0) {
      fd = q - q0;
      break;
    }
  is_acc[g] = fd > 0 ? 1 : 0;
  first_diff[g] = fd;
  aid[g] = t0;
}
void k_al_groups(Ctx* ctx, i64 ng, const i32* seg, const i32* order, const i32* ctid, int demote_any, i32* is_acc,
                 i32* first_diff, i32* aid) {
  LAUNCH_A(al_groups_kernel, ng, ng, seg, order, ctid, demote_any, is_acc, first_diff, aid);
}

// alias ids of the output blocks: accumulated block g -> base + rank[g] + 1
__global__ void al_final_ids_kernel(i64 ng, const i32* __restrict__ is_acc, const i64* __restrict__ rank,
                                    const i32* __restrict__ aid, i64 base, i32* __restrict__ out) {
  i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  out[g] = is_acc[g] ? (i32)(base + rank[g] + 1) : aid[g];
}
void k_al_final_ids(Ctx* ctx, i64 ng, const i32* is_acc, const i64* rank, const i32* aid, i64 base, i32* out) {
  LAUNCH_A(al_final_ids_kernel, ng, ng, is_acc, rank, aid, base, out);
}

// list of the accumulated groups in key order
__global__ void al_acc_list_kernel(i64 ng, const i32* __restrict__ is_acc, const i64* __restrict__ rank,
                                   i32* __restrict__ list) {
  i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  if (is_acc[g]) list[rank[g]] = (i32)g;
}
void k_al_acc_list(Ctx* ctx, i64 ng, const i32* is_acc, const i64* rank, i32* list) {
  LAUNCH_A(al_acc_list_kernel, ng, ng, is_acc, rank, list);
}

// alpha[p] = sA[pa[p]] * sB[pb[p]]  (sB == null: sA[pa[p]])
__global__ void al_alpha_kernel(const i32* __restrict__ pa, const i32* __restrict__ pb, i64 np,
                                const double* __restrict__ sA, const double* __restrict__ sB,
                                double* __restrict__ alpha) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  alpha[p] = sB ? sA[pa[p]] * sB[pb[p]] : sA[pa[p]];
}
void k_al_alpha(Ctx* ctx, const i32* pa, const i32* pb, i64 np, const double* sA, const double* sB, double* alpha) {
  LAUNCH_A(al_alpha_kernel, np, pa, pb, np, sA, sB, alpha);
}

// scalar of every output block: the in-order sum of its contributions' alphas
// when aliased (contract_aliased.jl:94-96), 1 when accumulated
__global__ void al_scalars_kernel(i64 ng, const i32* __restrict__ seg, const i32* __restrict__ order,
                                  const double* __restrict__ alpha, const i32* __restrict__ is_acc,
                                  double* __restrict__ sc) {
  i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  if (is_acc[g]) {
    sc[g] = 1.0;
    return;
  }
  const i32 q0 = seg[g], q1 = seg[g + 1];
  double s = alpha[order[q0]];
  for (i32 q = q0 + 1; q < q1; ++q) s = s + alpha[order[q]];
  sc[g] = s;
}
void k_al_scalars(Ctx* ctx, i64 ng, const i32* seg, const i32* order, const double* alpha, const i32* is_acc,
                  double* sc) {
  LAUNCH_A(al_scalars_kernel, ng, ng, seg, order, alpha, is_acc, sc);
}

// Accumulated block of output key list[a], element e, in the reference's order:
// the scalar of the leading run of one template, then the demoting
// contribution (acc = alpha_prev * T0 + alpha_j * Tj), then acc += alpha * T for
// the rest (contract_aliased.jl:84-112).  T of contribution c is
// base[toff[c] + e].
__global__ void al_accumulate_kernel(i64 nacc, i64 bs, const i32* __restrict__ list, const i32* __restrict__ seg,
                                     const i32* __restrict__ order, const i32* __restrict__ first_diff,
                                     const double* __restrict__ alpha, const double* __restrict__ base,
                                     const i64* __restrict__ toff, double* __restrict__ out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nacc * bs) return;
  const i64 a = t / bs, e = t - a * bs;
  const i32 g = list[a];
  const i32 q0 = seg[g], q1 = seg[g + 1], j = q0 + first_diff[g];
  double ap = alpha[order[q0]];
  for (i32 q = q0 + 1; q < j; ++q) ap = ap + alpha[order[q]];
  const i32 c0 = order[q0], cj = order[j];
  double acc = ap * base[toff[c0] + e] + alpha[cj] * base[toff[cj] + e];
  for (i32 q = j + 1; q < q1; ++q) {
    const i32 c = order[q];
    acc = acc + alpha[c] * base[toff[c] + e];
  }
  out[a * bs + e] = acc;
}
void k_al_accumulate(Ctx* ctx, i64 nacc, i64 bs, const i32* list, const i32* seg, const i32* order,
                     const i32* first_diff, const double* alpha, const double* base, const i64* toff, double* out) {
  LAUNCH_A(al_accumulate_kernel, nacc * bs, nacc, bs, list, seg, order, first_diff, alpha, base, toff, out);
}

// toff[p] = (ctid[p] - 1) * bs (+ add)
__global__ void al_ctid_off_kernel(const i32* __restrict__ ctid, i64 np, i64 bs, i64* __restrict__ toff) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  toff[p] = (i64)(ctid[p] - 1) * bs;
}
void k_al_ctid_off(Ctx* ctx, const i32* ctid, i64 np, i64 bs, i64* toff) {
  LAUNCH_A(al_ctid_off_kernel, np, ctid, np, bs, toff);
}

// small integer helpers of the aliased bookkeeping
__global__ void i32_to_u32_kernel(const i32* __restrict__ a, i64 n, unsigned* __restrict__ b) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (unsigned)a[i];
}
void k_i32_to_u32(Ctx* ctx, const i32* a, i64 n, unsigned* b) { LAUNCH_A(i32_to_u32_kernel, n, a, n, b); }
__global__ void i32_to_i64_kernel(const i32* __restrict__ a, i64 n, i64* __restrict__ b) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
void k_i32_to_i64(Ctx* ctx, const i32* a, i64 n, i64* b) { LAUNCH_A(i32_to_i64_kernel, n, a, n, b); }
// out[i] = keys[idx[i]*w + col0]
__global__ void take_col_kernel(const i32* __restrict__ keys, int w, int col0, const i32* __restrict__ idx, i64 n,
                                i32* __restrict__ out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = keys[(i64)idx[i] * w + col0];
}
void k_take_col(Ctx* ctx, const i32* keys, int w, int col0, const i32* idx, i64 n, i32* out) {
  LAUNCH_A(take_col_kernel, n, keys, w, col0, idx, n, out);
}
// lin[i] = v[i] - 1 (1-based values as a sort key)
__global__ void rv_lin_kernel(const i32* __restrict__ v, i64 n, u64* __restrict__ lin) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lin[i] = (u64)(v[i] - 1);
}
void k_rv_lin(Ctx* ctx, const i32* v, i64 n, u64* lin) { LAUNCH_A(rv_lin_kernel, n, v, n, lin); }
// out[idx[i]] = vals[i]
__global__ void scatter_i32_kernel(const i32* __restrict__ idx, const i32* __restrict__ vals, i64 n,
                                   i32* __restrict__ out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = vals[i];
}
void k_scatter_i32(Ctx* ctx, const i32* idx, const i32* vals, i64 n, i32* out) {
  LAUNCH_A(scatter_i32_kernel, n, idx, vals, n, out);
}

}  // namespace sb
