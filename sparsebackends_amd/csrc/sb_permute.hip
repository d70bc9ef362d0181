// Permute-into-GEMM-layout kernels: the dense-tail gather of
// permutedims(::NewBlockSparseSorted) (SparseBackends/src/blocksparse/storage.jl:217-262)
// and the dense N-D permutedims of the dense operand (contract_bs_dense.jl:67,280,493).
//
// One launch moves a batch of blocks.  Three shapes:
//  * identity tail: block copy (whole-block gather), coalesced both sides;
//  * fastest axis kept: one thread per output element, reads and writes both
//    run along the shared fastest axis;
//  * fastest axis moved: 32x32 tiles over (input-fastest, output-fastest)
//    staged through LDS so both the HBM read and the HBM write are coalesced.
#include "sb_kernels.h"

namespace sb {

struct FArr {
  float v[MAXR];
};

struct PermArgs {
  int nd;
  int dims[MAXR];   // old (input) order
  i64 istr[MAXR];   // input stride of old axis d
  i64 ostr[MAXR];   // output stride of old axis d
};

__global__ void gather_blocks_kernel(const double* __restrict__ in, const i64* __restrict__ src_off, i64 nb, i64 bs,
                                     double* __restrict__ out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  i64 stride = (i64)gridDim.x * blockDim.x;
  i64 tot = nb * bs;
  for (; t < tot; t += stride) {
    i64 o = t / bs, e = t - o * bs;
    out[t] = in[(src_off ? src_off[o] : o * bs) + e];
  }
}

// output-order decode: new axis j has old axis pj[j]; the thread owns one output element
__global__ void permute_elem_kernel(const double* __restrict__ in, const i64* __restrict__ src_off, i64 nb, i64 bs,
                                    PermArgs pa, IArr newdims, IArr pj, double* __restrict__ out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  i64 stride = (i64)gridDim.x * blockDim.x;
  i64 tot = nb * bs;
  for (; t < tot; t += stride) {
    i64 o = t / bs;
    i64 r = t - o * bs;
    i64 src = src_off ? src_off[o] : o * bs;
    for (int j = 0; j < pa.nd; ++j) {
      i64 d = newdims.v[j];
      i64 q = r / d;
      i64 c = r - q * d;
      r = q;
      src += c * pa.istr[pj.v[j]];
    }
    out[t] = in[src];
  }
}

// q = n / d for 0 <= n < 2^24, exactly: the float quotient is within one of
// the true one, then corrected with the remainder
__device__ inline int div_small(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  int r = n - q * d;
  if (r >= d) ++q;
  else if (r < 0) --q;
  return q;
}

// many blocks of a modest size: one workgroup per block (no division to find
// the block), the in-block index decoded with 32-bit exact division; when the
// fastest axis moves, a block of at most kStageMax values is read contiguously
// into LDS first, so both the HBM read and the HBM write are coalesced
constexpr i64 kStageMax = 8192;  // 64 KB of LDS
__global__ __launch_bounds__(256) void permute_block_kernel(const double* __restrict__ in,
                                                            const i64* __restrict__ src_off, i64 bs, PermArgs pa,
                                                            IArr newdims, IArr pj, FArr inv, int stage,
                                                            double* __restrict__ out) {
  extern __shared__ double blk[];  // the whole input block when it fits (bs <= kStageMax)
  const i64 o = blockIdx.x;
  const i64 src0 = src_off ? src_off[o] : o * bs;
  double* dst = out + o * bs;
  const bool staged = stage != 0;
  if (staged) {  // one contiguous read of the block, then the permuted order out of LDS
    for (int t = threadIdx.x; t < (int)bs; t += 256) blk[t] = in[src0 + t];
    __syncthreads();
  }
  for (int t = threadIdx.x; t < (int)bs; t += 256) {
    int r = t;
    i64 src = 0;
    for (int j = 0; j < pa.nd; ++j) {
      const int d = (int)newdims.v[j];
      const int q = div_small(r, d, inv.v[j]);
      src += (i64)(r - q * d) * pa.istr[pj.v[j]];
      r = q;
    }
    dst[t] = staged ? blk[src] : in[src0 + src];
  }
}

// tile over (a = input fastest axis 0, b = input axis that becomes output fastest)
constexpr int TT = 32;
__global__ __launch_bounds__(256) void permute_tiled_kernel(const double* __restrict__ in,
                                                            const i64* __restrict__ src_off, i64 nb, i64 bs,
                                                            PermArgs pa, int b, i64 ntile_a, i64 ntile_b,
                                                            i64 nbatch_inner, double* __restrict__ out) {
  __shared__ double tile[TT][TT + 1];
  i64 w = blockIdx.x;
  i64 ntiles = ntile_a * ntile_b;
  i64 tid = w % ntiles;
  i64 bat = w / ntiles;
  i64 o = bat / nbatch_inner;
  i64 rest = bat - o * nbatch_inner;
  if (o >= nb) return;
  i64 ta = tid % ntile_a, tb = tid / ntile_a;
  // decode the other axes (old order, skipping 0 and b)
  i64 ibase = src_off ? src_off[o] : o * bs;
  i64 obase = o * bs;
  for (int d = 1; d < pa.nd; ++d) {
    if (d == b) continue;
    i64 q = rest / pa.dims[d];
    i64 c = rest - q * pa.dims[d];
    rest = q;
    ibase += c * pa.istr[d];
    obase += c * pa.ostr[d];
  }
  int Da = pa.dims[0], Db = pa.dims[b];
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  i64 ia0 = ta * TT, ib0 = tb * TT;
  for (int yy = ty; yy < TT; yy += 8) {
    i64 ia = ia0 + tx, ib = ib0 + yy;
    if (ia < Da && ib < Db) tile[yy][tx] = in[ibase + ia * pa.istr[0] + ib * pa.istr[b]];
  }
  __syncthreads();
  for (int yy = ty; yy < TT; yy += 8) {
    i64 ia = ia0 + yy, ib = ib0 + tx;
    if (ia < Da && ib < Db) out[obase + ia * pa.ostr[0] + ib * pa.ostr[b]] = tile[tx][yy];
  }
}

static unsigned grid_for(i64 tot) {
  i64 g = (tot + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  return (unsigned)(g < 1 ? 1 : g);
}

void gather_blocks(Ctx* ctx, const double* in, const i64* src_off, i64 nblocks, i64 bs, double* out) {
  if (nblocks * bs == 0) return;
  hipLaunchKernelGGL(gather_blocks_kernel, dim3(grid_for(nblocks * bs)), dim3(256), 0, ctx->stream, in, src_off,
                     nblocks, bs, out);
  HIP_CHECK(hipGetLastError());
}

void permute_blocks(Ctx* ctx, const double* in, const i64* src_off, i64 nblocks, int N2, const i64* tdims,
                    const int* perm, double* out) {
  i64 bs = 1;
  for (int d = 0; d < N2; ++d) bs *= tdims[d];
  if (nblocks == 0 || bs == 0) return;
  // drop unit axes and check for identity
  bool ident = true;
  for (int j = 0; j < N2; ++j) ident = ident && perm[j] == j;
  int nz = 0, map[MAXR];
  for (int d = 0; d < N2; ++d) map[d] = tdims[d] > 1 ? nz++ : -1;
  int pj2[MAXR], k = 0;
  for (int j = 0; j < N2; ++j)
    if (map[perm[j]] >= 0) pj2[k++] = map[perm[j]];
  bool ident2 = true;
  for (int j = 0; j < nz; ++j) ident2 = ident2 && pj2[j] == j;
  if (ident || ident2) {
    gather_blocks(ctx, in, src_off, nblocks, bs, out);
    return;
  }
  PermArgs pa{};
  pa.nd = nz;
  i64 s = 1;
  for (int d = 0, e = 0; d < N2; ++d) {
    if (map[d] < 0) continue;
    pa.dims[e] = (int)tdims[d];
    pa.istr[e] = s;
    s *= tdims[d];
    ++e;
  }
  IArr newdims{}, pj{};
  s = 1;
  for (int j = 0; j < nz; ++j) {
    pj.v[j] = pj2[j];
    newdims.v[j] = pa.dims[pj2[j]];
    pa.ostr[pj2[j]] = s;
    s *= newdims.v[j];
  }
  int b = pj2[0];
  if ((b == 0 || pa.dims[0] < 8 || pa.dims[b] < 8) && bs < (1 << 24) && nblocks >= 256 && nblocks < (1ll << 31)) {
    FArr inv{};
    for (int j = 0; j < nz; ++j) inv.v[j] = 1.0f / (float)newdims.v[j];
    // staged through LDS when the fastest axis moves (strided reads otherwise); a
    // kept fastest axis already reads contiguously
    const int stage = b != 0 && bs <= kStageMax;
    const size_t lds = stage ? (size_t)bs * sizeof(double) : 0;
    hipLaunchKernelGGL(permute_block_kernel, dim3((unsigned)nblocks), dim3(256), lds, ctx->stream, in, src_off, bs,
                       pa, newdims, pj, inv, stage, out);
  } else if (b == 0 || pa.dims[0] < 8 || pa.dims[b] < 8) {
    hipLaunchKernelGGL(permute_elem_kernel, dim3(grid_for(nblocks * bs)), dim3(256), 0, ctx->stream, in, src_off,
                       nblocks, bs, pa, newdims, pj, out);
  } else {
    i64 nta = (pa.dims[0] + TT - 1) / TT, ntb = (pa.dims[b] + TT - 1) / TT;
    i64 inner = bs / ((i64)pa.dims[0] * pa.dims[b]);
    i64 nwg = nta * ntb * inner * nblocks;
    SB_ASSERT(nwg < (1ll << 31), "permute: grid too large");
    hipLaunchKernelGGL(permute_tiled_kernel, dim3((unsigned)nwg), dim3(256), 0, ctx->stream, in, src_off, nblocks,
                       bs, pa, b, nta, ntb, inner, out);
  }
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
