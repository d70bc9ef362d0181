// Two-site block-diagonal SVD with truncation (sb_factorize.cpp).
#pragma once

#include <array>
#include <vector>

#include "sb_ops.h"

namespace sb {

struct SvdResult {
  TP U, SV;                  // U (n_L, k, d1) keys (L, nb); SV (k, n_R, d2) keys (nb, R); N2 = 1
  std::vector<double> svs;   // kept singular values in new-bond order
  double truncerr = 0;       // sum of the discarded singular values squared
};

// blocksparse_svd (SparseBackends/src/ops_factorize.jl:50-148) of a
// NewBlockSparseSorted{T,4,2,2}; right: ortho "right" (else "left"); maxdim < 0: no limit
SvdResult bs_svd(Ctx*, const Tensor& A, bool right, i64 maxdim, i64 mindim, double cutoff);

}  // namespace sb
