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

struct SectSvdResult {
  TP A, B;                   // A [Ql, Qm | s1, l, m], B [Qm, Qr | s2, m, r]
  std::vector<i64> kept;     // kept singular values per fused sector
  std::vector<double> svs;   // kept singular values, new-bond order
  double truncerr = 0;
};

// Block-diagonal SVD of psi [Ql, Qr | s1, s2, l, r] by fused charge sector:
// rowsec[il + nl*s1] / colsec[ir + nr*s2] give the sector (0..nsec-1, or -1)
// of (Ql, s1) and (Qr, s2); dl / dr the actual extents of the bond sectors.
// Truncation as in bs_svd.  New bond: the sectors with kept values, ascending.
SectSvdResult bs_svd_sectors(Ctx*, const Tensor& psi, i64 nsec, const i64* rowsec, const i64* colsec, const i64* dl,
                             const i64* dr, bool right, i64 maxdim, i64 mindim, double cutoff);

}  // namespace sb
