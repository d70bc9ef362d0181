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

struct LeftQrResult {
  TP Q, R;                      // Q {3, N2 = 2} keyed by l; R dense (k x n_r), or (k x d_r) from left_factorize
  std::vector<i64> l_vals, r_vals;
};
// blocksparse_left_qr (SparseBackends/src/ops_factorize.jl:8-43) of a NewBlockSparseSorted{T,3,1,2}
LeftQrResult bs_left_qr(Ctx*, const Tensor& A);
// left_factorize (tensor_wrappers.jl:954-970) below the ITensor layer: R scattered to k x d_r
LeftQrResult bs_left_factorize(Ctx*, const Tensor& A);

struct SectQrResult {
  TP Q, R;                  // Q with A's block list; R [Qr, Qr | m, r] on the diagonal keys
  std::vector<i64> kept;    // new extent of every right sector (0: no rows)
};
// sector QR of a site tensor [Ql, Qr | s, l, r]: rowsec[il + nl s] = the right
// sector of (Ql, s) or -1; dl, dr the sector extents
SectQrResult bs_qr_sectors(Ctx*, const Tensor& A, const i64* rowsec, const i64* dl, const i64* dr);

}  // namespace sb
