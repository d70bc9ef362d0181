// Krylov vector operations on block-sparse tensors and the Lanczos
// eigensolver over a chain of contraction plans (sb_krylov.cpp).
#pragma once

#include <vector>

#include "sb_batch.h"
#include "sb_contract.h"

namespace sb {

double bs_norm(Ctx*, const Tensor& A);
TP bs_copy(Ctx*, const Tensor& A);
TP bs_similar(Ctx*, const Tensor& A);
TP bs_scale(Ctx*, const Tensor& A, double alpha);
void bs_fill(Ctx*, Tensor& A, double v);
void bs_axpby(Ctx*, double alpha, const Tensor& A, double beta, Tensor& R);
TP bs_add(Ctx*, const Tensor& A, const Tensor& B);
double bs_dot(Ctx*, const Tensor& A, const Tensor& B);
TP bs_restrict(Ctx*, const Tensor& A, const Tensor& B);

// allocation-free variants (sb_noalloc.hip): noalloc_setup finds the device's
// scratch array once per device (that device current); bs_norm_noalloc writes
// norm(A.data) into out[0], device memory, asynchronously
void noalloc_setup(int device);
void bs_norm_noalloc(Ctx*, const Tensor& A, double* out);

// eigenvalues (ascending) and eigenvectors (columns) of a symmetric
// column-major m x m matrix (Householder tridiagonalisation + implicit QL)
void sym_eig(int m, const std::vector<double>& A, std::vector<double>& w, std::vector<double>& U);

struct EigResult {
  double value = 0;           // lowest Ritz value
  std::vector<double> ritz;   // Ritz values of the last projected matrix, ascending
  double normres = 0;         // residual norm of the returned Ritz pair
  i64 numops = 0;             // operator applications
  int numiter = 0;            // Krylov factorizations (1 + restarts)
  bool converged = false;
};

// Lowest eigenpair of the Hermitian operator y = plans(x): the first plan
// reads x, the last writes y; y's blocks are matched to x's by key and those
// whose key x lacks are left out (the solve stays in x's block space).
// On return x holds the normalised Ritz vector.
EigResult eigsolve_lanczos(Ctx*, const std::vector<const Plan*>& plans, Tensor& x, Tensor& y, int krylovdim,
                           int maxiter, double tol);

// The same solve for every chain c of a batch (operator: chain c, from xs[c]
// to ys[c]), in lock step: one batched operator application per iteration for
// the chains still running and one host synchronisation for all of them.
// Each result (and x) is bit for bit what eigsolve_lanczos gives alone.
std::vector<EigResult> eigsolve_lanczos_batch(Ctx*, Batch& batch, const std::vector<Tensor*>& xs,
                                              const std::vector<Tensor*>& ys, int krylovdim, int maxiter, double tol);

}  // namespace sb
