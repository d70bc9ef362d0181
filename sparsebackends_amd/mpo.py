"""MPO products on COO and block-sparse storage, site by site, with the two
factors' links fused into one bond per storage region afterwards:

* sandwich_site: one site of P·H·P for a projector MPO P around a
  Hamiltonian MPO H, as the reference's sandwich_mpo builds it
  (test.jl: H1 = contract(P'', H', :coo, :dense); contract(P, H1, :coo,
  :blocksparse)): P (COO) x H (dense) -> block-sparse (contract_coo_dense.jl),
  then P (COO) x that (contract_coo_bs.jl, reduced index in the dense tail);
* mul_site: one site of an MPO x MPO product on COO (test.jl mulMPO,
  contract(A, B', :coo, :coo) -> contract_coo.jl);
* the link bookkeeping of contract_and_fuse_links
  (tensor_contraction.jl:352-412): the legs of one bond that sit in the same
  region (sparse head or dense tail of a block-sparse result; anywhere in a
  COO result) are fused into the first of them (fuse_two_axes!,
  blocksparse/conversions.jl:299-314; fuse_axes_coo, coo/conversions.jl:174-223),
  a leg alone in its region stays as it is (_canonicalize_bond_blocksparse!,
  :287-336).

Site tensors are [left link, s_out, s_in, right link] (the MPO layout of
heff.py's W).  Fused bond index = first leg + dim(first leg) * (second leg - 1)
(1-based; the convention of fuse_sparse_axes and fuse_axes_coo).
"""
from __future__ import annotations

import numpy as np

from .storage import NewBlockSparseSorted, COOTensor, coo_from_dense
from .tensoralgebra import contract_


def _lab(name):
    return (int.from_bytes(("m:" + name).encode()[:8].ljust(8, b"\0"), "little"), 0)


def sandwich_site(P, H, device=0):
    """P·H·P at one site.  P: dense projector site tensor [a, s', s, a'];
    H: dense Hamiltonian site tensor [b, s', s, b'].  Returns the block-sparse
    site tensor [L, s_out, s_in, R | b, b'] with L = (c, a), R = (c', a') the
    fused links of the two projector factors (head) and b, b' the
    Hamiltonian's links (tail)."""
    P = np.asarray(P, dtype=np.float64)
    H = np.asarray(H, dtype=np.float64)
    wP, d = P.shape[0], P.shape[1]
    wH = H.shape[0]
    Pc = coo_from_dense(P, atol=0.0, device=device)
    L = _lab
    # P (COO) [c, so, t, c'] x H (dense) [b, t, u, b'] -> X [c, so, c' | b, u, b']
    X = NewBlockSparseSorted.new((wP, d, wP, wH, d, wH), 3, device=device)
    contract_(X, [L("c"), L("so"), L("c2"), L("b"), L("u"), L("b2")], Pc, [L("c"), L("so"), L("t"), L("c2")],
              H, [L("b"), L("t"), L("u"), L("b2")])
    # P (COO) [a, u, si, a'] x X over u (in X's dense tail) -> Y [c, a, so, si, c', a' | b, b']
    Y = NewBlockSparseSorted.new((wP, wP, d, d, wP, wP, wH, wH), 2, device=device)
    contract_(Y, [L("c"), L("a"), L("so"), L("si"), L("c2"), L("a2"), L("b"), L("b2")], Pc,
              [L("a"), L("u"), L("si"), L("a2")], X, [L("c"), L("so"), L("c2"), L("b"), L("u"), L("b2")])
    # left bond: c and a share the head -> fused; b alone in the tail.  Same on the right.
    Y = Y.fuse_two_axes_(1, 2)   # [(c,a), so, si, c', a' | b, b']
    Y = Y.fuse_two_axes_(4, 5)   # [(c,a), so, si, (c',a') | b, b']
    return Y


def mul_site(A, B, device=0):
    """A x B at one site on COO: A [a, s', t, a'] and B [b, t, s, b'] dense
    site tensors, converted to COO; the product keeps every term (COO x COO
    does not coalesce) and its links are fused: [(a,b), s', s, (a',b')]."""
    Ac = coo_from_dense(np.asarray(A, dtype=np.float64), atol=0.0, device=device)
    Bc = coo_from_dense(np.asarray(B, dtype=np.float64), atol=0.0, device=device)
    wA, d = A.shape[0], A.shape[1]
    wB = B.shape[0]
    L = _lab
    C = COOTensor.new((wA, wB, d, d, wA, wB), device=device)
    contract_(C, [L("a"), L("b"), L("so"), L("si"), L("a2"), L("b2")], Ac, [L("a"), L("so"), L("t"), L("a2")],
              Bc, [L("b"), L("t"), L("si"), L("b2")])
    C = C.fuse_axes(1, 2)   # [(a,b), so, si, a', b']
    C = C.fuse_axes(4, 5)   # [(a,b), so, si, (a',b')]
    return C


# ---------------------------------------------------------------- models


def pxp_projector(d=2):
    """Rydberg-blockade projector MPO site tensor [a, s', s, a'] (w = 2): the
    link carries whether the previous site is excited (state 1); a site may be
    excited only if its left neighbour is not"""
    P = np.zeros((2, d, d, 2))
    for a in range(2):
        for s in range(d):
            if a == 1 and s == 1:
                continue
            P[a, s, s, s] = 1.0
    return P


def x_field(d=2, h=1.0):
    """sum_i h X_i as an MPO site tensor (w = 2, link 1 = start, 0 = done)"""
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    W = np.zeros((2, d, d, 2))
    W[0, :, :, 0] = np.eye(d)
    W[1, :, :, 1] = np.eye(d)
    W[1, :, :, 0] = h * X
    return W
