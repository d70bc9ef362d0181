"""The values of contract_prefix_outer_bd! (BS x Dense, the reduced label in
A's prefix), restated in numpy from the reference code alone, independently
of the C++ oracle: the summation order and the compensated (Kahan) update
decide the last bits of every output entry, so a kernel that sums the same
products in another order differs from this model bit for bit on inexact
inputs.

The rules restated (SparseBackends/src/tensoralgebra/contract_bs_dense.jl):
* A is permuted so the reduced label r is its last prefix axis (:56-58): the
  other prefix axes keep their order, so A's keys sort with r most
  significant;
* each A block's C key is its prefix restricted to C's prefix labels, in C's
  order (:105-113, :126-130); `sortperm` of these keys (tuples compared
  lexicographically; Julia's sortperm is stable, so equal keys keep A's
  order) gives the groups, one C block per group in that order (:133-147);
* within a group, block k contributes outer(A_k, B[:, r_k]) with B's r axis
  moved last and the rest flattened column-major (:61-69, :150-160);
* every entry is updated in k order by y = a*b - c; t = C + y;
  c = (t - C) - y; C = t, with c = 0 at the start of each group (:163-190);
  the product 1*a*b equals a*b in IEEE arithmetic.
numpy applies each operation elementwise with one rounding and no fused
multiply-add, which is the reference's operation sequence.
"""
import numpy as np


def _blocks(A, P):
    """blocksparse_from_dense: prefix keys (1-based) of the blocks holding a
    nonzero, with their column-major tail data"""
    tail = A.shape[P:]
    out = {}
    for idx in np.ndindex(*A.shape[:P]):
        blk = A[idx]
        if np.any(blk != 0):
            out[tuple(i + 1 for i in idx)] = np.asarray(blk, dtype=np.float64).reshape(-1, order="F") \
                if tail else np.asarray([blk], dtype=np.float64)
    return out


def prefix_outer_bd(A, la, N2A, B, lb, lc, N2C):
    """dict C key -> C block (column-major data) of contract!(C, lc, A, la, B, lb)
    through contract_prefix_outer_bd!"""
    PA = len(la) - N2A
    PC = len(lc) - N2C
    shared = [l for l in la if l in lb]
    assert len(shared) == 1 and la.index(shared[0]) < PA
    r = shared[0]
    blocks = _blocks(A, PA)
    pre = list(la[:PA])
    ra = pre.index(r)
    order = [i for i in range(PA) if i != ra] + [ra]  # r last, the others in their order
    pre_p = [pre[i] for i in order]
    # A's keys after the permute, sorted column-major (last axis most significant)
    items = sorted(((tuple(k[i] for i in order), v) for k, v in blocks.items()), key=lambda kv: kv[0][::-1])
    # B with r last, flattened column-major: column r is B's slice at r
    br = lb.index(r)
    Bp = np.moveaxis(B, br, -1)
    R = Bp.shape[-1]
    Bmat = np.asarray(Bp, dtype=np.float64).reshape(-1, R, order="F")
    src = [pre_p.index(l) for l in lc[:PC]]
    a_dense = la[PA:]
    b_wo = [l for l in lb if l != r]
    ab = list(lc[PC:]) == list(a_dense) + b_wo
    assert ab or list(lc[PC:]) == b_wo + list(a_dense)
    ckeys = [tuple(k[j] for j in src) for k, _ in items]
    perm = sorted(range(len(items)), key=lambda i: ckeys[i])  # stable
    out = {}
    i = 0
    while i < len(perm):
        ck = ckeys[perm[i]]
        j = i
        while j < len(perm) and ckeys[perm[j]] == ck:
            j += 1
        na = len(items[perm[i]][1])
        C = np.zeros((na, Bmat.shape[0]))
        comp = np.zeros_like(C)
        for k in range(i, j):
            key, a = items[perm[k]]
            b = Bmat[:, key[-1] - 1]
            y = np.multiply.outer(a, b) - comp
            t = C + y
            comp = (t - C) - y
            C = t
        out[ck] = C.reshape(-1, order="F") if ab else C.T.reshape(-1, order="F")
        i = j
    return out


def blocks_of(parts):
    """dict key -> block data of a block-sparse result's parts"""
    bs = int(parts["blksize"])
    return {tuple(int(x) for x in k): parts["data"][(int(i) - 1) * bs:int(i) * bs]
            for k, i in zip(parts["keys"], parts["ids"])}
