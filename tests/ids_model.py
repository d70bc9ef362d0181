"""Block-creation order of every contraction family, restated from the
reference's traversals in plain Python, independently of the C++ oracle.

A NewBlockSparseSorted keeps its keys sorted while `_alloc_block!` gives a new
block the next payload slot (SparseBackends/src/blocksparse/storage.jl:67-83),
so C.ids is fixed by the order in which a kernel's traversal first calls
`_ensure_block!` on each C key.  `creation_order(...)` returns that sequence of
C keys for one contract! call, from the operands as dense arrays and their
labels; `ids_order(parts)` reads the same sequence off a result's (keys, ids).
For a COO result it returns the key sequence the reference pushes
(contract_coo.jl:129-166), which is the result's key list itself.

Every rule below cites the reference lines it restates.  Operands are rebuilt
here from dense arrays with numpy (blocksparse_from_dense keeps the blocks
holding a nonzero, keys in column-major order, conversions.jl:47-72;
coo_from_dense keeps the entries with |v| > atol, default 1e-12, in
column-major order, coo/conversions.jl:13-26), so nothing here reads the oracle's or the GPU's
intermediate tensors.
"""
import itertools

import numpy as np


def _cm_key(k):
    """column-major order of a 1-based key: last axis most significant
    (_prefix_lin, storage.jl:16-18)"""
    return tuple(reversed(k))


class BSop:
    """A block-sparse operand: labels, prefix length P, keys in storage order
    and the tail array of every block (indexed by key)."""

    def __init__(self, labels, P, keys, blocks):
        self.labels, self.P, self.keys, self.blocks = list(labels), P, list(keys), blocks

    @classmethod
    def from_dense(cls, arr, labels, P):
        arr = np.asarray(arr)
        keys, blocks = [], {}
        for idx in itertools.product(*[range(n) for n in arr.shape[:P]]):
            blk = arr[idx]
            if np.any(blk != 0):
                k = tuple(i + 1 for i in idx)
                keys.append(k)
                blocks[k] = np.asarray(blk)
        keys.sort(key=_cm_key)
        return cls(labels, P, keys, blocks)

    def permute(self, perm):
        """permutedims(A, perm) (storage.jl:188-269): prefix axes stay prefix,
        keys remapped and stable-sorted by the new column-major order (:207),
        tails transposed"""
        P = self.P
        pp = [p for p in perm[:P]]
        tp = [p - P for p in perm[P:]]
        keys = [tuple(k[p] for p in pp) for k in self.keys]
        blocks = {tuple(k[p] for p in pp): np.transpose(self.blocks[k], tp) for k in self.keys}
        order = sorted(range(len(keys)), key=lambda i: _cm_key(keys[i]))  # Python's sort is stable
        return BSop([self.labels[p] for p in perm], P, [keys[i] for i in order], blocks)

    def axis(self, lab):
        return self.labels.index(lab)


class COOop:
    """A COO operand: labels, entries (coordinate tuples) and values in storage order."""

    def __init__(self, labels, keys, vals):
        self.labels, self.keys, self.vals = list(labels), list(keys), list(vals)

    @classmethod
    def from_dense(cls, arr, labels, atol=1e-12):
        arr = np.asarray(arr)
        ent = [(tuple(int(i) + 1 for i in idx), float(arr[tuple(idx)])) for idx in zip(*np.nonzero(np.abs(arr) > atol))]
        ent.sort(key=lambda e: _cm_key(e[0]))
        return cls(labels, [e[0] for e in ent], [e[1] for e in ent])

    def permute_sorted(self, perm):
        """permutedims (coo/storage.jl:153-162) then sort! = sort_coo!, a stable
        sort by the column-major order (coo/storage.jl:182-189)"""
        keys = [tuple(k[p] for p in perm) for k in self.keys]
        order = sorted(range(len(keys)), key=lambda i: _cm_key(keys[i]))
        return COOop([self.labels[p] for p in perm], [keys[i] for i in order], [self.vals[i] for i in order])


def _r_last_prefix(op, rlab):
    """_permute_r_to_last_prefix (contract_bs_bs.jl:77-95): no permute when r
    already is the last prefix axis; else the other prefix axes in order, then r"""
    ax = op.axis(rlab)
    if ax == op.P - 1:
        return op
    pref = [a for a in range(op.P) if a != ax] + [ax]
    return op.permute(pref + list(range(op.P, len(op.labels))))


def _ensure_rlab_last_and_sorted(op, rlab):
    """contract_coo.jl:48-82: unchanged (and NOT re-sorted) when r is already
    the last axis; else r moved last, the rest in order, and sorted"""
    if op.labels[-1] == rlab:
        return op
    perm = [i for i, l in enumerate(op.labels) if l != rlab] + [op.labels.index(rlab)]
    return op.permute_sorted(perm)


def _slice_nonzero(block, ax, r):
    """_slice_any_nonzero over the r-slice of a tail (contract_bs_bs.jl:280-285)"""
    return bool(np.any(np.take(block, r, axis=ax) != 0))


class _Creator:
    """_ensure_block!: first-creation order of keys (storage.jl:73-83)"""

    def __init__(self):
        self.seen, self.order = set(), []

    def ensure(self, k):
        if k not in self.seen:
            self.seen.add(k)
            self.order.append(k)


def _shared_labels(la, lb, lc):
    """contract.jl:63-81: shared = labels of A found in B, in A's order;
    reduced = shared labels absent from labelsC"""
    shared = [l for l in la if l in lb]
    reduced = [l for l in shared if l not in lc]
    return shared, reduced


# ---------------------------------------------------------------- BS x BS
def _prefix_outer_bb(A, B, lc, PC, rlab):
    """contract_prefix_outer_bb! (contract_bs_bs.jl:173-263): merge-join on the
    r runs; for every A block of a matching run, every B block of it (:232-252)"""
    A, B = _r_last_prefix(A, rlab), _r_last_prefix(B, rlab)
    src = [(0, A.axis(l)) if l in A.labels else (1, B.axis(l)) for l in lc[:PC]]
    c = _Creator()
    iA = iB = 0
    ra = lambda i: A.keys[i][A.P - 1]  # noqa: E731
    rb = lambda i: B.keys[i][B.P - 1]  # noqa: E731
    while iA < len(A.keys) and iB < len(B.keys):
        if ra(iA) < rb(iB):
            v = ra(iA)
            while iA < len(A.keys) and ra(iA) == v:
                iA += 1
            continue
        if rb(iB) < ra(iA):
            v = rb(iB)
            while iB < len(B.keys) and rb(iB) == v:
                iB += 1
            continue
        v = ra(iA)
        lo = iB
        while iB < len(B.keys) and rb(iB) == v:
            iB += 1
        while iA < len(A.keys) and ra(iA) == v:
            for j in range(lo, iB):
                c.ensure(tuple(A.keys[iA][a] if s == 0 else B.keys[j][a] for s, a in src))
            iA += 1
    return c.order


def _dense_bb(A, B, lc, PC, rlab):
    """contract_dense_bb! (contract_bs_bs.jl:288-423): the tail permutes keep
    the key order; every (A block, B block) in that order; the C block is
    created at the first r whose A and B slices both hold a nonzero (:374-379)"""
    axa, axb = A.axis(rlab) - A.P, B.axis(rlab) - B.P
    src = [(0, A.axis(l)) if l in A.labels else (1, B.axis(l)) for l in lc[:PC]]
    R = A.blocks[A.keys[0]].shape[axa] if A.keys else 0
    c = _Creator()
    for ka in A.keys:
        for kb in B.keys:
            if any(_slice_nonzero(A.blocks[ka], axa, r) and _slice_nonzero(B.blocks[kb], axb, r) for r in range(R)):
                c.ensure(tuple(ka[a] if s == 0 else kb[a] for s, a in src))
    return c.order


def _join_walk(A, B, posA, posB, exact):
    """The matched (A index, B index) pairs in the order contract_shared!
    visits them.  exact=False: the reference's merge walk comparing the join
    tuples lexicographically (_cmp_join_tuple / _advance_run,
    contract_bs_bs.jl:430-464, 684-728).  exact=True: every A block in order
    with every B block of an equal join tuple in B's order (the ordered
    equi-join the library runs under SB_FLAG_EXACT_JOIN)."""
    ja = lambda i: tuple(A.keys[i][p] for p in posA)  # noqa: E731
    jb = lambda j: tuple(B.keys[j][p] for p in posB)  # noqa: E731
    pairs = []
    if exact:
        for i in range(len(A.keys)):
            pairs += [(i, j) for j in range(len(B.keys)) if jb(j) == ja(i)]
        return pairs

    def adv(keys_j, i0, n):
        if not posA:
            return n
        i = i0 + 1
        while i < n and keys_j(i) == keys_j(i0):
            i += 1
        return i
    iA = iB = 0
    nA, nB = len(A.keys), len(B.keys)
    while iA < nA and iB < nB:
        a, b = ja(iA), jb(iB)
        if a < b:
            iA = adv(ja, iA, nA)
            continue
        if a > b:
            iB = adv(jb, iB, nB)
            continue
        iA2, iB2 = adv(ja, iA, nA), adv(jb, iB, nB)
        pairs += [(i, j) for i in range(iA, iA2) for j in range(iB, iB2)]
        iA, iB = iA2, iB2
    return pairs


def _shared_prefix_perm(op, shared_prefix):
    """prefix part of _find_perm_for_A_join_and_dense_order / ..._B_... (contract_bs_bs.jl:513-567):
    kept prefix axes in order, then the shared prefix axes in shared order; the
    tail order does not move keys"""
    sp = [op.axis(l) for l in shared_prefix]
    keep = [a for a in range(op.P) if a not in sp]
    return op.permute(keep + sp + list(range(op.P, len(op.labels))))


def _shared_bs_bs(A, B, lc, PC, shared, exact):
    """contract_shared! BS x BS (contract_bs_bs.jl:572-730): one _ensure_block!
    per matched pair, in the walk's order (:700-705)"""
    sp = [l for l in shared if A.axis(l) < A.P and B.axis(l) < B.P]
    A, B = _shared_prefix_perm(A, sp), _shared_prefix_perm(B, sp)
    src = [(0, A.axis(l)) if (l in A.labels and A.axis(l) < A.P) else (1, B.axis(l)) for l in lc[:PC]]
    c = _Creator()
    for i, j in _join_walk(A, B, [A.axis(l) for l in sp], [B.axis(l) for l in sp], exact):
        c.ensure(tuple(A.keys[i][a] if s == 0 else B.keys[j][a] for s, a in src))
    return c.order


# ---------------------------------------------------------------- BS x Dense
def _prefix_outer_bd(A, lc, PC, rlab):
    """contract_prefix_outer_bd! (contract_bs_dense.jl:34-204): A permuted so r
    is the last prefix axis, the C key of every A block, a stable sortperm of
    those keys (lexicographic tuple order, :121-126), one C block per distinct
    key in that order (:141-143)"""
    A = _r_last_prefix(A, rlab)
    src = [A.axis(l) for l in lc[:PC]]
    ck = [tuple(k[a] for a in src) for k in A.keys]
    c = _Creator()
    for k in sorted(ck):
        c.ensure(k)
    return c.order


def _dense_bd(A, B, lb, lc, PC, rlab):
    """contract_dense_bd! (contract_bs_dense.jl:248-392): per A block in key
    order, created at the first r whose A slice and B slice both hold a nonzero
    (:338-351); B's r slice is B with r fixed"""
    axa = A.axis(rlab) - A.P
    axb = list(lb).index(rlab)
    R = B.shape[axb]
    src = [A.axis(l) for l in lc[:PC]]
    c = _Creator()
    for ka in A.keys:
        if any(_slice_nonzero(A.blocks[ka], axa, r) and bool(np.any(np.take(B, r, axis=axb) != 0)) for r in range(R)):
            c.ensure(tuple(ka[a] for a in src))
    return c.order


def _shared_bs_dense(A, lc, PC, shared):
    """contract_shared! BS x Dense (contract_bs_dense.jl:411-619): A permuted with
    its shared prefix axes last (:473), then every A block in that order creates
    (or finds) C key akey[c_src_axes] (:561-566, 599-603)"""
    sp = [l for l in shared if A.axis(l) < A.P]
    A = _shared_prefix_perm(A, sp)
    src = [A.axis(l) for l in lc[:PC]]
    c = _Creator()
    for k in A.keys:
        c.ensure(tuple(k[a] for a in src))
    return c.order


# ---------------------------------------------------------------- COO x ...
def _coo_dense(A, B, lb, lc, PC, rlab, skip_zero_slices):
    """contract! COO x Dense (contract_coo_dense.jl:48-125): A permuted to
    [C prefix..., r] and sorted (aligned_A_to_Cprefix, :4-18); every entry with
    1 <= rv <= R creates its C key, whatever B's slice holds (:113-122).  With
    the library's zero-slice flag, entries whose B slice is all zero create
    nothing."""
    perm = [A.labels.index(l) for l in lc[:PC]] + [A.labels.index(rlab)]
    A = A.permute_sorted(perm)
    axb = list(lb).index(rlab)
    c = _Creator()
    for k in A.keys:
        rv = k[-1]
        if skip_zero_slices and not np.any(np.take(B, rv - 1, axis=axb) != 0):
            continue
        c.ensure(tuple(k[:PC]))
    return c.order


def _coo_bs_prefix(A, B, lc, PC, rlab):
    """contract_prefix! COO x BS (contract_coo_bs.jl:51-124): merge-join on r
    runs, every (entry, block) pair of a matching run (:97-119)"""
    A = _ensure_rlab_last_and_sorted(A, rlab)
    B = _r_last_prefix(B, rlab)
    src = [(0, A.labels.index(l)) if l in A.labels else (1, B.axis(l)) for l in lc[:PC]]
    c = _Creator()
    iA = iB = 0
    ra = lambda i: A.keys[i][-1]  # noqa: E731
    rb = lambda j: B.keys[j][B.P - 1]  # noqa: E731
    while iA < len(A.keys) and iB < len(B.keys):
        if ra(iA) < rb(iB):
            v = ra(iA)
            while iA < len(A.keys) and ra(iA) == v:
                iA += 1
        elif rb(iB) < ra(iA):
            v = rb(iB)
            while iB < len(B.keys) and rb(iB) == v:
                iB += 1
        else:
            v = ra(iA)
            lo = iB
            while iB < len(B.keys) and rb(iB) == v:
                iB += 1
            while iA < len(A.keys) and ra(iA) == v:
                for j in range(lo, iB):
                    c.ensure(tuple(A.keys[iA][a] if s == 0 else B.keys[j][a] for s, a in src))
                iA += 1
    return c.order


def _coo_bs_dense_index(A, B, lc, PC, rlab):
    """contract_dense_index! (contract_coo_bs.jl:140-241): every entry, every
    block of B in key order; created only where alpha * (B's r slice) holds a
    nonzero (:229-230)"""
    A = _ensure_rlab_last_and_sorted(A, rlab)
    axb = B.axis(rlab) - B.P
    src = [(0, A.labels.index(l)) if l in A.labels else (1, B.axis(l)) for l in lc[:PC]]
    c = _Creator()
    for k, v in zip(A.keys, A.vals):
        rv = k[-1]
        for kb in B.keys:
            if np.any(v * np.take(B.blocks[kb], rv - 1, axis=axb) != 0):
                c.ensure(tuple(k[a] if s == 0 else kb[a] for s, a in src))
    return c.order


def _coo_coo(A, B, lc, rlab):
    """contract! COO x COO (contract_coo.jl:93-170): keys pushed in walk order,
    one per nonzero product term (no coalescing, :84-91); the A cursor steps
    one entry at a time on a mismatch (:139-142)"""
    A = _ensure_rlab_last_and_sorted(A, rlab)
    B = _ensure_rlab_last_and_sorted(B, rlab)
    src = [(0, A.labels.index(l)) if l in A.labels else (1, B.labels.index(l)) for l in lc]
    out = []
    i = j = 0
    while i < len(A.keys) and j < len(B.keys):
        ra, rb = A.keys[i][-1], B.keys[j][-1]
        if ra < rb:
            i += 1
        elif rb < ra:
            j += 1
        else:
            lo = j
            while j < len(B.keys) and B.keys[j][-1] == rb:
                j += 1
            while i < len(A.keys) and A.keys[i][-1] == ra:
                for jj in range(lo, j):
                    if A.vals[i] * B.vals[jj] != 0.0:
                        out.append(tuple(A.keys[i][a] if s == 0 else B.keys[jj][a] for s, a in src))
                i += 1
    return out


# ---------------------------------------------------------------- entry points
def creation_order(ck, lc, N2C, ak, A, la, N2A, bk, B, lb, N2B, exact_join=False, skip_zero_slices=False,
                   atol=1e-12):
    """C keys in first-creation order for contract!(C, lc, A, la, B, lb), the
    operands given as dense arrays (kinds "bs", "coo", "dense"; N2 = dense-tail
    length of a block-sparse operand).  The dispatch is contract.jl:37-104."""
    la, lb, lc = list(la), list(lb), list(lc)
    PC = len(lc) - (N2C or 0) if ck == "bs" else len(lc)
    shared, reduced = _shared_labels(la, lb, lc)
    Aop = BSop.from_dense(A, la, len(la) - N2A) if ak == "bs" else COOop.from_dense(A, la, atol)
    if bk == "bs":
        Bop = BSop.from_dense(B, lb, len(lb) - N2B)
    elif bk == "coo":
        Bop = COOop.from_dense(B, lb, atol)
    else:
        Bop = np.asarray(B)
    single = len(reduced) == 1
    if ak == "bs" and bk == "bs":
        if not single:
            return _shared_bs_bs(Aop, Bop, lc, PC, shared, exact_join)
        r = reduced[0]
        if Bop.axis(r) < Bop.P:
            return _prefix_outer_bb(Aop, Bop, lc, PC, r)
        return _dense_bb(Aop, Bop, lc, PC, r)
    if ak == "bs" and bk == "dense":
        if not single:
            return _shared_bs_dense(Aop, lc, PC, shared)
        r = reduced[0]
        if Aop.axis(r) < Aop.P:
            return _prefix_outer_bd(Aop, lc, PC, r)
        return _dense_bd(Aop, Bop, lb, lc, PC, r)
    assert single, "COO kernels take exactly one reduced label"
    r = reduced[0]
    if ak == "coo" and bk == "dense":
        return _coo_dense(Aop, Bop, lb, lc, PC, r, skip_zero_slices)
    if ak == "coo" and bk == "bs":
        if Bop.axis(r) < Bop.P:
            return _coo_bs_prefix(Aop, Bop, lc, PC, r)
        return _coo_bs_dense_index(Aop, Bop, lc, PC, r)
    if ak == "coo" and bk == "coo":
        return _coo_coo(Aop, Bop, lc, r)
    raise ValueError((ak, bk))


def ids_order(parts):
    """The C keys ordered by their ids (payload slots), from a result's parts."""
    ids = np.asarray(parts["ids"])
    if len(ids) == 0:
        return []
    keys = np.asarray(parts["keys"]).reshape(len(ids), -1)
    out = [None] * len(ids)
    for k, i in zip(keys, ids):
        out[int(i) - 1] = tuple(int(x) for x in k)
    return out


def check_ids(parts, expected):
    """ids are 1..n and the key holding id j is the j-th key created"""
    ids = sorted(int(i) for i in np.asarray(parts["ids"]))
    assert ids == list(range(1, len(ids) + 1)), "ids are not a permutation of 1..n"
    got = ids_order(parts)
    assert got == list(expected), f"creation order differs:\n got {got}\n exp {list(expected)}"
