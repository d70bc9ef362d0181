"""CPU restatement of the reference's two-site block-diagonal SVD.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline): nothing in the product
imports this module.

blocksparse_svd (SparseBackends/src/ops_factorize.jl:50-148) on a
NewBlockSparseSorted{T,4,2,2} given as host parts (prefix (L, R), tail
(s1, s2)): one SVD per L value of the d_s1 x (n_R * d_s2) matrix of its
blocks in ascending R (:59-83), all singular values sorted descending by a
stable sort (:86-92), the kept count cut by maxdim and then lowered while
above mindim and below cutoff * max (:94-98), truncerr the sum of the
discarded squares (:98), new bond indices contiguous per L group in group
order (:106-112), U blocks (L, nb) of d_s1 values and SV blocks (nb, R) of
d_s2 values created in loop order with _ensure_block! (:118-143), the
singular value folded into SV (ortho "left") or U (ortho "right").

numpy's svd (LAPACK gesdd, the routine Julia's svd uses by default) does the
per-group factorisations.  Pinned by (tests/test_oracle_factorize.py):
exact reconstruction without truncation, orthonormal U within each group,
the per-group singular values of numpy.linalg.svd, the truncation rules on
hand-built spectra, and a literal key/id case worked by hand from the loop
order of the reference.

blocksparse_left_qr (ops_factorize.jl:8-43) and left_factorize
(tensor_wrappers.jl:954-970): the dense (n_l d_s) x n_r matrix of the
nonzero l and r values (row (li-1) d_s + s, column ri, :13-25), numpy's qr
(LAPACK geqrf / orgqr, the Householder QR with the sign conventions of the
reference's qr), thin to k = min(n_l d_s, n_r) (:28-31), Q keyed by l with one
block per l value in ascending order (setindex! creates every block, :34-41),
R scattered to k x d_r by left_factorize.  Pinned (tests/test_oracle_qr.py)
by reconstruction, orthonormal Q, a hand-worked literal case whose signs are
LAPACK's, and the key/id rule.

svd_sectors: the same factorisation for the sector format (psi
[Ql, Qr | s1, s2, l, r], blocks by fused charge), which the reference does not
have; pinned (tests/test_sector_svd.py) by exact reconstruction over the full
bonds, isometric factors, and its singular values equalling numpy's SVD of the
whole (s1, l) x (s2, r) matrix.
"""
from __future__ import annotations

import numpy as np


def blocksparse_left_qr(parts):
    """returns (Q parts, R (k x n_r), l_vals, r_vals); parts with dims
    (n_l, n_r, d_s), N2 = 1"""
    assert parts["N2"] == 1 and len(parts["dims"]) == 3
    keys = np.asarray(parts["keys"]).reshape(-1, 2)
    ids = np.asarray(parts["ids"])
    data = np.asarray(parts["data"], dtype=np.float64)
    ds = int(parts["dims"][2])
    l_vals = np.unique(keys[:, 0])
    r_vals = np.unique(keys[:, 1])
    li = {int(v): i for i, v in enumerate(l_vals)}
    ri = {int(v): i for i, v in enumerate(r_vals)}
    M = np.zeros((len(l_vals) * ds, len(r_vals)))
    for (l, r), i in zip(keys, ids):
        M[li[int(l)] * ds:(li[int(l)] + 1) * ds, ri[int(r)]] = data[(int(i) - 1) * ds:int(i) * ds]
    Q, R = np.linalg.qr(M, mode="reduced")
    k = min(M.shape)
    Q, R = Q[:, :k], R[:k, :]
    qdata = np.zeros(len(l_vals) * ds * k)
    for n, _ in enumerate(l_vals):
        qdata[n * ds * k:(n + 1) * ds * k] = Q[n * ds:(n + 1) * ds, :].reshape(-1, order="F")
    qparts = dict(dims=(int(parts["dims"][0]), ds, k), N2=2, keys=l_vals.reshape(-1, 1).astype(np.int64),
                  ids=np.arange(1, len(l_vals) + 1), data=qdata, blksize=ds * k)
    return qparts, R, l_vals.astype(np.int64), r_vals.astype(np.int64)


def left_factorize(parts):
    """returns (Q parts, R_full (k x d_r)): R's column ri at column r_vals[ri] - 1"""
    q, R, _, r_vals = blocksparse_left_qr(parts)
    full = np.zeros((R.shape[0], int(parts["dims"][1])))
    full[:, r_vals - 1] = R
    return q, full


def blocksparse_svd(parts, ortho="left", maxdim=None, mindim=1, cutoff=0.0):
    """returns (U parts, SV parts, svs_kept, eigs, truncerr); parts as in
    oracle.BS.parts() with dims (n_L, n_R, d_s1, d_s2), N2 = 2"""
    n_L, n_R, d1, d2 = (int(x) for x in parts["dims"])
    assert parts["N2"] == 2 and len(parts["dims"]) == 4
    keys = np.asarray(parts["keys"]).reshape(-1, 2)
    ids = np.asarray(parts["ids"])
    data = np.asarray(parts["data"], dtype=np.float64)
    bs = d1 * d2
    groups = {}
    for (L, R), i in zip(keys, ids):
        groups.setdefault(int(L), []).append((int(R), int(i)))
    block_svds = []
    for L in sorted(groups):
        rl = sorted(groups[L])
        M = np.zeros((d1, len(rl) * d2))
        for j, (R, i) in enumerate(rl):
            M[:, j * d2:(j + 1) * d2] = data[(i - 1) * bs:i * bs].reshape(d1, d2, order="F")
        U, S, Vt = np.linalg.svd(M, full_matrices=False)
        block_svds.append((L, U, S, Vt, [R for R, _ in rl]))
    all_svs = [(float(s), bi, c) for bi, (_, _, S, _, _) in enumerate(block_svds) for c, s in enumerate(S)]
    all_svs.sort(key=lambda x: -x[0])  # stable: ties keep group, then column order
    max_sv = all_svs[0][0] if all_svs else 1.0
    keep = len(all_svs) if maxdim is None else min(int(maxdim), len(all_svs))
    while keep > mindim and all_svs[keep - 1][0] < cutoff * max_sv:
        keep -= 1
    truncerr = 0.0
    for x in all_svs[keep:]:
        truncerr += x[0] ** 2
    k_kept = [0] * len(block_svds)
    for x in all_svs[:keep]:
        k_kept[x[1]] += 1
    k = sum(k_kept)
    ukeys, uids, udata = [], [], []
    skeys, sids, sdata = [], [], []
    svs = np.zeros(k)
    off = 0
    for bi, (L, U, S, Vt, rl) in enumerate(block_svds):
        for c in range(k_kept[bi]):
            nb = off + c + 1
            svs[nb - 1] = S[c]
            ukeys.append((L, nb))
            uids.append(len(uids) + 1)
            udata.append(U[:, c] * (S[c] if ortho == "right" else 1.0))
            for j, R in enumerate(rl):
                skeys.append((nb, R))
                sids.append(len(sids) + 1)
                sdata.append(Vt[c, j * d2:(j + 1) * d2] * (S[c] if ortho == "left" else 1.0))
        off += k_kept[bi]

    def pack(dims, keys_, ids_, blocks):
        P = 2
        order = sorted(range(len(keys_)), key=lambda t: tuple(reversed(keys_[t])))  # column-major lin
        return dict(dims=dims, N2=1, keys=np.array([keys_[t] for t in order], dtype=np.int64).reshape(-1, P),
                    ids=np.array([ids_[t] for t in order], dtype=np.int64),
                    data=np.concatenate(blocks) if blocks else np.zeros(0))

    Up = pack((n_L, k, d1), ukeys, uids, udata)
    SVp = pack((k, n_R, d2), skeys, sids, sdata)
    return Up, SVp, svs, svs ** 2, truncerr


def reconstruct(Up, SVp):
    """dense A[L, R, s1, s2] = sum_nb U[L, nb, s1] SV[nb, R, s2]"""
    n_L, k, d1 = Up["dims"]
    _, n_R, d2 = SVp["dims"]
    U = np.zeros((n_L, k, d1))
    for (L, nb), i in zip(Up["keys"], Up["ids"]):
        U[L - 1, nb - 1] = Up["data"][(i - 1) * d1:i * d1]
    S = np.zeros((k, n_R, d2))
    for (nb, R), i in zip(SVp["keys"], SVp["ids"]):
        S[nb - 1, R - 1] = SVp["data"][(i - 1) * d2:i * d2]
    return np.einsum("lka,krb->lrab", U, S)


def svd_sectors(parts, rowsec, colsec, dl, dr, nsec, ortho="left", maxdim=None, mindim=1, cutoff=0.0):
    """Block-diagonal SVD of psi [Ql, Qr | s1, s2, l, r] by fused sector (the
    sector-format counterpart of blocksparse_svd, with its truncation rules):
    sector m's matrix has rows (Ql, s1, l) and columns (Qr, s2, r) ascending;
    numpy's svd per sector; singular values sorted descending across sectors
    by a stable sort (sector, then index order for ties); new bond = sectors
    with kept values, ascending.  Returns (A parts, B parts, svs, kept, truncerr)."""
    nl, nr, d1, d2, Dl, Dr = (int(x) for x in parts["dims"])
    rowsec = np.asarray(rowsec).reshape(nl, d1)
    colsec = np.asarray(colsec).reshape(nr, d2)
    bs = d1 * d2 * Dl * Dr
    rows = [[] for _ in range(nsec)]
    cols = [[] for _ in range(nsec)]
    for il in range(nl):
        for s1 in range(d1):
            if rowsec[il, s1] >= 0:
                rows[rowsec[il, s1]].append((il, s1))
    for ir in range(nr):
        for s2 in range(d2):
            if colsec[ir, s2] >= 0:
                cols[colsec[ir, s2]].append((ir, s2))
    roff, coff = {}, {}
    R, Cn = [0] * nsec, [0] * nsec
    for m in range(nsec):
        for (il, s1) in rows[m]:
            roff[(il, s1)] = R[m]
            R[m] += int(dl[il])
        for (ir, s2) in cols[m]:
            coff[(ir, s2)] = Cn[m]
            Cn[m] += int(dr[ir])
    M = [np.zeros((R[m], Cn[m])) for m in range(nsec)]
    for (a, b), i in zip(parts["keys"], parts["ids"]):
        il, ir = int(a) - 1, int(b) - 1
        blk = np.asarray(parts["data"][(i - 1) * bs:i * bs]).reshape((d1, d2, Dl, Dr), order="F")
        for s1 in range(d1):
            m = rowsec[il, s1]
            if m < 0:
                continue
            for s2 in range(d2):
                if colsec[ir, s2] != m:
                    continue
                ro, co = roff[(il, s1)], coff[(ir, s2)]
                M[m][ro:ro + dl[il], co:co + dr[ir]] = blk[s1, s2, :dl[il], :dr[ir]]
    fac = {}
    all_svs = []
    for m in range(nsec):
        if R[m] and Cn[m]:
            U, S, Vt = np.linalg.svd(M[m], full_matrices=False)
            fac[m] = (U, S, Vt)
            all_svs += [(float(s), m) for s in S]
    all_svs.sort(key=lambda x: -x[0])
    max_sv = all_svs[0][0] if all_svs else 1.0
    keep = len(all_svs) if maxdim is None else min(int(maxdim), len(all_svs))
    while keep > mindim and all_svs[keep - 1][0] < cutoff * max_sv:
        keep -= 1
    truncerr = 0.0
    for x in all_svs[keep:]:
        truncerr += x[0] ** 2
    kept = [0] * nsec
    for x in all_svs[:keep]:
        kept[x[1]] += 1
    live = [m for m in range(nsec) if kept[m] > 0]
    qm = {m: i for i, m in enumerate(live)}
    Dm = max(kept) if live else 0
    svs = np.concatenate([fac[m][1][:kept[m]] for m in live]) if live else np.zeros(0)
    akeys, adata, bkeys, bdata = [], [], [], []
    for m in live:  # A keys (Ql, Qm): Qm major
        U, S, Vt = fac[m]
        k = kept[m]
        for il in range(nl):
            ss = [s1 for s1 in range(d1) if rowsec[il, s1] == m]
            if not ss or not dl[il]:
                continue
            blk = np.zeros((d1, Dl, Dm))
            for s1 in ss:
                ro = roff[(il, s1)]
                blk[s1, :dl[il], :k] = U[ro:ro + dl[il], :k] * (S[:k] if ortho == "right" else 1.0)
            akeys.append((il + 1, qm[m] + 1))
            adata.append(blk.reshape(-1, order="F"))
    for ir in range(nr):  # B keys (Qm, Qr): Qr major
        for m in live:
            U, S, Vt = fac[m]
            k = kept[m]
            ss = [s2 for s2 in range(d2) if colsec[ir, s2] == m]
            if not ss or not dr[ir]:
                continue
            blk = np.zeros((d2, Dm, Dr))
            for s2 in ss:
                co = coff[(ir, s2)]
                v = Vt[:k, co:co + dr[ir]]
                blk[s2, :k, :dr[ir]] = v * (S[:k, None] if ortho == "left" else 1.0)
            bkeys.append((qm[m] + 1, ir + 1))
            bdata.append(blk.reshape(-1, order="F"))
    A = dict(dims=(nl, len(live), d1, Dl, Dm), N2=3, keys=np.array(akeys, dtype=np.int64).reshape(-1, 2),
             ids=np.arange(1, len(akeys) + 1), data=np.concatenate(adata) if adata else np.zeros(0))
    B = dict(dims=(len(live), nr, d2, Dm, Dr), N2=3, keys=np.array(bkeys, dtype=np.int64).reshape(-1, 2),
             ids=np.arange(1, len(bkeys) + 1), data=np.concatenate(bdata) if bdata else np.zeros(0))
    return A, B, svs, np.array(kept), truncerr
