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
"""
from __future__ import annotations

import numpy as np


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
