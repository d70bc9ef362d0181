"""Dense restatement of a two-site H_eff*psi for checking the block chain.

A sector-format block-sparse tensor (key prefix = sector numbers, dense tail
= in-sector coordinates padded to the largest sector) is expanded to a plain
dense array over the full bond dimensions; H_eff*psi is then one einsum:

    Hpsi[s1', s2', l', r'] = sum L[wl, l', l] W[wl, s1', s1, wm] W[wm, s2', s2, wr]
                                 R[wr, r', r] psi[s1, s2, l, r]
"""
import numpy as np


def offsets(bond):
    return np.concatenate([[0], np.cumsum(bond.dims)]).astype(np.int64)


def expand(parts, bonds, tail_map):
    """parts: dims/N2/keys/ids/data; bonds[i]: the Bond of prefix axis i;
    tail_map[j]: the prefix axis whose sector fixes tail axis j, or None.
    Returns (dense array over the full bonds, max |padding entry|)."""
    P = len(bonds)
    tdims = tuple(parts["dims"][P:])
    bs = int(np.prod(tdims))
    offs = [offsets(b) for b in bonds]
    full = [bonds[m].dim if m is not None else tdims[j] for j, m in enumerate(tail_map)]
    out = np.zeros(full)
    pad = 0.0
    for k, i in zip(parts["keys"], parts["ids"]):
        blk = np.asarray(parts["data"][(i - 1) * bs:i * bs]).reshape(tdims, order="F")
        ext, dst = [], []
        for j, m in enumerate(tail_map):
            if m is None:
                ext.append(slice(0, tdims[j]))
                dst.append(slice(0, tdims[j]))
            else:
                s = int(k[m]) - 1
                n = bonds[m].dims[s]
                ext.append(slice(0, n))
                dst.append(slice(offs[m][s], offs[m][s] + n))
        out[tuple(dst)] += blk[tuple(ext)]
        mask = np.ones(tdims, dtype=bool)
        mask[tuple(ext)] = False
        if mask.any():
            pad = max(pad, float(np.abs(blk[mask]).max()))
    return out, pad


def heff_dense(prob):
    """(Hpsi over the full bonds, the four dense operands) of a HeffProblem's host parts"""
    Bl, Br = prob.Bl, prob.Br
    psi, _ = expand(prob.psi_parts, [Bl, Br], [None, None, 0, 1])  # [s1, s2, l, r]
    L, _ = expand(prob.L_parts, [Bl, Bl], [None, 0, 1])  # [wl, l', l]
    R, _ = expand(prob.R_parts, [Br, Br], [None, 0, 1])  # [wr, r', r]
    W = prob.m.W
    # pairwise, in the chain's order (one 5-operand einsum is far slower)
    T1 = np.einsum("wab,yube->awyue", L, psi)
    T2 = np.einsum("awyue,wxyv->axvue", T1, W)
    T3 = np.einsum("axvue,vzuq->axzqe", T2, W)
    H = np.einsum("axzqe,qce->axzc", T3, R)  # [l', s1', s2', r']
    return np.ascontiguousarray(H.transpose(1, 2, 0, 3))  # [s1', s2', l', r']


def expand_out(prob, parts):
    return expand(parts, [prob.Bl, prob.Br], [None, None, 0, 1])
