"""The block Jacobi schedule of the sector SVD (sparsebackends_amd/csrc/sb_svd.hip,
sect_bj_round_kernel): blocks of bs rows paired round-robin over B - 1 rounds;
in round 0 a workgroup sweeps every pair of its rows, in later rounds only the
pairs across its two blocks.  One-sided Jacobi converges when a sweep visits
every row pair, and a step may rotate a row in one pair only.  The index
formulas below are the kernel's; the test checks those two properties for
every sector size the kernel meets in practice.  The rotations of a
workgroup are recorded as one matrix T per (round, pair) and applied to the
accumulated rotation Q once per sweep (sect_bj_apply_kernel, in round order):
a model of that scheme must still give X_0 = Q X_final with Q orthogonal."""
from collections import Counter

import numpy as np
import pytest


def blocks(n, bs):
    B = (n + bs - 1) // bs
    B += B & 1
    return max(B, 2)


def sweep_steps(n, bs):
    """the row pairs of one sweep: (round, workgroup, pairs of one step)"""
    B = blocks(n, bs)
    mB = B - 1
    for rd in range(mB):
        for pi in range(B // 2):
            pb = 0 if pi == 0 else 1 + (rd + pi) % mB
            qb = 1 + rd if pi == 0 else 1 + (rd - pi + mB) % mB
            p0, p1 = pb * bs, min(n, pb * bs + bs)
            q0, q1 = qb * bs, min(n, qb * bs + bs)
            np_, nq_ = max(0, p1 - p0), max(0, q1 - q0)
            nl = np_ + nq_
            if nl < 2:
                continue

            def grow(l, p0=p0, q0=q0, np_=np_):
                return p0 + l if l < np_ else q0 + (l - np_)

            if rd == 0:
                n2 = nl + (nl & 1)
                m, npair = n2 - 1, n2 // 2
                for r in range(m):
                    step = []
                    for qi in range(npair):
                        p = 0 if qi == 0 else 1 + (r + qi) % m
                        q = 1 + r if qi == 0 else 1 + (r - qi + m) % m
                        if p < nl and q < nl:
                            step.append((grow(p), grow(q)))
                    yield rd, pi, step
            else:
                L = max(np_, nq_)
                for k in range(L):
                    step = []
                    for i in range(L):
                        j = (i + k) % L
                        if i < np_ and j < nq_:
                            step.append((grow(i), grow(np_ + j)))
                    yield rd, pi, step


@pytest.mark.parametrize("bs", [1, 2, 3, 5, 7, 8, 11, 16])
def test_every_pair_once_per_sweep(bs):
    for n in range(2, 160):
        seen = Counter()
        for _, _, step in sweep_steps(n, bs):
            rows = [r for pq in step for r in pq]
            assert len(rows) == len(set(rows)), (n, bs, step)  # no row twice in a step
            for p, q in step:
                assert p != q
                seen[(min(p, q), max(p, q))] += 1
        expect = {(i, j) for i in range(n) for j in range(i + 1, n)}
        assert set(seen) == expect, (n, bs)
        assert max(seen.values()) == 1, (n, bs)


def test_sequential_depth_148_rows():
    # 8-row blocks: 15 steps in the first round, 8 in each of the 18 others
    steps = Counter((rd, pi) for rd, pi, _ in sweep_steps(148, 8))
    depth = sum(max(c for (r, _), c in steps.items() if r == rd) for rd in {r for r, _ in steps})
    assert depth == 15 + 18 * 8


def rounds(n, bs):
    """per round: the list of (pi, rows of the pair in the kernel's local order, steps)"""
    B = blocks(n, bs)
    out = {}
    for rd, pi, step in sweep_steps(n, bs):
        out.setdefault(rd, {}).setdefault(pi, []).append(step)
    return B, out


def pair_rows(n, bs, B, rd, pi):
    mB = B - 1
    pb = 0 if pi == 0 else 1 + (rd + pi) % mB
    qb = 1 + rd if pi == 0 else 1 + (rd - pi + mB) % mB
    rows = list(range(pb * bs, min(n, pb * bs + bs))) + list(range(qb * bs, min(n, qb * bs + bs)))
    return rows


@pytest.mark.parametrize("n,bs,ncol", [(2, 16, 5), (17, 16, 20), (37, 5, 37), (40, 16, 45), (64, 16, 70), (23, 3, 30)])
def test_recorded_rotations_rebuild_q(n, bs, ncol):
    rng = np.random.default_rng(n * 100 + bs)
    X0 = rng.standard_normal((n, ncol)) * np.exp(-np.arange(n) / 7.0)[:, None]
    X = X0.copy()
    Q = np.eye(n)  # Q[:, c] rotates like row c of X
    tol = max(1e-15, 2 * ncol * 2.0 ** -53)
    null2 = 1e-30 * (X0 ** 2).sum()  # rows the kernel treats as zero (rows <= columns here, as in the kernel)
    for sweep in range(40):
        B, rnd = rounds(n, bs)
        Ts = {}
        rotated = False
        for rd in sorted(rnd):
            for pi, steps in rnd[rd].items():
                rows = pair_rows(n, bs, B, rd, pi)
                loc = {r: l for l, r in enumerate(rows)}
                T = np.eye(2 * bs)
                for step in steps:
                    for p, q in step:
                        x, y = X[p].copy(), X[q].copy()
                        a, b, c = x @ x, y @ y, x @ y
                        if c == 0 or a <= null2 or b <= null2 or abs(c) <= tol * np.sqrt(a) * np.sqrt(b):
                            continue
                        zeta = (b - a) / (2 * c)
                        t = (1.0 if zeta >= 0 else -1.0) / (abs(zeta) + np.sqrt(1 + zeta * zeta))
                        cs = 1 / np.sqrt(1 + t * t)
                        sn = cs * t
                        X[p], X[q] = cs * x - sn * y, sn * x + cs * y
                        tp, tq = T[loc[p]].copy(), T[loc[q]].copy()
                        T[loc[p]], T[loc[q]] = cs * tp - sn * tq, sn * tp + cs * tq
                        rotated = True
                Ts[(rd, pi)] = (rows, T)
        # the apply launch: every round in order, each pair's rows of Q^T get T times the old ones
        Qt = Q.T.copy()
        for rd in sorted(rnd):
            for pi in rnd[rd]:
                rows, T = Ts[(rd, pi)]
                nl = len(rows)
                Qt[rows] = T[:nl, :nl] @ Qt[rows]
        Q = Qt.T
        if not rotated:
            break
    np.testing.assert_allclose(Q @ X, X0, atol=1e-12 * np.abs(X0).max())
    np.testing.assert_allclose(Q.T @ Q, np.eye(n), atol=1e-12)
    G = X @ X.T
    off = G - np.diag(np.diag(G))
    d = np.sqrt(np.diag(G))
    assert np.all(np.abs(off) <= 1e-12 * np.outer(d, d) + 1e-300)
