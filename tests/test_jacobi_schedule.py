"""The block Jacobi schedule of the sector SVD (sparsebackends_amd/csrc/sb_svd.hip,
sect_bj_round_kernel): blocks of bs rows paired round-robin over B - 1 rounds;
in round 0 a workgroup sweeps every pair of its rows, in later rounds only the
pairs across its two blocks.  One-sided Jacobi converges when a sweep visits
every row pair, and a step may rotate a row in one pair only.  The index
formulas below are the kernel's; the test checks those two properties for
every sector size the kernel meets in practice."""
from collections import Counter

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


@pytest.mark.parametrize("bs", [1, 2, 3, 5, 7, 8])
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
