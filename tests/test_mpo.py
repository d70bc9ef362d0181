"""P·H·P and MPO x MPO on COO / block-sparse storage, site by site with the
links fused (the reference's sandwich_mpo / mulMPO flows, test.jl).

The oracle runs the same contract! and fusion calls; both are pinned by the
mathematics: the site tensors, chained with their boundary vectors, give the
dense P·H·P (and A·B) of a 6-site Rydberg-blockade chain.  The GPU runs the
same calls through the C ABI and must give the oracle's block list bit for
bit and its values exactly (0/1 projectors, integer couplings)."""
import numpy as np
import pytest

from sparsebackends_amd import mpo as M
from oracle import oracle as O


def oracle_sandwich(P, H):
    L = M._lab
    wP, d, wH = P.shape[0], P.shape[1], H.shape[0]
    Pc = O.COO.from_dense(P, atol=0.0)
    X = O.BS.new((wP, d, wP, wH, d, wH), 3)
    O.contract(X, [L("c"), L("so"), L("c2"), L("b"), L("u"), L("b2")], Pc, [L("c"), L("so"), L("t"), L("c2")],
               O.Dense.new(H), [L("b"), L("t"), L("u"), L("b2")], 0)
    Y = O.BS.new((wP, wP, d, d, wP, wP, wH, wH), 2)
    O.contract(Y, [L("c"), L("a"), L("so"), L("si"), L("c2"), L("a2"), L("b"), L("b2")], Pc,
               [L("a"), L("u"), L("si"), L("a2")], X, [L("c"), L("so"), L("c2"), L("b"), L("u"), L("b2")], 0)
    Y = Y.fuse_two_axes_(1, 2)
    return Y.fuse_two_axes_(4, 5)


def oracle_mul(A, B):
    L = M._lab
    wA, d, wB = A.shape[0], A.shape[1], B.shape[0]
    C = O.COO.new((wA, wB, d, d, wA, wB))
    O.contract(C, [L("a"), L("b"), L("so"), L("si"), L("a2"), L("b2")], O.COO.from_dense(A, atol=0.0),
               [L("a"), L("so"), L("t"), L("a2")], O.COO.from_dense(B, atol=0.0), [L("b"), L("t"), L("si"), L("b2")],
               0)
    return C.fuse_axes(1, 2).fuse_axes(4, 5)


def bs_site_dense(parts):
    """[L, so, si, R | b, b'] parts -> dense [(L, b), so, si, (R, b')] with L fastest"""
    dims = tuple(parts["dims"])
    D = np.zeros(dims)
    bs = int(np.prod(dims[4:]))
    for k, i in zip(parts["keys"], parts["ids"]):
        D[tuple(int(x) - 1 for x in k)] = np.asarray(parts["data"][(i - 1) * bs:i * bs]).reshape(dims[4:], order="F")
    wL, d, _, wR, wb, wb2 = dims
    return D.transpose(0, 4, 1, 2, 3, 5).reshape(wL * wb, d, d, wR * wb2, order="F")


def coo_site_dense(parts):
    """COO parts -> dense, duplicates summed (COO x COO keeps every term)"""
    D = np.zeros(tuple(parts["dims"]))
    for k, v in zip(parts["keys"], parts["vals"]):
        D[tuple(int(x) - 1 for x in k)] += v
    return D


def chain(W, N, lvec, rvec):
    """dense operator of N identical site tensors W [l, s', s, r] between boundary vectors"""
    d = W.shape[1]
    T = np.einsum("l,lxyr->xyr", lvec, W)  # [s', s, r]
    op = T
    for _ in range(N - 1):
        op = np.einsum("...r,rxyq->...xyq", op, W)
    op = np.einsum("...r,r->...", op, rvec)
    # op indices: s1', s1, s2', s2, ... -> matrix (s1'..sN') x (s1..sN)
    ax_out = list(range(0, 2 * N, 2))
    ax_in = list(range(1, 2 * N, 2))
    return op.transpose(ax_out + ax_in).reshape(d ** N, d ** N)


def full_pxp(N, h=1.0):
    d = 2
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Hf = np.zeros((d ** N, d ** N))
    for i in range(N):
        Hf += h * np.kron(np.kron(np.eye(d ** i), X), np.eye(d ** (N - i - 1)))
    allowed = np.array([0 if any((s >> (N - 1 - i)) & 1 and (s >> (N - 2 - i)) & 1 for i in range(N - 1)) else 1
                        for s in range(d ** N)], dtype=np.float64)
    Pf = np.diag(allowed)
    return Pf, Hf


def pxp_boundaries():
    # fused L index = c + 2 (a - 1) (1-based): both projector links start in state 0 -> index 0;
    # H starts at link 1 (start); combined index (L, b) = L + 4 b (0-based)
    lv = np.zeros(8)
    lv[0 + 4 * 1] = 1.0
    rv = np.zeros(8)
    for R in range(4):
        rv[R + 4 * 0] = 1.0  # any projector link state; H done
    return lv, rv


def test_oracle_sandwich_is_php():
    N = 6
    P, H = M.pxp_projector(), M.x_field(h=1.0)
    W = bs_site_dense(oracle_sandwich(P, H).parts())
    lv, rv = pxp_boundaries()
    Pf, Hf = full_pxp(N)
    np.testing.assert_array_equal(chain(W, N, lv, rv), Pf @ Hf @ Pf)


def test_oracle_mul_is_product():
    N = 5
    P, H = M.pxp_projector(), M.x_field(h=2.0)
    # (P x H) site by site = the operator P·H
    W = coo_site_dense(oracle_mul(P, H).parts())
    lv = np.zeros(4)
    lv[0 + 2 * 1] = 1.0  # P link 0, H start (fused index a + 2 b)
    rv = np.zeros(4)
    rv[0] = rv[1] = 1.0  # any P link, H done
    Pf, Hf = full_pxp(N, h=2.0)
    np.testing.assert_array_equal(chain(W, N, lv, rv), Pf @ Hf)


@pytest.mark.gpu
def test_gpu_sandwich_and_mul_vs_oracle():
    P, H = M.pxp_projector(), M.x_field(h=1.0)
    print("[mpo] sandwich", flush=True)
    g, o = M.sandwich_site(P, H).parts(), oracle_sandwich(P, H).parts()
    assert tuple(g["dims"]) == tuple(o["dims"])
    assert np.array_equal(g["keys"], o["keys"]) and np.array_equal(g["ids"], o["ids"])
    assert np.array_equal(g["data"], o["data"])
    print("[mpo] mul", flush=True)
    g, o = M.mul_site(P, H).parts(), oracle_mul(P, H).parts()
    assert tuple(g["dims"]) == tuple(o["dims"])
    assert np.array_equal(g["keys"], o["keys"]) and np.array_equal(g["vals"], o["vals"])
