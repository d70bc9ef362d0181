"""GPU: the grouped block-product kernels over every tile shape they select
(element kernel, 16x128, 128x16, 32x32, 64x64 MFMA tiles, K loops), FP64
against the oracle at 1e-12, and the precision forms (FP32 MFMA, bf16x3)
against the FP64 oracle at their stated tolerance."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from oracle import oracle as O
from both import compare_bs
from refs import labs, rel_frobenius

pytestmark = pytest.mark.gpu

SHAPES = [(5, 5, 128), (128, 5, 5), (30, 30, 30), (64, 64, 64), (100, 37, 70), (3, 200, 3), (1, 1, 300),
          (16, 4, 16), (17, 3, 129), (8, 8, 8), (2, 2, 2), (33, 65, 31)]


def _case(M, K, N, seed):
    """A[a, s | m, k] x B[b, s | k, n] -> C[a, b | m, n]: one shared key axis,
    one shared dense axis; 2 x 2 x 2 key blocks with some missing"""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((2, 2, M, K))
    B = rng.standard_normal((2, 2, K, N))
    A[1, 0] = 0.0
    B[0, 1] = 0.0
    return A, B


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_gemm_shapes_fp64(shape):
    M, K, N = shape
    A, B = _case(M, K, N, sum(shape))
    la, lb, lc = ["a", "s", "m", "k"], ["b", "s", "k", "n"], ["a", "b", "m", "n"]
    for lcv in (lc, ["b", "a", "n", "m"]):  # AthenB and BthenA output layouts
        dims = [{"a": 2, "b": 2, "m": M, "n": N}[x] for x in lcv]
        Cd = sb.NewBlockSparseSorted.new(dims, 2)
        sb.contract_(Cd, labs(lcv), sb.blocksparse_from_dense(A, 2), labs(la), sb.blocksparse_from_dense(B, 2),
                     labs(lb))
        Co = O.BS.new(dims, 2)
        O.contract(Co, labs(lcv), O.BS.from_dense(A, 2), labs(la), O.BS.from_dense(B, 2), labs(lb))
        compare_bs(Cd, Co, rtol=1e-12)


@pytest.mark.parametrize("prec,tol", [(sb.PREC_FP32, 2e-6), (sb.PREC_BF16X3, 2e-6)])
@pytest.mark.parametrize("shape", [(5, 5, 128), (64, 64, 64), (100, 37, 70), (8, 8, 8)],
                         ids=lambda s: "x".join(map(str, s)) if isinstance(s, tuple) else str(s))
def test_gemm_precision_forms(shape, prec, tol):
    M, K, N = shape
    A, B = _case(M, K, N, 7)
    la, lb, lc = ["a", "s", "m", "k"], ["b", "s", "k", "n"], ["a", "b", "m", "n"]
    Cd = sb.NewBlockSparseSorted.new((2, 2, M, N), 2)
    sb.contract_(Cd, labs(lc), sb.blocksparse_from_dense(A, 2), labs(la), sb.blocksparse_from_dense(B, 2),
                 labs(lb), prec)
    Co = O.BS.new((2, 2, M, N), 2)
    O.contract(Co, labs(lc), O.BS.from_dense(A, 2), labs(la), O.BS.from_dense(B, 2), labs(lb))
    p, q = Cd.parts(), Co.parts()
    assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
    err = rel_frobenius(p["data"], q["data"])
    assert err <= tol, err
    assert err > 0.0 or prec == sb.PREC_FP64


def test_bs_dense_shared_large_tiles():
    """the smoke case: W(s', s | wl, wr) x T(wl, s, chi, y) -> C(s' | wr, chi, y)"""
    rng = np.random.default_rng(0)
    W = rng.standard_normal((2, 2, 5, 5))
    T = rng.standard_normal((5, 2, 16, 8))
    la, lb, lc = ["sp", "s", "wl", "wr"], ["wl", "s", "chi", "y"], ["sp", "wr", "chi", "y"]
    Cd = sb.NewBlockSparseSorted.new((2, 5, 16, 8), 3)
    sb.contract_(Cd, labs(lc), sb.blocksparse_from_dense(W, 2), labs(la), sb.DenseTensor.from_numpy(T), labs(lb))
    Co = O.BS.new((2, 5, 16, 8), 3)
    O.contract(Co, labs(lc), O.BS.from_dense(W, 2), labs(la), O.Dense.new(T), labs(lb))
    compare_bs(Cd, Co, rtol=1e-12)


def test_rank1_dense_r_large():
    """dense_bb with the MFMA path (K = R large): chunk sizes >= 16"""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 20, 40))   # [pa | a1, r]
    B = rng.standard_normal((2, 24, 40))   # [pb | b1, r]
    A[1, :, 5:9] = 0.0
    Cd = sb.NewBlockSparseSorted.new((3, 2, 20, 24), 2)
    Co = O.BS.new((3, 2, 20, 24), 2)
    la, lb, lc = ["pa", "a1", "r"], ["pb", "b1", "r"], ["pa", "pb", "a1", "b1"]
    sb.contract_(Cd, labs(lc), sb.blocksparse_from_dense(A, 2), labs(la), sb.blocksparse_from_dense(B, 2), labs(lb))
    O.contract(Co, labs(lc), O.BS.from_dense(A, 2), labs(la), O.BS.from_dense(B, 2), labs(lb))
    compare_bs(Cd, Co, rtol=1e-12)
