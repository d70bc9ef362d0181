"""GPU: the block-creation order (C.ids) of every contraction family against
tests/ids_model.py, the restatement of the reference's traversals that is
independent of the C++ oracle, on every reference fixture and random case."""
import numpy as np
import pytest

from both import dev_tensor
from ids_model import creation_order, check_ids
from refs import labs
from sb_flags import EXACT_JOIN, SKIP_ZERO_SLICES
from test_ids_model import fixture_calls, random_calls

pytestmark = pytest.mark.gpu


def _dev_check(ck, cdims, N2C, lc, ak, A, N2A, la, bk, B, N2B, lb, flags=0):
    import sparsebackends_amd as sb
    Cd = dev_tensor(ck, dims=cdims, N2=N2C)
    sb.contract_(Cd, labs(lc), dev_tensor(ak, A, N2A, atol=1e-12), labs(la), dev_tensor(bk, B, N2B, atol=1e-12),
                 labs(lb), flags)
    exp = creation_order(ck, lc, N2C, ak, A, la, N2A, bk, B, lb, N2B, exact_join=bool(flags & EXACT_JOIN),
                         skip_zero_slices=bool(flags & SKIP_ZERO_SLICES))
    p = Cd.parts()
    if ck == "coo":
        assert [tuple(int(x) for x in k) for k in np.asarray(p["keys"]).reshape(-1, len(lc))] == exp
    else:
        check_ids(p, exp)
    return len(exp)


def test_gpu_fixtures_creation_order():
    n = [_dev_check(*call) for call in fixture_calls()]
    assert sum(1 for x in n if x > 1) > 20


def test_gpu_random_creation_order():
    n = [_dev_check(*call) for call in random_calls()]
    assert sum(1 for x in n if x > 2) > 100
