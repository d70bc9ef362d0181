"""GPU: sb_bs_norm_noalloc, the allocation-free sb_bs_norm, through the C ABI:
equal to sb_bs_norm bit for bit on two calls in a row (the second on a smaller
tensor, so stale scratch contents would show), with the context's device
memory cache untouched by the calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bs(rng, shape, N2):
    import sparsebackends_amd as sb
    a = rng.standard_normal(shape) * (rng.random(shape[: len(shape) - N2] + (1,) * N2) < 0.6)
    return sb.blocksparse_from_dense(np.asfortranarray(a), N2)


def test_norm_noalloc_matches_norm_twice():
    import sparsebackends_amd as sb
    from sparsebackends_amd._abi import load, check
    rng = np.random.default_rng(9100)
    big = _bs(rng, (12, 10, 9, 17, 13), 3)
    small = _bs(rng, (3, 2, 5, 4), 2)
    want = [big.norm(), small.norm()]  # the existing entry point
    out = sb.DenseTensor.from_numpy(np.full(1, -1.0))
    ptr = out.device_ptr()
    sb.synchronize()
    before = sb.memory()
    got = []
    for t in (big, small):
        check(load().sb_bs_norm_noalloc(t.ctx, t.handle, C.c_void_p(ptr)))
        sb.synchronize()
        assert sb.memory() == before  # no buffer taken from or added to the device memory cache
        got.append(float(out.numpy()[0]))
    assert got == want, (got, want)
    assert want[0] > want[1] > 0


def test_norm_noalloc_empty_and_errors():
    import sparsebackends_amd as sb
    out = sb.DenseTensor.from_numpy(np.full(1, -1.0))
    empty = sb.NewBlockSparseSorted.new((3, 4), 1)
    empty.norm_noalloc(out.device_ptr())
    assert float(out.numpy()[0]) == 0.0 == empty.norm()
    with pytest.raises(sb.SBError) as e:
        empty.norm_noalloc(0)
    assert e.value.kind == "ArgumentError"
    coo = sb.coo_from_dense(np.ones((2, 2)))
    from sparsebackends_amd._abi import load, check
    with pytest.raises(sb.SBError) as e:
        check(load().sb_bs_norm_noalloc(coo.ctx, coo.handle, C.c_void_p(out.device_ptr())))
    assert e.value.kind == "MethodError"
