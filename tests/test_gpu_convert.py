"""GPU: rekey / retypeval through the C ABI against the oracle restatement
(oracle/convert_oracle.py): accepted or refused alike, refusals as
ErrorException with the reference's message, accepted copies with the same
keys and values in the declared types."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from oracle.convert_oracle import rekey_failure, retypeval_failure

pytestmark = pytest.mark.gpu

INTS = [np.int64, np.int32, np.int16, np.int8, np.uint64, np.uint32, np.uint16, np.uint8]


def _bs(dims, N2, keys):
    tail = int(np.prod(dims[len(dims) - N2:]))
    data = np.arange(1, len(keys) * tail + 1, dtype=np.float64)
    return sb.NewBlockSparseSorted.from_parts(dims, N2, np.array(keys), np.arange(1, len(keys) + 1), data)


def _coo(dims, keys, vals):
    return sb.COOTensor.from_parts(dims, np.array(keys, dtype=np.int64).reshape(len(vals), len(dims)),
                                   np.asarray(vals, dtype=np.float64))


@pytest.mark.parametrize("K2", INTS)
@pytest.mark.parametrize("case", ["small", "dim300", "key256", "order"])
def test_rekey_matches_oracle(K2, case):
    print(f"[rekey] {np.dtype(K2).name} {case}", flush=True)
    if case == "small":
        dims, keys = (5, 7, 3), [[1, 2], [4, 7], [5, 1]]
    elif case == "dim300":
        dims, keys = (300, 4, 3), [[1, 1], [120, 4]]
    elif case == "key256":
        dims, keys = (300, 4, 3), [[256, 2], [3, 1]]
    else:  # two coordinates out of Int8's range: the first in stored order is named, not the largest
        dims, keys = (300, 4, 3), [[200, 1], [250, 2]]
    for kind, A in (("bs", _bs(dims, 1, keys)),
                    ("coo", _coo(dims[:2], keys, [1.5] * len(keys))),
                    ("aliased", sb.AliasedBlockSparse.new(dims, 1))):
        P = 2
        kk = keys if kind != "aliased" else []
        want = rekey_failure(dims if kind != "coo" else dims[:2], P, kk, K2, kind)
        if want is None:
            R = sb.rekey(A, K2)
            assert R.types()[0] is K2
            p, q = R.parts(), A.parts()
            assert p["keys"].dtype == np.dtype(K2)
            assert np.array_equal(p["keys"].astype(np.int64), q["keys"])
            for f in ("data", "vals"):
                if f in q:
                    assert np.array_equal(p[f], q[f])
        else:
            with pytest.raises(sb.SBError) as e:
                sb.rekey(A, K2)
            assert e.value.kind == "ErrorException"
            assert want in str(e.value), (want, str(e.value))
            R = sb.rekey(A, K2, check=False)  # no check: converted anyway
            assert R.types()[0] is K2


VALS = {
    "exact32": [1.5, -0.25, 2.0 ** 24, 3.0],
    "lossy32": [1.0, 2.0 ** 24 + 1, 0.1],
    "nan": [2.0, float("nan")],
    "f16": [65504.0, 0.5, -2.0],
    "f16inf": [1.0, 70000.0],
    "ints": [1.0, -128.0, 127.0, 5.0],
    "halfint": [3.0, 0.5],
    "big": [2.0 ** 31, 1.0],
}


@pytest.mark.parametrize("T2", [np.float64, np.float32, np.float16, np.int64, np.int32, np.int16, np.int8])
@pytest.mark.parametrize("name", sorted(VALS))
def test_retypeval_matches_oracle(T2, name):
    vals = VALS[name]
    print(f"[retypeval] {np.dtype(T2).name} {name}", flush=True)
    A = _coo((10, 10), [[i + 1, 2] for i in range(len(vals))], vals)
    stored = A.parts()["vals"]
    want = retypeval_failure(stored, T2)
    if want is None:
        R = sb.retypeval(A, T2)
        v = R.parts()["vals"]
        assert v.dtype == np.dtype(T2) and R.types()[1] is T2
        assert np.array_equal(v.astype(np.float64), stored)
    else:
        with pytest.raises(sb.SBError) as e:
            sb.retypeval(A, T2)
        assert e.value.kind == "ErrorException"
        msg = str(e.value)
        assert "Value " in msg
        assert {1: "non-integer", 2: "cannot be converted", 3: "roundtrip failed"}[want[1]] in msg


def test_retypeval_first_failure_in_stored_order():
    rng = np.random.default_rng(4)
    x = rng.standard_normal(5000).astype(np.float32).astype(np.float64)
    x[3777] = 0.1  # the only value Float32 cannot hold
    A = _coo((5000, 1), [[i + 1, 1] for i in range(5000)], x)
    with pytest.raises(sb.SBError) as e:
        sb.retypeval(A, np.float32)
    assert retypeval_failure(A.parts()["vals"], np.float32)[0] == 3777
    assert "0.10000000000000001" in str(e.value) or "0.1" in str(e.value)


@pytest.mark.parametrize("vals,T2,msg", [
    ([0.1], np.float32, "Value 0.1 → 0.1f0 → 0.10000000149011612: roundtrip failed; conversion to Float32 is lossy"),
    ([70000.0], np.float16, "Value 70000.0 → Inf16 → Inf: roundtrip failed; conversion to Float16 is lossy"),
    ([1.0, 0.5], np.int8, "Value 0.5 is non-integer; cannot convert to Int8 without loss"),
    ([300.0], np.int8, "Value 300.0 cannot be converted to Int8: InexactError(:Int8, Int8, 300.0)"),
])
def test_retypeval_messages(vals, T2, msg):
    """the reference's error strings with the values as Julia prints them"""
    A = _coo((10, 10), [[i + 1, 2] for i in range(len(vals))], vals)
    with pytest.raises(sb.SBError) as e:
        sb.retypeval(A, T2)
    assert str(e.value).endswith(msg), str(e.value)
