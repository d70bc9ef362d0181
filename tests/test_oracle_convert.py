"""CPU: the rekey / retypeval restatement (oracle/convert_oracle.py) pinned by
numpy's integer ranges and float round trips and by hand-worked cases."""
import numpy as np
import pytest

from oracle.convert_oracle import rekey_failure, retypeval_failure

INTS = [np.int64, np.int32, np.int16, np.int8, np.uint64, np.uint32, np.uint16, np.uint8]


@pytest.mark.parametrize("K2", INTS)
def test_rekey_ranges(K2):
    hi = np.iinfo(K2).max
    if hi < 2 ** 31:
        # a prefix dimension one past the maximum fails for block-sparse, not for COO
        assert rekey_failure((hi + 1, 3), 1, [[1], [2]], K2, "bs").startswith(f"Dimension {hi + 1} at prefix axis 1")
        assert rekey_failure((hi + 1, 3), 2, [[1, 1], [2, 3]], K2, "coo") is None
        # a coordinate one past the maximum
        assert rekey_failure((hi + 1, 3), 2, [[hi + 1, 1]], K2, "coo").startswith(f"Coordinate {hi + 1} is out of range")
    assert rekey_failure((hi if hi < 2 ** 31 else 2 ** 31 - 1, 3), 1, [[1], [2]], K2, "aliased") is None


def test_rekey_hand_cases():
    assert rekey_failure((300, 2), 1, [[5]], np.int8, "bs") == "Dimension 300 at prefix axis 1 cannot be indexed by Int8 (max=127)"
    assert rekey_failure((2, 300), 2, [[1, 1]], np.uint8, "bs") == "Dimension 300 at prefix axis 2 cannot be indexed by UInt8 (max=255)"
    assert rekey_failure((300, 2), 2, [[256, 1]], np.uint8, "coo") == "Coordinate 256 is out of range for UInt8 (0..255)"
    assert rekey_failure((10,), 1, [], np.int8, "coo") is None
    # two coordinates out of range: the first in stored order (key by key, axis by axis) is named,
    # not the smallest or the largest (block-sparse keys never exceed their prefix dimensions, which
    # are checked first, so only COO reaches this)
    assert rekey_failure((300, 300), 2, [[1, 200], [250, 2]], np.int8, "coo") == \
        "Coordinate 200 is out of range for Int8 (-128..127)"
    assert rekey_failure((300, 300), 2, [[250, 1], [1, 200]], np.int8, "coo") == \
        "Coordinate 250 is out of range for Int8 (-128..127)"


def test_retypeval_float_round_trips():
    # closed forms: 2^24 + 1 needs 25 bits (lossy in Float32), 2^24 does not;
    # 65504 is Float16's largest finite value, 70000 overflows to Inf
    assert retypeval_failure([2.0 ** 24, 1.5, -0.25], np.float32) is None
    assert retypeval_failure([1.0, 2.0 ** 24 + 1], np.float32) == (1, 3)
    assert retypeval_failure([0.1], np.float32) == (0, 3)
    assert retypeval_failure([65504.0, 0.5], np.float16) is None
    assert retypeval_failure([70000.0], np.float16) == (0, 3)
    assert retypeval_failure([np.inf], np.float32) is None
    assert retypeval_failure([np.nan], np.float32) == (0, 3)  # NaN != NaN
    rng = np.random.default_rng(0)
    x = rng.standard_normal(1000)
    exact32 = x.astype(np.float32).astype(np.float64)
    assert retypeval_failure(exact32, np.float32) is None
    assert retypeval_failure(np.concatenate([exact32, x[:1]]), np.float32) == (1000, 3)


def test_retypeval_integers():
    assert retypeval_failure([1.0, -128.0, 127.0], np.int8) is None
    assert retypeval_failure([1.0, 0.5], np.int8) == (1, 1)
    assert retypeval_failure([128.0], np.int8) == (0, 2)
    assert retypeval_failure([-129.0], np.int8) == (0, 2)
    assert retypeval_failure([np.inf], np.int32) == (0, 1)
    assert retypeval_failure([2.0 ** 31 - 1, -2.0 ** 31], np.int32) is None
    assert retypeval_failure([2.0 ** 31], np.int32) == (0, 2)
    assert retypeval_failure([2.0 ** 63], np.int64) == (0, 2)
    assert retypeval_failure([-2.0 ** 63], np.int64) is None
