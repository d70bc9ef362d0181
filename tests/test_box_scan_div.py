"""The 32-bit index decode of the box scan and of the Krylov vectors' compact
layout (sb_kernels.h:div24): quotient by a float reciprocal with one
correction step, used for blocks of fewer than 2^24 values.  Emulated here in IEEE single precision (numpy float32: the
reciprocal rounded on the host, the product rounded to nearest, truncation
toward zero, as the device's instructions do) and checked against integer
division on every divisor up to 4096, powers of two and large divisors, at
random and at the edges (multiples of d and their neighbours, values just
below 2^24)."""
import numpy as np


def scan_div(n, d):
    inv = np.float32(1.0) / np.float32(d)
    q = (n.astype(np.float32) * inv).astype(np.int64)  # float32 product, truncated
    r = n - q * d
    q = np.where(r >= d, q + 1, np.where(r < 0, q - 1, q))
    return q


def test_scan_div_exact_below_2_24():
    rng = np.random.default_rng(0)
    top = 1 << 24
    divisors = list(range(1, 4097)) + [1 << k for k in range(13, 24)] + \
        list(rng.integers(4097, top, 400)) + [top - 1, top // 3, 65537, 1000003]
    for d in divisors:
        d = int(d)
        n = np.concatenate([rng.integers(0, top, 2000),
                            np.arange(max(0, top - 64), top),
                            (np.arange(1, 64) * d) % top,
                            ((np.arange(1, 64) * d) - 1) % top,
                            ((np.arange(1, 64) * d) + 1) % top]).astype(np.int64)
        q = scan_div(n, d)
        assert np.array_equal(q, n // d), d
        r = n - q * d
        assert np.all((r >= 0) & (r < d)), d
