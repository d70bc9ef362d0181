"""GPU: contract_prefix_outer_bd! (BS x Dense, reduced label in A's prefix)
through the C ABI against tests/kahan_model.py, the restatement of the
reference's summation order and compensated update that is independent of
the C++ oracle, bit for bit on inexact (Gaussian) inputs."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from kahan_model import prefix_outer_bd, blocks_of
from refs import labs
from test_kahan_model import cases

pytestmark = pytest.mark.gpu


def test_gpu_equals_model_bitwise():
    print("[kahan model] 60 prefix_outer_bd cases", flush=True)
    for c in cases():
        want = prefix_outer_bd(c["A"], c["la"], c["N2A"], c["B"], c["lb"], c["lc"], c["N2C"])
        Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
        sb.contract_(Cd, labs(c["lc"]), sb.blocksparse_from_dense(c["A"], c["N2A"]), labs(c["la"]),
                     sb.DenseTensor.from_numpy(c["B"]), labs(c["lb"]))
        got = blocks_of(Cd.parts())
        assert set(got) == set(want)
        for key in want:
            assert np.array_equal(got[key], want[key]), key
