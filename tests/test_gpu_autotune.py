"""The box product autotuner (SB_FLAG_AUTOTUNE): an autotuned H_eff*psi plan
chain gives exactly the result of the cost-model plans (every tile shape and
grid density computes the same products in the same order), every plan keeps
a valid workgroup tile, and a second chain with the same block-size
histograms reuses the cached choices without timing."""
import copy

import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import heff as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [sb.PREC_FP64, sb.PREC_FP32])
def test_autotune_result_and_cache(prec):
    print(f"[autotune] prec {prec}", flush=True)
    sb.load().sb_autotune_reset()
    prob = H.heisenberg_chi256_problem(seed=1)
    prob.upload().outputs()
    prob.make_plans(prec)
    prob.execute()
    ref = prob.Hpsi.parts()
    tuned = copy.copy(prob).upload().outputs()
    plans = tuned.make_plans(prec | sb.AUTOTUNE)
    cfg = [p.box_config() for p in plans]
    assert all(0 <= c["tile"] < 14 for c in cfg), cfg
    assert all(c["candidates"] > 0 or c["best_ms"] > 0 for c in cfg), cfg
    assert sum(c["candidates"] for c in cfg) > 0
    tuned.execute()
    got = tuned.Hpsi.parts()
    assert np.array_equal(got["keys"], ref["keys"]) and np.array_equal(got["ids"], ref["ids"])
    assert np.array_equal(got["data"], ref["data"])
    # the same histograms again: cached, no timing, same choices
    again = copy.copy(prob).upload().outputs()
    plans2 = again.make_plans(prec | sb.AUTOTUNE)
    cfg2 = [p.box_config() for p in plans2]
    assert all(c["candidates"] == 0 for c in cfg2), cfg2
    assert [(c["tile"], c["wg_per_cu"]) for c in cfg2] == [(c["tile"], c["wg_per_cu"]) for c in cfg]
    again.execute()
    assert np.array_equal(again.Hpsi.parts()["data"], ref["data"])
