"""GPU: the SB_TRACE dispatch trace (the reference prints one line per
contract! dispatch under SB_TRACE=1, contract_bs_bs.jl:138-143,
contract_bs_dense.jl:225-230): a child process with SB_TRACE=1 plans a BS x Dense
contraction, the four contract! calls of a two-site H_eff*psi and the fused H_eff plan; each prints its family, block counts
and the plan's pair count."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import sparsebackends_amd as sb
from sparsebackends_amd import heff as H
lab = lambda n: (int.from_bytes(n.encode().ljust(8, b"\0"), "little"), 0)
rng = np.random.default_rng(0)
A = sb.blocksparse_from_dense(rng.standard_normal((3, 4, 5)) * (rng.random((3, 4, 1)) < 0.7), 1)  # [i, j | x]
C2 = sb.NewBlockSparseSorted.new((3, 5, 6), 2)  # [i | x, k]
sb.contract_(C2, [lab("i"), lab("x"), lab("k")], A, [lab("i"), lab("j"), lab("x")],
             sb.DenseTensor.from_numpy(rng.standard_normal((4, 6))), [lab("j"), lab("k")])
prob = H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                     H.heisenberg_bond(20, qmax=3, width=4.0), seed=3)
prob.upload().outputs()
prob.make_plans(0)
prob.make_fused_plan(0)
print("child done", flush=True)
"""


def test_sb_trace_lines():
    print("[trace] child process with SB_TRACE=1", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, SB_TRACE="1"))
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("[SB_TRACE]")]
    assert "child done" in r.stdout
    assert any("contract_shared!" in l and "shared_bs_bs" in l and "pairs=" in l for l in lines), lines
    assert any("contract_bs_dense.contract!" in l and "-> prefix_outer_bd" in l for l in lines), lines
    assert any(l.startswith("[SB_TRACE] heff_plan") and "-> fused" in l for l in lines), lines
    # without the variable: silent
    r2 = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=120,
                        env={k: v for k, v in os.environ.items() if k != "SB_TRACE"})
    assert r2.returncode == 0 and "[SB_TRACE]" not in r2.stdout
