"""GPU: a one-sided Jacobi SVD that reaches its sweep cap reports
LAPACKException (SB_E_NOCONV; the reference's LAPACK SVD raises
LAPACKException when it fails to converge) instead of returning unconverged
factors: the cap set to one sweep on inputs that need several, then back to
the default, where the same calls match the oracle.  Every kernel still runs
to its end; only the host's verdict differs."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import dmrg as D
from sparsebackends_amd import factorize as Fz
from sparsebackends_amd import heff as H
from oracle import factorize_oracle as F
from test_oracle_factorize import random_parts

pytestmark = pytest.mark.gpu


def _sector_problem(chi):
    return H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(chi, qmax=3, width=4.0),
                         H.heisenberg_bond(chi - 10, qmax=3, width=4.0), seed=9)


@pytest.mark.parametrize("chi", [24, 240], ids=["lds-kernel", "block-jacobi"])
def test_sector_svd_sweep_cap(chi):
    prob = _sector_problem(chi)
    rowsec, colsec, charges = D.fused_sectors(prob.m, prob.Bl, prob.Br)
    psi = D.upload(prob.psi_parts)
    print(f"[svd cap] sector chi={chi}", flush=True)
    try:
        sb.set_svd_sweeps(1)
        with pytest.raises(sb.SBError) as e:
            Fz.svd_sectors(psi, rowsec, colsec, prob.Bl.dims, prob.Br.dims, len(charges))
        assert e.value.kind == "LAPACKException"
    finally:
        sb.set_svd_sweeps(40)
    A, B, svs, kept, spec = Fz.svd_sectors(psi, rowsec, colsec, prob.Bl.dims, prob.Br.dims, len(charges))
    so = F.svd_sectors(prob.psi_parts, rowsec, colsec, prob.Bl.dims, prob.Br.dims, len(charges))[2]
    assert np.abs(np.sort(svs) - np.sort(so)).max() <= 1e-13 * so.max()


def test_blocksparse_svd_sweep_cap():
    p, _ = random_parts(np.random.default_rng(3), 6, 7, 4, 4, fill=0.6)
    A = sb.NewBlockSparseSorted.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])
    print("[svd cap] blocksparse_svd", flush=True)
    try:
        sb.set_svd_sweeps(1)
        with pytest.raises(sb.SBError) as e:
            sb.blocksparse_svd(A)
        assert e.value.kind == "LAPACKException"
    finally:
        sb.set_svd_sweeps(40)
    svs = sb.blocksparse_svd(A)[2]
    so = F.blocksparse_svd(p)[2]
    assert np.abs(svs - so).max() <= 1e-13 * so.max()


def test_sweep_cap_argument():
    with pytest.raises(sb.SBError) as e:
        sb.set_svd_sweeps(0)
    assert e.value.kind == "ArgumentError"
