"""Two-site DMRG on the GPU (H_eff plans, Lanczos, sector SVD, environment
updates through the C ABI) against exact diagonalisation of short chains:
with a bond dimension that holds the exact state the ground energy matches
ED to 1e-10; with truncation the energy is variational (above ED) and the
discarded weight is reported."""
import numpy as np
import pytest

from sparsebackends_amd import heff as H
from sparsebackends_amd import dmrg as D
from chain_ed import spin_chain, hubbard_chain, ground_energy, lowest_sparse

pytestmark = pytest.mark.gpu


def run(model, states, sweeps, **kw):
    dm = D.DMRG(model, states)
    es = [dm.sweep(**kw) for _ in range(sweeps)]
    return dm, es


@pytest.mark.parametrize("name", ["heisenberg-8", "xxz-10", "j1j2-8", "hubbard-6"])
def test_dmrg_matches_ed(name):
    if name == "heisenberg-8":
        model, states, ed, total, maxdim = H.heisenberg_model(1.0), [0, 1] * 4, spin_chain(8, [(1, 1.0, 1.0)]), (0,), 16
    elif name == "xxz-10":
        model, states, ed, total, maxdim = H.heisenberg_model(1.0, 0.7), [0, 1] * 5, \
            spin_chain(10, [(1, 1.0, 0.7)]), (0,), 32
    elif name == "j1j2-8":
        model, states, ed, total, maxdim = H.j1j2_model(1.0, 0.5), [0, 1] * 4, \
            spin_chain(8, [(1, 1.0, 1.0), (2, 0.5, 0.5)]), (0,), 16
    else:
        model, states, ed, total, maxdim = H.hubbard_model(1.0, 4.0), [1, 2] * 3, hubbard_chain(6, 1.0, 4.0), (6, 0), 64
    E0, _ = ground_energy(*ed, total)
    print(f"[dmrg] {name}", flush=True)
    dm, es = run(model, states, 6, maxdim=maxdim, cutoff=0.0, krylovdim=8, maxiter=3, tol=1e-13)
    assert abs(es[-1] - E0) <= 1e-10, (es, E0)
    assert dm.truncerr <= 1e-20


def test_dmrg_truncated_is_variational():
    """XXZ N=16 at bond dimension 8 (exact needs 2^8): the energy stays above
    the exact ground energy and approaches it as the bond dimension grows"""
    Hm, q = spin_chain(16, [(1, 1.0, 0.7)])
    E0 = lowest_sparse(Hm, q, (0,))
    model = H.heisenberg_model(1.0, 0.7)
    errs = []
    for chi in (4, 8, 16):
        dm, es = run(model, [0, 1] * 8, 4, maxdim=chi, cutoff=0.0, krylovdim=6, maxiter=2, tol=1e-12)
        assert es[-1] >= E0 - 1e-10
        assert dm.maxlinkdim() <= chi
        errs.append(es[-1] - E0)
    assert errs[2] < errs[1] < errs[0]
    assert errs[2] < 1e-4
