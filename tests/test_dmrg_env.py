"""Environment updates (makeL!/makeR!: three contract! calls each) on sector
tensors: the oracle's three calls against one dense einsum over the full
bonds (CPU), and the GPU's against the oracle (keys/ids bit-exact, values to
1e-12 relative Frobenius)."""
import numpy as np
import pytest

from sparsebackends_amd import heff as H
from sparsebackends_amd import dmrg as D
from oracle import oracle as O
from heff_dense import expand
from refs import rel_frobenius

L_ = D.lab


def bs(p):
    return O.BS.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])


def oracle_make_L(Lp, Ap, Wp):
    """the same three contractions on the oracle"""
    L, A, W = bs(Lp), bs(Ap), bs(Wp)
    dL, dA, dW = Lp["dims"], Ap["dims"], Wp["dims"]
    lL = [L_("Ql", 1), L_("Ql"), L_("wl"), L_("l", 1), L_("l")]
    lA = [L_("Ql"), L_("Qr"), L_("s"), L_("l"), L_("r")]
    lAp = [L_("Ql", 1), L_("Qr", 1), L_("s", 1), L_("l", 1), L_("r", 1)]
    lW = [L_("wl"), L_("s", 1), L_("s"), L_("wr")]
    l1 = [L_("Ql", 1), L_("Qr"), L_("wl"), L_("l", 1), L_("s"), L_("r")]
    X1 = O.BS.new((dL[0], dA[1], dL[2], dL[3], dA[2], dA[4]), 4)
    O.contract(X1, l1, L, lL, A, lA, 0)
    l2 = [L_("Ql", 1), L_("Qr"), L_("l", 1), L_("r"), L_("s", 1), L_("wr")]
    X2 = O.BS.new((dL[0], dA[1], dL[3], dA[4], dW[1], dW[3]), 4)
    O.contract(X2, l2, X1, l1, W, lW, 0)
    l3 = [L_("Qr", 1), L_("Qr"), L_("r", 1), L_("wr"), L_("r")]
    X3 = O.BS.new((dA[1], dA[1], dA[4], dW[3], dA[4]), 3)
    O.contract(X3, l3, X2, l2, A, lAp, 0)
    return X3.permutedims([1, 2, 4, 3, 5])


def oracle_make_R(Rp, Bp, Wp):
    R, B, W = bs(Rp), bs(Bp), bs(Wp)
    dR, dB, dW = Rp["dims"], Bp["dims"], Wp["dims"]
    lR = [L_("Qr", 1), L_("Qr"), L_("wr"), L_("r", 1), L_("r")]
    lB = [L_("Ql"), L_("Qr"), L_("s"), L_("l"), L_("r")]
    lBp = [L_("Ql", 1), L_("Qr", 1), L_("s", 1), L_("l", 1), L_("r", 1)]
    lW = [L_("wl"), L_("s", 1), L_("s"), L_("wr")]
    l1 = [L_("Qr", 1), L_("Ql"), L_("wr"), L_("r", 1), L_("s"), L_("l")]
    X1 = O.BS.new((dR[0], dB[0], dR[2], dR[3], dB[2], dB[3]), 4)
    O.contract(X1, l1, R, lR, B, lB, 0)
    l2 = [L_("Qr", 1), L_("Ql"), L_("r", 1), L_("l"), L_("s", 1), L_("wl")]
    X2 = O.BS.new((dR[0], dB[0], dR[3], dB[3], dW[1], dW[0]), 4)
    O.contract(X2, l2, X1, l1, W, lW, 0)
    l3 = [L_("Ql", 1), L_("Ql"), L_("l", 1), L_("wl"), L_("l")]
    X3 = O.BS.new((dB[0], dB[0], dB[3], dW[0], dB[3]), 3)
    O.contract(X3, l3, X2, l2, B, lBp, 0)
    return X3.permutedims([1, 2, 4, 3, 5])


def cases():
    yield "heisenberg", H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0), \
        H.heisenberg_bond(20, qmax=3, width=4.0)
    yield "hubbard", H.hubbard_model(1.0, 4.0), H.hubbard_bond(30, n0=0, sigma=1.2), H.hubbard_bond(28, n0=1, sigma=1.2)


@pytest.mark.parametrize("name,model,Bl,Br", list(cases()), ids=["heisenberg", "hubbard"])
def test_oracle_envs_vs_dense(name, model, Bl, Br):
    rng = np.random.default_rng(2)
    Ap = D.site_parts(model, Bl, Br, rng)
    Lp, Rp, Wp = D.env_parts(model, Bl, rng), D.env_parts(model, Br, rng), D.mpo_parts(model)
    A, _ = expand(Ap, [Bl, Br], [None, 0, 1])  # [s, l, r]
    L, _ = expand(Lp, [Bl, Bl], [None, 0, 1])  # [w, l', l]
    R, _ = expand(Rp, [Br, Br], [None, 0, 1])
    W = model.W
    Ln = oracle_make_L(Lp, Ap, Wp).parts()
    got, pad = expand(Ln, [Br, Br], [None, 0, 1])
    ref = np.einsum("wab,sbr,wtsv,tac->vcr", L, A, W, A)
    assert pad == 0.0 and np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    Rn = oracle_make_R(Rp, Ap, Wp).parts()
    got, pad = expand(Rn, [Bl, Bl], [None, 0, 1])
    ref = np.einsum("vcr,sbr,wtsv,tac->wab", R, A, W, A)
    assert pad == 0.0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name,model,Bl,Br", list(cases()), ids=["heisenberg", "hubbard"])
def test_gpu_envs_vs_oracle(name, model, Bl, Br):
    rng = np.random.default_rng(3)
    Ap = D.site_parts(model, Bl, Br, rng)
    Lp, Rp, Wp = D.env_parts(model, Bl, rng), D.env_parts(model, Br, rng), D.mpo_parts(model)
    print(f"[dmrg env] {name}", flush=True)
    A, L, R, W = D.upload(Ap), D.upload(Lp), D.upload(Rp), D.upload(Wp)
    for got, ref in ((D.make_L(L, A, W)[0], oracle_make_L(Lp, Ap, Wp)), (D.make_R(R, A, W)[0],
                                                                      oracle_make_R(Rp, Ap, Wp))):
        p, q = got.parts(), ref.parts()
        assert tuple(p["dims"]) == tuple(q["dims"])
        assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
        assert rel_frobenius(p["data"], q["data"]) <= 1e-12
