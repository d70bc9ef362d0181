"""The two-site H_eff*psi chain (sparsebackends_amd/heff.py) run on the oracle.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline): builds oracle tensors
from the same host parts the device problem uploads and runs the same four
contract! calls.
"""
from . import oracle as O


def _bs(p):
    return O.BS.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])


class HeffOracle:
    def __init__(self, prob):
        self.p = prob
        self.psi, self.L, self.R = _bs(prob.psi_parts), _bs(prob.L_parts), _bs(prob.R_parts)
        self.W1, self.W2 = _bs(prob.W_parts), _bs(prob.W_parts)

    def run(self, flags=0):
        p = self.p
        T1 = O.BS.new(p.dims_T1, 5)
        O.contract(T1, p.l_T1, self.psi, p.l_psi, self.L, p.l_L, flags)
        T2 = O.BS.new(p.dims_T2, 5)
        O.contract(T2, p.l_T2, T1, p.l_T1, self.W1, p.l_W1, flags)
        T3 = O.BS.new(p.dims_T3, 5)
        O.contract(T3, p.l_T3, T2, p.l_T2, self.W2, p.l_W2, flags)
        Hp = O.BS.new(p.dims_out, 4)
        O.contract(Hp, p.l_out, T3, p.l_T3, self.R, p.l_R, flags)
        self.T1, self.T2, self.T3, self.Hpsi = T1, T2, T3, Hp
        return Hp
