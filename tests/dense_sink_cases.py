"""Cases of contract_bs_bs_to_dense! (SparseBackends/src/tensoralgebra/
contract_bs_bs.jl:736-754), the P_C = 0 sink every fully dense output takes
(tensor_wrappers.jl:646-652): a BS x BS contraction into a dense array through
a P = 0 accumulator whose single block is copied out; a result without blocks
leaves the array untouched; NC = 0 writes the scalar C[] = blk[1]."""
import numpy as np

from gen_cases import dyadic


def cases():
    """(name, A, la, N2A, B, lb, N2B, lc): dyadic values, so every summation
    order gives the exact einsum"""
    out = []
    for seed in range(4):
        rng = np.random.default_rng(8000 + seed)
        s, k, x, y = (int(v) for v in rng.integers(2, 5, size=4))
        # two reduced labels (one key axis, one tail axis): contract_shared!
        A = dyadic(rng, (s, x, k), fill=1.0, P=1)
        B = dyadic(rng, (s, k, y), fill=0.7, P=1)
        out.append((f"shared{seed}", A, ["s", "x", "k"], 2, B, ["s", "k", "y"], 2, ["x", "y"]))
        # one reduced key label: contract_prefix_outer_bb! into the single block
        A1 = dyadic(rng, (s, x), fill=1.0, P=1)
        B1 = dyadic(rng, (s, y), fill=0.7, P=1)
        out.append((f"outer{seed}", A1, ["s", "x"], 1, B1, ["s", "y"], 1, ["y", "x"]))
        # NC = 0: the inner product of two tensors of one block structure
        A2 = dyadic(rng, (s, k, x, y), fill=1.0, P=2)
        B2 = dyadic(rng, (s, k, x, y), fill=1.0, P=2)
        out.append((f"scalar{seed}", A2, ["s", "k", "x", "y"], 2, B2, ["s", "k", "x", "y"], 2, []))
    return out


def empty_case():
    """A and B hold blocks at different key values: no pair matches, C keeps its contents"""
    A = np.zeros((3, 2, 2), order="F")
    A[0] = 1.0
    B = np.zeros((3, 2, 4), order="F")
    B[2] = 2.0
    return ("empty", A, ["s", "x", "k"], 2, B, ["s", "k", "y"], 2, ["x", "y"])
