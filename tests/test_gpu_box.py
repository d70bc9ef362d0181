"""GPU parity of the box-aware BS x BS block products (contract_shared!).

Operands carry zero padding in random boxes per block (the layout of padded
symmetry sectors), shuffled ids, stored all-zero blocks.  Checked against the
oracle, which multiplies the full blocks: keys and ids bit-exact; payloads
bit-exact for dyadic values, within 1e-12 relative Frobenius otherwise; every
tile shape; plan replays whose boxes shrink and grow; a C payload written
from outside; non-finite values (Inf/NaN must spread exactly as through the
full products).
"""
import ctypes

import numpy as np
import pytest

import sparsebackends_amd as sb
from oracle import oracle as O
from box_cases import gen_padded_case, padded_parts, gauss_vals, dyadic_vals
from refs import labs, rel_frobenius
from both import compare_bs

pytestmark = pytest.mark.gpu
TILE_SHIFT = 16
NTILES = 14


def dev(p):
    return sb.NewBlockSparseSorted.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])


def orc(p):
    return O.BS.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])


def run_case(c, flags=0, rtol=0.0):
    Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
    Co = O.BS.new(c["cdims"], c["N2C"])
    sb.contract_(Cd, labs(c["lc"]), dev(c["A"]), labs(c["la"]), dev(c["B"]), labs(c["lb"]), flags)
    O.contract(Co, labs(c["lc"]), orc(c["A"]), labs(c["la"]), orc(c["B"]), labs(c["lb"]), flags & 0xF)
    compare_bs(Cd, Co, rtol=rtol)
    return Cd, Co


@pytest.mark.parametrize("seed", range(60))
def test_padded_random_exact(seed):
    rng = np.random.default_rng(1000 + seed)
    c = gen_padded_case(rng, big=seed % 3 == 0)
    print(f"[box] seed {seed} la={c['la']} lb={c['lb']} lc={c['lc']}", flush=True)
    run_case(c)


@pytest.mark.parametrize("tile", range(1, NTILES + 1))
def test_every_tile_shape(tile):
    for seed in range(6):
        rng = np.random.default_rng(77 + 10 * tile + seed)
        c = gen_padded_case(rng, big=True)
        print(f"[box tile {tile}] seed {seed}", flush=True)
        run_case(c, flags=tile << TILE_SHIFT)


@pytest.mark.parametrize("prec", [sb.PREC_FP32, sb.PREC_BF16X3])
@pytest.mark.parametrize("tile", [0, 1, 2, 6, 9, 12])
def test_precision_forms_padded(prec, tile):
    rng = np.random.default_rng(5 + tile)
    c = gen_padded_case(rng, big=True, vals=gauss_vals)
    print(f"[box prec {prec} tile {tile}]", flush=True)
    Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
    Co = O.BS.new(c["cdims"], c["N2C"])
    sb.contract_(Cd, labs(c["lc"]), dev(c["A"]), labs(c["la"]), dev(c["B"]), labs(c["lb"]), prec | (tile << TILE_SHIFT))
    O.contract(Co, labs(c["lc"]), orc(c["A"]), labs(c["la"]), orc(c["B"]), labs(c["lb"]), 0)
    p, q = Cd.parts(), Co.parts()
    assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
    assert rel_frobenius(p["data"], q["data"]) <= 2e-6


def chain_case(rng, K1=24, K2=30, M=40, N=36, vals=gauss_vals):
    """A [a, s | k1, m, k2] x B [s, b | n, k2, k1] -> C [a, b | m, n]: long K
    (k1*k2 = 720 > one K table), several pairs per output block, multi-tile M, N"""
    A = padded_parts(rng, [3, 4], [K1, M, K2], 9, vals=vals, empty_frac=0.0)
    B = padded_parts(rng, [4, 3], [N, K2, K1], 9, vals=vals, empty_frac=0.0)
    la, lb, lc = ["a", "s", "k1", "m", "k2"], ["s", "b", "n", "k2", "k1"], ["a", "b", "m", "n"]
    return dict(A=A, la=la, B=B, lb=lb, lc=lc, cdims=(3, 3, M, N), N2C=2)


@pytest.mark.parametrize("tile", [0, 1, 2, 6, 9, 12, 14])
def test_long_k_many_pairs(tile):
    rng = np.random.default_rng(31)
    print(f"[box long K tile {tile}]", flush=True)
    run_case(chain_case(rng), flags=tile << TILE_SHIFT, rtol=1e-12)
    run_case(chain_case(np.random.default_rng(32), vals=dyadic_vals), flags=tile << TILE_SHIFT)


def test_replay_boxes_shrink_and_grow():
    """a plan replayed after the operand values change so that the boxes of C
    shrink (stale entries must be cleared) and grow again"""
    rng = np.random.default_rng(9)
    c = chain_case(rng, K1=6, K2=5, M=20, N=18, vals=dyadic_vals)
    A, B = dev(c["A"]), dev(c["B"])
    Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
    plan = sb.Plan(Cd, labs(c["lc"]), A, labs(c["la"]), B, labs(c["lb"]))

    def check(Ap):
        Co = O.BS.new(c["cdims"], c["N2C"])
        O.contract(Co, labs(c["lc"]), orc(Ap), labs(c["la"]), orc(c["B"]), labs(c["lb"]), 0)
        compare_bs(Cd, Co)

    check(c["A"])
    full = c["A"]["data"].copy()
    tdims = c["A"]["dims"][2:]
    bs = int(np.prod(tdims))
    for cut in (7, 3, 1, 20, 11):
        x = full.copy().reshape(-1, bs)
        blk = x.reshape(-1, *tdims[::-1])  # C-order view: axes (k2, m, k1) reversed
        blk[:, :, cut:, :] = 0.0  # m >= cut zero in every block
        A.set_data_(x.reshape(-1))
        plan.execute()
        check(dict(c["A"], data=x.reshape(-1)))


def test_replay_c_written_from_outside():
    rng = np.random.default_rng(10)
    c = chain_case(rng, K1=5, K2=4, M=24, N=20, vals=dyadic_vals)
    A, B = dev(c["A"]), dev(c["B"])
    Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
    plan = sb.Plan(Cd, labs(c["lc"]), A, labs(c["la"]), B, labs(c["lb"]))
    Co = O.BS.new(c["cdims"], c["N2C"])
    O.contract(Co, labs(c["lc"]), orc(c["A"]), labs(c["la"]), orc(c["B"]), labs(c["lb"]), 0)
    compare_bs(Cd, Co)
    Cd.set_data_(np.full(Cd.nblocks * Cd.blksize, 3.5))  # garbage everywhere, padding included
    plan.execute()
    compare_bs(Cd, Co)


def test_replay_operand_written_through_device_pointer():
    """writes through sb_bs_device_ptrs are not tracked: the payload is rescanned"""
    rng = np.random.default_rng(11)
    c = chain_case(rng, K1=5, K2=4, M=16, N=12, vals=dyadic_vals)
    A, B = dev(c["A"]), dev(c["B"])
    Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
    plan = sb.Plan(Cd, labs(c["lc"]), A, labs(c["la"]), B, labs(c["lb"]))
    _, _, dptr = A.device_ptrs()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    tdims = c["A"]["dims"][2:]
    bs = int(np.prod(tdims))
    for fill in (1.0, 0.0):
        x = c["A"]["data"].copy().reshape(-1, bs)
        blk = x.reshape(-1, *tdims[::-1])
        blk[:, :, 3:, :] = 0.0 if fill == 0.0 else 0.5  # grow the boxes to the full m range, then shrink
        x = np.ascontiguousarray(x.reshape(-1))
        sb.synchronize()
        assert hip.hipMemcpy(dptr, x.ctypes.data, x.nbytes, 1) == 0  # host to device
        sb.synchronize()
        plan.execute()
        Co = O.BS.new(c["cdims"], c["N2C"])
        O.contract(Co, labs(c["lc"]), orc(dict(c["A"], data=x)), labs(c["la"]), orc(c["B"]), labs(c["lb"]), 0)
        compare_bs(Cd, Co)


@pytest.mark.parametrize("bad", [np.inf, -np.inf, np.nan])
def test_non_finite_spreads_like_full_products(bad):
    """a non-finite entry of A inside its box meets zero padding of B: the full
    products give NaN there (inf * 0), and so must the box-aware product"""
    rng = np.random.default_rng(12)
    c = chain_case(rng, K1=5, K2=4, M=8, N=6, vals=dyadic_vals)
    a = c["A"]
    tdims = a["dims"][2:]
    bs = int(np.prod(tdims))
    x = a["data"].copy()
    j = int(a["ids"][0] - 1) * bs + int(np.ravel_multi_index((4, 2, 3), tdims, order="F"))
    x[j] = bad
    c["A"] = dict(a, data=x)
    b = c["B"]
    # B's boxes small along k1/k2 so the padding meets the non-finite entry
    y = b["data"].copy().reshape(-1, int(np.prod(b["dims"][2:])))
    yb = y.reshape(-1, *b["dims"][2:][::-1])
    yb[:, 3:, :, :] = 0.0  # k1 >= 3 zero
    c["B"] = dict(b, data=y.reshape(-1))
    Cd, Co = run_case_nan(c)


def run_case_nan(c):
    Cd = sb.NewBlockSparseSorted.new(c["cdims"], c["N2C"])
    Co = O.BS.new(c["cdims"], c["N2C"])
    sb.contract_(Cd, labs(c["lc"]), dev(c["A"]), labs(c["la"]), dev(c["B"]), labs(c["lb"]), 0)
    O.contract(Co, labs(c["lc"]), orc(c["A"]), labs(c["la"]), orc(c["B"]), labs(c["lb"]), 0)
    p, q = Cd.parts(), Co.parts()
    assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
    assert np.isnan(q["data"]).any() or np.isinf(q["data"]).any()
    assert np.array_equal(np.isnan(p["data"]), np.isnan(q["data"]))
    fin = np.isfinite(q["data"])
    assert np.array_equal(np.isfinite(p["data"]), fin)
    assert np.array_equal(p["data"][fin], q["data"][fin])
    return Cd, Co
