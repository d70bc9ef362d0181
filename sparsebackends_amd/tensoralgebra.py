"""contract! and contraction plans over the device storage types
(SparseBackends/src/tensoralgebra/contract.jl:37-104)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import load, check
from .storage import DenseTensor, _context_obj

# flags (include/sb_status.h)
EXACT_JOIN = 1 << 0
SKIP_ZERO_SLICES = 1 << 1
AUTOTUNE = 1 << 2
PREC_FP64 = 0 << 8
PREC_FP32 = 1 << 8
PREC_BF16X3 = 2 << 8


def labels_array(labels):
    """A label is (id, plev) (tensor_wrappers.jl:577-580) -> uint64 array of 2*n."""
    a = np.zeros(2 * len(labels), dtype=np.uint64)
    for i, (lid, plev) in enumerate(labels):
        a[2 * i] = lid
        a[2 * i + 1] = plev
    return a


def _as_operand(x, device):
    if isinstance(x, np.ndarray):
        return DenseTensor.from_numpy(x, device)
    return x


def contract_(C_, labelsC, A, labelsA, B, labelsB, flags=0):
    """contract!(C, labelsC, A, labelsA, B, labelsB): C (made with
    NewBlockSparseSorted.new / COOTensor.new) is overwritten; A and B are not
    modified.  B may be a numpy array (uploaded as a dense operand)."""
    dev = C_.device
    A = _as_operand(A, dev)
    B = _as_operand(B, dev)
    lc, la, lb = labels_array(labelsC), labels_array(labelsA), labels_array(labelsB)
    u = C.POINTER(C.c_uint64)
    check(load().sb_contract(C_.ctx, C_.handle, lc.ctypes.data_as(u), len(labelsC), A.handle, la.ctypes.data_as(u),
                             len(labelsA), B.handle, lb.ctypes.data_as(u), len(labelsB), flags))
    return C_


def contract_bs_bs_to_dense_(Cd, labelsC, A, labelsA, B, labelsB, flags=0):
    """contract_bs_bs_to_dense! (contract_bs_bs.jl:736-754): P_C = 0 sink."""
    lc, la, lb = labels_array(labelsC), labels_array(labelsA), labels_array(labelsB)
    u = C.POINTER(C.c_uint64)
    check(load().sb_contract_bs_bs_to_dense(Cd.ctx, Cd.handle, lc.ctypes.data_as(u), len(labelsC), A.handle,
                                            la.ctypes.data_as(u), len(labelsA), B.handle, lb.ctypes.data_as(u),
                                            len(labelsB), flags))
    return Cd


class Plan:
    """contract! with the block structure fixed: creation runs the contraction
    once; execute() recomputes C's values from the operands' current values."""

    def __init__(self, C_, labelsC, A, labelsA, B, labelsB, flags=0):
        dev = C_.device
        self._ops = (C_, _as_operand(A, dev), _as_operand(B, dev))  # keep alive: the plan holds their buffers
        lc, la, lb = labels_array(labelsC), labels_array(labelsA), labels_array(labelsB)
        u = C.POINTER(C.c_uint64)
        h = C.c_void_p()
        check(load().sb_plan_create(C_.ctx, C_.handle, lc.ctypes.data_as(u), len(labelsC), self._ops[1].handle,
                                    la.ctypes.data_as(u), len(labelsA), self._ops[2].handle, lb.ctypes.data_as(u),
                                    len(labelsB), flags, C.byref(h)))
        self._h = h
        self._owner = _context_obj(dev)  # the context the plan's buffers and stream belong to
        self.ctx = self._owner.h

    def execute(self):
        check(load().sb_plan_execute(self.ctx, self._h))

    def stats(self):
        f, gb, pb = C.c_double(), C.c_double(), C.c_double()
        npairs, nblk = C.c_int64(), C.c_int64()
        check(load().sb_plan_stats(self._h, C.byref(f), C.byref(gb), C.byref(pb), C.byref(npairs), C.byref(nblk)))
        return dict(flops=f.value, gemm_bytes=gb.value, perm_bytes=pb.value, npairs=npairs.value,
                    nblocks_out=nblk.value)

    def box_config(self):
        """the box product's launch shape: tile (-1: not a box product),
        workgroups per CU (0: all resident), autotuner candidates timed (0:
        not timed) and the best time measured (ms)"""
        t, w, n = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        check(load().sb_plan_box_config(self._h, C.byref(t), C.byref(w), C.byref(n), C.byref(ms)))
        return dict(tile=t.value, wg_per_cu=w.value, candidates=n.value, best_ms=ms.value)

    def mfma_flops(self):
        """matrix-core FLOPs one execution issues, tile padding included (0: not counted)"""
        v = C.c_double()
        check(load().sb_plan_mfma_flops(self._h, C.byref(v)))
        return v.value

    def operand_bytes(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(load().sb_plan_operand_bytes(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(a=a.value, b=b.value, c=c.value)

    def __del__(self):
        try:
            if self._h:
                load().sb_plan_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


class HeffPlan(Plan):
    """The two-site H_eff*psi of DMRG (the four contract! calls psi*L*W1*W2*R)
    as one plan executing one fused kernel (sb_heff_plan_create): Hpsi (made
    with NewBlockSparseSorted.new(psi.dims, 4)) gets the chain's block list and,
    at every execute(), the chain's value for psi's current values.  L, R, W1
    and W2 must keep their values; flags: PREC_FP64 or PREC_FP32.  Raises
    SBError of kind "LimitExceeded" (SB_E_LIMIT) when a sector block is too large for the
    fused kernel; the four Plan objects then do the work."""

    def __init__(self, Hpsi, psi, L, W1, W2, R, flags=0):
        self._ops = (Hpsi, psi, L, W1, W2, R)  # Hpsi first: eigsolve's default output
        h = C.c_void_p()
        check(load().sb_heff_plan_create(Hpsi.ctx, Hpsi.handle, psi.handle, L.handle, W1.handle, W2.handle,
                                         R.handle, flags, C.byref(h)))
        self._h = h
        self._owner = _context_obj(Hpsi.device)
        self.ctx = self._owner.h


class PlanBatch:
    """Independent plan chains executed together: chains[c] is the list of
    plans of problem c (every chain the same length).  Stage s of all chains
    runs as one box product launch when each of its plans is one (the strided
    BS x BS route), with results identical to executing the plans one by one;
    other stages run plan by plan.  execute(active) runs the chains numbered in
    active (default all)."""

    def __init__(self, chains):
        chains = [list(ch) for ch in chains]
        if not chains or not chains[0]:
            raise ValueError("a batch needs at least one chain of at least one plan")
        ns = len(chains[0])
        if any(len(ch) != ns for ch in chains):
            raise ValueError("every chain of a batch must have the same number of plans")
        self._chains = chains  # the batch refers to the plans
        self.ctx = chains[0][0].ctx
        arr = (C.c_void_p * (len(chains) * ns))(*[p._h.value for ch in chains for p in ch])
        h = C.c_void_p()
        check(load().sb_batch_create(self.ctx, C.cast(arr, C.POINTER(C.c_void_p)), len(chains), ns, C.byref(h)))
        self._h = h
        self.nchains, self.nstages = len(chains), ns

    def execute(self, active=None):
        if active is None:
            check(load().sb_batch_execute(self.ctx, self._h, None, 0))
        else:
            a = np.ascontiguousarray(active, dtype=np.int32)
            check(load().sb_batch_execute(self.ctx, self._h, a.ctypes.data_as(C.POINTER(C.c_int)), len(a)))

    def stats(self):
        f, n = C.c_double(), C.c_int()
        check(load().sb_batch_stats(self._h, C.byref(f), C.byref(n)))
        return dict(flops=f.value, fused_stages=n.value)

    def autotune(self):
        """time every workgroup tile shape for each fused stage over all chains, keep the fastest"""
        check(load().sb_batch_autotune(self.ctx, self._h))
        return self

    def stage_tiles(self):
        """per stage: (tile configuration or -1 if not fused, measured ms or 0)"""
        out = []
        for st in range(self.nstages):
            t, ms = C.c_int(), C.c_double()
            check(load().sb_batch_stage_tile(self._h, st, C.byref(t), C.byref(ms)))
            out.append((t.value, ms.value))
        return out

    def __del__(self):
        try:
            if self._h:
                load().sb_batch_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


class PlanGraph:
    """A chain of plans executed in order as one HIP graph (sb_graph_*): the
    first launch() runs eagerly, the second is captured, later ones replay the
    graph while the chain's buffers and box decisions match the capture.
    Results equal executing the plans one by one."""

    def __init__(self, plans):
        """plans: a list of Plan objects, or a PlanBatch (one execution of all
        its chains per launch)"""
        h = C.c_void_p()
        if isinstance(plans, PlanBatch):
            self._plans = plans  # the graph refers to the batch
            self.ctx = plans.ctx
            check(load().sb_batch_graph_create(self.ctx, plans._h, C.byref(h)))
        else:
            plans = list(plans)
            if not plans:
                raise ValueError("a plan chain needs at least one plan")
            self._plans = plans  # the graph refers to the plans
            self.ctx = plans[0].ctx
            arr = (C.c_void_p * len(plans))(*[p._h.value for p in plans])
            check(load().sb_graph_create(self.ctx, C.cast(arr, C.POINTER(C.c_void_p)), len(plans), C.byref(h)))
        self._h = h

    def launch(self):
        check(load().sb_graph_launch(self.ctx, self._h))

    def stats(self):
        e, c, r, n = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(load().sb_graph_stats(self._h, C.byref(e), C.byref(c), C.byref(r), C.byref(n)))
        return dict(eager=e.value, captures=c.value, replays=r.value, nodes=n.value)

    def __del__(self):
        try:
            if self._h:
                load().sb_graph_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def set_graphs(mode, device=0):
    """eigensolves replay their operator as a HIP graph: 1 always, 0 never,
    -1 (default) when krylovdim * maxiter >= 8"""
    from .storage import context
    check(load().sb_ctx_graphs(context(device), int(mode)))
