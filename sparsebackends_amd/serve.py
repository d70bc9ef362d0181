"""A serving loop for DMRG ground-state solves with continuous batching.

No counterpart in the reference, which runs one DMRG at a time from a script
(test_product_dmrg_state.jl).  Requests (XXZ chains of different lengths,
bond-dimension caps and anisotropies) arrive over time; the server admits them
as slots and device memory free up, keeps every admitted solve's MPS and
environments resident on the device, and co-batches the Krylov steps of the
solves in flight: a solve that reaches its eigensolve hands (plans, psi) to a
batcher, which runs the pending eigensolves of all slots as ONE lock-step
Lanczos (one batched operator application and one host synchronisation per
iteration, sb_eigsolve_lanczos_multi).  Everything else a solve does (its
H_eff plans, SVD and environment updates) runs in its slot's host thread on
the slot's own context and stream, so the device works on several solves at
once.

Each solve is deterministic and the batched eigensolve is bit-identical to
the solo one, so a request's result does not depend on what it was batched
with.

    reqs = poisson_requests(32, rate=8.0, seed=0)
    results = DMRGServer(slots=8).run(reqs)
    plain = run_plain(reqs)          # one solve at a time, first come first served
    summarize(results), summarize(plain)
"""
from __future__ import annotations

import gc
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import dmrg as D
from . import heff as H
from .krylov import eigsolve, eigsolve_multi
from .storage import memory, synchronize, trim


@dataclass
class SolveRequest:
    """A ground-state request: XXZ chain (J = 1, Jz = delta) of `sites`
    spins-1/2 in the Sz = 0 sector, two-site DMRG from the Neel state with
    bond dimension at most chi_max, sweeps until the energy changes by less
    than tol between sweeps (at most max_sweeps).  arrival: seconds after the
    start of the run."""
    rid: int
    sites: int
    chi_max: int
    delta: float
    arrival: float = 0.0
    tol: float = 1e-10
    max_sweeps: int = 12
    cutoff: float = 1e-12
    krylovdim: int = 3
    maxiter: int = 1
    flags: int = 0  # precision form of the block products (PREC_FP64 / PREC_FP32 / PREC_BF16X3)
    # a low-precision request sweeps with those products until the energy changes
    # by less than polish_tol * |E| between sweeps (or polish_sweeps sweeps are
    # left), then with FP64 products until tol: its energy is an FP64 result
    polish_tol: float = 1e-6
    polish_sweeps: int = 4


@dataclass
class SolveResult:
    rid: int
    energy: float
    sweeps: int
    converged: bool
    energies: list = field(repr=False)
    arrival: float = 0.0
    admitted: float = 0.0
    finish: float = 0.0
    truncerr: float = 0.0  # largest discarded weight of the last sweep
    footprint: int = 0     # admission estimate, bytes
    peak_bytes: int = 0    # device bytes the solve held after a sweep, largest over its sweeps
    maxlinkdim: int = 0

    @property
    def latency(self):
        return self.finish - self.arrival


def footprint_bytes(req: SolveRequest, d=2, w=5):
    """Upper estimate of the device bytes a solve holds: site tensors and
    environments of every site at the cap chi_max (a sweep keeps one side's
    environments), plus one step's working set (two-site psi, the three H_eff
    intermediates, Krylov vectors).  Blocks of the sector format are padded to
    the bond's largest sector; pad = 6 times the dense size covers the U(1)
    bond distributions of these chains: measured peaks (to convergence or 12-20
    sweeps) were 0.29-0.81 of it at (sites, chi_max) from (32, 128) to
    (128, 512) and (64, 1024), the largest at (128, 512)."""
    chi, n, pad = req.chi_max, req.sites, 6
    site = d * chi * chi
    env = w * chi * chi
    step = (req.krylovdim + 4) * d * d * chi * chi + 3 * w * d * d * chi * chi
    return int(8 * pad * (n * (site + env) + step))


def neel(n):
    return [0, 1] * (n // 2) + [0] * (n % 2)


def solve_dmrg(req: SolveRequest, eigsolver=None, device=0, on_sweep=None):
    """Runs one request to completion on the calling thread's context.
    Returns (energy, energies per sweep, converged, maxlinkdim, peak bytes,
    largest discarded weight of the last sweep):
    peak bytes = the context's device bytes in use after a sweep beyond what
    it held before the solve, at most over the sweeps.  on_sweep(req, k,
    energy, low): called after every sweep (low: it used the low-precision
    products)."""
    base = memory(device)[0]
    dm = D.DMRG(H.heisenberg_model(1.0, float(req.delta)), neel(req.sites), device=device, flags=req.flags,
                eigsolver=eigsolver)
    es, conv, peak = [], False, 0
    low = req.flags != 0  # in the low-precision stage
    prev = None  # the previous sweep's energy in the current stage
    for k in range(req.max_sweeps):
        e = dm.sweep(maxdim=req.chi_max, cutoff=req.cutoff, krylovdim=req.krylovdim, maxiter=req.maxiter, tol=1e-14)
        peak = max(peak, memory(device)[0] - base)
        es.append(e)
        if on_sweep is not None:
            on_sweep(req, k, e, low)
        settled = prev is not None and abs(e - prev) < (req.polish_tol * abs(e) if low else req.tol)
        prev = e
        if low and (settled or req.max_sweeps - 1 - k <= req.polish_sweeps):
            low, prev = False, None
            dm.flags = 0  # the FP64 polish: the next sweeps' plans use FP64 products
            continue
        if settled:
            conv = True
            break
    synchronize(device)
    return es[-1], es, conv, dm.maxlinkdim(), peak, dm.truncerr


class KrylovBatcher:
    """Co-batches the eigensolves that slots hand in, without a thread of its own.

    A slot that reaches its eigensolve queues it; the first queued slot leads:
    it waits up to `window` seconds for the other busy slots to queue theirs,
    takes every queued job, and runs them as one lock-step batched Lanczos on
    its own stream (the other slots' contexts lend theirs and the slots wait
    for their results).  Slots arriving meanwhile form the next batch under a
    new leader, so batches run concurrently on different streams.  A lone job
    runs as the plain eigensolve."""

    def __init__(self, busy, window=0.0005, max_batch=64, solo=None, multi=None, sync=None):
        """busy(): number of slots working on a request.  solo / multi / sync
        replace krylov.eigsolve, krylov.eigsolve_multi and the stream
        synchronisation of a handed-in vector (for tests of the schedule)."""
        self._busy = busy
        self.window, self.max_batch = float(window), int(max_batch)
        self._solo = solo if solo is not None else eigsolve
        self._multi = multi if multi is not None else eigsolve_multi
        self._sync = sync if sync is not None else (lambda x: synchronize(x.device))
        self._cv = threading.Condition()
        self._pending = []
        self._running = 0  # jobs taken by a leader and not finished
        self.sizes = []
        self.solve_s = 0.0  # time inside the eigensolves, summed over leaders

    def poke(self):
        """re-check the start condition (a slot became idle)"""
        with self._cv:
            self._cv.notify_all()

    def solve(self, plans, psi, krylovdim=3, maxiter=1, tol=1e-14):
        """the eigsolver hook a slot's DMRG calls (same contract as krylov.eigsolve)"""
        self._sync(psi)  # the slot's work is done before a leader borrows its stream
        job = dict(plans=list(plans), x=psi, key=(int(krylovdim), int(maxiter), float(tol)), done=threading.Event())
        jobs = None
        with self._cv:
            self._pending.append(job)
            self._cv.notify_all()
            while any(j is job for j in self._pending):  # until a leader (maybe this slot) takes it
                if self._pending[0] is job:
                    # only the first queued job's slot takes jobs, so it stays first while it waits
                    # for the busy slots that are not inside another batch
                    t0 = time.perf_counter()
                    while len(self._pending) < min(self._busy() - self._running, self.max_batch):
                        left = self.window - (time.perf_counter() - t0)
                        if left <= 0:
                            break
                        self._cv.wait(left)
                    jobs = [j for j in self._pending if j["key"] == job["key"]][:self.max_batch]
                    self._pending = [j for j in self._pending if not any(j is t for t in jobs)]
                    self._running += len(jobs)
                    self._cv.notify_all()  # the next queued job (if any) leads the next batch
                    break
                self._cv.wait()
        if jobs is not None:
            self._run(jobs)
        job["done"].wait()
        if "error" in job:
            raise job["error"]
        return [job["val"]], [psi], job["info"]

    def _run(self, jobs):
        k, mi, tol = jobs[0]["key"]
        t = time.perf_counter()
        try:
            if len(jobs) == 1:
                vals, _, info = self._solo(jobs[0]["plans"], jobs[0]["x"], krylovdim=k, maxiter=mi, tol=tol)
                jobs[0]["val"], jobs[0]["info"] = vals[0], info
            else:
                # on the leader's own context (the default of eigsolve_multi)
                vals, infos = self._multi([j["plans"] for j in jobs], [j["x"] for j in jobs], krylovdim=k,
                                          maxiter=mi, tol=tol)
                for j, v, inf in zip(jobs, vals, infos):
                    j["val"], j["info"] = v, inf
        except Exception as e:  # handed to every slot of the batch
            for j in jobs:
                j["error"] = e
        with self._cv:
            self.solve_s += time.perf_counter() - t
            self.sizes.append(len(jobs))
            self._running -= len(jobs)
            self._cv.notify_all()
        for j in jobs:
            j["done"].set()


class DMRGServer:
    """Continuous-batching server: `slots` solves in flight at most, admitted
    first come first served while the sum of their footprint estimates stays
    within hbm_budget bytes; co_batch=True routes their eigensolves through a
    KrylovBatcher (False: each slot solves its own)."""

    def __init__(self, slots=8, hbm_budget=200 << 30, co_batch=True, window=0.0005, device=0, solver=None,
                 batcher_args=None):
        """solver(request, eigsolver, device) -> (energy, energies, converged,
        maxlinkdim, peak bytes, truncation error) replaces solve_dmrg; batcher_args are extra
        KrylovBatcher arguments (both for tests of the schedule)."""
        if slots < 1:
            raise ValueError("a server needs at least one slot")
        self.slots, self.budget, self.co_batch = int(slots), int(hbm_budget), bool(co_batch)
        self.window, self.device = window, device
        self.solver = solver if solver is not None else solve_dmrg
        # what a slot thread does with its context when the run ends (the library's
        # solves: return the device memory its context caches)
        self.release = trim if solver is None else None
        self.batcher_args = dict(batcher_args or {})
        self.batch_sizes, self.batch_solve_s = [], 0.0
        self.max_reserved = 0
        self.max_inflight = 0

    def run(self, requests):
        """Serves the requests (arrival times relative to the call) and returns
        their results in the order the requests were given."""
        given = list(requests)
        reqs = sorted(given, key=lambda r: (r.arrival, r.rid))
        if len({r.rid for r in reqs}) != len(reqs):
            raise ValueError("request ids must be distinct")
        for r in reqs:
            if footprint_bytes(r) > self.budget:
                raise ValueError(f"request {r.rid} needs {footprint_bytes(r)} bytes, above the budget {self.budget}")
        cv = threading.Condition()
        state = dict(busy=0, reserved=0, done=0, error=None)
        inbox = [deque() for _ in range(self.slots)]
        used = [False] * self.slots
        idle = list(range(self.slots))
        results = {}
        batcher = KrylovBatcher(lambda: state["busy"], self.window, **self.batcher_args) if self.co_batch else None
        eig = batcher.solve if batcher else self.batcher_args.get("solo", eigsolve)
        t0 = time.perf_counter()

        def slot(i):  # a slot's tensors live on its thread's own context and stream
            while True:
                with cv:
                    while not inbox[i]:
                        cv.wait()
                    item = inbox[i].popleft()
                if item is None:
                    if used[i] and self.release:  # its solves' tensors are gone: give the cached memory back
                        gc.collect()
                        self.release(self.device)
                    return
                used[i] = True
                r, t_adm, fp = item
                try:
                    e, es, conv, mld, peak, terr = self.solver(r, eig, self.device)
                    res = SolveResult(r.rid, e, len(es), conv, es, r.arrival, t_adm, time.perf_counter() - t0, terr,
                                      fp, int(peak), mld)
                except Exception as ex:  # reported by run()
                    res = None
                    with cv:
                        state["error"] = ex
                with cv:
                    if res is not None:
                        results[r.rid] = res
                    state["busy"] -= 1
                    state["reserved"] -= fp
                    state["done"] += 1
                    idle.append(i)
                    cv.notify_all()
                if batcher:
                    batcher.poke()

        threads = [threading.Thread(target=slot, args=(i,), name=f"slot-{i}", daemon=True) for i in range(self.slots)]
        for t in threads:
            t.start()
        waiting = deque()
        nxt = 0
        try:
            with cv:
                while state["done"] < len(reqs) and state["error"] is None:
                    now = time.perf_counter() - t0
                    while nxt < len(reqs) and reqs[nxt].arrival <= now:
                        waiting.append(reqs[nxt])
                        nxt += 1
                    # admission: first come first served, by free slot and memory
                    while waiting and idle and state["reserved"] + footprint_bytes(waiting[0]) <= self.budget:
                        r = waiting.popleft()
                        fp = footprint_bytes(r)
                        i = idle.pop(0)
                        state["busy"] += 1
                        state["reserved"] += fp
                        self.max_reserved = max(self.max_reserved, state["reserved"])
                        self.max_inflight = max(self.max_inflight, state["busy"])
                        inbox[i].append((r, now, fp))
                        cv.notify_all()
                    wait = None if nxt >= len(reqs) else max(0.0, reqs[nxt].arrival - now)
                    cv.wait(wait)
        finally:
            with cv:
                for q in inbox:
                    q.append(None)
                cv.notify_all()
            for t in threads:
                t.join()
            if batcher:
                self.batch_sizes = list(batcher.sizes)
                self.batch_solve_s = batcher.solve_s
        if state["error"] is not None:
            raise state["error"]
        return [results[r.rid] for r in given]


def run_plain(requests, device=0, solver=None, eigsolver=None):
    """The plain GPU baseline: one solve at a time on the calling thread, first
    come first served, each eigensolve alone.  Same results and timing fields.
    solver / eigsolver replace solve_dmrg / krylov.eigsolve (tests).  Results in
    the order the requests were given."""
    solver = solver if solver is not None else solve_dmrg
    given = list(requests)
    if len({r.rid for r in given}) != len(given):
        raise ValueError("request ids must be distinct")
    reqs = sorted(given, key=lambda r: (r.arrival, r.rid))
    out = {}
    t0 = time.perf_counter()
    for r in reqs:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        t_adm = time.perf_counter() - t0
        e, es, conv, mld, peak, terr = solver(r, eigsolver if eigsolver is not None else eigsolve, device)
        out[r.rid] = SolveResult(r.rid, e, len(es), conv, es, r.arrival, t_adm, time.perf_counter() - t0, terr,
                                 footprint_bytes(r), int(peak), mld)
    return [out[r.rid] for r in given]


def poisson_requests(n, rate, seed=0, sites=(32, 64, 96, 128), chis=(128, 256, 512, 1024), delta=(0.5, 1.5),
                     tol=1e-10, max_sweeps=12):
    """n requests with exponential inter-arrival times of mean 1/rate (rate in
    requests per second; rate = inf: all at time 0), lengths and bond caps
    drawn uniformly from sites / chis, anisotropy uniform in delta.  One seed
    gives the same requests at every rate; only the arrival times scale."""
    rng = np.random.default_rng(seed)
    gaps = rng.standard_exponential(n)  # drawn at every rate, so the mix of a seed does not depend on the rate
    gaps = gaps / rate if np.isfinite(rate) else np.zeros(n)
    t = np.cumsum(gaps) - (gaps[0] if n else 0.0)
    return [SolveRequest(i, int(rng.choice(sites)), int(rng.choice(chis)), float(rng.uniform(*delta)), float(t[i]),
                         tol, max_sweeps) for i in range(n)]


def summarize(results):
    """solves/s over the run (first arrival to last finish), p50/p99 latency"""
    lat = np.array([r.latency for r in results])
    span = max(r.finish for r in results) - min(r.arrival for r in results)
    return dict(solves=len(results), span_s=float(span), solves_per_s=len(results) / span if span > 0 else 0.0,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), converged=int(sum(r.converged for r in results)))
