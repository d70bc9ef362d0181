"""Benchmark: two-site H_eff*psi on one MI355X.

Workload (default): the Hubbard chain of the precision configuration
(t=1, U=4, local dim 4, U(1)xU(1) charges (n, 2Sz), MPO bond dim 6,
chi=2048 in 520 sectors sized by a 2D Gaussian of width 7), one two-site
H_eff*psi = four contract! calls on NewBlockSparseSorted tensors
(psi*L*W1*W2*R).  Synthetic data: random psi / environments on the
charge-allowed entries, seed 0; the MPO is the Hubbard MPO.

A step is one H_eff*psi: every step recomputes all four contractions
(box-aware grouped block products, no operand permutes) from the inputs
resident in HBM.  The block structure is planned once (sb_plan_create), as a
Krylov solve keeps it fixed; the planning time (which includes the tile
autotuner's timing runs, SB_FLAG_AUTOTUNE) is reported separately, and the
step with the cost model's tile shapes is timed beside it.

The environments are symmetrised so that H_eff is symmetric; "eigsolve" then
reports the ground energy by FP64 Lanczos, the mixed-precision solve's
energy difference and time-to-energy against it, and solves/s with p50/p99
latency of the DMRG-step solve (krylovdim 3, maxiter 1); "dmrg_solve" times
whole two-site DMRG ground-state solves of XXZ chains (solves/s, p50/p99);
"batch" runs 64 independent J1-J2 H_eff*psi problems plan by plan and as one
batch (one box product launch per contraction stage over all problems);
"serve" runs a stream of DMRG requests through the serving loop against the
plain one-at-a-time queue (solves/s, p50/p99 latency, FP64 and FP32 products);
"coo" times the element-sparse workload (PXP chain, COO x Dense at chi=128);
"sweep" times one two-site DMRG sweep of the N=32 Heisenberg chain at chi=256;
"permute" times permutedims of the chi=2048 psi (GB/s against the HBM peak).

value = algorithmic FLOPs per H_eff*psi (sum of 2*M*K*N over the matched
block pairs with the actual sector extents) / time per step, in TFLOP/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hubbard-chi2048|heisenberg-chi256]
                       [--prec fp64|fp32|bf16x3] [--tile T] [--no-autotune] [--no-cpu] [--no-coo] [--no-sweep] [--no-mixed] [--no-eig]
                       [--krylovdim K] [--eig-tol T] [--eig-tol-lo T] [--solves N] [--dmrg-solves N]
                       [--dmrg-sites N] [--dmrg-chi X] [--batch-problems N] [--batch-reps R]
                       [--serve-requests N] [--serve-slots S] [--serve-sites L,..] [--serve-chis X,..]
                       [--calibrate-copy MB]
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 6.0}  # dense MFMA TFLOP/s (bf16x3: 6 bf16 products)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hubbard-chi2048", choices=["hubbard-chi2048", "heisenberg-chi256"])
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--tile", type=int, default=0,
                    help="force tile shape 1+c of the block products (0: chosen at plan time)")
    ap.add_argument("--path", default="fused", choices=["fused", "chain"],
                    help="the timed H_eff*psi: the fused plan (one kernel) or the four contract! plans")
    ap.add_argument("--no-dropin", action="store_true", help="skip the contract!-per-call (drop-in) eigensolve")
    ap.add_argument("--no-autotune", action="store_true",
                    help="tile shapes from the padded-work cost model instead of the measured autotuner")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-coo", action="store_true", help="skip the element-sparse (PXP, COO) workload")
    ap.add_argument("--no-sweep", action="store_true", help="skip the N=32 chi=256 Heisenberg DMRG sweep")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-eig", action="store_true", help="skip the eigensolver measurements")
    ap.add_argument("--krylovdim", type=int, default=30)
    ap.add_argument("--eig-tol", type=float, default=1e-8, help="residual norm of the FP64 solves")
    ap.add_argument("--eig-tol-lo", type=float, default=1e-6,
                    help="residual norm of the low-precision stage, relative to |E|")
    ap.add_argument("--solves", type=int, default=30, help="DMRG-step solves timed for solves/s and p99")
    ap.add_argument("--dmrg-solves", type=int, default=8, help="full DMRG solves (0: skip)")
    ap.add_argument("--dmrg-sites", type=int, default=16)
    ap.add_argument("--dmrg-chi", type=int, default=64)
    ap.add_argument("--batch-problems", type=int, default=64,
                    help="independent J1-J2 problems of the batching measurement (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=10)
    ap.add_argument("--batch-kdim", type=int, default=16)
    ap.add_argument("--batch-eig-tol", type=float, default=1e-8, help="0: skip the batched ground-state solves")
    ap.add_argument("--serve-requests", type=int, default=24, help="DMRG requests of the serving run (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-sites", default="16,24,32", help="chain lengths the requests draw from")
    ap.add_argument("--serve-chis", default="32,64,128", help="bond dimension caps the requests draw from")
    ap.add_argument("--serve-config5", type=int, default=0,
                    help="run only the serving measurement at the serving configuration's sizes with this many "
                         "requests (N in 32..128, chi_max in 128..1024; minutes per request) and print its JSON")
    ap.add_argument("--serve5-seed", type=int, default=0)
    ap.add_argument("--serve5-max-sweeps", type=int, default=20)
    ap.add_argument("--serve5-parts", default="plain,served,mixed")
    ap.add_argument("--serve5-rates", default="1,10,50", help="Poisson arrival rates (requests/s) of the served runs")
    ap.add_argument("--serve5-mixed-rates", default="10", help="rates of the served mixed-precision runs")
    ap.add_argument("--serve5-plain-json", default="",
                    help="reuse the plain run of an earlier --serve-config5 output (same requests) instead of "
                         "running it")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="internal: three fused applications and a 256 MB copy, for the profiled child runs")
    ap.add_argument("--no-traffic", action="store_true", help="skip the in-run rocprofv3 traffic measurement")
    ap.add_argument("--phase-sites", type=int, default=64, help="--only dmrg-phases: chain length")
    ap.add_argument("--phase-chi", type=int, default=512, help="--only dmrg-phases: bond dimension cap")
    ap.add_argument("--phase-sweeps", type=int, default=6, help="--only dmrg-phases: sweeps (the last one timed)")
    ap.add_argument("--only", default="", choices=["", "batch", "sweep", "dmrg-phases"],
                    help="run one section alone (batch: the batching configuration, for profiling)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="also run one plain device copy of this many MB (rocprof traffic calibration)")
    return ap.parse_args()


def make_problem(cfg):
    from sparsebackends_amd import heff as H
    return H.hubbard_chi2048_problem(0, hermitian=True) if cfg == "hubbard-chi2048" else \
        H.heisenberg_chi256_problem(0, hermitian=True)


def rel_err(x, y):
    n = np.linalg.norm(y)
    return float(np.linalg.norm(x - y) / (n if n else 1.0))


def dropin_eigsolve(sb, prob, x0d, krylovdim, tol, maxops=400):
    """The FP64 drop-in's time-to-energy: Lanczos (Gram-Schmidt twice,
    restart from the Ritz vector) whose every operator application is the four
    contract! calls ITensors issues through the C ABI (block structure derived
    on every call), the image kept in the vector's block space, and whose
    vector operations are the Krylov hooks (norm, dot, scale, key-union add:
    new tensors, as ITensors' out-of-place arithmetic makes them) -- what
    KrylovKit's eigsolve on H_eff runs with the backend dropped in.
    Returns (energy, seconds, applications)."""
    from sparsebackends_amd.storage import NewBlockSparseSorted as BS
    p = prob

    def apply(x):
        T1 = BS.new(p.dims_T1, 5)
        sb.contract_(T1, p.l_T1, x, p.l_psi, p.L, p.l_L)
        T2 = BS.new(p.dims_T2, 5)
        sb.contract_(T2, p.l_T2, T1, p.l_T1, p.W1, p.l_W1)
        T3 = BS.new(p.dims_T3, 5)
        sb.contract_(T3, p.l_T3, T2, p.l_T2, p.W2, p.l_W2)
        H = BS.new(p.dims_out, 4)
        sb.contract_(H, p.l_out, T3, p.l_T3, p.R, p.l_R)
        return H

    sb.synchronize()
    t0 = time.perf_counter()
    x = (1.0 / x0d.norm()) * x0d
    nops, theta, done = 0, None, False
    while nops < maxops and not done:
        V, al, be = [x], [], []
        for j in range(krylovdim):
            w = x0d.restrict(apply(V[j]))  # H v in the vector's block space, as the library's eigensolve keeps it
            nops += 1
            al.append(w.dot(V[j]))
            for _ in range(2):
                for v in V:
                    w = w + (-w.dot(v)) * v
            b = w.norm()
            ev, U = np.linalg.eigh(np.diag(al) + np.diag(be, 1) + np.diag(be, -1))
            theta = float(ev[0])
            done = abs(b * U[-1, 0]) <= tol or b <= 1e-13
            if done or j == krylovdim - 1:
                break
            be.append(b)
            V.append((1.0 / b) * w)
        y = float(U[0, 0]) * V[0]
        for c, v in zip(U[1:, 0], V[1:len(U)]):
            y = y + float(c) * v
        x = (1.0 / y.norm()) * y
    sb.synchronize()
    return theta, time.perf_counter() - t0, nops


def eig_section(sb, a, prob, plans, fused64, fused32):
    """Ground state of the (symmetric) H_eff.  FP64 Lanczos with the fused
    H_eff plan as the operator, and with the four contract! plans; the
    mixed-precision solve (FP32 products, then an FP64 second stage from its
    Ritz vector) on the fused plans; the drop-in's time-to-energy (contract!
    per call); and the DMRG-step solve (krylovdim 3, maxiter 1, the reference
    DMRG's call) as solves/s with its latency."""
    x0 = prob.psi_parts["data"]
    res = {}

    def timed(fn):
        sb.synchronize()
        t0 = time.perf_counter()
        r = fn()
        sb.synchronize()
        return r, time.perf_counter() - t0

    prob.psi.set_data_(x0)
    x0d = prob.psi.copy()  # the start vector, kept on the device
    op64 = [fused64] if fused64 is not None else plans
    (vals, _, info), t64 = timed(lambda: sb.eigsolve(op64, prob.psi, krylovdim=a.krylovdim, maxiter=50, tol=a.eig_tol))
    e64 = vals[0]
    res["tol"] = a.eig_tol
    res["krylovdim"] = a.krylovdim
    res["operator"] = "fused H_eff plan" if fused64 is not None else "four contract! plans"
    res["fp64"] = dict(energy=e64, seconds=t64, numops=info.numops, converged=info.converged, normres=info.normres)
    if fused64 is not None:
        prob.psi.axpby_(1.0, x0d, 0.0)
        (vc, _, ic), tc = timed(lambda: sb.eigsolve(plans, prob.psi, krylovdim=a.krylovdim, maxiter=50, tol=a.eig_tol))
        res["fp64_chain_plans"] = dict(energy=vc[0], seconds=tc, numops=ic.numops, abs_dE_vs_fused=abs(vc[0] - e64))
    lo_forms = [("fp32", sb.PREC_FP32, fused32)] if fused32 is not None else []
    lo_forms += [("fp32_chain", sb.PREC_FP32, None), ("bf16x3_chain", sb.PREC_BF16X3, None)]
    for name, fl, fplan in lo_forms:
        if fplan is not None:
            pm, plans_lo = fplan
        else:
            pm = copy.copy(prob).upload().outputs()
            plans_lo = pm.make_plans(fl)
        pm.psi.axpby_(1.0, x0d, 0.0)
        prob.psi.axpby_(1.0, x0d, 0.0)
        tol_lo = a.eig_tol_lo * max(1.0, abs(e64))
        (r, t) = timed(lambda: sb.eigsolve_refined(plans_lo, op64, pm.psi, prob.psi, krylovdim=a.krylovdim,
                                                   maxiter=50, tol_lo=tol_lo, tol=a.eig_tol))
        vals, _, ilo, ihi = r
        res[name] = dict(energy=vals[0], abs_dE_vs_fp64=abs(vals[0] - e64), seconds=t, time_to_energy_vs_fp64=t64 / t,
                         numops_low=ilo.numops, numops_fp64=ihi.numops, converged=ihi.converged, tol_low=tol_lo)
        if fplan is not None:
            # the same two stages eigsolve_refined runs, timed apart: the FP64 stage alone bounds what
            # any speed-up of the low-precision operator can give
            pm.psi.axpby_(1.0, x0d, 0.0)
            _, t_lo = timed(lambda: sb.eigsolve(plans_lo, pm.psi, krylovdim=a.krylovdim, maxiter=50, tol=tol_lo))
            prob.psi.axpby_(1.0, pm.psi, 0.0)
            (v2, _, i2), t_hi = timed(lambda: sb.eigsolve(op64, prob.psi, krylovdim=8, maxiter=10, tol=a.eig_tol))
            res[name].update(low_stage_s=t_lo, fp64_stage_s=t_hi, stages_energy_equal=bool(v2[0] == vals[0]),
                             fp64_iteration_ms=1e3 * t64 / info.numops, fp64_stage_iteration_ms=1e3 * t_hi / i2.numops,
                             low_iteration_ms=1e3 * t_lo / ilo.numops,
                             cap_if_low_stage_free=t64 / t_hi,
                             cap_note="time-to-energy ratio if the low-precision stage took no time: FP64 solve / "
                                      "its FP64 stage alone")
        del pm, plans_lo
    if not a.no_dropin:
        ed, td, nd = dropin_eigsolve(sb, prob, x0d, a.krylovdim, a.eig_tol)
        res["fp64_dropin"] = dict(energy=ed, seconds=td, numops=nd, abs_dE_vs_fp64=abs(ed - e64),
                                  note="contract! per call (block structure derived on every call) and the Krylov "
                                       "hooks, as KrylovKit drives the dropped-in backend")
        for name in ("fp32",):
            if name in res:
                res[name]["time_to_energy_vs_dropin"] = td / res[name]["seconds"]
        res["fp64"]["time_to_energy_vs_dropin"] = td / t64
    lat = []
    for _ in range(a.solves):
        prob.psi.axpby_(1.0, x0d, 0.0)
        _, t = timed(lambda: sb.eigsolve(op64, prob.psi, krylovdim=3, maxiter=1, tol=1e-14))
        lat.append(t)
    lat = np.array(lat)
    res["dmrg_step_solve"] = dict(krylovdim=3, maxiter=1, solves=a.solves, solves_per_s=len(lat) / lat.sum(),
                                  p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3))
    return res


def dmrg_phases_section(sb, a):
    """Where a two-site DMRG step at the serving sizes spends its time: an XXZ
    chain (Delta = 0.9) of --phase-sites sites from the Neel state, bond
    dimension cap --phase-chi, cutoff 1e-12, Krylov dimension 3; every sweep's
    steps timed by phase (two-site contraction, the four H_eff plans'
    creation, eigensolve, sector SVD, environment update), each phase ending
    with a device synchronisation."""
    from sparsebackends_amd import heff as H
    from sparsebackends_amd import dmrg as D
    from sparsebackends_amd import serve as S
    N, chi = a.phase_sites, a.phase_chi
    dm = D.DMRG(H.heisenberg_model(1.0, 0.9), S.neel(N))
    out = dict(model="XXZ J=1 Delta=0.9, Neel start, U(1) Sz, MPO w=5", sites=N, chi_max=chi, cutoff=1e-12,
               krylovdim=3, sweeps=[])
    for k in range(a.phase_sweeps):
        dm.phase_times = {}
        sb.synchronize()
        t0 = time.perf_counter()
        e = dm.sweep(maxdim=chi, cutoff=1e-12, krylovdim=3, maxiter=1, tol=1e-14)
        sb.synchronize()
        dt = time.perf_counter() - t0
        steps = 2 * (N - 1)
        row = dict(sweep=k, seconds=dt, ms_per_step=1e3 * dt / steps, maxlinkdim=dm.maxlinkdim(), energy=e,
                   largest_sector=max(max(b.dims) for b in dm.bonds),
                   phase_ms_per_step={p: 1e3 * v / steps for p, v in dm.phase_times.items()})
        out["sweeps"].append(row)
        print(f"[dmrg-phases] sweep {k}: {row['ms_per_step']:.2f} ms/step, maxlink {row['maxlinkdim']}, "
              + ", ".join(f"{p} {v:.2f}" for p, v in sorted(row["phase_ms_per_step"].items(), key=lambda x: -x[1])),
              flush=True)
    return out


def dmrg_section(sb, a):
    """Full two-site DMRG ground-state solves of XXZ chains (the serving
    configuration's model at its smallest length): N sites from the Neel state,
    bond dimension chi, 4 sweeps of (krylovdim 3, maxiter 1) steps, one solve
    per anisotropy Delta in [0.5, 1.5], run back to back.  Reported: solves/s
    and p50/p99 latency, and each energy against the same chain solved at bond
    dimension 256 (exact for 16 sites) with converged eigensolves."""
    from sparsebackends_amd import heff as H
    from sparsebackends_amd import dmrg as D
    N, chi, sweeps = a.dmrg_sites, a.dmrg_chi, 4
    deltas = np.linspace(0.5, 1.5, a.dmrg_solves)
    lat, es = [], []
    for dlt in deltas:
        sb.synchronize()
        t0 = time.perf_counter()
        dm = D.DMRG(H.heisenberg_model(1.0, float(dlt)), [0, 1] * (N // 2))
        for _ in range(sweeps):
            e = dm.sweep(maxdim=chi, cutoff=1e-10, krylovdim=3, maxiter=1, tol=1e-14)
        sb.synchronize()
        lat.append(time.perf_counter() - t0)
        es.append(e)
    errs = []
    for dlt, e in zip(deltas, es):
        ref = D.DMRG(H.heisenberg_model(1.0, float(dlt)), [0, 1] * (N // 2))
        for _ in range(5):
            er = ref.sweep(maxdim=256, cutoff=0.0, krylovdim=8, maxiter=3, tol=1e-13)
        errs.append(e - er)
    lat = np.array(lat)
    return dict(model="XXZ J=1", sites=N, chi=chi, sweeps=sweeps, krylovdim=3, maxiter=1, solves=len(lat),
                solves_per_s=len(lat) / lat.sum(), p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), max_energy_error_vs_chi256=float(np.max(np.abs(errs))),
                note="sequential solves, each recomputing its plans at every step")


def sweep_section(sb, a):
    """One two-site DMRG sweep of the plain configuration: S=1/2 Heisenberg
    J=1, N=32, U(1) Sz, MPO w=5, bond dimension 256 (sectors q = 2Sz of the
    bond's parity, dims ~ exp(-q^2/32)), the MPS i.i.d. N(0,1) from seed 0
    left-orthonormalised by QR (sector by sector, on the device), Krylov
    dimension 3 (maxiter 1), cutoff 1e-10.  The sweep (left to right and back,
    62 two-site steps) is timed.  CPU leg: the oracle (reference algorithm
    restated in C++, BLAS on one thread) on the three centre steps' work that
    dominates a step -- three H_eff*psi applications (krylovdim 3) and the
    sector SVD -- timed on the host, per step."""
    from sparsebackends_amd import heff as H
    from sparsebackends_amd import dmrg as D
    N, chi = 32, 256
    m = H.heisenberg_model(1.0)
    t0 = time.perf_counter()
    bonds = D.config_bonds(N, chi=chi)
    A, bonds = D.left_orthonormalize(m, D.random_mps(m, bonds, seed=0), bonds)
    dm = D.DMRG.from_mps(m, A, bonds)
    sb.synchronize()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    e = dm.sweep(maxdim=chi, cutoff=1e-10, krylovdim=3, maxiter=1, tol=1e-14)
    sb.synchronize()
    ts = time.perf_counter() - t0
    res = dict(model="Heisenberg S=1/2 J=1, N=32, U(1) Sz, MPO w=5", chi=chi, maxlinkdim=int(dm.maxlinkdim()),
               start="random N(0,1) MPS, seed 0, left-orthonormalised by sector QR on the device",
               init_s=t_init, krylovdim=3, maxiter=1, cutoff=1e-10, steps=2 * (N - 1), sweep_s=ts,
               ms_per_step=ts / (2 * (N - 1)) * 1e3, energy=e, truncerr=dm.truncerr)
    if not a.no_cpu:
        from oracle import oracle as O
        from oracle.factorize_oracle import svd_sectors as osvd
        O.use_blas(True, threads=1)

        def obs(t):
            p = t.parts()
            return O.BS.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"]), p

        # after the sweep the left environments are a half sweep old: left-to-right
        # steps 1..14 rebuild them (untimed), so that the sampled steps see the
        # consistent state the sweep's own forward steps see
        kw = dict(maxdim=chi, cutoff=1e-10, krylovdim=3, maxiter=1, tol=1e-14)
        for b in range(1, 15):
            dm.step(b, +1, **kw)
        tcpu, tgpu = [], []
        for b in (15, 16, 17):
            psi = D.two_site(dm.A[b - 1], dm.A[b])
            plans, Hp = H.heff_plans(psi, dm.L[b - 1], dm.W, dm.W, dm.R[b + 2])
            sb.synchronize()
            t1 = time.perf_counter()
            for _ in range(3):
                for pl in plans:
                    pl.execute()
            sb.synchronize()
            tgpu.append(time.perf_counter() - t1)
            (opsi, pp), (oL, _), (oW, _), (oR, _) = obs(psi), obs(dm.L[b - 1]), obs(dm.W), obs(dm.R[b + 2])
            nl, nr, d1, d2, Dl, Dr = psi.dims
            w = dm.W.dims[0]
            t1 = time.perf_counter()
            for _ in range(3):
                T1 = O.BS.new((nl, nr, d1, d2, Dr, w, Dl), 5)
                O.contract(T1, H.L_T1, opsi, H.L_PSI, oL, H.L_L)
                T2 = O.BS.new((nl, nr, d1, w, d2, Dr, Dl), 5)
                O.contract(T2, H.L_T2, T1, H.L_T1, oW, H.L_W1)
                T3 = O.BS.new((nl, nr, d2, w, d1, Dr, Dl), 5)
                O.contract(T3, H.L_T3, T2, H.L_T2, oW, H.L_W2)
                Ho = O.BS.new((nl, nr, d1, d2, Dl, Dr), 4)
                O.contract(Ho, H.L_OUT, T3, H.L_T3, oR, H.L_R)
            Bl, Br = dm.bonds[b - 1], dm.bonds[b + 1]
            rowsec, colsec, charges = D.fused_sectors(m, Bl, Br)
            osvd(pp, rowsec, colsec, Bl.dims, Br.dims, len(charges), maxdim=chi, cutoff=1e-10)
            tcpu.append(time.perf_counter() - t1)
            dm.step(b, +1, **kw)
        res["cpu_baseline_sample"] = dict(
            kind="port", cores=1, steps=[15, 16, 17], cpu_s_per_step=float(np.mean(tcpu)),
            gpu_heff_s_per_step=float(np.mean(tgpu)),
            sample="three centre two-site steps: 3 H_eff*psi applications (four contract! calls each, oracle "
                   "C++ restatement, BLAS on 1 thread) + the sector SVD (numpy LAPACK)",
            sweep_estimate_s=float(np.mean(tcpu)) * 2 * (N - 1),
            note="the estimate scales the centre steps (the largest bonds) to the sweep's 62 steps")
    return res


def batch_section(sb, a):
    """Many independent problems at once (the batching configuration): the
    two-site H_eff*psi of a J1-J2 chain's centre bond for each of
    a.batch_problems seeds (J2 uniform in [0, 0.5], chi log-uniform in
    [64, 1024] in 17 Sz sectors).  One matvec of the batch = the four
    contractions of every problem; run plan by plan (sequential), as a
    PlanBatch (one box product launch per stage over all problems) with the
    cost model's tiles, with tiles measured per stage over all problems
    (autotune), and as one HIP graph launch per matvec; FP64 and FP32 forms;
    all give identical bits."""
    from sparsebackends_amd import heff as H
    t0 = time.perf_counter()
    probs = H.j1j2_batch_problems(a.batch_problems)
    t_gen = time.perf_counter() - t0
    alg = sum(p.algorithmic_flops()["total"] for p, _, _ in probs)
    chis = [c for _, _, c in probs]
    res = dict(problems=len(probs), model="S=1/2 J1-J2 chains, J1=1, J2~U[0,0.5], U(1) Sz, MPO w=8, centre bond",
               chi_min=min(chis), chi_max=max(chis), algorithmic_gflop_per_matvec=alg / 1e9, generate_s=t_gen,
               reps=a.batch_reps)

    def timed(fn, reps):
        fn()
        sb.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        sb.synchronize()
        return (time.perf_counter() - t0) / reps

    for p, _, _ in probs:
        p.upload()
    for name, fl in (("fp64", sb.PREC_FP64), ("fp32", sb.PREC_FP32)):
        sb.synchronize()
        t0 = time.perf_counter()
        for p, _, _ in probs:
            p.outputs().make_plans(fl)
        batch = sb.PlanBatch([p.plans for p, _, _ in probs])
        sb.synchronize()
        t_plan = time.perf_counter() - t0

        def seq():
            for p, _, _ in probs:
                p.execute()
        t_seq = timed(seq, a.batch_reps)
        ref = [p.Hpsi.parts()["data"].copy() for p, _, _ in probs]
        t_bat = timed(batch.execute, a.batch_reps)
        same = all(np.array_equal(p.Hpsi.parts()["data"], r) for (p, _, _), r in zip(probs, ref))
        names = ["16x128", "128x16", "32x32", "64x32", "32x64", "64x64", "16x64", "64x16"]
        tiles_vote = [names[t] if t >= 0 else None for t, _ in batch.stage_tiles()]
        # the batch's tile shapes measured (every workgroup shape timed per stage over all problems)
        t0 = time.perf_counter()
        batch.autotune()
        t_tune = time.perf_counter() - t0
        sb.profile(True, products_only=True)
        sb.profile_reset()
        t_tuned = timed(batch.execute, a.batch_reps)
        prof = sb.profile_read()
        sb.profile(False)
        gemm_ms = prof.get("grouped_gemm", {}).get("total_ms", 0.0) / (a.batch_reps + 1)
        same_tuned = all(np.array_equal(p.Hpsi.parts()["data"], r) for (p, _, _), r in zip(probs, ref))
        # one HIP graph launch per matvec of the whole batch
        graph = sb.PlanGraph(batch)
        for _ in range(3):
            graph.launch()
        t_graph = timed(graph.launch, a.batch_reps)
        same_graph = all(np.array_equal(p.Hpsi.parts()["data"], r) for (p, _, _), r in zip(probs, ref))
        r = dict(sequential_ms=t_seq * 1e3, batched_ms=t_bat * 1e3, sequential_tflops=alg / t_seq / 1e12,
                 batched_tflops=alg / t_bat / 1e12, speedup=t_seq / t_bat, bit_identical=bool(same),
                 fused_stages=batch.stats()["fused_stages"], plan_s=t_plan,
                 kernels_ms_per_matvec={k: v["total_ms"] / (a.batch_reps + 1) for k, v in prof.items()},
                 tiles_cost_model=tiles_vote,
                 tiles_autotuned=[names[t] if t >= 0 else None for t, _ in batch.stage_tiles()],
                 autotune_s=t_tune, autotuned_ms=t_tuned * 1e3, autotuned_tflops=alg / t_tuned / 1e12,
                 autotuned_bit_identical=bool(same_tuned),
                 graph_ms=t_graph * 1e3, graph_tflops=alg / t_graph / 1e12, graph_bit_identical=bool(same_graph),
                 graph_stats=graph.stats())
        del graph
        if gemm_ms > 0:
            peak = PEAK[name]
            r["gemm_roofline"] = dict(bound="mfma", achieved=alg / (gemm_ms * 1e-3) / 1e12, peak=peak,
                                      unit="TFLOP/s", frac=alg / (gemm_ms * 1e-3) / 1e12 / peak,
                                      note="the four batched grouped_gemm launches of one matvec, autotuned tiles")
        if name == "fp64" and a.batch_eig_tol > 0:
            r["eigsolve"] = batch_eig(sb, a, probs, batch)
        res[name] = r
        del batch
        for p, _, _ in probs:
            p.plans = None
    return res


def batch_eig(sb, a, probs, batch):
    """Ground states of all problems of the batch: FP64 Lanczos (krylovdim
    a.batch_kdim, tol a.batch_eig_tol) one problem after the other, then all
    of them co-batched (one batched operator application per iteration for
    the solves still running); the energies agree bit for bit."""
    x0 = [p.psi.copy() for p, _, _ in probs]
    kw = dict(krylovdim=a.batch_kdim, maxiter=50, tol=a.batch_eig_tol)
    sb.synchronize()
    t0 = time.perf_counter()
    e_seq, ops = [], 0
    for (p, _, _), x in zip(probs, x0):
        p.psi.axpby_(1.0, x, 0.0)
        vals, _, info = sb.eigsolve(p.plans, p.psi, **kw)
        e_seq.append(vals[0])
        ops += info.numops
    sb.synchronize()
    t_seq = time.perf_counter() - t0
    for (p, _, _), x in zip(probs, x0):
        p.psi.axpby_(1.0, x, 0.0)
    sb.synchronize()
    t0 = time.perf_counter()
    e_bat, infos = sb.eigsolve_batch(batch, [p.psi for p, _, _ in probs], **kw)
    sb.synchronize()
    t_bat = time.perf_counter() - t0
    return dict(krylovdim=a.batch_kdim, tol=a.batch_eig_tol, solves=len(probs), operator_applications=ops,
                converged=int(sum(i.converged for i in infos)), sequential_s=t_seq, batched_s=t_bat,
                sequential_solves_per_s=len(probs) / t_seq, batched_solves_per_s=len(probs) / t_bat,
                speedup=t_seq / t_bat, energies_identical=bool(e_seq == e_bat))


def serve_section(sb, a):
    """The serving loop (sparsebackends_amd.serve) on DMRG ground-state requests
    of XXZ chains (J = 1, Delta uniform in [0.5, 1.5], Neel start, sweeps
    until the energy changes by less than 1e-10, at most 12), lengths and bond
    caps drawn from --serve-sites / --serve-chis.  Three measurements on the
    same request mix:
    * saturated: all requests at t = 0; plain (one solve at a time, first come
      first served) vs served (slots on their own contexts and streams), with
      and without co-batched Krylov steps; solves/s, p50/p99, and whether
      every energy equals the plain one bit for bit;
    * Poisson: plain at 0.7 x its saturated rate, served at 3 x and 4 x that
      rate; solves/s and p99 of each (the target: 3 x the solves/s at no
      higher p99);
    * mixed: the served run with FP32 block products (FP64 storage, Lanczos
      and SVD), solves/s and the largest energy difference to FP64."""
    import dataclasses
    from sparsebackends_amd import serve as S
    sites = tuple(int(x) for x in a.serve_sites.split(","))
    chis = tuple(int(x) for x in a.serve_chis.split(","))
    n, slots = a.serve_requests, a.serve_slots
    base = S.poisson_requests(n, float("inf"), seed=0, sites=sites, chis=chis)
    S.run_plain(base[:1])  # first-use costs (library load, contexts) out of the timed runs
    out = dict(model="XXZ J=1, Delta~U[0.5,1.5], Neel start, U(1) Sz, MPO w=5", requests=n, slots=slots,
               sites=list(sites), chi_max=list(chis), energy_tol=1e-10, krylovdim=3, maxiter=1,
               note="a reduced sample of the serving configuration (its lengths 32-128 and bond caps 128-1024 "
                    "take hours per 512 requests); same request distribution family")
    t0 = time.perf_counter()
    plain = S.run_plain(base)
    out["plain_saturated"] = dict(S.summarize(plain), wall_s=time.perf_counter() - t0)
    ref = {r.rid: r.energy for r in plain}
    srv = S.DMRGServer(slots=slots, co_batch=False)
    res = srv.run(base)
    out["served_saturated"] = dict(S.summarize(res), identical_to_plain=all(r.energy == ref[r.rid] for r in res),
                                   max_inflight=srv.max_inflight)
    srv = S.DMRGServer(slots=slots, co_batch=True, window=0.0005)
    res = srv.run(base)
    out["served_cobatched_saturated"] = dict(S.summarize(res),
                                             identical_to_plain=all(r.energy == ref[r.rid] for r in res),
                                             batches=len(srv.batch_sizes),
                                             mean_batch=float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0,
                                             max_batch=int(max(srv.batch_sizes)) if srv.batch_sizes else 0)
    out["saturated_speedup"] = out["served_saturated"]["solves_per_s"] / out["plain_saturated"]["solves_per_s"]
    # Poisson arrivals: the same mix (same seed), arrival times scaled by the rate
    mu = out["plain_saturated"]["solves_per_s"]
    rp = 0.7 * mu
    lp = S.summarize(S.run_plain(S.poisson_requests(n, rp, seed=1, sites=sites, chis=chis)))
    out["poisson"] = dict(plain=dict(lp, rate=rp))
    for mult in (3.0, 4.0):
        rs = mult * rp
        srv = S.DMRGServer(slots=slots, co_batch=False)
        ls = S.summarize(srv.run(S.poisson_requests(n, rs, seed=1, sites=sites, chis=chis)))
        out["poisson"][f"served_at_{mult:g}x"] = dict(ls, rate=rs,
                                                      solves_per_s_ratio=ls["solves_per_s"] / lp["solves_per_s"],
                                                      p99_ratio=ls["p99_s"] / lp["p99_s"])
    # mixed: FP32 block products
    mreq = [dataclasses.replace(r, flags=sb.PREC_FP32) for r in base]
    srv = S.DMRGServer(slots=slots, co_batch=False)
    res = srv.run(mreq)
    terr = {r.rid: r.truncerr for r in plain}
    out["served_mixed_fp32"] = dict(S.summarize(res),
                                    max_abs_dE_vs_fp64=float(max(abs(r.energy - ref[r.rid]) for r in res)),
                                    max_truncerr=float(max(r.truncerr for r in res)),
                                    max_truncerr_fp64=float(max(terr.values())),
                                    max_abs_dtruncerr_vs_fp64=float(max(abs(r.truncerr - terr[r.rid]) for r in res)))
    out["peak_over_footprint_estimate"] = float(max(r.peak_bytes / r.footprint for r in plain))
    out["largest_peak_mb"] = float(max(r.peak_bytes for r in plain) / 2 ** 20)
    return out


def plain_replay(reqs, service):
    """Latencies of the plain server (one solve at a time, first come first
    served) for the given arrivals: finish_i = max(arrival_i, finish_{i-1}) +
    service_i, with the service times measured in its saturated run."""
    order = sorted(reqs, key=lambda r: (r.arrival, r.rid))
    fin, lat = 0.0, {}
    for r in order:
        fin = max(r.arrival, fin) + service[r.rid]
        lat[r.rid] = fin - r.arrival
    la = np.array([lat[r.rid] for r in reqs])
    span = fin - min(r.arrival for r in reqs)
    return dict(solves=len(reqs), span_s=float(span), solves_per_s=len(reqs) / span, p50_s=float(np.percentile(la, 50)),
                p99_s=float(np.percentile(la, 99)), mean_s=float(la.mean()))


def serve5_section(sb, a):
    """The serving loop at the serving configuration's sizes: XXZ chains (J =
    1, Delta uniform in [0.5, 1.5], Neel start, U(1) Sz, MPO w=5), N drawn
    from {32, 64, 96, 128} and chi_max from {128, 256, 512, 1024}, energy
    tolerance 1e-10 between sweeps (at most --serve5-max-sweeps), cutoff 1e-12, Krylov
    dimension 3.  A sample of --serve-config5 requests (seed --serve5-seed)
    of the configuration's 512: one request there takes seconds to minutes.
    * plain: one solve at a time, FCFS, every request at t = 0 (its service
      times); at each Poisson rate its latencies follow exactly from those
      service times (plain_replay: the plain server idles only when no request
      is waiting);
    * served: the continuous-batching server (8 slots, admission by the HBM
      footprint estimate), measured at each Poisson rate;
    * mixed: the served run with FP32 block products and an FP64 polish
      (sweeps with FP64 products once the FP32 energy settles to 1e-6
      relative), measured at --serve5-mixed-rates.
    Served and plain energies are compared bit for bit, mixed ones against
    the FP64 plain energies."""
    import dataclasses
    from sparsebackends_amd import serve as S
    n = a.serve_config5
    parts = set(a.serve5_parts.split(","))
    base = S.poisson_requests(n, float("inf"), seed=a.serve5_seed, max_sweeps=a.serve5_max_sweeps)
    out = dict(model="XXZ J=1, Delta~U[0.5,1.5], Neel start, U(1) Sz, MPO w=5", requests=n,
               sample=f"{n} requests drawn with seed {a.serve5_seed} from N in (32,64,96,128) x chi_max in "
                      "(128,256,512,1024) (the configuration's stream has 512)",
               mix=[dict(rid=r.rid, sites=r.sites, chi_max=r.chi_max, delta=round(r.delta, 6)) for r in base],
               energy_tol=1e-10, max_sweeps=base[0].max_sweeps, cutoff=base[0].cutoff, krylovdim=3, maxiter=1,
               slots=8)
    S.run_plain([S.SolveRequest(10 ** 6, 8, 16, 1.0)])  # first-use costs out of the timed runs
    t_start = [time.perf_counter()]

    def sweep_line(r, k, e, low):  # long solves print as they go
        if r.sites * r.chi_max >= 128 * 256:
            print(f"[serve5]     request {r.rid} sweep {k} ({'FP32' if low else 'FP64'} products) E={e:.12f} at "
                  f"{time.perf_counter() - t_start[0]:.1f} s", flush=True)

    def solver(r, eig, dev):  # the library's solve, with a line per finished request
        res = S.solve_dmrg(r, eig, dev, on_sweep=sweep_line)
        print(f"[serve5]   request {r.rid} (N={r.sites}, chi_max={r.chi_max}) done at "
              f"{time.perf_counter() - t_start[0]:.1f} s, {len(res[1])} sweeps", flush=True)
        return res

    def server():
        srv = S.DMRGServer(slots=8, co_batch=False, solver=solver)
        srv.release = S.trim  # as the library's own solves do
        t_start[0] = time.perf_counter()
        return srv
    print(f"[serve5] {n} requests: " + ", ".join(f"({r.sites},{r.chi_max})" for r in base), flush=True)
    if a.serve5_plain_json:
        with open(a.serve5_plain_json) as f:
            pl = json.load(f)["plain"]
        service = {int(k): v for k, v in pl["service_s"].items()}
        ref = {int(k): v for k, v in pl["energies"].items()}
        out["plain"] = pl
    else:
        t0 = t_start[0] = time.perf_counter()
        plain = S.run_plain(base, solver=solver)
        service = {r.rid: r.finish - r.admitted for r in plain}
        ref = {r.rid: r.energy for r in plain}
        out["plain"] = dict(S.summarize(plain), wall_s=time.perf_counter() - t0, service_s=service, energies=ref,
                            sweeps={r.rid: r.sweeps for r in plain},
                            peak_gib={r.rid: r.peak_bytes / 2 ** 30 for r in plain},
                            footprint_gib={r.rid: r.footprint / 2 ** 30 for r in plain})
        print(f"[serve5] plain saturated: {out['plain']['solves_per_s']:.4f} solves/s, "
              f"converged {out['plain']['converged']}/{n}", flush=True)
    rates = [float(x) for x in a.serve5_rates.split(",") if x]
    out["rates"] = {}
    for rate in rates:
        reqs = S.poisson_requests(n, rate, seed=a.serve5_seed, max_sweeps=a.serve5_max_sweeps)
        row = dict(plain=plain_replay(reqs, service))
        if "served" in parts:
            srv = server()
            t0 = time.perf_counter()
            res = srv.run(reqs)
            row["served"] = dict(S.summarize(res), wall_s=time.perf_counter() - t0, max_inflight=srv.max_inflight,
                                 identical_to_plain=all(r.energy == ref[r.rid] for r in res))
            row["solves_per_s_ratio"] = row["served"]["solves_per_s"] / row["plain"]["solves_per_s"]
            row["p99_ratio"] = row["served"]["p99_s"] / row["plain"]["p99_s"]
            print(f"[serve5] rate {rate:g}/s served: {row['served']['solves_per_s']:.4f} solves/s "
                  f"(x{row['solves_per_s_ratio']:.2f}), p50 {row['served']['p50_s']:.1f} s, p99 "
                  f"{row['served']['p99_s']:.1f} s (x{row['p99_ratio']:.2f}), energies equal to plain: "
                  f"{row['served']['identical_to_plain']}", flush=True)
        out["rates"][f"{rate:g}"] = row
    if "mixed" in parts:
        for rate in [float(x) for x in a.serve5_mixed_rates.split(",") if x]:
            reqs = [dataclasses.replace(r, flags=sb.PREC_FP32) for r in S.poisson_requests(n, rate, seed=a.serve5_seed,
                                                                                     max_sweeps=a.serve5_max_sweeps)]
            srv = server()
            t0 = time.perf_counter()
            res = srv.run(reqs)
            pr = plain_replay(reqs, service)
            m = dict(S.summarize(res), wall_s=time.perf_counter() - t0,
                     max_abs_dE_vs_fp64=float(max(abs(r.energy - ref[r.rid]) for r in res)),
                     sweeps={r.rid: r.sweeps for r in res}, plain=pr,
                     solves_per_s_ratio_vs_plain=0.0, p99_ratio_vs_plain=0.0)
            m["solves_per_s_ratio_vs_plain"] = m["solves_per_s"] / pr["solves_per_s"]
            m["p99_ratio_vs_plain"] = m["p99_s"] / pr["p99_s"]
            out.setdefault("mixed", {})[f"{rate:g}"] = m
            print(f"[serve5] rate {rate:g}/s mixed: {m['solves_per_s']:.4f} solves/s, max |dE| "
                  f"{m['max_abs_dE_vs_fp64']:.2e}, converged {m['converged']}/{n}", flush=True)
    return out


def permute_section(sb, a, prob):
    """permutedims (storage.jl:188-269) of the chi=2048 Hubbard psi [Ql, Qr |
    s1, s2, l, r] (4430 blocks of 4 x 4 x 13 x 13): a tail permutation that
    keeps the fastest axis (-> s1, l, s2, r: element kernel) and one that
    moves it (-> l, r, s1, s2: LDS-tiled kernel).  Bytes = payload read and
    written + keys; time per call includes the key remap and sort."""
    nb, bs = int(prob.psi.nblocks), int(prob.psi.blksize)
    moved = 2 * nb * bs * 8 + 2 * nb * 2 * 4
    out = dict(blocks=nb, blksize=bs, bytes_per_call=moved)
    for name, perm in (("fastest_kept", [1, 2, 3, 5, 4, 6]), ("fastest_moved", [1, 2, 5, 6, 3, 4])):
        prob.psi.permutedims(perm)
        sb.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            prob.psi.permutedims(perm)
        sb.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[name] = dict(perm=perm, ms=dt * 1e3, gbs=moved / dt / 1e9, hbm_frac=moved / dt / 1e9 / HBM_PEAK_GBS)
    return out


def coo_section(sb, a):
    """The element-sparse workload: a PXP (Rydberg-blockade) chain's two-site
    psi at chi = 128 as a COOTensor (~30 % fill from the blockade mask), times
    the dense left environment (COO x Dense -> block-sparse), then the MPO
    (BS x Dense contract_shared!).  Reported: time per execution of each
    plan, algorithmic GB/s of the COO step against the HBM peak, the kernels
    it runs, parity with the oracle, and the oracle's time for the same step
    on one host core."""
    from sparsebackends_amd import pxp as X
    from oracle import oracle as O
    pr = X.problem(128, 0)
    chi = pr["chi"]
    A = sb.coo_from_dense(pr["psi"], atol=1e-12)
    nnz = int(len(A.parts()["vals"]))
    Ld, Wd = sb.DenseTensor.from_numpy(pr["L"]), sb.DenseTensor.from_numpy(pr["W"])
    C1 = sb.NewBlockSparseSorted.new((2, 2, chi, chi, 3), 2)
    p1 = sb.Plan(C1, X.L_C1, A, X.L_PSI, Ld, X.L_ENV)
    C2 = sb.NewBlockSparseSorted.new((2, chi, chi, 2, 3), 3)
    p2 = sb.Plan(C2, X.L_C2, C1, X.L_C1, Wd, X.L_W)
    reps = 200

    def timed(fn):
        fn()
        sb.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        sb.synchronize()
        return (time.perf_counter() - t0) / reps

    t1 = timed(p1.execute)
    t12 = timed(lambda: (p1.execute(), p2.execute()))
    sb.profile(True)
    sb.profile_reset()
    for _ in range(20):
        p1.execute()
    prof = sb.profile_read()
    sb.profile(False)
    nb = int(C1.nblocks)
    alg = X.algorithmic_bytes(nnz, chi, nb)
    # parity with the oracle on the same inputs
    Ao = O.COO.from_dense(pr["psi"], atol=1e-12)
    C1o = O.BS.new((2, 2, chi, chi, 3), 2)
    O.use_blas(True, threads=1)
    t0 = time.perf_counter()
    for _ in range(20):
        O.contract(C1o, X.L_C1, Ao, X.L_PSI, O.Dense.new(pr["L"]), X.L_ENV)
    tcpu = (time.perf_counter() - t0) / 20
    p, q = C1.parts(), C1o.parts()
    same = bool(np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"]) and
                np.array_equal(p["data"], q["data"]))
    dom = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    return dict(workload="PXP chain two-site psi (COO) x left environment (dense), chi=128, d=2, MPO w=3",
                fill=pr["fill"], nnz=nnz, out_blocks=nb, coo_dense_us=t1 * 1e6, coo_dense_then_mpo_us=t12 * 1e6,
                algorithmic_mb=alg / 1e6, achieved_gbs=alg / t1 / 1e9, hbm_frac=alg / t1 / 1e9 / HBM_PEAK_GBS,
                bound="latency (a few MB per execution: launch and synchronisation bound, far below the HBM roofline)",
                kernels_us={k: 1e3 * v["total_ms"] / v["count"] for k, v in prof.items()},
                dominant_kernel=dom[0], gpu_equals_oracle_bitwise=same,
                cpu_oracle_us=tcpu * 1e6, cpu_note="oracle C++ restatement, one host core")


def traffic_probe(sb, a):
    """the profiled child of measure_traffic: the fused plan applied three
    times on the workload, then a plain 256 MB copy (the counters' scale)"""
    prob = make_problem(a.workload)
    prob.upload().outputs()
    plan = prob.make_fused_plan(sb.PREC_FP64)
    for _ in range(3):
        plan.execute()
    sb.synchronize()
    calibration_copy(sb, 256)


def measure_traffic():
    """HBM bytes of one fused application, measured in this run: two
    rocprofv3 --pmc runs of this script in probe mode (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), summed over the last application's
    dispatches (the pack kernel and the fused kernel), each scaled by what the
    same counter reads for a plain copy of 256 MB.  None when rocprofv3 is
    missing or a run fails (the reason in the second value)."""
    import glob
    import shutil
    import sqlite3
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    out = {}
    for cn in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="sb_traffic_")
        cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", cn, "-d", d, "-o", "probe", "--", sys.executable,
               os.path.abspath(__file__), "--traffic-probe"]
        r = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ, TMPDIR=d))
        dbs = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)
        if r.returncode != 0 or not dbs:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"rocprofv3 --pmc {cn} failed (exit {r.returncode})"
        rows = []
        for db in dbs:
            c = sqlite3.connect(db)
            rows += list(c.execute("select dispatch_id, kernel_name, value from counters_collection "
                                   "where counter_name = ?", (cn,)))
            c.close()
        shutil.rmtree(d, ignore_errors=True)
        per = {}
        for did, kn, v in rows:
            k = per.setdefault(did, [kn, 0.0])
            k[1] += v
        disp = [per[i] for i in sorted(per)]
        copy = [v for kn, v in disp if "gather_blocks" in kn]
        packs = [i for i, (kn, v) in enumerate(disp) if "heff_pack_psi" in kn]
        if not copy or not packs:
            return None, f"rocprofv3 --pmc {cn}: dispatches not found"
        last = packs[-1]
        kib = disp[last][1] + sum(v for kn, v in disp[last + 1:] if "heff_kernel" in kn)
        scale = 256e6 / (copy[-1] * 1024.0)
        out[cn] = dict(kib=kib, scale=scale, bytes=kib * 1024.0 * scale)
    return out, None


def calibration_copy(sb, mb):
    """one plain copy kernel of mb MB each way (an identity permute), the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters"""
    n = int(mb * 1e6 / 8)
    T = sb.DenseTensor.from_numpy(np.zeros(n))
    sb.synchronize()
    T.permutedims([1])
    sb.synchronize()
    return n * 8


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {a.gpus} is not supported (this build measures one GPU's share; "
              "a multi-GPU run would shard the output block keys across GPUs)", file=sys.stderr)
        sys.exit(2)
    import sparsebackends_amd as sb

    if a.traffic_probe:
        traffic_probe(sb, a)
        return
    if a.serve_config5 > 0:
        print(json.dumps(dict(serving_config5=serve5_section(sb, a))), flush=True)
        return
    if a.only == "sweep":
        print(json.dumps(dict(sweep=sweep_section(sb, a))), flush=True)
        return
    if a.only == "dmrg-phases":
        print(json.dumps(dict(dmrg_phases=dmrg_phases_section(sb, a))), flush=True)
        return
    if a.only == "batch":
        res = dict(batch=batch_section(sb, a))
        if a.calibrate_copy > 0:
            res["calibration_copy_bytes"] = calibration_copy(sb, a.calibrate_copy)
        print(json.dumps(res), flush=True)
        return
    flags = {"fp64": sb.PREC_FP64, "fp32": sb.PREC_FP32, "bf16x3": sb.PREC_BF16X3}[a.prec] | (a.tile << 16)
    tune = not a.no_autotune and a.tile == 0
    if tune:
        flags |= sb.AUTOTUNE
    t0 = time.perf_counter()
    prob = make_problem(a.workload)
    t_gen = time.perf_counter() - t0
    alg = prob.algorithmic_flops()["total"]
    pad = prob.padded_flops()["total"]
    useful = prob.useful_flops()
    prob_desc = (("two-site H_eff*psi, Hubbard t=1 U=4, d=4, U(1)xU(1), MPO w=6, chi=2048 in "
                  f"{prob.Bl.nsec}+{prob.Br.nsec} sectors (2D Gaussian sigma=7)") if a.workload == "hubbard-chi2048" else
                 "two-site H_eff*psi, Heisenberg J=1, U(1) Sz, MPO w=5, chi=256 in 17 sectors")
    prob.upload().outputs()
    sb.synchronize()
    t0 = time.perf_counter()
    plans = prob.make_plans(flags)
    sb.synchronize()
    t_plan = time.perf_counter() - t0
    stats = [p.stats() for p in plans]
    tiles = [p.box_config() for p in plans]
    fprec = {"fp64": sb.PREC_FP64, "fp32": sb.PREC_FP32}.get(a.prec)
    fused, t_fplan = None, None
    if a.path == "fused" and fprec is not None:
        t0 = time.perf_counter()
        try:
            fused = prob.make_fused_plan(fprec)
        except sb.SBError as e:  # sectors too large for the fused kernel: the chain does the work
            print(f"fused plan unavailable ({e}); timing the contract! plans", file=sys.stderr)
        sb.synchronize()
        t_fplan = time.perf_counter() - t0
    step = fused.execute if fused is not None else prob.execute
    out_t = prob.Hpsi_fused if fused is not None else prob.Hpsi

    def timed_steps(fn, n, products_only=True):
        for _ in range(a.warmup):
            fn()
        sb.synchronize()
        sb.profile(True, products_only=products_only)  # events around the block products only
        sb.profile_reset()
        sb.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        sb.synchronize()
        d = time.perf_counter() - t0
        pr = sb.profile_read()
        sb.profile(False)
        return d, pr

    dt, prof = timed_steps(step, a.steps)
    # the plain eager steps without any timing events
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    sb.synchronize()
    dte = time.perf_counter() - t0
    ms_step = 1e3 * dt / a.steps
    value = alg / (dt / a.steps) / 1e12
    alg_bytes = prob.algorithmic_bytes()["total"]

    def chain_roofline(prof_c, dt_c):
        """the box products of the four contract! plans: HBM roofline of grouped_gemm"""
        rec = prof_c["grouped_gemm"]
        lps = rec["count"] / a.steps
        avg = rec["total_ms"] / rec["count"]
        ai = alg / alg_bytes
        r = dict(bound="hbm", achieved=alg_bytes / lps / (avg * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                 kernel="grouped_gemm", avg_launch_ms=avg,
                 note=f"average over the {int(lps)} grouped_gemm launches of a step; {alg_bytes / 1e9:.3f} GB "
                      f"algorithmic per step, arithmetic intensity {ai:.2f} FLOP/B")
        r["frac"] = r["achieved"] / r["peak"]
        return r

    if fused is not None:
        # dominant kernel: the fused application (the psi pack kernel and the
        # fused kernel, one plan step), timed by the events around it.  The
        # fused algorithm multiplies only psi's nonzero physical rows: its
        # useful FLOPs over its algorithmic bytes are far below the FP64 ridge
        # (78.6 TFLOP/s / 8 TB/s = 9.8 FLOP/B), so its roofline is HBM.
        fst = fused.stats()
        mf = fused.mfma_flops()
        rec = prof["heff_fused"]
        avg_ms = rec["total_ms"] / rec["count"]
        ab = fst["gemm_bytes"]
        ai = fst["flops"] / ab
        roof = dict(bound="hbm", achieved=ab / (avg_ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    kernel="heff_fused (heff_pack_psi + heff_kernel: one application)", avg_launch_ms=avg_ms,
                    algorithmic_bytes=ab, arithmetic_intensity_flop_per_byte=ai,
                    useful_gflop=fst["flops"] / 1e9, mfma_gflop=mf / 1e9,
                    note=(f"algorithmic bytes per application: psi read inside the environments' reach, L and R "
                          f"inside their reach, Hpsi written over every block's reach ({ab / 1e6:.1f} MB); useful "
                          f"FLOPs of the fused algorithm {fst['flops'] / 1e9:.4f} GFLOP "
                          f"({fst['flops'] / (avg_ms * 1e-3) / 1e12:.2f} TFLOP/s, {ai:.2f} FLOP/B), matrix-core "
                          f"FLOPs issued {mf / 1e9:.3f} GFLOP ({mf / (avg_ms * 1e-3) / 1e12:.2f} TFLOP/s, "
                          f"{mf / (avg_ms * 1e-3) / 1e12 / PEAK['fp64'] * 100:.1f} % of the FP64 matrix-core peak); "
                          f"the four-call chain's algorithmic {alg / 1e9:.3f} GFLOP per application "
                          f"({alg / (avg_ms * 1e-3) / 1e12:.2f} TFLOP/s, the work unit of value)"))
        roof["frac"] = roof["achieved"] / roof["peak"]
        roof["traffic"] = None
        if a.prec == "fp64" and not a.no_traffic:
            tr, why = measure_traffic()
            if tr is not None:
                roof["traffic"] = tr["FETCH_SIZE"]["bytes"] + tr["WRITE_SIZE"]["bytes"]
                roof["traffic_detail"] = dict(
                    fetch_bytes=tr["FETCH_SIZE"]["bytes"], write_bytes=tr["WRITE_SIZE"]["bytes"],
                    fetch_scale=tr["FETCH_SIZE"]["scale"], write_scale=tr["WRITE_SIZE"]["scale"],
                    source="measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py "
                           "--traffic-probe (child processes), the last application's dispatches, each counter "
                           "scaled by a plain 256 MB copy", traffic_over_algorithmic=roof["traffic"] / ab)
            else:
                roof["traffic_note"] = why
    else:
        roof = chain_roofline(prof, dt)
        roof["traffic"] = None

    out = {
        "metric": "H_eff·ψ TFLOP/s FP64 vs mixed at |ΔE|≤1e-10; DMRG solves/s + p99 latency",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (f64 storage)", "bf16x3": "bf16x3->f32 (f64 storage)"}[a.prec],
        "data": "synthetic: random psi and environments on the charge-allowed entries (seed 0), environments "
                "symmetrised with the MPO link involution (H_eff symmetric), exact model MPO",
        "config": {
            "workload": prob_desc,
            "path": "fused H_eff plan (one kernel per application)" if fused is not None else "four contract! plans",
            "format": "sector + in-sector axes, dense tail padded to the largest sector",
            "psi_blocks": int(len(prob.psi_parts["ids"])),
            "algorithmic_gflop_per_step": alg / 1e9,
            "algorithmic_gflop_note": "sum of 2*M*K*N over the matched block pairs of the four contract! calls, "
                                      "actual sector extents (the work unit of value)",
            "algorithmic_gb_per_step": alg_bytes / 1e9,
            "padded_gflop_per_step": pad / 1e9,
            "useful_gflop_per_step": useful["total"] / 1e9,
            "useful_gflop_note": "the same sum with the two MPO steps counted over the MPO's nonzero entries only "
                                 "(%.1f %% of the dense MPO block); useful TFLOP/s = %.2f" % (
                                     100 * useful["mpo_nonzero_fraction"], useful["total"] / (dt / a.steps) / 1e12),
            "block_pairs_per_step": int(sum(s["npairs"] for s in stats)),
            "plan_s": t_fplan if fused is not None else t_plan,
            "chain_plan_s": t_plan,
            "generate_s": t_gen,
        },
        "kernels": {k: dict(v, per_step_ms=v["total_ms"] / a.steps) for k, v in prof.items()},
        "roofline": roof,
        "eager_no_events": dict(ms_per_step=dte * 1e3 / a.steps, tflops=alg / (dte / a.steps) / 1e12,
                                note="the same steps without the HIP events of the timed region"),
    }
    dtc = dt
    if fused is not None:
        out["config"]["fused_useful_gflop_per_step"] = fst["flops"] / 1e9
        out["config"]["fused_sector_pairs"] = fst["npairs"]
        # the four contract! plans on the same problem (the reference's call pattern), and its HBM roofline
        dtc, profc = timed_steps(prob.execute, a.steps)
        graph = sb.PlanGraph(plans)
        for _ in range(3):
            graph.launch()
        sb.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            graph.launch()
        sb.synchronize()
        dtg = time.perf_counter() - t0
        out["chain"] = dict(ms_per_step=dtc * 1e3 / a.steps, tflops=alg / (dtc / a.steps) / 1e12,
                            graph_ms_per_step=dtg * 1e3 / a.steps, roofline=chain_roofline(profc, dtc),
                            kernels={k: dict(v, per_step_ms=v["total_ms"] / a.steps) for k, v in profc.items()},
                            fused_speedup=dtc / dt,
                            fused_vs_chain_rel_err=rel_err(prob.Hpsi_fused.parts()["data"], prob.Hpsi.parts()["data"]),
                            same_block_list=bool(np.array_equal(prob.Hpsi_fused.parts()["ids"],
                                                                prob.Hpsi.parts()["ids"])))

    names = ["16x128", "128x16", "32x32", "64x32", "32x64", "64x64", "16x64", "64x16",
             "wave16x16", "wave32x16", "wave16x32", "wave32x32", "wave64x16", "wave16x64"]
    out["tiles"] = dict(autotuned=tune, per_contraction=[
        dict(tile=names[t["tile"]] if 0 <= t["tile"] < len(names) else t["tile"], wg_per_cu=t["wg_per_cu"] or "resident",
             candidates_timed=t["candidates"], best_ms=t["best_ms"]) for t in tiles])
    if tune:
        # the same step with the cost model's tile shapes, for comparison
        pc = copy.copy(prob).upload().outputs()
        pc.make_plans(flags & ~sb.AUTOTUNE)
        pc.execute()
        sb.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            pc.execute()
        sb.synchronize()
        dcm = (time.perf_counter() - t0) / a.steps
        out["tiles"]["cost_model_ms_per_step"] = dcm * 1e3
        out["tiles"]["cost_model_tiles"] = [names[p.box_config()["tile"]] for p in pc.plans]
        out["tiles"]["autotuned_speedup"] = dcm / ((dtc if fused is not None else dt) / a.steps)
        out["tiles"]["same_result"] = bool(np.array_equal(pc.Hpsi.parts()["data"], prob.Hpsi.parts()["data"]))
        del pc

    fused32 = None
    if not a.no_mixed and a.prec == "fp64":
        # the precision forms on the same problem, with their error vs FP64
        ref = out_t.parts()["data"]
        ref_chain = prob.Hpsi.parts()["data"]
        mixed = {}
        forms = [("fp32_chain", sb.PREC_FP32, False), ("bf16x3_chain", sb.PREC_BF16X3, False)]
        if fused is not None:
            forms.insert(0, ("fp32", sb.PREC_FP32, True))
        for name, fl, fz in forms:
            pm = copy.copy(prob).upload().outputs()  # same host data, fresh device tensors
            if fz:
                plan_lo = pm.make_fused_plan(fl)
                run, res_t, r = plan_lo.execute, pm.Hpsi_fused, ref
            else:
                pm.make_plans(fl)
                run, res_t, r = pm.execute, pm.Hpsi, ref_chain
            run()
            sb.synchronize()
            t0 = time.perf_counter()
            for _ in range(max(3, a.steps // 2)):
                run()
            sb.synchronize()
            d2 = (time.perf_counter() - t0) / max(3, a.steps // 2)
            mixed[name] = dict(tflops=alg / d2 / 1e12, ms_per_step=d2 * 1e3,
                               rel_err_vs_fp64=rel_err(res_t.parts()["data"], r),
                               path="fused H_eff plan" if fz else "four contract! plans")
            if fz:
                fused32 = (pm, [plan_lo])
            else:
                del pm
        out["mixed"] = mixed

    if not a.no_cpu:
        # the oracle (reference algorithm restated in C++, OpenBLAS 0.3.29 one
        # thread) on one full H_eff*psi of the same problem: ~10-20 s of CPU work
        from oracle import oracle as O
        from oracle.heff_oracle import HeffOracle
        O.use_blas(True, threads=1)
        h = HeffOracle(prob)
        t0 = time.perf_counter()
        hp = h.run()
        tc = time.perf_counter() - t0
        blas = O.use_blas(True, threads=1)
        blas_cfg = None
        try:
            import ctypes
            f = ctypes.CDLL(blas).openblas_get_config
            f.restype = ctypes.c_char_p
            blas_cfg = f().decode()
        except Exception:
            pass
        cpu_model = None
        try:
            with open("/proc/cpuinfo") as fh:
                cpu_model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
        except Exception:
            pass
        cpu = dict(value=alg / tc / 1e12, unit="TFLOP/s", cores=1, kind="port",
                   sample="one full two-site H_eff*psi of the same workload (4 contract! calls), "
                          "oracle C++ restatement with the loaded BLAS's dgemm on 1 thread",
                   seconds=tc, blas_path=blas, blas_config=blas_cfg, cpu_model=cpu_model,
                   host_cores=os.cpu_count())
        q = hp.parts()
        p = out_t.parts()
        if np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"]):
            cpu["gpu_vs_cpu_rel_err"] = rel_err(p["data"], q["data"])
        else:
            cpu["gpu_vs_cpu_rel_err"] = "block lists differ"
        out["cpu_baseline"] = cpu
        # the same with the BLAS on every host core (the orchestration stays one thread, as in the reference)
        nth = min(16, os.cpu_count() or 1)
        O.use_blas(True, threads=nth)
        t0 = time.perf_counter()
        HeffOracle(prob).run()
        tc2 = time.perf_counter() - t0
        O.use_blas(True, threads=1)
        out["cpu_baseline_all_cores"] = dict(value=alg / tc2 / 1e12, unit="TFLOP/s", cores=nth, kind="port",
                                             seconds=tc2, sample="as cpu_baseline, OpenBLAS on %d threads" % nth)

    if not a.no_eig and a.prec == "fp64":
        # last: the solves overwrite psi and Hpsi
        out["eigsolve"] = eig_section(sb, a, prob, plans, fused, fused32)
    if a.dmrg_solves > 0 and a.prec == "fp64" and not a.no_eig:
        out["dmrg_solve"] = dmrg_section(sb, a)

    if a.batch_problems > 0 and a.prec == "fp64" and not a.no_eig:
        out["batch"] = batch_section(sb, a)

    if not a.no_sweep and a.prec == "fp64" and not a.no_eig:
        out["sweep"] = sweep_section(sb, a)

    if not a.no_coo and a.prec == "fp64":
        out["coo"] = coo_section(sb, a)
        out["permute"] = permute_section(sb, a, prob)

    if a.serve_requests > 0 and a.prec == "fp64" and not a.no_eig:
        out["serve"] = serve_section(sb, a)

    if a.calibrate_copy > 0:
        out["calibration_copy_bytes"] = calibration_copy(sb, a.calibrate_copy)

    print(json.dumps(out))


if __name__ == "__main__":
    main()
