"""Benchmark: solved RHS/s at backward error 1e-10 on one MI355X.

Workloads (--workload):
  c3 (default)  the 70,000-bus synthetic polar power-flow Jacobian (n ~ 126k,
                nnz ~ 850k, CSR float64) with 128 right-hand sides (8 N(0,1)
                columns + 120 unit injections).  One step = ILU(0) numeric
                factorization + flexible block GMRES (FP64 basis, FP64 Gram /
                Cholesky / Hessenberg, FP32 ILU(0) solves inside a GMRES
                polynomial preconditioner) to per-RHS backward error <= 1e-10
                + the final FP64 per-RHS backward-error check.
  c2            9,000-bus Jacobian, 64 N(0,1) RHS, same step.
  c4            256 independent systems, n log-uniform in [500, 20000], RHS
                counts in {8, 16, 32, 64}: one step = one batched solve of all
                of them (ILU(0) factorizations, Krylov, backward-error check),
                its four RHS-count bins solved at the same time.  The kernel
                timers come from one extra step with the bins one after
                another.
  c5            the 70,000-bus pattern with 32 Newton value sets and 256 RHS:
                one step = for each value set, device LU refactorization with
                fixed pivots + FP32 triangular solves + up to 2 FP64
                refinement steps, each residual checked for backward error
                <= 1e-10 on every RHS (refinement stops once all are).
  cases         the reference's own benchmark sweep over the case files it
                ships (tests/golden/mtx): per case ILU(0) + block GMRES with
                its settings (atol 1e-10, rtol 0, memory 1, itmax 50), then
                the same without a preconditioner.
Inputs are resident in HBM before timing.  The timed steps run with the
in-process kernel timers off; the roofline comes from one more, untimed step
with them on.  Beside each: the plain FP64 GPU form on the same input, and
the reference algorithm on CPU arrays (the oracle) on a bounded sample.

python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload c3|c2|c4|c5|cases]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (nbus, p, rhs kind, n_unit, block width of the mixed form)
    "c3": (70000, 128, "injection", 120, 128),
    "c2": (9000, 64, "normal", 0, 64),
}
BERR = 1e-10
PEAK_HBM = 8.0e12
METRIC = "solved RHS/sec at 1e-10 backward error: mixed+batched vs plain FP64, 1 GPU"
FAMILIES = ["spmm", "trsv", "proj", "gemm", "factor", "dense_lu", "proj16", "gemm16"]
PEAK_FP64_MFMA = 78.6e12
# family -> (kernels its timers cover, "mfma" for matrix-core kernels else
# "hbm", stem of its counter summary profiles/<workload>_<stem>_traffic.md).
KERNELS = {
    "spmm": ("spmm_rm_kernel", "hbm", "spmm_rm_kernel"),
    "trsv": ("tri_level_kernel, tri_tail_kernel", "hbm", "tri_level_kernel"),
    "factor": ("fac_level_kernel, fac_tail_kernel, fac_partial_kernel", "hbm", "fac_level_kernel"),
    "dense_lu": ("dlu_diag_kernel, dgemm_batch_kernel, copy_block_kernel, diag_inv_copy_kernel", "mfma",
                 "dense_lu"),
    "proj": ("proj_partial_kernel or proj16_kernel, then proj_reduce_kernel", "mfma", "projection"),
    "gemm": ("gemm128d_kernel (FP64 128-wide) or block_gemm_kernel", "mfma", "block_gemm_kernel"),
    "proj16": ("proj16_kernel or proj_partial_kernel (widths <= 32), then proj_reduce_kernel", "hbm",
               "projection_narrow"),
    "gemm16": ("gemm16f_kernel or block_gemm_kernel (widths <= 32)", "hbm", "block_gemm_narrow"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def family_stats(bpf, ctx):
    out = {}
    for f, nm in enumerate(FAMILIES):
        n = C.c_longlong()
        ms, by, fl = C.c_double(), C.c_double(), C.c_double()
        bpf._lib.check(bpf.L.bpf_profile_read(ctx.handle, f, C.byref(n), C.byref(ms), C.byref(by), C.byref(fl)))
        out[nm] = {"launches": n.value, "ms": ms.value, "bytes": by.value, "flops": fl.value}
    return out


def log_families(fam):
    for k, v in (fam or {}).items():
        if v["launches"]:
            log(f"[bench]   {k:6s} launches {v['launches']:7d} time {v['ms']:9.2f} ms  "
                f"{v['bytes'] / max(v['ms'], 1e-9) / 1e6:8.1f} GB/s  "
                f"{v['flops'] / max(v['ms'], 1e-9) / 1e9:8.2f} TFLOP/s")


def family_roofline(fam, name):
    """(bound, achieved, peak, unit) of one kernel family.  A matrix-core
    family whose arithmetic intensity (algorithmic FP64 flops per algorithmic
    byte) reaches the ridge point, 78.6 TFLOP/s / 8 TB/s = 9.8 flop/B, is
    priced as flops / time against the MFMA peak; below it (narrow blocks),
    and for the other families, as algorithmic bytes / time against HBM."""
    d = fam[name]
    sec = d["ms"] * 1e-3
    ridge = PEAK_FP64_MFMA / PEAK_HBM
    if KERNELS[name][1] == "mfma" and (d["bytes"] <= 0 or d["flops"] / d["bytes"] >= ridge):
        return "mfma", (d["flops"] / sec / 1e12 if sec > 0 else 0.0), PEAK_FP64_MFMA / 1e12, "TFLOP/s"
    return "hbm", (d["bytes"] / sec / 1e9 if sec > 0 else 0.0), PEAK_HBM / 1e9, "GB/s"


def roofline_from(fam, dt_prof, workload):
    """Roofline object for the dominant kernel family (largest share of the
    timed steps) from the in-process kernel timers (HIP events around every
    launch on the solver's stream during the timed steps).  The other
    families' figures are listed under "families"."""
    if fam is None:
        return None
    timed_fams = [k for k in KERNELS if fam[k]["launches"]]
    dom = max(timed_fams, key=lambda k: fam[k]["ms"])
    d = fam[dom]
    bound, achieved, peak, unit = family_roofline(fam, dom)
    others = {}
    for k in timed_fams:
        b, a, pk, u = family_roofline(fam, k)
        others[k] = {"kernel": KERNELS[k][0], "bound": b, "achieved": round(a, 2), "unit": u,
                     "frac": round(a / pk, 4), "share": round(fam[k]["ms"] / (dt_prof * 1e3), 3),
                     "traffic": measured_traffic(workload, KERNELS[k][2])}
    return {
        "bound": bound, "kernel": KERNELS[dom][0],
        "achieved": round(achieved, 2), "peak": peak, "unit": unit,
        "frac": round(achieved / peak, 4), "traffic": measured_traffic(workload, KERNELS[dom][2]),
        "launches": d["launches"], "avg_launch_us": round(d["ms"] * 1e3 / max(1, d["launches"]), 2),
        "families": others,
    }


NO_KERNEL_TIMERS = False


def timed(bpf, ctx, step, steps, warmup, label, profile=True):
    """warmup untimed steps, then `steps` timed steps with the kernel timers
    off; with profile=True one more, untimed step runs with the kernel timers
    (HIP events around every instrumented launch) on, so their own cost never
    touches the timed steps.  Returns (seconds of the timed steps, list of the
    timed steps' results, kernel families of the instrumented step or None,
    seconds of the instrumented step or None)."""
    for w in range(warmup):
        r = step()
        log(f"[bench] warmup {label} {w}: {r[0]}")
    ctx.synchronize()
    out = []
    t = time.perf_counter()
    for i in range(steps):
        out.append(step())
    ctx.synchronize()
    dt = time.perf_counter() - t
    if not profile or NO_KERNEL_TIMERS:
        return dt, out, None, None
    try:
        bpf._lib.check(bpf.L.bpf_profile_enable(ctx.handle, 1))
        t_prof = time.perf_counter()
        r = step()
        ctx.synchronize()
        dt_prof = time.perf_counter() - t_prof
        fam = family_stats(bpf, ctx)
    finally:
        bpf._lib.check(bpf.L.bpf_profile_enable(ctx.handle, 0))
    log(f"[bench] kernel-timer step {label}: {r[0]} in {dt_prof:.3f}s")
    return dt, out, fam, dt_prof


def result_line(args, value, dt, dtype, config, extra):
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
    }
    res.update(extra)
    return res


def cpu_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([t.get("num_threads", 1) for t in threadpool_info()] or [1])
    except Exception:
        return os.cpu_count() or 1


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


# ----------------------------------------------------------------- c2 / c3
def cpu_baseline_single(S, p, k_target, budget_s=20.0, k_known=True):
    """Reference algorithm (src/block_gmres.jl restated on CPU arrays: MGS,
    LAPACK Householder, ILU0 solves) timed on a bounded sample: the first
    block iterations until ~budget_s, plus one timed MGS projection and
    update, extrapolated to k_target iterations (MGS cost grows linearly with
    the iteration index)."""
    oracle = _oracle()
    A = S.scipy()
    t0 = time.perf_counter()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    t_ilu = time.perf_counter() - t0
    N = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    times = []
    m = 1
    while True:
        t0 = time.perf_counter()
        oracle.block_gmres(A, S.B[:, :p], memory=m, N=N, restart=False, atol=0.0, rtol=1e-30, itmax=m)
        dt = time.perf_counter() - t0
        times.append((m, dt))
        if dt > budget_s / 3 or m >= k_target:
            break
        m *= 2
    # Iteration j (1-based) costs a + b j: the preconditioner, SpMM and
    # Householder step (a) plus MGS against j basis blocks (b per block).
    # b is timed directly on one n x p projection and update (the probes are
    # too short to resolve it); a from the probes with b's share removed.
    ms = np.array([t[0] for t in times], dtype=float)
    ts = np.array([t[1] for t in times])
    rng = np.random.default_rng(0)
    Vb = rng.standard_normal((S.n, p))
    Wb = rng.standard_normal((S.n, p))
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        Hb = Vb.T @ Wb
        Wb -= Vb @ Hb
    b_pair = (time.perf_counter() - t0) / reps
    a_iter = max(0.0, float(np.mean((ts - b_pair * ms * (ms + 1) / 2) / ms)))
    total = a_iter * k_target + b_pair * k_target * (k_target + 1) / 2
    total = max(total, float(ts[-1] / ms[-1] * k_target))
    return {
        "value": p / (t_ilu + total),
        "unit": "RHS/s",
        "cores": int(cpu_threads()),
        "kind": "port",
        "sample": (f"reference algorithm restated (oracle/oracle.py block_gmres: block MGS, LAPACK "
                   f"Householder, ILU0 solves) on the same {S.n}-unknown system with {p} RHS; timed "
                   f"for {', '.join(str(int(x)) for x in ms)} block iterations ({ts.sum():.1f} s) and one "
                   f"MGS projection + update ({b_pair:.3f} s), "
                   + (f"extrapolated to the {k_target} iterations the FP64 GPU run needed" if k_known else
                      f"extrapolated to {k_target} iterations (the FP64 GPU run was skipped)")),
    }


def bench_single(args, bpf, gen):
    nbus, p, kind, nunit, s_mixed = WORKLOADS[args.workload]
    if args.block_width:
        s_mixed = args.block_width
    t0 = time.time()
    grid, S = gen.make_system(nbus, p, seed=1000, rhs_seed=2000, rhs=kind, nunit=nunit)
    log(f"[bench] generated {args.workload}: n={S.n} nnz={S.nnz} p={p} in {time.time() - t0:.1f}s")
    ctx = bpf.default_context()
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    B = bpf.HipMatrix.from_numpy(S.B)
    P = bpf.ilu0(A, p)                # symbolic/level analysis (untimed) + first factorization
    nl, nu = P.levels()
    log(f"[bench] ILU0 levels lower={nl} upper={nu}")

    def step(precision, bw):
        P.factor_(A)
        X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=args.memory, restart=args.restart, itmax=2000,
                                timemax=120.0, berr_tol=BERR, precision=precision, block_width=bw,
                                max_refine=8,
                                poly_degree=(args.poly_degree if precision == "mixed" else
                                             (args.poly_degree or 4) if precision == "mixed32" else 0),
                                bmgs_blocks=args.bmgs_blocks if precision == "mixed" else 1,
                                inner_its=args.inner_its if precision == "mixed" else 0,
                                inner_width=args.inner_width,
                                reorthogonalization=args.orth == "bcgs2")
        be = bpf.backward_error(A, X, B)
        return (f"its {st.niter} max berr {be.max():.2e}", st, be)

    dt, out, fam, dt_prof = timed(bpf, ctx, lambda: step("mixed", s_mixed), args.steps, args.warmup, "mixed")
    _, st_last, be_last = out[-1]
    solved = sum(int(np.sum(be <= BERR)) for _, _, be in out)
    value = solved / dt
    log(f"[bench] mixed: {args.steps} steps {dt:.3f}s solved {solved}/{args.steps * p} "
        f"its {st_last.niter} refinements {st_last.refine_steps} max berr {be_last.max():.2e}")
    log_families(fam)
    roofline = roofline_from(fam, dt_prof, args.workload)
    plain = None
    k_plain = None
    if not args.no_plain:
        dtp, outp, _, _ = timed(bpf, ctx, lambda: step("fp64", 0), 1, 1, "fp64", profile=False)
        _, stp, bep = outp[-1]
        k_plain = stp.niter
        solvedp = int(np.sum(bep <= BERR))
        plain = {"value": solvedp / dtp, "ms_per_step": dtp * 1e3, "niter": stp.niter,
                 "max_berr": float(bep.max()), "solved": solvedp}
        log(f"[bench] plain fp64: {dtp:.3f}s solved {solvedp}/{p} its {stp.niter} berr {bep.max():.2e}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_single(S, p, k_plain or 60, k_known=k_plain is not None)
        log(f"[bench] cpu baseline: {cpu['value']:.3f} RHS/s ({cpu['sample']})")
    mixed32 = None
    if not args.no_plain:
        # the north star's split (FP32 basis too, FP64 refinement outside),
        # timed once on the same input for comparison
        dtm, outm, _, _ = timed(bpf, ctx, lambda: step("mixed32", s_mixed), 1, 0, "mixed32", profile=False)
        _, stm, bem = outm[-1]
        mixed32 = {"value": int(np.sum(bem <= BERR)) / dtm, "ms_per_step": dtm * 1e3, "niter": stm.niter,
                   "refinement_steps": stm.refine_steps, "max_berr": float(bem.max()),
                   "form": "FP32 basis (two classical Gram-Schmidt passes), FP32 ILU0 + degree-4 polynomial, "
                           "FP64 Gram/Cholesky/Hessenberg, FP64 iterative refinement with the true residual"}
        log(f"[bench] mixed32: {dtm:.3f}s its {stm.niter} refinements {stm.refine_steps} berr {bem.max():.2e}")
    config = {
        "workload": (f"{args.workload}: {nbus}-bus synthetic polar NR Jacobian n={S.n} nnz={S.nnz}, "
                     f"{p} RHS ({p - nunit} N(0,1) + {nunit} unit injections); flexible block GMRES "
                     + ("restarted every " + str(args.memory) + " iterations" if args.restart else
                        f"without restart (memory {args.memory} is a storage hint: restarted block GMRES(40) "
                        "does not converge on this system)")
                     + f", block width {s_mixed}, {args.orth}"
                     + (f" over {args.bmgs_blocks}-block steps" if args.orth == "bmgs" else "")
                     + "; precision split: FP64 basis, Gram-Schmidt, CholQR2, Hessenberg least squares, "
                     "outer SpMM and true residual; FP32 in the preconditioner step: "
                     + (f"{args.inner_its} iterations of an inner FP32 block GMRES (groups of "
                        f"{args.inner_width} columns, one classical Gram-Schmidt and one CholQR pass) with "
                        "FP32 SpMM and FP32 ILU0 triangular solves (the inner solve is an addition that "
                        "cuts the outer iterations; plain_fp64 is the reference's form)"
                        if args.inner_its > 0 else
                        f"FP32 ILU0 triangular solves and the FP32 SpMMs of a degree-{args.poly_degree} "
                        "GMRES polynomial around them (an addition that cuts the iterations; plain_fp64 is "
                        "the reference's form)")
                     + ", to per-RHS backward error 1e-10"),
        "n": S.n, "nnz": S.nnz, "nrhs": p, "block_width": s_mixed, "poly_degree": args.poly_degree,
        "inner_its": args.inner_its, "inner_width": args.inner_width,
        "orth": args.orth, "bmgs_blocks": args.bmgs_blocks, "restart": bool(args.restart),
        "iterations": st_last.niter, "refinement_steps": st_last.refine_steps,
        "max_backward_error": float(be_last.max()),
    }
    return result_line(args, value, dt, "f32+f64", config, {
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(value / plain["value"], 3) if plain and plain["value"] > 0 else None,
        "mixed32": mixed32,
        "roofline": roofline,
        "cpu_baseline": cpu,
    })


# ----------------------------------------------------------------- c4
def bench_batch(args, bpf, gen):
    nsys = 256
    t0 = time.time()
    nbus, ps = gen.batch_sizes(nsys, 4000)
    systems = [gen.make_system(int(nbus[i]), int(ps[i]), seed=5000 + i, rhs_seed=6000 + i)[1]
               for i in range(nsys)]
    ntot = sum(S.n for S in systems)
    nnz = sum(S.nnz for S in systems)
    nrhs = int(sum(ps))
    log(f"[bench] generated c4: {nsys} systems, sum n={ntot} sum nnz={nnz} sum p={nrhs} "
        f"in {time.time() - t0:.1f}s")
    ctx = bpf.default_context()
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    Xs = [bpf.HipMatrix(B.shape, np.float64) for B in Bs]
    t0 = time.time()
    batch = bpf.BatchedSystems(As, [int(x) for x in ps])   # stacked operators + ILU0 analysis (untimed)
    log(f"[bench] batch analysis {time.time() - t0:.1f}s")

    def step():
        _, st, solved = batch.solve(Bs, Xs, precision="mixed", berr_tol=BERR, memory=40, itmax=1000,
                                    timemax=120.0, poly_degree=args.poly_degree,
                                    reorthogonalization=args.orth == "bcgs2", bmgs_blocks=args.bmgs_blocks,
                                    inner_its=args.inner_its, inner_width=args.inner_width)
        nsolved = int(sum(int(ps[i]) for i in range(nsys) if solved[i]))
        return (f"block its {st.niter} solved RHS {nsolved}/{nrhs} max berr {st.max_berr:.2e}", st, nsolved)

    bpf.set_option("batch_threads", args.batch_threads)
    dt, out, _, _ = timed(bpf, ctx, step, args.steps, args.warmup, "batched", profile=False)
    iw = batch.inner_widths()
    widths = {int(q): int(w) for q, w in sorted(set(zip(np.asarray(ps).tolist(), iw.tolist())))}
    widths_txt = ", ".join(f"{w} for {q} RHS" for q, w in widths.items())
    log(f"[bench] inner Krylov block widths by RHS count: {widths_txt}"
        + (" (chosen by timing on the first solve)" if args.inner_width < 0 else ""))
    solved = sum(r[2] for r in out)
    value = solved / dt
    st_last = out[-1][1]
    # Independent check of the last step's solutions.
    be_max = max(float(bpf.backward_error(A, X, B).max()) for A, X, B in zip(As, Xs, Bs))
    log(f"[bench] batched: {args.steps} steps {dt:.3f}s solved {solved}/{args.steps * nrhs} "
        f"max berr (rechecked) {be_max:.2e} ({args.batch_threads} bins at a time)")
    # Kernel timers: the extra, untimed step runs the bins one after another.
    # Bins solved at the same time interleave their kernels, so an event pair
    # around one launch would also enclose other bins' work.
    bpf.set_option("batch_threads", 1)
    try:
        _, _, fam, dt_prof = timed(bpf, ctx, step, 0, 0, "batched, bins one after another")
    finally:
        bpf.set_option("batch_threads", args.batch_threads)
    log_families(fam)
    roofline = roofline_from(fam, dt_prof, "c4")
    plain = None
    if not args.no_plain:
        Ps = [bpf.ilu0(A) for A in As]

        def step_plain():
            ns = 0
            its = 0
            for i in range(nsys):
                Ps[i].factor_(As[i])
                X, st = bpf.block_gmres(As[i], Bs[i], N=Ps[i], ldiv=True, memory=40, itmax=1000,
                                        berr_tol=BERR, precision="fp64")
                ns += int(np.sum(bpf.backward_error(As[i], X, Bs[i]) <= BERR))
                its += st.niter
            return (f"block its {its} solved RHS {ns}/{nrhs}", its, ns)

        dtp, outp, _, _ = timed(bpf, ctx, step_plain, 1, 0, "fp64 per system", profile=False)
        plain = {"value": outp[-1][2] / dtp, "ms_per_step": dtp * 1e3, "niter": outp[-1][1],
                 "solved": outp[-1][2],
                 "form": "plain FP64 block GMRES + ILU0, one system after another"}
        log(f"[bench] plain fp64 per system: {dtp:.3f}s {outp[-1][0]}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(systems, budget_s=20.0)
        log(f"[bench] cpu baseline: {cpu['value']:.3f} RHS/s ({cpu['sample']})")
    config = {
        "workload": (f"c4: {nsys} independent synthetic NR Jacobians, n log-uniform in [500, 20000] "
                     f"(sum n={ntot}, sum nnz={nnz}), RHS counts in {{8,16,32,64}} (sum {nrhs}); batched "
                     f"flexible block GMRES, FP64 basis ({args.orth}"
                     + (f" over {args.bmgs_blocks}-block steps" if args.orth == "bmgs" else "")
                     + "), FP32 ILU0 solves"
                     f"{f' in a degree-{args.poly_degree} GMRES polynomial' if args.poly_degree else ''}, "
                     + (f"; preconditioner step: {args.inner_its} iterations of an inner FP32 block GMRES "
                        f"(each system's columns in inner Krylov blocks of {widths_txt} columns; the inner "
                        "chain captured in a HIP graph and replayed)" if args.inner_its > 0 else "")
                     + f"; per-system stopping, to per-RHS backward error 1e-10; the four RHS-count bins "
                     f"solved {args.batch_threads} at a time from separate host threads"),
        "inner_its": args.inner_its, "inner_width": args.inner_width, "inner_widths_by_rhs": widths,
        "nsys": nsys, "sum_n": ntot, "sum_nnz": nnz, "nrhs": nrhs, "orth": args.orth,
        "bmgs_blocks": args.bmgs_blocks, "batch_threads": args.batch_threads,
        "poly_degree": args.poly_degree,
        "block_iterations": st_last.niter, "max_backward_error": be_max,
    }
    return result_line(args, value, dt, "f32+f64", config, {
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(value / plain["value"], 3) if plain and plain["value"] > 0 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    })


def cpu_baseline_batch(systems, budget_s=20.0):
    """Reference algorithm on CPU arrays (oracle block_gmres + ILU0 solves)
    on a seeded random sample of the batch's systems, solved one after
    another until the time budget is spent; rate = solved RHS / time."""
    oracle = _oracle()
    rng = np.random.default_rng(9)
    order = rng.permutation(len(systems))
    t_all = 0.0
    solved = 0
    tried = []
    for i in order:
        S = systems[i]
        t0 = time.perf_counter()
        F = oracle.ilu0(S.indptr, S.indices, S.data)
        N = lambda Z, S=S, F=F: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
        X, st = oracle.block_gmres(S.scipy(), S.B, memory=40, N=N, atol=0.0, rtol=1e-12, itmax=1000)
        be = oracle.backward_error(S.indptr, S.indices, S.data, X, S.B)
        t_all += time.perf_counter() - t0
        solved += int(np.sum(be <= BERR))
        tried.append(int(i))
        if t_all > budget_s:
            break
    return {
        "value": solved / t_all, "unit": "RHS/s", "cores": int(cpu_threads()), "kind": "port",
        "sample": (f"reference algorithm restated (oracle/oracle.py block_gmres + ILU0, rtol 1e-12) on "
                   f"{len(tried)} of the 256 systems drawn at random (seed 9), solved one after another "
                   f"in {t_all:.1f} s"),
    }


# ----------------------------------------------------------------- c5
def bench_refactor(args, bpf, gen):
    p, nsets, nbus = 256, 32, 70000
    t0 = time.time()
    grid, S = gen.make_system(nbus, p, seed=1000, rhs_seed=2000, rhs="injection", nunit=240)
    vals = gen.newton_value_sets(grid, nsets, seed=7000)
    log(f"[bench] generated c5: n={S.n} nnz={S.nnz} p={p} {nsets} value sets in {time.time() - t0:.1f}s")
    ctx = bpf.default_context()
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    B = bpf.HipMatrix.from_numpy(S.B)
    X = bpf.HipMatrix(B.shape, np.float64)
    V = bpf.HipMatrix.from_numpy(np.column_stack(vals))     # nnz x nsets, resident in HBM
    t0 = time.time()
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, ordering=":AMD")    # host ordering + pivoted LU (untimed)
    log(f"[bench] rf analysis {time.time() - t0:.1f}s, factor nnz {rf.nnz_factor}, levels {rf.levels()}")
    CU = bpf.CUSOLVERRF

    def step_mixed():
        ns = 0
        worst = 0.0
        steps = 0
        for k in range(nsets):
            A.set_values_device(V, k)
            CU.update_(rf, A)
            _, berr = rf.solve_refined(A, B, refine=2, X=X, berr_tol=BERR)
            steps += rf.refine_steps
            ns += int(np.sum(berr <= BERR))
            worst = max(worst, float(berr.max()))
        return (f"solved RHS {ns}/{nsets * p} max berr {worst:.2e} refinement steps {steps}", worst, ns, steps)

    def step_plain():
        ns = 0
        worst = 0.0
        for k in range(nsets):
            A.set_values_device(V, k)
            CU.update_(rf, A)
            CU._ldiv(X, rf, B)
            be = bpf.backward_error(A, X, B)
            ns += int(np.sum(be <= BERR))
            worst = max(worst, float(be.max()))
        return (f"solved RHS {ns}/{nsets * p} max berr {worst:.2e}", worst, ns)

    dt, out, fam, dt_prof = timed(bpf, ctx, step_mixed, args.steps, args.warmup, "refactor+fp32 solve+IR")
    solved = sum(r[2] for r in out)
    value = solved / dt
    log(f"[bench] refactor mixed: {args.steps} steps {dt:.3f}s solved {solved}/{args.steps * nsets * p} "
        f"{out[-1][0]}")
    log_families(fam)
    roofline = roofline_from(fam, dt_prof, "c5")
    plain = None
    if not args.no_plain:
        dtp, outp, _, _ = timed(bpf, ctx, step_plain, 1, 1, "refactor+fp64 solve", profile=False)
        plain = {"value": outp[-1][2] / dtp, "ms_per_step": dtp * 1e3, "solved": outp[-1][2],
                 "max_berr": outp[-1][1], "form": "FP64 refactorization + FP64 triangular solves"}
        log(f"[bench] refactor plain fp64: {dtp:.3f}s {outp[-1][0]}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_refactor(S, vals, rf, p)
        log(f"[bench] cpu baseline: {cpu['value']:.3f} RHS/s ({cpu['sample']})")
    config = {
        "workload": (f"c5: {nbus}-bus synthetic NR Jacobian pattern n={S.n} nnz={S.nnz}, {nsets} Newton "
                     f"value sets, {p} RHS; per set device LU refactorization (AMD ordering, fixed "
                     f"pivots) + FP32 triangular solves + up to 2 FP64 refinement steps, stopping once every RHS "
                     f"is at backward error {BERR:g}"),
        "n": S.n, "nnz": S.nnz, "nrhs": p, "value_sets": nsets, "nnz_factor": rf.nnz_factor,
        "max_backward_error": out[-1][1], "refinement_steps": out[-1][3],
    }
    return result_line(args, value, dt, "f32+f64", config, {
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(value / plain["value"], 3) if plain and plain["value"] > 0 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    })


def cpu_baseline_refactor(S, vals, rf, p, budget_s=20.0):
    """Reference algorithm on CPU arrays: refactorization with the same
    P, Q and factor pattern (oracle lu_refactor) and FP64 permuted
    triangular solves of all RHS, per value set, until the budget is spent."""
    oracle = _oracle()
    P, Q, rp, ci, _ = rf.factors()
    t_all = 0.0
    solved = 0
    done = 0
    for k in range(len(vals)):
        t0 = time.perf_counter()
        f = oracle.lu_refactor(S.indptr, S.indices, vals[k], P, Q, rp, ci)
        Y = oracle.trsm(oracle.TRSM_UNIT_LOWER, rp, ci, f, S.B[P, :])
        Y = oracle.trsm(oracle.TRSM_UPPER, rp, ci, f, Y)
        X = np.empty_like(Y)
        X[Q, :] = Y
        be = oracle.backward_error(S.indptr, S.indices, vals[k], X, S.B)
        t_all += time.perf_counter() - t0
        solved += int(np.sum(be <= BERR))
        done += 1
        if t_all > budget_s:
            break
    return {
        "value": solved / t_all, "unit": "RHS/s", "cores": 1, "kind": "port",
        "sample": (f"oracle lu_refactor (same P, Q, pattern) + FP64 permuted triangular solves of {p} RHS "
                   f"for the first {done} of the value sets ({t_all:.1f} s)"),
    }


# ----------------------------------------------------------------- cases
# The reference's benchmark sweep (benchmarks/benchmarks_amd.jl:24-66) over
# the case files it ships with a Jacobian (tests/golden/mtx, copies of its
# benchmarks/dump).
CASES = ["case9", "case57", "case118", "case300", "case1354pegase", "case_ACTIVSg2000",
         "case2869pegase", "case9241pegase", "case_ACTIVSg10k"]
CASE_KW = dict(atol=1e-10, rtol=0.0, memory=1, restart=False, itmax=50)


def bench_cases(args, bpf, gen):
    mtx = os.path.join(ROOT, "tests", "golden", "mtx")
    ctx = bpf.default_context()
    cases = []
    t0 = time.time()
    for name in CASES:
        Ah = bpf.mmread(os.path.join(mtx, f"Gx_{name}.mtx.gz"))
        Bh = bpf.mmread_dense(os.path.join(mtx, f"Gu_{name}.mtx.gz"))
        A = bpf.HipSparseMatrixCSR(Ah.indptr, Ah.indices, Ah.data)
        B = bpf.HipMatrix.from_numpy(Bh)
        # the reference's block-Jacobi leg: size_block = 32, nblocks = div(n, 32)
        # (benchmarks/benchmarks_nvidia.jl:46-51)
        BJ = bpf.BlockJacobiPreconditioner(A, max(1, Ah.shape[0] // 32))
        cases.append({"name": name, "Ah": Ah, "Bh": Bh, "A": A, "B": B, "P": bpf.ilu0(A), "BJ": BJ})
    log(f"[bench] read {len(cases)} cases from tests/golden/mtx in {time.time() - t0:.1f}s")

    def solve(c, precond):
        if precond == "ilu0":
            c["P"].factor_(c["A"])
            return bpf.block_gmres(c["A"], c["B"], N=c["P"], ldiv=True, **CASE_KW)
        if precond == "bj":
            bpf.update(c["BJ"], c["A"])
            return bpf.block_gmres(c["A"], c["B"], N=c["BJ"], **CASE_KW)
        return bpf.block_gmres(c["A"], c["B"], **CASE_KW)

    def step(full=True):
        # the sweep as the reference runs it: per case the ILU(0)-preconditioned
        # solve (ILU(0) factorization included), the block-Jacobi one (block
        # inverses included), then the unpreconditioned one; full=False runs
        # the ILU(0) legs only (the warmup and the kernel-timer sweep: the two
        # largest cases' one-block legs take about two minutes per sweep)
        rows = []
        legs = (("ilu0", "ilu0"), ("bj", "bj"), (None, "none")) if full else (("ilu0", "ilu0"),)
        for c in cases:
            r = {}
            for precond, key in legs:
                t = time.perf_counter()
                X, st = solve(c, precond)
                ctx.synchronize()
                r[key] = (time.perf_counter() - t, X, st)
            rows.append(r)
            # one line per case: the larger cases take seconds each
            log(f"[bench]   {c['name']}: " + ", ".join(
                f"{k} {v[2].niter} its {v[0] * 1e3:.1f} ms" for k, v in r.items()))
        ns = sum(c["Bh"].shape[1] for c, r in zip(cases, rows) if r["ilu0"][2].solved)
        return (f"{ns} RHS solved with ILU(0)", rows)

    for w in range(args.warmup):
        log(f"[bench] warmup sweep (ILU(0) legs) {w}: {step(False)[0]}")
    ctx.synchronize()
    t = time.perf_counter()
    out = [step(True) for _ in range(args.steps)]
    ctx.synchronize()
    dt = time.perf_counter() - t
    fam, dt_prof = None, None
    if not NO_KERNEL_TIMERS:
        try:
            bpf._lib.check(bpf.L.bpf_profile_enable(ctx.handle, 1))
            t_prof = time.perf_counter()
            step(False)
            ctx.synchronize()
            dt_prof = time.perf_counter() - t_prof
            fam = family_stats(bpf, ctx)
        finally:
            bpf._lib.check(bpf.L.bpf_profile_enable(ctx.handle, 0))
    rows = out[-1][1]
    table = []
    solved_rhs, t_prec = 0, 0.0
    for c, r in zip(cases, rows):
        Ah, Bh = c["Ah"], c["Bh"]
        ent = {"case": c["name"], "n": Ah.shape[0], "nnz": Ah.nnz, "p": Bh.shape[1]}
        for key in ("ilu0", "bj", "none"):
            sec, X, st = r[key]
            be = bpf.backward_error(c["A"], X, c["B"])
            R = Bh - _csr_matvec(Ah, X.to_numpy())
            ent[key] = {"ms": round(sec * 1e3, 3), "niter": st.niter, "solved": bool(st.solved),
                        "rnorm": float(np.linalg.norm(R)), "max_berr": float(be.max())}
        table.append(ent)
        if ent["ilu0"]["solved"] and ent["ilu0"]["max_berr"] <= BERR:
            solved_rhs += ent["p"]
        t_prec += ent["ilu0"]["ms"] * 1e-3
        log(f"[bench]   {c['name']:18s} n={ent['n']:6d} p={ent['p']:5d}  ILU(0): its {ent['ilu0']['niter']:3d} "
            f"solved {ent['ilu0']['solved']!s:5s} |R| {ent['ilu0']['rnorm']:.2e} {ent['ilu0']['ms']:9.2f} ms"
            f"   block Jacobi: its {ent['bj']['niter']:3d} solved {ent['bj']['solved']!s:5s} "
            f"|R| {ent['bj']['rnorm']:.2e} {ent['bj']['ms']:9.2f} ms"
            f"   none: its {ent['none']['niter']:3d} solved {ent['none']['solved']!s:5s} "
            f"|R| {ent['none']['rnorm']:.2e} {ent['none']['ms']:9.2f} ms")
    value = solved_rhs / t_prec
    log_families(fam)
    roofline = roofline_from(fam, dt_prof, "cases")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_cases(cases)
        log(f"[bench] cpu baseline: {cpu['value']:.3f} RHS/s ({cpu['sample']})")
    config = {
        "workload": ("cases: the reference benchmark sweep (benchmarks/benchmarks_amd.jl) over its "
                     f"{len(cases)} case files with a Jacobian, A = Gx (CSR FP64), B = dense Gu; per case "
                     "ILU(0) + block GMRES(atol=1e-10, rtol=0, memory=1, restart=false, itmax=50), then "
                     "the same with the block-Jacobi preconditioner (blocks of 32) and without a "
                     "preconditioner; value = RHS of the cases the ILU(0) run solves "
                     "(||R||_F <= 1e-10 and per-RHS backward error <= 1e-10) / the ILU(0) runs' time; "
                     "the warmup and the kernel-timer sweep run the ILU(0) legs only"),
        "cases": table,
    }
    res = result_line(args, value, dt, "f64", config, {
        "plain_fp64": None, "speedup_vs_plain_fp64": None, "roofline": roofline, "cpu_baseline": cpu})
    res["data"] = "reference case files (benchmarks/dump)"
    return res


def _csr_matvec(Ah, X):
    import scipy.sparse as sp
    return sp.csr_matrix((Ah.data, Ah.indices, Ah.indptr), shape=Ah.shape) @ X


def cpu_baseline_cases(cases, budget_s=20.0):
    """The reference algorithm restated on CPU arrays (the oracle's block
    GMRES + ILU(0), same settings) over the cases with n p <= 6e6 in sweep
    order, stopping once the budget is spent."""
    import scipy.sparse as sp
    o = _oracle()
    nrhs, t_all, done = 0, 0.0, []
    for c in cases:
        Ah, Bh = c["Ah"], c["Bh"]
        if Ah.shape[0] * Bh.shape[1] > 6e6 or t_all > budget_s:
            break
        As = sp.csr_matrix((Ah.data, Ah.indices, Ah.indptr), shape=Ah.shape)
        t = time.perf_counter()
        F = o.ilu0(Ah.indptr, Ah.indices, Ah.data)
        _, st = o.block_gmres(As, Bh, N=lambda Z: o.ilu0_apply(Ah.indptr, Ah.indices, F, Z), **CASE_KW)
        t_all += time.perf_counter() - t
        nrhs += Bh.shape[1]
        done.append(c["name"])
    return {"value": nrhs / t_all, "unit": "RHS/s", "cores": cpu_threads(), "kind": "port",
            "sample": (f"reference algorithm restated (oracle/oracle.py block_gmres + ILU0, same settings) on "
                       f"{', '.join(done)} in {t_all:.1f} s, RHS counted whether or not solved")}


def copy_probe(nbytes=1 << 30, copies=3):
    """Known-size copies for calibrating rocprofv3's FETCH_SIZE / WRITE_SIZE
    counters: each copy reads and writes nbytes of HBM."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(copies):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": copies}))


def measured_traffic(workload, stem):
    """HBM bytes per timed launch of the dominant family from the committed
    counter summary (profiles/<workload>_<stem>_traffic.md, line 'bytes per
    launch: N'), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{stem}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().lower().startswith("bytes per launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c4", "c5", "cases"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block-width", type=int, default=0)
    ap.add_argument("--memory", type=int, default=40,
                    help="Krylov basis blocks kept (c2/c3); with --restart the restart length")
    ap.add_argument("--restart", action="store_true", help="restarted block GMRES (c2/c3)")
    ap.add_argument("--bmgs-blocks", type=int, default=None,
                    help="basis blocks projected together per block-MGS step of the chosen form "
                         "(default 8 for c2/c3, 16 for c4; the plain FP64 form keeps the reference's "
                         "one-block loop)")
    ap.add_argument("--poly-degree", type=int, default=None,
                    help="GMRES-polynomial degree of the mixed form (default 0; used when --inner-its is 0)")
    ap.add_argument("--inner-its", type=int, default=None,
                    help="mixed form: iterations of the inner FP32 block GMRES that is the preconditioner "
                         "step (default 8 for c2/c3, 6 for c4)")
    ap.add_argument("--inner-width", type=int, default=None,
                    help="columns per independent Krylov block of the inner solve (default 16 for c2/c3; "
                         "c4: -1 = chosen per RHS-count bin by timing on the first, untimed solve)")
    ap.add_argument("--batch-threads", type=int, default=4,
                    help="c4: RHS-count bins solved at the same time, each from its own host thread")
    ap.add_argument("--orth", choices=["bmgs", "bcgs2"], default="bmgs",
                    help="block Gram-Schmidt: bmgs (the reference's loop, over --bmgs-blocks steps) or "
                         "bcgs2 (two classical passes)")
    ap.add_argument("--no-graph", action="store_true",
                    help="issue the inner solves' launches from the host instead of replaying a HIP graph "
                         "(the same kernels; for rocprofv3 runs, whose kernel tracing does not survive graph "
                         "instantiation here)")
    ap.add_argument("--no-kernel-timers", action="store_true",
                    help="skip the extra kernel-timer step (no roofline families; for kernel-trace runs)")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="process-wide library option (bpf.set_option) set before the run; repeatable")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three device-to-device copies of 1 GiB (the known-size kernel that "
                         "scales the HBM byte counters) and exit")
    args = ap.parse_args()
    global NO_KERNEL_TIMERS
    NO_KERNEL_TIMERS = args.no_kernel_timers
    if args.copy_probe:
        copy_probe()
        return
    batched = args.workload == "c4"
    if args.inner_its is None:
        args.inner_its = 6 if batched else 8
    if args.poly_degree is None:
        args.poly_degree = 0
    if args.bmgs_blocks is None:
        args.bmgs_blocks = 16 if batched else 8
    if args.inner_width is None:
        args.inner_width = -1 if batched else 16
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    bpf = ge.package()
    if args.no_graph:
        bpf.set_option("graph", 0)
    for opt in args.option:
        name, _, value = opt.partition("=")
        bpf.set_option(name, int(value))
    gen = ge._load("bpf_generator", os.path.join(ge.PKG_DIR, "generator.py"))
    if args.workload == "c4":
        res = bench_batch(args, bpf, gen)
    elif args.workload == "c5":
        res = bench_refactor(args, bpf, gen)
    elif args.workload == "cases":
        res = bench_cases(args, bpf, gen)
    else:
        res = bench_single(args, bpf, gen)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
