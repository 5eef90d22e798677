"""Benchmark: solved DOF/s of the assembly + Jacobi-PCG hot path.

Workload (one GPU): 3-D scalar diffusion -div(k grad u) = 1 on the unit cube,
128^3 Q1 hex mesh (2,130,048 active DoF, 56,621,950 CSR entries), 2x2x2 Gauss,
k_e = exp(g_e), g_e ~ N(0,1) per element (seed 12345), u = 0 on x = 0,
relative residual 1e-8.

A step = one full pass of the hot path on inputs resident in HBM: element
assembly into the CSR values and RHS, then Jacobi-PCG to |b - A x|/|b| <= 1e-8.
The mesh, DoF numbering and CSR pattern are set up once before the timed
region (reported as setup_ms).  value = mixed-precision path (FP32 PCG sweeps,
FP64 residual replacement, SELL-64 storage); fp64_value = the plain FP64 path
(FP64 assembly + FP64 PCG on the exported CSR).

Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "solved DOF/s at 1e-8 rel. residual (mixed vs FP64 PCG); p50/p99 solve latency"


def spmv_bytes(n, nnz, nslice_vals, fmt, vbytes, xbytes):
    """Algorithmic bytes of one SpMV launch of the given storage."""
    if fmt == "sym":
        # upper-triangle values of the symmetric slices (each read once from
        # HBM; its mirror read hits the caches) + x read + q write
        return nslice_vals * vbytes + 2 * n * xbytes
    if fmt == "sell":
        # values (padded slices) + pattern id per row + x read + q write
        return nslice_vals * vbytes + n * 4 + 2 * n * xbytes
    # values + column indices + row pointers + x read + q write
    return nnz * (vbytes + 4) + n * 4 + 2 * n * xbytes


def run_path(ctx, prob, mode, fmt, precision, rtol, steps, warmup, profile=False):
    """warmup + steps passes of assembly + solve; with profile, every SpMV
    launch of the timed steps is bracketed by HIP events on the context
    stream (recorded back to back, read after each solve)."""
    from insilico_amd import abi
    lat, its, rels, asm, sol = [], [], [], [], []
    spmv_ms, spmv_calls, rep_ms, rep_calls = 0.0, 0, 0.0, 0
    run_path.element_ms = []
    ctx.set_profiling(profile)
    for s in range(warmup + steps):
        ctx.sync()
        t0 = time.perf_counter()
        ctx.assemble(precision, abi.ASM_MATRIX | abi.ASM_RHS)
        it, rel, _ = ctx.solve(mode, fmt, rtol)
        ctx.sync()
        t1 = time.perf_counter()
        if s >= warmup:
            st = ctx.stats()
            lat.append((t1 - t0) * 1e3)
            its.append(it)
            rels.append(rel)
            asm.append(st["assemble_ms"])
            sol.append(st["solve_ms"])
            spmv_ms += st["spmv_ms"]
            spmv_calls += st["spmv_calls"]
            rep_ms += st["rep_ms"]
            rep_calls += st["rep_calls"]
            if profile:
                run_path.element_ms.append(st["element_ms"])
    ctx.set_profiling(False)
    return lat, its, rels, asm, sol, (spmv_ms / max(spmv_calls, 1), spmv_calls,
                                      rep_ms / max(rep_calls, 1), rep_calls)


def _cpu_rate(prob, O, rp, ci, val, b, iters_fp64, nthreads, n_elem_sample, k_iters):
    """Seconds per element of the oracle assembly (first n_elem_sample
    elements, contiguous z-slabs) and seconds per Jacobi-CG iteration on the
    full matrix, with nthreads OpenMP threads."""
    from insilico_amd import problems as P
    import dataclasses
    E_s = min(prob.n_elem, n_elem_sample)
    sub = dataclasses.replace(prob, conn=prob.conn[:E_s], coef=prob.coef[:E_s], faces=None,
                              fconn=prob.fconn[:E_s] if prob.deg == 2 else None)
    rp_s, ci_s = O.pattern(sub)
    _, _, _, _, t_asm_s = O.assemble(sub, rp_s, ci_s, nthreads=nthreads)
    t0 = time.perf_counter()
    O.pcg(rp, ci, val, b, tol=1e-30, maxit=k_iters, nthreads=nthreads)
    t_it = (time.perf_counter() - t0) / (k_iters + 1)
    t_total = prob.n_elem * (t_asm_s / E_s) + iters_fp64 * t_it
    return t_total, t_asm_s, E_s, t_it


def cpu_baseline(prob, ctx, iters_fp64, repeats=3):
    """The oracle (reference algorithm restated in C, -O3 -fopenmp) on the host
    cores: assembly of a bounded element sample and a bounded number of
    Jacobi-CG iterations on the full matrix, extrapolated to the full job;
    median of `repeats` samples.  Two settings: all host cores (the
    NTHREADS=0 convention of the reference's element loop) and one thread
    (its default NTHREADS=1)."""
    from oracle import pyoracle as O
    ncores = len(os.sched_getaffinity(0))
    rp, ci = ctx.get_pattern()
    val = ctx.get_csr_values()
    b = ctx.get_rhs()
    # about 10-15 s of host work in all: ~3 s per all-core sample, ~0.8 s per
    # one-thread sample at 128^3
    k_all, k_one = 200, 10
    alls = [_cpu_rate(prob, O, rp, ci, val, b, iters_fp64, ncores, 1048576, k_all)
            for _ in range(repeats)]
    ones = [_cpu_rate(prob, O, rp, ci, val, b, iters_fp64, 1, 65536, k_one)
            for _ in range(repeats)]
    t_all, t_asm_s, E_s, t_it = sorted(alls)[repeats // 2]
    t_one, t_asm1, E_1, t_it1 = sorted(ones)[repeats // 2]
    return {
        "value": prob.n_active / t_all, "unit": "DOF/s", "cores": ncores, "kind": "port",
        "sample": "oracle (reference algorithm in C, -O3 -fopenmp, %d threads): assembly of "
                  "%d of %d elements (%.2f s) and %d Jacobi-CG iterations (OpenMP CSR SpMV) on "
                  "the full matrix (%.3f s/iter), extrapolated to %d elements + %d iterations; "
                  "median of %d" % (ncores, E_s, prob.n_elem, t_asm_s, k_all, t_it, prob.n_elem,
                                    iters_fp64, repeats),
        "t_assemble_s": prob.n_elem * t_asm_s / E_s, "t_iter_s": t_it,
        "value_1thread": prob.n_active / t_one,
        "sample_1thread": "1 thread: assembly of %d elements (%.2f s) and %d CG iterations "
                          "(%.3f s/iter), extrapolated likewise; median of %d"
                          % (E_1, t_asm1, k_one, t_it1, repeats),
    }


def read_traffic(workload_tag, kernel_tag):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary of this workload (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated); None when there is none."""
    path = os.path.join(ROOT, "profiles", "spmv_counters_%s.md" % workload_tag)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch[%s]" % kernel_tag):
            try:
                return float(line.split("=")[1].split()[0])
            except (IndexError, ValueError):
                return None
    return None


def copy_probe(n_bytes=1 << 30, reps=3):
    """A plain device copy of known size (reads and writes n_bytes per launch),
    run under the same rocprofv3 counter passes as the SpMV to calibrate
    FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    src = torch.empty(n_bytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n_bytes, "copy_probe_bytes_written": n_bytes,
                      "launches": reps}))


def run_newton(ctx, prob, pull, steps, tangent_precision, mode, fmt, tol=1e-8, max_newton=25):
    """Load stepping + full Newton of reference/06-elastic/compressible.cpp
    :245-335 on the device: per iteration the FP64 internal-force residual
    and the tangent (tangent_precision) with incremental lifting g - u, a
    Jacobi-PCG solve of the increment, u += du.  A step ends when
    |R|/|R0| <= tol (R0: the step's first residual).  Linear tolerance
    eta = min(1e-3, |R|/|R0|) (inexact Newton; the final accuracy is set by
    the FP64 residual)."""
    from insilico_amd import abi
    right = prob.meta["right_x"]
    g = np.zeros(prob.n_nodes * 3)
    g[right * 3] = pull / steps
    ctx.set_prescribed(g)
    ctx.set_dof_values(None)
    flags = abi.ASM_MATRIX | abi.ASM_RHS | abi.ASM_RESIDUAL | abi.ASM_INCREMENTAL
    hist, lin_its, t_asm, t_sol = [], 0, 0.0, 0.0
    ctx.sync()
    t0 = time.perf_counter()
    for step in range(steps):
        if step > 0:
            ctx.scale_prescribed((step + 1.0) / step)
        r0 = None
        for it in range(max_newton + 1):
            ctx.assemble(tangent_precision, flags)
            t_asm += ctx.stats()["assemble_ms"]
            r = ctx.rhs_norm() * prob.n_active
            r0 = r if r0 is None else r0
            if r <= tol * r0 or it == max_newton:
                hist.append((step, it, r / r0))
                break
            eta = min(1e-3, r / r0)
            k, _, st = ctx.solve(mode, fmt, eta, check=False)
            if st not in (abi.INS_OK,):
                raise RuntimeError("linear solve failed: %s" % abi.ERRORS.get(st, st))
            lin_its += k
            t_sol += ctx.stats()["solve_ms"]
            ctx.distribute(abi.DIST_ADD)
    ctx.sync()
    total = time.perf_counter() - t0
    u = ctx.get_dof_values()
    newton = sum(h[1] for h in hist)
    return {"seconds": round(total, 3), "newton_iterations": newton,
            "per_step": [h[1] for h in hist], "final_rel_residual": [float("%.3e" % h[2]) for h in hist],
            "linear_iterations": lin_its, "assemble_ms_total": round(t_asm, 1),
            "solve_ms_total": round(t_sol, 1), "max_u_x": float(np.abs(u[0::3]).max()),
            "max_u_y": float(np.abs(u[1::3]).max()),
            "dof_per_s": round(prob.n_active * newton / total, 1)}


def newton_main(args):
    from insilico_amd import abi, problems as P
    n = args.n or 64
    prob = P.neohooke3d(n, mu=1.0, lam=10.0)
    log("neo-Hookean %d^3: %d dof" % (n, prob.n_active))
    ctx = abi.Context(0)
    nnz = ctx.load_problem(prob)
    res = {}
    for name, prec, mode, fmt in (("fp32_tangent_mixed", abi.FP32, abi.SOLVE_MIXED, abi.FMT_SELL),
                                  ("fp64", abi.FP64, abi.SOLVE_FP64, abi.FMT_CSR)):
        res[name] = run_newton(ctx, prob, 0.2, 10, prec, mode, fmt)
        log("%s: %.2f s, %d Newton iterations" % (name, res[name]["seconds"],
                                                  res[name]["newton_iterations"]))
    a, b = res["fp32_tangent_mixed"], res["fp64"]
    out = {"metric": "Newton load stepping: solved DOF/s (Newton iterations x DoF / time)",
           "value": a["dof_per_s"], "unit": "DOF/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f32 tangent + f32 sweeps / f64 residual", "data": "synthetic",
           "config": {"workload": "compressible neo-Hookean (mu=1, lambda=10) %d^3 Q1 hex, clamped "
                                  "x=0, u_x on x=1 ramped to 0.2 over 10 steps, full Newton to "
                                  "|R|/|R0| <= 1e-8 per step" % n,
                      "n_dof": prob.n_active, "nnz": nnz},
           "fp32_tangent_mixed": a, "fp64": b,
           "speedup_vs_fp64": round(b["seconds"] / a["seconds"], 3),
           "max_u_y_rel_diff": abs(a["max_u_y"] - b["max_u_y"]) / max(b["max_u_y"], 1e-300)}
    print(json.dumps(out))
    ctx.close()


def transient_main(args):
    """heat::DiffusionDriver::advanceInTime steps (heat/DiffusionDriver.hpp
    :88-118) on the 3-D diffusion problem: BDF2 with its BDF1 start-up, each
    step = assembly of K + (a0/dt) M and of the RHS with the mass history
    on the device, Jacobi-PCG to rtol, solution pushed into the history,
    all on the device (ins_bdf_step)."""
    from insilico_amd import abi, problems as P
    n = args.n or 128
    prob = P.diffusion3d(n, seed=args.seed)
    dt, nsteps = 1e-3, max(args.steps, 2)
    res = {}
    for name, mode, fmt in (("mixed_sell", abi.SOLVE_MIXED, abi.FMT_SELL),
                            ("fp64_csr", abi.SOLVE_FP64, abi.FMT_CSR)):
        ctx = abi.Context(0)
        ctx.load_problem(prob)
        ctx.bdf_init(2, 1.0)
        ctx.bdf_step(dt, abi.FP64, mode, fmt, args.rtol)   # warm-up step (BDF1 start-up)
        ctx.sync()
        t0 = time.perf_counter()
        its, asm = [], []
        for _ in range(nsteps):
            it, _ = ctx.bdf_step(dt, abi.FP64, mode, fmt, args.rtol)
            its.append(it)
            asm.append(ctx.stats()["assemble_ms"])
        ctx.sync()
        sec = time.perf_counter() - t0
        u = ctx.get_dof_values()
        res[name] = {"ms_per_step": round(sec / nsteps * 1e3, 3), "iterations": its,
                     "assemble_ms": round(statistics.mean(asm), 3),
                     "max_u": float(np.abs(u).max())}
        log("transient %s: %.1f ms per time step, iterations %s" % (name, sec / nsteps * 1e3, its))
        ctx.close()
    a, b = res["mixed_sell"], res["fp64_csr"]
    out = {"metric": "time steps per second (BDF2 diffusion, assemble + solve per step)",
           "value": round(1e3 / a["ms_per_step"], 3), "unit": "steps/s", "n_gpus": 1,
           "steps": nsteps, "higher_is_better": True, "dtype": "f32 sweeps / f64 residual",
           "data": "synthetic",
           "config": {"workload": "3-D diffusion %d^3 Q1 hex, k=exp(N(0,1)), u=0 on x=0, f=1, "
                                  "rho=1, dt=%g, BDF2 (BDF1 first step), u0=0, Jacobi-PCG to %g "
                                  "per step" % (n, dt, args.rtol), "n_dof": prob.n_active},
           "mixed_sell": a, "fp64_csr": b,
           "speedup_vs_fp64": round(b["ms_per_step"] / a["ms_per_step"], 3),
           "max_u_rel_diff": abs(a["max_u"] - b["max_u"]) / max(b["max_u"], 1e-300)}
    print(json.dumps(out))


def fp32_integration_run(ctx, prob, fmt, args):
    """FP32 element integration feeding the same mixed solve; its result is
    measured against the FP64-integrated system with a host SpMV.  Not the
    headline: the refinement converges to the FP32-integrated matrix, so the
    error against the FP64 system is stated instead of a tolerance."""
    from insilico_amd import abi
    import scipy.sparse as sp
    lat32, its32, rel32, asm32, _, _ = run_path(ctx, prob, abi.SOLVE_MIXED, fmt, abi.FP32,
                                                args.rtol, args.steps, args.warmup)
    x32 = ctx.get_solution()
    ctx.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
    rp, ci = ctx.get_pattern()
    A = sp.csr_matrix((ctx.get_csr_values(), ci, rp), shape=(prob.n_active, prob.n_active))
    b64 = ctx.get_rhs()
    rel_vs64 = float(np.linalg.norm(b64 - A @ x32) / np.linalg.norm(b64))
    log("fp32 integration: %.1f ms/step, relres vs FP64 system %.2e"
        % (statistics.mean(lat32), rel_vs64))
    return {"ms_per_step": round(statistics.mean(lat32), 3),
            "value": round(prob.n_active / (statistics.mean(lat32) / 1e3), 1),
            "assemble_ms": round(statistics.mean(asm32), 3),
            "iterations": int(statistics.median(its32)),
            "relres_own_system": max(rel32),
            "relres_vs_fp64_system": rel_vs64}


def elasticity_run(n, seed, rtol, steps=2, warmup=1):
    """Linear elasticity n^3 (E_e ~ U[0.5, 2], nu = 0.3, clamped z = 0,
    traction on z = 1): assembly + PCG to rtol on three paths -- mixed
    precision (FP32 sweeps, FP64 residual replacement), plain FP64 over the
    CSR, FP64 over the SELL slices."""
    from insilico_amd import abi, problems as P
    t0 = time.perf_counter()
    prob = P.elasticity3d(n, seed=seed)
    t_gen = time.perf_counter() - t0
    out = {"workload": "3-D linear elasticity (StVenant at u=0) %d^3 Q1 hex, E~U[0.5,2] per "
                       "element, nu=0.3, clamped z=0, traction (0,0,-0.01) on z=1, assemble + "
                       "Jacobi-PCG to %g" % (n, rtol),
           "n_dof": prob.n_active, "host_generate_s": round(t_gen, 3)}
    with abi.Context(0) as ctx:
        out["nnz"] = ctx.load_problem(prob)
        for name, mode, fmt in (("mixed", abi.SOLVE_MIXED, abi.FMT_SELL),
                                ("fp64_csr", abi.SOLVE_FP64, abi.FMT_CSR),
                                ("fp64_sell", abi.SOLVE_FP64, abi.FMT_SELL)):
            lat, its, rels, asm, sol, _ = run_path(ctx, prob, mode, fmt, abi.FP64, rtol, steps,
                                                   warmup)
            ms = statistics.mean(lat)
            out[name] = {"ms_per_step": round(ms, 3), "value": round(prob.n_active / (ms / 1e3), 1),
                         "iterations": int(statistics.median(its)), "relres": max(rels),
                         "assemble_ms": round(statistics.mean(asm), 3)}
            if name == "mixed":
                out[name]["format"] = "sym" if ctx.stats()["sell_sym_offsets"] else "sell"
            log("elasticity %d^3 %s: %.1f ms/step, %d iterations" % (n, name, ms, its[0]))
    out["mixed_over_fp64"] = round(out["fp64_csr"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    out["mixed_over_fp64_sell"] = round(out["fp64_sell"]["ms_per_step"] /
                                        out["mixed"]["ms_per_step"], 3)
    return out


def newton_run(n):
    """Neo-Hookean n^3 load stepping (10 steps to u_x = 0.2) with full
    Newton: FP32 tangent + mixed PCG against FP64 tangent + FP64 PCG."""
    from insilico_amd import abi, problems as P
    prob = P.neohooke3d(n, mu=1.0, lam=10.0)
    res = {"workload": "compressible neo-Hookean (mu=1, lambda=10) %d^3 Q1 hex, clamped x=0, u_x "
                       "on x=1 ramped to 0.2 over 10 steps, full Newton to |R|/|R0| <= 1e-8 per "
                       "step" % n, "n_dof": prob.n_active}
    with abi.Context(0) as ctx:
        res["nnz"] = ctx.load_problem(prob)
        for name, prec, mode, fmt in (("fp32_tangent_mixed", abi.FP32, abi.SOLVE_MIXED, abi.FMT_SELL),
                                      ("fp64", abi.FP64, abi.SOLVE_FP64, abi.FMT_CSR)):
            res[name] = run_newton(ctx, prob, 0.2, 10, prec, mode, fmt)
            log("newton %d^3 %s: %.2f s, %d Newton iterations" % (
                n, name, res[name]["seconds"], res[name]["newton_iterations"]))
    a, b = res["fp32_tangent_mixed"], res["fp64"]
    res["speedup_vs_fp64"] = round(b["seconds"] / a["seconds"], 3)
    res["max_u_y_rel_diff"] = abs(a["max_u_y"] - b["max_u_y"]) / max(b["max_u_y"], 1e-300)
    return res


def serving_run(requests, calib, seed):
    """The serving driver (insilico_amd/apps/serve): a stream of 3-D diffusion
    and elasticity requests of uneven sizes (n^3, n log-uniform in [16, 96])
    with Poisson arrivals at 0.7 x the capacity measured on the first
    `calib` requests, admitted by free HBM, solved by continuous batching;
    arrival-to-result latencies and the alone-versus-mix bit check."""
    import subprocess
    exe = os.path.join(ROOT, "insilico_amd", "apps", "serve")
    if not os.path.exists(exe):
        return {"error": "apps/serve not built"}
    t0 = time.perf_counter()
    out = subprocess.run([exe, "--requests", str(requests), "--calib", str(calib), "--seed",
                          str(seed), "--verify", "4"], capture_output=True, text=True, timeout=400)
    if out.returncode != 0:
        return {"error": "serve exited %d: %s" % (out.returncode, out.stderr[-400:])}
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1]
    res = json.loads(line)
    res["wall_s"] = round(time.perf_counter() - t0, 2)
    log("serving %d requests: p50 %.1f ms, p99 %.1f ms, %d converged, %d/%d bitwise equal alone"
        % (requests, res["p50_latency_ms"], res["p99_latency_ms"], res["converged"],
           res["alone_vs_mix_bitwise_equal"], res["alone_vs_mix_checked"]))
    return res


MFMA_FP64_PEAK_TFS = 78.6


def element_roofline(prob, element_ms):
    """Element-kernel figures of the FP64 assembly.  For the degree-2
    Laplace kernel (one matrix-core product Ke = G^T diag(s) G per element)
    the algorithmic count is 2 NF^2 NQ dim flops per element (NF = 27 field
    functions, NQ Gauss points); the bound is the FP64 matrix-core peak."""
    out = {"avg_launch_ms": round(element_ms, 4)}
    if prob.deg == 2 and prob.kernel == 1:
        nf = 3 ** prob.dim
        nq = prob.nq ** prob.dim
        flops = 2.0 * nf * nf * nq * prob.dim * prob.n_elem
        nk = nf * (nf + 1) // 2
        out.update({"bound": "mfma", "kernel": "k_elem_laplace_q2<%d, %d, double>" % (prob.dim, prob.nq),
                    "flops_per_launch": flops, "unit": "TFLOP/s",
                    "achieved": round(flops / (element_ms / 1e3) / 1e12, 2),
                    "peak": MFMA_FP64_PEAK_TFS,
                    "frac": round(flops / (element_ms / 1e3) / 1e12 / MFMA_FP64_PEAK_TFS, 4),
                    "bytes_written": prob.n_elem * nk * 8})
    return out


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="diffusion",
                    choices=["diffusion", "elasticity", "newton", "q2", "transient"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-launch SpMV events in the timed steps")
    ap.add_argument("--fmt", default="sell", choices=["sell", "csr"])
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only the copy kernel that calibrates the HBM counters")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only: no elasticity, Newton or serving figures")
    ap.add_argument("--serve-requests", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.config == "newton":
        newton_main(args)
        return
    if args.config == "transient":
        transient_main(args)
        return

    from insilico_amd import abi, problems as P

    t0 = time.perf_counter()
    if args.config == "diffusion":
        prob = P.diffusion3d(args.n or 128, seed=args.seed)
        workload = ("3-D diffusion %d^3 Q1 hex, k=exp(N(0,1)) per element, u=0 on x=0, "
                    "assemble + Jacobi-PCG to %g" % (args.n or 128, args.rtol))
    elif args.config == "q2":
        prob = P.diffusion3d_q2(args.n or 64, seed=args.seed)
        workload = ("3-D diffusion %d^3 Q1 hex with a Q2 (HEX27) field, k=exp(N(0,1)) per "
                    "element, u=0 on x=0, 3x3x3 Gauss, assemble + Jacobi-PCG to %g"
                    % (args.n or 64, args.rtol))
    else:
        prob = P.elasticity3d(args.n or 96, seed=args.seed)
        workload = ("3-D linear elasticity (StVenant at u=0) %d^3 Q1 hex, E~U[0.5,2] per "
                    "element, nu=0.3, clamped z=0, traction (0,0,-0.01) on z=1, assemble + "
                    "Jacobi-PCG to %g" % (args.n or 96, args.rtol))
    t_gen = time.perf_counter() - t0
    log("generated %s: %d elements, %d dof (%.1f s)" % (prob.name, prob.n_elem, prob.n_active, t_gen))
    ctx = abi.Context(0)
    nnz = ctx.load_problem(prob)
    log("pattern: nnz=%d" % nnz)
    setup_ms = ctx.stats()["setup_ms"]
    fmt = abi.FMT_SELL if args.fmt == "sell" else abi.FMT_CSR

    # plain FP64 path
    lat64, its64, rel64, asm64, sol64, _ = run_path(ctx, prob, abi.SOLVE_FP64, abi.FMT_CSR,
                                                    abi.FP64, args.rtol, args.steps, args.warmup)
    log("fp64 path: %.1f ms/step, %d iterations" % (statistics.mean(lat64), its64[0]))
    # FP64 PCG over the SELL storage (the best FP64 storage, for reference)
    lat64s, its64s, rel64s, _, _, _ = run_path(ctx, prob, abi.SOLVE_FP64, abi.FMT_SELL, abi.FP64,
                                               args.rtol, args.steps, args.warmup)
    log("fp64 sell path: %.1f ms/step, %d iterations" % (statistics.mean(lat64s), its64s[0]))
    # mixed-precision path (headline)
    latm, itsm, relm, asmm, solm, _ = run_path(ctx, prob, abi.SOLVE_MIXED, fmt, abi.FP64,
                                               args.rtol, args.steps, args.warmup)
    log("mixed path: %.1f ms/step, %d iterations" % (statistics.mean(latm), itsm[0]))
    st = ctx.stats()
    ms_step = statistics.mean(latm)
    value = prob.n_active / (ms_step / 1e3)
    value64 = prob.n_active / (statistics.mean(lat64) / 1e3)

    # dominant kernel: the SpMV of the mixed path, timed with HIP event pairs
    # around each launch over as many further steps (the events add about 8%
    # to a step, so the steps that give `value` run without them)
    spmv_ms, spmv_calls, rep_ms, rep_calls = 0.0, 0, 0.0, 0
    if not args.no_profile:
        _, _, _, _, _, (spmv_ms, spmv_calls, rep_ms, rep_calls) = run_path(
            ctx, prob, abi.SOLVE_MIXED, fmt, abi.FP64, args.rtol, args.steps, 0, profile=True)
    element_ms = statistics.mean(run_path.element_ms) if run_path.element_ms else None
    fp32_integration = fp32_integration_run(ctx, prob, fmt, args)
    n = prob.n_active
    stn = ctx.stats()
    sell_vals = stn["sell_values"]
    # the library falls back to CSR when the SELL padding would not pay, and
    # runs the FP32 sweeps on the symmetric slices (scalar rows, or D x D
    # node blocks of vector problems) where the pattern allows them
    fmt_name = "sell" if (fmt == abi.FMT_SELL and sell_vals) else "csr"
    if fmt_name == "sell" and stn["sell_sym_offsets"]:
        fmt_name = "sym"
        sell_vals = stn["sym_values"]
    alg = spmv_bytes(n, nnz, sell_vals, fmt_name, 4, 4)
    achieved = round(alg / (spmv_ms / 1e3) / 1e9, 1) if spmv_calls else None
    default_n = {"diffusion": 128, "elasticity": 96, "q2": 64}[args.config]
    workload_tag = "%s%d" % (args.config, args.n or default_n)
    traffic = read_traffic(workload_tag, "mixed_" + fmt_name)
    sym_kind = "node-block" if stn["sym_block"] > 1 else "scalar"
    sym_kernel = ("k_spmv_cap<5, float, float, float, 4, double, 3> (stencil-line rows)"
                  if stn["sym_block"] == 1 and stn["sym_lines"] else
                  "k_spmv_sell<float, float, float, 4, double, %d>"
                  % (4 if stn["sym_block"] > 1 else 1))
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic,
            "traffic_source": ("profiles/spmv_counters_%s.md, traffic_bytes_per_launch[mixed_%s]"
                               % (workload_tag, fmt_name)) if traffic else None,
            "kernel": {"sym": "%s on the %s symmetric slices" % (sym_kernel, sym_kind),
                       "sell": "k_spmv_sell<float, float, float, 4>",
                       "csr": "k_spmv_csr<LPR, float, float, float>"}[fmt_name]
                      + " (q = A p with p.q partials)",
            "bytes_per_launch": alg, "avg_launch_ms": round(spmv_ms, 5),
            "launches_timed": spmv_calls,
            "replacement_launch_ms": round(rep_ms, 5), "replacement_launches": rep_calls}

    log("spmv: %.4f ms/launch over %d launches" % (spmv_ms, spmv_calls))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (bounded sample) ...")
        cpu = cpu_baseline(prob, ctx, int(statistics.median(its64)))

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 sweeps / f64 residual", "data": "synthetic",
        "config": {"workload": workload,
                   "n_elements": prob.n_elem, "n_dof": n, "nnz": nnz, "rtol": args.rtol,
                   "format": fmt_name},
        "p50_ms": round(statistics.median(latm), 3),
        "p99_ms": round(float(np.percentile(latm, 99)), 3),
        "iterations": int(statistics.median(itsm)), "relres": max(relm),
        "assemble_ms": round(statistics.mean(asmm), 3), "solve_ms": round(statistics.mean(solm), 3),
        "refinements": st["refinements"],
        "fp64_value": round(value64, 1), "fp64_ms_per_step": round(statistics.mean(lat64), 3),
        "fp64_iterations": int(statistics.median(its64)), "fp64_relres": max(rel64),
        "fp64_assemble_ms": round(statistics.mean(asm64), 3),
        "mixed_over_fp64": round(value / value64, 3),
        "fp64_sell_value": round(prob.n_active / (statistics.mean(lat64s) / 1e3), 1),
        "fp64_sell_ms_per_step": round(statistics.mean(lat64s), 3),
        "fp64_sell_relres": max(rel64s),
        "mixed_over_fp64_sell": round(statistics.mean(lat64s) / ms_step, 3),
        "setup_ms": round(setup_ms, 3), "host_generate_s": round(t_gen, 3),
        "roofline": roof, "cpu_baseline": cpu,
    }
    if fp32_integration is not None:
        out["fp32_integration"] = fp32_integration
    if element_ms:
        out["element_kernel"] = element_roofline(prob, element_ms)
    ctx.close()
    if args.config == "diffusion" and not args.no_extra:
        # the other workloads of the metric, each in its own context
        for key, fn in (("elasticity96", lambda: elasticity_run(96, args.seed, args.rtol)),
                        ("newton64", lambda: newton_run(64)),
                        ("serving", lambda: serving_run(args.serve_requests, 150, args.seed))):
            try:
                out[key] = fn()
            except Exception as e:  # reported in the line, the headline stands
                out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
