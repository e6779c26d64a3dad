"""Benchmark: solve + gradient of topomax's elasticity hot path at N=1000 on one MI355X.

Workload (one "step"): the 3x1 cantilever (Left side fixed, traction (0,-1) on the Right side
centred at 0.5 with length 0.1; E=1, nu=0.3, SIMP p=3, m=1e-6, filter radius 0.05) meshed
with N=1000 cells per unit length (3000 x 1000 cells, 6,000,000 P1 triangles, 3,004,001
vertices, 6,008,002 DOFs); design = Gaussian-smoothed (sigma = 5 cells) seeded uniform noise
mapped to [1e-3, 1].  One step = Helmholtz filter solve -> SIMP elasticity solve to 1e-8
relative residual -> compliance -> sensitivity -> filter adjoint solve -> M w, with every
input resident in HBM.  The headline line is the mixed-precision path (FP64 PCG preconditioned
by an FP32 multigrid V-cycle, for the elasticity and both filter solves); the FP64 path and
FP32 inner solves inside FP64 iterative refinement ("mixed-ir") are timed beside it.

Prints one JSON line (rank 0, N=1).  `value` = GDOF/s = DOFs / (solve+grad seconds) / 1e9.

Beside it: `roofline` (the PCG direction update + FP64 operator, timed with HIP events around
each of its launches inside the timed steps), `kernels` (fine-grid kernels, each launch timed
alone after 512 MiB of other writes), `step_bytes` (the elasticity PCG's algorithmic streams
per step over the step time), `p2` (the reference's P2 element at N=400: FP64, and `mixed` with
the preconditioner in FP32), `batch` (the
64-design workload) and `cpu_baseline` (the CPU oracle at N=400).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the roofline kernel at N=1000, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) of `bench.py --roofline-only`, each scaled by the same counter's
# reading for a plain copy of 48,064,016 bytes (FETCH x1.999, WRITE x1.0).
PMC_TRAFFIC_BYTES = {1000: 308.42e6}
# tmx_bench_kernel id of the roofline kernel: the mixed PCG's direction update + FP64 operator
ROOFLINE_KERNEL = 5


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_steps(problem, rho, steps, warmup, precision, rtol):
    import torch
    for _ in range(warmup):
        problem.objective_and_gradient(rho, rtol=rtol, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    J = None
    for _ in range(steps):
        J, w, mw = problem.objective_and_gradient(rho, rtol=rtol, precision=precision)
        stats.append(dict(problem.last_stats))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, J, w, stats


def bench_kernel(problem, kernel, precision, reps=50):
    import ctypes
    from topomax_amd import _lib
    h = problem._h
    us, nbytes = ctypes.c_double(), ctypes.c_double()
    h.stream()
    prec = {"fp64": _lib.TMX_FP64, "fp32": _lib.TMX_FP32}[precision]
    _lib.check(h.lib.tmx_bench_kernel(h.ptr, kernel, prec, reps, ctypes.byref(us),
                                      ctypes.byref(nbytes)))
    return us.value, nbytes.value


def batch_throughput(n_designs: int, steps: int, rtol: float, merged_only: bool = False):
    """The batched workload: n_designs seeded designs (N in {40, ..., 320}, aspect 1-3, p 1-5,
    1-3 random tractions, optional body force, random fixed sides;
    topomax_amd.workloads.batch_cases), mixed precision with autotuned tiles; one
    objective+gradient each, timed three ways: one design after another, as one tmx_batch call
    of single-design calls on worker threads (mode "alone"), and as one merged batch step
    (mode "merged": one launch per PCG phase and V-cycle stage across all designs)."""
    import torch
    from topomax_amd import ProblemBatch
    from topomax_amd.workloads import batch_cases, problem_from_case, smoothed_density
    probs, rhos = [], []
    ndof = 0
    for k, (case, N, p) in enumerate(batch_cases(n_designs)):
        nx, ny = int(case["width"] * N), N
        gp = problem_from_case(case, nx, ny, precision="mixed", rtol=rtol)
        gp.set_penalization(p)
        gp.autotune()
        probs.append(gp)
        rhos.append(torch.as_tensor(smoothed_density(nx, ny, seed=k), device="cuda:0"))
        ndof += 2 * (nx + 1) * (ny + 1)
    alone = ProblemBatch(probs, mode="alone")  # default workers: min(16, cores)
    merged = ProblemBatch(probs, mode="merged")
    for gp, r in zip(probs, rhos):
        gp.objective_and_gradient(r)
    if not merged_only:
        alone.objective_and_gradient(rhos)
    merged.objective_and_gradient(rhos)  # (records and merges the launch plan)
    torch.cuda.synchronize()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_merged = timed(lambda: merged.objective_and_gradient(rhos))
    t_serial = t_alone = None
    if not merged_only:
        t_serial = timed(lambda: [gp.objective_and_gradient(r) for gp, r in zip(probs, rhos)])
        t_alone = timed(lambda: alone.objective_and_gradient(rhos))
    rep = merged.last_report
    its = [st["iterations"] for st in merged.last_stats]
    return {"designs": n_designs, "workload": "seeded random designs, N 40-320, aspect 1-3, "
                                              "p 1-5, mixed precision, autotuned tiles",
            "dofs": ndof, "ms_serial": t_serial and t_serial * 1e3,
            "ms_alone": t_alone and t_alone * 1e3, "ms_batched": t_merged * 1e3,
            "GDOF_s_serial": t_serial and ndof / t_serial / 1e9,
            "GDOF_s_alone": t_alone and ndof / t_alone / 1e9,
            "GDOF_s_batched": ndof / t_merged / 1e9,
            "speedup": t_serial and t_serial / t_merged,
            "speedup_over_alone": t_alone and t_alone / t_merged,
            "merged_designs": rep["merged_designs"], "launches_per_step": rep["launches"],
            "merged_iterations": rep["iterations"],
            "elasticity_iterations": {"min": min(its), "mean": sum(its) / len(its),
                                      "max": max(its)},
            "mode": "merged (one launch per PCG phase and V-cycle stage across the designs); "
                    "alone: worker threads, min(16, cores), one single-design call each"}


def instep_kernel_us(problem, rho, steps, precision, rtol, kernel=ROOFLINE_KERNEL):
    """Average duration of `kernel`'s launches inside `steps` full steps, from HIP events the
    library records around each launch on the handle's stream."""
    import ctypes
    from topomax_amd import _lib
    h = problem._h
    _lib.check(h.lib.tmx_kernel_timing(h.ptr, kernel, 1))
    total, count = ctypes.c_double(), ctypes.c_int()
    try:
        _lib.check(h.lib.tmx_kernel_timing_read(h.ptr, kernel, ctypes.byref(total),
                                                ctypes.byref(count)))
        for _ in range(steps):
            problem.objective_and_gradient(rho, rtol=rtol, precision=precision)
        _lib.check(h.lib.tmx_kernel_timing_read(h.ptr, kernel, ctypes.byref(total),
                                                ctypes.byref(count)))
    finally:
        _lib.check(h.lib.tmx_kernel_timing(h.ptr, kernel, 0))
    return total.value / max(1, count.value), count.value


def p2_throughput(N: int, steps: int, rtol: float, mixed: bool = True):
    """The reference's own element (P2 displacement, FEM_src/elasisity_problem.py:194-196) on
    the same cantilever at N: FP64 MINRES preconditioned by P2 block Jacobi over the P1
    multigrid, one objective + gradient per step; with mixed, also the same solve with the
    preconditioner's P1 V-cycle in FP32 (precision "mixed")."""
    import torch
    from topomax_amd.workloads import CANTILEVER, problem_from_case, smoothed_density
    nx, ny = 3 * N, N
    rho = torch.as_tensor(smoothed_density(nx, ny, seed=0, sigma=5.0), device="cuda:0")

    def run(precision):
        gp = problem_from_case(CANTILEVER, nx, ny, precision=precision, rtol=rtol, degree=2)
        gp.set_penalization(3)
        gp.objective_and_gradient(rho)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            J, w, _ = gp.objective_and_gradient(rho)
        torch.cuda.synchronize()
        return gp, J, w, (time.perf_counter() - t0) / steps

    gp, J, w, dt = run("fp64")
    mixed_part = None
    if mixed:
        gm, Jm, wm, dtm = run("mixed")
        mixed_part = {"ms_per_step": dtm * 1e3, "iterations": gm.last_stats["iterations"],
                      "compliance_rel_err_vs_fp64": abs(Jm - J) / abs(J),
                      "gradient_rel_err_vs_fp64":
                          float(torch.linalg.norm(wm - w) / torch.linalg.norm(w)),
                      "preconditioner": "two-level cycle in FP32 (P2 sweeps, transfers, P1 V-cycle)"}
    ndof = 2 * (2 * nx + 1) * (2 * ny + 1)
    st = gp.last_stats
    return {"workload": f"cantilever 3x1, N={N}, P2 displacement ({ndof} DOFs), p=3, FP64, "
                        f"rtol {rtol:g}", "dofs": ndof, "ms_per_step": dt * 1e3,
            "GDOF_s": ndof / dt / 1e9, "iterations": st["iterations"],
            "rel_residual": st["rel_residual"], "compliance": J, "mixed": mixed_part}


def cpu_baseline(n_cpu: int):
    """The CPU oracle (numpy/scipy, SuperLU direct solves in nested-dissection order, single
    thread) on the same workload at N=n_cpu: one objective + gradient (filter solve, elasticity
    solve, compliance, sensitivity, filter adjoint solve, M w) of the same cantilever; the
    assembly of the filter and its factorisation (done once, as the oracle caches it) are not
    timed, so the CPU figure is favourable to the CPU."""
    from oracle.fem import ElasticityOracle, StructuredMesh
    from topomax_amd.workloads import CANTILEVER as C, smoothed_density
    nx, ny = 3 * n_cpu, n_cpu
    ElasticityOracle.permc_spec = "ND"
    orc = ElasticityOracle(StructuredMesh(C["width"], C["height"], nx, ny),
                           young_modulus=C["young_modulus"], poisson_ratio=C["poisson_ratio"],
                           filter_radius=C["filter_radius"], fixed_sides=C["fixed_sides"],
                           tractions=C["tractions"], body_force=None)
    orc.set_penalization(3)
    rho = smoothed_density(nx, ny, seed=0, sigma=5.0)
    t0 = time.perf_counter()
    orc.calculate_objective(rho)
    w = orc.calculate_objective_gradient()
    _ = orc.mass @ w
    dt = time.perf_counter() - t0
    ndof = 2 * (nx + 1) * (ny + 1)
    return {
        "value": ndof / dt / 1e9,
        "unit": "GDOF/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"CPU oracle (numpy/scipy SuperLU in nested-dissection order, "
                   f"single-threaded, not FEniCS) P1 FP64 filter+solve+compliance+gradient of "
                   f"the same cantilever at N={n_cpu} ({ndof} DOFs; the N=400 configuration of "
                   f"the workload): {dt:.2f} s per step on {cpu_model()} "
                   f"({len(os.sched_getaffinity(0))} cores visible)"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=1000)
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--cpu-N", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-p2", action="store_true")
    ap.add_argument("--p2-N", type=int, default=400)
    ap.add_argument("--p2-only", action="store_true",
                    help="time only the P2 element's steps (for a kernel trace of them)")
    ap.add_argument("--batch-designs", type=int, default=64)
    ap.add_argument("--batch-only", action="store_true",
                    help="time only the batched workload (for a kernel trace of it)")
    ap.add_argument("--batch-merged-only", action="store_true",
                    help="with --batch-only: only the merged batch step")
    ap.add_argument("--roofline-only", action="store_true",
                    help="set up the N operator, then launch only the roofline kernel (PCG "
                         "direction update + FP64 operator, 20x) and a plain copy of the 2 n_v "
                         "load vector "
                         "(calibration for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent designs would shard across GPUs with zero communication)",
              file=sys.stderr)
        sys.exit(2)

    if args.p2_only:
        print(json.dumps({"p2": p2_throughput(args.p2_N, args.steps, args.rtol)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_throughput(args.batch_designs, args.steps, args.rtol,
                                                    args.batch_merged_only)}), flush=True)
        return

    import torch
    from topomax_amd.workloads import CANTILEVER, problem_from_case, smoothed_density

    N = args.N
    nx, ny = 3 * N, N
    problem = problem_from_case(CANTILEVER, nx, ny, precision="mixed", rtol=args.rtol)
    problem.set_penalization(args.p)
    rho = torch.as_tensor(smoothed_density(nx, ny, seed=0, sigma=5.0), device="cuda:0")
    ndof = 2 * (nx + 1) * (ny + 1)
    if args.roofline_only:
        problem.autotune()  # the same tiles as the timed run
        problem.objective_and_gradient(rho)
        torch.cuda.synchronize()
        us, nbytes = bench_kernel(problem, ROOFLINE_KERNEL, "fp64", reps=20)
        F = problem.load_vector()  # one k_copy<double> of 2 n_v doubles
        torch.cuda.synchronize()
        print(json.dumps({"roofline_kernel_us": us, "algorithmic_bytes": nbytes,
                          "copy_bytes_each_way": F.numel() * 8}), flush=True)
        return
    tune = problem.autotune()

    t_mixed, J_mixed, w_mixed, st_mixed = time_steps(problem, rho, args.steps, args.warmup,
                                                     "mixed", args.rtol)
    # the roofline kernel's launches inside the timed steps (events around each launch)
    au_us, au_launches = instep_kernel_us(problem, rho, args.steps, "mixed", args.rtol)
    # fine-grid kernels, each launch alone after 512 MiB of other writes (cold caches)
    kernels = {}
    problem.set_option("bench_cold", 1)
    for name, kid, prec in (("apply_fp32", 0, "fp32"), ("smooth_fp32", 1, "fp32"),
                            ("apply_fp64", 0, "fp64"), ("smooth_fp64", 1, "fp64"),
                            ("sweep_pair_restrict_fp32", 6, "fp32"),
                            ("sweep_pair_prolong_fp32", 4, "fp32"),
                            ("update_apply_fp64", ROOFLINE_KERNEL, "fp64")):
        us, nbytes = bench_kernel(problem, kid, prec, reps=20)
        kernels[name] = {"us": us, "bytes": nbytes, "GBs": nbytes / us / 1e3}
    problem.set_option("bench_cold", 0)
    b2b_us, _ = bench_kernel(problem, ROOFLINE_KERNEL, "fp64")

    result_fp64 = result_ir = result_ru = None
    if not args.no_fp64:
        t64, J64, w64, st64 = time_steps(problem, rho, max(2, args.steps // 2), 1, "fp64",
                                         args.rtol)
        dw = float(torch.linalg.norm(w_mixed - w64) / torch.linalg.norm(w64))
        result_fp64 = {"ms_per_step": t64 * 1e3, "GDOF_s": ndof / t64 / 1e9,
                       "iterations": st64[-1]["iterations"],
                       "compliance_rel_err_mixed_vs_fp64": abs(J_mixed - J64) / abs(J64),
                       "gradient_rel_err_mixed_vs_fp64": dw,
                       "speedup_mixed_over_fp64": t64 / t_mixed,
                       "filter_preconditioner": "DCT-I / tridiagonal partition solver (FP64)",
                       "mixed_filter_preconditioner": "DCT-I / tridiagonal partition solver (FP32)"}
        tir, Jir, wir, stir = time_steps(problem, rho, max(2, args.steps // 2), 1, "mixed-ir",
                                         args.rtol)
        result_ir = {"ms_per_step": tir * 1e3, "GDOF_s": ndof / tir / 1e9,
                     "iterations": stir[-1]["iterations"],
                     "outer_iterations": stir[-1]["outer_iterations"],
                     "compliance_rel_err_vs_fp64": abs(Jir - J64) / abs(J64),
                     "gradient_rel_err_vs_fp64":
                         float(torch.linalg.norm(wir - w64) / torch.linalg.norm(w64))}
        tru, Jru, wru, stru = time_steps(problem, rho, max(2, args.steps // 2), 1, "mixed-ru",
                                         args.rtol)
        result_ru = {"ms_per_step": tru * 1e3, "GDOF_s": ndof / tru / 1e9,
                     "iterations": stru[-1]["iterations"],
                     "outer_iterations": stru[-1]["outer_iterations"],
                     "rel_residual": stru[-1]["rel_residual"],
                     "speedup_over_fp64": t64 / tru,
                     "compliance_rel_err_vs_fp64": abs(Jru - J64) / abs(J64),
                     "gradient_rel_err_vs_fp64":
                         float(torch.linalg.norm(wru - w64) / torch.linalg.norm(w64))}

    # the largest elasticity kernel of the mixed step: the PCG direction update fused with the
    # FP64 operator application (one launch per iteration), timed inside the steps
    au_bytes = kernels["update_apply_fp64"]["bytes"]
    achieved = au_bytes / au_us / 1e3
    roofline = {"kernel": "k_elast_sweep<double,true,kApplyUpdate> (PCG direction update "
                          "p = z + beta p_old with q = K(xi) p in FP64, p.q partials and the "
                          "deferred x += alpha p_old)",
                "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": PMC_TRAFFIC_BYTES.get(N),
                "algorithmic_bytes_per_launch": au_bytes, "avg_launch_us": au_us,
                "timing": f"HIP events around each of its {au_launches} launches inside "
                          f"{args.steps} timed mixed steps",
                "back_to_back_us": b2b_us,
                "cold_us": kernels["update_apply_fp64"]["us"]}
    # whole step: the elasticity PCG's algorithmic streams per iteration (direction update +
    # operator 96 B, x/r update 56 B, the finest V-cycle level's two paired launches 26 + 34 B
    # per vertex) over the step time; the filter solves and the coarse levels are left out, so
    # this is a lower bound of the step's algorithmic bytes
    nv = (nx + 1) * (ny + 1)
    its = st_mixed[-1]["iterations"]
    step_bytes = its * (96 + 56 + 26 + 34) * nv
    step = {"bytes": step_bytes, "GBs": step_bytes / t_mixed / 1e9,
            "frac": step_bytes / t_mixed / 1e9 / PEAK_HBM_GBS,
            "streams": f"{its} PCG iterations x 212 B per vertex (direction update + operator, "
                       "x/r update, finest-level paired V-cycle launches)"}

    p2 = None
    if not args.no_p2:
        p2 = p2_throughput(args.p2_N, max(2, args.steps // 5), args.rtol)

    batch = None
    if not args.no_batch:
        batch = batch_throughput(args.batch_designs, max(2, args.steps // 2), args.rtol)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_N)

    line = {
        "metric": "GDOF/s & solve+grad time at N=1000 to 1e-8 rel. residual, FP64 vs mixed",
        "value": ndof / t_mixed / 1e9,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 PCG + f32 multigrid (mixed)",
        "data": "synthetic (seeded smoothed random density)",
        "config": {"workload": f"cantilever 3x1, N={N} ({nx}x{ny} cells, {ndof} P1 DOFs), "
                               f"p={args.p}, filter->solve->compliance->gradient, "
                               f"rtol {args.rtol:g}",
                   "N": N, "dofs": ndof, "p": args.p, "rtol": args.rtol,
                   "precision": "mixed", "tiles": tune},
        "iterations": st_mixed[-1]["iterations"],
        "outer_iterations": st_mixed[-1]["outer_iterations"],
        "rel_residual": st_mixed[-1]["rel_residual"],
        "compliance": J_mixed,
        "fp64": result_fp64,
        "mixed_ir": result_ir,
        "mixed_ru": result_ru,
        "batch": batch,
        "p2": p2,
        "step_bytes": step,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
