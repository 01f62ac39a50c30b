"""Benchmark: batched Levenberg-Marquardt iterations per second at 4096 variants x 48 variables on
the 24-surface NA 0.95 objective (5 fields x 7 wavelengths x 16 rings x 32 spokes = 17920 rays
per variant, hard ray failure), adjoint Jacobian + mixed precision, next to plain FP64
finite-difference (FD) Jacobian LM on the same starts.

A step is one LM cycle over all 4096 runs (src/optim/levenberg_marquardt.cpp:398-572): normal
equations for the runs that need a new Jacobian, the damped step solve, one trial-point MERIT.
Inputs are generated in host memory, uploaded once per backend call by the engine (inputs are
the template system and the run points; the timed region includes that small H2D traffic).

Prints one JSON line: the metric and its roofline and CPU baseline, with the FD, MERIT-rate,
constraint-row, soft-failure-mode and uneven-topology figures beside it. Single GPU only.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fd-steps", type=int, default=None,
                    help="timed FD cycles (default: --steps, after the same --warmup)")
    ap.add_argument("--variants", type=int, default=4096)
    ap.add_argument("--threads", type=int, default=None,
                    help="host threads of the CPU baseline and of the LM's host work (default: "
                         "the process's CPU share, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--convergence-cycles", type=int, default=20,
                    help="LM cycles of the time-to-merit comparison")
    ap.add_argument("--no-uneven", action="store_true",
                    help="leave out the uneven-topology measurement (for short profiled runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with three 1 GiB device copies (scale for rocprofv3 byte counters)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the device LM cycles kernel by kernel instead of replaying the "
                         "captured hipGraph (for profiles: launches then cover the active runs "
                         "only, so per-dispatch counters divide by the grid's runs)")
    return ap.parse_args()


def sync():
    import torch

    torch.cuda.synchronize()


def cpu_share():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS where that
    is set (a GPU box's share of a shared host is set there, whatever the core count shows)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def convergence(eng, base, settings, variables, x0, method, precision, cycles, threads,
                graph=True):
    """Mean merit over the runs after each LM cycle and the device time of the cycles so far
    (each cycle timed alone; the merit download between cycles is not counted)."""
    from mikroptim_amd import abi

    opt = abi.lm_options(max_iterations=cycles, threads=threads, jacobian_method=method,
                         reject_new_failures=True, graph=graph)
    sess = eng.lm(base, settings, variables, x0, opt, precision)
    elapsed, trace = 0.0, []
    for _ in range(cycles):
        t0 = time.perf_counter()
        alive = sess.cycle()
        sync()
        elapsed += time.perf_counter() - t0
        trace.append((elapsed, float(np.mean(sess.results().merit))))
        if not alive:
            break
    sess.close()
    return trace


def time_to_merit(traces):
    """Cycles and seconds each LM needs to bring the mean merit to the larger of their final
    mean merits (a level both reach)."""
    target = max(t[-1][1] for t in traces.values())
    out = {"target_mean_merit": target}
    for name, t in traces.items():
        k = next(i for i, (_, m) in enumerate(t) if m <= target)
        out[name] = {"cycles": k + 1, "seconds": t[k][0], "final_mean_merit": t[-1][1],
                     "mean_merit_per_cycle": [m for _, m in t]}
    return out


def run_lm(eng, base, settings, variables, x0, method, precision, warmup, steps, threads,
           graph=True):
    """Times `steps` LM cycles after `warmup` cycles. Returns the timed cycle count, seconds,
    device time per kernel family, the results and the Jacobians computed in the timed cycles."""
    from mikroptim_amd import abi

    opt = abi.lm_options(max_iterations=warmup + steps + 5, threads=threads,
                         jacobian_method=method, reject_new_failures=True, graph=graph)
    sess = eng.lm(base, settings, variables, x0, opt, precision)
    for _ in range(warmup):
        sess.cycle()
    sync()
    jac0 = int(sum(r.jacobians for r in sess.results().runs))
    eng.kernel_timing(True)
    t0 = time.perf_counter()
    done = 0
    for _ in range(steps):
        if not sess.cycle():
            break
        done += 1
    sync()
    dt = time.perf_counter() - t0
    kt = eng.kernel_timing(False)
    res = sess.results()
    run_lm.graph_cycles = sess.graph_cycles()
    sess.close()
    jacobians = int(sum(r.jacobians for r in res.runs)) - jac0
    return done, dt, kt, res, jacobians


def uneven_groups(eng, topologies=64, per_topology=128, max_iterations=20, threads=16):
    """The uneven workload: 8192 variants of 64 random topologies (8-32 refracting surfaces,
    3-5 fields, 3-7 wavelengths, 10-96 variables drawn (curvatures, thicknesses, working
    distance, then the glass media's Buchdahl coefficients), 6x8 Half pupil; 128 jittered
    variants each), 20 LM iterations (adjoint + mixed, device control), one batched LM per
    topology. Timed with the topologies' cycles run concurrently (optimize_groups) and one
    topology after another."""
    from mikroptim_amd import abi, systems

    groups = []
    for i in range(topologies):
        b, v = systems.uneven_topology(i, variants=per_topology)
        groups.append((b.take([0]), systems.uneven_settings(abi.FAILURE_HARD), v, v.extract(b)))
    opt = abi.lm_options(max_iterations=max_iterations, threads=threads,
                         jacobian_method=abi.JACOBIAN_ADJOINT)
    eng.optimize_groups(groups[:4], opt, precision="mixed")
    sync()
    t0 = time.perf_counter()
    together = eng.optimize_groups(groups, opt, precision="mixed")
    sync()
    t_together = time.perf_counter() - t0
    t0 = time.perf_counter()
    alone = [eng.optimize(*g, opt, precision="mixed") for g in groups]
    sync()
    t_alone = time.perf_counter() - t0
    same = all(np.array_equal(a.x, b.x) for a, b in zip(together, alone))
    # autotuned launch shapes per topology: the first pass times the candidates, the second
    # runs on the cached choices
    previous = eng.tuning()
    t0 = time.perf_counter()
    eng.optimize_groups(groups, opt, precision="mixed", autotune=True)
    sync()
    t_tune_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    tuned = eng.optimize_groups(groups, opt, precision="mixed", autotune=True)
    sync()
    t_tuned = time.perf_counter() - t0
    chosen = {}
    for g in groups:
        key = "pack%d_waves%d" % eng.autotune(*g)
        chosen[key] = chosen.get(key, 0) + 1
    eng.tuning(*previous)
    tuned_iters = int(sum(int(np.sum(r.iterations)) for r in tuned))
    iters = int(sum(int(np.sum(r.iterations)) for r in together))
    shapes = [(g[0].surfaces, g[2].size, g[0].fields, g[0].wavelengths)
              for g in groups]
    return {"variants": topologies * per_topology, "topologies": topologies,
            "max_iterations": max_iterations,
            "surfaces": [min(x[0] for x in shapes), max(x[0] for x in shapes)],
            "variables": [min(x[1] for x in shapes), max(x[1] for x in shapes)],
            "topologies_with_glass_variables": sum(int(g[2].has_glass) for g in groups),
            "fields": [min(x[2] for x in shapes), max(x[2] for x in shapes)],
            "wavelengths": [min(x[3] for x in shapes), max(x[3] for x in shapes)],
            "run_iterations": iters, "seconds_concurrent": t_together,
            "run_iter_per_s": iters / t_together, "seconds_one_after_another": t_alone,
            "concurrency_speedup": t_alone / t_together, "results_equal": same,
            "autotune": {"chosen": chosen, "seconds_first_with_trials": t_tune_first,
                         "seconds_cached": t_tuned, "run_iterations": tuned_iters,
                         "run_iter_per_s": tuned_iters / t_tuned},
            "lm_cycles_max": max(r.cycles for r in together)}


def profiled_traffic(runs_per_launch):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same counters on a 1 GiB copy, as
    bytes per run), times the runs of this launch; None without the summary."""
    import re
    path = os.path.join(ROOT, "profiles", "adjoint_normal_kernel_objective_4096x48.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    f = re.search(r"fetch_bytes_per_run\s*=\s*([0-9.e+]+)", text)
    w = re.search(r"write_bytes_per_run\s*=\s*([0-9.e+]+)", text)
    if not (f and w):
        return None
    return (float(f.group(1)) + float(w.group(1))) * runs_per_launch


def cpu_baseline(batch, variables, settings, x_gpu_final, eng, threads, sample_runs=64,
                 fd_sample_runs=16):
    """Reference CPU LM (the reference's own CpuLmBackend, built from source into oracle/_ref; the
    oracle's restatement if that build is absent) timed on bounded samples of runs for one LM
    cycle each, scaled to the batch's runs: with its adjoint Jacobian (the value) and with its FD
    Jacobian. Also checks the GPU forward trace against the FP64 oracle on the final points of a
    few runs (OPD error in waves)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from mikroptim_amd import abi
    from oracle_lib import oracle, reference

    lib, kind = reference(), "reference"
    if lib is None:
        lib, kind = oracle(), "port"
    base = batch.take([0])
    rates = {}
    for method, runs in ((abi.JACOBIAN_ADJOINT, sample_runs), (abi.JACOBIAN_FD, fd_sample_runs)):
        x0 = variables.extract(batch)[:runs]
        opt = abi.lm_options(max_iterations=1, threads=threads, jacobian_method=method)
        t0 = time.perf_counter()
        lib.optimize_lm(base, settings, variables, x0, opt)
        dt = time.perf_counter() - t0
        rates[method] = (1.0 / (dt * batch.variants / runs), dt)
    # OPD error of the GPU forward trace vs the FP64 oracle at the GPU's final points
    pts = base.take(np.zeros(4, dtype=np.int64))
    variables.apply(pts, x_gpu_final[:4])
    _, opd_g, st_g = eng.residuals(pts, settings)
    _, opd_o, st_o = oracle().residuals(pts, settings, threads=threads)
    lam = batch.reference_wavelength_nm * 1e-6
    ok = (st_g == 0) & (st_o == 0)
    err = float(np.sqrt(np.mean(((opd_g - opd_o)[ok] / lam) ** 2)))
    adj, fd = rates[abi.JACOBIAN_ADJOINT], rates[abi.JACOBIAN_FD]
    return {"value": adj[0], "unit": "LM iter/s @4096 variants", "cores": threads, "kind": kind,
            "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
            "sample": f"{sample_runs} runs x 1 LM cycle (adjoint Jacobian, FP64, {threads} host "
                      f"threads), scaled to {batch.variants} runs",
            "seconds": adj[1],
            "fd_lm_iter_per_s": fd[0],
            "fd_sample": f"{fd_sample_runs} runs x 1 LM cycle (FD Jacobian, FP64, {threads} host "
                         f"threads), scaled to {batch.variants} runs",
            "fd_seconds": fd[1]}, err, bool(np.array_equal(st_g, st_o))


def singlet_rates(eng, threads):
    """Config 1 (the singlet, 1024 variants x 648 rays, FP64 RESIDUALS): GPU trace-kernel rays/s
    and the reference's residuals_cpu (one thread, as the reference runs it) and the oracle's
    restatement on the host threads, in rays/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from mikroptim_amd import abi, systems
    from mikroptim_amd.batch import rays_per_variant
    from oracle_lib import oracle, reference

    b, _ = systems.singlet(1024)
    s = systems.singlet_settings(abi.FAILURE_HARD)
    rays = b.variants * rays_per_variant(b, s)
    eng.residuals(b, s)
    eng.kernel_timing(True)
    eng.residuals(b, s)
    ms, _ = eng.kernel_timing(False)["trace"]
    out = {"workload": "singlet, 1024 variants x 3 fields x 3 wavelengths x 6x12 Half pupil "
                       "(648 rays/variant), FP64 RESIDUALS, hard failure",
           "rays": rays, "gpu_rays_per_s": rays / (ms * 1e-3) if ms > 0 else None}
    ref = reference()
    if ref is not None:
        t0 = time.perf_counter()
        ref.residuals(b, s)
        out["cpu_reference_rays_per_s"] = rays / (time.perf_counter() - t0)
        out["cpu_reference_threads"] = 1
    t0 = time.perf_counter()
    oracle().residuals(b, s, threads=threads)
    out["cpu_port_rays_per_s"] = rays / (time.perf_counter() - t0)
    out["cpu_port_threads"] = threads
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards "
              "by variant ranges with no inter-GPU traffic", file=sys.stderr)
        sys.exit(2)
    from mikroptim_amd import abi, systems
    from mikroptim_amd.engine import TraceEngine

    if args.threads is None:
        args.threads = cpu_share()
    batch, variables = systems.objective(variants=args.variants, seed=2)
    settings = systems.objective_settings(abi.FAILURE_HARD)
    base = batch.take([0])
    x0 = variables.extract(batch)
    eng = TraceEngine(0)

    n_adj, t_adj, kt_adj, res_adj, jac_adj = run_lm(eng, base, settings, variables, x0,
                                           abi.JACOBIAN_ADJOINT, "mixed", args.warmup,
                                           args.steps, args.threads, graph=not args.no_graph)
    it_adj = n_adj / t_adj
    graph_adj = run_lm.graph_cycles
    fd_steps = args.steps if args.fd_steps is None else args.fd_steps
    n_fd, t_fd, kt_fd, res_fd, jac_fd = run_lm(eng, base, settings, variables, x0, abi.JACOBIAN_FD,
                                       "fp64", args.warmup, fd_steps, args.threads)
    it_fd = n_fd / t_fd

    # roofline of the dominant kernel of the measured path, the fused adjoint + normal equations.
    # Algorithmic FLOP (div/sqrt counted as 1, FMA as 2): FLOP_fwd = 59 (S-1) + 60 per ray for a
    # forward trace, FLOP_adj = FLOP_fwd + 210 (S-1) + 6 P per ray for an adjoint Jacobian (its
    # recording forward pass plus the reverse sweep). In the LM the fused kernel starts from the
    # ray states that the MERIT trace of the same point stored, so its work is the reverse part,
    # 210 (S-1) + 6 P per ray, plus the MFMA products (3 x 2 P^2 per ray, bf16). The MERIT trace
    # (FLOP_fwd per ray) is reported beside it.
    S, P = batch.surfaces, variables.size
    rays = 17920
    flop_fwd = 59 * (S - 1) + 60
    flop_rev = 210 * (S - 1) + 6 * P
    adj_ms, adj_n = kt_adj["adjoint"]
    achieved = None
    runs_per_launch = jac_adj / adj_n if adj_n else 0.0
    if adj_n:
        achieved = runs_per_launch * rays * flop_rev / (adj_ms / adj_n * 1e-3) / 1e12
    # bound by FP64 arithmetic on the vector ALU (the MFMA share is 6 P^2 bf16 FLOP per ray, which
    # MI355X's 2.5 PFLOP/s bf16 matrix rate makes negligible); roof: the FP64 vector peak
    roof = {"bound": "valu", "roof": "FP64 vector ALU, 78.6 TFLOP/s",
            "kernel": "adjoint_normal_kernel (fused adjoint + normal equations)",
            "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s",
            "frac": (achieved / 78.6) if achieved else None,
            "traffic": profiled_traffic(runs_per_launch),
            "avg_launch_ms": (adj_ms / adj_n) if adj_n else None,
            "runs_per_launch": runs_per_launch,
            "flop_per_ray": flop_rev,
            "mfma_bf16_flop_per_ray": 6 * P * P}

    # forward trace throughput of the two number types (MERIT of all variants at the starts):
    # FP64 is the trace of every mode; df64 is the reference's second number type
    import copy
    trace_rate = {}
    for name, prec in (("fp64", abi.PRECISION_FP64), ("df64", abi.PRECISION_DF64)):
        st = copy.copy(settings)
        st.precision = prec
        eng.merit(batch, st)
        eng.kernel_timing(True)
        eng.merit(batch, st)
        ms, n = eng.kernel_timing(False)["trace"]
        trace_rate[name] = batch.variants * rays / (ms * 1e-3) if ms > 0 else None
    # FP32 prefilter (float trace with boundary margins) of all variants, against the FP64 MERIT
    eng.prefilter(batch, settings)
    eng.kernel_timing(True)
    eng.prefilter(batch, settings)
    ms, _ = eng.kernel_timing(False)["trace"]
    trace_rate["prefilter_fp32"] = batch.variants * rays / (ms * 1e-3) if ms > 0 else None
    merit_roof = None
    if trace_rate.get("fp64"):
        ach = trace_rate["fp64"] * flop_fwd / 1e12
        merit_roof = {"kernel": "trace_kernel (MERIT, FP64, all 4096 start points)",
                      "achieved": ach, "peak": 78.6, "unit": "TFLOP/s", "frac": ach / 78.6,
                      "flop_per_ray": flop_fwd}

    # constraint rows (paraxial and mechanical, the reference's extra residuals) on the same
    # workload: LM cycles/s with the rows that need no extrema, and with rows from the surface
    # extrema as well (their columns come from central differences of each run's FD batch,
    # traced with extrema, as in the reference); and the MERIT trace rate with extrema
    # with extrema, the differences trace only the candidate rays at the 2P difference points
    # (checked; the same rows as tracing every ray), and, for comparison, every ray
    ext0 = eng.surface_extrema(base, settings)
    rows = {}
    for name, ext, full in (("without_extrema", None, False), ("with_extrema", ext0, False),
                            ("with_extrema_every_tuple", ext0, True)):
        st = copy.copy(settings)
        st.constraints = systems.objective_constraints(base, ext)
        steps = args.steps if ext is None else max(1, min(args.steps, 3 if not full else 1))
        eng.tuning(full_row_traces=full)
        stats0 = eng.row_trace_stats()
        try:
            n_r, t_r, kt_r, res_r, jac_r = run_lm(eng, base, st, variables, x0,
                                                  abi.JACOBIAN_ADJOINT, "mixed", 1, steps,
                                                  args.threads)
        finally:
            eng.tuning(full_row_traces=False)
        stats1 = eng.row_trace_stats()
        rows[name] = {"lm_iter_per_s": n_r / t_r, "steps": n_r, "rows": len(st.constraints),
                      "jacobians": jac_r,
                      "kernel_ms": {k: v[0] for k, v in kt_r.items()}}
        if ext is not None and not full:
            rows[name]["candidate_chunks"] = stats1[0] - stats0[0]
            rows[name]["chunks_traced_whole_after_check"] = stats1[1] - stats0[1]
    if rows["with_extrema_every_tuple"]["lm_iter_per_s"]:
        rows["candidate_speedup"] = (rows["with_extrema"]["lm_iter_per_s"]
                                     / rows["with_extrema_every_tuple"]["lm_iter_per_s"])
    st = copy.copy(settings)
    eng.surface_extrema(batch, st)
    eng.kernel_timing(True)
    eng.surface_extrema(batch, st)
    ms, _ = eng.kernel_timing(False)["trace"]
    rows["merit_trace_with_extrema_rays_per_s"] = batch.variants * rays / (ms * 1e-3) if ms else None

    # per-result error of the mixed-precision normal equations against the FP64 stored-J path
    # (itself within 1e-10 of the FP64 oracle) at 8 final points
    xs = res_adj.x[:8]
    _, _, am, gm = eng.jacobian(base, settings, variables, xs, method="adjoint",
                                precision="mixed", want_jacobian=False)
    _J8, _r8, ad, gd = eng.jacobian(base, settings, variables, xs, method="adjoint",
                                    precision="fp64")
    nJ = np.linalg.norm(_J8.reshape(len(xs), -1), axis=1)
    nr = np.linalg.norm(_r8, axis=1)
    mixed_err = {
        "A_rel_frobenius": float(np.max(np.linalg.norm(am - ad, axis=(1, 2)) /
                                        np.linalg.norm(ad, axis=(1, 2)))),
        "A_diag_rel": float(np.max(np.abs(np.diagonal(am, axis1=1, axis2=2) /
                                          np.diagonal(ad, axis1=1, axis2=2) - 1))),
        "g_rel": float(np.max(np.linalg.norm(gm - gd, axis=1) / np.linalg.norm(gd, axis=1))),
        # against the absolute bounds of the split-bf16 products and the float staging:
        # ||dA||_F <= (3 u_bf^2 + 514 u_f) ||J||_F^2, ||dg|| <= 4 u_f ||J||_F ||r||
        "A_vs_bound": float(np.max(np.linalg.norm(am - ad, axis=(1, 2)) /
                                   ((3 * 2.0 ** -16 + 514 * 2.0 ** -24) * nJ ** 2))),
        "g_vs_bound": float(np.max(np.linalg.norm(gm - gd, axis=1) /
                                   (4 * 2.0 ** -24 * nJ * nr))),
        "runs": 8}

    # config 4: J^T J and J^T r0 on FP64 MFMA and as split-bf16 x3 products (the fused path),
    # each against the FP64 FMA chain (bitwise the host's normal_equations_host) at the same 8
    # final points; and the two Gram kernels' device time on the stored FP64 Jacobians of 256 runs
    _, _, a64m, g64m = eng.jacobian(base, settings, variables, xs, method="adjoint",
                                    precision="fp64_mfma", want_jacobian=False)
    normal_eq = {
        "fp64_mfma": {
            "A_rel_frobenius": float(np.max(np.linalg.norm(a64m - ad, axis=(1, 2)) /
                                            np.linalg.norm(ad, axis=(1, 2)))),
            "g_rel": float(np.max(np.linalg.norm(g64m - gd, axis=1) / np.linalg.norm(gd, axis=1)))},
        "split_bf16x3": {"A_rel_frobenius": mixed_err["A_rel_frobenius"],
                         "g_rel": mixed_err["g_rel"]},
        "reference": "FP64 FMA chain (normal_equations_host order), runs 8"}
    xg = res_adj.x[:256]
    for name, prec in (("chain_fp64", "fp64"), ("mfma_fp64", "fp64_mfma")):
        eng.kernel_timing(True)
        eng.jacobian(base, settings, variables, xg, method="adjoint", precision=prec,
                     want_jacobian=False)
        ms, n = eng.kernel_timing(False)["gram"]
        flop = len(xg) * rays * (P + 1) * (P + 2)
        normal_eq[name + "_ms_256_runs"] = ms
        normal_eq[name + "_tflops"] = flop / (ms * 1e-3) / 1e12 if ms > 0 else None

    uneven = None if args.no_uneven else uneven_groups(eng, threads=args.threads)

    # per-cycle progress of the two LMs on the same starts: cycles and device seconds each needs
    # to reach a mean merit both reach (LM iter/s alone does not show what a cycle buys)
    traces = {name: convergence(eng, base, settings, variables, x0, method, prec,
                                args.convergence_cycles, args.threads, graph=not args.no_graph)
              for name, method, prec in (("adjoint_mixed", abi.JACOBIAN_ADJOINT, "mixed"),
                                         ("fd_fp64", abi.JACOBIAN_FD, "fp64"))}
    to_merit = time_to_merit(traces)
    config1 = singlet_rates(eng, args.threads)

    # soft failure mode (the reference CLI's default) on the same workload: the runs whose point
    # has no failed ray take the fused kernel on the MERIT trace's stored ray states, the others
    # the FP64 path
    st = copy.copy(settings)
    st.failure = abi.FAILURE_SOFT
    n_s, t_s, kt_s, res_s, jac_s = run_lm(eng, base, st, variables, x0, abi.JACOBIAN_ADJOINT,
                                          "mixed", 1, max(1, min(args.steps, 5)), args.threads,
                                          graph=not args.no_graph)
    soft = {"lm_iter_per_s": n_s / t_s, "steps": n_s, "jacobians": jac_s,
            "failed_rays_final": int(sum(r.failed_rays for r in res_s.runs)),
            "kernel_ms": {k: v[0] for k, v in kt_s.items()}}

    out = {
        "metric": "LM iter/s @4096 variants: adjoint+mixed vs FP64 FD; err <=1e-4 waves RMS",
        "value": it_adj, "unit": "LM iter/s", "n_gpus": 1, "steps": n_adj,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_adj / max(1, n_adj),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (seeded 24-surface NA 0.95 objective, N(0,1%) jitter)",
        "config": {"workload": "4096 variants x 48 variables, 24-surface NA 0.95 objective, "
                               "5 fields x 7 wavelengths x 16x32 pupil (17920 rays/variant), "
                               "hard failure", "variants": args.variants, "variables": P,
                   "rays_per_variant": rays},
        "fd_fp64_iter_per_s": it_fd, "fd_steps": n_fd,
        "jacobians_in_timed_cycles": {"adjoint_mixed": jac_adj, "fd_fp64": jac_fd},
        # cycles of the adjoint + mixed LM session that ran as replays of its captured hipGraph
        "lm_graph_cycles_adjoint_mixed": graph_adj,
        "mixed_normal_equations_error_vs_fp64": mixed_err,
        "normal_equations_fp64_mfma_vs_bf16x3": normal_eq,
        "merit_trace_rays_per_s": trace_rate,
        "constraint_rows": rows,
        "uneven_topologies": uneven,
        "time_to_merit": to_merit,
        "singlet_config1": config1,
        "soft_failure_mode": soft,
        "speedup_vs_fd_fp64": it_adj / it_fd if it_fd > 0 else None,
        "kernel_ms_adjoint_path": {k: v[0] for k, v in kt_adj.items()},
        "kernel_ms_fd_path": {k: v[0] for k, v in kt_fd.items()},
        "final_mean_merit": {"adjoint_mixed": float(np.mean(res_adj.merit)),
                             "fd_fp64": float(np.mean(res_fd.merit))},
        "roofline": roof,
        "merit_trace": merit_roof,
    }
    if not args.no_cpu_baseline:
        base_cpu, err, same_status = cpu_baseline(batch, variables, settings, res_adj.x, eng,
                                                  args.threads)
        out["cpu_baseline"] = base_cpu
        out["opd_err_waves_rms"] = err
        out["opd_status_equal"] = same_status
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        import torch

        x = torch.empty(256 * 2 ** 20, dtype=torch.float32, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        sync()


if __name__ == "__main__":
    main()
