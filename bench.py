"""Benchmark: DoF-updates/s of 3D Euler ADER-DG at order 5 on one MI355X.

Workload: 81^3 periodic cells on the unit cube, order N=5 (6^3 Gauss-Legendre
nodes per cell, 5 variables), isentropic vortex plus 8 seeded density modes,
CFL 0.9; one step = predictor (Picard + time averages + face extrapolation +
volume integral), Rusanov on every face, surface integral + update + global
CFL minimum, with inputs resident in HBM.

Prints ONE JSON line (rank 0, N=1).  Fields beyond the driver contract:
  roofline     dominant kernel (the predictor): algorithmic FLOP per launch
               divided by its HIP-event-timed average launch duration
  cpu_baseline the C oracle (OpenMP) timed on this host on a bounded sample
               of the same workload
  precisions   per-precision throughput and the L2 difference of the mixed
               path against the FP64 path after the timed steps
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DoF-updates/s, 3D Euler ADER-DG N=5, mixed vs FP64 GPU at L2 err ≤1e-6"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBPS = 8000.0
# mixed-precision variants measured: (name, precision, FP64 correction sweeps)
MIXED_VARIANTS = [("fp32", 1, 1), ("fp32_mfma", 2, 1), ("bf16_mfma", 3, 2), ("bf16_mfma_3", 3, 3)]
PHI_FLUX = 40   # flops of one 3D Euler flux evaluation (1 division counted as 1)
NVAR = 5


def predictor_flops_per_cell(N, iters):
    """Algorithmic FLOP of one predictor launch per cell: I Picard sweeps of
    8 V B^5 contraction flops + B^4 (flux + 5V), then time averages and the
    volume integral (14 V B^4) and the face extrapolation (24 V B^3)."""
    B = N + 1
    return iters * (8 * NVAR * B ** 5 + B ** 4 * (PHI_FLUX + 5 * NVAR)) + 14 * NVAR * B ** 4 + 24 * NVAR * B ** 3


def measured_traffic(kernel, n, N):
    """HBM bytes per launch of `kernel` from the counter runs summarised in
    profiles/traffic.json (FETCH_SIZE and WRITE_SIZE, each scaled by what the
    counter reads for the streaming copy kernel), or None if that kernel and
    workload were not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("cells") != n ** 3 or rec.get("order") != N:
        return None
    return rec["bytes_per_launch"]


ERROR_STEPS = 50  # the mixed-precision error is stated after this many steps from the initial state


def state_after(order, n, nsteps, precision, waves_per_simd=0, contraction=0, fp64_sweeps=1):
    """The grid state after nsteps steps from the headline initial state
    (graphed steps, untimed): the input of the mixed-precision error norms."""
    import exahype_amd as E
    s = E.ADERDGSolver(order, precision=precision, fp64_sweeps=fp64_sweeps, waves_per_simd=waves_per_simd,
                       contraction=contraction)
    g = E.Grid(s, n, n, n)
    g.set_initial_conditions(E.make_ic(0, **vortex_ic_kwargs()))
    g.compute_dt()
    g.use_graph(True)
    g.step(nsteps)
    g.sync()
    out = g.download()
    g.close()
    s.close()
    return out


def vortex_ic_kwargs(seed=0):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 2 * np.pi, 8))


def run_gpu(order, n, steps, warmup, precision, waves_per_simd=0, contraction=0, fp64_sweeps=1, keep_state=True,
            timing_steps=5):
    """Times `steps` steps of the device grid as the product runs them: one
    captured hipGraph per step, no per-kernel events (the wall clock of the
    steady-state loop).  Then runs `timing_steps` more steps with HIP events
    around every launch (graph off) for the per-phase kernel times."""
    import exahype_amd as E
    s = E.ADERDGSolver(order, precision=precision, fp64_sweeps=fp64_sweeps, waves_per_simd=waves_per_simd,
                       contraction=contraction)
    g = E.Grid(s, n, n, n)
    g.set_initial_conditions(E.make_ic(0, **vortex_ic_kwargs()))
    g.compute_dt()
    g.use_graph(True)
    g.step(warmup)
    g.sync()
    g.total_picard_iterations()  # reset the device sweep counter
    E.device_synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    E.device_synchronize()
    wall = time.perf_counter() - t0
    sweeps = g.total_picard_iterations() / (g.ncells * steps)
    state = g.download() if keep_state else None
    g.kernel_times()  # discard
    g.enable_timing(True)
    g.step(timing_steps)
    g.sync()
    times, counts = g.kernel_times()
    g.close()
    s.close()
    return dict(wall=wall, times=times, counts=counts, iters=sweeps, state=state)


def error_norms(a, b, N, ncells):
    """Domain L1, L2 (GL quadrature per cell, unit cube) and max norms of a-b
    per variable."""
    from exahype_amd.tables import dg_math
    w = dg_math.gl_nodes_weights(N)[1]
    w3 = np.einsum("z,y,x->zyx", w, w, w).ravel()
    d = (a - b).reshape(ncells, (N + 1) ** 3, NVAR)
    return dict(l1=np.einsum("n,cnv->v", w3, np.abs(d)) / ncells,
                l2=np.sqrt(np.einsum("n,cnv->v", w3, d ** 2) / ncells),
                linf=np.abs(d).max(axis=(0, 1)))


def l2_per_variable(a, b, N, ncells):
    return error_norms(a, b, N, ncells)["l2"]


def host_info():
    """CPU model, logical CPUs usable by this process and the oracle's compiler."""
    import platform
    import subprocess
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.check_output(["gcc", "--version"], text=True).splitlines()[0]
    except (OSError, subprocess.CalledProcessError):
        cc = "unknown"
    return dict(cpu_model=model, affinity_cpus=len(os.sched_getaffinity(0)), os_cpu_count=os.cpu_count(),
                cgroup_cpu_quota=cgroup_cpus(), omp_num_threads=os.environ.get("OMP_NUM_THREADS"), compiler=cc,
                oracle_flags="-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (built in the build container; native would target its CPU, not this host)")


def cgroup_cpus():
    """CPUs of this process's cgroup quota (cpu.max), or None without one."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        return None if quota == "max" else float(quota) / float(period)
    except (OSError, ValueError):
        return None


def host_threads():
    """The oracle's OpenMP team: all the CPUs this process may use -- the
    affinity mask, capped by the cgroup CPU quota and by OMP_NUM_THREADS
    where the host sets them (threads beyond the quota only time-slice it)."""
    n = len(os.sched_getaffinity(0))
    q = cgroup_cpus()
    if q:
        n = min(n, max(1, int(q)))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(order, n_full, threads, sample_cells=27, steps=1):
    """Oracle (the C restatement, OpenMP) on a sample_cells^3 sub-box with the
    same cell size, order and initial data; returns DoF-updates/s."""
    from oracle import oracle as O
    c = O.cfg(order)
    og = O.OracleGrid(c, sample_cells, sample_cells, sample_cells)
    for d in range(3):
        og.g.width[d] = sample_cells / n_full
    og.init(O.make_ic(0, **vortex_ic_kwargs()))
    dt = og.stable_dt(threads)
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = og.step(dt, threads)
    wall = time.perf_counter() - t0
    B = order + 1
    return og.ncells * NVAR * B ** 3 * steps / wall, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--order", type=int, default=5)
    ap.add_argument("--cells", type=int, default=81)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--waves-per-simd", type=int, default=0, help="predictor occupancy target (0 = tuned)")
    ap.add_argument("--contraction", type=int, default=0,
                    help="predictor sweep contraction: 0 tuned, 1 pencil owners, 2 node owners, "
                         "3 pencil owners two slices at a time, 4 staged iterate (FP64, orders 5..9)")
    ap.add_argument("--workload", default="aderdg", choices=["aderdg", "fv2d", "ensemble", "limiter", "muscl", "amr",
                                                          "ncp", "linear", "amr_limiter", "parity81"],
                    help="aderdg: the headline ADER-DG step; fv2d: the EulerFV demonstrator "
                         "(2D Godunov, 81x81 patches of 10x10 volumes, 200 steps); ensemble: independent "
                         "grids of random order/size/gamma advanced together; limiter: troubled-cell detection and "
                         "subcell projection on the Sod sphere (27^3 cells, N=5); muscl: 3D MUSCL-Hancock FV on "
                         "27^3 patches of 11^3 subcells (the limiter's FV scheme of the N=5 Sod sphere); amr: batched "
                         "volume prolongation/restriction of 27^3 N=5 cells one level down and back; amr_limiter: "
                         "limiting ADER-DG (N=5) on a 3-level adaptive mesh refined in a spherical shell, Sod sphere, "
                         "100 steps")
    ap.add_argument("--instances", type=int, default=512, help="ensemble members (--workload ensemble)")
    ap.add_argument("--point-source", action="store_true",
                    help="--workload linear: with a Ricker point source at the domain centre (evaluated per node "
                         "in every cell's predictor)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="also time a plain device copy of this many GiB (HBM reference, counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the path runs on a single MI355X in this build "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "fv2d":
        print(json.dumps(run_fv2d(args)))
        return
    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args)))
        return
    if args.workload == "limiter":
        print(json.dumps(run_limiter(args)))
        return
    if args.workload == "muscl":
        print(json.dumps(run_muscl(args)))
        return
    if args.workload == "amr":
        print(json.dumps(run_amr(args)))
        return
    if args.workload == "ncp":
        print(json.dumps(run_ncp(args)))
        return
    if args.workload == "linear":
        print(json.dumps(run_linear(args)))
        return
    if args.workload == "amr_limiter":
        print(json.dumps(run_amr_limiter(args)))
        return
    if args.workload == "parity81":
        print(json.dumps(run_parity81(args)))
        return
    import exahype_amd as E
    N, n = args.order, args.cells
    ncells = n ** 3
    dof = ncells * NVAR * (N + 1) ** 3

    # the FP64 state is the reference of the mixed runs (the FP64 path equals
    # the oracle to 1e-12); errors after ERROR_STEPS steps from the initial state
    r64 = run_gpu(N, n, args.steps, args.warmup, E.PREC_FP64, args.waves_per_simd, args.contraction,
                  keep_state=False)
    rate64 = dof * args.steps / r64["wall"]
    ref50 = state_after(N, n, ERROR_STEPS, E.PREC_FP64, args.waves_per_simd, args.contraction)
    mixed = {}
    for name, prec, sweeps in MIXED_VARIANTS:
        if N > 7 and prec != E.PREC_FP32:  # orders 8 and 9: FP32 sweeps on the staged predictor only
            continue
        # the staged form takes FP64 and FP32 sweeps; the matrix-core forms their tuned form
        cm = 0 if (args.contraction == 4 and prec != E.PREC_FP32) else args.contraction
        r = run_gpu(N, n, args.steps, args.warmup, prec, args.waves_per_simd, cm, sweeps, keep_state=False)
        err = error_norms(state_after(N, n, ERROR_STEPS, prec, args.waves_per_simd, cm, sweeps), ref50, N, ncells)
        l2 = err["l2"]
        r.update(rate=dof * args.steps / r["wall"], l2=l2, err=err, ok=bool(l2.max() <= 1e-6), name=name,
                 sweeps_fp64=sweeps, prec=prec)
        del r["state"]
        mixed[name] = r
    passing = [r for r in mixed.values() if r["ok"]]
    best = max(passing, key=lambda r: r["rate"]) if passing else None
    mixed_ok = best is not None
    head = best if mixed_ok else r64
    rate = best["rate"] if mixed_ok else rate64

    B = N + 1
    pred_ms = head["times"][0] / max(head["counts"][0], 1)
    # the fused Riemann + surface integral + update + CFL kernel: reads the
    # folded predictor output and the cell's own face states and fluxes,
    # writes the solution (the neighbours' face blocks are the other cells'
    # own, counted once)
    tail_ms = head["times"][2] / max(head["counts"][2], 1)
    tail_bytes = ncells * 8 * 5 * (2 * (N + 1) ** 3 + 12 * (N + 1) ** 2)
    sweeps = head["iters"]
    if mixed_ok:
        # low-precision sweeps at their peak (FP32 vector ALU 157.3 TF/s with
        # packed math, BF16 MFMA 2.5 PF/s dense), the FP64 correction sweeps
        # and the FP64 tail at the FP64 peak: peak = total flop / minimum time
        f_sweep = 8 * NVAR * B ** 5 + B ** 4 * (PHI_FLUX + 5 * NVAR)
        n_lp = max(sweeps - best["sweeps_fp64"], 0.0)
        f_lp = n_lp * f_sweep * ncells
        f64 = (predictor_flops_per_cell(N, sweeps) * ncells) - f_lp
        flops = f_lp + f64
        peak_lp = PEAK_BF16_TFLOPS if best["prec"] == E.PREC_BF16_MFMA else PEAK_FP32_TFLOPS
        peak = flops / (f_lp / peak_lp + f64 / PEAK_FP64_TFLOPS)
        kname = "k_predictor<%d, %s + %d FP64 sweep(s)>" % (N, best["name"], best["sweeps_fp64"])
        # what bounds each part: FP32 sweeps on the vector ALU (no packed
        # FP32), MFMA forms on the matrix cores, FP64 sweeps and tail on the
        # FP64 vector ALU
        bound = {E.PREC_FP32: "fp32_valu+fp64_valu", E.PREC_FP32_MFMA: "mfma_fp32+fp64_valu",
                 E.PREC_BF16_MFMA: "mfma_bf16+fp64_valu"}[best["prec"]]
    else:
        flops = predictor_flops_per_cell(N, sweeps) * ncells
        peak = PEAK_FP64_TFLOPS
        kname = "k_predictor<%d, FP64>" % N
        bound = "fp64_valu"
    achieved = flops / (pred_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC if N == 5 else METRIC.replace("N=5", "N=%d" % N),
        "value": rate,
        "unit": "DoF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (("bf16+f32+f64" if best["prec"] == E.PREC_BF16_MFMA else "f32+f64") if mixed_ok else "f64"),
        "data": "synthetic",
        "config": {"workload": "3D Euler ADER-DG, order %d, %d^3 periodic cells, vortex + 8 density modes, CFL 0.9"
                               % (N, n), "cells": ncells, "order": N, "dof_per_step": dof,
                   "predictor": ("mixed: %s Picard sweeps + %d FP64 correction sweep(s)"
                                 % (best["name"], best["sweeps_fp64"]) if mixed_ok else "fp64"),
                   "waves_per_simd": args.waves_per_simd or "tuned", "contraction": args.contraction or "tuned",
                   "parallelism": "single GPU"},
        "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(kname, n, N),
                     "kernel": kname, "avg_launch_ms": pred_ms,
                     "picard_sweeps_per_cell_step": sweeps, "flop_per_launch": flops},
        "phase_ms_per_step": {"predictor": pred_ms,
                              "boundary_faces": head["times"][1] / max(head["counts"][1], 1),
                              "riemann_surface_update_dt": tail_ms},
        "tail_roofline": {"bound": "hbm", "achieved": tail_bytes / tail_ms * 1e-6, "peak": PEAK_HBM_GBPS,
                          "unit": "GB/s", "frac": tail_bytes / tail_ms * 1e-6 / PEAK_HBM_GBPS,
                          "traffic": measured_traffic("k_riemann_update<%d>" % N, n, N),
                          "kernel": "k_riemann_update<%d>" % N, "avg_launch_ms": tail_ms,
                          "bytes_per_launch": tail_bytes,
                          "bytes_with_neighbour_face_reads": tail_bytes + ncells * 8 * 5 * 12 * (N + 1) ** 2},
        "precisions": dict(
            fp64={"dof_updates_per_s": rate64, "ms_per_step": r64["wall"] / args.steps * 1e3,
                  "predictor_ms": r64["times"][0] / max(r64["counts"][0], 1), "sweeps_per_cell": r64["iters"]},
            **{k: {"dof_updates_per_s": r["rate"], "ms_per_step": r["wall"] / args.steps * 1e3,
                   "predictor_ms": r["times"][0] / max(r["counts"][0], 1), "sweeps_per_cell": r["iters"],
                   "fp64_correction_sweeps": r["sweeps_fp64"], "l2_vs_fp64_per_variable": r["l2"].tolist(),
                   "l1_vs_fp64_per_variable": r["err"]["l1"].tolist(),
                   "linf_vs_fp64_per_variable": r["err"]["linf"].tolist(),
                   "speedup_vs_fp64": r["rate"] / rate64, "l2_ok": r["ok"]} for k, r in mixed.items()}),
        "timing": "value and ms_per_step: wall clock of the steps replayed as one hipGraph per step, no events; "
                  "phase_ms_per_step and roofline: HIP events around every launch of %d further steps "
                  "(graph off)" % 5,
        "error_steps": ERROR_STEPS,
        "errors": "L1/L2/Linf per variable of each precision's state against the FP64 path's after error_steps "
                  "steps from the same initial state (the FP64 path equals the oracle to 1e-12)",
    }
    if not args.no_cpu_baseline:
        threads = host_threads()
        # bounded CPU sample: the work per cell-step grows ~B^5, so the step
        # counts shrink with the order (30 and 20 steps at N=5)
        scale = (6.0 / (N + 1)) ** 5
        st, st1 = max(1, min(30, int(round(30 * scale)))), max(1, min(20, int(round(20 * scale))))
        v, wall = cpu_baseline(N, n, threads, sample_cells=27, steps=st)
        v1, wall1 = cpu_baseline(N, n, 1, sample_cells=9, steps=st1)
        out["cpu_baseline"] = {"value": v, "unit": "DoF-updates/s", "cores": threads, "kind": "port",
                               "host": host_info(),
                               "sample": "oracle (C, OpenMP), order %d, 27^3 periodic cells of size 1/%d, same initial "
                                         "data, %d steps (%.1f s wall)" % (N, n, st, wall),
                               "single_thread": {"value": v1, "sample": "9^3 cells, %d steps (%.1f s)" % (st1, wall1)}}
    if args.copy_gib > 0:
        out["hbm_copy"] = copy_reference(args.copy_gib)
    print(json.dumps(out))


NCP_VELOCITY = (0.3, -0.7, 0.5)
PHI_NCP = 6 * NVAR  # a . grad Q: three multiply-adds per variable


def advection_ic_kwargs():
    return dict(nmodes=2, k=[(1, 0, 1), (0, 1, 1)], a=[0.2, 0.1], phi=[0.3, 1.1], uinf=NCP_VELOCITY)


# A Ricker-wavelet point force at the domain centre (inside cell 40 of 81 per
# axis), peaking during the first steps (--workload linear --point-source)
POINT_SOURCE = dict(x0=(0.5, 0.5, 0.5), amp=(1.0, 0.5, -0.5, 0.25, 2.0), f0=10.0, t0=0.005)


def run_advection(pde, N, n, steps, warmup, linear=False, point_source=None):
    import exahype_amd as E
    s = E.ADERDGSolver(N, pde=pde, velocity=NCP_VELOCITY, linear=linear, point_source=point_source)
    g = E.Grid(s, n, n, n)
    g.set_initial_conditions(E.make_ic(3, **advection_ic_kwargs()))
    g.compute_dt()
    g.step(warmup)
    g.sync()
    g.kernel_times()
    g.total_picard_iterations()
    g.enable_timing(True)
    E.device_synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    E.device_synchronize()
    wall = time.perf_counter() - t0
    times, counts = g.kernel_times()
    sweeps = g.total_picard_iterations() / (g.ncells * steps)
    g.close()
    s.close()
    return dict(wall=wall, times=times, counts=counts, iters=sweeps)


def run_ncp(args):
    """Linear advection of 5 variables in non-conservative form (no flux,
    nonConservativeProduct = a . grad Q: the useFlux = false / useNCP = true
    arms) on n^3 periodic cells at order N, with the flux form of the same
    PDE timed beside it; the predictor is the dominant kernel."""
    import exahype_amd as E
    N, n = args.order, args.cells
    ncells = n ** 3
    dof = ncells * NVAR * (N + 1) ** 3
    r = run_advection(E.PDE_ADVECTION_NCP, N, n, args.steps, args.warmup)
    rf = run_advection(E.PDE_ADVECTION, N, n, args.steps, args.warmup)
    B = N + 1
    pred_ms = r["times"][0] / max(r["counts"][0], 1)
    # per sweep: the dudx gradients (6 V B^5) and the iK1 contraction (2 V B^5),
    # B^4 products a . grad Q and fused-source updates; then time averages,
    # volume integral (source part) and face extrapolation
    flops = (r["iters"] * (8 * NVAR * B ** 5 + B ** 4 * (PHI_NCP + 5 * NVAR)) + 14 * NVAR * B ** 4
             + 24 * NVAR * B ** 3) * ncells
    achieved = flops / (pred_ms * 1e-3) / 1e12
    out = {
        "metric": "DoF-updates/s, 3D linear advection in non-conservative form, ADER-DG N=%d" % N,
        "value": dof * args.steps / r["wall"],
        "unit": "DoF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D linear advection of 5 variables, non-conservative form (no flux, B_d = a_d I), "
                               "ADER-DG order %d, %d^3 periodic cells, 2 sine modes, CFL 0.9" % (N, n),
                   "velocity": list(NCP_VELOCITY), "cells": ncells, "order": N, "dof_per_step": dof,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "k_predictor_pde<%d, PdeAdvectionNcp>" % N, "avg_launch_ms": pred_ms,
                     "picard_sweeps_per_cell_step": r["iters"], "flop_per_launch": flops},
        "phase_ms_per_step": {"predictor": pred_ms, "riemann": r["times"][1] / max(r["counts"][1], 1),
                              "update_surface_dt": r["times"][2] / max(r["counts"][2], 1)},
        "flux_form": {"dof_updates_per_s": dof * args.steps / rf["wall"], "ms_per_step": rf["wall"] / args.steps * 1e3,
                      "predictor_ms": rf["times"][0] / max(rf["counts"][0], 1), "sweeps_per_cell": rf["iters"]},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = host_threads()
        steps = max(1, min(8, int(round(8 * (6.0 / (N + 1)) ** 5))))
        og = O.OracleGrid(O.cfg(N, pde=4, adv=NCP_VELOCITY), 27, 27, 27)
        for d in range(3):
            og.g.width[d] = 27 / n
        og.init(O.make_ic(kind=3, **advection_ic_kwargs()))
        dt = og.stable_dt(threads)
        t0 = time.perf_counter()
        for _ in range(steps):
            dt = og.step(dt, threads)
        wall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": og.ncells * NVAR * B ** 3 * steps / wall, "unit": "DoF-updates/s",
                               "cores": threads, "kind": "port",
                               "sample": "oracle (C, OpenMP), order %d, 27^3 periodic cells of size 1/%d, same "
                                         "initial data, %d steps (%.1f s wall)" % (N, n, steps, wall)}
    return out


def run_linear(args):
    """The linear kernel family (Cauchy-Kowalevski predictor, coefficient-
    matrix Riemann solver, fluctuation-form surface integral) on linear
    advection in non-conservative form, n^3 periodic cells at order N."""
    import exahype_amd as E
    N, n = args.order, args.cells
    ncells = n ** 3
    B = N + 1
    dof = ncells * NVAR * B ** 3
    ps = POINT_SOURCE if args.point_source else None
    r = run_advection(E.PDE_ADVECTION_NCP, N, n, args.steps, args.warmup, linear=True, point_source=ps)
    pred_ms = r["times"][0] / max(r["counts"][0], 1)
    # per time level: dudx gradients 6 V B^4, products and the next level
    # 6 V B^3, two Taylor accumulations 8 V B^3; then the volume integral
    # (6 V B^3) and the face extrapolation (24 V B^3)
    flops = (B * (6 * NVAR * B ** 4 + 14 * NVAR * B ** 3) + 30 * NVAR * B ** 3) * ncells
    # compulsory bytes: luh read, lduh written, both extrapolated face arrays written
    nbytes = 8 * (2 * NVAR * B ** 3 + 12 * NVAR * B ** 2) * ncells
    t = pred_ms * 1e-3
    ai = flops / nbytes
    if ai >= PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBPS * 1e9):
        roof = {"bound": "fp64_valu", "achieved": flops / t / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": nbytes / t / 1e9, "peak": PEAK_HBM_GBPS, "unit": "GB/s"}
    kname = "k_predictor_linear<%d, PdeAdvectionNcp%s>" % (N, ", point source" if ps else "")
    roof.update(frac=roof["achieved"] / roof["peak"], traffic=measured_traffic(kname, n, N), kernel=kname,
                avg_launch_ms=pred_ms, flop_per_launch=flops,
                bytes_per_launch=nbytes, flop_per_byte=ai)
    out = {
        "metric": "DoF-updates/s, 3D linear advection%s, linear ADER-DG kernels N=%d" % (
            " with a point source" if ps else "", N),
        "value": dof * args.steps / r["wall"],
        "unit": "DoF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D linear advection of 5 variables, linear ADER-DG kernels (Cauchy-Kowalevski "
                               "predictor, coefficient-matrix Rusanov), order %d, %d^3 periodic cells, 2 sine modes, "
                               "CFL 0.9" % (N, n),
                   "velocity": list(NCP_VELOCITY), "cells": ncells, "order": N, "dof_per_step": dof,
                   "point_source": ps, "parallelism": "single GPU"},
        "roofline": roof,
        "phase_ms_per_step": {"predictor": pred_ms, "riemann": r["times"][1] / max(r["counts"][1], 1),
                              "update_surface_dt": r["times"][2] / max(r["counts"][2], 1)},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = host_threads()
        steps = 20
        # the point source of the sample sits at the same place of the
        # (smaller) domain: the centre of its middle cell
        pso = None if ps is None else dict(ps, x0=(13.5 / n,) * 3)
        og = O.OracleGrid(O.cfg(N, pde=4, adv=NCP_VELOCITY, linear=1, point_source=pso), 27, 27, 27)
        for d in range(3):
            og.g.width[d] = 27 / n
        og.init(O.make_ic(kind=3, **advection_ic_kwargs()))
        dt = og.stable_dt(threads)
        t0 = time.perf_counter()
        for _ in range(steps):
            dt = og.step(dt, threads)
        wall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": og.ncells * NVAR * B ** 3 * steps / wall, "unit": "DoF-updates/s",
                               "cores": threads, "kind": "port",
                               "sample": "oracle (C, OpenMP), order %d, 27^3 periodic cells of size 1/%d, same "
                                         "initial data, %d steps (%.1f s wall)" % (N, n, steps, wall)}
    return out


def fv2d_ic(mod, seed=0):
    """EulerFV demonstrator state: rho = 1 + 0.5 exp(-|x-x0|^2/0.01), u = v = w = 0.1, p = 1."""
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(0.3, 0.7, 2)
    return mod.make_ic(5, xc=(x0[0], x0[1], 0.5), L=0.01, amp=0.5, rho_inf=1.0, p_inf=1.0, uinf=(0.1, 0.1, 0.1))


def run_fv2d(args):
    """EulerFV demonstrator (Demonstrators/EulerFV/EulerFV.exahype): 2D Euler,
    first-order Godunov with Rusanov fluxes, 81x81 patches of 10x10 finite
    volumes (810^2 volumes), periodic, CFL 0.9, 200 steps.  GPU steps timed;
    the oracle runs the same 200 steps on the host and the final states are
    compared (max over variables of ||gpu - oracle||_inf / ||oracle||_inf)."""
    import exahype_amd as E
    from oracle import oracle as O
    n, patch, steps = 81, 10, args.steps if args.steps != 10 else 200
    s = E.FiniteVolumesSolver(2, patch, E.FV_GODUNOV)
    g = E.FVGrid(s, n, n, 1)
    g.set_initial_conditions(fv2d_ic(E))
    dt0 = g.compute_dt()
    g.step(1)
    g.sync()
    g.set_initial_conditions(fv2d_ic(E))
    g.compute_dt()
    E.device_synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    t_end, _, _ = g.sync()
    wall = time.perf_counter() - t0
    gpu = g.download()
    vols = (n * patch) ** 2
    rate = vols * NVAR * steps / wall
    out = {"metric": "DoF-updates/s, 2D Euler FV Godunov (EulerFV demonstrator, 810^2 volumes)", "value": rate,
           "unit": "DoF-updates/s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": wall / steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "EulerFV demonstrator: 2D Euler, Godunov + Rusanov, 81x81 patches of 10x10, "
                                  "periodic, CFL 0.9", "volumes": vols, "first_dt": dt0, "t_end": t_end}}
    if not args.no_cpu_baseline:
        threads = host_threads()
        c = O.cfg(1, pde=2)
        og = O.OracleFVGrid(c, 2, 0, patch, n, n, 1)
        og.init(fv2d_ic(O))
        dt = og.stable_dt(threads)
        t0 = time.perf_counter()
        for _ in range(steps):
            dt, _ = og.step(dt, threads)
        cwall = time.perf_counter() - t0
        S = patch + 2
        sl = (slice(None), slice(1, 1 + patch), slice(1, 1 + patch))
        a = gpu.reshape(n * n, S, S, NVAR)[sl].reshape(-1, NVAR)
        o = og.luh().reshape(n * n, S, S, NVAR)[sl].reshape(-1, NVAR)
        scale = np.abs(o).max(axis=0)
        err = float((np.abs(a - o).max(axis=0) / np.maximum(scale, 1e-3 * scale.max())).max())
        out["cpu_baseline"] = {"value": vols * NVAR * steps / cwall, "unit": "DoF-updates/s", "cores": threads,
                               "kind": "port", "sample": "oracle (C, OpenMP), the whole workload, %d steps (%.1f s)"
                                                         % (steps, cwall)}
        out["parity_vs_oracle_rel_inf"] = err
    return out


def sod_ic(mod):
    return mod.make_ic(1, xc=(0.5, 0.5, 0.5), r0=0.25, rho_in=1.0, p_in=1.0, rho_out=0.125, p_out=0.1)


def run_muscl(args):
    """3D MUSCL-Hancock finite volumes (kernels/finitevolumes/musclhancock/c/3d/
    musclhancock.cpph) on the limiter's subcell grid of the N=5 Sod sphere: 27^3
    patches of 11^3 volumes with 2 ghost layers (297^3 volumes), periodic,
    20 steps.  Roofline: HBM bytes of the MUSCL-Hancock kernel (slopes, half
    step and Riemann fluxes in one launch), counted as the compulsory read of
    every patch (ghosts included) and write of its interior, over the kernel's
    HIP-event time."""
    import exahype_amd as E
    from oracle import oracle as O
    n, patch, steps = 27, 11, args.steps if args.steps != 10 else 20
    S = patch + 4
    s = E.FiniteVolumesSolver(3, patch, E.FV_MUSCLHANCOCK)
    g = E.FVGrid(s, n, n, n)
    g.set_initial_conditions(sod_ic(E))
    g.compute_dt()
    g.step(1)
    g.sync()
    g.kernel_times()
    g.enable_timing(True)
    E.device_synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    t_end, _, _ = g.sync()
    wall = time.perf_counter() - t0
    times, counts = g.kernel_times()
    ms = times / np.maximum(counts, 1)
    vols = (n * patch) ** 3
    rate = vols * NVAR * steps / wall
    bytes_upd = n ** 3 * NVAR * 8 * (S ** 3 + patch ** 3)
    upd_ms = ms[1] + ms[2]
    achieved = bytes_upd / (upd_ms * 1e-3) / 1e9
    out = {"metric": "DoF-updates/s, 3D Euler FV MUSCL-Hancock (27^3 patches of 11^3 volumes)", "value": rate,
           "unit": "DoF-updates/s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": wall / steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "3D Euler MUSCL-Hancock + Rusanov, minmod slopes, CFL 0.99, 27^3 patches of 11^3 "
                                  "volumes (2 ghost layers), periodic, Sod sphere r0=0.25", "volumes": vols,
                      "t_end": t_end},
           "phase_ms_per_step": {"ghosts": ms[0], "update": ms[2], "dt": ms[3]},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                        "kernel": "k_muscl", "avg_launch_ms": upd_ms,
                        "bytes_per_launch": bytes_upd}}
    if not args.no_cpu_baseline:
        threads = host_threads()
        c = O.cfg(1)
        ns, st = 27, 8
        og = O.OracleFVGrid(c, 3, 1, patch, ns, ns, ns)
        og.init(sod_ic(O))
        dt = og.stable_dt(threads)
        t0 = time.perf_counter()
        for _ in range(st):
            dt, _ = og.step(dt, threads)
        cwall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": (ns * patch) ** 3 * NVAR * st / cwall, "unit": "DoF-updates/s",
                               "cores": threads, "kind": "port",
                               "sample": "oracle (C, OpenMP), the same 27^3 patches of 11^3 volumes, Sod sphere, "
                                         "%d steps (%.1f s)" % (st, cwall)}
    return out


def run_amr(args):
    """Batched AMR level transfers of volume unknowns (amrRoutines.cpph:
    volumeUnknownsProlongation / Restriction): every cell of a 27^3 N=5 grid
    prolongated to its 27 children one level down (531441 transfers), then the
    children restricted back (27 consecutive transfers per coarse cell).
    Roofline: compulsory HBM bytes (read the inputs once, write the outputs
    once) over the kernel's HIP-event time, median of 5 launches."""
    import exahype_amd as E
    from oracle import oracle as O
    N, n = 5, 27
    B = N + 1
    blk = B ** 3 * NVAR
    nc = n ** 3
    nf = nc * 27
    rng = np.random.default_rng(1)
    coarse_h = 1.0 + rng.random(nc * blk)
    coarse = E.DeviceArray(coarse_h)
    fine = E.DeviceArray(nf * blk)
    back = E.DeviceArray(nc * blk)
    child = np.arange(nf)
    par = child // 27
    sub = child % 27
    idx = np.stack([sub % 3, (sub // 3) % 3, sub // 9], axis=1)
    tr = E.amr_transfer_array(child * blk, par * blk, 1, idx)
    res = {}
    for name, fn, args_ in (("prolongation", E.amrProlongationBatched, (3, N, NVAR, fine, coarse, tr)),
                            ("restriction", E.amrRestrictionBatched, (3, N, NVAR, back, fine, tr))):
        t = []
        for _ in range(6):
            if name == "restriction":
                back.upload(np.zeros(nc * blk))
            fn(*args_)
            t.append(E.amr_last_kernel_ms())
        ms = float(np.median(t[1:]))
        byts = (nc + nf) * blk * 8
        res[name] = {"kernel_ms": ms, "bytes": byts, "GBps": byts / (ms * 1e-3) / 1e9}
    # identity check: restrict(prolong(u)) == u for every cell (L2 projection of an exact embedding)
    err = float(np.abs(back.numpy() - coarse_h).max() / np.abs(coarse_h).max())
    ms_tot = res["prolongation"]["kernel_ms"] + res["restriction"]["kernel_ms"]
    rate = 2 * nf / (ms_tot * 1e-3)
    dom = max(res, key=lambda k: res[k]["kernel_ms"])
    achieved = res[dom]["GBps"]
    out = {"metric": "AMR volume transfers/s (N=5, one level, prolongation + restriction)", "value": rate,
           "unit": "transfers/s", "n_gpus": 1, "steps": 5, "warmup": 1, "ms_per_step": ms_tot,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "27^3 cells N=5 -> 27 children each (531441 prolongations), then 531441 "
                                  "restrictions back", "transfers_each": nf},
           "kernels": res, "restrict_prolong_identity_rel_inf": err,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                        "kernel": "k_amr_transfer<3, %s>" % ("true" if dom == "prolongation" else "false"),
                        "avg_launch_ms": res[dom]["kernel_ms"]}}
    if not args.no_cpu_baseline:
        # the oracle's per-cell transfers, one thread, on a sample of the list
        cnt = 50000
        ch = coarse_h.reshape(nc, blk)
        t0 = time.perf_counter()
        for k in range(cnt):
            O.volume_unknowns_prolongation(N, ch[par[k]], 0, 1, tuple(idx[k]))
        cwall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cnt / cwall, "unit": "transfers/s", "cores": 1, "kind": "port",
                               "sample": "oracle volume prolongation, %d transfers through the Python wrapper "
                                         "(%.2f s)" % (cnt, cwall)}
    for a in (coarse, fine, back):
        a.close()
    return out


def ensemble_members(n, seed=5):
    """n members: order ~ U{3..7}, cells ~ U{9^3, 27^3, 45^3}, gamma ~ U[1.3, 1.67],
    smooth seeded initial data (vortex of random strength/drift + 8 density modes)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(rng.choice([9, 27, 45]))
        ic = dict(xc=tuple(rng.uniform(0.3, 0.7, 3)), L=0.1, eps=float(rng.uniform(1, 5)),
                  uinf=tuple(rng.uniform(-1, 1, 3)), rho_inf=1.0, p_inf=1.0, amp=1e-2, nmodes=8,
                  k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                  phi=rng.uniform(0, 2 * np.pi, 8))
        out.append(dict(order=int(rng.integers(3, 8)), cells=(side, side, side), gamma=float(rng.uniform(1.3, 1.67)),
                        ic=ic))
    return out


def member_bytes(m):
    B = m["order"] + 1
    nc = m["cells"][0] * m["cells"][1] * m["cells"][2]
    return nc * (2 * NVAR * B ** 3 + 12 * NVAR * B ** 2) * 8 + nc * 20


def run_ensemble(args):
    """Ensemble batching: members of random order, size and gamma advanced
    20 steps by shared launches (one hipGraph per step), in waves that fit
    the device; each wave runs the mixed-precision path and the FP64 path
    side by side, and every member's L2 difference between them is reported."""
    import exahype_amd as E
    steps = args.steps if args.steps != 10 else 20
    members = ensemble_members(args.instances)
    cap = 100e9  # bytes per ensemble of a wave (two are resident: mixed and FP64)
    waves, cur, size = [], [], 0
    for i, m in enumerate(members):
        if cur and size + member_bytes(m) > cap:
            waves.append(cur)
            cur, size = [], 0
        cur.append(i)
        size += member_bytes(m)
    waves.append(cur)
    prec, sweeps = E.PREC_FP32, 1
    dof_steps, wall_mx, wall_64, l2s = 0, 0.0, 0.0, [None] * len(members)
    for wi, w in enumerate(waves):
        print("ensemble wave %d/%d: %d members" % (wi + 1, len(waves), len(w)), file=sys.stderr, flush=True)
        ms = [members[i] for i in w]
        ics = [E.make_ic(0, **m["ic"]) for m in ms]
        runs = {}
        for p, sw in ((prec, sweeps), (E.PREC_FP64, 1)):
            ens = E.Ensemble(ms, precision=p, fp64_sweeps=sw)
            ens.set_initial_conditions(ics)
            ens.compute_dt()
            ens.step(1)  # graph capture + warm-up
            ens.sync()
            ens.set_initial_conditions(ics)
            ens.compute_dt()
            E.device_synchronize()
            t0 = time.perf_counter()
            ens.step(steps)
            ens.sync()
            wall = time.perf_counter() - t0
            runs[p] = (ens, wall)
            print("  precision %d: %d steps in %.2f s" % (p, steps, wall), file=sys.stderr, flush=True)
        wall_mx += runs[prec][1]
        wall_64 += runs[E.PREC_FP64][1]
        for j, i in enumerate(w):
            m = members[i]
            nc = m["cells"][0] * m["cells"][1] * m["cells"][2]
            dof_steps += nc * NVAR * (m["order"] + 1) ** 3 * steps
            a = runs[prec][0].download(j)
            b = runs[E.PREC_FP64][0].download(j)
            l2s[i] = float(l2_per_variable(a, b, m["order"], nc).max())
            if j % 32 == 31:
                print("  compared %d members" % (j + 1), file=sys.stderr, flush=True)
        for ens, _ in runs.values():
            ens.close()
    l2a = np.array(l2s)
    return {"metric": "DoF-updates/s, ensemble of %d independent 3D Euler ADER-DG grids (N 3..7, 9^3..45^3 cells)"
                      % len(members), "value": dof_steps / wall_mx, "unit": "DoF-updates/s", "n_gpus": 1,
            "steps": steps, "warmup": 1, "ms_per_step": wall_mx / steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": "ensemble: order ~ U{3..7}, cells ~ U{9^3,27^3,45^3}, gamma ~ U[1.3,1.67], "
                                   "seeded smooth states, 20 steps, shared launches per order, one hipGraph per step",
                       "members": len(members), "waves": len(waves), "predictor": "fp32 + 1 FP64 sweep",
                       "orders": {str(N): sum(1 for m in members if m["order"] == N) for N in range(3, 8)}},
            "fp64": {"dof_updates_per_s": dof_steps / wall_64, "speedup_mixed_vs_fp64": wall_64 / wall_mx},
            "l2_mixed_vs_fp64": {"max": float(l2a.max()), "median": float(np.median(l2a)),
                                 "members_above_1e-6": int((l2a > 1e-6).sum())}}


def run_limiter(args):
    """The limiter batch of the AMR-with-limiter configuration on its base
    level: a sphere of (rho, p) = (1, 1) inside r < 0.25 and (0.3, 0.3)
    outside on 27^3 cells, N=5; per step the snapshot extrema, the ADER-DG
    step, the DMP + positivity check and the projection of the troubled cells
    onto 11^3 subcells with ghost width 2 (one launch per stage).  The batch
    does not replace the troubled cells' solution, so the jump is one that
    the unlimited scheme keeps positive (the 1 : 0.125 Sod sphere drives its
    density negative in the first step; the limiting step with FV
    recomputation is the amr_limiter workload).  Reports the troubled
    fraction per step and the time of each stage."""
    import exahype_amd as E
    n, N, steps = 27, 5, args.steps if args.steps != 10 else 20
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.set_initial_conditions(E.make_ic(1, xc=(0.5, 0.5, 0.5), r0=0.25, rho_in=1.0, p_in=1.0, rho_out=0.3,
                                       p_out=0.3))
    g.compute_dt()
    patches = E.DeviceArray(n ** 3 // 4 * NVAR * (2 * N + 5) ** 3)  # room for a quarter of the cells
    t_snap = t_step = t_check = t_proj = 0.0
    counts = []
    for k in range(steps + 1):
        E.device_synchronize()
        t0 = time.perf_counter()
        g.limiter_snapshot(5)
        t1 = time.perf_counter()
        g.step(1)
        g.sync()
        t2 = time.perf_counter()
        ntr, _ = g.limiter_check()
        t3 = time.perf_counter()
        g.limiter_project(2, ntr, out=patches)
        t4 = time.perf_counter()
        if k == 0:
            continue  # warm-up
        counts.append(ntr)
        t_snap += t1 - t0
        t_step += t2 - t1
        t_check += t3 - t2
        t_proj += t4 - t3
    total = t_snap + t_step + t_check + t_proj
    dof = n ** 3 * NVAR * (N + 1) ** 3 * steps
    return {"metric": "DoF-updates/s, 3D Euler ADER-DG N=5 with troubled-cell detection and subcell projection "
                      "(density sphere 1 : 0.3, 27^3 cells)", "value": dof / total, "unit": "DoF-updates/s", "n_gpus": 1,
            "steps": steps, "warmup": 1, "ms_per_step": total / steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "density/pressure sphere r0=0.25 (1, 1 inside; 0.3, 0.3 outside), 27^3 periodic "
                                   "cells, N=5, DMP relax 1e-4, scaling 1e-3, 11^3 subcells + 2 ghost layers per "
                                   "troubled cell", "cells": n ** 3},
            "troubled_fraction": {"mean": float(np.mean(counts)) / n ** 3, "min": min(counts) / n ** 3,
                                  "max": max(counts) / n ** 3},
            "stage_ms_per_step": {"snapshot": t_snap / steps * 1e3, "aderdg_step": t_step / steps * 1e3,
                                  "check": t_check / steps * 1e3, "project": t_proj / steps * 1e3},
            "note": "host-timed stages incl. synchronisation and the troubled count / cell list read back; this "
                    "base-level batch does not replace the troubled cells' solution (the limiting step with FV "
                    "recomputation is the amr_limiter workload)"}


AMR_BASE, AMR_LEVELS, AMR_RADIUS, AMR_WIDTHS = 27, 3, 0.25, (0.06, 0.031)


def amr_centre(seed=4):
    """The seeded centre of the refinement shell and of the Sod sphere."""
    rng = np.random.default_rng(seed)
    return tuple(0.5 + rng.uniform(-0.01, 0.01, 3))


def run_parity81(args):
    """The plain FP64 drop-in checked on its full workload: 3D Euler N=3 on
    81^3 periodic cells, vortex + 8 density modes, CFL 0.9, 50 steps -- the
    GPU grid (one hipGraph per step, wall clock) and the oracle (OpenMP on
    the host's CPU share) from the same initial state.  Reports the GPU
    throughput, the oracle's on the whole workload, and the final-state
    parity (max |a-b| / max |b| per variable) and Picard counts."""
    import exahype_amd as E
    from oracle import oracle as O
    N, n = 3, args.cells
    steps = args.steps if args.steps != 10 else 50
    threads = host_threads()
    c = O.cfg(N)
    og = O.OracleGrid(c, n, n, n)
    og.init(O.make_ic(0, **vortex_ic_kwargs()))
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.upload(og.luh().copy())
    dt_g = g.compute_dt()
    g.use_graph(True)
    E.device_synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    E.device_synchronize()
    wall_g = time.perf_counter() - t0
    a = g.download()
    it_g = g.picard_iterations()
    print("parity81: GPU %d steps in %.3f s" % (steps, wall_g), file=sys.stderr, flush=True)
    dt = og.stable_dt(threads)
    t0 = time.perf_counter()
    for k in range(steps):
        dt = og.step(dt, threads)
        if (k + 1) % 5 == 0:
            print("parity81: oracle %d steps, %.1f s" % (k + 1, time.perf_counter() - t0), file=sys.stderr, flush=True)
    wall_o = time.perf_counter() - t0
    b = og.luh()
    A, Bm = a.reshape(-1, NVAR), b.reshape(-1, NVAR)
    scale = np.abs(Bm).max(axis=0)
    # the parity tests' measure: a variable whose scale is below 1e-3 of the
    # largest (the z momentum, zero at t = 0 here) is measured against that floor
    rel = (np.abs(A - Bm).max(axis=0) / np.maximum(scale, 1e-3 * scale.max())).tolist()
    rel_own = (np.abs(A - Bm).max(axis=0) / scale).tolist()
    dof = n ** 3 * NVAR * (N + 1) ** 3 * steps
    return {"metric": "DoF-updates/s, plain FP64 ADER-DG N=3 drop-in, parity against the oracle on the full workload",
            "value": dof / wall_g, "unit": "DoF-updates/s", "n_gpus": 1, "steps": steps, "warmup": 0,
            "ms_per_step": wall_g / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "3D Euler ADER-DG, order 3, %d^3 periodic cells, vortex + 8 density modes, CFL 0.9, "
                                   "%d steps, GPU and oracle from the same initial state" % (n, steps),
                       "parallelism": "single GPU"},
            "parity_vs_oracle_rel_inf": rel, "parity_vs_oracle_rel_own_scale": rel_own,
            "variable_scales": scale.tolist(), "parity_bar": 1e-12, "parity_ok": bool(max(rel) <= 1e-12),
            "picard_counts_equal_last_step": bool(np.array_equal(it_g, og.iters())),
            "cpu_baseline": {"value": dof / wall_o, "unit": "DoF-updates/s", "cores": threads, "kind": "port",
                             "host": host_info(),
                             "sample": "oracle (C, OpenMP), the whole workload: %d^3 cells, %d steps (%.1f s)"
                                       % (n, steps, wall_o)},
            "timing": "GPU: wall clock of the steps replayed as one hipGraph per step; oracle: wall clock of its "
                      "steps"}


def run_amr_limiter(args):
    """Limiting ADER-DG on a 3-level adaptive mesh: 3D Euler N=5, base 27^3
    cells refined by 3 twice where a cell meets a spherical shell around the
    Sod discontinuity (seeded centre, radius 0.25), 2:1 balanced; Sod sphere
    (rho, p = 1, 1 inside r < 0.25; 0.125, 0.1 outside); MUSCL-Hancock on
    11^3 subcells for the troubled cells and their neighbours on the finest
    level; 100 steps, one hipGraph per step.  The predictor runs in FP64 and in
    FP32 with one FP64 correction sweep (update, Riemann fluxes and the time
    step stay FP64); value = the mixed run.  Roofline: the predictor's
    algorithmic FLOP (Picard sweeps measured) over its HIP-event time."""
    import exahype_amd as E
    N, steps = 5, args.steps if args.steps != 10 else 100
    centre = amr_centre()
    cfg = E.amr_config(AMR_BASE, AMR_LEVELS, centre, AMR_RADIUS, AMR_WIDTHS, relax=1e-4, scaling=1e-3,
                       steps_till_cured=0)
    ic = E.make_ic(1, xc=centre, r0=AMR_RADIUS, rho_in=1.0, p_in=1.0, rho_out=0.125, p_out=0.1)
    runs, states = {}, {}
    for name, prec in (("fp64", E.PREC_FP64), ("fp32", E.PREC_FP32)):
        s = E.ADERDGSolver(N, precision=prec, fp64_sweeps=1)
        g = E.AMRGrid(s, cfg)
        cells = g.cells()
        g.set_initial_conditions(ic)
        g.compute_dt()
        print("amr_limiter %s: %d leaves, first step" % (name, g.ncells), file=sys.stderr, flush=True)
        g.step(1)  # warm-up (graph capture)
        g.sync()
        it0 = g.total_picard_iterations()
        E.device_synchronize()
        t0 = time.perf_counter()
        chunk = 20
        for k in range(0, steps, chunk):
            g.step(min(chunk, steps - k))
            g.sync()
            print("amr_limiter %s: %d steps" % (name, min(k + chunk, steps)), file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        t_end, dt = g.sync()
        sweeps = (g.total_picard_iterations() - it0) / (steps * g.ncells)
        status = g.limiter_status()
        cnt = g.counters()
        # phase times of a few further steps, graph off
        g.enable_timing(True)
        g.step(3)
        ms, nst = g.phase_times()
        g.enable_timing(False)
        fin = cells[:, 0] == AMR_LEVELS - 1
        runs[name] = {"dof_updates_per_s": g.ncells * NVAR * (N + 1) ** 3 * steps / wall,
                      "ms_per_step": wall / steps * 1e3, "sweeps_per_cell": sweeps, "t_end": t_end,
                      "phase_ms_per_step": dict(zip(("predictor", "transfers_merge", "riemann", "dg_update",
                                                     "limiter", "time_step"), (ms / max(nst, 1)).tolist())),
                      "counters": cnt,
                      "troubled_fraction_per_step": cnt["troubled"] / (steps + 1) / g.ncells,
                      "final_status": {E.AMR_STATUS_NAMES[i]: int((status[fin] == i).sum()) for i in range(6)}}
        states[name] = g.download()
        levels = np.bincount(cells[:, 0], minlength=AMR_LEVELS)
        ncells = g.ncells
        g.close()
        s.close() if hasattr(s, "close") else None
    norms = error_norms(states["fp32"], states["fp64"], N, ncells)
    m = runs["fp32"]
    m["l1_vs_fp64_per_variable"] = norms["l1"].tolist()
    m["l2_vs_fp64_per_variable"] = norms["l2"].tolist()
    m["linf_vs_fp64_per_variable"] = norms["linf"].tolist()
    m["speedup_vs_fp64"] = m["dof_updates_per_s"] / runs["fp64"]["dof_updates_per_s"]
    pred_ms = m["phase_ms_per_step"]["predictor"]
    flop = predictor_flops_per_cell(N, m["sweeps_per_cell"]) * ncells
    achieved = flop / (pred_ms * 1e-3) / 1e12
    # flop-weighted peak: the FP64 correction sweep and the tail in FP64, the rest in FP32
    sw = m["sweeps_per_cell"]
    f64 = predictor_flops_per_cell(N, 1.0) * ncells
    frac64 = min(1.0, f64 / flop)
    peak = 1.0 / (frac64 / PEAK_FP64_TFLOPS + (1 - frac64) / PEAK_FP32_TFLOPS)
    out = {"metric": "DoF-updates/s, limiting ADER-DG N=5 on a 3-level adaptive mesh (Sod sphere), mixed predictor",
           "value": m["dof_updates_per_s"], "unit": "DoF-updates/s", "n_gpus": 1, "steps": steps, "warmup": 1,
           "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "3D Euler limiting ADER-DG N=5, base 27^3 cells + 2 levels refined by 3 in a "
                                  "spherical shell (radius 0.25, half widths 0.06 / 0.031, seeded centre), 2:1 "
                                  "balanced, Sod sphere, MUSCL-Hancock 11^3 subcells on the finest level, DMP relax "
                                  "1e-4 / scaling 1e-3, global time step, 100 steps",
                      "cells": int(ncells), "cells_per_level": levels.tolist(), "centre": list(centre),
                      "order": N, "predictor": "fp32 Picard sweeps + 1 FP64 correction sweep; FP64 update",
                      "parallelism": "single GPU"},
           "precisions": runs,
           "roofline": {"bound": "fp32_valu+fp64", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                        "frac": achieved / peak, "traffic": None, "kernel": "k_predictor<5, fp32 + 1 FP64 sweep>",
                        "avg_launch_ms": pred_ms, "picard_sweeps_per_cell_step": sw, "flop_per_launch": flop},
           "timing": "value: wall clock of the steps replayed as one hipGraph per step; phase_ms_per_step: HIP "
                     "events around the stages of 3 further steps (graph off)"}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        from oracle import amr_limiting as A
        og = A.LimitingGrid(A.Mesh(3, AMR_LEVELS, centre, AMR_RADIUS, list(AMR_WIDTHS)), N)
        og.init(O.make_ic(1, xc=centre, r0=AMR_RADIUS, rho_in=1.0, p_in=1.0, rho_out=0.125, p_out=0.1))
        dt = og.stable_dt()
        t0 = time.perf_counter()
        og.step(dt)
        cwall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": og.nc * NVAR * (N + 1) ** 3 / cwall, "unit": "DoF-updates/s", "cores": 1,
                               "kind": "port",
                               "sample": "oracle (per-cell C kernels driven from Python, one thread), the same "
                                         "shell and Sod sphere on a base-3^3 mesh (%d leaves), 1 step (%.1f s)"
                                         % (og.nc, cwall)}
    return out


def copy_reference(gib):
    """A plain streaming device-to-device copy of `gib` GiB by the library's
    copy kernel: bytes moved and GB/s (HBM reference, counter calibration)."""
    import exahype_amd as E
    n = int(gib * (1 << 30)) // 8
    a = E.DeviceArray(n)
    b = E.DeviceArray(n)
    E.device_copy(b, a, 8 * n)
    ms = min(E.device_copy(b, a, 8 * n) for _ in range(3))
    a.close()
    b.close()
    return {"kernel": "k_stream_copy", "bytes_read": 8 * n, "bytes_written": 8 * n, "ms": ms,
            "GBps": 16 * n / (ms * 1e-3) / 1e9}


if __name__ == "__main__":
    main()
