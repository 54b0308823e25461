#!/usr/bin/env python3
"""Benchmark of the EbE linear-elasticity solve path on one MI355X.

Workload (N=1): 3D static tetrahedra on the unit cube, 64x64x64 hexes split
into 6 Kuhn tets each (1,572,864 elements, 274,625 nodes, 823,875 DOF),
z=0 face clamped, uniform pressure 1e3 on the z=1 face, E=2e7, nu=0.25.
One "step" is one complete static solve to ||f - K u||_2 / ||f||_2 <= 1e-8
with mixed precision (FP32 EbE matvec + FP32 Jacobi-PCG inside FP64
iterative refinement), mesh and right-hand side already resident in HBM.

Printed (one JSON line on stdout): time-to-solve of the mixed solve (the
headline value, seconds, lower is better) plus the FP64 PCG time-to-solve,
the mixed/FP64 speedup and displacement difference, PCG iterations/s, EbE
matvec DOF-matvecs/s and GB/s in FP32 and FP64, the roofline of the dominant
kernel and the CPU baseline (the oracle's FP32 restatement of the reference's
single-threaded CPU PCG, timed on a bounded sample and extrapolated).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public spec


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_matvec(ne, nn, ndof, s):
    """Algorithmic bytes of one EbE matvec y = K x: int32 connectivity
    (4 per node of each element), node coordinates, read x, write y; each
    array counted once."""
    return 4 * 4 * ne + s * 3 * nn + 2 * s * ndof


def read_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, or None."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for ln in f:
            if ":" in ln:
                k, v = ln.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    if "fetch_bytes_per_launch" in vals and "write_bytes_per_launch" in vals:
        return vals["fetch_bytes_per_launch"] + vals["write_bytes_per_launch"]
    return None


def traffic_profile(threads, elements):
    """The committed PMC summary of the FP32 cube64 matvec at this patch
    tuning (the default 256 / 768 keeps the plain name)."""
    tag = "" if (threads, elements) == (256, 768) else f"_{threads}x{elements}"
    return os.path.join(ROOT, "profiles", f"ebe_matvec_fp32_cube64{tag}_traffic.txt")


def parse_patch(v):
    t, e = v.lower().split("x")
    return int(t), int(e)


def cpu_baseline(pb, n_fp64_iters, cpu_iters):
    """The oracle's FP32 restatement of the reference's CPU path (element-local
    vectors, per-DOF sums in element order, single thread): setup + a bounded
    number of PCG iterations, extrapolated to the iteration count of the
    converged solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    setup_s, iter_s = O.static_time_f32(pb, cpu_iters)
    per_it = iter_s / cpu_iters
    est = setup_s + per_it * n_fp64_iters
    return {
        "value": est, "unit": "s", "cores": 1, "kind": "port",
        "sample": (f"cube64 (1,572,864 tets): setup {setup_s:.2f} s + {cpu_iters} PCG iterations "
                   f"{iter_s:.2f} s on 1 core, extrapolated to {n_fp64_iters} iterations"),
    }


def cpu_efficient_baseline(pb, n_fp64_iters, cpu_iters):
    """The "efficient CPU" line: the FP64 restatement with its loops spread
    over OpenMP threads (one per host core share), setup + a bounded number
    of PCG iterations, extrapolated like cpu_baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or min(os.cpu_count() or 1, 16)
    setup_s, iter_s = O.static_time_omp_f64(pb, cpu_iters, threads)
    est = setup_s + iter_s / cpu_iters * n_fp64_iters
    return {
        "value": est, "unit": "s", "cores": threads, "kind": "port-openmp-fp64",
        "sample": (f"cube64: FP64 restatement, {threads} OpenMP threads: setup {setup_s:.2f} s + "
                   f"{cpu_iters} PCG iterations {iter_s:.2f} s, extrapolated to "
                   f"{n_fp64_iters} iterations"),
    }


def copy_gbs(reps=20):
    """Device-to-device copy bandwidth (read + write bytes / time) of a 256 MiB
    float32 tensor, event-timed: the measured-copy ceiling.  Runs in a child
    process: torch ships its own HIP runtime, which cannot open the GPU in a
    process where the library's runtime already did."""
    code = (
        "import torch\n"
        "x = torch.empty(64 << 20, dtype=torch.float32, device='cuda'); y = torch.empty_like(x)\n"
        "x.fill_(1.0)\n"
        "[y.copy_(x) for _ in range(3)]\n"
        "e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)\n"
        f"e0.record(); [y.copy_(x) for _ in range({reps})]; e1.record()\n"
        "torch.cuda.synchronize()\n"
        f"print(2 * (256 << 20) / (e0.elapsed_time(e1) / 1e3 / {reps}) / 1e9)\n")
    import subprocess
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        log(f"device copy timing failed: {r.stderr.strip()[-300:]}")
        return None
    return float(r.stdout.strip().splitlines()[-1])


def lumped_mass(pb, rho):
    """Physical lumped mass per node: rho V / 4 of every tet at each corner."""
    X = pb.nodes[pb.elems]
    V = np.abs(np.einsum("ei,ei->e", X[:, 1] - X[:, 0],
                         np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6.0
    m = np.zeros(pb.n_nodes)
    np.add.at(m, pb.elems.ravel(), np.repeat(rho * V / 4.0, 4))
    return m


def omega_max(ctx, pb, m, iters=60):
    """Largest eigenvalue of M^-1 K by power iteration (free DOFs), as omega."""
    free = ~pb.fixed_dofs()
    x = np.random.default_rng(0).standard_normal((pb.n_nodes, 3)) * free
    lam = 0.0
    for _ in range(iters):
        y = ctx.matvec(x, 0) / m[:, None] * free
        lam = float(np.vdot(x, y) / np.vdot(x, x))
        x = y / np.linalg.norm(y)
    return np.sqrt(lam)


def bench_column(args, F):
    """Explicit GN22 (central difference, lumped mass) on the 16x16x256 column,
    Ricker 50 Hz top pressure, dt = 0.5 * 2/omega_max, 20,000 steps; the patch
    tuning of each precision from femgpu_autotune_dyn (--no-autotune: the
    default)."""
    from cudafemproject_amd.problem import column_problem
    pb = column_problem(16, 16, 256, 0.1, 0.1, 1.6, E=2e7, nu=0.25, pressure=1.0,
                        convention=F.LOADS_PHYSICAL)
    rho = 1000.0
    ctx = F.Context(0)
    ctx.set_problem(pb)
    w = omega_max(ctx, pb, lumped_mass(pb, rho))
    dt = 0.5 * 2.0 / w
    nsteps = args.dyn_steps
    log(f"column: {pb.n_elems} elems, {pb.n_dof} dof, omega_max {w:.4e}, dt {dt:.4e}, "
        f"{nsteps} steps")
    res = {}
    tuning = {}
    for name, prec in (("fp64", F.FP64), ("fp32", F.FP32)):
        kw = dict(rho=rho, beta1=0.5, beta2=0.0, dt=dt, endtime=(nsteps + 0.5) * dt,
                  wavelet_type=F.WAVELET_RICKER, freq=50.0, timeshift=0.0, ampl=1e3,
                  precision=prec, mass_convention=F.MASS_PHYSICAL)
        if not args.no_autotune:
            # the patch tuning of this stepper and precision (femgpu_autotune_dyn)
            tuning[name] = ctx.autotune_dyn(**{k: v for k, v in kw.items() if k != "endtime"})
            log(f"{name}: autotuned patches {tuning[name]}")
        for _ in range(args.warmup):
            ctx.dyn_run(**dict(kw, endtime=200 * dt))
        runs = [ctx.dyn_run(**kw) for _ in range(args.steps)]
        u, st = runs[-1]
        ms = np.mean([r[1]["run_ms"] for r in runs])
        res[name] = (ms, st["steps"], u)
        log(f"{name}: {st['steps']} steps in {ms:.1f} ms = {ms * 1e3 / st['steps']:.2f} us/step, "
            f"max |u| {np.abs(u).max():.3e}")
    ms64, n64, u64 = res["fp64"]
    ms32, n32, u32 = res["fp32"]
    N = pb.n_dof
    b64 = bytes_per_matvec(pb.n_elems, pb.n_nodes, N, 8) + 11 * 8 * N
    b32 = bytes_per_matvec(pb.n_elems, pb.n_nodes, N, 4) + 11 * 4 * N
    us64, us32 = ms64 * 1e3 / n64, ms32 * 1e3 / n32
    ach = b64 / (us64 * 1e-6) / 1e9
    line = {
        "metric": "explicit GN22 step time (column 16x16x256, 20,000 steps)",
        "value": us64, "unit": "us/step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "column 0.1x0.1x1.6, 16x16x256 hexes x 6 tets, bottom clamped, "
                               "Ricker 50 Hz x 1e3 top pressure, lumped mass rho=1000, "
                               "dt = 0.5 * 2/omega_max", "n_elems": pb.n_elems, "n_dof": N,
                   "dt": dt, "time_steps": n64},
        "us_per_step_fp32": us32, "steps_per_s_fp64": 1e6 / us64, "steps_per_s_fp32": 1e6 / us32,
        "patch_tuning": tuning or "default (256 threads, 768 elements)",
        "fp32_vs_fp64_rel_displacement_diff": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one explicit step fp64 (k_ebe_patch MODE 4: predictor, EbE matvec, corrector)",
                     "bytes_per_launch": b64},
        "fp32_step_gbs": b32 / (us32 * 1e-6) / 1e9,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_implicit(args, F):
    """Implicit GN22 (Newmark, beta1 = 0.5, beta2 = 0.5) on the 16x16x256
    column, physical operator M + beta2 dt^2/2 K with a nested Jacobi-PCG per
    step (eps 1e-8 on gamma), dt = 4x the explicit step, 200 steps."""
    from cudafemproject_amd.problem import column_problem
    pb = column_problem(16, 16, 256, 0.1, 0.1, 1.6, E=2e7, nu=0.25, pressure=1.0,
                        convention=F.LOADS_PHYSICAL)
    rho = 1000.0
    ctx = F.Context(0)
    ctx.set_problem(pb)
    dt = 4.0 * 0.5 * 2.0 / omega_max(ctx, pb, lumped_mass(pb, rho))
    nsteps = 200
    res = {}
    tuning = {}
    for name, prec in (("fp64", F.FP64), ("fp32", F.FP32)):
        kw = dict(rho=rho, beta1=0.5, beta2=0.5, dt=dt, endtime=(nsteps + 0.5) * dt,
                  wavelet_type=F.WAVELET_RICKER, freq=50.0, timeshift=0.0, ampl=1e3,
                  precision=prec, mass_convention=F.MASS_PHYSICAL,
                  implicit_mode=F.IMPLICIT_PHYSICAL, pcg_eps=1e-8)
        if not args.no_autotune:
            tuning[name] = ctx.autotune_dyn(**{k: v for k, v in kw.items() if k != "endtime"})
            log(f"implicit {name}: autotuned patches {tuning[name]}")
        for _ in range(args.warmup):
            ctx.dyn_run(**dict(kw, endtime=20.5 * dt))
        runs = [ctx.dyn_run(**kw) for _ in range(args.steps)]
        u, st = runs[-1]
        ms = np.mean([r[1]["run_ms"] for r in runs])
        res[name] = (ms, st["steps"], st["pcg_iterations"], u)
        log(f"implicit {name}: {st['steps']} steps, {st['pcg_iterations']} PCG iterations in "
            f"{ms:.1f} ms = {ms * 1e3 / st['steps']:.1f} us/step")
    ms64, n64, it64, u64 = res["fp64"]
    ms32, n32, it32, u32 = res["fp32"]
    line = {
        "metric": "implicit GN22 step time (column 16x16x256, 200 steps, PCG per step)",
        "value": ms64 * 1e3 / n64, "unit": "us/step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "column 0.1x0.1x1.6, 16x16x256 hexes x 6 tets, bottom clamped, "
                               "Ricker 50 Hz x 1e3 top pressure, rho=1000, Newmark beta1 0.5 "
                               "beta2 0.5, operator M + beta2 dt^2/2 K, PCG eps 1e-8, "
                               "dt = 4 x 0.5 * 2/omega_max", "n_elems": pb.n_elems,
                   "n_dof": pb.n_dof, "dt": dt, "time_steps": n64},
        "pcg_iterations_per_step_fp64": it64 / n64, "pcg_iterations_per_step_fp32": it32 / n32,
        "us_per_step_fp32": ms32 * 1e3 / n32,
        "us_per_pcg_iteration_fp64": ms64 * 1e3 / it64,
        "patch_tuning": tuning or "default (256 threads, 768 elements)",
        "fp32_vs_fp64_rel_displacement_diff": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_cantilever(args, F):
    """Compliance and its adjoint gradient on the 96x32x32 cantilever; central
    FD (h = 1e-6 in rho_e) on 64 seeded elements.  The FD evaluates compliance
    in its stationary form 2 f'u - u'Ku (the PCG error enters quadratically)
    and sums the terms of C(+h) - C(-h) exactly (math.fsum), so it is limited
    by the rounding of individual terms, not by the PCG tolerance or by the
    rounding of C itself (a one-element change moves C by ~1e-11)."""
    from cudafemproject_amd.problem import cantilever_problem
    pb = cantilever_problem()
    ctx = F.Context(0)
    ctx.set_problem(pb)
    rho = pb.rho
    ctx.compliance_grad(F.FP64, tol=1e-3)
    times = {}
    tuning = {}
    for name, prec in (("fp64", F.FP64), ("mixed", F.MIXED)):
        if not args.no_autotune:
            # each precision at the patch tuning femgpu_autotune picks for it
            tuning[name] = ctx.autotune(prec)
            log(f"{name}: autotuned patches {tuning[name]}")
        for _ in range(args.warmup):
            ctx.compliance_grad(prec, tol=1e-8)
        t = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            c, g, u, st = ctx.compliance_grad(prec, tol=1e-8)
            t.append(time.perf_counter() - t0)
        times[name] = (float(np.mean(t)), st, c, g)
        log(f"{name}: C={c:.9e} {st['iterations']} it, solve {st['solve_ms']:.1f} ms, "
            f"with gradient and host copies {np.mean(t) * 1e3:.1f} ms")
    t64, s64, c64, g64 = times["fp64"]
    tmx, smx, cmx, gmx = times["mixed"]

    f = ctx.rhs().ravel()

    def c_stat_terms(scale):
        """Terms of the stationary compliance 2 f'u - u'Ku = sum(2 f u) + sum(s g)."""
        ctx.set_material(pb.E, pb.nu, scale)
        _, g, u, _ = ctx.compliance_grad(F.FP64, tol=1e-12)
        return np.concatenate([2.0 * f * u.ravel(), scale * g]), g

    _, g0 = c_stat_terms(rho ** 3)
    adj = 3.0 * rho ** 2 * g0
    h = 1e-6
    sample = np.random.default_rng(0).choice(pb.n_elems, args.fd_elems, replace=False)
    fd = []
    for i, k in enumerate(sample):
        if i % 16 == 0:
            log(f"FD element {i}/{len(sample)}")
        ts = []
        for sgn in (1.0, -1.0):
            r = rho.copy()
            r[k] += sgn * h
            ts.append(c_stat_terms(r ** 3)[0])
        # exactly rounded sum of all terms of C(+h) - C(-h): no cancellation
        # error beyond the rounding of the individual terms
        fd.append(math.fsum(np.concatenate([ts[0], -ts[1]])) / (2 * h))
    fd = np.array(fd)
    err = np.abs(fd - adj[sample])
    fd_rel_max = float(err.max() / np.abs(fd).max())
    fd_rel_med = float(np.median(err / np.abs(fd)))
    log(f"FD check on {len(sample)} elements: max|adj-fd|/max|fd| {fd_rel_max:.2e}, "
        f"median per-element {fd_rel_med:.2e}")
    line = {
        "metric": "compliance + adjoint gradient time (cantilever 96x32x32)",
        "value": tmx, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tmx * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "cantilever 3x1x1, 96x32x32 hexes x 6 tets, rho_e ~ U[0.05,1] "
                               "seed 5, E_e = 2e7 rho_e^3, nu 0.3, x=0 clamped, tip pressure",
                   "n_elems": pb.n_elems, "n_dof": pb.n_dof, "tol": 1e-8},
        "fp64_s": t64, "iterations_fp64": s64["iterations"], "iterations_mixed": smx["iterations"],
        "patch_tuning": tuning or "default (256 threads, 768 elements)",
        "mixed_vs_fp64_grad_rel_l2": float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64)),
        "fd_elements": int(len(sample)), "fd_h": h,
        "fd_max_abs_err_over_max_fd": fd_rel_max, "fd_median_rel_err": fd_rel_med,
        "homogeneity_rel": float(abs(np.dot(pb.elem_scale, g64) + c64) / abs(c64)),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_batch(args, F):
    """512 independent plates x 4 load cases (lockstep schedule; the
    persistent one-workgroup-per-instance schedule is timed beside it)."""
    from cudafemproject_amd.problem import load_batch, plate_batch
    insts = plate_batch(512, seed=2024)
    ctx = F.Context(0)
    load_batch(ctx, insts)
    rhs = [b.f for b in insts]
    n_sys = len(insts) * 4
    res = {}
    runs = [("mixed", F.MIXED, F.BATCH_LOCKSTEP), ("fp64", F.FP64, F.BATCH_LOCKSTEP)]
    if not args.no_persistent:
        runs += [("mixed_persistent", F.MIXED, F.BATCH_PERSISTENT),
                 ("fp64_persistent", F.FP64, F.BATCH_PERSISTENT)]
    for name, prec, sched in runs:
        for _ in range(args.warmup):
            ctx.batch_solve(rhs, prec, tol=1e-8, schedule=sched)
        ms = []
        for _ in range(args.steps):
            us, st, m = ctx.batch_solve(rhs, prec, tol=1e-8, schedule=sched)
            ms.append(m)
        res[name] = (float(np.mean(ms)), us, st)
        log(f"{name}: {np.mean(ms):.2f} ms for {n_sys} systems, iterations max "
            f"{st['iterations'].max()} mean {st['iterations'].mean():.0f}, max rel_res "
            f"{st['rel_res'].max():.2e}")
    tmx, umx, smx = res["mixed"]
    t64, u64, s64 = res["fp64"]
    diff = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(umx, u64))
    # independent check: four instances through the single-problem solver
    chk = []
    for i in (0, 1, 2, int(np.argmax([len(b.elems) for b in insts]))):
        b = insts[i]
        ctx.set_problem(b.problem())
        for r in range(4):
            u1, _ = ctx.solve_rhs(b.f[:, r], F.FP64, tol=1e-12)
            chk.append(float(np.linalg.norm(umx[i][:, r] - u1) / np.linalg.norm(u1)))
    log(f"mixed vs fp64 max rel diff {diff:.2e}; vs single-problem solver max {max(chk):.2e}")
    # algorithmic bytes of the inner iterations actually run (FP32 inner vectors)
    tot = 0.0
    for b, it in zip(insts, smx["iterations"].max(1)):
        ne, nn = len(b.elems), len(b.nodes)
        N = 2 * nn
        tot += int(it) * (4 * 3 * ne + 4 * 2 * nn + 2 * 4 * N * 4 + 11 * 4 * N * 4)
    ach = tot / (tmx * 1e-3) / 1e9
    line = {
        "metric": "batched static solves: time for 512 plates x 4 load cases to rel-res 1e-8",
        "value": tmx / 1e3, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tmx, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "512 2D plates 1.0x0.5, nx~U{16..160}, ny~U{16..80} (seed 2024), "
                               "E~LogNormal(ln 2e7, 0.5), nu~U[0.2,0.35], left edge fixed, "
                               "4 load cases (x/y traction, shear, body force)",
                   "n_instances": len(insts), "n_systems": n_sys,
                   "n_elems": int(sum(len(b.elems) for b in insts)),
                   "n_dof": int(sum(2 * len(b.nodes) for b in insts)), "tol": 1e-8},
        "fp64_s": t64 / 1e3, "speedup_mixed_vs_fp64": t64 / tmx,
        "schedule": "lockstep (every system over the whole GPU, graph replays)",
        "persistent_schedule_s": ({k: res[k][0] / 1e3 for k in ("mixed_persistent", "fp64_persistent")}
                                  if "mixed_persistent" in res else None),
        "systems_per_s_mixed": n_sys / (tmx * 1e-3),
        "iterations_max_mixed": int(smx["iterations"].max()),
        "iterations_mean_mixed": float(smx["iterations"].mean()),
        "rel_res_max_mixed": float(smx["rel_res"].max()),
        "mixed_vs_fp64_max_rel_diff": diff, "vs_single_solver_max_rel_diff": max(chk),
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_bl_patch<float,2,4,1> + k_bl_update<float,2,4> (all iterations)",
                     "bytes_per_launch": tot},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_contraction(args, F):
    """The element contraction of the batch workload's triangles, F_e = K_e U_e
    for its four load cases, on the fly (the solver's form) against K_e and
    v_mfma_f32_4x4x1f32 (libfemgpu_probe.so); same forces, event-timed.  The
    timed output is each load case's element energy u' F (16 B per element);
    with the full forces written (96 B per element) the launches are
    store-bound and are reported beside it."""
    import torch
    from cudafemproject_amd import contraction_probe as CP
    from cudafemproject_amd.problem import plate_batch
    insts = plate_batch(512, seed=2024)
    coords, elems = CP.flat_plates(insts)
    m, n = len(elems), len(coords)
    rng = np.random.default_rng(11)
    u = rng.standard_normal((n, CP.N_CASES, 2)).astype(np.float32)
    E, nu = 2e7, 0.3
    lam, mu = E * nu / (1 - nu * nu), E / (2 * (1 + nu))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    te, tc, tu = dev(elems), dev(coords), dev(u)
    st = torch.cuda.current_stream()
    reps = max(args.matvec_reps, 20)

    def timed(v, out):
        CP.contract(v, te, tc, tu, lam, mu, out, reps=5)
        torch.cuda.synchronize()
        best = None
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            CP.contract(v, te, tc, tu, lam, mu, out, reps=reps)
            e1.record(st)
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) / reps * 1e3
            best = us if best is None else min(best, us)
        return best

    # energies u_case' F_case per element (16 B out per element): the contraction
    g = [torch.empty((m, CP.N_CASES), device="cuda") for _ in range(2)]
    t = {v: timed(v, g[v]) for v in (0, 1)}
    # full forces (96 B out per element): store-bound
    f = [torch.empty((m, CP.N_CASES, 6), device="cuda") for _ in range(2)]
    tf = {v: timed(v, f[v]) for v in (0, 1)}
    a, b = f[0].cpu().numpy(), f[1].cpu().numpy()
    diff = float((np.abs(a - b) / np.abs(a).max(axis=2, keepdims=True)).max())
    log(f"{m} triangles x {CP.N_CASES} load cases, energies out: on the fly {t[0]:.2f} us, "
        f"K_e + MFMA {t[1]:.2f} us; forces out: {tf[0]:.2f} / {tf[1]:.2f} us; "
        f"max relative force difference {diff:.2e}")
    # bytes a launch must move: connectivity, coordinates and load vectors
    # once each, energies written
    byt = 12 * m + 8 * n + 4 * CP.N_CASES * 2 * n + 4 * CP.N_CASES * m
    line = {
        "metric": "stacked-load-case element contraction F_e = K_e U_e (batch plates' triangles, "
                  "4 load cases): time per launch",
        "value": t[0], "unit": "us", "n_gpus": 1, "steps": 3, "warmup": 5,
        "ms_per_step": t[0] * 1e-3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "the 512 plates of the batch workload (seed 2024) as one mesh, "
                               "random load vectors", "n_elems": m, "n_nodes": n,
                   "n_cases": CP.N_CASES},
        "on_the_fly_us": t[0], "ke_mfma_us": t[1], "mfma_over_on_the_fly": t[1] / t[0],
        "forces_out_on_the_fly_us": tf[0], "forces_out_ke_mfma_us": tf[1],
        "max_rel_force_diff": diff,
        "roofline": {"bound": "hbm", "achieved": byt / (t[0] * 1e-6) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": byt / (t[0] * 1e-6) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_contract_valu", "bytes_per_launch": byt},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_cube_mixed_only(args, F, ctx, tune, run):
    """--only mixed (for profilers): the mixed solves and the FP32 matvec at
    the mixed solve's tuning, in a process of their own."""
    tune(F.MIXED, "mixed")
    for _ in range(args.warmup):
        run(F.MIXED, 1)
    rmx = run(F.MIXED, args.steps)
    ctx.sync()
    tmx = np.mean([s["solve_ms"] for _, s in rmx])
    mv32 = ctx.time_matvec(F.FP32, args.matvec_reps)
    log(f"mixed only: {tmx:.2f} ms, {rmx[-1][1]['iterations']} it; matvec fp32 {mv32 * 1e3:.2f} us")


def traffic_probe(args, F):
    """The command profiled for the HBM traffic figures: 1 + 5 FP32 cube
    matvecs (at --patch, else the default tuning), then 3 device copies of a
    256 MiB buffer as the counters' calibration (run under rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE)."""
    import torch
    from cudafemproject_amd.problem import cube_problem
    pb = cube_problem(args.n, E=2e7, nu=0.25, pressure=1e3, convention=F.LOADS_PHYSICAL)
    ctx = F.Context(0)
    if args.patch:
        ctx.set_tuning(args.patch[1], args.patch[0])
    ctx.set_problem(pb)
    ctx.matvec_repeat(F.FP32, 1)
    ctx.sync()
    ctx.matvec_repeat(F.FP32, 5)
    ctx.sync()
    x = torch.ones(64 * 1024 * 1024, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log("traffic probe done")
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["cube", "column", "implicit", "cantilever", "batch",
                                           "contraction"],
                    default="cube",
                    help="cube: static solve (default, headline); column: 20,000-step explicit "
                         "dynamics; implicit: 200 implicit Newmark steps on the same column; "
                         "cantilever: compliance + adjoint gradient; batch: 512 "
                         "plates x 4 load cases in one launch")
    ap.add_argument("--dyn-steps", type=int, default=20000)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the FP32 matvec + copy sequence profiled for HBM traffic")
    ap.add_argument("--fd-elems", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default patch tuning instead of the autotuners "
                         "(femgpu_autotune for the cube, femgpu_autotune_dyn for the "
                         "column and implicit workloads)")
    ap.add_argument("--patch", type=parse_patch, default=None, metavar="THREADSxELEMENTS",
                    help="cube / traffic probe: this patch tuning for every precision")
    ap.add_argument("--patch-fp64", type=parse_patch, default=None, metavar="THREADSxELEMENTS",
                    help="cube: the FP64 solves' patch tuning")
    ap.add_argument("--patch-mixed", type=parse_patch, default=None, metavar="THREADSxELEMENTS",
                    help="cube: the mixed solves' patch tuning")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="cube, for profilers: one precision's solves and matvec in this "
                         "process, no JSON line")
    ap.add_argument("--no-persistent", action="store_true",
                    help="batch workload: skip the persistent-schedule comparison")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="hexes per cube edge")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy-probe", action="store_true",
                    help="skip the device-copy bandwidth child process")
    ap.add_argument("--no-single-reduction", action="store_true",
                    help="cube: skip the single-reduction recurrence comparison")
    ap.add_argument("--matvec-reps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import cudafemproject_amd as F
    from cudafemproject_amd.problem import cube_problem
    if args.traffic_probe:
        return traffic_probe(args, F)
    if args.workload == "column":
        return bench_column(args, F)
    if args.workload == "implicit":
        return bench_implicit(args, F)
    if args.workload == "cantilever":
        return bench_cantilever(args, F)
    if args.workload == "batch":
        return bench_batch(args, F)
    if args.workload == "contraction":
        return bench_contraction(args, F)

    t0 = time.perf_counter()
    pb = cube_problem(args.n, E=2e7, nu=0.25, pressure=1e3, convention=F.LOADS_PHYSICAL)
    ctx = F.Context(0)
    ctx.set_problem(pb)
    log(f"setup {time.perf_counter() - t0:.2f} s: {pb.n_elems} elems, {pb.n_nodes} nodes, "
        f"{pb.n_dof} dof")
    # Each precision runs at the patch tuning the library's autotuner picks
    # for it (femgpu_autotune: timed PCG iterations per candidate, FP32 ones
    # for the mixed solve); --no-autotune keeps the default; --patch,
    # --patch-fp64, --patch-mixed set them.
    tuning = {}
    fixed = {"fp64": args.patch_fp64 or args.patch, "mixed": args.patch_mixed or args.patch}

    def tune(prec, name):
        if fixed[name]:
            tuning[name] = {"patch_threads": fixed[name][0], "elems_per_patch": fixed[name][1]}
            ctx.set_tuning(fixed[name][1], fixed[name][0])
            ctx.set_problem(pb)
            return
        if args.no_autotune:
            return
        t = time.perf_counter()
        tuning[name] = ctx.autotune(prec)
        log(f"{name}: autotuned patches {tuning[name]} in {time.perf_counter() - t:.1f} s")

    def run(prec, k, rec=F.CG_STANDARD):
        out = []
        for _ in range(k):
            u, st = ctx.solve(prec, tol=1e-8, tol_kind=F.TOL_REL_RES, recurrence=rec)
            out.append((u, st))
        return out

    b32 = bytes_per_matvec(pb.n_elems, pb.n_nodes, pb.n_dof, 4)
    b64 = bytes_per_matvec(pb.n_elems, pb.n_nodes, pb.n_dof, 8)
    # FP64: both recurrences and the FP64 matvec at the FP64 tuning
    if args.only == "mixed":
        return bench_cube_mixed_only(args, F, ctx, tune, run)
    tune(F.FP64, "fp64")
    for _ in range(args.warmup):
        run(F.FP64, 1)
    ctx.sync()
    r64 = run(F.FP64, args.steps)
    ctx.sync()
    t64 = np.mean([s["solve_ms"] for _, s in r64]) / 1e3
    u64, s64 = r64[-1]
    log(f"fp64: {t64 * 1e3:.2f} ms, {s64['iterations']} it, rel_res {s64['rel_res']:.2e}")
    # the same solves with the single-reduction recurrence (one launch per
    # iteration); the FP64 baseline of the speedup is the faster of the two
    t64s = tmxs = None
    it64s = itmxs = None
    if not args.no_single_reduction:
        run(F.FP64, 1, F.CG_SINGLE_REDUCTION)
        r64s = run(F.FP64, args.steps, F.CG_SINGLE_REDUCTION)
        t64s = float(np.mean([s["solve_ms"] for _, s in r64s]) / 1e3)
        it64s = r64s[-1][1]["iterations"]
    mv64 = ctx.time_matvec(F.FP64, args.matvec_reps) / 1e3
    if args.only == "fp64":
        log(f"fp64 only: matvec fp64 {mv64 * 1e6:.2f} us")
        return
    # mixed: both recurrences and the FP32 matvec (the roofline kernel) at
    # the mixed solve's tuning
    tune(F.MIXED, "mixed")
    for _ in range(args.warmup):
        run(F.MIXED, 1)
    ctx.sync()
    t_wall = time.perf_counter()
    rmx = run(F.MIXED, args.steps)
    ctx.sync()
    t_wall = time.perf_counter() - t_wall
    tmx = np.mean([s["solve_ms"] for _, s in rmx]) / 1e3
    umx, smx = rmx[-1]
    err = float(np.linalg.norm(umx - u64) / np.linalg.norm(u64))
    log(f"mixed: {tmx * 1e3:.2f} ms, {smx['iterations']} it ({smx['outer_iterations']} outer), "
        f"rel_res {smx['rel_res']:.2e}, |u_mixed-u_fp64|/|u_fp64| {err:.2e}")
    if not args.no_single_reduction:
        run(F.MIXED, 1, F.CG_SINGLE_REDUCTION)
        rmxs = run(F.MIXED, args.steps, F.CG_SINGLE_REDUCTION)
        tmxs = float(np.mean([s["solve_ms"] for _, s in rmxs]) / 1e3)
        itmxs = rmxs[-1][1]["iterations"]
        log(f"single-reduction recurrence: fp64 {t64s * 1e3:.2f} ms ({it64s} it), "
            f"mixed {tmxs * 1e3:.2f} ms ({itmxs} it)")
    t64_best = min(t64, t64s) if t64s else t64
    mv32 = ctx.time_matvec(F.FP32, args.matvec_reps) / 1e3
    log(f"matvec fp32 {mv32 * 1e6:.2f} us ({b32 / mv32 / 1e9:.0f} GB/s), "
        f"fp64 {mv64 * 1e6:.2f} us ({b64 / mv64 / 1e9:.0f} GB/s)")

    t32 = ((tuning["mixed"]["patch_threads"], tuning["mixed"]["elems_per_patch"])
           if "mixed" in tuning else (256, 768))
    traffic = read_traffic(traffic_profile(*t32))
    achieved = b32 / mv32 / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": f"EbE matvec fp32 (k_ebe_patch + k_iface_gather), cube64, "
                          f"{t32[0]}-thread patches of {t32[1]} elements",
                "bytes_per_launch": b32}

    cgbs = None if args.no_copy_probe else copy_gbs()
    if cgbs:
        log(f"device copy {cgbs:.0f} GB/s; fp32 matvec at {achieved / cgbs:.3f} of it")

    cpu = cpu_eff = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(pb, s64["iterations"], args.cpu_iters)
        log(f"cpu baseline: {cpu['value']:.1f} s ({cpu['sample']})")
        cpu_eff = cpu_efficient_baseline(pb, s64["iterations"], args.cpu_iters)
        log(f"efficient cpu baseline: {cpu_eff['value']:.1f} s ({cpu_eff['sample']})")

    line = {
        "metric": "3D PCG time-to-solve @ rel-res 1e-8, mixed vs FP64; EbE matvec HBM GB/s",
        "value": float(tmx), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": float(tmx * 1e3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"cube {args.n}^3 hexes x 6 Kuhn tets, z=0 clamped, p=1e3 on z=1, "
                               "static PCG to rel-res 1e-8",
                   "n_elems": pb.n_elems, "n_nodes": pb.n_nodes, "n_dof": pb.n_dof,
                   "tol": 1e-8, "solver": "mixed (fp32 pcg in fp64 refinement)",
                   "patch_tuning": ({k: {"threads": v["patch_threads"],
                                         "elements": v["elems_per_patch"]}
                                     for k, v in tuning.items()}
                                    or {"threads": 256, "elements": 768})},
        "fp64_time_to_solve_s": float(t64),
        "speedup_mixed_vs_fp64": float(t64_best / tmx),
        # the same (standard) recurrence in both precisions
        "speedup_mixed_vs_fp64_standard_recurrence": float(t64 / tmx),
        "fp64_single_reduction_time_to_solve_s": t64s,
        "mixed_single_reduction_time_to_solve_s": tmxs,
        "iterations_single_reduction_fp64": it64s,
        "iterations_single_reduction_mixed": itmxs,
        "mixed_vs_fp64_rel_displacement_diff": err,
        "iterations_fp64": s64["iterations"], "iterations_mixed": smx["iterations"],
        "outer_iterations_mixed": smx["outer_iterations"],
        "rel_res_fp64": s64["rel_res"], "rel_res_mixed": smx["rel_res"],
        "pcg_it_per_s_fp64": s64["iterations"] / t64, "pcg_it_per_s_mixed": smx["iterations"] / tmx,
        "wall_s_per_mixed_solve_incl_u_to_host": t_wall / args.steps,
        "matvec_us_fp32": mv32 * 1e6, "matvec_us_fp64": mv64 * 1e6,
        "matvec_gbs_fp32": b32 / mv32 / 1e9, "matvec_gbs_fp64": b64 / mv64 / 1e9,
        "dof_matvecs_per_s_fp32": pb.n_dof / mv32, "dof_matvecs_per_s_fp64": pb.n_dof / mv64,
        "copy_gbs": cgbs, "matvec_frac_of_copy_fp32": achieved / cgbs if cgbs else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_efficient_baseline": cpu_eff,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
