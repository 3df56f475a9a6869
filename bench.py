"""Benchmark of the IMEX-DG hot path on one MI355X.

Workload: 2D nonhydrostatic flow over a Witch-of-Agnesi hill, 512x100 terrain-following
cells, degree 4 (N = 1.28 M nodes per field, 5.12 M DoF in rho, u, w, p), hydrostatic
constant-N background plus seeded N(0, 0.01 K) theta noise, dt = 1 s, Schur GMRES tolerance
1e-10, fixed-point tolerance 1e-10. A "step" is one full IMEX-ARK2 step on the device.

Prints one JSON line: IMEX-step DoF/s (value), isolated explicit-stage and Schur applies,
the roofline of the dominant kernel (HIP events on the library's stream) and the CPU
baseline (the oracle's FP64 restatement on the host cores, bounded sample).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FR = 1.0 / 9.80616 ** 0.5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reps", type=int, default=200, help="isolated-apply repetitions")
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--degree", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-cols", type=int, default=64, help=argparse.SUPPRESS)
    return ap.parse_args()


def workload(nx, nz, k):
    from atmospheric_flows_amd import make_mesh, make_hill_state, make_solver_params
    mesh = make_mesh(dim=2, degree=k, n_cells=(nx, 1, nz), x_range=(-72e3, 72e3), z_top=25e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=2024)
    sp = make_solver_params(gmres_eps=1e-10, gmres_restart=28, fp_tol=1e-10, fp_max_iter=100)
    return mesh, st, sp


def cpu_baseline(args):
    """Oracle (FP64 restatement) IMEX step on a bounded sample of the same workload: a periodic
    slab of cpu_cols x nz cells at the same resolution (dx = 144 km / nx), timed on the host."""
    import numpy as np
    from atmospheric_flows_amd import make_mesh
    from oracle.pyoracle import OracleOperator
    nx = args.cpu_cols
    half = 72e3 * nx / args.nx
    _, st, sp = workload(args.nx, args.nz, args.degree)
    mesh = make_mesh(dim=2, degree=args.degree, n_cells=(nx, 1, args.nz), x_range=(-half, half), z_top=25e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    op = OracleOperator(mesh, mach=1.0, froude=FR, dt=1.0)
    op.set_solver_params(sp)
    rho, u, p = op.init_hill_state(st)
    t0 = time.perf_counter()
    op.imex_step(rho, u, p)
    dt = time.perf_counter() - t0
    dofs = 4 * op.N
    return {"value": dofs / dt, "unit": "DoF/s", "seconds": dt,
            "sample": f"1 IMEX step of the same hill problem on a {nx}x{args.nz}-cell periodic slab "
                      f"(k={args.degree}, {dofs} DoF, same dx, dz, dt and tolerances)"}


def run_cpu_subprocess(args, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-only", "--nx", str(args.nx), "--nz", str(args.nz),
           "--degree", str(args.degree), "--cpu-cols", str(args.cpu_cols)]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        return {"value": None, "unit": "DoF/s", "cores": threads, "kind": "port", "error": r.stderr[-400:]}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out.update({"cores": threads, "kind": "port"})
    return out


def profiled_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of this workload (calibrated against an L2-cold copy of known size); None if not profiled."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{args.nx}x{args.nz}_k{args.degree}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel + ":"):
            for tok in line.split():
                if tok.startswith("bytes_per_launch="):
                    return float(tok.split("=")[1])
    return None


def main():
    args = parse()
    if args.cpu_only:
        print(json.dumps(cpu_baseline(args)))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    import torch
    from atmospheric_flows_amd import EulerOperator
    mesh, st, sp = workload(args.nx, args.nz, args.degree)
    op = EulerOperator(mesh, mach=1.0, froude=FR, dt=1.0)
    op.set_solver_params(sp)
    N, dim = op.N, op.dim
    rho, u, p = op.init_hill_state(st)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3  # seconds per launch

    # calibration for the rocprofv3 FETCH_SIZE / WRITE_SIZE passes: an L2-cold scale-by-one of
    # 4N doubles (torch elementwise kernel; reads and writes exactly 32 N bytes)
    cal_src = torch.rand(4 * N, device="cuda", dtype=torch.float64)
    flush = torch.empty(16 * 1024 * 1024, device="cuda", dtype=torch.float64).fill_(1.0)
    cal_dst = torch.mul(cal_src, 1.0)
    torch.cuda.synchronize()
    del cal_src, cal_dst, flush

    # isolated applies on the initial state (stage-3 explicit applies, Schur pieces)
    r2, u2, p2 = rho.clone(), u.clone(), p.clone()
    dst_s, dst_v = op.new_scalar(), op.new_vector()
    op.set_HYPERBOLIC_stage(3)
    src_u = [rho, u, p, r2, u2, p2, rho, u, p]
    t_e2 = timed(lambda: op.vmult_rhs_velocity_fixed(dst_v, src_u), args.reps)
    src_p = src_u + [r2, u2]
    t_e3 = timed(lambda: op.vmult_rhs_pressure(dst_s, src_p), args.reps)
    src_r = [rho, u, r2, u2, rho, u]
    t_e1 = timed(lambda: op.vmult_rhs_rho_update(dst_s, src_r), args.reps)
    op.set_HYPERBOLIC_stage(1)
    op.set_rho_for_fixed(rho)
    op.set_pres_fixed(p)
    op.set_NS_stage(2)
    t_schur = timed(lambda: op.vmult(dst_s, p), args.reps)
    t_sw = timed(lambda: op.vmult_schur_velocity(dst_v, p), args.reps)
    # the FP32 level operator of the preconditioner: w = A^-1 B p in single precision (the kernel
    # that dominates the FP32-preconditioned step) and one full V-cycle
    from atmospheric_flows_amd import make_solver_params
    from atmospheric_flows_amd._abi import AFL_SCHUR_JACOBI, AFL_SCHUR_MG_FP32, AFL_TIME_LEVEL_W, AFL_TIME_LEVEL_C_CHEB
    op.set_solver_params(make_solver_params(gmres_eps=1e-10, gmres_restart=28, fp_tol=1e-10,
                                            schur_precond=AFL_SCHUR_MG_FP32))
    p32 = p.float()
    w32 = torch.empty(dim * N, device="cuda", dtype=torch.float32)
    t_sw32 = timed(lambda: op.vmult_schur_velocity_f32(w32, p32), args.reps)
    t_mg = timed(lambda: op.mg_vmult(dst_s, p), max(1, args.reps // 10))

    # IMEX steps: FP64 plain path (Jacobi GMRES) and FP32 p-multigrid FGMRES, same initial
    # state, same true-residual tolerance 1e-10

    def run_path(precond):
        o = EulerOperator(mesh, mach=1.0, froude=FR, dt=1.0)
        o.set_solver_params(make_solver_params(gmres_eps=1e-10, gmres_restart=28, fp_tol=1e-10,
                                               schur_precond=precond))
        r0, u0, p0 = o.init_hill_state(st)
        infos = []
        for _ in range(args.warmup):
            o.imex_step(r0, u0, p0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            infos.append(o.imex_step(r0, u0, p0))
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps, infos, (r0, u0, p0), o

    step_fp64, infos64, s64, _ = run_path(AFL_SCHUR_JACOBI)
    step_s, infos, s32, o32 = run_path(AFL_SCHUR_MG_FP32)
    dev = {name: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
           for name, a, b in (("rho", s32[0], s64[0]), ("u", s32[1], s64[1]), ("p", s32[2], s64[2]))}
    dofs = (2 + dim) * N

    # per-kernel time of the two multigrid level kernels inside the same FP32 steps (HIP events
    # recorded by the library on its stream around every launch; a second pass of args.steps
    # steps continuing from the timed state)
    ncells = N // (args.degree + 1) ** dim
    level = {}
    for kid, name, fields in ((AFL_TIME_LEVEL_W, "k_schur_w_lvl<float>", 1 + dim),
                              (AFL_TIME_LEVEL_C_CHEB, "k_apply_C_cheb<float>", 7 + dim)):
        o32.kernel_timer_start(kid)
        for _ in range(args.steps):
            o32.imex_step(*s32)
        n, ms = o32.kernel_timer_stop()
        # algorithmic bytes: each distinct field once (FP32), plus one 4-byte scale per cell
        level[name] = {"launches_per_step": n / args.steps, "avg_us": 1e3 * ms / max(n, 1),
                       "ms_per_step": ms / args.steps, "bytes_per_launch": fields * N * 4 + ncells * 4}
    # gradient of a final-time functional <c, x_T> with respect to x_0 over the same number of
    # steps (binomial checkpointing with two snapshots; FP32-preconditioned solves, also in the
    # adjoint stage solves), against the forward run of those steps
    gen = torch.Generator(device="cuda").manual_seed(1)
    cbar = (torch.randn(N, dtype=torch.float64, device="cuda", generator=gen),
            torch.randn(dim * N, dtype=torch.float64, device="cuda", generator=gen),
            1e-5 * torch.randn(N, dtype=torch.float64, device="cuda", generator=gen))
    x0 = tuple(t.clone() for t in s32)
    adjoint = []
    # two snapshots, a snapshot per step (each reverse step recomputes its own step), and every
    # step's stage vectors kept from one forward sweep (-1: no recomputation)
    for snaps in sorted({min(2, args.steps - 1), args.steps - 1}) + [-1]:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        *_, adj_stats = o32.adjoint_run(*x0, args.steps, snaps, *cbar)
        torch.cuda.synchronize()
        t_adj = time.perf_counter() - t0
        adjoint.append({"steps": args.steps, "snapshots": snaps, "ms": t_adj * 1e3,
                        "forward_run_ms": step_s * args.steps * 1e3,
                        "cost_vs_forward_run": t_adj / (step_s * args.steps), "schedule": adj_stats})
    dom = max(level, key=lambda k: level[k]["ms_per_step"])
    bytes_dom = level[dom]["bytes_per_launch"]
    ach = bytes_dom / (level[dom]["avg_us"] * 1e-6) / 1e9
    field_bytes = N * 8
    res = {
        "metric": "IMEX step wall-time, FP32-MG vs FP64 path at equal residual; DoF/s, % HBM peak",
        "value": dofs / step_s,
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D Agnesi hill {args.nx}x{args.nz} cells, k={args.degree}, FP64 state, "
                               f"FP32 p-multigrid FGMRES Schur solve, dt=1 s, true residual 1e-10",
                   "n_per_field": N, "dof": dofs, "solver": "fgmres+fp32-pmg"},
        "fp_iters_per_step": [list(i.fp_iters) for i in infos],
        "gmres_iters_per_step": [list(i.gmres_iters) for i in infos],
        "fp64_plain_path": {"ms_per_step": step_fp64 * 1e3, "value": dofs / step_fp64,
                            "gmres_iters_per_step": [list(i.gmres_iters) for i in infos64],
                            "fp_iters_per_step": [list(i.fp_iters) for i in infos64]},
        "speedup_fp32mg_vs_fp64": step_fp64 / step_s,
        "rel_l2_fp32mg_vs_fp64_after_steps": dev,
        "isolated_applies_us": {"E1_stage3": t_e1 * 1e6, "E2_stage3": t_e2 * 1e6, "E3_stage3": t_e3 * 1e6,
                                "schur_vmult": t_schur * 1e6, "schur_velocity_kernel": t_sw * 1e6,
                                "schur_velocity_kernel_fp32": t_sw32 * 1e6, "fp32_vcycle": t_mg * 1e6},
        "isolated_applies_GBs": {"E1_stage3": 10 * field_bytes / t_e1 / 1e9,
                                 "E2_stage3": 14 * field_bytes / t_e2 / 1e9,
                                 "E3_stage3": 16 * field_bytes / t_e3 / 1e9,
                                 "schur_vmult": 9 * field_bytes / t_schur / 1e9,
                                 "schur_velocity_kernel": (2 + dim) * field_bytes / t_sw / 1e9},
        "level_kernels_in_step": level,
        "adjoint_gradient": adjoint,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": ach, "bytes_per_launch": bytes_dom,
                     "avg_launch_us": level[dom]["avg_us"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(dom.replace("<float>", f"<{dim}, {args.degree}, float, {args.degree + 1}>"), args)},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = run_cpu_subprocess(args, threads)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
