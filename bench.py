"""Benchmark of the IMEX-DG hot path on one MI355X.

Headline workload (the mixed-precision 3D configuration): isolated Agnesi hill
h = 400 m / (1 + (r / 1 km)^2) in a 96x96x40-cell terrain-following box of 144 x 144 x 25 km,
degree 3 (N = 23.6 M nodes per field, 118 M DoF in rho, u, v, w, p), hydrostatic constant-N
background (N = 0.01 1/s, theta0 = 280 K, p0 = 1e5 Pa, u = 10 m/s) plus seeded N(0, 0.01 K) theta
noise, dt = 1 s. A "step" is one full IMEX-ARK2 step on the device. The FP64 state, explicit stages
and outer flexible GMRES (true residual 1e-8) are preconditioned by an FP32 multigrid V-cycle; the
all-FP64 Jacobi-GMRES path and the FP64 V-cycle are timed on the same inputs, and the theta' and w
differences of the FP32-preconditioned state from the all-FP64 one are reported.

Further legs: isolated operator applies on the headline mesh (HIP events on the library's stream);
per-kernel HIP-event timing inside the FP32-preconditioned steps (the roofline object is quoted for
the kernel with the most time per step); the 2D 512x100 degree-4 hill (FP32-MG vs FP64 steps,
isolated applies over buffer sets rotated beyond the 256 MB Infinity Cache); the adjoint gradient on
the 2D 256x60 degree-3 hill (200 steps, 20 snapshots and every-step storage, against the forward
run); and the CPU baseline, the oracle's FP64 restatement on the host cores (bounded sample).

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FR = 1.0 / 9.80616 ** 0.5
R_GAS, CP = 287.0, 1.4 * 287.0 / 0.4
HEADLINE = "c3"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed IMEX steps per path")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reps", type=int, default=20, help="isolated-apply repetitions")
    ap.add_argument("--legs", default="c3,applies,kernels,c2,adjoint,mfma",
                    help="comma list of c3 (the headline; applies and kernels need it), applies, kernels, c2, adjoint, mfma, c5")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-snapshots", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate", action="store_true",
                    help="first scale a 4N-double buffer by one in place (a kernel of known HBM traffic for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE calibration)")
    ap.add_argument("--cpu-only", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


def c3_workload():
    from atmospheric_flows_amd import make_mesh, make_hill_state
    mesh = make_mesh(dim=3, degree=3, n_cells=(96, 96, 40), x_range=(-72e3, 72e3), y_range=(-72e3, 72e3),
                     z_top=25e3, hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=2024)
    return mesh, st


def c2_workload():
    from atmospheric_flows_amd import make_mesh, make_hill_state
    mesh = make_mesh(dim=2, degree=4, n_cells=(512, 1, 100), x_range=(-72e3, 72e3), z_top=25e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=2024)
    return mesh, st


def params(precond, eps, restart=28):
    from atmospheric_flows_amd import make_solver_params
    return make_solver_params(gmres_eps=eps, gmres_restart=restart, fp_tol=1e-10, fp_max_iter=100,
                              schur_precond=precond)


# ---------------------------------------------------------------------------------------------
# CPU baseline: the oracle (FP64 restatement, OpenMP over cells) on a bounded sample of the headline
# workload, timed on the host cores in a subprocess
def cpu_sample():
    import numpy as np
    from atmospheric_flows_amd import make_mesh
    from atmospheric_flows_amd._abi import AFL_SCHUR_JACOBI
    from oracle.pyoracle import OracleOperator
    mesh, st = c3_workload()
    # an 8x8x40-cell periodic column of the same mesh spacing (dx = dy = 1.5 km, 40 levels)
    half = 72e3 * 8 / 96
    sample = make_mesh(dim=3, degree=3, n_cells=(8, 8, 40), x_range=(-half, half), y_range=(-half, half),
                       z_top=25e3, hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    op = OracleOperator(sample, mach=1.0, froude=FR, dt=1.0)
    op.set_solver_params(params(AFL_SCHUR_JACOBI, 1e-8))
    rho, u, p = op.init_hill_state(st)
    t0 = time.perf_counter()
    op.imex_step(rho, u, p)
    dt = time.perf_counter() - t0
    dofs = 5 * op.N
    assert np.isfinite(p).all()
    return {"value": dofs / dt, "unit": "DoF/s", "seconds": dt,
            "sample": f"1 IMEX step (plain FP64 Jacobi-GMRES, tolerance 1e-8) of the same 3D hill problem on an "
                      f"8x8x40-cell periodic column (k=3, {dofs} DoF, same dx, dy, dz, dt)"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_cpu_subprocess():
    threads = min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-only"], capture_output=True, text=True,
                       env=env, timeout=900)
    base = {"unit": "DoF/s", "cores": threads, "kind": "port", "host_nproc": os.cpu_count(),
            "cpu_model": cpu_model(),
            "build": "g++ -O3 -march=x86-64-v4 (AVX-512) -fopenmp -ffp-contract=off, OpenMP over cells"}
    if r.returncode != 0:
        base.update({"value": None, "error": r.stderr[-400:]})
        return base
    base.update(json.loads(r.stdout.strip().splitlines()[-1]))
    return base


# ---------------------------------------------------------------------------------------------
def main():
    args = parse()
    if args.cpu_only:
        print(json.dumps(cpu_sample()))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    import torch
    from atmospheric_flows_amd import EulerOperator
    from atmospheric_flows_amd import _abi as A

    stream = torch.cuda.current_stream()

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    def events_time(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(reps):
            fn(i)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3  # seconds per call

    def run_path(mesh, st, precond, eps, steps, warmup):
        o = EulerOperator(mesh, mach=1.0, froude=FR, dt=1.0)
        o.set_solver_params(params(precond, eps))
        r0, u0, p0 = o.init_hill_state(st)
        for _ in range(warmup):
            o.imex_step(r0, u0, p0)
        torch.cuda.synchronize()
        infos = []
        t0 = time.perf_counter()
        for _ in range(steps):
            infos.append(o.imex_step(r0, u0, p0))
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, infos, (r0, u0, p0), o

    def theta_prime(o, st, rho, p):
        _, _, th_bar = o.background(st)
        th = p / (rho * R_GAS) * torch.pow(st.p0 / p, R_GAS / CP)
        return th - th_bar

    def rel_l2(o, ncomp, a, b):
        e, r = o.l2_difference(ncomp, a, b)
        return e / r

    res = {}
    if "c3" in legs:  # the headline (omitted only for runs that profile other legs)
        # ---- headline: 3D 96x96x40, k=3, FP32-MG FGMRES vs FP64 Jacobi GMRES vs FP64-MG FGMRES -------------
        mesh, st = c3_workload()
        N = A.n_per_field(mesh)
        dim, dofs = 3, 5 * N
        if args.calibrate:  # reads and writes exactly 4 N doubles (755 MB, beyond the 256 MB Infinity Cache)
            buf = torch.ones(4 * N, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            buf.mul_(1.0)
            torch.cuda.synchronize()
            del buf
        paths = {}
        for name, pc in (("fp32_mg", A.AFL_SCHUR_MG_FP32), ("fp64_jacobi", A.AFL_SCHUR_JACOBI),
                         ("fp64_mg", A.AFL_SCHUR_MG_FP64)):
            log(f"c3 path {name}: {args.warmup} warm-up + {args.steps} timed steps")
            t, infos, state, o = run_path(mesh, st, pc, 1e-8, args.steps, args.warmup)
            paths[name] = {"t": t, "infos": infos, "state": state, "op": o}
            if name != "fp32_mg":
                del o
                paths[name]["op"] = None
        o32 = paths["fp32_mg"]["op"]
        ref = paths["fp64_jacobi"]["state"]
        th_ref = theta_prime(o32, st, ref[0], ref[2])
        errs = {}
        for name in ("fp32_mg", "fp64_mg"):
            s = paths[name]["state"]
            th = theta_prime(o32, st, s[0], s[2])
            errs[name] = {"theta_prime": rel_l2(o32, 1, th, th_ref),
                          "w": rel_l2(o32, 1, s[1][2 * N:].contiguous(), ref[1][2 * N:].contiguous()),
                          "rho": rel_l2(o32, 1, s[0], ref[0]), "u": rel_l2(o32, dim, s[1], ref[1]),
                          "p": rel_l2(o32, 1, s[2], ref[2])}
        step_s = paths["fp32_mg"]["t"]
        res.update({
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
            "config": {"workload": "3D Agnesi hill 96x96x40 cells (144x144x25 km), k=3, FP64 state and explicit stages, "
                                   "FP64 flexible GMRES (true residual 1e-8) with an FP32 p-multigrid Chebyshev V-cycle, "
                                   "dt=1 s, seeded N(0, 0.01 K) theta noise",
                       "n_per_field": N, "dof": dofs, "solver": "fgmres+fp32-mg"},
            "fp_iters_per_step": [list(i.fp_iters) for i in paths["fp32_mg"]["infos"]],
            "gmres_iters_per_step": [list(i.gmres_iters) for i in paths["fp32_mg"]["infos"]],
            "fp64_plain_path": {"ms_per_step": paths["fp64_jacobi"]["t"] * 1e3, "value": dofs / paths["fp64_jacobi"]["t"],
                                "gmres_iters_per_step": [list(i.gmres_iters) for i in paths["fp64_jacobi"]["infos"]]},
            "fp64_mg_path": {"ms_per_step": paths["fp64_mg"]["t"] * 1e3, "value": dofs / paths["fp64_mg"]["t"],
                             "gmres_iters_per_step": [list(i.gmres_iters) for i in paths["fp64_mg"]["infos"]]},
            "speedup_fp32mg_vs_fp64": paths["fp64_jacobi"]["t"] / step_s,
            "speedup_fp32mg_vs_fp64mg": paths["fp64_mg"]["t"] / step_s,
            "rel_l2_vs_fp64_path_after_steps": {**errs, "steps": args.steps + args.warmup},
        })
        for name in ("fp64_jacobi", "fp64_mg"):
            paths[name]["state"] = None

        # ---- isolated applies on the headline mesh (every working set is far beyond the Infinity Cache) ----
        if "applies" in legs:
            log("c3 isolated applies")
            rho, u, p = paths["fp32_mg"]["state"]
            o = o32
            dst_s, dst_v = o.new_scalar(), o.new_vector()
            o.set_HYPERBOLIC_stage(3)
            r2, u2, p2 = rho.clone(), u.clone(), p.clone()
            src_u = [rho, u, p, r2, u2, p2, rho, u, p]
            ap = {}
            fb = N * 8
            ap["E2_stage3"] = (events_time(lambda i=0: o.vmult_rhs_velocity_fixed(dst_v, src_u), args.reps), 18 * fb)
            src_p = src_u + [r2, u2]
            ap["E3_stage3"] = (events_time(lambda i=0: o.vmult_rhs_pressure(dst_s, src_p), args.reps), 20 * fb)
            src_r = [rho, u, r2, u2, rho, u]
            ap["E1_stage3"] = (events_time(lambda i=0: o.vmult_rhs_rho_update(dst_s, src_r), args.reps), 13 * fb)
            o.set_HYPERBOLIC_stage(1)
            o.set_rho_for_fixed(rho)
            o.set_pres_fixed(p)
            o.set_NS_stage(2)
            ap["schur_vmult_fp64"] = (events_time(lambda i=0: o.vmult(dst_s, p), args.reps), 11 * fb)
            ap["schur_w_fp64"] = (events_time(lambda i=0: o.vmult_schur_velocity(dst_v, p), args.reps), 5 * fb)
            ap["apply_B"] = (events_time(lambda i=0: o.vmult_pressure(dst_v, p), args.reps), 4 * fb)
            ap["apply_C"] = (events_time(lambda i=0: o.vmult_enthalpy(dst_s, u), args.reps), 5 * fb)
            o.set_NS_stage(3)
            ap["A_solve"] = (events_time(lambda i=0: o.solve_mass(dst_v, u), args.reps), 7 * fb)
            p32, w32 = p.float(), torch.empty(dim * N, device="cuda", dtype=torch.float32)
            ap["schur_w_fp32_level"] = (events_time(lambda i=0: o.vmult_schur_velocity_f32(w32, p32), args.reps), 5 * N * 4)
            o.set_NS_stage(2)
            ap["fp32_vcycle"] = (events_time(lambda i=0: o.mg_vmult(dst_s, p), max(2, args.reps // 4)), None)
            res["isolated_applies_c3"] = {k: {"us": t * 1e6, **({"GBs": b / t / 1e9, "hbm_frac": b / t / 1e9 / HBM_PEAK_GBS}
                                                                 if b else {})} for k, (t, b) in ap.items()}
            del dst_s, dst_v, r2, u2, p2, p32, w32
            torch.cuda.synchronize()

        # ---- per-kernel times inside the FP32-preconditioned steps; roofline of the dominant kernel --------
        if "kernels" in legs:
            log("c3 per-kernel timing inside FP32-MG steps")
            kern, dom = kernels_in_steps(o32, A, paths["fp32_mg"]["state"], args.steps)
            res["kernels_in_fp32_step"] = kern
            d = kern[dom]
            res["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": d["GBs"], "bytes_per_launch": d["bytes_per_launch"],
                               "avg_launch_us": d["avg_us"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": d["GBs"] / HBM_PEAK_GBS, "traffic": profiled_traffic(dom),
                               "fp64_vector": profiled_fp64(dom)}
        res["fp32_step_state_finite"] = bool(torch.isfinite(paths["fp32_mg"]["state"][2]).all())
        del paths, o32
        torch.cuda.synchronize()

    # ---- 2D 512x100, k=4 (c2): steps and isolated applies over rotated buffer sets ------------------------
    if "c2" in legs:
        log("c2 leg")
        mesh2, st2 = c2_workload()
        t32, inf32, s32, o = run_path(mesh2, st2, A.AFL_SCHUR_MG_FP32, 1e-10, args.steps, args.warmup)
        t64, inf64, s64, _ = run_path(mesh2, st2, A.AFL_SCHUR_JACOBI, 1e-10, args.steps, args.warmup)
        N2 = o.N
        sets = 4  # E2 stage-3 sources: 9 fields x 10 MB per set, 4 sets = 368 MB > 256 MB Infinity Cache
        gen = torch.Generator(device="cuda").manual_seed(3)
        rot = [[t * (1 + 1e-6 * torch.randn(t.numel(), dtype=torch.float64, device="cuda", generator=gen))
                for t in s32] for _ in range(sets)]
        dsts, dstv = [o.new_scalar() for _ in range(sets)], [o.new_vector() for _ in range(sets)]
        fb = N2 * 8
        o.set_HYPERBOLIC_stage(3)
        ap2 = {}
        ap2["E2_stage3"] = (events_time(lambda i=0: o.vmult_rhs_velocity_fixed(
            dstv[i % sets], [*rot[i % sets], *rot[(i + 1) % sets], *rot[(i + 2) % sets]]), args.reps * 5), 14 * fb)
        ap2["E3_stage3"] = (events_time(lambda i=0: o.vmult_rhs_pressure(
            dsts[i % sets], [*rot[i % sets], *rot[(i + 1) % sets], *rot[(i + 2) % sets],
                             rot[(i + 3) % sets][0], rot[(i + 3) % sets][1]]), args.reps * 5), 16 * fb)
        o.set_HYPERBOLIC_stage(1)
        o.set_rho_for_fixed(s32[0])
        o.set_pres_fixed(s32[2])
        o.set_NS_stage(2)
        ap2["schur_vmult_fp64"] = (events_time(lambda i=0: o.vmult(dsts[i % sets], rot[i % sets][2]),
                                               args.reps * 5), 9 * fb)
        res["c2"] = {"workload": "2D Agnesi hill 512x100 cells, k=4, dt=1 s, true residual 1e-10",
                     "dof": 4 * N2, "fp32_mg_ms_per_step": t32 * 1e3, "fp64_ms_per_step": t64 * 1e3,
                     "value_fp32_mg": 4 * N2 / t32, "speedup_fp32mg_vs_fp64": t64 / t32,
                     "gmres_iters_per_step_fp32_mg": [list(i.gmres_iters) for i in inf32],
                     "gmres_iters_per_step_fp64": [list(i.gmres_iters) for i in inf64],
                     "isolated_applies_rotated": {k: {"us": t * 1e6, "GBs": b / t / 1e9} for k, (t, b) in ap2.items()}}
        del o, rot, dsts, dstv, s32, s64
        torch.cuda.synchronize()

    # ---- adjoint gradient on the 2D 256x60, k=3 hill --------------------------------------------------------
    if "adjoint" in legs:
        log(f"adjoint leg: {args.adjoint_steps} steps")
        res["adjoint"] = adjoint_leg(args, torch, EulerOperator, A)

    # ---- one GPU's share of the 3D 192x192x64, k=4 configuration: forward steps and adjoint gradient ----
    if "c5" in legs:
        log("c5 shard leg: 24x192x64 cells, k=4")
        res["c5_shard"] = c5_shard_leg(torch, EulerOperator, A, log)

    if "mfma" in legs:
        log("MFMA vs VALU contraction probe")
        res["mfma_probe"] = mfma_probe_leg(torch)

    if not args.no_cpu:
        log("CPU baseline")
        res["cpu_baseline"] = run_cpu_subprocess()
    print(json.dumps(res))


def kernels_in_steps(o, A, state, steps):
    """HIP-event time, launches and algorithmic bytes of every timed kernel over `steps` IMEX steps
    of operator o from `state`; returns the table and the kernel with the most time per step."""
    mask = 0
    for f in A.TIMED_KERNELS:
        mask |= f
    o.kernel_timer_start(mask, 2000 + 1000 * steps)
    for _ in range(steps):
        o.imex_step(*state)
    o.kernel_timer_stop()
    kern = {}
    for f, name in A.TIMED_KERNELS.items():
        n, ms, by = o.kernel_timer_result(f, 0)
        if n:
            kern[name] = {"launches_per_step": n / steps, "ms_per_step": ms / steps,
                          "avg_us": 1e3 * ms / n, "bytes_per_launch": by / n,
                          "GBs": by / (ms * 1e-3) / 1e9}
            for stg in (1, 2, 3):
                ns_, mss, bys = o.kernel_timer_result(f, stg)
                if ns_:
                    kern[name][f"stage{stg}"] = {"launches": ns_, "avg_us": 1e3 * mss / ns_,
                                                 "bytes_per_launch": bys / ns_, "GBs": bys / (mss * 1e-3) / 1e9}
    return kern, max(kern, key=lambda k: kern[k]["ms_per_step"])


def adjoint_leg(args, torch, EulerOperator, A):
    """J = 1/2 sum_k (theta'(x_k, T) - obs_k)^2 over 512 seeded-uniform points, obs from a seed-7
    perturbed initial state; gradient with respect to the initial state by binomial checkpointing
    and by every-step storage, against the forward run of the same steps."""
    import numpy as np
    from atmospheric_flows_amd import make_mesh, make_hill_state
    mesh = make_mesh(dim=2, degree=3, n_cells=(256, 1, 60), x_range=(-72e3, 72e3), z_top=25e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    st = make_hill_state(seed=2024)
    o = EulerOperator(mesh, mach=1.0, froude=FR, dt=1.0)
    o.set_solver_params(params(A.AFL_SCHUR_MG_FP32, 1e-10))
    rng = np.random.default_rng(11)
    pts = np.stack([rng.uniform(-70e3, 70e3, 512), rng.uniform(500.0, 20e3, 512)], axis=1)
    o.set_observation_points(pts)
    n = args.adjoint_steps
    truth = o.init_hill_state(make_hill_state(seed=7))
    for _ in range(n):
        o.imex_step(*truth)
    obs = o.observe_theta(st, truth[0], truth[2])
    x0 = o.init_hill_state(st)
    x = tuple(t.clone() for t in x0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        o.imex_step(*x)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    resid = o.observe_theta(st, x[0], x[2]) - obs
    rb, pb = o.observe_theta_transpose(st, x[0], x[2], resid)
    ub = torch.zeros_like(x[1])
    out = {"workload": f"2D Agnesi hill 256x60 cells, k=3, {n} steps at dt=1 s, 512 observation points, "
                       "FP32-MG preconditioned stage solves", "forward_run_s": t_fwd, "modes": []}
    for snaps in (args.adjoint_snapshots, A_STORE_ALL):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = o.adjoint_run(*x0, n, snaps, rb, ub, pb)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out["modes"].append({"snapshots": snaps, "seconds": t, "cost_vs_forward_run": t / t_fwd,
                             "schedule": g[3], "gradient_finite": bool(torch.isfinite(g[2]).all())})
    return out


A_STORE_ALL = -1


def mfma_probe_leg(torch, reps=10000, waves=8192):
    """FP64 throughput of the 3D degree-3 sum-factorisation contraction (the 7x4 value and derivative
    tables on Gauss-Lobatto nodes -> 7 Gauss points, applied to the 4 nodal values of a line) on the
    vector ALUs (v_fma_f64, lane = line) and on the matrix cores (v_mfma_f64_16x16x4f64, [S; D] as the
    16x4 A operand with 14 rows used, 16 lines per instruction), from libafl_probe.so. The MFMA output
    placement is checked on known operands first."""
    import ctypes as C
    import numpy as np
    lib = C.CDLL(os.path.join(ROOT, "atmospheric_flows_amd", "libafl_probe.so"))
    vp = C.c_void_p
    lib.afl_probe_set_tables.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.afl_probe_contraction.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, C.POINTER(C.c_float)]
    lib.afl_probe_mfma_layout.argtypes = [vp, vp, vp]
    # k = 3 tables: Lagrange basis on the Gauss-Lobatto nodes of [0, 1], values and derivatives at the
    # 7 Gauss points
    xn = np.array([0.0, 0.5 - 0.5 / 5 ** 0.5, 0.5 + 0.5 / 5 ** 0.5, 1.0])
    xq = 0.5 * (np.polynomial.legendre.leggauss(7)[0] + 1.0)
    S = np.zeros((7, 4))
    D = np.zeros((7, 4))
    for i in range(4):
        c = np.poly1d([1.0])
        for j in range(4):
            if j != i:
                c = c * np.poly1d([1.0, -xn[j]]) / (xn[i] - xn[j])
        S[:, i], D[:, i] = c(xq), c.deriv()(xq)
    S, D = np.ascontiguousarray(S), np.ascontiguousarray(D)
    assert lib.afl_probe_set_tables(S.ctypes.data_as(C.POINTER(C.c_double)), D.ctypes.data_as(C.POINTER(C.c_double))) == 0
    # output placement of v_mfma_f64_16x16x4f64 on known operands: every (lane, register) output is
    # matched to the entry of A B it equals, and the placement must be a bijection onto the 16x16 result
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((16, 4)), rng.standard_normal((4, 16))
    da, db = torch.from_numpy(a.ravel()).cuda(), torch.from_numpy(b.ravel()).cuda()
    dd = torch.zeros(256, dtype=torch.float64, device="cuda")
    assert lib.afl_probe_mfma_layout(da.data_ptr(), db.data_ptr(), dd.data_ptr()) == 0
    got = dd.cpu().numpy().reshape(64, 4)
    ref = a @ b
    place, layout_err = {}, 0.0
    for lane in range(64):
        for j in range(4):
            k = int(np.argmin(np.abs(ref - got[lane, j])))
            layout_err = max(layout_err, float(np.abs(ref.flat[k] - got[lane, j])))
            place[(lane, j)] = divmod(k, 16)
    bijective = len(set(place.values())) == 256
    rows_rule = all(place[(l, j)] == (l // 16 + 4 * j, l % 16) for l in range(64) for j in range(4))
    tab = torch.from_numpy(np.concatenate([S.ravel(), D.ravel(), S.ravel(), D.ravel()])).cuda()
    lines = waves * 64
    x = torch.from_numpy(rng.uniform(-1.0, 1.0, lines * 4)).cuda()
    out = torch.empty(lines, dtype=torch.float64, device="cuda")
    res = {"workload": f"{lines} lines x {reps} repetitions of the k=3 value+derivative contraction (56 FMAs per line)",
           "mfma_layout": {"max_err": layout_err, "bijective": bijective,
                           "lane_l_reg_j_holds_row_l/16+4j_col_l%16": rows_rule,
                           "sample": {f"lane{l}": [place[(l, j)] for j in range(4)] for l in (0, 1, 16, 17, 63)}}}
    for mode, name in ((0, "valu"), (1, "mfma")):
        ms = C.c_float()
        assert lib.afl_probe_contraction(mode, 10, waves, tab.data_ptr(), x.data_ptr(), out.data_ptr(), C.byref(ms)) == 0
        assert lib.afl_probe_contraction(mode, reps, waves, tab.data_ptr(), x.data_ptr(), out.data_ptr(), C.byref(ms)) == 0
        useful = 2.0 * 56 * lines * reps
        issued = 2.0 * (56 if mode == 0 else 64) * lines * reps
        res[name] = {"ms": ms.value, "useful_tflops": useful / (ms.value * 1e-3) / 1e12,
                     "issued_tflops": issued / (ms.value * 1e-3) / 1e12,
                     "frac_of_fp64_peak_useful": useful / (ms.value * 1e-3) / 1e12 / FP64_PEAK_TFS,
                     "checksum_finite": bool(torch.isfinite(out).all())}
    res["mfma_over_valu_useful"] = res["mfma"]["useful_tflops"] / res["valu"]["useful_tflops"]
    return res


def c5_shard_leg(torch, EulerOperator, A, log, nsteps=20, npts=4096, snapshots=20):
    """The 3D 192x192x64-cell, k=4 configuration (288 x 288 x 25 km, hill h0 = 600 m, a = 2 km) split over
    eight GPUs along x gives each a 24x192x64-cell slab (36 x 288 x 25 km, 184 M DoF); this runs one slab at
    its real shape with periodic x (the neighbours' halo faces replaced by the slab's own): 20 FP32-MG steps
    timed, the per-kernel HIP-event times of two more and the HBM roofline of the kernel with the most
    time per step, the FP64-MG path over the same steps for the theta' and w errors, and the gradient of
    the final-time theta' misfit at 4096 seeded points of the slab by binomial checkpointing."""
    import numpy as np
    from atmospheric_flows_amd import make_mesh, make_hill_state
    half = 24 * 1500.0 / 2
    mesh = make_mesh(dim=3, degree=4, n_cells=(24, 192, 64), x_range=(-half, half), y_range=(-144e3, 144e3),
                     z_top=25e3, hill_height=600.0, hill_width=2000.0, periodic=(1, 1))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=2024)
    rng = np.random.default_rng(4096)
    pts = np.stack([rng.uniform(-half, half, npts), rng.uniform(-143e3, 143e3, npts),
                    rng.uniform(500.0, 20e3, npts)], axis=1)
    states, times, iters = {}, {}, {}
    free0 = torch.cuda.mem_get_info()[0]
    for name, pc in (("fp64_mg", A.AFL_SCHUR_MG_FP64), ("fp32_mg", A.AFL_SCHUR_MG_FP32)):
        log(f"c5 shard: {nsteps} {name} steps")
        o = EulerOperator(mesh, mach=1.0, froude=FR, dt=1.0)
        o.set_solver_params(params(pc, 1e-8, restart=8))
        x = o.init_hill_state(st)
        o.imex_step(*tuple(t.clone() for t in x))  # set-up (multigrid levels, eigenvalue estimates)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [o.imex_step(*x) for _ in range(nsteps)]
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / nsteps
        states[name] = x
        iters[name] = {"fp": [list(i.fp_iters) for i in infos], "gmres": [list(i.gmres_iters) for i in infos]}
        if name == "fp64_mg":
            del o
        else:
            # per-kernel HIP-event times of two more steps (from a copy of the state) and the HBM
            # roofline of the kernel with the most time per step
            log("c5 shard: per-kernel timing of 2 FP32-MG steps")
            kern, dom = kernels_in_steps(o, A, tuple(t.clone() for t in x), 2)
            d = kern[dom]
            roof = {"kernel": dom, "bound": "hbm", "achieved": d["GBs"], "bytes_per_launch": d["bytes_per_launch"],
                    "avg_launch_us": d["avg_us"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": d["GBs"] / HBM_PEAK_GBS}
    used = free0 - torch.cuda.mem_get_info()[0]
    N = A.n_per_field(mesh)
    x32, x64 = states["fp32_mg"], states["fp64_mg"]

    def rl2(a, b):
        e, r = o.l2_difference(1, a.contiguous(), b.contiguous())
        return e / r

    def theta_prime(rho, p):
        _, _, th_bar = o.background(st)
        return p / (rho * R_GAS) * torch.pow(st.p0 / p, R_GAS / CP) - th_bar

    errs = {"theta_prime": rl2(theta_prime(x32[0], x32[2]), theta_prime(x64[0], x64[2])),
            "w": rl2(x32[1][2 * N:], x64[1][2 * N:]), "rho": rl2(x32[0], x64[0]), "p": rl2(x32[2], x64[2])}
    del states, x32, x64
    # adjoint: obs from a seed-7 perturbed initial state, gradient of J at the seed-2024 state
    log("c5 shard: adjoint gradient")
    o.set_verbose(1)  # the reversal reports each reversed step
    o.set_observation_points(pts)
    truth = o.init_hill_state(make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=7))
    for _ in range(nsteps):
        o.imex_step(*truth)
    obs = o.observe_theta(st, truth[0], truth[2])
    del truth
    log("c5 shard: observations ready; timed forward run")
    x0 = o.init_hill_state(st)
    x = tuple(t.clone() for t in x0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.imex_step(*x)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    d = o.observe_theta(st, x[0], x[2]) - obs
    rb, pb = o.observe_theta_transpose(st, x[0], x[2], d)
    free1 = torch.cuda.mem_get_info()[0]
    # a two-step reversal first (its cost per step bounds the full run's duration)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    o.adjoint_run(*x0, 2, snapshots, rb, torch.zeros_like(x[1]), pb)
    torch.cuda.synchronize()
    log(f"c5 shard: two-step reversal {time.perf_counter() - t0:.1f} s; {nsteps}-step reversal")
    t0 = time.perf_counter()
    g = o.adjoint_run(*x0, nsteps, snapshots, rb, torch.zeros_like(x[1]), pb)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    log(f"c5 shard: reversal {t_adj:.1f} s")
    used_adj = free1 - torch.cuda.mem_get_info()[0]
    dofs = 5 * N
    return {"workload": "3D hill 192x192x64 cells, k=4 (288x288x25 km, h0=600 m, a=2 km), one GPU's 24x192x64-cell "
                        "x-slab at its real shape (periodic x), FP32-MG FGMRES (restart 8, true residual 1e-8)",
            "n_per_field": N, "dof": dofs, "steps": nsteps,
            "fp32_mg_ms_per_step": times["fp32_mg"] * 1e3, "value_fp32_mg": dofs / times["fp32_mg"],
            "fp64_mg_ms_per_step": times["fp64_mg"] * 1e3,
            "kernels_in_fp32_step": kern, "roofline": roof,
            "iters_per_step": iters,
            "rel_l2_fp32mg_vs_fp64mg_after_steps": errs,
            "adjoint": {"observation_points": npts, "snapshots": snapshots, "forward_run_s": t_fwd,
                        "adjoint_s": t_adj, "cost_vs_forward_run": t_adj / t_fwd, "schedule": g[3],
                        "gradient_finite": bool(torch.isfinite(g[2]).all() and torch.isfinite(g[0]).all())},
            "device_memory_gb": {"solver_paths": used / 1e9, "adjoint_run_extra": used_adj / 1e9,
                                 "full_mesh_estimate_x8": 8 * (used + used_adj) / 1e9, "hbm": 288 * 1.073741824}}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    the headline workload (calibrated against a copy of known size); None if not profiled."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_hill3d_96x96x40_k3.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel + ":"):
            for tok in line.split():
                if tok.startswith("bytes_per_launch="):
                    return float(tok.split("=")[1])
    return None


def profiled_fp64(kernel):
    """FP64 vector rate of `kernel` from the committed instruction-counter summary of the headline workload
    ((2 FMA + ADD + MUL) x 64 lanes x waves / duration, every lane counted), against the 78.6 TFLOP/s FP64
    vector peak, and its VALU issue-slot use; None if not profiled."""
    traffic = os.path.join(ROOT, "profiles", "pmc_traffic_hill3d_96x96x40_k3.txt")
    mix = os.path.join(ROOT, "profiles", "pmc_valu_hill3d_96x96x40_k3.txt")
    if not (os.path.exists(traffic) and os.path.exists(mix)):
        return None
    name = None
    for line in open(traffic):
        if line.startswith(kernel + ":"):
            name = [t for t in line.split() if t.startswith("kernel=")][0][len("kernel="):]
    if name is None:
        return None
    for line in open(mix):
        tok = line[42:].split()
        if len(tok) == 12 and "".join(line[:42].split()) == name:
            return {"tflops": float(tok[-3]), "peak_tflops": 78.6, "frac": float(tok[-2]) / 100.0,
                    "valu_issue_frac": float(tok[-1]) / 100.0, "source": "profiles/pmc_valu_hill3d_96x96x40_k3.txt"}
    return None


if __name__ == "__main__":
    main()
