"""Adjoint of whole IMEX steps on the GPU (afl_imex_step_vjp) against the CPU oracle.

The checker is the oracle's forward step: for a scalar J(x) = <c, step(x)> the GPU gradient
g = step^T c must reproduce the oracle's central differences <c, step(x + h v) - step(x - h v)> / 2h
along seeded random directions v, i.e. the transposed Jacobian of the CPU step applied to c, probed
column-direction by column-direction. A second test chains step adjoints over several steps (the
gradient of a final-time functional with respect to the initial state); a third runs at the size of
the reference's CPU configuration (64x25, k = 2) against the GPU forward map itself.

The initial states carry velocity noise: on the plain hill state u = 10 m/s exactly, both traces of
every vertical face have the same normal velocity and the Rusanov speed max(|a|, |b|) is at its kink,
where the step is directionally differentiable but not differentiable. Away from exact ties the step
is still only piecewise smooth: the Schur operator's Rusanov speed acts on w = A^{-1} B p, whose
horizontal component is nearly zero on the vertical faces of a near-hydrostatic state, so a
perturbation of a difference quotient can cross a kink. Such a crossing spoils that one quotient
(errors of 1e-5..1e-3 were seen), not the gradient, so each test takes several directions and bounds
the median error (1e-6; the smooth quotients agree to 1e-7..1e-10) and every error (1e-3).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from atmospheric_flows_amd import EulerOperator, make_mesh, make_hill_state, make_solver_params

FR = 1.0 / np.sqrt(9.80616)


def say(msg):
    print(msg, flush=True)


def _pair(oracle_mod, dim, k, nc, dt):
    xr = (-6e3, 6e3)
    mesh = make_mesh(dim=dim, degree=k, n_cells=nc, x_range=xr, y_range=xr, z_top=8e3,
                     hill_height=500.0, hill_width=1000.0, periodic=(1, 1))
    cpu = oracle_mod.OracleOperator(mesh, mach=1.0, froude=FR, dt=dt)
    gpu = EulerOperator(mesh, mach=1.0, froude=FR, dt=dt)
    sp = make_solver_params(gmres_eps=1e-13, fp_tol=1e-13)
    cpu.set_solver_params(sp)
    gpu.set_solver_params(sp)
    return cpu, gpu


def _state(cpu, seed):
    rng = np.random.default_rng(seed)
    rho, u, p = cpu.init_hill_state(make_hill_state(theta_sigma=0.3, seed=seed))
    u = u + 0.5 * rng.standard_normal(u.size)
    return rho, u, p, rng


def _weights(rng, N, dim):
    """Output adjoint c and direction scalings that balance rho (~1), u (~10) and p (~1e5)."""
    c = (rng.standard_normal(N), rng.standard_normal(dim * N), 1e-5 * rng.standard_normal(N))
    scale = (1e-3, 1e-2, 1e2)
    return c, scale


def _cpu_steps(cpu, x, n):
    r, u, p = (a.copy() for a in x)
    for _ in range(n):
        cpu.imex_step(r, u, p)
    return r, u, p


def _J(c, x):
    return sum(float(a @ b) for a, b in zip(c, x))


def _check_directions(cpu, c, g, x, dirs, nsteps, h=1e-3):
    errs = []
    for i, v in enumerate(dirs):
        jp = _J(c, _cpu_steps(cpu, tuple(a + h * b for a, b in zip(x, v)), nsteps))
        jm = _J(c, _cpu_steps(cpu, tuple(a - h * b for a, b in zip(x, v)), nsteps))
        fdv = (jp - jm) / (2 * h)
        gv = _J(g, v)
        errs.append(abs(fdv - gv) / abs(fdv))
        say(f"  direction {i}: FD {fdv:.12e} adjoint {gv:.12e} rel {errs[-1]:.2e}")
    assert np.median(errs) <= 1e-6 and max(errs) <= 1e-3


@pytest.mark.parametrize("dim,k,nc", [(2, 2, (6, 1, 4)), (2, 3, (5, 1, 3)), (3, 2, (3, 2, 2))])
def test_step_vjp_matches_oracle_differences(oracle_mod, dim, k, nc):
    import torch
    say(f"[gpu step adjoint] dim={dim} k={k} cells={nc}")
    cpu, gpu = _pair(oracle_mod, dim, k, nc, dt=2.0)
    N = cpu.N
    rho, u, p, rng = _state(cpu, 17 + dim + k)
    c, scale = _weights(rng, N, dim)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rb, ub, pb, info = gpu.imex_step_vjp(dev(rho), dev(u), dev(p), dev(c[0]), dev(c[1]), dev(c[2]))
    torch.cuda.synchronize()
    g = (rb.cpu().numpy(), ub.cpu().numpy(), pb.cpu().numpy())
    say(f"  forward recompute: fp {list(info.fp_iters)} gmres {list(info.gmres_iters)}")
    x = (rho, u, p)
    dirs = [tuple(s * rng.standard_normal(a.size) for s, a in zip(scale, x)) for _ in range(3)]
    # one field at a time as well (a wrong block shows even where the others dominate)
    for f in range(3):
        v = [np.zeros_like(a) for a in x]
        v[f] = scale[f] * rng.standard_normal(x[f].size)
        dirs.append(tuple(v))
    _check_directions(cpu, c, g, x, dirs, 1)


def test_multistep_gradient_matches_oracle_differences(oracle_mod):
    """dJ/dx0 for J = <c, x_3> over three steps: chained step adjoints in reverse order."""
    import torch
    say("[gpu step adjoint] three steps, dim=2 k=2 cells=(6, 1, 4)")
    cpu, gpu = _pair(oracle_mod, 2, 2, (6, 1, 4), dt=2.0)
    N = cpu.N
    rho, u, p, rng = _state(cpu, 29)
    c, scale = _weights(rng, N, 2)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    xs = [tuple(dev(a) for a in (rho, u, p))]
    for _ in range(3):
        r, uu, pp = (t.clone() for t in xs[-1])
        gpu.imex_step(r, uu, pp)
        xs.append((r, uu, pp))
    lam = tuple(dev(a) for a in c)
    for n in (2, 1, 0):
        rb, ub, pb, _ = gpu.imex_step_vjp(*xs[n], *lam)
        lam = (rb, ub, pb)
    torch.cuda.synchronize()
    g = tuple(t.cpu().numpy() for t in lam)
    x = (rho, u, p)
    dirs = [tuple(s * rng.standard_normal(a.size) for s, a in zip(scale, x)) for _ in range(5)]
    # kinks multiply over steps: a smaller quotient step keeps crossings rare
    _check_directions(cpu, c, g, x, dirs, 3, h=1e-4)


def test_step_vjp_at_size():
    """64x25 k=2 hill: the step adjoint is linear in the output adjoint (calls share only the
    initial guess of the stage solves) and agrees with GPU difference quotients along density
    directions."""
    import torch
    say("[gpu step adjoint] hill 64x25 k=2")
    mesh = make_mesh(dim=2, degree=2, n_cells=(64, 1, 25), x_range=(-72e3, 72e3), z_top=25e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    op = EulerOperator(mesh, mach=1.0, froude=FR, dt=2.0)
    op.set_solver_params(make_solver_params(gmres_eps=1e-13, fp_tol=1e-13))
    rho, u, p = op.init_hill_state(make_hill_state(theta_sigma=0.01, seed=2024))
    N = op.N
    gen = torch.Generator(device="cuda").manual_seed(5)
    rnd = lambda n, s: s * torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    u = u + rnd(2 * N, 0.5)
    c1 = (rnd(N, 1.0), rnd(2 * N, 1.0), rnd(N, 1e-5))
    c2 = (rnd(N, 1.0), rnd(2 * N, 1.0), rnd(N, 1e-5))
    g1 = op.imex_step_vjp(rho, u, p, *c1)[:3]
    g2 = op.imex_step_vjp(rho, u, p, *c2)[:3]
    g12 = op.imex_step_vjp(rho, u, p, *(2.0 * a - 3.0 * b for a, b in zip(c1, c2)))[:3]
    for a, b, ab in zip(g1, g2, g12):
        e = float((ab - (2.0 * a - 3.0 * b)).norm() / ab.norm())
        say(f"  linearity: rel {e:.2e}")
        assert e <= 1e-10

    def J(x):
        r, uu, pp = (t.clone() for t in x)
        op.imex_step(r, uu, pp)
        return sum(float(torch.dot(a, b)) for a, b in zip(c1, (r, uu, pp)))

    errs = []
    for t in range(3):
        v = rnd(N, 1e-3)
        h = 1e-3
        fdv = (J((rho + h * v, u, p)) - J((rho - h * v, u, p))) / (2 * h)
        gv = float(torch.dot(g1[0], v))
        errs.append(abs(fdv - gv) / abs(fdv))
        say(f"  density direction {t}: FD {fdv:.10e} adjoint {gv:.10e} rel {errs[-1]:.2e}")
    assert np.median(errs) <= 1e-4 and max(errs) <= 1e-2


def _binomial_cost(t, s, memo={}):
    """Forward steps of the optimal checkpointed reversal of t steps with s free snapshots."""
    if t == 1:
        return 0
    if s == 0:
        return t * (t - 1) // 2
    key = (t, s)
    if key not in memo:
        memo[key] = min(m + _binomial_cost(t - m, s - 1) + _binomial_cost(m, s) for m in range(1, t))
    return memo[key]


def _reversal_recomputes(t, s):
    """Replays the reversal order of the checkpointed run (split = the first optimal m) with a stage
    set kept at the end of every advance to a split point: returns (steps reversed without a set,
    most sets held at once). The reversal must visit the steps in reverse order."""
    held, order, most, recomputed = set(), [], [0], [0]

    def split(t, s):
        return min(range(1, t), key=lambda m: m + _binomial_cost(t - m, s - 1) + _binomial_cost(m, s))

    def leaf(j):
        order.append(j)
        if j in held:
            held.remove(j)
        else:
            recomputed[0] += 1

    def rev(a, b, s):
        if b - a == 1:
            return leaf(a)
        if s == 0:
            for k in range(b - 1, a - 1, -1):
                leaf(k)
            return
        m = split(b - a, s)
        held.add(a + m - 1)
        most[0] = max(most[0], len(held))
        rev(a + m, b, s - 1)
        rev(a, a + m, s)

    rev(0, t, s)
    assert order == list(range(t - 1, -1, -1))
    return recomputed[0], most[0]


@pytest.mark.parametrize("snapshots,keep", [(0, True), (1, True), (2, True), (7, True), (2, False), (7, False),
                                            (-1, True)])
def test_adjoint_run_checkpointing_matches_stored_states(snapshots, keep):
    """Seven steps reversed with binomial checkpointing (or with every step's stages kept, -1) give
    the gradient of the chain of step adjoints over stored states (to the solver tolerance), with
    the optimal number of recomputed forward steps; with stage sets kept, only the steps reversed
    without one recompute their forward step."""
    import torch
    say(f"[gpu adjoint run] 7 steps, {snapshots} snapshots, stage sets {keep}, hill 32x12 k=2")
    mesh = make_mesh(dim=2, degree=2, n_cells=(32, 1, 12), x_range=(-36e3, 36e3), z_top=12e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    op = EulerOperator(mesh, mach=1.0, froude=FR, dt=2.0)
    op.set_solver_params(make_solver_params(gmres_eps=1e-12, fp_tol=1e-12))
    rho, u, p = op.init_hill_state(make_hill_state(theta_sigma=0.01, seed=7))
    N = op.N
    gen = torch.Generator(device="cuda").manual_seed(3)
    rnd = lambda n, s: s * torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    u = u + rnd(2 * N, 0.5)
    c = (rnd(N, 1.0), rnd(2 * N, 1.0), rnd(N, 1e-5))
    nsteps = 7
    xs = [(rho.clone(), u.clone(), p.clone())]
    for _ in range(nsteps):
        r, uu, pp = (t.clone() for t in xs[-1])
        op.imex_step(r, uu, pp)
        xs.append((r, uu, pp))
    lam = c
    for n in reversed(range(nsteps)):
        lam = op.imex_step_vjp(*xs[n], *lam)[:3]
    op.set_adjoint_stage_sets(keep)
    rb, ub, pb, stats = op.adjoint_run(rho, u, p, nsteps, snapshots, *c)
    say(f"  stats {stats}")
    # the adjoint stage solves start from the previous call's multiplier, so the two runs agree to
    # the solver tolerance rather than bitwise
    for a, b in zip((rb, ub, pb), lam):
        e = float((a - b).norm() / b.norm())
        say(f"  checkpointed vs stored: rel {e:.2e}")
        assert e <= 1e-10
    assert stats["adjoint_steps"] == nsteps
    # the stage solves' Krylov iterations are counted (a solve whose warm start already meets the
    # tolerance takes none, so no per-step lower bound holds)
    assert stats["stage_solve_iterations"] > 0
    if snapshots < 0:  # every step's stages kept: one forward sweep, no recomputation
        assert stats["forward_steps"] == nsteps and stats["snapshots_used"] == nsteps
    else:
        s = min(snapshots, nsteps - 1)
        assert stats["forward_steps"] == _binomial_cost(nsteps, s)
        assert stats["snapshots_used"] <= s + 1
        if keep:
            recomputed, most = _reversal_recomputes(nsteps, s)
            assert (stats["recomputed_steps"], stats["stage_sets_used"]) == (recomputed, most)
            assert most <= s + 1
        else:
            assert stats["recomputed_steps"] == nsteps and stats["stage_sets_used"] == 0


def test_step_vjp_with_fp32_multigrid_solves(oracle_mod):
    """The adjoint stage solves preconditioned by the FP32 multigrid V-cycle (FGMRES) give the
    Jacobi-GMRES gradient to the solver tolerance."""
    import torch
    from atmospheric_flows_amd._abi import AFL_SCHUR_MG_FP32
    say("[gpu step adjoint] FP32 multigrid vs Jacobi stage solves, dim=2 k=4 cells=(12, 1, 6)")
    cpu, gpu = _pair(oracle_mod, 2, 4, (12, 1, 6), dt=2.0)
    N = cpu.N
    rho, u, p, rng = _state(cpu, 41)
    c, _ = _weights(rng, N, 2)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g_j = gpu.imex_step_vjp(dev(rho), dev(u), dev(p), *(dev(a) for a in c))[:3]
    gpu.set_solver_params(make_solver_params(gmres_eps=1e-13, fp_tol=1e-13, schur_precond=AFL_SCHUR_MG_FP32))
    g_m = gpu.imex_step_vjp(dev(rho), dev(u), dev(p), *(dev(a) for a in c))[:3]
    for a, b in zip(g_m, g_j):
        e = float((a - b).norm() / b.norm())
        say(f"  multigrid vs Jacobi: rel {e:.2e}")
        assert e <= 1e-9


def test_step_vjp_with_sponge(oracle_mod):
    """The Rayleigh sponge (u <- (u + a u_ref) / (1 + a) above z_start) enters the step adjoint as
    the diagonal 1 / (1 + a): checked against oracle differences with the sponge on."""
    import torch
    from atmospheric_flows_amd._abi import make_sponge
    say("[gpu step adjoint] sponge above 4 km, dim=2 k=2 cells=(6, 1, 4)")
    cpu, gpu = _pair(oracle_mod, 2, 2, (6, 1, 4), dt=2.0)
    sponge = make_sponge(z_start=4e3, alpha_max=0.05, u_ref=(10.0, 0.0, 0.0), enabled=True)
    cpu.set_sponge(sponge)
    gpu.set_sponge(sponge)
    N = cpu.N
    rho, u, p, rng = _state(cpu, 53)
    c, scale = _weights(rng, N, 2)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rb, ub, pb, _ = gpu.imex_step_vjp(dev(rho), dev(u), dev(p), *(dev(a) for a in c))
    g = (rb.cpu().numpy(), ub.cpu().numpy(), pb.cpu().numpy())
    x = (rho, u, p)
    dirs = [tuple(s * rng.standard_normal(a.size) for s, a in zip(scale, x)) for _ in range(3)]
    v = [np.zeros_like(a) for a in x]
    v[1] = scale[1] * rng.standard_normal(u.size)
    dirs.append(tuple(v))
    _check_directions(cpu, c, g, x, dirs, 1)
