"""Taylor test of the misfit gradient in the adjoint configuration's stated form.

2D hill, 256x60 cells, degree 3, dt = 1 s; J = 1/2 sum_k (theta'(x_k, T) - obs_k)^2 over 512 seeded
uniform observation points, obs from a seed-7 perturbed initial state; 200 steps, binomial
checkpointing with 20 snapshots, FP32-preconditioned stage solves at GMRES tolerance 1e-12. Along
eight seeded random directions v (scaled per field) the remainder |J(x + e v) - J(x) - e g.v| is
taken at e = 1e-2 .. 1e-7. It must fall by 100x per decade (second order, ratios within 60..160)
over all five decades; a wrong gradient leaves a first-order remainder that falls by 10x. (Seen:
ratios 83..114 in the first decade, 96..102 in the others, for every direction.)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from atmospheric_flows_amd import EulerOperator, make_mesh, make_hill_state, make_solver_params

FR = 1.0 / np.sqrt(9.80616)


def say(msg):
    print(msg, flush=True)


def test_taylor_remainder_is_second_order():
    import torch
    from atmospheric_flows_amd._abi import AFL_SCHUR_MG_FP32
    nsteps, nsnap, ndir = 200, 20, 8
    eps_list = (1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7)
    say(f"[gpu taylor] hill 256x60 k=3, {nsteps} steps, {nsnap} snapshots, 512 points, {ndir} directions")
    mesh = make_mesh(dim=2, degree=3, n_cells=(256, 1, 60), x_range=(-72e3, 72e3), z_top=25e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    op = EulerOperator(mesh, mach=1.0, froude=FR, dt=1.0)
    op.set_solver_params(make_solver_params(gmres_eps=1e-12, fp_tol=1e-12, schur_precond=AFL_SCHUR_MG_FP32))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=2024)
    rho, u, p = op.init_hill_state(st)
    N = op.N
    rng = np.random.default_rng(512)
    x = rng.uniform(-72e3 + 1, 72e3 - 1, 512)
    zeta = rng.uniform(1, 25e3 - 1, 512)
    h = 400.0 / (1 + (x / 1000.0) ** 2)
    op.set_observation_points(np.stack([x, zeta + h * (1 - zeta / 25e3)], 1))
    ro, uo, po = op.init_hill_state(make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0,
                                                    theta_sigma=0.01, seed=7))
    for _ in range(nsteps):
        op.imex_step(ro, uo, po)
    obs = op.observe_theta(st, ro, po)

    def J(r, uu, pp):
        r, uu, pp = r.clone(), uu.clone(), pp.clone()
        for _ in range(nsteps):
            op.imex_step(r, uu, pp)
        d = op.observe_theta(st, r, pp) - obs
        return 0.5 * float(torch.dot(d, d)), (r, uu, pp), d

    j0, xT, d = J(rho, u, p)
    rb, pb = op.observe_theta_transpose(st, xT[0], xT[2], d)
    *g, stats = op.adjoint_run(rho, u, p, nsteps, nsnap, rb, torch.zeros_like(u), pb)
    say(f"  J = {j0:.6e}, schedule {stats}")
    assert stats["adjoint_steps"] == nsteps and stats["snapshots_used"] <= nsnap + 1
    gen = torch.Generator(device="cuda").manual_seed(99)
    for t in range(ndir):
        v = (1e-3 * torch.randn(N, dtype=torch.float64, device="cuda", generator=gen),
             1e-2 * torch.randn(2 * N, dtype=torch.float64, device="cuda", generator=gen),
             1e2 * torch.randn(N, dtype=torch.float64, device="cuda", generator=gen))
        gv = sum(float(torch.dot(a, b)) for a, b in zip(g, v))
        rems = []
        for eps in eps_list:
            je = J(rho + eps * v[0], u + eps * v[1], p + eps * v[2])[0]
            rems.append(abs(je - j0 - eps * gv))
            say(f"  direction {t} eps {eps:.0e}: remainder {rems[-1]:.3e} (first-order term {abs(eps * gv):.3e})")
        ratios = [a / b for a, b in zip(rems[:-1], rems[1:])]
        say(f"  ratios per decade {['%.1f' % r for r in ratios]}")
        assert all(60.0 <= r <= 160.0 for r in ratios)
