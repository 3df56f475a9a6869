"""Gradient-check harness: the C-ABI adjoint wrapped as a torch.autograd.Function.

forward runs n IMEX steps on the device; backward hands the output adjoint to afl_adjoint_run
(binomial checkpointing) and returns the initial-state adjoint. The check is that autograd's
gradient of a scalar of the final state equals the chain of step adjoints over stored states,
and that a difference quotient of the forward along a density direction agrees with it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from atmospheric_flows_amd import EulerOperator, make_mesh, make_hill_state, make_solver_params

FR = 1.0 / np.sqrt(9.80616)


def say(msg):
    print(msg, flush=True)


def _imex_steps_function():
    import torch

    class ImexSteps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, u, p, op, nsteps, snapshots):
            ctx.op, ctx.nsteps, ctx.snapshots = op, nsteps, snapshots
            ctx.save_for_backward(rho, u, p)
            r, uu, pp = rho.detach().clone(), u.detach().clone(), p.detach().clone()
            for _ in range(nsteps):
                op.imex_step(r, uu, pp)
            return r, uu, pp

        @staticmethod
        def backward(ctx, gr, gu, gp):
            rho, u, p = ctx.saved_tensors
            rb, ub, pb, _ = ctx.op.adjoint_run(rho, u, p, ctx.nsteps, ctx.snapshots,
                                               gr.contiguous(), gu.contiguous(), gp.contiguous())
            return rb, ub, pb, None, None, None

    return ImexSteps


def test_autograd_function_gradient():
    import torch
    say("[gpu autograd] 4 steps, hill 32x12 k=3")
    mesh = make_mesh(dim=2, degree=3, n_cells=(32, 1, 12), x_range=(-36e3, 36e3), z_top=12e3,
                     hill_height=400.0, hill_width=1000.0, periodic=(1, 1))
    op = EulerOperator(mesh, mach=1.0, froude=FR, dt=2.0)
    op.set_solver_params(make_solver_params(gmres_eps=1e-12, fp_tol=1e-12))
    rho, u, p = op.init_hill_state(make_hill_state(theta_sigma=0.01, seed=11))
    N = op.N
    gen = torch.Generator(device="cuda").manual_seed(2)
    u = u + 0.5 * torch.randn(2 * N, dtype=torch.float64, device="cuda", generator=gen)
    w = [torch.randn(n, dtype=torch.float64, device="cuda", generator=gen) * s
         for n, s in ((N, 1.0), (2 * N, 1.0), (N, 1e-5))]
    ImexSteps = _imex_steps_function()
    x0 = [t.clone().requires_grad_(True) for t in (rho, u, p)]
    xT = ImexSteps.apply(*x0, op, 4, 2)
    J = sum((a * b).sum() for a, b in zip(xT, w))
    J.backward()
    # the chain of step adjoints over stored states
    xs = [(rho.clone(), u.clone(), p.clone())]
    for _ in range(4):
        r, uu, pp = (t.clone() for t in xs[-1])
        op.imex_step(r, uu, pp)
        xs.append((r, uu, pp))
    lam = tuple(w)
    for n in reversed(range(4)):
        lam = op.imex_step_vjp(*xs[n], *lam)[:3]
    for a, b in zip(x0, lam):
        e = float((a.grad - b).norm() / b.norm())
        say(f"  autograd vs step-adjoint chain: rel {e:.2e}")
        assert e <= 1e-10  # the stage solves' initial guesses differ: agreement to the tolerance
    # a difference quotient along a density direction
    v = 1e-3 * torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)

    def Jx(r0):
        with torch.no_grad():
            out = ImexSteps.apply(r0, u, p, op, 4, 2)
            return float(sum((a * b).sum() for a, b in zip(out, w)))

    h = 1e-3
    fdv = (Jx(rho + h * v) - Jx(rho - h * v)) / (2 * h)
    gv = float(torch.dot(x0[0].grad, v))
    say(f"  density direction: FD {fdv:.10e} gradient {gv:.10e} rel {abs(fdv - gv) / abs(fdv):.2e}")
    assert abs(fdv - gv) <= 1e-4 * abs(fdv)
