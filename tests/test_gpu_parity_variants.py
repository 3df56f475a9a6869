"""Parity of the HIP path against the oracle on mesh variants the other parity tests leave out:
lateral walls instead of periodic sides (wall faces with horizontal normals in B, E2 and their
transposes) and equidistant support points instead of Gauss-Lobatto ones. Same bars: relative L2
<= 1e-12 per apply, <= 1e-10 after three IMEX steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from atmospheric_flows_amd import EulerOperator, make_mesh, make_hill_state, make_solver_params
from atmospheric_flows_amd._abi import AFL_NODES_EQUIDISTANT, AFL_NODES_GAUSS_LOBATTO

FR = 1.0 / np.sqrt(9.80616)
TOL = 1e-12

VARIANTS = [
    (2, 3, (6, 1, 4), (0, 0), AFL_NODES_GAUSS_LOBATTO),
    (3, 2, (3, 3, 2), (0, 1), AFL_NODES_GAUSS_LOBATTO),
    (2, 3, (6, 1, 4), (1, 1), AFL_NODES_EQUIDISTANT),
    (3, 3, (2, 2, 2), (0, 0), AFL_NODES_EQUIDISTANT),
]


def _ids(v):
    return f"d{v[0]}k{v[1]}per{v[3][0]}{v[3][1]}nodes{v[4]}"


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(params=VARIANTS, ids=_ids)
def setup(request, oracle_mod):
    import torch
    dim, k, nc, per, nodes = request.param
    say(f"[gpu parity variants] dim={dim} k={k} cells={nc} periodic={per} nodes={nodes}")
    xr = (-6e3, 6e3)
    mesh = make_mesh(dim=dim, degree=k, n_cells=nc, x_range=xr, y_range=xr, z_top=8e3, hill_height=500.0,
                     hill_width=1000.0, periodic=per, node_type=nodes)
    cpu = oracle_mod.OracleOperator(mesh, mach=1.0, froude=FR, dt=2.0)
    gpu = EulerOperator(mesh, mach=1.0, froude=FR, dt=2.0)
    rng = np.random.default_rng(17 + dim + k)
    rho, u, p = cpu.init_hill_state(make_hill_state(theta_sigma=0.3, seed=17))
    states = [(rho * (1 + 1e-3 * rng.standard_normal(rho.size)), u + 2.0 * rng.standard_normal(u.size),
               p * (1 + 1e-4 * rng.standard_normal(p.size))) for _ in range(4)]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return cpu, gpu, states, rng, dev, dim


def test_operators(setup):
    import torch
    cpu, gpu, S, rng, dev, dim = setup
    N = cpu.N
    cpu.set_hyperbolic_stage(3)
    gpu.set_HYPERBOLIC_stage(3)
    flat = [S[0][0], S[0][1], S[0][2], S[1][0], S[1][1], S[1][2], S[2][0], S[2][1], S[2][2]]
    cases = [("E1", cpu.vmult_rhs_rho_update, gpu.vmult_rhs_rho_update, [S[0][0], S[0][1], S[1][0], S[1][1], S[2][0], S[2][1]], N),
             ("E2", cpu.vmult_rhs_velocity_fixed, gpu.vmult_rhs_velocity_fixed, flat, dim * N),
             ("E3", cpu.vmult_rhs_pressure, gpu.vmult_rhs_pressure, flat + [S[3][0], S[3][1]], N)]
    for name, fc, fg, srcs, n in cases:
        ref = fc(srcs)
        out = fg(torch.zeros(n, dtype=torch.float64, device="cuda"), [dev(a) for a in srcs])
        e = rel(out.cpu().numpy(), ref)
        say(f"  {name}: rel {e:.2e}")
        assert e <= TOL
    cpu.set_hyperbolic_stage(1)
    gpu.set_HYPERBOLIC_stage(1)
    cpu.set_pres_fixed(S[2][2]); gpu.set_pres_fixed(dev(S[2][2]))
    cpu.set_rho_for_fixed(S[3][0]); gpu.set_rho_for_fixed(dev(S[3][0]))
    e = rel(gpu.vmult_pressure(gpu.new_vector(), dev(S[0][2])).cpu().numpy(), cpu.vmult_pressure(S[0][2]))
    say(f"  B: rel {e:.2e}"); assert e <= TOL
    e = rel(gpu.vmult_enthalpy(gpu.new_scalar(), dev(S[1][1])).cpu().numpy(), cpu.vmult_enthalpy(S[1][1]))
    say(f"  C: rel {e:.2e}"); assert e <= TOL
    cpu.set_ns_stage(2); gpu.set_NS_stage(2)
    e = rel(gpu.vmult(gpu.new_scalar(), dev(S[0][2])).cpu().numpy(), cpu.vmult(S[0][2]))
    say(f"  S: rel {e:.2e}"); assert e <= TOL
    # transposes with wall faces
    y = rng.standard_normal(dim * N)
    e = rel(gpu.vmult_pressure_transpose(gpu.new_scalar(), dev(y)).cpu().numpy(), cpu.vmult_pressure_transpose(y))
    say(f"  B^T: rel {e:.2e}"); assert e <= TOL
    rbar = rng.standard_normal(dim * N)
    src4 = [S[0][0], S[0][1], S[0][2], S[1][0]]
    ref = cpu.vjp_rhs_velocity_fixed(rbar, src4)
    outs = gpu.vjp_rhs_velocity_fixed(dev(rbar), [dev(a) for a in src4],
                                      [torch.zeros(a.size, dtype=torch.float64, device="cuda") for a in src4])
    e = max(rel(o.cpu().numpy(), r) for o, r in zip(outs, ref))
    say(f"  E2 adjoint: rel {e:.2e}"); assert e <= TOL


def test_imex_steps(setup):
    import torch
    cpu, gpu, S, rng, dev, dim = setup
    sp = make_solver_params(gmres_eps=1e-12, fp_tol=1e-10)
    cpu.set_solver_params(sp); gpu.set_solver_params(sp)
    rho, u, p = (a.copy() for a in S[0])
    dr, du, dp = dev(rho), dev(u), dev(p)
    for _ in range(3):
        cpu.imex_step(rho, u, p)
        gpu.imex_step(dr, du, dp)
    errs = [rel(t.cpu().numpy(), a) for t, a in ((dr, rho), (du, u), (dp, p))]
    say(f"  3 steps: rel {['%.2e' % x for x in errs]}")
    assert max(errs) <= 1e-10
