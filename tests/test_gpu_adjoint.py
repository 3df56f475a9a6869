"""Parity of the HIP adjoint building blocks (through the C-ABI) against the CPU oracle.

The oracle's transposes are pinned by finite differences of the forward oracle and by
<y, L x> = <L^T y, x> (test_oracle_adjoint.py). Bar: relative L2 <= 1e-12 per apply (FP64),
the forward operators' bar. At a larger size, where the oracle is not run, the GPU transposes
must satisfy the same duality identities against the GPU forward operators to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from atmospheric_flows_amd import EulerOperator, make_mesh, make_hill_state

TOL = 1e-12
FR = 1.0 / np.sqrt(9.80616)

CASES = [
    (2, 1, (6, 1, 4)),
    (2, 2, (7, 1, 5)),
    (2, 3, (5, 1, 3)),
    (2, 4, (6, 1, 4)),
    (3, 1, (3, 3, 2)),
    (3, 2, (3, 2, 2)),
    (3, 3, (2, 2, 2)),
    (3, 4, (2, 2, 2)),
]


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _ids(c):
    return f"d{c[0]}k{c[1]}n{'x'.join(map(str, c[2]))}"


class Pair:
    def __init__(self, oracle_mod, dim, k, nc, dt=2.0, seed=5):
        import torch
        self.torch = torch
        xr = (-6e3, 6e3)
        self.mesh = make_mesh(dim=dim, degree=k, n_cells=nc, x_range=xr, y_range=xr, z_top=8e3,
                              hill_height=500.0, hill_width=1000.0, periodic=(1, 1))
        self.cpu = oracle_mod.OracleOperator(self.mesh, mach=1.0, froude=FR, dt=dt)
        self.gpu = EulerOperator(self.mesh, mach=1.0, froude=FR, dt=dt)
        self.N, self.dim = self.cpu.N, dim
        self.rng = np.random.default_rng(seed)
        st = make_hill_state(theta_sigma=0.5, seed=seed)
        self.rho, self.u, self.p = self.cpu.init_hill_state(st)

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def host(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()

    def zeros(self, n):
        return self.torch.zeros(n, dtype=self.torch.float64, device="cuda")

    def set_both(self, stage, rho_fix=None, p_fix=None):
        self.cpu.set_hyperbolic_stage(stage)
        self.gpu.set_HYPERBOLIC_stage(stage)
        if rho_fix is not None:
            self.cpu.set_rho_for_fixed(rho_fix); self.gpu.set_rho_for_fixed(self.dev(rho_fix))
        if p_fix is not None:
            self.cpu.set_pres_fixed(p_fix); self.gpu.set_pres_fixed(self.dev(p_fix))


@pytest.fixture(params=CASES, ids=_ids)
def pair(request, oracle_mod):
    dim, k, nc = request.param
    say(f"[gpu adjoint] case dim={dim} k={k} cells={nc}")
    return Pair(oracle_mod, dim, k, nc)


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_vjp_rhs_rho(pair, stage):
    P = pair
    P.set_both(stage)
    srcs = []
    for _ in range(stage):
        srcs += [P.rho * (1 + 1e-3 * P.rng.standard_normal(P.N)), P.u + 3.0 * P.rng.standard_normal(P.u.size)]
    rbar = P.rng.standard_normal(P.N)
    ref = P.cpu.vjp_rhs_rho_update(rbar, srcs)
    outs = [P.zeros(s.size) for s in srcs]
    P.gpu.vjp_rhs_rho_update(P.dev(rbar), [P.dev(s) for s in srcs], outs)
    for i, (o, r) in enumerate(zip(outs, ref)):
        e = rel(P.host(o), r)
        say(f"  stage {stage} source {i}: rel {e:.2e}")
        assert e <= TOL


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_vjp_rhs_velocity(pair, stage):
    P = pair
    P.set_both(stage)
    n = {1: 4, 2: 7, 3: 9}[stage]
    srcs = []
    for i in range(n):
        if i % 3 == 0:
            srcs.append(P.rho * (1 + 1e-3 * P.rng.standard_normal(P.N)))
        elif i % 3 == 1:
            srcs.append(P.u + 3.0 * P.rng.standard_normal(P.u.size))
        else:
            srcs.append(P.p * (1 + 1e-4 * P.rng.standard_normal(P.N)))
    rbar = P.rng.standard_normal(P.dim * P.N)
    ref = P.cpu.vjp_rhs_velocity_fixed(rbar, srcs)
    outs = [P.zeros(s.size) for s in srcs]
    P.gpu.vjp_rhs_velocity_fixed(P.dev(rbar), [P.dev(s) for s in srcs], outs)
    for i, (o, r) in enumerate(zip(outs, ref)):
        e = rel(P.host(o), r)
        say(f"  stage {stage} source {i}: rel {e:.2e}")
        assert e <= TOL


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_vjp_rhs_pressure(pair, stage):
    P = pair
    P.set_both(stage)
    n = {1: 5, 2: 8, 3: 11}[stage]
    nexp = stage
    srcs = []
    for i in range(n):
        kind = i % 3 if i < 3 * nexp else (0 if i == 3 * nexp else 1)
        if kind == 0:
            srcs.append(P.rho * (1 + 1e-3 * P.rng.standard_normal(P.N)))
        elif kind == 1:
            srcs.append(P.u + 3.0 * P.rng.standard_normal(P.u.size))
        else:
            srcs.append(P.p * (1 + 1e-4 * P.rng.standard_normal(P.N)))
    rbar = P.rng.standard_normal(P.N)
    ref = P.cpu.vjp_rhs_pressure(rbar, srcs)
    outs = [P.zeros(s.size) for s in srcs]
    P.gpu.vjp_rhs_pressure(P.dev(rbar), [P.dev(s) for s in srcs], outs)
    for i, (o, r) in enumerate(zip(outs, ref)):
        e = rel(P.host(o), r)
        say(f"  stage {stage} source {i}: rel {e:.2e}")
        assert e <= TOL


@pytest.mark.parametrize("stage", [1, 2])
def test_coefficient_transposes(pair, stage):
    P = pair
    pf = P.p * (1 + 1e-3 * P.rng.standard_normal(P.N))
    P.set_both(stage, p_fix=pf)
    w, cb = P.rng.standard_normal(P.dim * P.N), P.rng.standard_normal(P.N)
    ref = P.cpu.vmult_enthalpy_coefficient_transpose(cb, w)
    out = P.gpu.vmult_enthalpy_coefficient_transpose(P.zeros(P.N), P.dev(cb), P.dev(w))
    e1 = rel(P.host(out), ref)
    a, b = P.rng.standard_normal(P.dim * P.N), P.rng.standard_normal(P.dim * P.N)
    ref = P.cpu.vjp_rho_for_fixed(a, b)
    out = P.gpu.vjp_rho_for_fixed(P.zeros(P.N), P.dev(a), P.dev(b))
    e2 = rel(P.host(out), ref)
    # E3 with respect to u_fixed: the last output of the oracle's E3 adjoint
    n = {1: 5, 2: 8}[stage]
    srcs = []
    for i in range(n):
        kind = i % 3 if i < 3 * stage else (0 if i == 3 * stage else 1)
        srcs.append({0: P.rho * (1 + 1e-3 * P.rng.standard_normal(P.N)),
                     1: P.u + 3.0 * P.rng.standard_normal(P.u.size),
                     2: P.p * (1 + 1e-4 * P.rng.standard_normal(P.N))}[kind])
    rbar = P.rng.standard_normal(P.N)
    ref = P.cpu.vjp_rhs_pressure(rbar, srcs)[-1]
    out = P.gpu.vjp_rhs_pressure_u_fixed(P.zeros(P.dim * P.N), P.dev(rbar), P.dev(srcs[-2]), P.dev(srcs[-1]))
    e3 = rel(P.host(out), ref)
    say(f"  stage {stage}: C_pf^T rel {e1:.2e}; A_rho^T rel {e2:.2e}; E3_u^T rel {e3:.2e}")
    assert e1 <= TOL and e2 <= TOL and e3 <= TOL


@pytest.mark.parametrize("stage", [1, 2])
def test_pressure_transpose(pair, stage):
    P = pair
    P.set_both(stage)
    y = P.rng.standard_normal(P.dim * P.N)
    ref = P.cpu.vmult_pressure_transpose(y)
    out = P.gpu.vmult_pressure_transpose(P.zeros(P.N), P.dev(y))
    e = rel(P.host(out), ref)
    say(f"  B^T stage {stage}: rel {e:.2e}")
    assert e <= TOL


@pytest.mark.parametrize("stage", [1, 2])
def test_enthalpy_tangent_and_transpose(pair, stage):
    P = pair
    pf = P.p * (1 + 1e-3 * P.rng.standard_normal(P.N))
    P.set_both(stage, p_fix=pf)
    w0 = P.rng.standard_normal(P.dim * P.N)
    dw = P.rng.standard_normal(P.dim * P.N)
    cb = P.rng.standard_normal(P.N)
    ref_t = P.cpu.vmult_enthalpy_tangent(dw, w0)
    ref_tt = P.cpu.vmult_enthalpy_tangent_transpose(cb, w0)
    out_t = P.gpu.vmult_enthalpy_tangent(P.zeros(P.N), P.dev(dw), P.dev(w0))
    out_tt = P.gpu.vmult_enthalpy_tangent_transpose(P.zeros(P.dim * P.N), P.dev(cb), P.dev(w0))
    e1, e2 = rel(P.host(out_t), ref_t), rel(P.host(out_tt), ref_tt)
    say(f"  C' stage {stage}: rel {e1:.2e}; C'^T: rel {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL


@pytest.mark.parametrize("stage", [1, 2])
def test_schur_tangent_and_transpose(pair, stage):
    P = pair
    rf = P.rho * (1 + 1e-3 * P.rng.standard_normal(P.N))
    pf = P.p * (1 + 1e-3 * P.rng.standard_normal(P.N))
    P.set_both(stage, rho_fix=rf, p_fix=pf)
    p0 = 1e2 * P.rng.standard_normal(P.N)
    x = 1e2 * P.rng.standard_normal(P.N)
    ref_t = P.cpu.vmult_schur_tangent(x, p0)
    ref_tt = P.cpu.vmult_schur_tangent_transpose(x, p0)
    out_t = P.gpu.vmult_schur_tangent(P.zeros(P.N), P.dev(x), P.dev(p0))
    out_tt = P.gpu.vmult_schur_tangent_transpose(P.zeros(P.N), P.dev(x), P.dev(p0))
    e1, e2 = rel(P.host(out_t), ref_t), rel(P.host(out_tt), ref_tt)
    say(f"  S' stage {stage}: rel {e1:.2e}; S'^T: rel {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL


@pytest.mark.parametrize("dim,k,nc", [(2, 4, (128, 1, 40)), (3, 3, (12, 12, 6))])
def test_duality_at_size(dim, k, nc):
    """<y, L x> = <L^T y, x> on the GPU alone for B, C'(w0), S'(p0) and the E1 Jacobian."""
    import torch
    say(f"[gpu adjoint] duality dim={dim} k={k} cells={nc}")
    mesh = make_mesh(dim=dim, degree=k, n_cells=nc, x_range=(-30e3, 30e3), y_range=(-30e3, 30e3), z_top=20e3,
                     hill_height=400.0, hill_width=2000.0, periodic=(1, 1))
    op = EulerOperator(mesh, mach=1.0, froude=FR, dt=2.0)
    rho, u, p = op.init_hill_state(make_hill_state(theta_sigma=0.5, seed=9))
    N, D = op.N, dim
    g = torch.Generator(device="cuda").manual_seed(11)

    def rnd(n, s=1.0):
        return s * torch.randn(n, dtype=torch.float64, device="cuda", generator=g)

    def dot(a, b):
        return float(torch.dot(a, b))

    op.set_HYPERBOLIC_stage(2)
    op.set_rho_for_fixed(rho * (1 + 1e-3 * rnd(N)))
    op.set_pres_fixed(p * (1 + 1e-3 * rnd(N)))
    # B
    x, y = rnd(N), rnd(D * N)
    Bx = op.vmult_pressure(torch.empty(D * N, dtype=torch.float64, device="cuda"), x)
    Bty = op.vmult_pressure_transpose(torch.empty(N, dtype=torch.float64, device="cuda"), y)
    a, b = dot(y, Bx), dot(Bty, x)
    say(f"  B: {a:.12e} {b:.12e}")
    assert abs(a - b) <= TOL * float(y.norm() * Bx.norm())
    # C'(w0)
    w0, dw, cb = rnd(D * N), rnd(D * N), rnd(N)
    Cd = op.vmult_enthalpy_tangent(torch.empty(N, dtype=torch.float64, device="cuda"), dw, w0)
    Ctc = op.vmult_enthalpy_tangent_transpose(torch.empty(D * N, dtype=torch.float64, device="cuda"), cb, w0)
    a, b = dot(cb, Cd), dot(Ctc, dw)
    say(f"  C': {a:.12e} {b:.12e}")
    assert abs(a - b) <= TOL * float(cb.norm() * Cd.norm())
    # S'(p0)
    p0, xs, ys = rnd(N, 1e2), rnd(N, 1e2), rnd(N)
    Sx = op.vmult_schur_tangent(torch.empty(N, dtype=torch.float64, device="cuda"), xs, p0)
    Sty = op.vmult_schur_tangent_transpose(torch.empty(N, dtype=torch.float64, device="cuda"), ys, p0)
    a, b = dot(ys, Sx), dot(Sty, xs)
    say(f"  S': {a:.12e} {b:.12e}")
    assert abs(a - b) <= 1e-11 * float(ys.norm() * Sx.norm())
    # E1: the VJP against a directional derivative of the forward apply (exact: E1 is bilinear
    # in (rho_j, u_j) when the direction leaves the Rusanov speeds' order unchanged; a direction in
    # the densities only does)
    srcs = [rho * (1 + 1e-3 * rnd(N)), u + rnd(D * N), rho * (1 + 1e-3 * rnd(N)), u + rnd(D * N)]
    dirs = [rnd(N, 1e-3), torch.zeros(D * N, dtype=torch.float64, device="cuda"), rnd(N, 1e-3),
            torch.zeros(D * N, dtype=torch.float64, device="cuda")]
    rbar = rnd(N)
    outs = [torch.empty_like(s) for s in srcs]
    op.vjp_rhs_rho_update(rbar, srcs, outs)
    e_p = op.vmult_rhs_rho_update(torch.empty(N, dtype=torch.float64, device="cuda"),
                                  [s + d for s, d in zip(srcs, dirs)])
    e_0 = op.vmult_rhs_rho_update(torch.empty(N, dtype=torch.float64, device="cuda"), srcs)
    a = dot(rbar, e_p - e_0)
    b = sum(dot(o, d) for o, d in zip(outs, dirs))
    say(f"  E1 (density directions): {a:.12e} {b:.12e}")
    assert abs(a - b) <= 1e-9 * abs(a)
