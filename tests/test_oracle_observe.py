"""Pins of the oracle's observation operator (theta' at points and its transpose).

The reference has no point evaluation. Pins: at interior nodes the evaluation returns the nodal
values (Lagrange property, checks the inverse map and the basis); fields linear in the horizontal
coordinate or in the terrain-following vertical coordinate zeta are reproduced exactly at random
points (checks the column search and the inverse z -> zeta); the hydrostatic background has
theta' = 0 at the nodes; the transpose agrees with central differences (theta is smooth in rho, p).
"""
import numpy as np
import pytest

from atmospheric_flows_amd import make_mesh, make_hill_state

R, G, GAMMA = 287.0, 9.80616, 1.4
KAP = (GAMMA - 1.0) / GAMMA


def _mesh(dim, k):
    if dim == 2:
        return make_mesh(dim=2, degree=k, n_cells=(8, 1, 5), x_range=(-8e3, 8e3), z_top=10e3,
                         hill_height=600.0, hill_width=1500.0, periodic=(1, 1))
    return make_mesh(dim=3, degree=k, n_cells=(4, 3, 3), x_range=(-8e3, 8e3), y_range=(-6e3, 6e3), z_top=10e3,
                     hill_height=600.0, hill_width=1500.0, periodic=(1, 1))


def _random_points(rng, n, dim, H=10e3):
    x = rng.uniform(-8e3 + 1, 8e3 - 1, n)
    y = rng.uniform(-6e3 + 1, 6e3 - 1, n)
    zeta = rng.uniform(1.0, H - 1.0, n)
    r2 = x * x + (y * y if dim == 3 else 0.0)
    h = 600.0 / (1.0 + r2 / 1500.0 ** 2)
    z = zeta + h * (1 - zeta / H)
    pts = np.stack([x, y, z] if dim == 3 else [x, z], axis=1)
    return pts, zeta


def _rho_from_theta(st, tp, z):
    """p = p0 everywhere: theta = p0 / (rho R), so rho = p0 / (R (theta' + theta_bar))."""
    return st.p0 / (R * (tp + st.theta0 * np.exp(st.brunt ** 2 * z / G)))


@pytest.mark.parametrize("dim,k", [(2, 2), (2, 4), (3, 3)])
def test_evaluation_pins(oracle_mod, dim, k):
    op = oracle_mod.OracleOperator(_mesh(dim, k))
    st = make_hill_state()
    N = op.N
    X = op.node_coords()
    n1 = k + 1
    # interior nodes of every cell
    ii = np.arange(N) % (n1 ** dim)
    idx = [(ii // n1 ** d) % n1 for d in range(dim)]
    interior = np.all([(a > 0) & (a < k) for a in idx], axis=0)
    sel = np.flatnonzero(interior)[::3]
    rng = np.random.default_rng(5)
    f = 1.0 + 0.1 * rng.random(N)
    op.set_observation_points(X[:, sel].T)
    p0 = np.full(N, st.p0)
    tp = op.observe_theta(st, f, p0)
    assert np.allclose(_rho_from_theta(st, tp, X[dim - 1, sel]), f[sel], rtol=1e-13, atol=0)
    # linear in x, and linear in the terrain-following coordinate zeta
    pts, zeta = _random_points(rng, 200, dim)
    op.set_observation_points(pts)
    fx = 1.0 + 1e-5 * X[0]
    tp = op.observe_theta(st, fx, p0)
    assert np.allclose(_rho_from_theta(st, tp, pts[:, -1]), 1.0 + 1e-5 * pts[:, 0], rtol=1e-13, atol=0)
    r2 = X[0] ** 2 + (X[1] ** 2 if dim == 3 else 0.0)
    hN = 600.0 / (1.0 + r2 / 1500.0 ** 2)
    zetaN = (X[dim - 1] - hN) / (1.0 - hN / 10e3)
    fz = 1.0 + 1e-5 * zetaN
    tp = op.observe_theta(st, fz, p0)
    assert np.allclose(_rho_from_theta(st, tp, pts[:, -1]), 1.0 + 1e-5 * zeta, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dim,k", [(2, 3), (3, 2)])
def test_background_has_no_perturbation_at_nodes(oracle_mod, dim, k):
    op = oracle_mod.OracleOperator(_mesh(dim, k))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5)
    rho_b, p_b, _ = op.background(st)
    X = op.node_coords()
    n1 = k + 1
    ii = np.arange(op.N) % (n1 ** dim)
    interior = np.all([((ii // n1 ** d) % n1 > 0) & ((ii // n1 ** d) % n1 < k) for d in range(dim)], axis=0)
    op.set_observation_points(X[:, interior].T)
    tp = op.observe_theta(st, rho_b, p_b)
    assert np.max(np.abs(tp)) <= 1e-10


@pytest.mark.parametrize("dim,k", [(2, 3), (3, 2)])
def test_transpose_matches_differences(oracle_mod, dim, k):
    op = oracle_mod.OracleOperator(_mesh(dim, k))
    st = make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, theta_sigma=0.3, seed=4)
    rho, _, p = op.init_hill_state(st)
    rng = np.random.default_rng(9)
    pts, _ = _random_points(rng, 64, dim)
    op.set_observation_points(pts)
    wbar = rng.standard_normal(64)
    rb, pb = op.observe_theta_transpose(st, rho, p, wbar)
    dr, dp = 1e-3 * rng.standard_normal(op.N), 1e2 * rng.standard_normal(op.N)
    h = 1e-2
    fd = (op.observe_theta(st, rho + h * dr, p + h * dp) - op.observe_theta(st, rho - h * dr, p - h * dp)) / (2 * h)
    a, b = float(wbar @ fd), float(rb @ dr + pb @ dp)
    assert abs(a - b) <= 1e-8 * abs(a), (a, b)
    with pytest.raises(RuntimeError):
        op.set_observation_points(np.array([[1e9] * dim]))
