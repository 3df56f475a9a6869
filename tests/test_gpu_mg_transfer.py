"""The p-multigrid transfers on the GPU (afl_mg_transfer), pinned by what the mathematics fixes:
  - prolongation is the nodal interpolation of the coarse polynomial, so a field that is a degree-kc
    polynomial in each reference coordinate of every cell (random per cell) comes out as the same
    polynomial at the fine nodes, to rounding;
  - restriction is its transpose: <P c, f> = <c, R f> for random c, f (Euclidean inner products);
  - P maps the constant 1 to 1 and is injective (P^T P is symmetric positive definite).
The V-cycle uses the same matrices and kernel (interp_matrix, k_tensor_map).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from numpy.polynomial import legendre as npleg

from atmospheric_flows_amd import EulerOperator, make_mesh


def say(msg):
    print(msg, flush=True)


def gl_nodes(k):
    inner = np.sort(npleg.Legendre.basis(k).deriv().roots().real)
    return np.concatenate([[0.0], (inner + 1) / 2, [1.0]])


def _poly_values(coef, k_nodes, dim):
    """coef: (ncells, kc+1, ..., kc+1) monomial coefficients per cell; values at the degree-k_nodes GL nodes,
    node-lexicographic with x fastest."""
    x = gl_nodes(k_nodes)
    kc1 = coef.shape[1]
    V = np.vander(x, kc1, increasing=True)  # V[i, m] = x_i^m
    if dim == 2:
        vals = np.einsum("cmn,im,jn->cji", coef, V, V)   # m: x power, n: z power; out (cell, iz, ix)
    else:
        vals = np.einsum("cmno,im,jn,lo->clji", coef, V, V, V)  # out (cell, iz, iy, ix)
    return vals.reshape(coef.shape[0], -1).ravel()


@pytest.mark.parametrize("dim,kf,nc", [(2, 4, (6, 1, 4)), (2, 3, (5, 1, 3)), (3, 4, (3, 2, 2)), (3, 3, (3, 3, 2))])
def test_prolongation_reproduces_coarse_polynomials_and_restriction_is_its_transpose(dim, kf, nc):
    import torch
    mesh = make_mesh(dim=dim, degree=kf, n_cells=nc, x_range=(-6e3, 6e3), y_range=(-6e3, 6e3), z_top=8e3,
                     hill_height=400.0)
    op = EulerOperator(mesh)
    ncells = int(np.prod([nc[0], nc[1] if dim == 3 else 1, nc[2]]))
    rng = np.random.default_rng(10 * dim + kf)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    for kc in range(1, kf):
        say(f"[gpu mg transfer] dim={dim} degrees {kc} -> {kf}")
        Nc, Nf = ncells * (kc + 1) ** dim, ncells * (kf + 1) ** dim
        coef = rng.standard_normal((ncells,) + (kc + 1,) * dim)
        coarse = _poly_values(coef, kc, dim)
        fine = _poly_values(coef, kf, dim)
        out = op.mg_transfer(kc, kf, False, dev(coarse), torch.zeros(Nf, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        err = np.abs(got - fine).max() / np.abs(fine).max()
        say(f"  polynomial reproduction: max rel {err:.2e}")
        assert err <= 1e-13
        ones = op.mg_transfer(kc, kf, False, dev(np.ones(Nc)), torch.zeros(Nf, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        assert np.abs(ones.cpu().numpy() - 1.0).max() <= 1e-14
        c = rng.standard_normal(Nc)
        f = rng.standard_normal(Nf)
        Pc = op.mg_transfer(kc, kf, False, dev(c), torch.zeros(Nf, dtype=torch.float64, device="cuda"))
        Rf = op.mg_transfer(kc, kf, True, dev(f), torch.zeros(Nc, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        Pc, Rf = Pc.cpu().numpy(), Rf.cpu().numpy()
        lhs, rhs = float(Pc @ f), float(c @ Rf)
        say(f"  <Pc, f> = {lhs:.15e}, <c, Rf> = {rhs:.15e}")
        assert abs(lhs - rhs) <= 1e-13 * np.linalg.norm(Pc) * np.linalg.norm(f)
        # P^T P c: positive definite (injective prolongation)
        PtPc = op.mg_transfer(kc, kf, True, dev(Pc), torch.zeros(Nc, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        assert float(c @ PtPc.cpu().numpy()) > 0.0
    with pytest.raises(Exception):
        op.mg_transfer(kf, kf, False, dev(np.zeros(ncells * (kf + 1) ** dim)),
                       torch.zeros(ncells * (kf + 1) ** dim, dtype=torch.float64, device="cuda"))
