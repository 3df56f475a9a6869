"""The boundary's layout adapter (afl_pack_dofs / afl_unpack_dofs) in three caller layouts.

Caller index of (cell c, component a, node i) = cell_perm[c] cell_stride + a comp_stride + i node_stride:
  - component-blocked: cell_stride = dim (k+1)^d, comp_stride = (k+1)^d, node_stride = 1. This is
    deal.II's DoF order for FESystem(FE_DGQ(k), dim) on a DoFHandler: FE_DGQ has all its DoFs on the
    cell interior, and FESystem numbers an object's DoFs base element by base element (copy 0's
    (k+1)^d nodes, then copy 1's, ...), the nodes of each in FE_DGQ's lexicographic order;
  - interleaved: cell_stride = dim (k+1)^d, comp_stride = 1, node_stride = dim;
  - component-blocked with a random cell permutation (deal.II's active-cell order is not the
    structured x-y-z order), written into a buffer with padding between cells.
Pack and unpack are exact inverses; pack -> afl_vmult -> unpack equals the apply in the caller's
layout (bitwise: the same arithmetic on the same values); and out-of-range layouts are refused
before any device access.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from atmospheric_flows_amd import AflError, EulerOperator, make_mesh, make_hill_state


def say(msg):
    print(msg, flush=True)


def _caller_layout(afl, N, npc, ncomp, perm, cs, as_, ns, size):
    out = np.full(size, np.nan)
    ncells = N // npc
    for c in range(ncells):
        for a in range(ncomp):
            out[perm[c] * cs + a * as_ + np.arange(npc) * ns] = afl[a * N + c * npc: a * N + (c + 1) * npc]
    return out


@pytest.mark.parametrize("dim,k,nc", [(2, 3, (5, 1, 3)), (3, 2, (3, 2, 2))])
def test_pack_apply_unpack_in_caller_layouts(dim, k, nc):
    import torch
    say(f"[gpu layout] dim={dim} k={k} cells={nc}")
    mesh = make_mesh(dim=dim, degree=k, n_cells=nc, x_range=(-6e3, 6e3), y_range=(-6e3, 6e3), z_top=8e3,
                     hill_height=400.0)
    op = EulerOperator(mesh, dt=1.5)
    N = op.N
    npc = (k + 1) ** dim
    ncells = N // npc
    rng = np.random.default_rng(3 + dim)
    rho, _, _ = op.init_hill_state(make_hill_state(seed=4))
    op.set_rho_for_fixed(rho)
    op.set_NS_stage(3)  # A: rho-weighted velocity mass matrix, a dim-component apply
    u = rng.standard_normal(dim * N)
    ref = op.vmult(op.new_vector(), torch.from_numpy(u).cuda())
    torch.cuda.synchronize()
    ref = ref.cpu().numpy()
    ident = np.arange(ncells)
    shuffled = rng.permutation(ncells)
    layouts = {
        "component-blocked (FESystem(FE_DGQ(k), dim))": (ident, dim * npc, npc, 1, dim * N, False),
        "interleaved": (ident, dim * npc, 1, dim, dim * N, False),
        "permuted cells, padded": (shuffled, dim * npc + 5, npc, 1, ncells * (dim * npc + 5), True),
    }
    for name, (perm, cs, as_, ns, size, use_perm) in layouts.items():
        caller = _caller_layout(u, N, npc, dim, perm, cs, as_, ns, size)
        expect = _caller_layout(ref, N, npc, dim, perm, cs, as_, ns, size)
        dperm = torch.from_numpy(perm.astype(np.int64)).cuda() if use_perm else None
        dcaller = torch.from_numpy(np.nan_to_num(caller, nan=0.0)).cuda()
        packed = op.pack_dofs(dim, dcaller, dperm, cs, as_, ns, op.new_vector())
        torch.cuda.synchronize()
        assert np.array_equal(packed.cpu().numpy(), u), name
        y = op.vmult(op.new_vector(), packed)
        out = torch.zeros(size, dtype=torch.float64, device="cuda")
        op.unpack_dofs(dim, y, dperm, cs, as_, ns, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        mask = ~np.isnan(expect)
        assert np.array_equal(got[mask], expect[mask]), name
        assert not got[~mask].any(), name  # padding untouched
        back = op.unpack_dofs(dim, packed, dperm, cs, as_, ns, torch.zeros(size, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        assert np.array_equal(back.cpu().numpy()[mask], caller[mask]), name
        say(f"  {name}: pack, apply, unpack exact")
    # a layout that reaches past the buffer, a short permutation, a negative stride: refused on the host
    small = torch.zeros(dim * N - 1, dtype=torch.float64, device="cuda")
    with pytest.raises(AflError):
        op.pack_dofs(dim, small, None, dim * npc, npc, 1, op.new_vector())
    with pytest.raises(AflError):
        op.pack_dofs(dim, torch.zeros(dim * N, dtype=torch.float64, device="cuda"),
                     torch.zeros(ncells - 1, dtype=torch.int64, device="cuda"), dim * npc, npc, 1, op.new_vector())
    with pytest.raises(AflError):
        op.pack_dofs(dim, torch.zeros(dim * N, dtype=torch.float64, device="cuda"), None, dim * npc, -1, 1,
                     op.new_vector())
