"""Host-side argument checks of the ctypes mirror (no GPU): every vector argument is checked for
device placement, precision, contiguity and its exact element count before the C-ABI is called,
so an undersized buffer raises instead of becoming an out-of-bounds access on the device."""
import pytest
import torch

from atmospheric_flows_amd import AflError, EulerOperator
from atmospheric_flows_amd.euler_operator import _ptr


class FakeDeviceTensor:
    """Stands in for a CUDA tensor on a machine without a GPU (only the attributes _ptr reads)."""

    def __init__(self, n, dtype=torch.float64, cuda=True, contiguous=True):
        self.n, self.dtype, self.is_cuda, self._contig = n, dtype, cuda, contiguous

    def element_size(self):
        return torch.empty(0, dtype=self.dtype).element_size()

    def is_contiguous(self):
        return self._contig

    def numel(self):
        return self.n

    def data_ptr(self):
        return 0x1000


class FakeLib:
    """Records C-ABI calls; a test fails if any call reaches it."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def call(*args):
            self.calls.append(name)
            return 0
        return call


def host_only_operator(N=12, dim=2):
    op = object.__new__(EulerOperator)
    op.L, op.h, op.N, op.dim, op.npts = FakeLib(), None, N, dim, 5
    op.m = lambda: N  # NS_stage 1 (mass) rows
    return op


def test_ptr_checks_placement_precision_layout_and_count():
    assert _ptr(FakeDeviceTensor(7), 8, 7).value == 0x1000
    with pytest.raises(AflError, match="expected 7 elements"):
        _ptr(FakeDeviceTensor(6), 8, 7)
    with pytest.raises(AflError, match="contiguous float64"):
        _ptr(FakeDeviceTensor(7, cuda=False), 8, 7)
    with pytest.raises(AflError, match="contiguous float64"):
        _ptr(FakeDeviceTensor(7, dtype=torch.float32), 8, 7)
    with pytest.raises(AflError, match="contiguous float64"):
        _ptr(FakeDeviceTensor(7, contiguous=False), 8, 7)
    assert _ptr(None) is None


@pytest.mark.parametrize("method,args", [
    ("imex_step", (12, 23, 12)),                   # u one short of dim * N
    ("imex_step", (11, 24, 12)),                   # rho one short
    ("vmult_pressure", (23, 12)),                  # B: dim*N <- N
    ("vmult_enthalpy", (12, 25)),                  # C: N <- dim*N
    ("vmult_schur_velocity", (24, 13)),
    ("schur_solve", (12, 10)),
    ("set_u_fixed", (12,)),
    ("set_rho_for_fixed", (24,)),
    ("mg_vmult", (12, 11)),
    ("vmult_pressure_transpose", (12, 12)),
    ("vmult_enthalpy_tangent", (12, 24, 23)),
])
def test_wrong_sizes_raise_before_any_library_call(method, args):
    op = host_only_operator()
    tensors = [FakeDeviceTensor(a) for a in args]
    with pytest.raises(AflError):
        getattr(op, method)(*tensors)
    assert op.L.calls == []


def test_explicit_stage_sources_are_sized_by_position():
    op = host_only_operator()
    N, NU = 12, 24
    good = [FakeDeviceTensor(n) for n in (N, NU, N, N)]  # E2 stage 1: rho, u, p, rho_new
    op.vmult_rhs_velocity_fixed(FakeDeviceTensor(NU), good)
    assert op.L.calls == ["afl_vmult_rhs_velocity_fixed"]
    op.L.calls.clear()
    bad = [FakeDeviceTensor(n) for n in (N, N, N, N)]  # u passed as a scalar field
    with pytest.raises(AflError, match="expected 24 elements"):
        op.vmult_rhs_velocity_fixed(FakeDeviceTensor(NU), bad)
    e1 = [FakeDeviceTensor(n) for n in (N, NU, N, NU)]  # E1 stage 2 alternates (rho_j, u_j)
    op.vmult_rhs_rho_update(FakeDeviceTensor(N), e1)
    with pytest.raises(AflError):
        op.vmult_rhs_rho_update(FakeDeviceTensor(NU), e1)  # destination is a scalar field
    assert op.L.calls == ["afl_vmult_rhs_rho_update"]
