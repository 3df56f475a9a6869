"""Host-side mirror of the reference's EULEROperator (euler_operator.h:19-98) over the C-ABI
of libafl_hip.so. Device memory and streams come from PyTorch (ROCm); every compute call
goes to the hand-written HIP kernels. There is no CPU fallback: constructing an operator
without the built library raises.

Method names, argument meaning and stage semantics follow the reference:
  set_dt / set_Mach / set_Froude / set_HYPERBOLIC_stage / set_NS_stage      (:34-53)
  set_rho_for_fixed / set_pres_fixed / set_u_fixed                           (:310-339)
  vmult_rhs_rho_update / vmult_rhs_velocity_fixed / vmult_rhs_pressure       (:55-62)
  vmult (apply_add dispatch on NS_stage), vmult_pressure (B), vmult_enthalpy (C)
  compute_diagonal / get_matrix_diagonal_inverse                             (:2589-2626)
Vectors are torch.float64 CUDA tensors in the layout documented in include/afl.h.
"""
import ctypes as C
import os

from ._abi import MeshDesc, SolverParams, StepInfo, HillState, Sponge, Diagnostics, AdjointStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafl_hip.so")
_lib = None


class AflError(RuntimeError):
    pass


def load_library():
    """Load libafl_hip.so (built in-tree by atmospheric_flows_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AflError(f"{LIB_PATH} is missing: build it with `python -m atmospheric_flows_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, dp, i64p = C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)
    L.afl_create.argtypes = [C.POINTER(MeshDesc), C.c_double, C.c_double, C.c_double, C.POINTER(C.c_void_p)]
    L.afl_destroy.argtypes = [vp]
    L.afl_last_error.restype = C.c_char_p
    L.afl_set_stream.argtypes = [vp, C.c_void_p]
    L.afl_n_dofs.argtypes = [vp, C.c_int, i64p]
    L.afl_m.argtypes = [vp, i64p]
    for n in ("afl_set_dt", "afl_set_mach", "afl_set_froude"):
        getattr(L, n).argtypes = [vp, C.c_double]
    for n in ("afl_set_hyperbolic_stage", "afl_set_ns_stage"):
        getattr(L, n).argtypes = [vp, C.c_uint]
    for n in ("afl_set_rho_for_fixed", "afl_set_pres_fixed", "afl_set_u_fixed"):
        getattr(L, n).argtypes = [vp, dp]
    L.afl_set_solver_params.argtypes = [vp, C.POINTER(SolverParams)]
    L.afl_set_sponge.argtypes = [vp, C.POINTER(Sponge)]
    for n in ("afl_vmult_rhs_rho_update", "afl_vmult_rhs_velocity_fixed", "afl_vmult_rhs_pressure"):
        getattr(L, n).argtypes = [vp, dp, C.POINTER(C.c_void_p), C.c_int]
    for n in ("afl_vmult", "afl_vmult_pressure", "afl_vmult_enthalpy", "afl_solve_mass", "afl_vmult_schur_velocity",
              "afl_mg_vmult", "afl_mg_level_vmult", "afl_vmult_schur_velocity_f32"):
        getattr(L, n).argtypes = [vp, dp, dp]
    L.afl_compute_diagonal.argtypes = [vp]
    L.afl_get_matrix_diagonal_inverse.argtypes = [vp, dp]
    L.afl_schur_solve.argtypes = [vp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.afl_imex_step.argtypes = [vp, dp, dp, dp, C.POINTER(StepInfo)]
    L.afl_init_hill_state.argtypes = [vp, C.POINTER(HillState), dp, dp, dp]
    L.afl_background.argtypes = [vp, C.POINTER(HillState), dp, dp, dp]
    L.afl_write_vtu.argtypes = [vp, C.c_char_p, dp, dp, dp, C.c_int]
    L.afl_kernel_timer_start.argtypes = [vp, C.c_int, C.c_int]
    L.afl_kernel_timer_stop.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.afl_kernel_timer_result.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
    L.afl_compute_diagnostics.argtypes = [vp, dp, dp, dp, C.POINTER(Diagnostics)]
    L.afl_l2_difference.argtypes = [vp, C.c_int, dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.afl_l2_difference_noalloc.argtypes = [vp, C.c_int, dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    for n in ("afl_vjp_rhs_rho_update", "afl_vjp_rhs_velocity_fixed", "afl_vjp_rhs_pressure"):
        getattr(L, n).argtypes = [vp, dp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]
    L.afl_vmult_pressure_transpose.argtypes = [vp, dp, dp]
    L.afl_vjp_rhs_pressure_u_fixed.argtypes = [vp, dp, dp, dp, dp]
    L.afl_set_observation_points.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.afl_observe_theta.argtypes = [vp, C.POINTER(HillState), dp, dp, dp]
    L.afl_observe_theta_transpose.argtypes = [vp, C.POINTER(HillState), dp, dp, dp, dp, dp]
    L.afl_adjoint_run.argtypes = [vp, dp, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, C.POINTER(AdjointStats)]
    L.afl_set_adjoint_stage_sets.argtypes = [vp, C.c_int]
    L.afl_imex_step_vjp.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.POINTER(StepInfo)]
    for n in ("afl_vmult_enthalpy_coefficient_transpose", "afl_vjp_rho_for_fixed",
              "afl_vmult_enthalpy_tangent", "afl_vmult_enthalpy_tangent_transpose", "afl_vmult_schur_tangent",
              "afl_vmult_schur_tangent_transpose"):
        getattr(L, n).argtypes = [vp, dp, dp, dp]
    L.afl_mg_transfer.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
    L.afl_set_verbose.argtypes = [vp, C.c_int]
    for n in ("afl_section_timer_start", "afl_section_timer_stop"):
        getattr(L, n).argtypes = [vp]
    L.afl_section_timer_result.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    for n in ("afl_pack_dofs", "afl_unpack_dofs"):
        getattr(L, n).argtypes = [vp, C.c_int, vp, vp, C.c_int64, C.c_int64, C.c_int64, vp]
    _lib = L
    return L


def _ptr(t, size=8, numel=None):
    """Device address of a contiguous floating-point tensor of element size `size`, checked to
    hold exactly `numel` elements when given (an undersized buffer would become an out-of-bounds
    access on the device)."""
    if t is None:
        return None
    if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == size and t.is_contiguous()):
        raise AflError(f"expected a contiguous float{8 * size} device tensor")
    if numel is not None and t.numel() != numel:
        raise AflError(f"expected {numel} elements, got a tensor of {t.numel()}")
    return C.c_void_p(t.data_ptr())


class EulerOperator:
    """EULEROperator on one MI355X; one handle per host thread (not reentrant)."""

    def __init__(self, mesh: MeshDesc, mach=1.0, froude=1.0 / 9.80616 ** 0.5, dt=1.0, stream=None):
        import torch
        self.torch = torch
        self.L = load_library()
        self.mesh = mesh
        h = C.c_void_p()
        self._chk(self.L.afl_create(C.byref(mesh), mach, froude, dt, C.byref(h)))
        self.h = h
        s = stream if stream is not None else torch.cuda.current_stream()
        self._chk(self.L.afl_set_stream(self.h, C.c_void_p(s.cuda_stream)))
        self.dim = mesh.dim
        n = C.c_int64()
        self._chk(self.L.afl_n_dofs(self.h, 1, C.byref(n)))
        self.N = n.value

    def set_stream(self, stream):
        """Order every later call on this torch stream (afl_set_stream)."""
        self._chk(self.L.afl_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.afl_destroy(self.h)
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise AflError(f"afl error {rc}: {self.L.afl_last_error().decode()}")

    # device addresses checked against the field sizes: scalar (N), velocity (dim N), one value per
    # observation point, and the FP32 variants
    def _s(self, t): return _ptr(t, 8, self.N)
    def _v(self, t): return _ptr(t, 8, self.dim * self.N)
    def _sv(self, t, vec): return self._v(t) if vec else self._s(t)
    def _obs(self, t): return _ptr(t, 8, self.npts)

    # ---- allocation helpers -----------------------------------------------------------------
    def new_scalar(self):
        return self.torch.zeros(self.N, dtype=self.torch.float64, device="cuda")

    def new_vector(self):
        return self.torch.zeros(self.dim * self.N, dtype=self.torch.float64, device="cuda")

    # ---- setters -----------------------------------------------------------------------------
    def set_dt(self, v): self._chk(self.L.afl_set_dt(self.h, v))
    def set_Mach(self, v): self._chk(self.L.afl_set_mach(self.h, v))
    def set_Froude(self, v): self._chk(self.L.afl_set_froude(self.h, v))
    def set_HYPERBOLIC_stage(self, s): self._chk(self.L.afl_set_hyperbolic_stage(self.h, s))
    def set_NS_stage(self, s): self._chk(self.L.afl_set_ns_stage(self.h, s))
    def set_rho_for_fixed(self, t): self._chk(self.L.afl_set_rho_for_fixed(self.h, self._s(t)))
    def set_pres_fixed(self, t): self._chk(self.L.afl_set_pres_fixed(self.h, self._s(t)))
    def set_u_fixed(self, t): self._chk(self.L.afl_set_u_fixed(self.h, self._v(t)))
    def set_solver_params(self, sp: SolverParams): self._chk(self.L.afl_set_solver_params(self.h, C.byref(sp)))
    def set_sponge(self, s: Sponge): self._chk(self.L.afl_set_sponge(self.h, C.byref(s)))

    def n_dofs(self, dof_index):
        n = C.c_int64()
        self._chk(self.L.afl_n_dofs(self.h, dof_index, C.byref(n)))
        return n.value

    def m(self):
        n = C.c_int64()
        self._chk(self.L.afl_m(self.h, C.byref(n)))
        return n.value

    # ---- explicit-stage applies ---------------------------------------------------------------
    # positional sources: E1 alternates (rho_j, u_j); E2 and E3 take (rho, u, p) triples and a trailing
    # (rho, u) pair, so position i is a velocity exactly when i % 3 == 1
    @staticmethod
    def _is_vec(kind, i):
        return i % 2 == 1 if kind == "E1" else i % 3 == 1

    def _rhs(self, fn, kind, dst, srcs, dst_vec):
        ptrs = [self._sv(t, self._is_vec(kind, i)) for i, t in enumerate(srcs)]
        arr = (C.c_void_p * len(srcs))(*[p.value for p in ptrs])
        self._chk(fn(self.h, self._sv(dst, dst_vec), arr, len(srcs)))
        return dst

    def vmult_rhs_rho_update(self, dst, srcs): return self._rhs(self.L.afl_vmult_rhs_rho_update, "E1", dst, srcs, False)
    def vmult_rhs_velocity_fixed(self, dst, srcs):
        return self._rhs(self.L.afl_vmult_rhs_velocity_fixed, "E2", dst, srcs, True)
    def vmult_rhs_pressure(self, dst, srcs): return self._rhs(self.L.afl_vmult_rhs_pressure, "E3", dst, srcs, False)

    # ---- linear blocks -------------------------------------------------------------------------
    def vmult(self, dst, src):
        n = self.m()
        self._chk(self.L.afl_vmult(self.h, _ptr(dst, 8, n), _ptr(src, 8, n)))
        return dst

    def vmult_pressure(self, dst, src):
        self._chk(self.L.afl_vmult_pressure(self.h, self._v(dst), self._s(src)))
        return dst

    def vmult_enthalpy(self, dst, src):
        self._chk(self.L.afl_vmult_enthalpy(self.h, self._s(dst), self._v(src)))
        return dst

    def vmult_schur_velocity(self, w, p):
        self._chk(self.L.afl_vmult_schur_velocity(self.h, self._v(w), self._s(p)))
        return w

    # ---- discrete-adjoint building blocks -------------------------------------------------------
    def _vjp(self, fn, kind, rbar, srcs, outs, rbar_vec):
        if len(outs) != len(srcs):
            raise AflError("vector-Jacobian product: one output per source")
        sp = [self._sv(t, self._is_vec(kind, i)) for i, t in enumerate(srcs)]
        op_ = [self._sv(t, self._is_vec(kind, i)) for i, t in enumerate(outs)]
        sarr = (C.c_void_p * len(srcs))(*[p.value for p in sp])
        oarr = (C.c_void_p * len(outs))(*[p.value for p in op_])
        self._chk(fn(self.h, self._sv(rbar, rbar_vec), oarr, sarr, len(srcs)))
        return outs

    def vjp_rhs_rho_update(self, rbar, srcs, outs):
        """outs[i] = (d rhs_rho / d srcs[i])^T rbar for the E1 sources of the current stage."""
        return self._vjp(self.L.afl_vjp_rhs_rho_update, "E1", rbar, srcs, outs, False)

    def vjp_rhs_pressure(self, rbar, srcs, outs):
        """outs[i] = (d rhs_p / d srcs[i])^T rbar for the E3 sources of the current stage."""
        return self._vjp(self.L.afl_vjp_rhs_pressure, "E3", rbar, srcs, outs, False)

    def vjp_rhs_velocity_fixed(self, rbar, srcs, outs):
        """outs[i] = (d rhs_u / d srcs[i])^T rbar for the E2 sources of the current stage."""
        return self._vjp(self.L.afl_vjp_rhs_velocity_fixed, "E2", rbar, srcs, outs, True)

    def imex_step_vjp(self, rho, u, p, rho_bar_out, u_bar_out, p_bar_out):
        """Adjoint of one IMEX step at the input state (rho, u, p): returns (rho_bar, u_bar, p_bar),
        the transposed step Jacobian applied to the output adjoint, and the recomputed step's info."""
        import torch
        rb, ub, pb = torch.empty_like(rho), torch.empty_like(u), torch.empty_like(p)
        info = StepInfo()
        self._chk(self.L.afl_imex_step_vjp(self.h, self._s(rho), self._v(u), self._s(p), self._s(rho_bar_out),
                                           self._v(u_bar_out), self._s(p_bar_out), self._s(rb), self._v(ub),
                                           self._s(pb), C.byref(info)))
        return rb, ub, pb, info

    def adjoint_run(self, rho0, u0, p0, nsteps, snapshots, rho_bar_T, u_bar_T, p_bar_T):
        """Gradient of <xbar_T, x_T> with respect to the initial state over nsteps steps, binomial
        checkpointing with `snapshots` snapshots. Returns (rho_bar_0, u_bar_0, p_bar_0, stats dict)."""
        import torch
        rb, ub, pb = torch.empty_like(rho0), torch.empty_like(u0), torch.empty_like(p0)
        st = AdjointStats()
        self._chk(self.L.afl_adjoint_run(self.h, self._s(rho0), self._v(u0), self._s(p0), int(nsteps), int(snapshots),
                                         self._s(rho_bar_T), self._v(u_bar_T), self._s(p_bar_T), self._s(rb),
                                         self._v(ub), self._s(pb),
                                         C.byref(st)))
        return rb, ub, pb, {"forward_steps": st.forward_steps, "adjoint_steps": st.adjoint_steps,
                            "snapshots_used": st.snapshots_used, "recomputed_steps": st.recomputed_steps,
                            "stage_sets_used": st.stage_sets_used,
                            "stage_solve_iterations": st.stage_solve_iterations}

    def set_adjoint_stage_sets(self, keep):
        """Whether adjoint_run keeps the stage vectors of each advance's last step (default True)."""
        self._chk(self.L.afl_set_adjoint_stage_sets(self.h, 1 if keep else 0))

    # ---- observation operator -------------------------------------------------------------------
    def set_observation_points(self, xyz):
        """Points as a host array of shape (npts, dim): x[, y], z."""
        import numpy as np
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.npts = xyz.size // self.dim
        self._chk(self.L.afl_set_observation_points(self.h, self.npts, xyz.ctypes.data_as(C.POINTER(C.c_double))))

    def observe_theta(self, st, rho, p):
        import torch
        out = torch.empty(self.npts, dtype=torch.float64, device="cuda")
        self._chk(self.L.afl_observe_theta(self.h, C.byref(st), self._s(rho), self._s(p), self._obs(out)))
        return out

    def observe_theta_transpose(self, st, rho, p, wbar):
        import torch
        rb, pb = torch.empty_like(rho), torch.empty_like(p)
        self._chk(self.L.afl_observe_theta_transpose(self.h, C.byref(st), self._s(rho), self._s(p), self._obs(wbar),
                                                     self._s(rb), self._s(pb)))
        return rb, pb

    def vmult_pressure_transpose(self, dst, src):
        self._chk(self.L.afl_vmult_pressure_transpose(self.h, self._s(dst), self._v(src)))
        return dst

    def vmult_enthalpy_tangent(self, dst, dw, w0):
        self._chk(self.L.afl_vmult_enthalpy_tangent(self.h, self._s(dst), self._v(dw), self._v(w0)))
        return dst

    def vmult_enthalpy_tangent_transpose(self, dst, cbar, w0):
        self._chk(self.L.afl_vmult_enthalpy_tangent_transpose(self.h, self._v(dst), self._s(cbar), self._v(w0)))
        return dst

    def vmult_enthalpy_coefficient_transpose(self, dst, cbar, w):
        self._chk(self.L.afl_vmult_enthalpy_coefficient_transpose(self.h, self._s(dst), self._s(cbar), self._v(w)))
        return dst

    def vjp_rho_for_fixed(self, dst, ybar, u):
        self._chk(self.L.afl_vjp_rho_for_fixed(self.h, self._s(dst), self._v(ybar), self._v(u)))
        return dst

    def vjp_rhs_pressure_u_fixed(self, dst, rbar, rho_new, u_fixed):
        self._chk(self.L.afl_vjp_rhs_pressure_u_fixed(self.h, self._v(dst), self._s(rbar), self._s(rho_new),
                                                      self._v(u_fixed)))
        return dst

    def vmult_schur_tangent(self, dst, src, p0):
        self._chk(self.L.afl_vmult_schur_tangent(self.h, self._s(dst), self._s(src), self._s(p0)))
        return dst

    def vmult_schur_tangent_transpose(self, dst, src, p0):
        self._chk(self.L.afl_vmult_schur_tangent_transpose(self.h, self._s(dst), self._s(src), self._s(p0)))
        return dst

    def mg_vmult(self, dst, src):
        """Schur preconditioner apply (p-multigrid V-cycle of the selected schur_precond)."""
        self._chk(self.L.afl_mg_vmult(self.h, self._s(dst), self._s(src)))
        return dst

    def mg_level_vmult(self, dst, src):
        """The finest multigrid level's Schur operator (level precision and quadrature) applied to src."""
        self._chk(self.L.afl_mg_level_vmult(self.h, self._s(dst), self._s(src)))
        return dst

    def vmult_schur_velocity_f32(self, w, p):
        """FP32 fine-level w = A^-1 B p of the FP32 V-cycle (float32 tensors)."""
        self._chk(self.L.afl_vmult_schur_velocity_f32(self.h, _ptr(w, 4, self.dim * self.N), _ptr(p, 4, self.N)))
        return w

    # ---- benchmark instrumentation -------------------------------------------------------------
    def set_verbose(self, level): self._chk(self.L.afl_set_verbose(self.h, int(level)))

    SECTIONS = ("Update density", "Fixed point pressure", "Update velocity", "Update pressure")

    def section_timer_start(self): self._chk(self.L.afl_section_timer_start(self.h))

    def section_timer_stop(self): self._chk(self.L.afl_section_timer_stop(self.h))

    def section_timer_result(self, section):
        """(timed phases, summed milliseconds) of one afl_imex_step section (index into SECTIONS)."""
        n, ms = C.c_int(), C.c_double()
        self._chk(self.L.afl_section_timer_result(self.h, section, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def kernel_timer_start(self, kernels, max_launches=100000):
        """HIP events around every launch of the kernels in `kernels` (an or-ed mask of AFL_TIME_*)."""
        self._chk(self.L.afl_kernel_timer_start(self.h, kernels, max_launches))

    def kernel_timer_stop(self):
        """(launches, total milliseconds) of every timed kernel since kernel_timer_start."""
        n, ms = C.c_int(), C.c_double()
        self._chk(self.L.afl_kernel_timer_stop(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def kernel_timer_result(self, kernel, stage=0):
        """(launches, total milliseconds, total algorithmic bytes) of one timed kernel after
        kernel_timer_stop, over IMEX stage `stage` (0: all)."""
        n, ms, b = C.c_int(), C.c_double(), C.c_double()
        self._chk(self.L.afl_kernel_timer_result(self.h, kernel, stage, C.byref(n), C.byref(ms), C.byref(b)))
        return n.value, ms.value, b.value

    # ---- diagnostics (Atmospheric_Flow.cc:925-1087, 1225-1235) ---------------------------------
    def diagnostics(self, rho, u=None, p=None):
        """Density bounds, maximum celerity and speed, minimal cell diameter and CFL numbers."""
        d = Diagnostics()
        self._chk(self.L.afl_compute_diagnostics(self.h, self._s(rho), None if u is None else self._v(u),
                                                   None if p is None else self._s(p), C.byref(d)))
        return d.as_dict()

    def write_vtu(self, path, rho, u, p, high_order=False):
        """VTU file of the state (point data rho, u, p on the degree-k patch points of each cell)."""
        self._chk(self.L.afl_write_vtu(self.h, os.fsencode(path), self._s(rho), self._v(u), self._s(p),
                                         1 if high_order else 0))

    def l2_difference(self, ncomp, a, b=None):
        """(||a - b||_L2, ||b||_L2) of ncomp-component fields."""
        e, r = C.c_double(), C.c_double()
        self._chk(self.L.afl_l2_difference(self.h, ncomp, _ptr(a, 8, ncomp * self.N),
                                             None if b is None else _ptr(b, 8, ncomp * self.N), C.byref(e), C.byref(r)))
        return e.value, r.value

    def l2_difference_noalloc(self, ncomp, a, b=None):
        """l2_difference through the entry point that allocates no device memory."""
        e, r = C.c_double(), C.c_double()
        self._chk(self.L.afl_l2_difference_noalloc(self.h, ncomp, _ptr(a, 8, ncomp * self.N),
                                                     None if b is None else _ptr(b, 8, ncomp * self.N),
                                                     C.byref(e), C.byref(r)))
        return e.value, r.value

    def mg_transfer(self, k_coarse, k_fine, transpose, src, dst):
        """Multigrid transfer: dst (k_fine) = P src (k_coarse), or with transpose dst (k_coarse) = P^T src."""
        ncells = self.N // (self.mesh.degree + 1) ** self.dim
        nc, nf = ncells * (k_coarse + 1) ** self.dim, ncells * (k_fine + 1) ** self.dim
        n_in, n_out = (nf, nc) if transpose else (nc, nf)
        self._chk(self.L.afl_mg_transfer(self.h, k_coarse, k_fine, 1 if transpose else 0, _ptr(src, 8, n_in),
                                         _ptr(dst, 8, n_out)))
        return dst

    # ---- caller layouts (afl_pack_dofs / afl_unpack_dofs) ---------------------------------------
    def _layout_args(self, ncomp, caller, cell_perm, cell_stride, comp_stride, node_stride):
        """Checks a caller layout against the buffers: caller index of (cell c, component a, node i) =
        cell_perm[c] cell_stride + a comp_stride + i node_stride must stay inside `caller`."""
        torch = self.torch
        npc = (self.mesh.degree + 1) ** self.dim
        ncells = self.N // npc
        if ncomp not in (1, self.dim):
            raise AflError("ncomp must be 1 or dim")
        if min(cell_stride, comp_stride, node_stride) < 0:
            raise AflError("strides must be non-negative")
        if cell_perm is None:
            top = ncells - 1
            pp = None
        else:
            if not (cell_perm.is_cuda and cell_perm.dtype == torch.int64 and cell_perm.is_contiguous()
                    and cell_perm.numel() == ncells):
                raise AflError(f"cell_perm must be a contiguous int64 device tensor of {ncells} entries")
            lo, top = int(cell_perm.min()), int(cell_perm.max())
            if lo < 0:
                raise AflError("cell_perm holds a negative cell index")
            pp = C.c_void_p(cell_perm.data_ptr())
        extent = top * cell_stride + (ncomp - 1) * comp_stride + (npc - 1) * node_stride + 1
        if not (caller.is_cuda and caller.dtype == torch.float64 and caller.is_contiguous()):
            raise AflError("expected a contiguous float64 device tensor")
        if caller.numel() < extent:
            raise AflError(f"caller layout reaches index {extent - 1}, the buffer holds {caller.numel()}")
        return C.c_void_p(caller.data_ptr()), pp

    def pack_dofs(self, ncomp, caller, cell_perm, cell_stride, comp_stride, node_stride, out):
        """Caller layout -> this library's layout (out: ncomp * N)."""
        cp, pp = self._layout_args(ncomp, caller, cell_perm, cell_stride, comp_stride, node_stride)
        self._chk(self.L.afl_pack_dofs(self.h, ncomp, cp, pp, cell_stride, comp_stride, node_stride,
                                       _ptr(out, 8, ncomp * self.N)))
        return out

    def unpack_dofs(self, ncomp, src, cell_perm, cell_stride, comp_stride, node_stride, caller):
        """This library's layout (src: ncomp * N) -> caller layout."""
        cp, pp = self._layout_args(ncomp, caller, cell_perm, cell_stride, comp_stride, node_stride)
        self._chk(self.L.afl_unpack_dofs(self.h, ncomp, _ptr(src, 8, ncomp * self.N), pp, cell_stride, comp_stride,
                                         node_stride, cp))
        return caller

    def compute_diagonal(self): self._chk(self.L.afl_compute_diagonal(self.h))

    def get_matrix_diagonal_inverse(self, dst):
        self._chk(self.L.afl_get_matrix_diagonal_inverse(self.h, _ptr(dst, 8, self.m())))
        return dst

    # ---- solves --------------------------------------------------------------------------------
    def solve_mass(self, x, b): self._chk(self.L.afl_solve_mass(self.h, _ptr(x, 8, self.m()), _ptr(b, 8, self.m()))); return x

    def schur_solve(self, p, rhs):
        it = C.c_int(0)
        res = C.c_double(0.0)
        self._chk(self.L.afl_schur_solve(self.h, self._s(p), self._s(rhs), C.byref(it), C.byref(res)))
        return it.value, res.value

    def imex_step(self, rho, u, p):
        info = StepInfo()
        self._chk(self.L.afl_imex_step(self.h, self._s(rho), self._v(u), self._s(p), C.byref(info)))
        return info

    def init_hill_state(self, st: HillState):
        rho, u, p = self.new_scalar(), self.new_vector(), self.new_scalar()
        self._chk(self.L.afl_init_hill_state(self.h, C.byref(st), self._s(rho), self._v(u), self._s(p)))
        return rho, u, p

    def background(self, st: HillState):
        rho, p, th = self.new_scalar(), self.new_scalar(), self.new_scalar()
        self._chk(self.L.afl_background(self.h, C.byref(st), self._s(rho), self._s(p), self._s(th)))
        return rho, p, th
