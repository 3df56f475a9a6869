"""ctypes mirrors of the plain-C descriptor types in include/afl_types.h."""
import ctypes as C

AFL_NODES_GAUSS_LOBATTO = 0
AFL_NODES_EQUIDISTANT = 1


class MeshDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("degree", C.c_int32), ("node_type", C.c_int32),
        ("n_cells", C.c_int32 * 3),
        ("x_min", C.c_double), ("x_max", C.c_double),
        ("y_min", C.c_double), ("y_max", C.c_double),
        ("z_top", C.c_double), ("hill_height", C.c_double), ("hill_width", C.c_double),
        ("periodic", C.c_int32 * 2),
    ]


class SolverParams(C.Structure):
    _fields_ = [
        ("gmres_eps", C.c_double), ("gmres_max_iter", C.c_int32), ("gmres_restart", C.c_int32),
        ("fp_tol", C.c_double), ("fp_max_iter", C.c_int32), ("mg_quad_points", C.c_int32),
        ("schur_precond", C.c_int32), ("mg_smoother_degree", C.c_int32), ("mg_coarse_degree", C.c_int32),
        ("mg_levels", C.c_int32), ("mg_smoothing_range", C.c_double), ("mg_cell_tol", C.c_double),
    ]


AFL_SCHUR_JACOBI, AFL_SCHUR_MG_FP32, AFL_SCHUR_MG_FP64 = 0, 1, 2
AFL_TIME_LEVEL_W, AFL_TIME_LEVEL_C_CHEB = 1, 2
AFL_TIME_RHS_RHO, AFL_TIME_RHS_U, AFL_TIME_RHS_P, AFL_TIME_RHS_P_UFIX = 4, 8, 16, 32
AFL_TIME_SCHUR_W, AFL_TIME_APPLY_C, AFL_TIME_MASS_SOLVE, AFL_TIME_APPLY_B, AFL_TIME_A_SOLVE = 64, 128, 256, 512, 1024
TIMED_KERNELS = {AFL_TIME_LEVEL_W: "level_w", AFL_TIME_LEVEL_C_CHEB: "level_c_cheb", AFL_TIME_RHS_RHO: "rhs_rho",
                 AFL_TIME_RHS_U: "rhs_u", AFL_TIME_RHS_P: "rhs_p", AFL_TIME_RHS_P_UFIX: "rhs_p_ufix",
                 AFL_TIME_SCHUR_W: "schur_w", AFL_TIME_APPLY_C: "apply_C", AFL_TIME_MASS_SOLVE: "mass_solve",
                 AFL_TIME_APPLY_B: "apply_B", AFL_TIME_A_SOLVE: "A_solve"}


class StepInfo(C.Structure):
    _fields_ = [
        ("fp_iters", C.c_int32 * 2), ("gmres_iters", C.c_int32 * 2),
        ("fp_converged", C.c_int32 * 2), ("pad_", C.c_int32 * 2),
        ("schur_residual", C.c_double),
    ]


class AdjointStats(C.Structure):
    _fields_ = [("forward_steps", C.c_int32), ("adjoint_steps", C.c_int32), ("snapshots_used", C.c_int32),
                ("recomputed_steps", C.c_int32), ("stage_sets_used", C.c_int32),
                ("stage_solve_iterations", C.c_int32)]


class Diagnostics(C.Structure):
    _fields_ = [
        ("min_density", C.c_double), ("max_density", C.c_double), ("max_celerity", C.c_double),
        ("max_velocity", C.c_double), ("min_cell_diameter", C.c_double), ("cfl_c", C.c_double),
        ("cfl_u", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class HillState(C.Structure):
    _fields_ = [
        ("theta0", C.c_double), ("brunt", C.c_double), ("p0", C.c_double), ("u0", C.c_double),
        ("theta_sigma", C.c_double), ("seed", C.c_uint64),
    ]


class Sponge(C.Structure):
    _fields_ = [
        ("z_start", C.c_double), ("alpha_max", C.c_double), ("u_ref", C.c_double * 3),
        ("enabled", C.c_int32), ("pad_", C.c_int32),
    ]


def make_mesh(dim=2, degree=2, n_cells=(8, 1, 4), x_range=(-72e3, 72e3), y_range=(-72e3, 72e3),
              z_top=25e3, hill_height=400.0, hill_width=1000.0, periodic=(1, 1),
              node_type=AFL_NODES_GAUSS_LOBATTO):
    m = MeshDesc()
    m.dim, m.degree, m.node_type = dim, degree, node_type
    for i in range(3):
        m.n_cells[i] = int(n_cells[i])
    m.x_min, m.x_max = x_range
    m.y_min, m.y_max = y_range
    m.z_top, m.hill_height, m.hill_width = z_top, hill_height, hill_width
    m.periodic[0], m.periodic[1] = periodic
    return m


def make_hill_state(theta0=280.0, brunt=0.01, p0=1e5, u0=10.0, theta_sigma=0.01, seed=1234):
    s = HillState()
    s.theta0, s.brunt, s.p0, s.u0, s.theta_sigma, s.seed = theta0, brunt, p0, u0, theta_sigma, seed
    return s


def make_solver_params(gmres_eps=1e-10, gmres_max_iter=30000, gmres_restart=28, fp_tol=1e-10, fp_max_iter=100,
                       schur_precond=0, mg_smoother_degree=3, mg_coarse_degree=0, mg_levels=1,
                       mg_smoothing_range=15.0, mg_cell_tol=0.05, mg_quad_points=0):
    p = SolverParams()
    p.gmres_eps, p.gmres_max_iter, p.gmres_restart = gmres_eps, gmres_max_iter, gmres_restart
    p.fp_tol, p.fp_max_iter = fp_tol, fp_max_iter
    p.schur_precond, p.mg_smoother_degree, p.mg_coarse_degree = schur_precond, mg_smoother_degree, mg_coarse_degree
    p.mg_levels, p.mg_smoothing_range, p.mg_cell_tol = mg_levels, mg_smoothing_range, mg_cell_tol
    p.mg_quad_points = mg_quad_points
    return p


def make_sponge(z_start=15e3, alpha_max=0.0, u_ref=(10.0, 0.0, 0.0), enabled=False):
    s = Sponge()
    s.z_start, s.alpha_max = z_start, alpha_max
    for i in range(3):
        s.u_ref[i] = u_ref[i]
    s.enabled = 1 if enabled else 0
    return s


def n_per_field(mesh):
    k1 = mesh.degree + 1
    if mesh.dim == 2:
        return mesh.n_cells[0] * mesh.n_cells[2] * k1 * k1
    return mesh.n_cells[0] * mesh.n_cells[1] * mesh.n_cells[2] * k1 ** 3
