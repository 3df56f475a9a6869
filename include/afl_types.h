/*
 * Plain-C descriptor types shared by the C-ABI (afl.h) and its callers.
 *
 * The mesh descriptor replaces what the reference builds through deal.II's
 * Triangulation + MappingQ + MatrixFree::reinit (Atmospheric_Flow.cc:280-348):
 * a structured, terrain-following mesh over a Witch-of-Agnesi hill, with the
 * exact analytic mapping
 *     x = x_min + (ix + xi) hx,   y = y_min + (iy + eta) hy,
 *     zeta = (iz + tau) hz,       z = zeta + h(x, y) (1 - zeta / z_top),
 *     h(x, y) = h0 / (1 + (x^2 + y^2) / a^2)          (y absent in 2D).
 * Cells are numbered x fastest, then y, then z; nodes inside a cell are
 * lexicographic (x fastest) on the 1D support points of FE_DGQ(k).
 */
#ifndef AFL_TYPES_H
#define AFL_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1D support points of the nodal DG basis (FE_DGQ(k) in the reference). */
enum afl_node_type {
  AFL_NODES_GAUSS_LOBATTO = 0, /* FE_DGQArbitraryNodes(QGaussLobatto(k+1)) */
  AFL_NODES_EQUIDISTANT = 1    /* equispaced support points incl. end points */
};

typedef struct afl_mesh_desc {
  int32_t dim;          /* 2 or 3 */
  int32_t degree;       /* polynomial degree k, 1..4; quadrature has 2k+1 Gauss points per direction */
  int32_t node_type;    /* enum afl_node_type */
  int32_t n_cells[3];   /* cells along x, y, z (n_cells[1] ignored in 2D) */
  double x_min, x_max;  /* horizontal extent [m] */
  double y_min, y_max;  /* second horizontal extent [m], 3D only */
  double z_top;         /* height H of the model top [m] */
  double hill_height;   /* h0 [m] */
  double hill_width;    /* a [m] */
  int32_t periodic[2];  /* 1: periodic in x (resp. y); 0: wall. Bottom and top are walls. */
} afl_mesh_desc;

/* Solver controls of one IMEX step (reference values in brackets). */
typedef struct afl_solver_params {
  double gmres_eps;        /* relative tolerance of the Schur GMRES [prm eps = 1e-8] */
  int32_t gmres_max_iter;  /* [prm max_iterations = 30000] */
  int32_t gmres_restart;   /* Krylov basis size before restart [deal.II default 28] */
  double fp_tol;           /* fixed-point tolerance on the relative inf-norm change of p [1e-10] */
  int32_t fp_max_iter;     /* [100] */
  int32_t mg_quad_points;  /* Gauss points per direction of the multigrid level operators: 0 or k + 1
                            * [default; the level operator of the preconditioner only], or 2k + 1 (the
                            * operator's own rule). Results are unaffected: the outer FGMRES applies
                            * the exact Schur operator. */
  /* Schur preconditioner (zero-initialised fields select the reference's plain path):
   * AFL_SCHUR_JACOBI: GMRES + point Jacobi on D, all FP64;
   * AFL_SCHUR_MG_FP32 / AFL_SCHUR_MG_FP64: flexible GMRES (FP64) + p-multigrid V-cycle with
   * Chebyshev smoothers run in FP32 / FP64. Non-positive MG fields take the defaults below. */
  int32_t schur_precond;
  int32_t mg_smoother_degree;  /* Chebyshev degree of pre/post smoothing [3] */
  int32_t mg_coarse_degree;    /* Chebyshev degree on the coarsest level [2D: 16 at k = 4, 8 at k <= 3;
                                * 3D: 12 at k = 4, 5 at k <= 3] */
  int32_t mg_levels;           /* maximum number of degree levels k, k/2, ..., 1 [1]; with one level the
                                * V-cycle is the Chebyshev polynomial preconditioner on the fine level */
  double mg_smoothing_range;   /* lambda_max / lambda_min of the smoothed interval [15] */
  double mg_cell_tol;          /* cell-local A-solve tolerance inside the levels [0.05] */
} afl_solver_params;

enum afl_schur_precond { AFL_SCHUR_JACOBI = 0, AFL_SCHUR_MG_FP32 = 1, AFL_SCHUR_MG_FP64 = 2 };

/* Per-step diagnostics (the reference prints these through pcout). */
typedef struct afl_step_info {
  int32_t fp_iters[2];      /* fixed-point iterations in stages 1 and 2 */
  int32_t gmres_iters[2];   /* total Schur GMRES iterations in stages 1 and 2 */
  int32_t fp_converged[2];  /* 1 if the fixed point met fp_tol */
  int32_t pad_[2];
  double schur_residual;    /* last GMRES residual estimate */
} afl_step_info;

/* afl_adjoint_run's snapshots argument for keeping every step's state and stage vectors
 * (15 scalar fields per step in 2D, 20 in 3D): no recomputation at all. */
#define AFL_ADJOINT_STORE_ALL (-1)

/* Counters of an adjoint run with binomial checkpointing (afl_adjoint_run). */
typedef struct afl_adjoint_stats {
  int32_t forward_steps;     /* IMEX steps taken to reach checkpoints and re-reach states */
  int32_t adjoint_steps;     /* steps reversed */
  int32_t snapshots_used;    /* most snapshots held at once (including the initial state) */
  int32_t recomputed_steps;  /* reversed steps that recomputed their forward step (no stage set held) */
  int32_t stage_sets_used;   /* most stage sets held at once (8 fields each: 5 scalar + 3 velocity) */
  int32_t stage_solve_iterations; /* Krylov iterations of all adjoint stage solves (two per reversed step) */
} afl_adjoint_stats;

/* Hydrostatic constant-N background plus seeded theta noise. */
typedef struct afl_hill_state {
  double theta0;      /* surface potential temperature [K] (280) */
  double brunt;       /* Brunt-Vaisala frequency N [1/s] (0.01) */
  double p0;          /* surface pressure [Pa] (1e5) */
  double u0;          /* uniform inflow [m/s] (10) */
  double theta_sigma; /* std-dev of the nodal theta perturbation [K] (0.01) */
  uint64_t seed;      /* noise seed */
} afl_hill_state;

/* Rayleigh sponge applied to the velocity after every step. */
typedef struct afl_sponge {
  double z_start;     /* bottom of the sponge [m] (15 km) */
  double alpha_max;   /* relaxation rate at the top [1/s] */
  double u_ref[3];    /* velocity the sponge relaxes to */
  int32_t enabled;
  int32_t pad_;
} afl_sponge;

/* Kernels that afl_kernel_timer_start can time (benchmark instrumentation); flags, or-ed into a mask. */
enum afl_timed_kernel {
  AFL_TIME_LEVEL_W = 1,        /* multigrid level operator, first half: w = P^-1 B p */
  AFL_TIME_LEVEL_C_CHEB = 2,   /* level C/D apply fused with one Chebyshev smoothing step */
  AFL_TIME_RHS_RHO = 4,        /* E1, vmult_rhs_rho_update */
  AFL_TIME_RHS_U = 8,          /* E2, vmult_rhs_velocity_fixed */
  AFL_TIME_RHS_P = 16,         /* E3, vmult_rhs_pressure (whole, or its u_fixed-free part) */
  AFL_TIME_RHS_P_UFIX = 32,    /* E3's u_fixed terms inside the fixed point */
  AFL_TIME_SCHUR_W = 64,       /* FP64 Schur first half w = A^-1 B p */
  AFL_TIME_APPLY_C = 128,      /* FP64 C apply (with D p: the Schur second half) */
  AFL_TIME_MASS_SOLVE = 256,   /* FP64 cell-local M and D solves (one component) */
  AFL_TIME_APPLY_B = 512,      /* FP64 B apply */
  AFL_TIME_A_SOLVE = 1024      /* FP64 cell-local rho-weighted A solves (dim components) */
};

/* Phases of afl_imex_step timed by afl_section_timer_*: the reference's TimerOutput sections
 * (Atmospheric_Flow.cc:432, 546, 682, 790). */
enum afl_timed_section {
  AFL_SECTION_UPDATE_DENSITY = 0,         /* E1 and M^-1, every stage */
  AFL_SECTION_FIXED_POINT_PRESSURE = 1,   /* the fixed-point loop of an implicit stage (its velocity updates included) */
  AFL_SECTION_UPDATE_VELOCITY = 2,        /* stage 3: E2 and A^-1 */
  AFL_SECTION_UPDATE_PRESSURE = 3,        /* stage 3: E3 and D^-1 */
  AFL_N_SECTIONS = 4
};

/* Diagnostics of a state (EulerSolver::get_min_density / get_max_density /
 * compute_max_celerity / get_max_velocity and the Courant numbers printed in run,
 * Atmospheric_Flow.cc:992-1087, 1225-1235). */
typedef struct afl_diagnostics {
  double min_density;        /* min nodal rho */
  double max_density;        /* max |rho| (linfty_norm) */
  double max_celerity;       /* max sqrt(gamma p / rho); 0 if p was not given */
  double max_velocity;       /* max sqrt(u_0^2 + u_1^2) (the reference's formula, also in 3D); 0 if u not given */
  double min_cell_diameter;  /* minimal_cell_diameter: min over cells of the longest vertex diagonal */
  double cfl_c;              /* dt max_celerity k sqrt(dim) / (Ma min_cell_diameter) */
  double cfl_u;              /* dt max_velocity k sqrt(dim) / min_cell_diameter */
} afl_diagnostics;

#ifdef __cplusplus
}
#endif

#endif /* AFL_TYPES_H */
