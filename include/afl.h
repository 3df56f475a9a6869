/*
 * C-ABI of libafl_hip.so: the MI355X (gfx950) drop-in for the matrix-free cell and face
 * loops of the reference's EULEROperator and for the solves around them.
 *
 * Reference interface (letiziabottani/Atmospheric_Flows):
 *   class EULEROperator : MatrixFreeOperators::Base<dim, Vec>    euler_operator.h:19-98
 *   consumed as a matrix by SolverCG / SolverGMRES / PreconditionJacobi / mg::Matrix
 *   (Atmospheric_Flow.cc:518, 527, 642, 672).
 *
 * Conventions
 *  - Every vector argument is a DEVICE pointer (hipMalloc'ed or any device allocation).
 *    Scalar fields hold N = cells*(k+1)^dim doubles, cell-major, nodes lexicographic with
 *    x fastest; vector fields (velocity) hold dim consecutive blocks of N doubles.
 *    afl_pack_dofs / afl_unpack_dofs convert from/to a caller layout with explicit strides
 *    and a cell permutation (e.g. deal.II's FESystem DoF order).
 *  - Calls are ordered on the handle's stream (afl_set_stream; default: the null stream).
 *    Solver entry points synchronise that stream where they need a scalar on the host.
 *  - A handle is not reentrant (the reference's NS_stage is mutable, euler_operator.h:91):
 *    use one handle per host thread.
 *  - Errors never cross the ABI as exceptions: every call returns an afl_status and
 *    afl_last_error() returns a thread-local message.
 *  - Setters copy their vector into device memory owned by the handle (the reference
 *    deep-copies, euler_operator.h:312-339); the caller keeps ownership of its buffers.
 */
#ifndef AFL_H
#define AFL_H

#include <stdint.h>

#include "afl_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum afl_status {
  AFL_OK = 0,
  AFL_ERR_INVALID_ARGUMENT = 1,
  AFL_ERR_NOT_CONVERGED = 2, /* SolverControl::NoConvergence in the reference */
  AFL_ERR_DEVICE = 3,
  AFL_ERR_STATE = 4          /* call not valid for the current NS_stage / setup */
} afl_status;

typedef struct afl_operator afl_operator;

/* ---- lifetime ------------------------------------------------------------------------ */
/* Replaces EULEROperator(data) (euler_operator.h:228-240) + MatrixFree::reinit
 * (Atmospheric_Flow.cc:325-348): mesh/basis descriptor, Ma, Fr, dt. */
int afl_create(const afl_mesh_desc* mesh, double mach, double froude, double dt, afl_operator** out);
int afl_destroy(afl_operator* op);
const char* afl_last_error(void);
int afl_set_stream(afl_operator* op, void* hip_stream);
/* Number of unknowns of DoFHandler index 0 (velocity), 1 (pressure), 2 (density),
 * as MatrixFree::initialize_dof_vector (Atmospheric_Flow.cc:351-375). */
int afl_n_dofs(const afl_operator* op, int dof_index, int64_t* n);
/* Base::m(): rows of the operator selected by NS_stage. */
int afl_m(const afl_operator* op, int64_t* m);

/* ---- setters, euler_operator.h:34-53 / 243-339 ------------------------------------- */
int afl_set_dt(afl_operator* op, double dt);
int afl_set_mach(afl_operator* op, double mach);
int afl_set_froude(afl_operator* op, double froude);
int afl_set_hyperbolic_stage(afl_operator* op, unsigned stage); /* 1..3, :283-288 */
int afl_set_ns_stage(afl_operator* op, unsigned stage);         /* 1..6, :296-301 */
int afl_set_rho_for_fixed(afl_operator* op, const double* rho); /* :310-315 */
int afl_set_pres_fixed(afl_operator* op, const double* p);      /* :321-327 */
int afl_set_u_fixed(afl_operator* op, const double* u);         /* :333-339 */
int afl_set_solver_params(afl_operator* op, const afl_solver_params* sp);
int afl_set_sponge(afl_operator* op, const afl_sponge* sponge);

/* ---- explicit-stage applies (rows E1-E3), euler_operator.h:55-62 ---------------------
 * src holds device pointers in the reference's positional order
 * (Atmospheric_Flow.cc:438-450, 555-568, 696-698, 796-799); n_src is checked per stage. */
int afl_vmult_rhs_rho_update(afl_operator* op, double* dst, const double* const* src, int n_src);     /* :697-712 */
int afl_vmult_rhs_velocity_fixed(afl_operator* op, double* dst, const double* const* src, int n_src); /* :2310-2325 */
int afl_vmult_rhs_pressure(afl_operator* op, double* dst, const double* const* src, int n_src);       /* :1366-1381 */

/* ---- linear blocks ------------------------------------------------------------------- */
/* Base::vmult -> apply_add (euler_operator.h:2370-2416): NS 1/4 M, 2 Schur S = D - C A^-1 B,
 * 3/5 A (rho_for_fixed), 6 D. The A^-1 inside S is an exact cell-local solve. */
int afl_vmult(afl_operator* op, double* dst, const double* src);
int afl_vmult_pressure(afl_operator* op, double* dst, const double* src); /* B, :2420-2435 */
int afl_vmult_enthalpy(afl_operator* op, double* dst, const double* src); /* C, :2438-2453 (uses pres_fixed) */
/* First half of the Schur apply, w = A^{-1} B p (euler_operator.h:2380-2397: vmult_pressure
 * followed by the inner CG on A), one fused kernel with an exact cell-local A solve. */
int afl_vmult_schur_velocity(afl_operator* op, double* w, const double* p);
/* Preconditioner apply z = P^{-1} v of the Schur solve: the p-multigrid V-cycle with Chebyshev
 * smoothers (the reference wires deal.II PreconditionMG + PreconditionChebyshev into its
 * mass-type solves, Atmospheric_Flow.cc:455-527; here the V-cycle smooths the Schur operator of
 * the current stage coefficients). Requires solver_params.schur_precond = AFL_SCHUR_MG_FP32 or
 * AFL_SCHUR_MG_FP64; FP64 vectors in and out, level arithmetic in the selected precision. */
int afl_mg_vmult(afl_operator* op, double* dst, const double* src);
/* The operator that V-cycle smooths on its finest level, dst = S_0 src: the Schur operator of
 * the current stage coefficients rediscretised in the level precision on mg_quad_points Gauss
 * points per direction, with the level cell solves (mean-density block when mg_cell_tol >= 0.05).
 * No counterpart in the reference (whose preconditioners are deal.II's); it exposes the
 * preconditioner's operator for testing. Requires schur_precond = AFL_SCHUR_MG_FP32 or _FP64. */
int afl_mg_level_vmult(afl_operator* op, double* dst, const double* src);
/* The FP32 fine-level half of that V-cycle's operator, w = A^{-1} B p in single precision on the
 * FP32 copies of rho_for_fixed (requires schur_precond = AFL_SCHUR_MG_FP32). */
int afl_vmult_schur_velocity_f32(afl_operator* op, float* w, const float* p);

/* ---- discrete-adjoint building blocks ------------------------------------------------------
 * The reference has no adjoint; these are the "adjoint entry points: transpose applies" that a
 * gradient of a final-time misfit with respect to the initial state is assembled from. Each is
 * the exact transpose (or tangent) of the linearisation of the forward operator named. */
/* E1 vector-Jacobian product (vmult_rhs_rho_update, euler_operator.h:342-712):
 * src_bar[i] = (d rhs_rho / d src[i])^T rbar for the E1 sources of the current HYPERBOLIC_stage
 * in their positional order (rho_j: N, u_j: dim*N). Outputs are overwritten. */
int afl_vjp_rhs_rho_update(afl_operator* op, const double* rbar, double* const* src_bar, const double* const* src,
                           int n_src);
/* E2 vector-Jacobian product (vmult_rhs_velocity_fixed, euler_operator.h:1620-2325): rbar has dim*N
 * entries; src_bar[i] matches src[i] (rho, p: N; u: dim*N), E2 positional order of the stage. */
int afl_vjp_rhs_velocity_fixed(afl_operator* op, const double* rbar, double* const* src_bar, const double* const* src,
                               int n_src);
/* E3 vector-Jacobian product (vmult_rhs_pressure, euler_operator.h:743-1381): rbar has N entries;
 * src_bar[i] matches src[i] in the E3 positional order of the stage. Sources that alias one vector
 * (u_old read twice at stage 2, rho_new / u_fix read by the correction and by gravity) get the
 * adjoint of every use at their position. */
int afl_vjp_rhs_pressure(afl_operator* op, const double* rbar, double* const* src_bar, const double* const* src,
                         int n_src);
/* B^T (vmult_pressure, :2420-2435): dim*N -> N. */
int afl_vmult_pressure_transpose(afl_operator* op, double* dst, const double* src);
/* Tangent of vmult_enthalpy (:2438-2453) at w0 with the current pres_fixed: C'(w0) dw, N <- dim*N.
 * C is nonlinear in w through the Rusanov speed max(|w+.n|, |w-.n|); at ties the subgradient is
 * split evenly between the two traces. */
int afl_vmult_enthalpy_tangent(afl_operator* op, double* dst, const double* dw, const double* w0);
/* C'(w0)^T cbar: N -> dim*N. */
int afl_vmult_enthalpy_tangent_transpose(afl_operator* op, double* dst, const double* cbar, const double* w0);
/* Coefficient transposes for the implicit stages: (d C(w; pres_fixed) / d pres_fixed)^T cbar (C is
 * linear in its coefficient, :1411-1507), N -> N; (d A(rho) u / d rho)^T ybar = int phi ybar . u
 * (assemble_cell_term_velocity_fixed, :2328-2357), dim*N -> N; and the adjoint of E3 with respect to
 * u_fixed at stages 1 and 2, whose kinetic-energy correction and gravity work read (rho_new, u_fixed)
 * (:743-1007), N -> dim*N. */
int afl_vmult_enthalpy_coefficient_transpose(afl_operator* op, double* dst, const double* cbar, const double* w);
int afl_vjp_rho_for_fixed(afl_operator* op, double* dst, const double* ybar, const double* u);
int afl_vjp_rhs_pressure_u_fixed(afl_operator* op, double* dst, const double* rbar, const double* rho_new,
                                 const double* u_fixed);
/* Tangent of the Schur apply (apply_add NS 2, :2374-2408) at p0: S'(p0) x = D x - C'(w0) A^{-1} B x
 * with w0 = A^{-1} B p0, and its transpose D x - B^T A^{-1} C'(w0)^T x (rho_for_fixed, pres_fixed). */
int afl_vmult_schur_tangent(afl_operator* op, double* dst, const double* src, const double* p0);
int afl_vmult_schur_tangent_transpose(afl_operator* op, double* dst, const double* src, const double* p0);

/* Adjoint of one IMEX step (afl_imex_step, i.e. EulerSolver::run's step, Atmospheric_Flow.cc:1110-1251):
 * given the step's input state (rho, u, p) and the adjoint of its output (rho_bar_out, ...), writes
 * the adjoint of its input (rho_bar, ...): the transpose of the step's Jacobian applied to the output
 * adjoint. Recomputes the forward step internally (info: that step's solver counts), then runs the
 * stages backwards; the implicit stages are differentiated at their converged fixed point and need
 * one GMRES solve each (gmres_eps, gmres_restart of the solver parameters; Jacobi(D) preconditioner).
 * Uses rho_for_fixed, pres_fixed and the stage as scratch, like afl_imex_step. */
int afl_imex_step_vjp(afl_operator* op, const double* rho, const double* u, const double* p, const double* rho_bar_out,
                      const double* u_bar_out, const double* p_bar_out, double* rho_bar, double* u_bar, double* p_bar,
                      afl_step_info* info);

/* Gradient of a final-time functional with respect to the initial state over nsteps IMEX steps:
 * (rho_bar_0, ...) = (d x_T / d x_0)^T (rho_bar_T, ...), x_T = nsteps steps of afl_imex_step from
 * (rho0, u0, p0). States are kept in at most `snapshots` device snapshots besides the initial one
 * (each 2+dim fields) and recomputed from them with binomial checkpointing (Griewank), with the
 * split points that minimise the forward steps for that memory. Each advance to a split point also
 * keeps the stage vectors of its last step (a stage set: 5 + 3 dim fields; at most snapshots + 1
 * held), so that step is reversed without recomputing it; the other reverse steps recompute their
 * own step to restore the stage vectors (afl_set_adjoint_stage_sets(op, 0): every reverse step
 * does, and no stage set is held). snapshots = AFL_ADJOINT_STORE_ALL instead keeps every step's
 * state and stage vectors from one forward sweep (no recomputation; fails with AFL_ERR_DEVICE when
 * they do not fit). stats may be NULL. No reference counterpart (the reference has no adjoint). */
int afl_set_adjoint_stage_sets(afl_operator* op, int keep);
int afl_adjoint_run(afl_operator* op, const double* rho0, const double* u0, const double* p0, int nsteps,
                    int snapshots, const double* rho_bar_T, const double* u_bar_T, const double* p_bar_T,
                    double* rho_bar_0, double* u_bar_0, double* p_bar_0, afl_adjoint_stats* stats);

/* Observation operator (no reference counterpart; the misfit of the gradient configurations):
 * afl_set_observation_points copies npts points (host array, dim coordinates each: x[, y], z) and
 * locates them on the mapped mesh (error if one lies outside). afl_observe_theta writes, per point,
 * theta' = theta(rho, p) - theta0 exp(N^2 z / g) with theta = p / (rho R (p / p0)^(R/cp)), rho and p
 * evaluated at the point with the cell's Lagrange basis, background from bg (theta0, brunt, p0).
 * afl_observe_theta_transpose overwrites (rho_bar, p_bar) with the transpose of its linearisation
 * applied to wbar (device, npts). Deterministic: points are grouped by cell on the host. */
int afl_set_observation_points(afl_operator* op, int npts, const double* xyz);
int afl_observe_theta(afl_operator* op, const afl_hill_state* bg, const double* rho, const double* p,
                      double* theta_prime);
int afl_observe_theta_transpose(afl_operator* op, const afl_hill_state* bg, const double* rho, const double* p,
                                const double* wbar, double* rho_bar, double* p_bar);

/* Benchmark instrumentation, no counterpart in the reference (deal.II's TimerOutput times whole
 * phases, not kernels): HIP events recorded on the operator's stream around every launch of the
 * kernels in `kernels` (an or-ed mask of enum afl_timed_kernel) until afl_kernel_timer_stop, at most
 * max_launches launches in all. Stop synchronises the stream and returns the launch count and summed
 * duration over every timed kernel; afl_kernel_timer_result then gives one kernel's launches, summed
 * duration and summed algorithmic bytes (each distinct input and output field once), restricted to
 * one IMEX stage (1..3; 0: every stage). */
/* Progress lines on stderr (the reference's `verbose` argument of EulerSolver::run and its verbose_cout,
 * Atmospheric_Flow.cc:1101-1102): level >= 1 reports every reversed step of
 * afl_adjoint_run. 0 (the default) prints nothing. */
int afl_set_verbose(afl_operator* op, int level);
/* Phase timing of afl_imex_step, the sections of the reference's TimerOutput time_table
 * (Atmospheric_Flow.cc:246-249, 432, 546, 682, 790; enum afl_timed_section): HIP events on the
 * operator's stream around every phase of every step between start and stop (start clears earlier
 * records). afl_section_timer_result synchronises the stream and returns a section's number of timed
 * phases and their summed duration. */
int afl_section_timer_start(afl_operator* op);
int afl_section_timer_stop(afl_operator* op);
int afl_section_timer_result(afl_operator* op, int section, int* calls, double* total_ms);
int afl_kernel_timer_start(afl_operator* op, int kernels, int max_launches);
int afl_kernel_timer_stop(afl_operator* op, int* launches, double* total_ms);
int afl_kernel_timer_result(afl_operator* op, int kernel, int stage, int* launches, double* total_ms,
                            double* bytes);

/* ---- diagnostics ------------------------------------------------------------------------ */
/* VTU output of a state (EulerSolver::output_results, Atmospheric_Flow.cc:886-918): point data
 * "rho", "u" (3 components) and "p" on k+1 equidistant patch points per direction of every cell
 * (DataOut::build_patches with n_subdivisions = degree, mapped with the terrain-following map),
 * written as k^dim linear sub-cells (high_order = 0) or one VTK Lagrange quadrilateral (2D) or
 * hexahedron (3D, VTK file version 2.2 point order) per cell (high_order = 1, the reference's second
 * file, :909-913). Binary (base64) XML, UInt64 headers.
 * Copies the fields to the host; synchronises the stream. */
int afl_write_vtu(afl_operator* op, const char* path, const double* rho, const double* u, const double* p,
                  int high_order);

/* Density bounds, maximum celerity and speed of a state and the two Courant numbers
 * (Atmospheric_Flow.cc:992-1087 get_max_velocity / get_min_density / get_max_density /
 * compute_max_celerity; CFL_c and CFL_u as printed at :1225-1235). The reference evaluates
 * these with QGaussLobatto(k+1) on Gauss-Lobatto support points, i.e. at the nodes, which is
 * what this computes. u and p may be NULL (their entries are then 0). Synchronises the stream. */
int afl_compute_diagnostics(afl_operator* op, const double* rho, const double* u, const double* p,
                            afl_diagnostics* out);
/* L2 norms ||a - b|| and ||b|| of ncomp-component fields (1 or dim), the quantities of
 * VectorTools::integrate_difference + compute_global_error in analyze_results (:925-985),
 * with b a discrete reference field (b == NULL: b = 0). Gauss quadrature with 2k+1 points on
 * the mapped cells. Synchronises the stream. */
int afl_l2_difference(afl_operator* op, int ncomp, const double* a, const double* b, double* err_l2,
                      double* ref_l2);
/* afl_l2_difference without any device allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments and results; the partial sums live in a fixed 32.8 KB device array
 * of the library (one copy per GPU, its address resolved at afl_create or on the first call on a
 * device), so no size limit applies beyond afl_create's. Calls that overlap on one GPU share that
 * array and race. Synchronises the stream. */
int afl_l2_difference_noalloc(afl_operator* op, int ncomp, const double* a, const double* b, double* err_l2,
                              double* ref_l2);
/* compute_diagonal / get_matrix_diagonal_inverse (euler_operator.h:2589-2626) */
int afl_compute_diagonal(afl_operator* op);
int afl_get_matrix_diagonal_inverse(afl_operator* op, double* dst);

/* ---- solves -------------------------------------------------------------------------- */
/* Exact inverse of the block-diagonal mass-type operator selected by NS_stage (1/4 M,
 * 3/5 A, 6 D). Replaces SolverCG + PreconditionMG on those systems
 * (Atmospheric_Flow.cc:452-535, 571-642, 701-779, 801-870). */
int afl_solve_mass(afl_operator* op, double* x, const double* b);
/* Restarted GMRES on the Schur system S p = rhs with Jacobi(D) (Atmospheric_Flow.cc:652-672),
 * right-preconditioned so that the stopping test ||rhs - S p|| <= gmres_eps * ||rhs|| is on the
 * true residual. p holds the initial guess on entry; restart and iteration cap from
 * afl_set_solver_params. */
int afl_schur_solve(afl_operator* op, double* p, const double* rhs, int* iters, double* residual);
/* One full IMEX-ARK2 step (Atmospheric_Flow.cc:1110-1251): density solves, fixed-point
 * Schur loops of stages 1-2, explicit stage 3, then the sponge. Updates rho, u, p in place. */
int afl_imex_step(afl_operator* op, double* rho, double* u, double* p, afl_step_info* info);

/* ---- problem set-up helpers ---------------------------------------------------------- */
/* Hydrostatic constant-N background + seeded nodal theta noise (replaces
 * EulerSolver::initialize, Atmospheric_Flow.cc:417-423, for the hill problems). */
int afl_init_hill_state(afl_operator* op, const afl_hill_state* st, double* rho, double* u, double* p);
/* Noise-free background (rho_bar, p_bar, theta_bar); any output may be NULL. */
int afl_background(afl_operator* op, const afl_hill_state* st, double* rho_bar, double* p_bar, double* theta_bar);

/* The p-multigrid transfer between degrees of the preconditioner's hierarchy, exposed so that callers and tests
 * can check it (no counterpart in the reference, whose multigrid is geometric: MGTransferMatrixFree,
 * Atmospheric_Flow.cc:478-486). One component on this operator's mesh and node family:
 * transpose = 0: dst (degree k_fine) = P src (degree k_coarse), the nodal interpolation of the coarse
 * polynomial (the V-cycle's prolongation); transpose = 1: dst (degree k_coarse) = P^T src (degree k_fine),
 * its restriction. 1 <= k_coarse < k_fine <= 4. */
int afl_mg_transfer(afl_operator* op, int k_coarse, int k_fine, int transpose, const double* src, double* dst);

/* ---- layout conversion ---------------------------------------------------------------
 * dst[c*NPC*ncomp + ...] style caller layouts: caller index of (cell c, component a,
 * node i) = cell_perm[c]*cell_stride + a*comp_stride + i*node_stride (cell_perm may be
 * NULL for the identity; it is a device array of int64). */
int afl_pack_dofs(afl_operator* op, int ncomp, const double* caller, const int64_t* cell_perm, int64_t cell_stride,
                  int64_t comp_stride, int64_t node_stride, double* afl_layout);
int afl_unpack_dofs(afl_operator* op, int ncomp, const double* afl_layout, const int64_t* cell_perm,
                    int64_t cell_stride, int64_t comp_stride, int64_t node_stride, double* caller);

#ifdef __cplusplus
}
#endif

#endif /* AFL_H */
