/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference's EULEROperator (euler_operator.h) and of
 * the IMEX-ARK2 stage logic of EulerSolver (Atmospheric_Flow.cc:426-871, 1100-1257),
 * on the terrain-following hill mesh of afl_types.h. Only tests/, the smoke check and
 * the bench's cpu_baseline leg may load it, and only as the checker / CPU baseline.
 *
 * Vectors are host arrays in structure-of-arrays order: a scalar field is N doubles
 * (N = cells * (k+1)^dim, cell-major, nodes lexicographic x fastest); a vector field
 * is dim consecutive blocks of N doubles.
 */
#ifndef AFL_ORACLE_H
#define AFL_ORACLE_H

#include "../include/afl_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_operator orc_operator;

orc_operator* orc_create(const afl_mesh_desc* mesh, double mach, double froude, double dt);
void orc_destroy(orc_operator* op);
const char* orc_last_error(void);
/* Gauss points per direction of every volume and face integral: 0 restores 2k + 1 (the
 * reference's QGauss(2k+1)); k + 1 gives the level operator of the GPU multigrid
 * preconditioner. Resets the fixed fields' factorisations. */
int orc_set_quadrature(orc_operator* op, int nq);

int orc_set_dt(orc_operator* op, double dt);
int orc_set_mach(orc_operator* op, double mach);
int orc_set_froude(orc_operator* op, double froude);
int orc_set_hyperbolic_stage(orc_operator* op, unsigned stage);
int orc_set_ns_stage(orc_operator* op, unsigned stage);
int orc_set_rho_for_fixed(orc_operator* op, const double* src);
int orc_set_pres_fixed(orc_operator* op, const double* src);
int orc_set_u_fixed(orc_operator* op, const double* src);
int orc_set_solver_params(orc_operator* op, const afl_solver_params* sp);
int orc_set_sponge(orc_operator* op, const afl_sponge* sponge);
/* Planted faults for the oracle's own tests (0 = none; see g_fault in afl_oracle.cpp). */
int orc_set_fault(int fault);

int64_t orc_n_dofs(const orc_operator* op, int dof_index); /* 0 velocity, 1 pressure, 2 density */

int orc_vmult_rhs_rho_update(orc_operator* op, double* dst, const double* const* src, int n_src);
int orc_vmult_rhs_velocity_fixed(orc_operator* op, double* dst, const double* const* src, int n_src);
int orc_vmult_rhs_pressure(orc_operator* op, double* dst, const double* const* src, int n_src);

int orc_vmult(orc_operator* op, double* dst, const double* src);          /* NS_stage dispatch */
int orc_vmult_pressure(orc_operator* op, double* dst, const double* src); /* B */
int orc_vmult_enthalpy(orc_operator* op, double* dst, const double* src); /* C */
int orc_compute_diagonal(orc_operator* op);

/* Adjoint building blocks (no reference counterpart): transposes of the linearised forward
 * operators. sbar follows the E1 source order (rho_j: N, u_j: dim*N). */
int orc_vjp_rhs_rho_update(orc_operator* op, const double* rbar, double* const* sbar, const double* const* src, int n_src);
int orc_vjp_rhs_velocity_fixed(orc_operator* op, const double* rbar, double* const* sbar, const double* const* src,
                               int n_src); /* E2; sbar sizes follow the sources (rho, p: N, u: dim*N) */
int orc_vjp_rhs_pressure(orc_operator* op, const double* rbar, double* const* sbar, const double* const* src,
                         int n_src); /* E3 */
int orc_vmult_pressure_transpose(orc_operator* op, double* dst, const double* src);        /* B^T: dim*N -> N */
int orc_vmult_enthalpy_tangent(orc_operator* op, double* dst, const double* dw, const double* w0);  /* C'(w0) dw */
int orc_vmult_enthalpy_tangent_transpose(orc_operator* op, double* dst, const double* cb, const double* w0);
int orc_vmult_enthalpy_coefficient_transpose(orc_operator* op, double* dst, const double* cb, const double* w);
int orc_vjp_rho_for_fixed(orc_operator* op, double* dst, const double* a, const double* b); /* int phi a.b */
int orc_vmult_schur_tangent(orc_operator* op, double* dst, const double* src, const double* p0);
int orc_vmult_schur_tangent_transpose(orc_operator* op, double* dst, const double* src, const double* p0);
int orc_get_matrix_diagonal_inverse(orc_operator* op, double* dst);

int orc_solve_mass(orc_operator* op, double* x, const double* b); /* exact inverse of M / A / D */
int orc_schur_solve(orc_operator* op, double* p, const double* rhs, int* iters, double* res);
int orc_imex_step(orc_operator* op, double* rho, double* u, double* p, afl_step_info* info);

int orc_init_hill_state(orc_operator* op, const afl_hill_state* st, double* rho, double* u, double* p);
int orc_background(orc_operator* op, const afl_hill_state* st, double* rho_bar, double* p_bar, double* theta_bar);

/* Tables for pinning tests. */
int orc_diagnostics(orc_operator* op, const double* rho, const double* u, const double* p, double* out7);
int orc_l2_difference(orc_operator* op, int ncomp, const double* a, const double* b, double* err, double* ref);
int orc_basis_tables(const orc_operator* op, double* nodes, double* qp, double* qw, double* S, double* D);
int orc_node_coords(const orc_operator* op, double* X); /* dim blocks of N */
/* observation operator: points (dim coords each), theta' at the points (out may be NULL) and,
 * when wbar != NULL, the transpose of its linearisation into rho_bar, p_bar (overwritten) */
int orc_set_observation_points(orc_operator* op, int npts, const double* xyz);
int orc_observe_theta(orc_operator* op, const afl_hill_state* bg, const double* rho, const double* p, double* out,
                      const double* wbar, double* rho_bar, double* p_bar);

#ifdef __cplusplus
}
#endif

#endif
