// p-multigrid preconditioner for the Schur system (see afl_mg.hip).
#pragma once

#include <vector>

#include "afl_internal.h"

namespace afl {

struct MgConfig {
  int smoother_degree = 3;      // Chebyshev degree on every level but the coarsest
  int coarse_degree = 16;       // Chebyshev degree on the coarsest level
  double smoothing_range = 15;  // lambda_max / lambda_min of the smoothed interval
  int max_levels = 1;
  double cell_tol = 0.05;       // cell-local A solve tolerance on the levels (preconditioner accuracy)
  int cell_maxit = 40;
  int eig_iters = 12;           // power iterations for lambda_max
  bool reduced_quadrature = true;  // level operators on k + 1 Gauss points per direction (else 2k + 1)
};

template <typename R>
struct MgLevelT {
  int dim = 2, k = 1;
  int64_t ncells = 0, N = 0;
  Geo geo{};
  TabT<R> T{};
  double MI[5][5] = {};
  double E[5][5] = {};   // nodal interpolation fine level -> this level
  double P[5][5] = {};   // prolongation this (coarse) level -> next finer level
  double Pt[5][5] = {};  // restriction (transpose of P)
  double lam_max = 1.0;
  double cell_tol = 0.05;
  int cell_maxit = 40;
  bool rq = true;  // quadrature of the level operator: k + 1 Gauss points per direction (else 2k + 1)
  hipStream_t stream = nullptr;
  R *rho = nullptr, *pfix = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *r2 = nullptr, *z = nullptr,
    *d = nullptr, *t = nullptr, *w = nullptr, *jbar = nullptr, *ev = nullptr, *sinv = nullptr;
  // preconditioner-accuracy cell solves (tolerance >= 0.05) use the mean-density block P = (sum
  // rho JxW) (M1 x M1 [x M1]) with sinv = 1 / sum rho JxW per cell instead of iterating on A
  bool lean() const { return cell_tol >= 0.05; }
  bool ev_valid = false;
  TabT<R>* dT = nullptr;  // device copy of T (kernels read the tables through it)
  Tab* dTd = nullptr;     // device FP64 tables of this degree  // ev holds the power-iteration vector of an earlier estimate
};

template <typename R>
struct MgT {
  std::vector<MgLevelT<R>> levels;
  int smoother_degree = 3, coarse_degree = 8;
  // power iterations when restarting from the previous eigenvector (rho_for_fixed changes by about
  // 1e-4 between stages; the smoothed interval reaches 1.2 lambda_max)
  int warm_eig_iters = 2;
  KernelTimer* timer = nullptr;  // optional launch timing of the level kernels
  double smoothing_range = 15.0;
  // tabs_by_k: 1D tables per degree with the operator's 2k + 1 Gauss points; tabs_rq_by_k: the same
  // with k + 1 points (used when cfg.reduced_quadrature)
  int setup_levels(const OpCtx& fine, const MgConfig& cfg, const Tab* tabs_by_k, const Tab* tabs_rq_by_k);
  void release();
  // nodal interpolation of pres_fixed (and of rho_for_fixed when rho_changed) to every level
  void set_coefficients(const double* rho_fine, const double* pfix_fine, bool rho_changed);
  int estimate_eigenvalues(const OpCtx& fine, int iters, double* work_d, double* scal_d, double* scal_h);
  void apply(const OpCtx& fine, const double* v, double* z);
  // z = S_0 v with the finest level's operator (its precision, quadrature and cell solves)
  void level_vmult(const OpCtx& fine, const double* v, double* z);
  // fine-level w = A^{-1} B p in precision R on the level's coefficient copies
  void schur_velocity(const OpCtx& fine, const R* p, R* w);
  ~MgT() { release(); }
};

// The multigrid's transfer between degrees kc < kf (same mesh and node family, one component):
// transpose = false: out (degree kf) = P in (degree kc), the nodal interpolation of the coarse polynomial
// (the V-cycle's prolongation); true: out (degree kc) = P^T in (degree kf), its restriction. Returns
// nonzero for degrees outside 1 <= kc < kf <= 4.
int mg_transfer(const OpCtx& fine, const Tab* tabs_by_k, int kc, int kf, bool transpose, const double* in,
                double* out);

}  // namespace afl
