// Host/device shared internals of libafl_hip.so (not part of the C-ABI).
#pragma once

#include <cstdlib>

#include <hip/hip_ext.h>

#include <vector>

#include "afl_common.h"

namespace afl {

// ARK2 tableau, euler_operator.h:214-240.
struct TableauC {
  double a21, a22, a31, a32, a33;
  double at21, at22, at31, at32, at33;
  double b1, b2, b3;
};

// HIP-event timing of every launch of the kernels in a mask of AFL_TIME_* flags (benchmark
// instrumentation). The launch itself records the pair (hipExtLaunchKernelGGL start/stop events), so
// the events bracket exactly the kernel's execution. Each timed launch also records its flag, its
// IMEX stage and its algorithmic bytes.
struct KernelTimer {
  unsigned mask = 0;  // 0: off
  int used = 0;
  int dropped = 0;
  std::vector<hipEvent_t> ev;  // pairs (start, stop)
  std::vector<int> flag, stage;
  std::vector<double> bytes;
  // the event pair for the next launch of kernel f, or nullptr when not timing it
  hipEvent_t* next(int f, int stg = 0, double b = 0.0) {
    if (!(mask & (unsigned)f)) return nullptr;
    if (2 * used + 1 >= (int)ev.size()) { ++dropped; return nullptr; }
    flag[used] = f;
    stage[used] = stg;
    bytes[used] = b;
    return &ev[2 * used++];
  }
};
enum KernelId { kTimeLevelW = 1, kTimeLevelCCheb = 2, kTimeRhsRho = 4, kTimeRhsU = 8, kTimeRhsP = 16,
                kTimeRhsPUfix = 32, kTimeSchurW = 64, kTimeApplyC = 128, kTimeMassSolve = 256, kTimeApplyB = 512,
                kTimeASolve = 1024 };

// Positional sources of an explicit-stage apply (kernel argument by value).
struct Srcs { const double* p[12]; };

// src.p[i] for a runtime i without indexing the kernel argument through memory
__device__ __attribute__((always_inline)) inline const double* src_at(const Srcs& s, int i) {
  const double* v = s.p[0];
#pragma unroll
  for (int j = 1; j < 12; ++j) v = (i == j) ? s.p[j] : v;
  return v;
}

// Stage weights of E2: advective w[j], pressure wp[j] per explicit state, gravity gw[j] per density.
struct E2Coef { double w[3], wp[3], gw[3]; };
// Stage weights of E3: kinetic w[j], enthalpy wt[j] per explicit state, gravity work gw[j].
struct E3Coef { double w[3], wt[3], gw[3]; };

// Everything a launcher needs; a plain copy of the operator state.
struct OpCtx {
  int dim = 2, k = 1;
  Geo geo{};
  Tab T{};          // host copy of the basis tables
  const Tab* dT = nullptr;  // device copy (kernels read the tables through this pointer)
  int64_t N = 0;          // nodes per scalar field
  double Ma = 1, Fr = 1, dt = 1;
  unsigned hyp = 1;       // HYPERBOLIC_stage
  TableauC tab{};
  hipStream_t stream = nullptr;
  // relative tolerance of the cell-local mass-type solves (the reference's inner A solve: CG to 1e-12,
  // euler_operator.h:2391); on the 3D hill it takes 3 Chebyshev steps in most cells, 4 at 1e-15
  double cell_tol = 1e-13;
  int cell_maxit = 40;
  KernelTimer* timer = nullptr;  // set while afl_kernel_timer_* is active
};

// Launch on the context's stream; when the timer times kernel `flag`, bracket it with its events.
template <typename F, typename... Args>
inline void launch_timed(const OpCtx& o, int flag, int stage, double bytes, F kern, dim3 gr, dim3 bl, Args... args) {
  hipEvent_t* ev = o.timer ? o.timer->next(flag, stage, bytes) : nullptr;
  hipExtLaunchKernelGGL(kern, gr, bl, 0, o.stream, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, 0, args...);
}

// operator kernels (afl_ops.hip)
void launch_rhs_rho(const OpCtx& o, int stage, const double* const* src, double* dst);
void launch_rhs_u(const OpCtx& o, int stage, const double* const* src, double* dst);
// full = false at stages 1-2: E3 without its two u_fixed terms (launch_rhs_p_ufix adds them)
void launch_rhs_p(const OpCtx& o, int stage, const double* const* src, double* dst, bool full = true);
void launch_rhs_p_ufix(const OpCtx& o, int stage, const double* rho_new, const double* u_fix, const double* base,
                       double* dst);
void launch_mass_apply(const OpCtx& o, const double* rho, double scale, int ncomp, const double* src, double* dst);
// 3D, k <= 3: wave-per-cell kernels (afl_ops3d.hip); `by` is the launch's algorithmic bytes
void launch3_mass_solve(const OpCtx& o, double by, const double* rho, double scale, int ncomp, const double* b,
                        double* x);
void launch3_apply_B(const OpCtx& o, double by, double f, const double* p, double* dst);
void launch3_schur_w(const OpCtx& o, double by, double f, const double* rho, const double* p, double* w);
void launch3_apply_C(const OpCtx& o, double by, double cdt, const double* pfix, const double* w, const double* pD,
                     double* dst);
void launch3_rhs_rho(const OpCtx& o, double by, int stage, const Srcs& S, double w0, double w1, double w2,
                     double* dst);
void launch3_rhs_u(const OpCtx& o, double by, int stage, const Srcs& S, const E2Coef& cf, double iMa2, double iFr2,
                   double* dst);
void launch3_rhs_p(const OpCtx& o, double by, int stage, bool full, const Srcs& S, const E3Coef& cf, double Ma2,
                   double iFr2, double* dst);
void launch3_rhs_p_ufix(const OpCtx& o, double by, int stage, const double* rho_new, const double* u_fix,
                        double half_ma2, double cg, const double* base, double* dst);
inline bool use_w3(const OpCtx& o) { return o.dim == 3 && o.k <= 3; }
void launch_inv_diag(const OpCtx& o, const double* rho, double scale, int ncomp, double* dst);
void launch_apply_B(const OpCtx& o, const double* p, double* dst);
void launch_apply_C(const OpCtx& o, const double* pfix, const double* w, const double* pD, double* dst);
void launch_mass_solve(const OpCtx& o, const double* rho, double scale, int ncomp, const double* b, double* x);
void launch_schur_w(const OpCtx& o, const double* rho, const double* p, double* w);
// observation points (afl_observe.hip): located on the host, device copies here
struct ObsDev {
  int npts = 0, ncp = 0;            // points; cells that hold points
  int64_t* cell = nullptr;          // cell of point k
  double* ref = nullptr;            // reference coordinates, 3 per point
  double* z = nullptr;              // physical height of point k
  int64_t* ccell = nullptr;         // cells with points (ncp)
  int* cstart = nullptr;            // CSR offsets into order (ncp + 1)
  int* order = nullptr;             // point indices grouped by cell
};
int locate_points(const Geo& g, int dim, int npts, const double* xyz, std::vector<int64_t>& cell,
                  std::vector<double>& ref, std::vector<double>& zpt);
void launch_observe(const OpCtx& o, const ObsDev& ob, const double* rho, const double* p, double theta0, double n2,
                    double p0, double* out, const double* wbar, double* fr, double* fp);
void launch_observe_t(const OpCtx& o, const ObsDev& ob, const double* fr, const double* fp, double* rbar,
                      double* pbar);

// discrete-adjoint building blocks (afl_adjoint.hip)
void launch_vjp_rhs_rho(const OpCtx& o, int stage, const double* rbar, const double* const* src, double* const* out);
void launch_vjp_rhs_u(const OpCtx& o, int stage, const double* rbar, const double* const* src, double* const* out);
void launch_vjp_rhs_p(const OpCtx& o, int stage, const double* rbar, const double* const* src, double* const* out);
void launch_apply_Bt(const OpCtx& o, const double* y, double* dst);
void launch_apply_C_coef_t(const OpCtx& o, const double* w, const double* cb, double* dst);
// dst = scaleD M cb + C_coef_t(wE, cb) - C_coef_t(wB, cb) - acc (acc may be dst)
void launch_apply_C_coef_pair_t(const OpCtx& o, const double* wE, const double* wB, const double* cb, double scaleD,
                                const double* acc, double* dst);
// dst = int phi (a . b) over DIM-vector fields a, b
void launch_dot_mass(const OpCtx& o, const double* a, const double* b, double* dst);
// dst_d = int phi lam rho (alpha u_d + beta [d = vertical])
void launch_scaled_velocity_mass(const OpCtx& o, const double* lam, const double* rho, const double* u, double alpha,
                                 double beta, double* dst);
void launch_apply_C_tangent(const OpCtx& o, const double* pf, const double* dw, const double* w0, double* dst);
void launch_apply_C_tangent_t(const OpCtx& o, const double* pf, const double* cb, const double* w0, double* dst);
// launch_apply_C_tangent_t minus launch_scaled_velocity_mass(cb, rho, u, alpha, beta) in one pass
void launch_apply_C_tangent_t_svm(const OpCtx& o, const double* pf, const double* cb, const double* w0, const double* rho,
                                  const double* u, double alpha, double beta, double* dst);
// Multigrid level operator in precision R. sinv != nullptr selects the mean-density block
// solve (w = P^{-1} B p, sinv per cell from launch_cell_rho_mass_inv); otherwise the cell-local
// A solve runs to tol with rho. rq selects the level quadrature: k + 1 Gauss points per direction
// (rq) or the operator's 2k + 1.
template <typename R>
void launch_schur_pair(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st, const R* rho,
                       const R* sinv, const R* pfix, R f, R cdt, R tol, int maxit, int64_t N, const R* p, R* w, R* y);
template <typename R>
void launch_schur_w_r(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st, const R* rho,
                      const R* sinv, R f, R tol, int maxit, int64_t N, const R* p, R* w, hipEvent_t* ev = nullptr);
template <typename R>
void launch_apply_C_cheb(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st, const R* pfix,
                         R cdt, int64_t N, const R* w, const R* p, const R* jbar, const R* rin, R* rout, R c1, R c2,
                         R* d, R* x, hipEvent_t* ev = nullptr);
template <typename R>
void launch_cell_rho_mass_inv(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st,
                              const R* rho, R* sinv);

// diagnostics (afl_diag.hip): out4 = {-min rho, max |rho|, max celerity, max speed};
// out2 = {sum JxW (a-b)^2, sum JxW b^2}; work holds diag_work_size(o) doubles
int64_t diag_work_size(const OpCtx& o);
void launch_nodal_stats(const OpCtx& o, const double* rho, const double* u, const double* p, double* work,
                        double* out4);
// L2 difference sums into out2[0..1]; work holds 2 x (grid blocks) partial sums, the grid capped at
// max_blocks (> 0; 0: one block per cell block). l2_blocks gives the grid it uses.
void launch_l2_difference(const OpCtx& o, int ncomp, const double* a, const double* b, double* work, double* out2,
                          int64_t max_blocks = 0);
int64_t l2_blocks(const OpCtx& o, int64_t max_blocks);

// vector kernels (afl_blas.hip); reductions leave their result in device memory
void vec_copy(hipStream_t st, int64_t n, const double* x, double* y);
void vec_axpby(hipStream_t st, int64_t n, double a, const double* x, double b, double* y);  // y = a x + b y
void vec_sub_mul(hipStream_t st, int64_t n, const double* b, const double* y, const double* d, double* out);  // (b-y)*d
void vec_sub(hipStream_t st, int64_t n, const double* a, const double* b, double* out);  // out = a - b
// dst[p][0..n[p]) += src[p][...] for p < count, in one launch (distinct dst buffers)
struct AccList {
  static constexpr int kMax = 12;
  double* dst[kMax];
  const double* src[kMax];
  int64_t n[kMax];
  int count = 0;
  void add(double* d, const double* s, int64_t len) {
    if (count >= kMax) std::abort();  // the callers add at most 11 pairs
    dst[count] = d;
    src[count] = s;
    n[count] = len;
    ++count;
  }
};
void vec_acc_multi(hipStream_t st, const AccList& L);
void vec_mul(hipStream_t st, int64_t n, const double* x, const double* d, double* out);
void vec_addmul(hipStream_t st, int64_t n, const double* d, const double* z, double* x);  // x += d*z
void vec_scale_by_inv(hipStream_t st, int64_t n, const double* s, double* x);  // x /= *s
void vec_fill(hipStream_t st, int64_t n, double v, double* x);
// h[j] = <V_j, w> for j < m (V_j = V + j*ldv); partial sums use `work` (>= red_work_size(m) doubles)
void vec_mdot(hipStream_t st, int64_t n, const double* V, int64_t ldv, int m, const double* w, double* work, double* h);
// w -= sum_j h[j] V_j  (h on device);  x += sum_j y[j] V_j with y on device
void vec_maxpy(hipStream_t st, int64_t n, const double* V, int64_t ldv, int m, const double* h, double sign, double* w);
void vec_norm2(hipStream_t st, int64_t n, const double* x, double* work, double* out);  // sqrt(<x,x>)
void vec_norm2_f32(hipStream_t st, int64_t n, const float* x, double* work, double* out);  // FP64 accumulation
void vec_linf(hipStream_t st, int64_t n, const double* x, double* work, double* out);
void vec_linf_diff(hipStream_t st, int64_t n, const double* x, const double* y, double* work, double* out);
int64_t red_work_size(int m);
// state initialisation and sponge (afl_state.hip)
struct HillParams { double theta0, N2, p0, u0, sigma; uint64_t seed; };
void launch_init_state(const OpCtx& o, const HillParams& hp, double* rho, double* u, double* p);
void launch_background(const OpCtx& o, const HillParams& hp, double* rho, double* p, double* th);
void launch_sponge(const OpCtx& o, double z_start, double alpha_dt, const double* uref, double* u);

}  // namespace afl
