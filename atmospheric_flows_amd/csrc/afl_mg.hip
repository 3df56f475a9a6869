// p-multigrid preconditioner for the Schur (pressure Helmholtz) system S = D - C A^{-1} B,
// run in FP32 (or FP64) inside the FP64 flexible GMRES of afl_capi.hip.
//
// Levels share the mesh and lower the degree k -> max(1, k/2) -> ... -> 1. A level operator
// is the same Schur apply (k_schur_w + k_apply_C<WITH_D>) at the level's degree, with the
// coefficients rho_for_fixed and pres_fixed interpolated nodally from the fine level. Smoother:
// Chebyshev iteration of degree nu on P_D^{-1} S over [lambda_max / range, 1.1 lambda_max]
// (the reference's MG smoother settings: degree 3, smoothing range 15,
// Atmospheric_Flow.cc:487-499), preconditioned by the cell-local tensor-product inverse of
// D = M/(gamma-1) with the cell-mean Jacobian. lambda_max comes from a power iteration at
// set-up. Transfers: nodal interpolation of the coarse polynomial (prolongation) and its
// transpose (restriction). Coarsest level: Chebyshev over [0.9, 1.1 lambda_max].
#include <cmath>
#include <vector>

#include "afl_mg.h"

namespace afl {

template <typename R>
struct Mat5 { R m[5][5]; };

// ---- kernels ------------------------------------------------------------------------------
// out[node of level (KO)] = sum_j E[i][j] in[node of level (KI)] per cell, tensor product
// (E: N1o x N1i 1D matrix). One thread per output node. Used for coefficient interpolation
// (ADD = false, E = interpolation) and restriction (E = P^T).
template <int DIM, int KO, int KI, bool ADD, typename RI, typename RO>
__global__ __launch_bounds__(AFL_NT) void k_tensor_map(int64_t ncells, int ncomp, int64_t Nin, int64_t Nout,
                                                       Mat5<double> E, const RI* __restrict__ in,
                                                       RO* __restrict__ out, const RI* __restrict__ sub = nullptr) {
  constexpr int NO = KO + 1, NI = KI + 1;
  constexpr int PO = DIM == 2 ? NO * NO : NO * NO * NO;
  constexpr int PI = DIM == 2 ? NI * NI : NI * NI * NI;
  const int64_t total = ncells * PO * ncomp;
  for (int64_t idx = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * AFL_NT) {
    const int a = (int)(idx / (ncells * PO));
    const int64_t r = idx % (ncells * PO);
    const int64_t c = r / PO;
    const int i = (int)(r % PO);
    const int ix = i % NO, iy = (i / NO) % NO, iz = i / (NO * NO);
    const RI* src = in + a * Nin + c * PI;
    const RI* sb = sub ? sub + a * Nin + c * PI : nullptr;
    auto val = [&](int j) -> double { return sb ? (double)src[j] - (double)sb[j] : (double)src[j]; };
    double acc = 0.0;
    if (DIM == 2) {
#pragma unroll
      for (int jz = 0; jz < NI; ++jz) {
        double t = 0.0;
#pragma unroll
        for (int jx = 0; jx < NI; ++jx) t += E.m[ix][jx] * val(jz * NI + jx);
        acc += E.m[iy][jz] * t;
      }
    } else {
#pragma unroll
      for (int jz = 0; jz < NI; ++jz) {
        double tz = 0.0;
#pragma unroll
        for (int jy = 0; jy < NI; ++jy) {
          double t = 0.0;
#pragma unroll
          for (int jx = 0; jx < NI; ++jx) t += E.m[ix][jx] * val((jz * NI + jy) * NI + jx);
          tz += E.m[iy][jy] * t;
        }
        acc += E.m[iz][jz] * tz;
      }
    }
    RO* o = out + a * Nout + c * PO + i;
    if (ADD) *o += (RO)acc;
    else *o = (RO)acc;
  }
}

// z = (gamma - 1) / Jbar_c (MI x MI (x MI)) r : exact inverse of D on an affine cell with the
// cell's mean Jacobian (cell-local smoother preconditioner). One thread per node.
template <int DIM, int K, typename R>
__global__ __launch_bounds__(AFL_NT) void k_pd_inv(int64_t ncells, const R* __restrict__ jbar, Mat5<double> MI,
                                                   const R* __restrict__ r, R* __restrict__ z) {
  constexpr int N1 = K + 1;
  constexpr int NPC = DIM == 2 ? N1 * N1 : N1 * N1 * N1;
  const int64_t total = ncells * NPC;
  const double gm1 = kGamma - 1.0;
  for (int64_t idx = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * AFL_NT) {
    const int64_t c = idx / NPC;
    const int i = (int)(idx % NPC);
    const int ix = i % N1, iy = (i / N1) % N1, iz = i / (N1 * N1);
    const R* src = r + c * NPC;
    double acc = 0.0;
    if (DIM == 2) {
#pragma unroll
      for (int jz = 0; jz < N1; ++jz) {
        double t = 0.0;
#pragma unroll
        for (int jx = 0; jx < N1; ++jx) t += MI.m[ix][jx] * (double)src[jz * N1 + jx];
        acc += MI.m[iy][jz] * t;
      }
    } else {
#pragma unroll
      for (int jz = 0; jz < N1; ++jz) {
        double tz = 0.0;
#pragma unroll
        for (int jy = 0; jy < N1; ++jy) {
          double t = 0.0;
#pragma unroll
          for (int jx = 0; jx < N1; ++jx) t += MI.m[ix][jx] * (double)src[(jz * N1 + jy) * N1 + jx];
          tz += MI.m[iy][jy] * t;
        }
        acc += MI.m[iz][jz] * tz;
      }
    }
    z[idx] = (R)(gm1 / (double)jbar[c] * acc);
  }
}

// One fused Chebyshev smoothing step per node i (one thread per node):
//   FIRST:  z = P_D^{-1} rin;               d = c2 z;          x = d
//   else:   r' = rin - t (written to rout); z = P_D^{-1} r';  d = c1 d + c2 z;  x += d
// P_D^{-1} is cell-local, so every thread reads its cell's rin/t; the residual is written to a
// separate buffer (ping-pong) so no thread overwrites a value another thread still reads.
template <int DIM, int K, typename R, bool FIRST>
__global__ __launch_bounds__(AFL_NT) void k_cheb_step(int64_t ncells, const R* __restrict__ jbar, Mat5<double> MI,
                                                      const R* __restrict__ rin, const R* __restrict__ t,
                                                      R* __restrict__ rout, R c1, R c2, R* __restrict__ d,
                                                      R* __restrict__ x) {
  constexpr int N1 = K + 1;
  constexpr int NPC = DIM == 2 ? N1 * N1 : N1 * N1 * N1;
  const int64_t total = ncells * NPC;
  const double gm1 = kGamma - 1.0;
  for (int64_t idx = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * AFL_NT) {
    const int64_t c = idx / NPC;
    const int i = (int)(idx % NPC);
    const int ix = i % N1, iy = (i / N1) % N1, iz = i / (N1 * N1);
    const R* rc = rin + c * NPC;
    const R* tc = FIRST ? nullptr : t + c * NPC;
    auto res = [&](int j) -> double { return FIRST ? (double)rc[j] : (double)rc[j] - (double)tc[j]; };
    double acc = 0.0;
    if (DIM == 2) {
#pragma unroll
      for (int jz = 0; jz < N1; ++jz) {
        double a = 0.0;
#pragma unroll
        for (int jx = 0; jx < N1; ++jx) a += MI.m[ix][jx] * res(jz * N1 + jx);
        acc += MI.m[iy][jz] * a;
      }
    } else {
#pragma unroll
      for (int jz = 0; jz < N1; ++jz) {
        double tz = 0.0;
#pragma unroll
        for (int jy = 0; jy < N1; ++jy) {
          double a = 0.0;
#pragma unroll
          for (int jx = 0; jx < N1; ++jx) a += MI.m[ix][jx] * res((jz * N1 + jy) * N1 + jx);
          tz += MI.m[iy][jy] * a;
        }
        acc += MI.m[iz][jz] * tz;
      }
    }
    const R z = (R)(gm1 / (double)jbar[c] * acc);
    if (FIRST) {
      const R dn = c2 * z;
      d[idx] = dn;
      x[idx] = dn;
    } else {
      rout[idx] = (R)res(i);
      const R dn = (c1 != (R)0 ? c1 * d[idx] : (R)0) + c2 * z;  // d is unset before the first step
      d[idx] = dn;
      x[idx] += dn;
    }
  }
}

// cell-mean Jacobian: sum_q JxW_q (reference weights sum to one)
template <int DIM, int K, typename R>
__global__ __launch_bounds__(AFL_NT) void k_jbar(Geo g, const Tab* __restrict__ Tp, R* __restrict__ jbar) {
  const auto& T = *Tp;
  constexpr int NQ = 2 * K + 1;
  for (int64_t c = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; c < g.ncells; c += (int64_t)gridDim.x * AFL_NT) {
    int ic[3];
    ic[0] = (int)(c % g.nc[0]);
    if (DIM == 2) { ic[1] = (int)(c / g.nc[0]); ic[2] = 0; }
    else { ic[1] = (int)((c / g.nc[0]) % g.nc[1]); ic[2] = (int)(c / ((int64_t)g.nc[0] * g.nc[1])); }
    double acc = 0.0;
    const int nl = DIM == 2 ? NQ : NQ * NQ;
    for (int l = 0; l < nl; ++l) {
      LineGeo<DIM, K, double> lg;
      lg.init(g, T, ic, l);
      acc += (DIM == 2 ? lg.hx : lg.hx * lg.hy) * lg.c * lg.wxy;
    }
    jbar[c] = (R)acc;
  }
}

template <typename R>
__global__ void k_vec_axpby(int64_t n, R a, const R* __restrict__ x, R b, R* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
// Chebyshev update: x += d (if ax) ; d = c1 d + c2 z
template <typename R>
__global__ void k_cheb_d(int64_t n, R c1, R c2, const R* __restrict__ z, R* __restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = c1 * d[i] + c2 * z[i];
}
template <typename R>
__global__ void k_vec_scale(int64_t n, R a, const R* __restrict__ x, R* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}
// y = x / s[0] (x unchanged when s[0] is not positive), the norm s read on the device
template <typename R>
__global__ void k_vec_scale_inv_dev(int64_t n, const double* __restrict__ s, const R* __restrict__ x,
                                    R* __restrict__ y) {
  const R a = (R)(1.0 / (s[0] > 0 ? s[0] : 1.0));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}
template <typename R>
__global__ void k_vec_fill(int64_t n, R v, R* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}
template <typename R>
__global__ void k_vec_sub(int64_t n, const R* __restrict__ b, const R* __restrict__ t, R* __restrict__ r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - t[i];
}
template <typename RI, typename RO>
__global__ void k_convert(int64_t n, const RI* __restrict__ x, RO* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (RO)x[i];
}
template <typename R>
__global__ void k_hash_fill(int64_t n, uint64_t seed, R* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t v = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    v ^= v >> 31;
    x[i] = (R)((double)(v >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  }
}

static int grid_n(int64_t n) {
  int64_t b = (n + AFL_NT - 1) / AFL_NT;
  return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

#define AFL_DISPATCH_K(DIMV, KV, ...)                          \
  do {                                                         \
    if ((DIMV) == 2) {                                         \
      switch (KV) {                                            \
        case 1: { constexpr int D_ = 2, K_ = 1; __VA_ARGS__; } break; \
        case 2: { constexpr int D_ = 2, K_ = 2; __VA_ARGS__; } break; \
        case 3: { constexpr int D_ = 2, K_ = 3; __VA_ARGS__; } break; \
        case 4: { constexpr int D_ = 2, K_ = 4; __VA_ARGS__; } break; \
      }                                                        \
    } else {                                                   \
      switch (KV) {                                            \
        case 1: { constexpr int D_ = 3, K_ = 1; __VA_ARGS__; } break; \
        case 2: { constexpr int D_ = 3, K_ = 2; __VA_ARGS__; } break; \
        case 3: { constexpr int D_ = 3, K_ = 3; __VA_ARGS__; } break; \
        case 4: { constexpr int D_ = 3, K_ = 4; __VA_ARGS__; } break; \
      }                                                        \
    }                                                          \
  } while (0)

// (KO, KI) pairs with KO, KI in 1..4
#define AFL_DISPATCH_PAIR(DIMV, KOV, KIV, ...)                                  \
  do {                                                                          \
    const int key_ = (DIMV) * 100 + (KOV) * 10 + (KIV);                         \
    switch (key_) {                                                             \
      case 221: { constexpr int D_ = 2, KO_ = 2, KI_ = 1; __VA_ARGS__; } break; \
      case 231: { constexpr int D_ = 2, KO_ = 3, KI_ = 1; __VA_ARGS__; } break; \
      case 232: { constexpr int D_ = 2, KO_ = 3, KI_ = 2; __VA_ARGS__; } break; \
      case 241: { constexpr int D_ = 2, KO_ = 4, KI_ = 1; __VA_ARGS__; } break; \
      case 242: { constexpr int D_ = 2, KO_ = 4, KI_ = 2; __VA_ARGS__; } break; \
      case 243: { constexpr int D_ = 2, KO_ = 4, KI_ = 3; __VA_ARGS__; } break; \
      case 212: { constexpr int D_ = 2, KO_ = 1, KI_ = 2; __VA_ARGS__; } break; \
      case 213: { constexpr int D_ = 2, KO_ = 1, KI_ = 3; __VA_ARGS__; } break; \
      case 214: { constexpr int D_ = 2, KO_ = 1, KI_ = 4; __VA_ARGS__; } break; \
      case 223: { constexpr int D_ = 2, KO_ = 2, KI_ = 3; __VA_ARGS__; } break; \
      case 224: { constexpr int D_ = 2, KO_ = 2, KI_ = 4; __VA_ARGS__; } break; \
      case 234: { constexpr int D_ = 2, KO_ = 3, KI_ = 4; __VA_ARGS__; } break; \
      case 321: { constexpr int D_ = 3, KO_ = 2, KI_ = 1; __VA_ARGS__; } break; \
      case 331: { constexpr int D_ = 3, KO_ = 3, KI_ = 1; __VA_ARGS__; } break; \
      case 332: { constexpr int D_ = 3, KO_ = 3, KI_ = 2; __VA_ARGS__; } break; \
      case 341: { constexpr int D_ = 3, KO_ = 4, KI_ = 1; __VA_ARGS__; } break; \
      case 342: { constexpr int D_ = 3, KO_ = 4, KI_ = 2; __VA_ARGS__; } break; \
      case 343: { constexpr int D_ = 3, KO_ = 4, KI_ = 3; __VA_ARGS__; } break; \
      case 312: { constexpr int D_ = 3, KO_ = 1, KI_ = 2; __VA_ARGS__; } break; \
      case 313: { constexpr int D_ = 3, KO_ = 1, KI_ = 3; __VA_ARGS__; } break; \
      case 314: { constexpr int D_ = 3, KO_ = 1, KI_ = 4; __VA_ARGS__; } break; \
      case 323: { constexpr int D_ = 3, KO_ = 2, KI_ = 3; __VA_ARGS__; } break; \
      case 324: { constexpr int D_ = 3, KO_ = 2, KI_ = 4; __VA_ARGS__; } break; \
      case 334: { constexpr int D_ = 3, KO_ = 3, KI_ = 4; __VA_ARGS__; } break; \
      default: break;                                                           \
    }                                                                           \
  } while (0)

// ---- level operations (host) -------------------------------------------------------------
template <typename R>
static void level_schur(const MgLevelT<R>& L, const OpCtx& fine, const R* p, R* y) {
  const double coeff = fine.hyp == 1 ? fine.tab.at22 : fine.tab.at33;
  const R f = (R)(coeff * fine.dt / (fine.Ma * fine.Ma));
  const R cdt = (R)(coeff * fine.dt);
  const R tol = (R)L.cell_tol;
  launch_schur_pair<R>(L.dim, L.k, L.rq, L.geo, L.dT, L.stream, L.rho, L.lean() ? L.sinv : nullptr, L.pfix, f, cdt, tol,
                       L.cell_maxit, L.N, p, L.w, y);
}

template <typename R>
static void level_pinv(const MgLevelT<R>& L, const R* r, R* z) {
  Mat5<double> MI;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) MI.m[i][j] = L.MI[i][j];
  AFL_DISPATCH_K(L.dim, L.k, {
    k_pd_inv<D_, K_, R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.ncells, L.jbar, MI, r, z);
  });
}

// Chebyshev smoothing x <- x + p(P^{-1} S) P^{-1} (b - S x) of the given degree (number of x
// updates); zero_init starts from x = 0 and skips the first residual. One level Schur apply and
// one fused step kernel per update.
template <typename R>
static void cheb_step(const MgLevelT<R>& L, bool first, const R* rin, const R* t, R* rout, double c1, double c2,
                      R* d, R* x) {
  Mat5<double> MI;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) MI.m[i][j] = L.MI[i][j];
  AFL_DISPATCH_K(L.dim, L.k, {
    if (first)
      k_cheb_step<D_, K_, R, true><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.ncells, L.jbar, MI, rin, t, rout,
                                                                                    (R)c1, (R)c2, d, x);
    else
      k_cheb_step<D_, K_, R, false><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.ncells, L.jbar, MI, rin, t, rout,
                                                                                     (R)c1, (R)c2, d, x);
  });
}

// S p followed by one fused Chebyshev step: r' = rin - S p -> rout, d = c1 d + c2 P^{-1} r', x += d
template <typename R>
static void level_schur_cheb(const MgLevelT<R>& L, const OpCtx& fine, const R* p, const R* rin, R* rout, double c1,
                             double c2, R* d, R* x, KernelTimer* tm) {
  const double coeff = fine.hyp == 1 ? fine.tab.at22 : fine.tab.at33;
  const R f = (R)(coeff * fine.dt / (fine.Ma * fine.Ma));
  const R cdt = (R)(coeff * fine.dt);
  // algorithmic bytes: each distinct field once in the level precision, plus one scale per cell
  const double bw = (double)sizeof(R) * (double)((1 + L.dim) * L.N + L.ncells);
  const double bc = (double)sizeof(R) * (double)((7 + L.dim) * L.N + L.ncells);
  launch_schur_w_r<R>(L.dim, L.k, L.rq, L.geo, L.dT, L.stream, L.rho, L.lean() ? L.sinv : nullptr, f, (R)L.cell_tol,
                      L.cell_maxit, L.N, p, L.w, tm && L.lean() ? tm->next(kTimeLevelW, fine.hyp, bw) : nullptr);
  launch_apply_C_cheb<R>(L.dim, L.k, L.rq, L.geo, L.dT, L.stream, L.pfix, cdt, L.N, L.w, p, L.jbar, rin, rout, (R)c1, (R)c2,
                         d, x, tm ? tm->next(kTimeLevelCCheb, fine.hyp, bc) : nullptr);
}

template <typename R>
static void smooth(const MgLevelT<R>& L, const OpCtx& fine, const R* b, R* x, bool zero_init, int degree,
                   double lo, double hi, KernelTimer* tm) {
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma1 = theta / delta;
  const R* rcur = b;
  R* rnext = L.r;
  R* ralt = L.r2;
  if (zero_init) {
    cheb_step(L, true, b, (const R*)nullptr, (R*)nullptr, 0.0, 1.0 / theta, L.d, x);
  } else {
    level_schur_cheb(L, fine, x, b, rnext, 0.0, 1.0 / theta, L.d, x, tm);
    rcur = rnext;
    std::swap(rnext, ralt);
  }
  double rho = 1.0 / sigma1;
  for (int k = 1; k < degree; ++k) {
    const double rn = 1.0 / (2.0 * sigma1 - rho);
    level_schur_cheb(L, fine, L.d, rcur, rnext, rn * rho, 2.0 * rn / delta, L.d, x, tm);
    rcur = rnext;
    std::swap(rnext, ralt);
    rho = rn;
  }
}

template <typename R>
static void tensor_map(int dim, int ko, int ki, bool add, int64_t ncells, int64_t Nin, int64_t Nout,
                       const double (&E)[5][5], const R* in, R* out, hipStream_t st, const R* sub = nullptr) {
  Mat5<double> M;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) M.m[i][j] = E[i][j];
  const int npo = dim == 2 ? (ko + 1) * (ko + 1) : (ko + 1) * (ko + 1) * (ko + 1);
  const int g = grid_n(ncells * npo);
  AFL_DISPATCH_PAIR(dim, ko, ki, {
    if (add) k_tensor_map<D_, KO_, KI_, true, R, R><<<dim3(g), dim3(AFL_NT), 0, st>>>(ncells, 1, Nin, Nout, M, in, out, sub);
    else k_tensor_map<D_, KO_, KI_, false, R, R><<<dim3(g), dim3(AFL_NT), 0, st>>>(ncells, 1, Nin, Nout, M, in, out, sub);
  });
}

template <typename R>
static void vcycle(MgT<R>& mg, const OpCtx& fine, int l) {
  MgLevelT<R>& L = mg.levels[l];
  const int nl = (int)mg.levels.size();
  if (l == nl - 1) {
    smooth(L, fine, L.b, L.x, true, mg.coarse_degree, 0.9, 1.2 * L.lam_max, mg.timer);
    return;
  }
  const double hi = 1.2 * L.lam_max, lo = L.lam_max / mg.smoothing_range;
  smooth(L, fine, L.b, L.x, true, mg.smoother_degree, lo, hi, mg.timer);
  // residual b - S x, restricted to the coarser level (subtraction fused into the restriction)
  level_schur(L, fine, L.x, L.t);
  MgLevelT<R>& Cl = mg.levels[l + 1];
  tensor_map<R>(L.dim, Cl.k, L.k, false, L.ncells, L.N, Cl.N, Cl.Pt, L.b, Cl.b, L.stream, L.t);
  vcycle(mg, fine, l + 1);
  tensor_map<R>(L.dim, L.k, Cl.k, true, L.ncells, Cl.N, L.N, Cl.P, Cl.x, L.x, L.stream);
  smooth(L, fine, L.b, L.x, false, mg.smoother_degree, lo, hi, mg.timer);
}

// host 1D tables of a level (same node family as the fine level)
static void interp_matrix(const Tab& to, int kto, const Tab& from, int kfrom, double (&E)[5][5]);

template <typename R>
int MgT<R>::setup_levels(const OpCtx& fine, const MgConfig& cfg, const Tab* tabs_by_k, const Tab* tabs_rq_by_k) {
  smoother_degree = cfg.smoother_degree;
  coarse_degree = cfg.coarse_degree;
  smoothing_range = cfg.smoothing_range;
  std::vector<int> ks;
  int k = fine.k;
  ks.push_back(k);
  while (k > 1 && (int)ks.size() < cfg.max_levels) {
    k = k / 2 > 0 ? k / 2 : 1;
    ks.push_back(k);
  }
  release();
  levels.resize(ks.size());
  for (size_t l = 0; l < ks.size(); ++l) {
    MgLevelT<R>& L = levels[l];
    L.dim = fine.dim;
    L.k = ks[l];
    L.geo = fine.geo;
    L.ncells = fine.geo.ncells;
    const int n1 = L.k + 1;
    L.N = (int64_t)L.ncells * (L.dim == 2 ? n1 * n1 : n1 * n1 * n1);
    L.stream = fine.stream;
    L.cell_tol = cfg.cell_tol;
    L.cell_maxit = cfg.cell_maxit;
    L.rq = cfg.reduced_quadrature;
    const Tab& tb = tabs_by_k[L.k];
    const Tab& tq = L.rq ? tabs_rq_by_k[L.k] : tb;  // quadrature of the level operator
    for (int q = 0; q < 9; ++q) {
      for (int i = 0; i < 5; ++i) { L.T.S[q][i] = (R)tq.S[q][i]; L.T.D[q][i] = (R)tq.D[q][i]; }
      L.T.W[q] = (R)tq.W[q];
      L.T.QP[q] = (R)tq.QP[q];
    }
    for (int i = 0; i < 5; ++i) {
      L.T.ND[i] = (R)tb.ND[i];
      for (int j = 0; j < 5; ++j) { L.T.MI[i][j] = (R)tb.MI[i][j]; L.MI[i][j] = tb.MI[i][j]; }
    }
    interp_matrix(tb, L.k, tabs_by_k[fine.k], fine.k, L.E);  // fine -> this level
    if (l > 0) {
      interp_matrix(tabs_by_k[ks[l - 1]], ks[l - 1], tb, L.k, L.P);  // this (coarse) -> finer
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) L.Pt[i][j] = L.P[j][i];
    }
    if (hipMalloc((void**)&L.dT, sizeof(TabT<R>)) != hipSuccess ||
        hipMemcpy(L.dT, &L.T, sizeof(TabT<R>), hipMemcpyHostToDevice) != hipSuccess)
      return 3;
    if (hipMalloc((void**)&L.dTd, sizeof(Tab)) != hipSuccess ||
        hipMemcpy(L.dTd, &tb, sizeof(Tab), hipMemcpyHostToDevice) != hipSuccess)
      return 3;
    const int64_t nv = L.N * L.dim;
    R** bufs[] = {&L.rho, &L.pfix, &L.x, &L.b, &L.r, &L.r2, &L.z, &L.d, &L.t, &L.ev};
    if (hipMalloc((void**)&L.sinv, L.ncells * sizeof(R)) != hipSuccess) return 3;
    L.ev_valid = false;
    for (R** b : bufs)
      if (hipMalloc((void**)b, L.N * sizeof(R)) != hipSuccess) return 3;
    if (hipMalloc((void**)&L.w, nv * sizeof(R)) != hipSuccess) return 3;
    if (hipMalloc((void**)&L.jbar, L.ncells * sizeof(R)) != hipSuccess) return 3;
    AFL_DISPATCH_K(L.dim, L.k, {
      k_jbar<D_, K_, R><<<dim3(grid_n(L.ncells)), dim3(AFL_NT), 0, L.stream>>>(L.geo, L.dTd, L.jbar);
    });
  }
  return 0;
}

template <typename R>
void MgT<R>::release() {
  for (auto& L : levels) {
    R* bufs[] = {L.rho, L.pfix, L.x, L.b, L.r, L.r2, L.z, L.d, L.t, L.w, L.jbar, L.ev, L.sinv};
    for (R* b : bufs)
      if (b) hipFree(b);
    if (L.dT) hipFree(L.dT);
    if (L.dTd) hipFree(L.dTd);
  }
  levels.clear();
}

// coefficients of the current stage: nodal interpolation of rho_for_fixed and pres_fixed
template <typename R>
void MgT<R>::set_coefficients(const double* rho_fine, const double* pfix_fine, bool rho_changed) {
  for (auto& L : levels) {
    Mat5<double> M;
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) M.m[i][j] = L.E[i][j];
    const int npo = L.dim == 2 ? (L.k + 1) * (L.k + 1) : (L.k + 1) * (L.k + 1) * (L.k + 1);
    const int g = grid_n(L.ncells * npo);
    const int kf = levels[0].k;
    const int64_t Nf = levels[0].N;
    if (L.k == kf) {
      if (rho_changed) k_convert<double, R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, rho_fine, L.rho);
      k_convert<double, R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, pfix_fine, L.pfix);
    } else {
      AFL_DISPATCH_PAIR(L.dim, L.k, kf, {
        if (rho_changed)
          k_tensor_map<D_, KO_, KI_, false, double, R><<<dim3(g), dim3(AFL_NT), 0, L.stream>>>(L.ncells, 1, Nf, L.N, M,
                                                                                              rho_fine, L.rho);
        k_tensor_map<D_, KO_, KI_, false, double, R><<<dim3(g), dim3(AFL_NT), 0, L.stream>>>(L.ncells, 1, Nf, L.N, M,
                                                                                            pfix_fine, L.pfix);
      });
    }
    if (rho_changed && L.lean()) launch_cell_rho_mass_inv<R>(L.dim, L.k, L.rq, L.geo, L.dT, L.stream, L.rho, L.sinv);
  }
}

// lambda_max of P_D^{-1} S on every level by power iteration (norms read on the host). The
// iteration vector persists per level, so a re-estimate after a change of the stage
// coefficients starts from the previous eigenvector and needs only a few steps.
template <typename R>
int MgT<R>::estimate_eigenvalues(const OpCtx& fine, int iters, double* work_d, double* scal_d, double* scal_h) {
  for (auto& L : levels) {
    int its = iters;
    if (!L.ev_valid) {
      k_hash_fill<R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, 12345, L.ev);
    } else {
      its = warm_eig_iters < iters ? warm_eig_iters : iters;
    }
    for (int it = 0; it < its; ++it) {
      level_schur(L, fine, L.ev, L.t);
      level_pinv(L, L.t, L.z);
      // |z| / |x| (norms accumulated in double)
      if constexpr (sizeof(R) == 4) {
        vec_norm2_f32(L.stream, L.N, (const float*)L.z, work_d + L.N, scal_d);
        vec_norm2_f32(L.stream, L.N, (const float*)L.ev, work_d + L.N, scal_d + 1);
      } else {
        vec_norm2(L.stream, L.N, (const double*)L.z, work_d + L.N, scal_d);
        vec_norm2(L.stream, L.N, (const double*)L.ev, work_d + L.N, scal_d + 1);
      }
      // x = z / |z| with the norm read on the device: the host waits only for the last quotient
      k_vec_scale_inv_dev<R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, scal_d, L.z, L.ev);
    }
    double lam = 1.0;
    if (its > 0) {
      if (hipMemcpyAsync(scal_h, scal_d, 2 * sizeof(double), hipMemcpyDeviceToHost, L.stream) != hipSuccess) return 3;
      if (hipStreamSynchronize(L.stream) != hipSuccess) return 3;
      lam = scal_h[1] > 0 ? scal_h[0] / scal_h[1] : 1.0;
    }
    L.lam_max = lam > 1.0 ? lam : 1.0;
    L.ev_valid = true;
  }
  return 0;
}

// z = MG(v): v and z are FP64 fine-level vectors
template <typename R>
void MgT<R>::apply(const OpCtx& fine, const double* v, double* z) {
  MgLevelT<R>& L = levels[0];
  k_convert<double, R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, v, L.b);
  vcycle(*this, fine, 0);
  k_convert<R, double><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, L.x, z);
}

template <typename R>
void MgT<R>::level_vmult(const OpCtx& fine, const double* v, double* z) {
  MgLevelT<R>& L = levels[0];
  k_convert<double, R><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, v, L.b);
  level_schur(L, fine, L.b, L.t);
  k_convert<R, double><<<dim3(grid_n(L.N)), dim3(AFL_NT), 0, L.stream>>>(L.N, L.t, z);
}

template <typename R>
void MgT<R>::schur_velocity(const OpCtx& fine, const R* p, R* w) {
  const MgLevelT<R>& L = levels[0];
  const double coeff = fine.hyp == 1 ? fine.tab.at22 : fine.tab.at33;
  launch_schur_w_r<R>(L.dim, L.k, L.rq, L.geo, L.dT, L.stream, L.rho, L.lean() ? L.sinv : nullptr,
                      (R)(coeff * fine.dt / (fine.Ma * fine.Ma)),
                      (R)L.cell_tol, L.cell_maxit, L.N, p, w);
}

static void interp_matrix(const Tab& to, int kto, const Tab& from, int kfrom, double (&E)[5][5]) {
  // E[i][j] = phi^{from}_j(x^{to}_i), Lagrange basis on `from`'s nodes
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) E[i][j] = 0.0;
  for (int i = 0; i <= kto; ++i)
    for (int j = 0; j <= kfrom; ++j) {
      double v = 1.0;
      for (int m = 0; m <= kfrom; ++m)
        if (m != j) v *= (to.ND[i] - from.ND[m]) / (from.ND[j] - from.ND[m]);
      E[i][j] = v;
    }
}

int mg_transfer(const OpCtx& fine, const Tab* tabs_by_k, int kc, int kf, bool transpose, const double* in,
                double* out) {
  if (kc < 1 || kf > 4 || kc >= kf) return 1;
  double P[5][5], Pt[5][5];
  interp_matrix(tabs_by_k[kf], kf, tabs_by_k[kc], kc, P);  // the V-cycle's prolongation matrix
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) Pt[i][j] = P[j][i];
  const int64_t nc = fine.geo.ncells;
  const int64_t Nc = nc * (fine.dim == 2 ? (kc + 1) * (kc + 1) : (kc + 1) * (kc + 1) * (kc + 1));
  const int64_t Nf = nc * (fine.dim == 2 ? (kf + 1) * (kf + 1) : (kf + 1) * (kf + 1) * (kf + 1));
  if (!transpose) tensor_map<double>(fine.dim, kf, kc, false, nc, Nc, Nf, P, in, out, fine.stream);
  else tensor_map<double>(fine.dim, kc, kf, false, nc, Nf, Nc, Pt, in, out, fine.stream);
  return 0;
}

template struct MgT<float>;
template struct MgT<double>;

}  // namespace afl
