// Device diagnostics of a state (Atmospheric_Flow.cc:992-1087 and 925-985): nodal density
// bounds, maximum celerity and speed (the reference evaluates them with QGaussLobatto(k+1) on
// FE_DGQ(k) with Gauss-Lobatto support points, i.e. at the nodes), and the L2 norm of a
// difference of two discrete fields with the operators' Gauss quadrature.
#include "afl_internal.h"

namespace afl {

static constexpr int kDiagBlocks = 1024;

template <bool MAX>
__device__ inline double diag_block_reduce(double v) {
  __shared__ double sh[AFL_NT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = MAX ? fmax(v, o) : v + o;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int i = 1; i < (int)(blockDim.x / 64); ++i) r = MAX ? fmax(r, sh[i]) : r + sh[i];
  }
  __syncthreads();
  return r;
}

// part[j * kDiagBlocks + b], j = 0: max(-rho) (i.e. -min rho), 1: max |rho|, 2: max sqrt(gamma p / rho),
// 3: max sqrt(u0^2 + u1^2) (the reference's get_max_velocity uses the first two components)
__global__ __launch_bounds__(AFL_NT) void k_nodal_stats(int64_t n, const double* __restrict__ rho,
                                                        const double* __restrict__ u, const double* __restrict__ p,
                                                        double* __restrict__ part) {
  double a0 = -1e300, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AFL_NT) {
    const double r = rho[i];
    a0 = fmax(a0, -r);
    a1 = fmax(a1, fabs(r));
    if (p) a2 = fmax(a2, sqrt(kGamma * p[i] / r));
    if (u) {
      const double u0 = u[i], u1 = u[n + i];
      a3 = fmax(a3, sqrt(u0 * u0 + u1 * u1));
    }
  }
  a0 = diag_block_reduce<true>(a0);
  a1 = diag_block_reduce<true>(a1);
  a2 = diag_block_reduce<true>(a2);
  a3 = diag_block_reduce<true>(a3);
  if (threadIdx.x == 0) {
    part[0 * kDiagBlocks + blockIdx.x] = a0;
    part[1 * kDiagBlocks + blockIdx.x] = a1;
    part[2 * kDiagBlocks + blockIdx.x] = a2;
    part[3 * kDiagBlocks + blockIdx.x] = a3;
  }
}

// out[j] = max_b part[j * stride + b] (MAX) or sum (!MAX), j = blockIdx.x
template <bool MAX>
__global__ __launch_bounds__(AFL_NT) void k_diag_finish(const double* __restrict__ part, int64_t nb, int64_t stride,
                                                        double* __restrict__ out) {
  const int j = blockIdx.x;
  double a = MAX ? -1e300 : 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += AFL_NT) {
    const double v = part[j * stride + b];
    a = MAX ? fmax(a, v) : a + v;
  }
  a = diag_block_reduce<MAX>(a);
  if (threadIdx.x == 0) out[j] = a;
}

// Per-block partial sums of sum_q JxW (a - b)^2 and sum_q JxW b^2 over ncomp components
// (b == nullptr: b = 0): block i of the grid takes the cell blocks i, i + grid, ... < ncb and writes
// part[i] and part[gridDim.x + i] (every block the same trip count, so the cell barriers match).
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_l2_partial(Geo g, const Tab* __restrict__ Tp, int ncomp, int64_t N,
                                                       const double* __restrict__ a, const double* __restrict__ b,
                                                       int64_t ncb, double* __restrict__ part) {
  const auto& T = *Tp;
  using C = Cfg<DIM, K>;
  constexpr int NQ = C::NQ;
  __shared__ double shw[C::CPB * (C::EV1 + C::EV2)];
  double se = 0.0, sn = 0.0;
  for (int64_t vb = blockIdx.x; vb < ncb; vb += gridDim.x) {
    Slot<DIM, K> s(g, vb);
    const int cbs = s.cb < C::CPB ? s.cb : 0;
    double* t1 = shw + cbs * (C::EV1 + C::EV2);
    double* t2 = t1 + C::EV1;
    LineGeo<DIM, K> lg;
    lg.init(g, T, s.ic, s.l);
    for (int c = 0; c < ncomp; ++c) {
      double va[NQ], vbq[NQ];
      eval_line<DIM, K>(a + c * N, s, t1, t2, T, va);
      if (b) eval_line<DIM, K>(b + c * N, s, t1, t2, T, vbq);
      else for (int q = 0; q < NQ; ++q) vbq[q] = 0.0;
      if (s.act) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const double w = lg.jxw(g, T, q), d = va[q] - vbq[q];
          se += w * d * d;
          sn += w * vbq[q] * vbq[q];
        }
      }
    }
  }
  se = diag_block_reduce<false>(se);
  sn = diag_block_reduce<false>(sn);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = se;
    part[gridDim.x + blockIdx.x] = sn;
  }
}

int64_t l2_blocks(const OpCtx& o, int64_t max_blocks) {
  int cpb = 1;
  switch (o.dim * 10 + o.k) {
#define AFL_L2_CPB(D_, K_) case D_ * 10 + K_: cpb = Cfg<D_, K_>::CPB; break;
    AFL_L2_CPB(2, 1) AFL_L2_CPB(2, 2) AFL_L2_CPB(2, 3) AFL_L2_CPB(2, 4)
    AFL_L2_CPB(3, 1) AFL_L2_CPB(3, 2) AFL_L2_CPB(3, 3) AFL_L2_CPB(3, 4)
#undef AFL_L2_CPB
    default: break;
  }
  const int64_t ncb = (o.geo.ncells + cpb - 1) / cpb;
  return max_blocks > 0 && ncb > max_blocks ? max_blocks : ncb;
}

// Work of the diagnostics: 4 x kDiagBlocks nodal-statistics partials, or 2 partial sums per block of
// the L2 pass (one block per block of Cfg::CPB cells)
int64_t diag_work_size(const OpCtx& o) {
  const int64_t nb = l2_blocks(o, 0);
  return 2 * nb > 4 * kDiagBlocks ? 2 * nb : 4 * kDiagBlocks;
}

void launch_nodal_stats(const OpCtx& o, const double* rho, const double* u, const double* p, double* work,
                        double* out4) {
  const int64_t nb64 = (o.N + AFL_NT - 1) / AFL_NT;
  const int nb = nb64 < kDiagBlocks ? (int)nb64 : kDiagBlocks;
  k_nodal_stats<<<dim3(nb), dim3(AFL_NT), 0, o.stream>>>(o.N, rho, u, p, work);
  k_diag_finish<true><<<dim3(4), dim3(AFL_NT), 0, o.stream>>>(work, nb, kDiagBlocks, out4);
}

void launch_l2_difference(const OpCtx& o, int ncomp, const double* a, const double* b, double* work, double* out2,
                          int64_t max_blocks) {
  switch (o.dim * 10 + o.k) {
#define AFL_L2_CASE(D_, K_)                                                                            \
  case D_ * 10 + K_: {                                                                                 \
    using C = Cfg<D_, K_>;                                                                             \
    const int64_t ncb = (o.geo.ncells + C::CPB - 1) / C::CPB;                                          \
    const int64_t nb = max_blocks > 0 && ncb > max_blocks ? max_blocks : ncb;                          \
    k_l2_partial<D_, K_><<<dim3((unsigned)nb), dim3(C::NT), 0, o.stream>>>(o.geo, o.dT, ncomp, o.N, a, b, ncb, \
                                                                            work);                     \
    k_diag_finish<false><<<dim3(2), dim3(AFL_NT), 0, o.stream>>>(work, nb, nb, out2);                  \
  } break;
    AFL_L2_CASE(2, 1) AFL_L2_CASE(2, 2) AFL_L2_CASE(2, 3) AFL_L2_CASE(2, 4)
    AFL_L2_CASE(3, 1) AFL_L2_CASE(3, 2) AFL_L2_CASE(3, 3) AFL_L2_CASE(3, 4)
#undef AFL_L2_CASE
    default: break;
  }
}

}  // namespace afl
