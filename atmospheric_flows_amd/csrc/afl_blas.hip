// Vector kernels for the device-resident Krylov solvers: streaming updates and
// deterministic two-stage reductions (wavefront shuffles, then one block per result).
#include "afl_internal.h"

namespace afl {

static constexpr int kRedBlocks = 1024;

static inline int grid_1d(int64_t n) {
  int64_t b = (n + AFL_NT - 1) / AFL_NT;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

__global__ void k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = x[i];
}
__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
__global__ void k_sub_mul(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                          const double* __restrict__ d, double* __restrict__ o) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    o[i] = (b[i] - y[i]) * d[i];
}
__global__ void k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    o[i] = a[i] - b[i];
}
__global__ void k_acc_multi(AccList L) {
  const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  for (int p = 0; p < L.count; ++p) {
    double* __restrict__ d = L.dst[p];
    const double* __restrict__ s = L.src[p];
    for (int64_t i = i0; i < L.n[p]; i += step) d[i] += s[i];
  }
}
__global__ void k_mul(int64_t n, const double* __restrict__ x, const double* __restrict__ d, double* __restrict__ o) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    o[i] = x[i] * d[i];
}
__global__ void k_addmul(int64_t n, const double* __restrict__ d, const double* __restrict__ z, double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] += d[i] * z[i];
}
__global__ void k_scale_by_inv(int64_t n, const double* __restrict__ s, double* __restrict__ x) {
  const double v = *s;
  const double f = v != 0.0 ? 1.0 / v : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= f;
}
__global__ void k_fill(int64_t n, double v, double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = v;
}

// block-wide sum/max of one value per thread (wave64 shuffles, then LDS across 4 waves)
template <bool MAX>
__device__ inline double block_reduce(double v) {
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
    for (int i = 1; i < AFL_NT / 64; ++i) r = MAX ? fmax(r, sh[i]) : r + sh[i];
  }
  __syncthreads();
  return r;
}

// partial[j*kRedBlocks + b] = sum over block b's slice of V_j . w
__global__ __launch_bounds__(AFL_NT) void k_mdot_partial(int64_t n, const double* __restrict__ V, int64_t ldv,
                                                         const double* __restrict__ w, double* __restrict__ part) {
  const int j = blockIdx.y;
  const double* v = V + j * ldv;
  double a = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AFL_NT) a += v[i] * w[i];
  a = block_reduce<false>(a);
  if (threadIdx.x == 0) part[j * (int64_t)kRedBlocks + blockIdx.x] = a;
}

template <bool MAX, bool DIFF>
__global__ __launch_bounds__(AFL_NT) void k_norm_partial(int64_t n, const double* __restrict__ x,
                                                         const double* __restrict__ y, double* __restrict__ part) {
  double a = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AFL_NT) {
    const double v = DIFF ? x[i] - y[i] : x[i];
    a = MAX ? fmax(a, fabs(v)) : a + v * v;
  }
  a = block_reduce<MAX>(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// FP32 vector, FP64 accumulation: partial sums of x . x
__global__ __launch_bounds__(AFL_NT) void k_norm_partial_f32(int64_t n, const float* __restrict__ x,
                                                             double* __restrict__ part) {
  double a = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AFL_NT) {
    const double v = (double)x[i];
    a += v * v;
  }
  a = block_reduce<false>(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// out[j] = fin(sum_b part[j*kRedBlocks + b]) for j < m
template <int MODE>  // 0 sum, 1 sqrt(sum), 2 max
__global__ __launch_bounds__(AFL_NT) void k_finish(const double* __restrict__ part, int nb, double* __restrict__ out) {
  const int j = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += AFL_NT) {
    const double v = part[j * (int64_t)kRedBlocks + b];
    a = MODE == 2 ? fmax(a, v) : a + v;
  }
  a = block_reduce<MODE == 2>(a);
  if (threadIdx.x == 0) out[j] = MODE == 1 ? sqrt(a) : a;
}

__global__ __launch_bounds__(AFL_NT) void k_maxpy(int64_t n, const double* __restrict__ V, int64_t ldv, int m,
                                                  const double* __restrict__ h, double sign, double* __restrict__ w) {
  for (int64_t i = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AFL_NT) {
    double a = 0.0;
    for (int j = 0; j < m; ++j) a += h[j] * V[j * ldv + i];
    w[i] += sign * a;
  }
}

int64_t red_work_size(int m) { return (int64_t)(m < 1 ? 1 : m) * kRedBlocks; }

static int red_blocks(int64_t n) {
  int64_t b = (n + AFL_NT - 1) / AFL_NT;
  return (int)(b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks);
}

void vec_copy(hipStream_t st, int64_t n, const double* x, double* y) {
  k_copy<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, x, y);
}
void vec_axpby(hipStream_t st, int64_t n, double a, const double* x, double b, double* y) {
  k_axpby<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, a, x, b, y);
}
void vec_sub_mul(hipStream_t st, int64_t n, const double* b, const double* y, const double* d, double* out) {
  k_sub_mul<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, b, y, d, out);
}
void vec_sub(hipStream_t st, int64_t n, const double* a, const double* b, double* out) {
  k_sub<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, a, b, out);
}
void vec_acc_multi(hipStream_t st, const AccList& L) {
  if (L.count <= 0) return;
  int64_t n = 0;
  for (int p = 0; p < L.count; ++p) n = L.n[p] > n ? L.n[p] : n;
  k_acc_multi<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(L);
}
void vec_mul(hipStream_t st, int64_t n, const double* x, const double* d, double* out) {
  k_mul<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, x, d, out);
}
void vec_addmul(hipStream_t st, int64_t n, const double* d, const double* z, double* x) {
  k_addmul<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, d, z, x);
}
void vec_scale_by_inv(hipStream_t st, int64_t n, const double* s, double* x) {
  k_scale_by_inv<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, s, x);
}
void vec_fill(hipStream_t st, int64_t n, double v, double* x) {
  k_fill<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, v, x);
}
void vec_mdot(hipStream_t st, int64_t n, const double* V, int64_t ldv, int m, const double* w, double* work, double* h) {
  const int nb = red_blocks(n);
  k_mdot_partial<<<dim3(nb, m), dim3(AFL_NT), 0, st>>>(n, V, ldv, w, work);
  k_finish<0><<<dim3(m), dim3(AFL_NT), 0, st>>>(work, nb, h);
}
void vec_maxpy(hipStream_t st, int64_t n, const double* V, int64_t ldv, int m, const double* h, double sign, double* w) {
  k_maxpy<<<dim3(grid_1d(n)), dim3(AFL_NT), 0, st>>>(n, V, ldv, m, h, sign, w);
}
void vec_norm2(hipStream_t st, int64_t n, const double* x, double* work, double* out) {
  const int nb = red_blocks(n);
  k_norm_partial<false, false><<<dim3(nb), dim3(AFL_NT), 0, st>>>(n, x, nullptr, work);
  k_finish<1><<<dim3(1), dim3(AFL_NT), 0, st>>>(work, nb, out);
}
void vec_norm2_f32(hipStream_t st, int64_t n, const float* x, double* work, double* out) {
  const int nb = red_blocks(n);
  k_norm_partial_f32<<<dim3(nb), dim3(AFL_NT), 0, st>>>(n, x, work);
  k_finish<1><<<dim3(1), dim3(AFL_NT), 0, st>>>(work, nb, out);
}
void vec_linf(hipStream_t st, int64_t n, const double* x, double* work, double* out) {
  const int nb = red_blocks(n);
  k_norm_partial<true, false><<<dim3(nb), dim3(AFL_NT), 0, st>>>(n, x, nullptr, work);
  k_finish<2><<<dim3(1), dim3(AFL_NT), 0, st>>>(work, nb, out);
}
void vec_linf_diff(hipStream_t st, int64_t n, const double* x, const double* y, double* work, double* out) {
  const int nb = red_blocks(n);
  k_norm_partial<true, true><<<dim3(nb), dim3(AFL_NT), 0, st>>>(n, x, y, work);
  k_finish<2><<<dim3(1), dim3(AFL_NT), 0, st>>>(work, nb, out);
}

}  // namespace afl
