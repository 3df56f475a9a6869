// Measurement probe (not part of the operator library): FP64 throughput of the 3D degree-3
// sum-factorisation contraction on the vector ALUs and on the matrix cores of gfx950.
//
// The contraction is one sweep of a cell evaluation at k = 3 on Q = 7 Gauss points: the 7x4 value
// table S and the 7x4 derivative table D applied to the 4 nodal values of a line, 14 outputs per line
// (56 FMAs). Lines come from many cells; a wavefront takes 64 lines at a time.
//  - VALU: lane = line, 56 FMAs from the lane's 4 inputs into 14 accumulators (v_fma_f64).
//  - MFMA: v_mfma_f64_16x16x4f64 with A = [S; D] (14 of 16 rows used, 4 columns) and B = the 4 nodal
//    values of 16 lines, so four instructions cover 64 lines (4096 FMAs, 3584 of them useful), each
//    into its own accumulator.
// Each repetition accumulates the contraction again with inputs that drift by 1e-6, so no repetition
// can be folded into another; the kernels differ only in how the contraction is done. A layout check
// computes one MFMA on known operands so the host can read off the output placement.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__constant__ double cS[7][4];
__constant__ double cD[7][4];

// tab: two identical copies of [S; D] (2 x 56 doubles); each lane reads copy (lane & 1), so the
// tables sit in vector registers as in the operator kernels (a lane-uniform table would be held
// in scalar registers and moved per use)
__global__ __launch_bounds__(256) void k_probe_valu(int reps, const double* __restrict__ tab,
                                                    const double* __restrict__ in, double* __restrict__ out) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const double* tb = tab + (threadIdx.x & 1) * 56;
  double S[7][4], D[7][4];
#pragma unroll
  for (int q = 0; q < 7; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      S[q][i] = tb[q * 4 + i];
      D[q][i] = tb[28 + q * 4 + i];
    }
  double x[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = in[gid * 4 + i];
  double y[14];
#pragma unroll
  for (int q = 0; q < 14; ++q) y[q] = 0.0;
  for (int r = 0; r < reps; ++r) {
    // the line's 14 outputs accumulated (the inputs drift so that no repetition repeats another)
#pragma unroll
    for (int q = 0; q < 7; ++q)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        y[q] = fma(S[q][i], x[i], y[q]);
        y[7 + q] = fma(D[q][i], x[i], y[7 + q]);
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] += 1e-6;
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 14; ++q) acc += y[q];
  out[gid] = acc;
}

// lane l: A operand A[l % 16][l / 16], B operand B[l / 16][l % 16]; four independent accumulators
// (16 lines each) per wavefront
__global__ __launch_bounds__(256) void k_probe_mfma(int reps, const double* __restrict__ in, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int row = lane % 16, kk = lane / 16;
  const double a = row < 7 ? cS[row][kk] : (row < 14 ? cD[row - 7][kk] : 0.0);
  double b[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) b[g] = in[(wave * 64 + g * 16 + row) * 4 + kk];
  v4d d[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) d[g] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < reps; ++r) {
#pragma unroll
    for (int g = 0; g < 4; ++g) d[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[g], d[g], 0, 0, 0);
#pragma unroll
    for (int g = 0; g < 4; ++g) b[g] += 1e-6;
  }
  double acc = 0.0;
#pragma unroll
  for (int g = 0; g < 4; ++g) acc += d[g][0] + d[g][1] + d[g][2] + d[g][3];
  out[wave * 64 + lane] = acc;
}

// one MFMA on A[i][k] = a_in[i * 4 + k], B[k][j] = b_in[k * 16 + j]; lane l writes its 4 outputs
__global__ void k_probe_layout(const double* __restrict__ a_in, const double* __restrict__ b_in, double* __restrict__ d_out) {
  const int lane = threadIdx.x & 63;
  const double a = a_in[(lane % 16) * 4 + lane / 16];
  const double b = b_in[(lane / 16) * 16 + lane % 16];
  const v4d z = {0.0, 0.0, 0.0, 0.0};
  const v4d d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, z, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) d_out[lane * 4 + j] = d[j];
}

}  // namespace

extern "C" {

// tables S, D: 7x4 row-major each
int afl_probe_set_tables(const double* S, const double* D) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(cS), S, 28 * sizeof(double)) != hipSuccess) return 1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(cD), D, 28 * sizeof(double)) != hipSuccess) return 1;
  return 0;
}

// mode 0: VALU, 1: MFMA. lines = waves * 64; tab: 112 doubles (two copies of [S; D]), in: lines * 4
// doubles, out: lines doubles (device).
// Returns the kernel time in ms (HIP events on the null stream) in *ms.
int afl_probe_contraction(int mode, int reps, int waves, const double* tab, const double* in, double* out, float* ms) {
  if (waves < 4 || waves % 4 || reps < 1) return 2;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const dim3 grid(waves / 4), block(256);
  (void)hipEventRecord(e0, 0);
  if (mode == 0) k_probe_valu<<<grid, block, 0, 0>>>(reps, tab, in, out);
  else k_probe_mfma<<<grid, block, 0, 0>>>(reps, in, out);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  (void)hipEventElapsedTime(ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return hipGetLastError() == hipSuccess ? 0 : 3;
}

// a_in: 16x4, b_in: 4x16, d_out: 64 lanes x 4 (device pointers)
int afl_probe_mfma_layout(const double* a_in, const double* b_in, double* d_out) {
  k_probe_layout<<<dim3(1), dim3(64), 0, 0>>>(a_in, b_in, d_out);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

}  // extern "C"
