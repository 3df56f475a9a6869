// Shared device building blocks of the IMEX-DG kernels (gfx950).
//
// Work decomposition ("thread per quadrature line"): a cell is handled by LT = NQ^(dim-1)
// threads; thread l owns the quadrature line parallel to the vertical axis through the
// horizontal quadrature index l (2D: l = qx; 3D: l = qx + NQ*qy). Sum-factorised
// interpolation runs along x (and y) with one LDS transpose per axis and along z inside
// registers, so every thread ends up with the NQ values of its own line. Integration runs
// the same sweeps transposed and leaves node row l (2D: z-row l; 3D: (iy, iz) = l) of the
// cell's result in the registers of thread l, which then writes N1 contiguous doubles.
//
// Faces are computed cell-centrically from both sides (each cell evaluates the fluxes on
// its own 2*dim faces with its own outward normal), so every output entry is written by
// exactly one thread, without atomics. The fluxes of the reference are antisymmetric in
// (+,-,n), so the two sides agree to the last bit except for the periodic wrap geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define AFL_NT 256
// wavefronts per block of the cell kernels whose cells are wave-local
#ifndef AFL_CELL_WAVES
#define AFL_CELL_WAVES 4
#endif

namespace afl {

constexpr double kGamma = 1.4;  // EquationData::Cp_Cv

template <typename R>
struct TabT {
  R S[9][5];   // S[q][i]  = phi_i(x_q), Gauss point q of QGauss(2k+1) on [0,1]
  R D[9][5];   // D[q][i]  = phi_i'(x_q)
  R W[9];      // Gauss weights on [0,1]
  R QP[9];     // Gauss points on [0,1]
  R MI[5][5];  // inverse of the reference 1D mass matrix (cell-local preconditioner)
  R ND[5];     // 1D support points
};
using Tab = TabT<double>;

struct Geo {
  int nc[3];        // cells along x, y, z (ny = 1 in 2D)
  int ncells;
  int per[2];       // periodic x, y
  double lo[2];     // x_min, y_min
  double hc[3];     // hx, hy, hz
  double H, h0, inv_a2;
};

// Runtime-indexed reads of the basis tables (uniform values in device memory, read through the
// kernel's table pointer). A select chain over them is turned by the compiler into one per-lane
// indexed load, a memory round trip on the critical path; a blend of the uniform entries costs
// CNT multiply-adds instead. The FP32 multigrid level kernels blend; the FP64 kernels keep the
// select chain, since there the blend's extra FP64 work and registers cost more than it saves
// (AFL_PICK_BLEND_F64=1 at 512x100, k = 4: E1 stage 3 102 -> 97 us, E2 140 -> 162 us,
// E3 171 -> 197 us).
#ifndef AFL_PICK_BLEND_F64
#define AFL_PICK_BLEND_F64 0
#endif
template <int CNT, int N, typename R>
__device__ __attribute__((always_inline)) inline R pick(const R (&a)[N], int i) {
  if constexpr (sizeof(R) == 4 || AFL_PICK_BLEND_F64) {
    R v = R(0);
#pragma unroll
    for (int j = 0; j < CNT; ++j) v += (i == j ? R(1) : R(0)) * a[j];
    return v;
  } else {
    R v = a[0];
#pragma unroll
    for (int j = 1; j < CNT; ++j) v = (i == j) ? a[j] : v;
    return v;
  }
}
// S[r][col] for a runtime row r < ROWS and a compile-time column
template <int ROWS, typename R>
__device__ __attribute__((always_inline)) inline R pick_row(const R (&S)[9][5], int r, int col) {
  if constexpr (sizeof(R) == 4 || AFL_PICK_BLEND_F64) {
    R v = R(0);
#pragma unroll
    for (int j = 0; j < ROWS; ++j) v += (r == j ? R(1) : R(0)) * S[j][col];
    return v;
  } else {
    R v = S[0][col];
#pragma unroll
    for (int j = 1; j < ROWS; ++j) v = (r == j) ? S[j][col] : v;
    return v;
  }
}
// S[row][c] for a compile-time row and a runtime column c < COLS
template <int COLS, typename R>
__device__ __attribute__((always_inline)) inline R pick_col(const R (&S)[9][5], int row, int c) {
  if constexpr (sizeof(R) == 4 || AFL_PICK_BLEND_F64) {
    R v = R(0);
#pragma unroll
    for (int j = 0; j < COLS; ++j) v += (c == j ? R(1) : R(0)) * S[row][j];
    return v;
  } else {
    R v = S[row][0];
#pragma unroll
    for (int j = 1; j < COLS; ++j) v = (c == j) ? S[row][j] : v;
    return v;
  }
}

template <int DIM, int K, int Q = 2 * K + 1>
struct Cfg {
  static constexpr int N1 = K + 1;
  static constexpr int NQ = Q;  // quadrature points per direction (2k+1: the operator's rule)
  static constexpr int LT = (DIM == 2) ? NQ : NQ * NQ;        // threads per cell
  // Cells that fit in one 64-wide wavefront are packed whole into waves (CPW per wave, the few
  // spare lanes idle): all exchanges inside a cell are then wave-local and need no block barrier.
  static constexpr bool WAVE_LOCAL = LT <= 64;
  static constexpr int CPW = WAVE_LOCAL ? 64 / LT : 0;         // cells per wavefront
  static constexpr int WPB = AFL_CELL_WAVES;                   // wavefronts per block (wave-local)
  static constexpr int NT = WAVE_LOCAL ? 64 * WPB : AFL_NT;    // threads per block
  static constexpr int CPB = WAVE_LOCAL ? WPB * CPW : AFL_NT / LT;  // cells per block
  static constexpr int NPC = (DIM == 2) ? N1 * N1 : N1 * N1 * N1;
  static constexpr int NFQ = (DIM == 2) ? NQ : NQ * NQ;         // points per face
  static constexpr int ROWS = (DIM == 2) ? N1 : N1 * N1;        // node rows per cell
  // LDS doubles per cell for one evaluation (t1 + t2)
  static constexpr int EV1 = (DIM == 2) ? N1 * NQ : N1 * N1 * NQ;
  static constexpr int EV2 = (DIM == 2) ? 0 : N1 * NQ * NQ;
  // LDS doubles per cell for one integration (2D: a,b [N1][NQ]; 3D: a,x,y [N1][NQ][NQ] + a2,x2 [N1][N1][NQ])
  static constexpr int IN1 = (DIM == 2) ? 2 * N1 * NQ : 3 * N1 * NQ * NQ;
  static constexpr int IN2 = (DIM == 2) ? 0 : 2 * N1 * N1 * NQ;
};

// Thread position inside the block.
template <int DIM, int K, int Q = 2 * K + 1>
struct Slot {
  int cb;        // cell slot in block
  int l;         // line index in cell
  int64_t cell;  // global cell index
  bool act;      // slot maps to an existing cell
  int ic[3];     // cell coordinates
  __device__ Slot(const Geo& g) : Slot(g, blockIdx.x) {}
  // the slot in cell block vb (a kernel that strides over cell blocks)
  __device__ Slot(const Geo& g, int64_t vb) {
    using C = Cfg<DIM, K, Q>;
    if (C::WAVE_LOCAL) {
      const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, lc = lane / C::LT;
      l = lane % C::LT;
      cb = lc < C::CPW ? wave * C::CPW + lc : C::CPB;  // spare lanes: no cell
    } else {
      cb = threadIdx.x / C::LT;
      l = threadIdx.x % C::LT;
    }
    cell = vb * C::CPB + cb;
    act = (cb < C::CPB) && (cell < g.ncells);
    const int c = act ? (int)cell : 0;
    ic[0] = c % g.nc[0];
    if (DIM == 2) { ic[1] = c / g.nc[0]; ic[2] = 0; }
    else { ic[1] = (c / g.nc[0]) % g.nc[1]; ic[2] = c / (g.nc[0] * g.nc[1]); }
  }
};

// Barrier between the phases of a cell's computation (LDS written by some of the cell's threads,
// read by others). Wave-local cells need only ordering inside the wavefront; otherwise the block.
template <int DIM, int K, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void cell_sync() {
  if constexpr (Cfg<DIM, K, Q>::WAVE_LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <int DIM>
__device__ __attribute__((always_inline)) inline int vaxis() { return DIM - 1; }

// Hill h(x,y) = h0 / (1 + r^2/a^2) and its horizontal derivatives (arithmetic in R).
template <typename R>
__device__ __attribute__((always_inline)) inline void hill(const Geo& g, R x, R y, R& h, R& hx, R& hy) {
  const R s = R(1) + (x * x + y * y) * R(g.inv_a2);
  const R is = R(1) / s;
  h = R(g.h0) * is;
  const R f = R(-2) * R(g.h0) * R(g.inv_a2) * is * is;
  hx = f * x;
  hy = f * y;
}

// Per-line geometry: everything that is constant along the thread's vertical line.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
struct LineGeo {
  R c;      // hz (1 - h/H) = dz/dzeta
  R dhx, dhy;
  R wxy;    // horizontal quadrature weight product
  R hx, hy, hz, iH, zeta0;
  __device__ __attribute__((always_inline)) void init(const Geo& g, const TabT<R>& T, const int* ic, int l) {
    using C = Cfg<DIM, K, Q>;
    hx = R(g.hc[0]);
    hy = DIM == 3 ? R(g.hc[1]) : R(1);
    hz = R(g.hc[DIM - 1]);
    iH = R(1.0 / g.H);
    zeta0 = R(ic[DIM - 1] * g.hc[DIM - 1]);
    R x, y = R(0);
    if (DIM == 2) {
      x = R(g.lo[0] + (ic[0] + (double)pick<C::NQ>(T.QP, l)) * g.hc[0]);
      wxy = pick<C::NQ>(T.W, l);
    } else {
      const int qx = l % C::NQ, qy = l / C::NQ;
      x = R(g.lo[0] + (ic[0] + (double)pick<C::NQ>(T.QP, qx)) * g.hc[0]);
      y = R(g.lo[1] + (ic[1] + (double)pick<C::NQ>(T.QP, qy)) * g.hc[1]);
      wxy = pick<C::NQ>(T.W, qx) * pick<C::NQ>(T.W, qy);
    }
    R h;
    hill(g, x, y, h, dhx, dhy);
    c = hz * (R(1) - h * iH);
  }
  // JxW at vertical quadrature index qz
  __device__ __attribute__((always_inline)) R jxw(const Geo& g, const TabT<R>& T, int qz) const {
    return (DIM == 2 ? hx : hx * hy) * c * wxy * T.W[qz];
  }
  // gradient contribution (physical G, without JxW) -> reference coordinates * JxW
  __device__ __attribute__((always_inline)) void gref(const Geo& g, const TabT<R>& T, const int* ic, int qz, const R* G, R* Rf) const {
    const R zeta = zeta0 + T.QP[qz] * hz;
    const R fz = R(1) - zeta * iH;
    const R w = wxy * T.W[qz];
    if (DIM == 2) {
      const R dx = dhx * fz * hx;
      Rf[0] = G[0] * c * w;
      Rf[1] = (-dx * G[0] + hx * G[1]) * w;
    } else {
      const R dx = dhx * fz * hx, dy = dhy * fz * hy;
      Rf[0] = G[0] * hy * c * w;
      Rf[1] = G[1] * hx * c * w;
      Rf[2] = (-dx * hy * G[0] - dy * hx * G[1] + hx * hy * G[DIM - 1]) * w;
    }
  }
  // transpose of gref: reference gradient gr -> physical gradient * JxW (P = Mq^T gr)
  __device__ __attribute__((always_inline)) void gref_t(const Geo& g, const TabT<R>& T, int qz, const R* gr, R* P) const {
    const R zeta = zeta0 + T.QP[qz] * hz;
    const R fz = R(1) - zeta * iH;
    const R w = wxy * T.W[qz];
    if (DIM == 2) {
      const R dx = dhx * fz * hx;
      P[0] = (gr[0] * c - dx * gr[1]) * w;
      P[1] = hx * gr[1] * w;
    } else {
      const R dx = dhx * fz * hx, dy = dhy * fz * hy;
      P[0] = (gr[0] * c - dx * gr[2]) * hy * w;
      P[1] = (gr[1] * c - dy * gr[2]) * hx * w;
      P[2] = hx * hy * gr[2] * w;
    }
  }
};

// Geometry of face f = 2*dir + side of cell ic at face point fq (2D: fq = q; 3D: fq = qa + NQ qb,
// qa/qb along the two tangential axes in increasing order). Returns outward normal and JxW.
template <int DIM, int K, int Q = 2 * K + 1, typename R = double>
__device__ __attribute__((always_inline)) inline void face_geo(const Geo& g, const TabT<R>& T, const int* ic, int f, int fq, R* n, R& jxw) {
  using C = Cfg<DIM, K, Q>;
  const int dir = f >> 1, side = f & 1;
  const R sg = side ? R(1) : R(-1);
  const R iH = R(1.0 / g.H);
  if (DIM == 2) {
    if (dir == 0) {
      const R x = R(g.lo[0] + (ic[0] + side) * g.hc[0]);
      R h, hx, hy;
      hill(g, x, R(0), h, hx, hy);
      n[0] = sg; n[1] = R(0);
      jxw = R(g.hc[1]) * (R(1) - h * iH) * pick<C::NQ>(T.W, fq);
    } else {
      const R x = R(g.lo[0] + (ic[0] + (double)pick<C::NQ>(T.QP, fq)) * g.hc[0]);
      R h, hx, hy;
      hill(g, x, R(0), h, hx, hy);
      const R zeta = R((ic[1] + side) * g.hc[1]);
      const R vx = -hx * (R(1) - zeta * iH);
      const R nv = sqrt(R(1) + vx * vx);
      n[0] = sg * vx / nv; n[1] = sg / nv;
      jxw = R(g.hc[0]) * nv * pick<C::NQ>(T.W, fq);
    }
  } else {
    const int qa = fq % C::NQ, qb = fq / C::NQ;
    const R w = pick<C::NQ>(T.W, qa) * pick<C::NQ>(T.W, qb);
    const double xa = (double)pick<C::NQ>(T.QP, qa), xb = (double)pick<C::NQ>(T.QP, qb);
    double x, y, zeta;
    if (dir == 0) {
      x = g.lo[0] + (ic[0] + side) * g.hc[0];
      y = g.lo[1] + (ic[1] + xa) * g.hc[1];
      zeta = (ic[2] + xb) * g.hc[2];
    } else if (dir == 1) {
      x = g.lo[0] + (ic[0] + xa) * g.hc[0];
      y = g.lo[1] + (ic[1] + side) * g.hc[1];
      zeta = (ic[2] + xb) * g.hc[2];
    } else {
      x = g.lo[0] + (ic[0] + xa) * g.hc[0];
      y = g.lo[1] + (ic[1] + xb) * g.hc[1];
      zeta = (ic[2] + side) * g.hc[2];
    }
    R h, hx, hy;
    hill(g, R(x), R(y), h, hx, hy);
    const R c = R(g.hc[2]) * (R(1) - h * iH);
    if (dir == 0) { n[0] = sg; n[1] = R(0); n[2] = R(0); jxw = R(g.hc[1]) * c * w; }
    else if (dir == 1) { n[0] = R(0); n[1] = sg; n[2] = R(0); jxw = R(g.hc[0]) * c * w; }
    else {
      const R fz = R(1) - R(zeta) * iH;
      const R vx = -hx * fz, vy = -hy * fz;
      const R nv = sqrt(R(1) + vx * vx + vy * vy);
      n[0] = sg * vx / nv; n[1] = sg * vy / nv; n[2] = sg / nv;
      jxw = R(g.hc[0]) * R(g.hc[1]) * nv * w;
    }
  }
}

// Neighbour across face f (-1 for a wall).
template <int DIM>
__device__ __attribute__((always_inline)) inline int64_t neighbour(const Geo& g, const int* ic, int f) {
  const int dir = f >> 1, side = f & 1;
  const int n = g.nc[dir];
  int j[3] = {ic[0], ic[1], ic[2]};
  int v = j[dir] + (side ? 1 : -1);
  if (v < 0 || v >= n) {
    if (dir < DIM - 1 && g.per[dir]) v = (v + n) % n;
    else return -1;
  }
  j[dir] = v;
  return DIM == 2 ? (int64_t)j[0] + (int64_t)g.nc[0] * j[1]
                  : (int64_t)j[0] + (int64_t)g.nc[0] * (j[1] + (int64_t)g.nc[1] * j[2]);
}

// Node index of face node t (2D: t along the face; 3D: t = ta + N1 tb) on face (dir, side).
template <int DIM, int K>
__device__ __attribute__((always_inline)) inline int face_node(int dir, int side, int t) {
  constexpr int N1 = K + 1;
  const int s = side ? K : 0;
  if (DIM == 2) return dir == 0 ? t * N1 + s : s * N1 + t;
  const int ta = t % N1, tb = t / N1;
  if (dir == 0) return s + N1 * (ta + N1 * tb);
  if (dir == 1) return ta + N1 * (s + N1 * tb);
  return ta + N1 * (tb + N1 * s);
}

// ---------------------------------------------------------------------------------------
// The thread's node row of a scalar field (global dofs): row l of its cell for the row owners
// (l < ROWS), zeros elsewhere. Loading rows ahead of the barriers of a kernel lets their memory
// latency overlap earlier work.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void load_row(const R* __restrict__ f, const Slot<DIM, K, Q>& s,
                                                                R (&d)[K + 1]) {
  using C = Cfg<DIM, K, Q>;
  const R* u = f + (s.act ? s.cell : 0) * C::NPC;
#pragma unroll
  for (int i = 0; i < C::N1; ++i) d[i] = R(0);
  if (s.act && s.l < C::ROWS)
#pragma unroll
    for (int i = 0; i < C::N1; ++i) d[i] = u[s.l * C::N1 + i];
}

// Evaluate a scalar field at the quadrature points of the thread's line from the node rows that
// load_row gave the row owners. All threads of the block must call it (it contains barriers).
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_line_row(const R (&d)[K + 1], const Slot<DIM, K, Q>& s, R* t1,
                                                                     R* t2, const TabT<R>& T, R (&v)[Q]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ;
  if (DIM == 2) {
    if (s.act && s.l < N1) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * d[i];
        t1[s.l * NQ + q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
      R c[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) c[i] = t1[i * NQ + s.l];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
        v[q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    if (s.act && s.l < N1 * N1) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * d[i];
        t1[s.l * NQ + q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
      R c[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) c[i] = t1[(iz * N1 + i) * NQ + qx];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
        t2[(iz * NQ + q) * NQ + qx] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
      R c[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) c[i] = t2[i * NQ * NQ + s.l];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
        v[q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Evaluate a scalar field (global dofs) at the quadrature points of the thread's line.
// All threads of the block must call it (it contains barriers).
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_line(const R* __restrict__ f, const Slot<DIM, K, Q>& s, R* t1, R* t2,
                                 const TabT<R>& T, R (&v)[Q]) {
  using C = Cfg<DIM, K, Q>;
  const R* u = f + (s.act ? s.cell : 0) * C::NPC;
  R d[K + 1];  // read by the row owners only
  if (s.act && s.l < C::ROWS)
#pragma unroll
    for (int i = 0; i < C::N1; ++i) d[i] = u[s.l * C::N1 + i];
  eval_line_row<DIM, K, R>(d, s, t1, t2, T, v);
}

// Reference-coordinate gradient of a scalar field at the quadrature points of the thread's line:
// gr[b][q] = d/dxi_b (D along axis b, S along the others). Work area per cell: 2D 2*N1*NQ,
// 3D 2*N1*N1*NQ + 3*N1*NQ*NQ values from t1 on (no t2). Contains barriers.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_line_grad(const R* __restrict__ f, const Slot<DIM, K, Q>& s, R* t1,
                                      const TabT<R>& T, R (&gr)[DIM][Q]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ;
  const R* u = f + (s.act ? s.cell : 0) * C::NPC;
  constexpr int X1 = (DIM == 2) ? N1 * NQ : N1 * N1 * NQ;  // one x-sweep array
  R* tS = t1;
  R* tD = t1 + X1;
  if (s.act && s.l < C::ROWS) {
    R d[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) d[i] = u[s.l * N1 + i];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      R a = 0.0, b = 0.0;
#pragma unroll
      for (int i = 0; i < N1; ++i) { a += T.S[q][i] * d[i]; b += T.D[q][i] * d[i]; }
      tS[s.l * NQ + q] = a;
      tD[s.l * NQ + q] = b;
    }
  }
  cell_sync<DIM, K, Q>();
  if (DIM == 2) {
    if (s.act) {
      R cs[N1], cd[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) { cs[i] = tS[i * NQ + s.l]; cd[i] = tD[i * NQ + s.l]; }
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0, b = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) { a += T.S[q][i] * cd[i]; b += T.D[q][i] * cs[i]; }
        gr[0][q] = a;
        gr[DIM - 1][q] = b;
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    constexpr int Y1 = N1 * NQ * NQ;
    R* uSS = t1 + 2 * X1;  // S_x S_y: for the z derivative
    R* uSD = uSS + Y1;     // S_x D_y: y derivative
    R* uDS = uSD + Y1;     // D_x S_y: x derivative
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
      R cs[N1], cd[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) { cs[i] = tS[(iz * N1 + i) * NQ + qx]; cd[i] = tD[(iz * N1 + i) * NQ + qx]; }
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) { a += T.S[q][i] * cs[i]; b += T.D[q][i] * cs[i]; c += T.S[q][i] * cd[i]; }
        const int o = (iz * NQ + q) * NQ + qx;
        uSS[o] = a; uSD[o] = b; uDS[o] = c;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
      R a0[N1], a1[N1], a2[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) { a0[i] = uDS[i * NQ * NQ + s.l]; a1[i] = uSD[i * NQ * NQ + s.l]; a2[i] = uSS[i * NQ * NQ + s.l]; }
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) { x += T.S[q][i] * a0[i]; y += T.S[q][i] * a1[i]; z += T.D[q][i] * a2[i]; }
        gr[0][q] = x; gr[1][q] = y; gr[DIM - 1][q] = z;
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Neighbour cells of the thread's cell, face order f = 2*dir + side (-1 for walls).
template <int DIM>
struct Nbrs {
  int64_t c[2 * DIM];
  __device__ Nbrs(const Geo& g, const int* ic) {
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) c[f] = neighbour<DIM>(g, ic, f);
  }
};

// Neighbour-side face nodes of node row l for every face of the cell: nbv[fc] is the field at the
// neighbour's node that matches the row's own face node on face fc. All faces are loaded before
// any is used, and stored unconditionally (a wall reads the own cell; its neighbour-side trace is
// replaced by the own trace after the interpolation), so the loads are in flight together
// instead of each waiting, inside its own branch, before its LDS store.
template <int DIM, int K, typename R>
__device__ __attribute__((always_inline)) inline void nbr_face_nodes(const R* __restrict__ f, const Nbrs<DIM>& nb,
                                                                      int64_t cell, int l, R (&nbv)[2 * DIM]) {
  constexpr int N1 = K + 1, NPC = DIM == 2 ? N1 * N1 : N1 * N1 * N1;
  int64_t c[2 * DIM];
#pragma unroll
  for (int fc = 0; fc < 2 * DIM; ++fc) c[fc] = (nb.c[fc] >= 0 ? nb.c[fc] : cell) * NPC;
  // x faces: the neighbour's opposite node of the same row
  nbv[0] = f[c[0] + l * N1 + K];
  nbv[1] = f[c[1] + l * N1];
  if (DIM == 2) {
    // z faces: the neighbour's row K (below) / 0 (above)
    nbv[2] = f[c[2] + K * N1 + l];
    nbv[3] = f[c[3] + l];
  } else {
    const int iy = l % N1, iz = l / N1;
    nbv[2] = f[c[2] + (iz * N1 + K) * N1 + iy];
    nbv[3] = f[c[3] + (iz * N1) * N1 + iy];
    nbv[4] = f[c[4] + (K * N1 + iz) * N1 + iy];
    nbv[5] = f[c[5] + iz * N1 + iy];
  }
}

// LDS doubles per cell for one field's face nodes: [face][own|neighbour][face node]
template <int DIM, int K>
struct FaceLds {
  static constexpr int N1 = K + 1;
  static constexpr int NFN = DIM == 2 ? N1 : N1 * N1;
  static constexpr int FN = 2 * 2 * DIM * NFN;
};

// eval_line plus the traces of the same field at face point l of every face of the cell:
// tr[f][0] own side, tr[f][1] neighbour side (at walls the neighbour entry is the own trace).
// Row owners copy their own face nodes and fetch the neighbour's matching face nodes (2 scalar
// loads per face direction) into fn during the x sweep; face-point threads interpolate after the
// first barrier, so no barrier is added. fn: FaceLds::FN doubles per cell.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_line_faces(const R* __restrict__ f, const Slot<DIM, K, Q>& s, const Nbrs<DIM>& nb,
                                       R* t1, R* t2, R* fn, const TabT<R>& T, R (&v)[Q],
                                       R (&tr)[2 * DIM][2]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ, NPC = C::NPC, NFN = FaceLds<DIM, K>::NFN;
  const R* u = f + (s.act ? s.cell : 0) * NPC;
  auto FN = [&](int face, int side, int t) -> R& { return fn[(face * 2 + side) * NFN + t]; };
  if (s.act && s.l < C::ROWS) {
    R nbv[2 * DIM];
    nbr_face_nodes<DIM, K, R>(f, nb, s.cell, s.l, nbv);
    R d[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) d[i] = u[s.l * N1 + i];
    // the neighbour sides are stored first, right after their loads: every load of the staging
    // is in flight before the first wait
    FN(0, 1, s.l) = nb.c[0] >= 0 ? nbv[0] : d[0];
    FN(1, 1, s.l) = nb.c[1] >= 0 ? nbv[1] : d[K];
    if (DIM == 2) {
      FN(2, 1, s.l) = nbv[2];
      FN(3, 1, s.l) = nbv[3];
    } else {
      const int iy = s.l % N1, iz = s.l / N1;
      FN(2, 1, iy + N1 * iz) = nbv[2];
      FN(3, 1, iy + N1 * iz) = nbv[3];
      FN(4, 1, iy + N1 * iz) = nbv[4];
      FN(5, 1, iy + N1 * iz) = nbv[5];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      R a = 0.0;
#pragma unroll
      for (int i = 0; i < N1; ++i) a += T.S[q][i] * d[i];
      t1[s.l * NQ + q] = a;
    }
    // x faces: node (0|K, row l); neighbour's opposite node of the same row
    FN(0, 0, s.l) = d[0];
    FN(1, 0, s.l) = d[K];
    if (DIM == 2) {
      // z faces: own rows 0 / K; neighbour rows K (below) / 0 (above), one node per row owner
      if (s.l == 0) for (int i = 0; i < N1; ++i) FN(2, 0, i) = d[i];
      if (s.l == K) for (int i = 0; i < N1; ++i) FN(3, 0, i) = d[i];
    } else {
      const int iy = s.l % N1, iz = s.l / N1;
      // y faces: own nodes (ix, 0|K, iz) -> t = ix + N1 iz; neighbour node (ix = iy, K|0, iz)
      if (iy == 0) for (int i = 0; i < N1; ++i) FN(2, 0, i + N1 * iz) = d[i];
      if (iy == K) for (int i = 0; i < N1; ++i) FN(3, 0, i + N1 * iz) = d[i];
      // z faces: own nodes (ix, iy, 0|K) -> t = ix + N1 iy; neighbour node (ix = iy, iy = iz, K|0)
      if (iz == 0) for (int i = 0; i < N1; ++i) FN(4, 0, i + N1 * iy) = d[i];
      if (iz == K) for (int i = 0; i < N1; ++i) FN(5, 0, i + N1 * iy) = d[i];
    }
  }
  cell_sync<DIM, K, Q>();
  // traces at face point l (every thread of an active cell owns one face point per face)
  if (s.act) {
#pragma unroll
    for (int fc = 0; fc < 2 * DIM; ++fc)
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        R a = 0.0;
        if (DIM == 2) {
#pragma unroll
          for (int t = 0; t < N1; ++t) a += pick_row<NQ>(T.S, s.l, t) * FN(fc, sd, t);
        } else {
          const int qa = s.l % NQ, qb = s.l / NQ;
#pragma unroll
          for (int tb = 0; tb < N1; ++tb) {
            R r = 0.0;
#pragma unroll
            for (int ta = 0; ta < N1; ++ta) r += pick_row<NQ>(T.S, qa, ta) * FN(fc, sd, ta + N1 * tb);
            a += pick_row<NQ>(T.S, qb, tb) * r;
          }
        }
        tr[fc][sd] = a;
      }
#pragma unroll
    for (int fc = 0; fc < 2 * DIM; ++fc)
      if (nb.c[fc] < 0) tr[fc][1] = tr[fc][0];
  }
  if (DIM == 2) {
    if (s.act) {
      R c[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) c[i] = t1[i * NQ + s.l];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
        v[q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
      R c[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) c[i] = t1[(iz * N1 + i) * NQ + qx];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
        t2[(iz * NQ + q) * NQ + qx] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
      R c[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) c[i] = t2[i * NQ * NQ + s.l];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
        v[q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Traces only: the face-point values of eval_line_faces (tr[f][0] own side, tr[f][1] neighbour side,
// own trace at walls) without the cell interpolation. Row owners stage their row's face nodes and
// the neighbours' matching nodes in fn; one barrier; every thread interpolates at its face point.
// Ends with a barrier (fn is reused by the next call). fn: FaceLds::FN values per cell.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void trace_line_faces(const R* __restrict__ f, const Slot<DIM, K, Q>& s,
                                                                       const Nbrs<DIM>& nb, R* fn, const TabT<R>& T,
                                                                       R (&tr)[2 * DIM][2]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ, NPC = C::NPC, NFN = FaceLds<DIM, K>::NFN;
  const R* u = f + (s.act ? s.cell : 0) * NPC;
  auto FN = [&](int face, int side, int t) -> R& { return fn[(face * 2 + side) * NFN + t]; };
  if (s.act && s.l < C::ROWS) {
    R nbv[2 * DIM];
    nbr_face_nodes<DIM, K, R>(f, nb, s.cell, s.l, nbv);
    R d[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) d[i] = u[s.l * N1 + i];
    // the neighbour sides are stored first, right after their loads: every load of the staging
    // is in flight before the first wait
    FN(0, 1, s.l) = nb.c[0] >= 0 ? nbv[0] : d[0];
    FN(1, 1, s.l) = nb.c[1] >= 0 ? nbv[1] : d[K];
    if (DIM == 2) {
      FN(2, 1, s.l) = nbv[2];
      FN(3, 1, s.l) = nbv[3];
    } else {
      const int iy = s.l % N1, iz = s.l / N1;
      FN(2, 1, iy + N1 * iz) = nbv[2];
      FN(3, 1, iy + N1 * iz) = nbv[3];
      FN(4, 1, iy + N1 * iz) = nbv[4];
      FN(5, 1, iy + N1 * iz) = nbv[5];
    }
    FN(0, 0, s.l) = d[0];
    FN(1, 0, s.l) = d[K];
    if (DIM == 2) {
      if (s.l == 0) for (int i = 0; i < N1; ++i) FN(2, 0, i) = d[i];
      if (s.l == K) for (int i = 0; i < N1; ++i) FN(3, 0, i) = d[i];
    } else {
      const int iy = s.l % N1, iz = s.l / N1;
      if (iy == 0) for (int i = 0; i < N1; ++i) FN(2, 0, i + N1 * iz) = d[i];
      if (iy == K) for (int i = 0; i < N1; ++i) FN(3, 0, i + N1 * iz) = d[i];
      if (iz == 0) for (int i = 0; i < N1; ++i) FN(4, 0, i + N1 * iy) = d[i];
      if (iz == K) for (int i = 0; i < N1; ++i) FN(5, 0, i + N1 * iy) = d[i];
    }
  }
  cell_sync<DIM, K, Q>();
  if (s.act) {
#pragma unroll
    for (int fc = 0; fc < 2 * DIM; ++fc)
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        R a = 0.0;
        if (DIM == 2) {
#pragma unroll
          for (int t = 0; t < N1; ++t) a += pick_row<NQ>(T.S, s.l, t) * FN(fc, sd, t);
        } else {
          const int qa = s.l % NQ, qb = s.l / NQ;
#pragma unroll
          for (int tb = 0; tb < N1; ++tb) {
            R r = 0.0;
#pragma unroll
            for (int ta = 0; ta < N1; ++ta) r += pick_row<NQ>(T.S, qa, ta) * FN(fc, sd, ta + N1 * tb);
            a += pick_row<NQ>(T.S, qb, tb) * r;
          }
        }
        tr[fc][sd] = a;
      }
#pragma unroll
    for (int fc = 0; fc < 2 * DIM; ++fc)
      if (nb.c[fc] < 0) tr[fc][1] = tr[fc][0];
  }
  cell_sync<DIM, K, Q>();
}

// Batched eval_line_faces: NF scalar fields evaluated together (one barrier per sweep for all of
// them), the first NFF of them also at the face points. Per cell: t1 NF*EV1, t2 NF*EV2 and
// fn NFF*FaceLds::FN values. Row owners issue all NF row loads before the sweeps.
template <int DIM, int K, int NF, int NFF, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_lines_faces(const R* const (&f)[NF], const Slot<DIM, K, Q>& s, const Nbrs<DIM>& nb,
                                        R* t1, R* t2, R* fn, const TabT<R>& T, R (&v)[NF][Q],
                                        R (&tr)[NFF > 0 ? NFF : 1][2 * DIM][2]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ, NPC = C::NPC, NFN = FaceLds<DIM, K>::NFN, FNS = FaceLds<DIM, K>::FN;
  constexpr int E1 = C::EV1, E2 = C::EV2;
  const int64_t cell = s.act ? s.cell : 0;
  auto FN = [&](int a, int face, int side, int t) -> R& { return fn[a * FNS + (face * 2 + side) * NFN + t]; };
  if (s.act && s.l < C::ROWS) {
    R nbv[NFF > 0 ? NFF : 1][2 * DIM];
#pragma unroll
    for (int a = 0; a < NFF; ++a) nbr_face_nodes<DIM, K, R>(f[a], nb, cell, s.l, nbv[a]);
    R d[NF][N1];
#pragma unroll
    for (int a = 0; a < NF; ++a)
#pragma unroll
      for (int i = 0; i < N1; ++i) d[a][i] = f[a][cell * NPC + s.l * N1 + i];
    // the neighbour sides are stored first, right after their loads: every load of the staging
    // is in flight before the first wait
#pragma unroll
    for (int a = 0; a < NFF; ++a) {
      FN(a, 0, 1, s.l) = nb.c[0] >= 0 ? nbv[a][0] : d[a][0];
      FN(a, 1, 1, s.l) = nb.c[1] >= 0 ? nbv[a][1] : d[a][K];
      if (DIM == 2) {
        FN(a, 2, 1, s.l) = nbv[a][2];
        FN(a, 3, 1, s.l) = nbv[a][3];
      } else {
        const int iy = s.l % N1, iz = s.l / N1;
        FN(a, 2, 1, iy + N1 * iz) = nbv[a][2];
        FN(a, 3, 1, iy + N1 * iz) = nbv[a][3];
        FN(a, 4, 1, iy + N1 * iz) = nbv[a][4];
        FN(a, 5, 1, iy + N1 * iz) = nbv[a][5];
      }
    }
#pragma unroll
    for (int a = 0; a < NF; ++a)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R acc = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) acc += T.S[q][i] * d[a][i];
        t1[a * E1 + s.l * NQ + q] = acc;
      }
#pragma unroll
    for (int a = 0; a < NFF; ++a) {
      FN(a, 0, 0, s.l) = d[a][0];
      FN(a, 1, 0, s.l) = d[a][K];
      if (DIM == 2) {
        if (s.l == 0) for (int i = 0; i < N1; ++i) FN(a, 2, 0, i) = d[a][i];
        if (s.l == K) for (int i = 0; i < N1; ++i) FN(a, 3, 0, i) = d[a][i];
      } else {
        const int iy = s.l % N1, iz = s.l / N1;
        if (iy == 0) for (int i = 0; i < N1; ++i) FN(a, 2, 0, i + N1 * iz) = d[a][i];
        if (iy == K) for (int i = 0; i < N1; ++i) FN(a, 3, 0, i + N1 * iz) = d[a][i];
        if (iz == 0) for (int i = 0; i < N1; ++i) FN(a, 4, 0, i + N1 * iy) = d[a][i];
        if (iz == K) for (int i = 0; i < N1; ++i) FN(a, 5, 0, i + N1 * iy) = d[a][i];
      }
    }
  }
  cell_sync<DIM, K, Q>();
  if (s.act) {
    // the thread's rows of the interpolation table (face point l), read once
    const int qa = DIM == 2 ? s.l : s.l % NQ, qb = DIM == 2 ? 0 : s.l / NQ;
    R sa[N1], sb[N1];
#pragma unroll
    for (int t = 0; t < N1; ++t) { sa[t] = pick_row<NQ>(T.S, qa, t); sb[t] = pick_row<NQ>(T.S, qb, t); }
#pragma unroll
    for (int a = 0; a < NFF; ++a) {
#pragma unroll
      for (int fc = 0; fc < 2 * DIM; ++fc)
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
          R acc = 0.0;
          if (DIM == 2) {
#pragma unroll
            for (int t = 0; t < N1; ++t) acc += sa[t] * FN(a, fc, sd, t);
          } else {
#pragma unroll
            for (int tb = 0; tb < N1; ++tb) {
              R r = 0.0;
#pragma unroll
              for (int ta = 0; ta < N1; ++ta) r += sa[ta] * FN(a, fc, sd, ta + N1 * tb);
              acc += sb[tb] * r;
            }
          }
          tr[a][fc][sd] = acc;
        }
#pragma unroll
      for (int fc = 0; fc < 2 * DIM; ++fc)
        if (nb.c[fc] < 0) tr[a][fc][1] = tr[a][fc][0];
    }
  }
  if (DIM == 2) {
    if (s.act) {
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        R c[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) c[i] = t1[a * E1 + i * NQ + s.l];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += T.S[q][i] * c[i];
          v[a][q] = acc;
        }
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        R c[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) c[i] = t1[a * E1 + (iz * N1 + i) * NQ + qx];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += T.S[q][i] * c[i];
          t2[a * E2 + (iz * NQ + q) * NQ + qx] = acc;
        }
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        R c[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) c[i] = t2[a * E2 + i * NQ * NQ + s.l];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += T.S[q][i] * c[i];
          v[a][q] = acc;
        }
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Same as eval_line but the dofs come from an LDS copy of the cell (u = lds + cb*NPC).
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_line_lds(const R* u, const Slot<DIM, K, Q>& s, R* t1, R* t2,
                                     const TabT<R>& T, R (&v)[Q]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ;
  if (DIM == 2) {
    if (s.act && s.l < N1) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * u[s.l * N1 + i];
        t1[s.l * NQ + q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * t1[i * NQ + s.l];
        v[q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    if (s.act && s.l < N1 * N1) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * u[s.l * N1 + i];
        t1[s.l * NQ + q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * t1[(iz * N1 + i) * NQ + qx];
        t2[(iz * NQ + q) * NQ + qx] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        R a = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) a += T.S[q][i] * t2[i * NQ * NQ + s.l];
        v[q] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Batched evaluation of NF scalar fields whose cell dofs sit in LDS (field f at u + f*NPC).
// t1 needs NF*EV1 doubles, t2 NF*EV2 doubles per cell. Contains barriers.
template <int DIM, int K, int NF, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void eval_lines_lds(const R* u, const Slot<DIM, K, Q>& s, R* t1, R* t2, const TabT<R>& T,
                                      R (&v)[NF][Q]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ, NPC = C::NPC, E1 = C::EV1, E2 = C::EV2;
  if (DIM == 2) {
    if (s.act && s.l < N1) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R a = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) a += T.S[q][i] * u[f * NPC + s.l * N1 + i];
          t1[f * E1 + s.l * NQ + q] = a;
        }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        R c[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) c[i] = t1[f * E1 + i * NQ + s.l];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R a = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
          v[f][q] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    if (s.act && s.l < N1 * N1) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R a = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) a += T.S[q][i] * u[f * NPC + s.l * N1 + i];
          t1[f * E1 + s.l * NQ + q] = a;
        }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        R c[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) c[i] = t1[f * E1 + (iz * N1 + i) * NQ + qx];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R a = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
          t2[f * E2 + (iz * NQ + q) * NQ + qx] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        R c[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) c[i] = t2[f * E2 + i * NQ * NQ + s.l];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          R a = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) a += T.S[q][i] * c[i];
          v[f][q] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Batched value-only integration of NF line fields; row l of field f ends in out[f].
// i1 needs NF*I1V doubles, i2 NF*I2V doubles per cell (2D: I1V = N1*NQ, I2V = 0;
// 3D: I1V = N1*NQ*NQ, I2V = N1*N1*NQ). Contains barriers.
template <int DIM, int K, int Q = 2 * K + 1>
struct IntV {
  static constexpr int N1 = K + 1, NQ = Q;
  static constexpr int I1 = DIM == 2 ? N1 * NQ : N1 * NQ * NQ;
  static constexpr int I2 = DIM == 2 ? 0 : N1 * N1 * NQ;
};

template <int DIM, int K, int NF, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void integrate_lines_v(const Slot<DIM, K, Q>& s, R* i1, R* i2, const TabT<R>& T,
                                         const R (&V)[NF][Q], R (&out)[NF][K + 1]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ, A1 = IntV<DIM, K, Q>::I1, A2 = IntV<DIM, K, Q>::I2;
  if (DIM == 2) {
    if (s.act) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = 0.0;
#pragma unroll
          for (int q = 0; q < NQ; ++q) a += T.S[q][i] * V[f][q];
          i1[f * A1 + i * NQ + s.l] = a;
        }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < N1) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = 0.0;
#pragma unroll
          for (int q = 0; q < NQ; ++q) a += T.S[q][i] * i1[f * A1 + s.l * NQ + q];
          out[f][i] = a;
        }
    }
    cell_sync<DIM, K, Q>();
  } else {
    if (s.act) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = 0.0;
#pragma unroll
          for (int q = 0; q < NQ; ++q) a += T.S[q][i] * V[f][q];
          i1[f * A1 + i * NQ * NQ + s.l] = a;
        }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = 0.0;
#pragma unroll
          for (int q = 0; q < NQ; ++q) a += T.S[q][i] * i1[f * A1 + (iz * NQ + q) * NQ + qx];
          i2[f * A2 + (iz * N1 + i) * NQ + qx] = a;
        }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < N1 * N1) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = 0.0;
#pragma unroll
          for (int q = 0; q < NQ; ++q) a += T.S[q][i] * i2[f * A2 + s.l * NQ + q];
          out[f][i] = a;
        }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Integrate line data (values V and reference-gradient contributions G[d], all already
// multiplied by JxW) against the test functions. Row `l` (l < ROWS) of the result ends up
// in out[N1]. HAS_V / HAS_G select the terms. Contains barriers.
template <int DIM, int K, bool HAS_V, bool HAS_G, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void integrate_line(const Slot<DIM, K, Q>& s, R* i1, R* i2, const TabT<R>& T,
                                      const R (&V)[Q], const R (&G)[DIM][Q],
                                      R (&out)[K + 1]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ;
  if (DIM == 2) {
    R* ta = i1;
    R* tb = i1 + N1 * NQ;
    if (s.act) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0, b = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (HAS_V) a += T.S[q][i] * V[q];
          if (HAS_G) { a += T.D[q][i] * G[1][q]; b += T.S[q][i] * G[0][q]; }
        }
        ta[i * NQ + s.l] = a;
        if (HAS_G) tb[i * NQ + s.l] = b;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < N1) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          a += T.S[q][i] * ta[s.l * NQ + q];
          if (HAS_G) a += T.D[q][i] * tb[s.l * NQ + q];
        }
        out[i] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  } else {
    R* ta = i1;
    R* tx = i1 + N1 * NQ * NQ;
    R* ty = i1 + 2 * N1 * NQ * NQ;
    R* ua = i2;
    R* ux = i2 + N1 * N1 * NQ;
    if (s.act) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (HAS_V) a += T.S[q][i] * V[q];
          if (HAS_G) { a += T.D[q][i] * G[2][q]; bx += T.S[q][i] * G[0][q]; by += T.S[q][i] * G[1][q]; }
        }
        ta[i * NQ * NQ + s.l] = a;
        if (HAS_G) { tx[i * NQ * NQ + s.l] = bx; ty[i * NQ * NQ + s.l] = by; }
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0, bx = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int idx = (iz * NQ + q) * NQ + qx;
          a += T.S[q][i] * ta[idx];
          if (HAS_G) { a += T.D[q][i] * ty[idx]; bx += T.S[q][i] * tx[idx]; }
        }
        ua[(iz * N1 + i) * NQ + qx] = a;
        if (HAS_G) ux[(iz * N1 + i) * NQ + qx] = bx;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < N1 * N1) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          a += T.S[q][i] * ua[s.l * NQ + q];
          if (HAS_G) a += T.D[q][i] * ux[s.l * NQ + q];
        }
        out[i] = a;
      }
    }
    cell_sync<DIM, K, Q>();
  }
}

// Add face-flux integrals to node row l. Fq points to the LDS block [2*DIM][NFQ] of this cell
// (values already multiplied by JxW_f and oriented for this cell). Only rows on faces change.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void add_faces(const Slot<DIM, K, Q>& s, const R* Fq, const TabT<R>& T, R (&out)[K + 1],
                                 unsigned face_mask) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NQ = C::NQ, NFQ = C::NFQ;
  if (!(s.act && s.l < C::ROWS)) return;
  if (DIM == 2) {
    const int iz = s.l;
    // x faces: nodes (0, iz) and (K, iz)
    R sz[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) sz[q] = pick_col<N1>(T.S, q, iz);
    if (face_mask & 1u) { R a = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) a += sz[q] * Fq[0 * NFQ + q]; out[0] += a; }
    if (face_mask & 2u) { R a = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) a += sz[q] * Fq[1 * NFQ + q]; out[K] += a; }
    const int fz = (iz == 0) ? 2 : ((iz == K) ? 3 : -1);
    if (fz >= 0 && (face_mask & (1u << fz))) {
#pragma unroll
      for (int i = 0; i < N1; ++i) { R a = 0.0; for (int q = 0; q < NQ; ++q) a += T.S[q][i] * Fq[fz * NFQ + q]; out[i] += a; }
    }
  } else {
    const int iy = s.l % N1, iz = s.l / N1;
    R sy[NQ], sz[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) { sy[q] = pick_col<N1>(T.S, q, iy); sz[q] = pick_col<N1>(T.S, q, iz); }
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (!(face_mask & (1u << side))) continue;
      const R* F = Fq + side * NFQ;
      R a = 0.0;
#pragma unroll
      for (int qb = 0; qb < NQ; ++qb) {
        R r = 0.0;
#pragma unroll
        for (int qa = 0; qa < NQ; ++qa) r += sy[qa] * F[qb * NQ + qa];
        a += sz[qb] * r;
      }
      out[side ? K : 0] += a;
    }
    const int fy = (iy == 0) ? 2 : ((iy == K) ? 3 : -1);
    if (fy >= 0 && (face_mask & (1u << fy))) {
      const R* F = Fq + fy * NFQ;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0;
        for (int qb = 0; qb < NQ; ++qb) {
          R r = 0.0;
          for (int qa = 0; qa < NQ; ++qa) r += T.S[qa][i] * F[qb * NQ + qa];
          a += sz[qb] * r;
        }
        out[i] += a;
      }
    }
    const int fz = (iz == 0) ? 4 : ((iz == K) ? 5 : -1);
    if (fz >= 0 && (face_mask & (1u << fz))) {
      const R* F = Fq + fz * NFQ;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = 0.0;
        for (int qb = 0; qb < NQ; ++qb) {
          R r = 0.0;
          for (int qa = 0; qa < NQ; ++qa) r += T.S[qa][i] * F[qb * NQ + qa];
          a += sy[qb] * r;
        }
        out[i] += a;
      }
    }
  }
}

// Write node row l of a cell.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void store_row(R* __restrict__ dst, const Slot<DIM, K, Q>& s, const R (&out)[K + 1]) {
  using C = Cfg<DIM, K, Q>;
  if (s.act && s.l < C::ROWS) {
    R* o = dst + s.cell * C::NPC + s.l * C::N1;
#pragma unroll
    for (int i = 0; i < C::N1; ++i) o[i] = out[i];
  }
}

}  // namespace afl
