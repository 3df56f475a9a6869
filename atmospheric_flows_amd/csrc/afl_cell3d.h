// Wave-per-cell building blocks of the 3D FP64 operator kernels (gfx950), degree k <= 3 on Q Gauss
// points per direction (the operators' 2k + 1).
//
// One cell per 64-lane wavefront, so every exchange inside a cell is ordered by the wavefront alone
// (no block barrier after the kernel's first table load):
//  - lane = node for global traffic: a field's 64 cell values are one coalesced 512-byte load, and a
//    result is one coalesced store. All of a kernel's own-cell fields are loaded at its start, with
//    every load in flight together, and staged in LDS; the neighbours' face nodes (6 (k+1)^2 values
//    per field, the cell's own face nodes at walls) are loaded one state ahead of their use.
//  - sum-factorised sweeps are spread over all lanes: lane = (qa = lane % Q, group = lane / Q), each
//    group taking one line of the sweep per round, with the lane's table row S[qa][.] in registers.
//  - lane l < Q^2 owns the vertical quadrature line (qx, qy) = (l % Q, l / Q) for the pointwise work
//    (the last sweep of an evaluation and the first of an integration stay in its registers), and
//    the face point (qa, qb) = (l % Q, l / Q) of every face.
//  - face traces and face integrals are sum-factorised across the lanes as well: 12 face sides
//    (own and neighbour side of 6 faces) x (k+1) x Q first-sweep values, then Q^2 points per side.
// Fields staged in LDS are node-lexicographic (x fastest), like the global layout.
#pragma once

#include "afl_common.h"

namespace afl {
namespace w3 {

template <int K, int Q = 2 * K + 1>
struct Sz {
  static constexpr int N1 = K + 1, NQ = Q, NPC = N1 * N1 * N1, LT = Q * Q, NFN = N1 * N1;
  static constexpr int GQ = 64 / Q;             // lane groups of Q lanes
  static constexpr int NF2 = 2 * NFN;           // face nodes of the two faces normal to one direction
  static_assert(NPC <= 64 && LT <= 64 && NF2 <= 64, "one cell per wavefront");
  // work-area sizes in values per wavefront
  static constexpr int EV = N1 * N1 * Q + N1 * Q * Q;              // evaluation of one field
  static constexpr int TR = 4 * N1 * Q;                            // traces of one field on one direction's faces
  static constexpr int INV = N1 * Q * Q + N1 * N1 * Q;             // values-only integration, one field
  static constexpr int ING = 3 * N1 * Q * Q + 2 * N1 * N1 * Q;     // value + gradient integration, one field
  static constexpr int INGZ = 3 * N1 * Q * Q;                      // integrate_from_z (second pass in place)
  static constexpr int FI = 2 * LT + 2 * N1 * Q + 2 * NFN;         // one direction's face integrals, one component
  static constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
};

__device__ __attribute__((always_inline)) inline void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The block's copy of the 1D tables S[q][i], D[q][i] (q < Q, i <= K) for lane-dependent indices.
template <int K, int Q>
struct Tabs {
  double S[Q * (K + 1)];
  double D[Q * (K + 1)];
};

// Loads the tables into LDS; every thread of the block calls it (contains the block's only barrier).
template <int K, int Q>
__device__ __attribute__((always_inline)) inline void load_tabs(const Tab& T, Tabs<K, Q>& tb) {
  constexpr int N1 = K + 1;
  for (int t = threadIdx.x; t < Q * N1; t += blockDim.x) {
    tb.S[t] = T.S[t / N1][t % N1];
    tb.D[t] = T.D[t / N1][t % N1];
  }
  __syncthreads();
}

// The wavefront's cell and the lane's roles in it.
template <int K, int Q = 2 * K + 1>
struct Cell {
  using Z = Sz<K, Q>;
  static constexpr int N1 = Z::N1, NPC = Z::NPC, LT = Z::LT, NFN = Z::NFN, GQ = Z::GQ, NF2 = Z::NF2;
  int lane, qa, qb;      // lane % Q, lane / Q (line / face point (qa, qb) for lane < LT)
  bool line;             // lane < Q^2
  bool grp_ok;           // lane belongs to one of the GQ full groups of Q lanes
  int64_t cell;
  int ic[3];
  bool wall[6];          // face f = 2 dir + side is a wall
  int64_t noff[3];       // per direction: global offset of the lane's neighbour face node (lane < NF2)
  double Sa[N1];         // S[qa][.]
  __device__ __attribute__((always_inline)) Cell(const Geo& g, const Tabs<K, Q>& tb, int64_t c) {
    lane = threadIdx.x & 63;
    qa = lane % Q;
    qb = lane / Q;
    line = lane < LT;
    grp_ok = qb < GQ;
    cell = c;
    const int ci = (int)c;
    ic[0] = ci % g.nc[0];
    ic[1] = (ci / g.nc[0]) % g.nc[1];
    ic[2] = ci / (g.nc[0] * g.nc[1]);
    const int side = lane < NF2 ? lane / NFN : 0, t = lane % NFN;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        const int n = g.nc[dir];
        const int v = ic[dir] + (sd ? 1 : -1);
        wall[2 * dir + sd] = (v < 0 || v >= n) && !(dir < 2 && g.per[dir]);
      }
      // the neighbour across face (dir, side) of the lane's face node, and its matching node on the
      // neighbour's opposite face; at a wall the cell's own face node (scalar selects only: a runtime
      // index into a private array would place the array in scratch memory)
      int j[3] = {ic[0], ic[1], ic[2]};
      const int n = g.nc[dir];
      int v = ic[dir] + (side ? 1 : -1);
      const bool w = (v < 0 || v >= n) && !(dir < 2 && g.per[dir]);
      v = v < 0 ? v + n : (v >= n ? v - n : v);
      j[dir] = v;
      const int64_t nc = (int64_t)j[0] + (int64_t)g.nc[0] * (j[1] + (int64_t)g.nc[1] * j[2]);
      noff[dir] = !w ? nc * NPC + face_node<3, K>(dir, 1 - side, t) : c * NPC + face_node<3, K>(dir, side, t);
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) Sa[i] = tb.S[qa * N1 + i];
  }
  __device__ __attribute__((always_inline)) bool is_wall(int f) const {
    bool r = wall[0];
#pragma unroll
    for (int i = 1; i < 6; ++i) r = f == i ? wall[i] : r;
    return r;
  }
  // the cell's value of field f at node `lane` (lanes >= NPC: 0)
  __device__ __attribute__((always_inline)) double own(const double* __restrict__ f) const {
    return lane < NPC ? f[cell * NPC + lane] : 0.0;
  }
  // the lane's neighbour face node of field f on the faces normal to direction dir
  __device__ __attribute__((always_inline)) double nbr(const double* __restrict__ f, int dir) const {
    const int64_t o = dir == 0 ? noff[0] : (dir == 1 ? noff[1] : noff[2]);
    return lane < NF2 ? f[o] : 0.0;
  }
  // staged neighbour face nodes of one direction: nbf[side * NFN + t]
  __device__ __attribute__((always_inline)) void put_nbr(double* nbf, double v) const {
    if (lane < NF2) nbf[lane] = v;
  }
  __device__ __attribute__((always_inline)) void store(double* __restrict__ f, double v) const {
    if (lane < NPC) f[cell * NPC + lane] = v;
  }
};

// Outward normals n[side] and JxW of the two faces normal to direction dir at the lane's face point
// (lane < Q^2), the values face_geo gives. The x and y faces have constant normals; the z faces lie
// over the lane's own vertical line, so their hill slope is the line's (no second hill evaluation).
template <int K>
__device__ __attribute__((always_inline)) inline void dir_geo(const Geo& g, const Tab& T, const int* ic,
                                                               const LineGeo<3, K>& lg, int dir, int lane,
                                                               double (&n)[2][3], double (&jxw)[2]) {
  if (dir < 2) {
#pragma unroll
    for (int s = 0; s < 2; ++s) face_geo<3, K>(g, T, ic, 2 * dir + s, lane, n[s], jxw[s]);
  } else {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const double zeta = (ic[2] + s) * g.hc[2];
      const double fz = 1.0 - zeta * lg.iH;
      const double vx = -lg.dhx * fz, vy = -lg.dhy * fz;
      const double nv = sqrt(1.0 + vx * vx + vy * vy);
      const double sg = s ? 1.0 : -1.0;
      n[s][0] = sg * vx / nv;
      n[s][1] = sg * vy / nv;
      n[s][2] = sg / nv;
      jxw[s] = lg.hx * lg.hy * nv * lg.wxy;
    }
  }
}

// Values at the quadrature points of the lane's vertical line of NF fields staged in LDS
// (field a at nod + a * fs). wk: NF * Sz::EV values.
template <int K, int Q, int NF>
__device__ __attribute__((always_inline)) inline void eval(const Cell<K, Q>& c, const Tab& T, const double* nod, int fs,
                                                            double* wk, double (&v)[NF][Q]) {
  using Z = Sz<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ;
  double* t1 = wk;                       // [NF][N1*N1 rows (iz, iy)][Q qx]
  double* t2 = wk + NF * N1 * N1 * Q;    // [NF][N1 iz][Q qy][Q qx]
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(N1 * N1, GQ); ++r) {
      const int row = c.qb + GQ * r;
      if (row < N1 * N1)
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += c.Sa[i] * nod[a * fs + row * N1 + i];
          t1[(a * N1 * N1 + row) * Q + c.qa] = acc;
        }
    }
  }
  wsync();
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(N1 * Q, GQ); ++r) {
      const int pr = c.qb + GQ * r;  // (iz, qx)
      if (pr < N1 * Q) {
        const int iz = pr / Q, qx = pr % Q;
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += c.Sa[i] * t1[(a * N1 * N1 + iz * N1 + i) * Q + qx];
          t2[((a * N1 + iz) * Q + c.qa) * Q + qx] = acc;
        }
      }
    }
  }
  wsync();
  if (c.line) {
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      double z[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) z[i] = t2[((a * N1 + i) * Q + c.qb) * Q + c.qa];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) acc += T.S[q][i] * z[i];
        v[a][q] = acc;
      }
    }
  }
  wsync();
}

// The x and y sweeps of eval only: zl[a][i] holds field a after them at (qx, qy) = the lane's line and
// vertical node i, so that the value at vertical quadrature point q is line_value(T, zl[a], q). Keeps
// (k+1) instead of Q values per field live in the pointwise work.
template <int K, int Q, int NF>
__device__ __attribute__((always_inline)) inline void eval_lines(const Cell<K, Q>& c, const double* nod, int fs,
                                                                  double* wk, double (&zl)[NF][K + 1]) {
  using Z = Sz<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ;
  double* t1 = wk;
  double* t2 = wk + NF * N1 * N1 * Q;
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(N1 * N1, GQ); ++r) {
      const int row = c.qb + GQ * r;
      if (row < N1 * N1)
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += c.Sa[i] * nod[a * fs + row * N1 + i];
          t1[(a * N1 * N1 + row) * Q + c.qa] = acc;
        }
    }
  }
  wsync();
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(N1 * Q, GQ); ++r) {
      const int pr = c.qb + GQ * r;
      if (pr < N1 * Q) {
        const int iz = pr / Q, qx = pr % Q;
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < N1; ++i) acc += c.Sa[i] * t1[(a * N1 * N1 + iz * N1 + i) * Q + qx];
          t2[((a * N1 + iz) * Q + c.qa) * Q + qx] = acc;
        }
      }
    }
  }
  wsync();
#pragma unroll
  for (int a = 0; a < NF; ++a)
#pragma unroll
    for (int i = 0; i < N1; ++i) zl[a][i] = c.line ? t2[((a * N1 + i) * Q + c.qb) * Q + c.qa] : 0.0;
  wsync();
}

template <int K>
__device__ __attribute__((always_inline)) inline double line_value(const Tab& T, const double (&z)[K + 1], int q) {
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i <= K; ++i) acc += T.S[q][i] * z[i];
  return acc;
}

// Traces of NF fields at the lane's face point of the two faces normal to direction dir:
// tr[a][side][0] own side, [1] neighbour side (at walls the staged neighbour nodes are the cell's
// own, so the two agree). Own nodes of field a at nod + a * fs, its staged neighbour face nodes at
// nbf + a * fn (Sz::NF2 each). wk: NF * Sz::TR.
template <int K, int Q, int NF>
__device__ __attribute__((always_inline)) inline void traces(const Cell<K, Q>& c, const Tabs<K, Q>& tb, int dir,
                                                              const double* nod, int fs, const double* nbf, int fn,
                                                              double* wk, double (&tr)[NF][2][2]) {
  using Z = Sz<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ, NFN = Z::NFN;
  double* r1 = wk;  // [NF][4 sides (side, own|nbr)][N1 tb][Q qa]
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(4 * N1, GQ); ++r) {
      const int pr = c.qb + GQ * r;  // (side s = 2 fside + sd, tb)
      if (pr < 4 * N1) {
        const int sdx = pr / N1, tbi = pr % N1, fside = sdx >> 1, sd = sdx & 1;
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int ta = 0; ta < N1; ++ta) {
            const int t = ta + N1 * tbi;
            const double val = sd == 0 ? nod[a * fs + face_node<3, K>(dir, fside, t)] : nbf[a * fn + fside * NFN + t];
            acc += c.Sa[ta] * val;
          }
          r1[((a * 4 + sdx) * N1 + tbi) * Q + c.qa] = acc;
        }
      }
    }
  }
  wsync();
  if (c.line) {
    double Sb[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) Sb[i] = tb.S[c.qb * N1 + i];
#pragma unroll
    for (int a = 0; a < NF; ++a)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) acc += Sb[i] * r1[((a * 4 + s) * N1 + i) * Q + c.qa];
        tr[a][s >> 1][s & 1] = acc;
      }
  }
  wsync();
}

// Integration of NC line fields against the test functions: out[c] = the cell's result at node
// `lane`. V: values, G: reference-coordinate gradient contributions, all multiplied by JxW.
// wk: NC * (HAS_G ? Sz::ING : Sz::INV).
template <int K, int Q, int NC, bool HAS_V, bool HAS_G>
__device__ __attribute__((always_inline)) inline void integrate(const Cell<K, Q>& c, const Tab& T, const Tabs<K, Q>& tb,
                                                                 double* wk, const double (&V)[NC][Q],
                                                                 const double (&G)[NC][3][Q], double (&out)[NC]) {
  using Z = Sz<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ, NPC = Z::NPC, A1 = N1 * Q * Q, A2 = N1 * N1 * Q;
  double* ta = wk;                               // [NC][N1 iz][Q qy][Q qx]
  double* tx = ta + NC * A1;
  double* ty = tx + (HAS_G ? NC * A1 : 0);
  double* ua = ty + (HAS_G ? NC * A1 : 0);       // [NC][N1 iz][N1 iy][Q qx]
  double* ux = ua + NC * A2;
  if (c.line) {
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        double s0 = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (HAS_V) s0 += T.S[q][i] * V[a][q];
          if (HAS_G) {
            s0 += T.D[q][i] * G[a][2][q];
            sx += T.S[q][i] * G[a][0][q];
            sy += T.S[q][i] * G[a][1][q];
          }
        }
        const int o = a * A1 + (i * Q + c.qb) * Q + c.qa;
        ta[o] = s0;
        if (HAS_G) { tx[o] = sx; ty[o] = sy; }
      }
  }
  wsync();
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(N1 * N1, GQ); ++r) {
      const int pr = c.qb + GQ * r;  // (iz, iy)
      if (pr < N1 * N1) {
        const int iz = pr / N1, iy = pr % N1;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
          double s0 = 0.0, sx = 0.0;
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const int o = a * A1 + (iz * Q + q) * Q + c.qa;
            s0 += tb.S[q * N1 + iy] * ta[o];
            if (HAS_G) {
              s0 += tb.D[q * N1 + iy] * ty[o];
              sx += tb.S[q * N1 + iy] * tx[o];
            }
          }
          const int o = a * A2 + pr * Q + c.qa;
          ua[o] = s0;
          if (HAS_G) ux[o] = sx;
        }
      }
    }
  }
  wsync();
  if (c.lane < NPC) {
    const int ix = c.lane % N1, row = c.lane / N1;
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      double s0 = 0.0;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        s0 += tb.S[q * N1 + ix] * ua[a * A2 + row * Q + q];
        if (HAS_G) s0 += tb.D[q * N1 + ix] * ux[a * A2 + row * Q + q];
      }
      out[a] = s0;
    }
  } else {
#pragma unroll
    for (int a = 0; a < NC; ++a) out[a] = 0.0;
  }
  wsync();
}

// The y and x sweeps of integrate for one component, from the z sweep already applied on the lane's
// line: sa[i] = sum_q S[q][i] V[q] + D[q][i] Gz[q], sx[i] = sum_q S[q][i] Gx[q], sy[i] likewise
// (vertical node i). Lets a kernel accumulate the z-swept sums state by state instead of holding
// every quadrature-point contribution. Returns the result at node `lane`. The y-swept values stay in
// registers across a wavefront barrier and then overwrite the line arrays. wk: Sz::INGZ.
template <int K, int Q>
__device__ __attribute__((always_inline)) inline double integrate_from_z(const Cell<K, Q>& c, const Tabs<K, Q>& tb,
                                                                          double* wk, const double (&sa)[K + 1],
                                                                          const double (&sx)[K + 1],
                                                                          const double (&sy)[K + 1]) {
  using Z = Sz<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ, NPC = Z::NPC, A1 = N1 * Q * Q, A2 = N1 * N1 * Q;
  constexpr int NR = Z::cdiv(N1 * N1, GQ);  // rounds of the y sweep
  static_assert(2 * A2 <= 3 * A1, "y-swept arrays fit the line arrays' area");
  double* ta = wk;
  double* tx = ta + A1;
  double* ty = tx + A1;
  double* ua = wk;  // written once every line array has been read
  double* ux = ua + A2;
  if (c.line) {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const int o = (i * Q + c.qb) * Q + c.qa;
      ta[o] = sa[i];
      tx[o] = sx[i];
      ty[o] = sy[i];
    }
  }
  wsync();
  double sv[NR][2];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int pr = c.qb + GQ * r;  // (iz, iy)
    sv[r][0] = sv[r][1] = 0.0;
    if (c.grp_ok && pr < N1 * N1) {
      const int iz = pr / N1, iy = pr % N1;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int o = (iz * Q + q) * Q + c.qa;
        sv[r][0] += tb.S[q * N1 + iy] * ta[o] + tb.D[q * N1 + iy] * ty[o];
        sv[r][1] += tb.S[q * N1 + iy] * tx[o];
      }
    }
  }
  wsync();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int pr = c.qb + GQ * r;
    if (c.grp_ok && pr < N1 * N1) {
      ua[pr * Q + c.qa] = sv[r][0];
      ux[pr * Q + c.qa] = sv[r][1];
    }
  }
  wsync();
  double out = 0.0;
  if (c.lane < NPC) {
    const int ix = c.lane % N1, row = c.lane / N1;
#pragma unroll
    for (int q = 0; q < Q; ++q) out += tb.S[q * N1 + ix] * ua[row * Q + q] + tb.D[q * N1 + ix] * ux[row * Q + q];
  }
  wsync();
  return out;
}

// Face integrals of NC components over the two faces normal to direction dir: F[a][side] is the
// flux at the lane's face point, multiplied by JxW and oriented for this cell; add[a] += the integral
// against the test function of node `lane` (nodes on those faces). wk: NC * Sz::FI.
template <int K, int Q, int NC>
__device__ __attribute__((always_inline)) inline void face_integrate(const Cell<K, Q>& c, const Tabs<K, Q>& tb, int dir,
                                                                      double* wk, const double (&F)[NC][2],
                                                                      double (&add)[NC]) {
  using Z = Sz<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ, LT = Z::LT, NFN = Z::NFN, NPC = Z::NPC;
  double* Fq = wk;                        // [NC][2][LT]
  double* rr = Fq + NC * 2 * LT;          // [NC][2][N1 tb][Q qa]
  double* cc = rr + NC * 2 * N1 * Q;      // [NC][2][NFN]
  if (c.line)
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
      for (int f = 0; f < 2; ++f) Fq[(a * 2 + f) * LT + c.lane] = F[a][f];
  wsync();
  if (c.grp_ok) {
#pragma unroll 1
    for (int r = 0; r < Z::cdiv(2 * N1, GQ); ++r) {
      const int pr = c.qb + GQ * r;  // (side, tb)
      if (pr < 2 * N1) {
        const int f = pr / N1, tbi = pr % N1;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
          double acc = 0.0;
#pragma unroll
          for (int q = 0; q < Q; ++q) acc += tb.S[q * N1 + tbi] * Fq[(a * 2 + f) * LT + q * Q + c.qa];
          rr[((a * 2 + f) * N1 + tbi) * Q + c.qa] = acc;
        }
      }
    }
  }
  wsync();
  if (c.lane < 2 * NFN) {
    const int f = c.lane / NFN, t = c.lane % NFN, tai = t % N1, tbi = t / N1;
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < Q; ++q) acc += tb.S[q * N1 + tai] * rr[((a * 2 + f) * N1 + tbi) * Q + q];
      cc[(a * 2 + f) * NFN + t] = acc;
    }
  }
  wsync();
  if (c.lane < NPC) {
    const int ix = c.lane % N1, iy = (c.lane / N1) % N1, iz = c.lane / (N1 * N1);
    const int inorm = dir == 0 ? ix : (dir == 1 ? iy : iz);
    const int t = dir == 0 ? iy + N1 * iz : (dir == 1 ? ix + N1 * iz : ix + N1 * iy);
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      if (inorm == 0) add[a] += cc[(a * 2 + 0) * NFN + t];
      if (inorm == K) add[a] += cc[(a * 2 + 1) * NFN + t];
    }
  }
  wsync();
}

// In-place inverse of a small SPD matrix (Gauss-Jordan without pivoting, fully unrolled).
template <int N>
__device__ __attribute__((always_inline)) inline void inv_spd_w3(double (&A)[N][N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double p = 1.0 / A[k][k];
#pragma unroll
    for (int j = 0; j < N; ++j) A[k][j] = j == k ? p : A[k][j] * p;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == k) continue;
      const double f = A[i][k];
#pragma unroll
      for (int j = 0; j < N; ++j) A[i][j] = j == k ? -f * p : A[i][j] - f * A[k][j];
    }
  }
}

template <int K, int Q, int NC>
struct SolveWk {
  using Z = Sz<K, Q>;
  static constexpr int MI = 3 * Z::N1 * Z::N1;                               // the 1D inverses (persist)
  static constexpr int SETUP = Z::LT * Q + Z::LT + 4 * Q + Z::LT;             // co per line, line sums, marginals,
                                                                               // partial vertical marginals
  static constexpr int EVI = Z::EV > Z::INV ? Z::EV : Z::INV;
  static constexpr int RUN = 2 * NC * Z::NPC + NC * EVI;                      // staged vectors + sweeps
  static constexpr int SIZE = MI + (SETUP > RUN ? SETUP : RUN);
};

// Cell-local solve of A x = b, A = S^T diag(co) S for NC components sharing co (co: the lane's line;
// b, x: the value at node `lane`). The method of cell_cheb in afl_ops.hip: Chebyshev iteration (Saad,
// Iterative Methods, Alg. 12.1) preconditioned with the separable P = S^T diag(a (x) b (x) c / t^2) S
// built from the marginal sums a, b, c of co (t their total), whose inverse is the Kronecker product
// of three (k+1)x(k+1) inverses; the spectrum of P^{-1} A lies in [min co / c~, max co / c~], which
// fixes the interval and the iteration count for the A-norm tolerance tol before the first step.
// wk: SolveWk::SIZE values.
template <int K, int Q, int NC>
__device__ __attribute__((always_inline)) inline void cheb_solve(const Cell<K, Q>& c, const Tab& T, const Tabs<K, Q>& tb,
                                                                  const double (&co)[Q], double* wk, double tol, int maxit,
                                                                  const double (&b)[NC], double (&x)[NC]) {
  using Z = Sz<K, Q>;
  using SW = SolveWk<K, Q, NC>;
  constexpr int N1 = Z::N1, LT = Z::LT, NPC = Z::NPC;
  double* Mi = wk;                  // [3 axes][N1][N1]
  double* sco = wk + SW::MI;        // [LT][Q]
  double* sL = sco + LT * Q;        // [LT] line sums
  double* smg = sL + LT;            // [4][Q]: x, y, z marginals, 1 / z marginal
  double* part = smg + 4 * Q;       // [LT]: vertical marginal of each row of lines
  double lsum = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) lsum += co[q];
  if (c.line) {
#pragma unroll
    for (int q = 0; q < Q; ++q) sco[c.lane * Q + q] = co[q];
    sL[c.lane] = lsum;
  }
  wsync();
  // vertical marginal in two levels: lane qy Q + q first sums the Q lines of row qy at height q
  if (c.line) {
    const int q = c.lane % Q, qy = c.lane / Q;
    double p = 0.0;
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) p += sco[(qy * Q + qx) * Q + q];
    part[c.lane] = p;
  }
  wsync();
  if (c.lane < Q) {
    double cz = 0.0, ax = 0.0, by = 0.0;
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) cz += part[qy * Q + c.lane];
#pragma unroll
    for (int t = 0; t < Q; ++t) { ax += sL[t * Q + c.lane]; by += sL[c.lane * Q + t]; }
    smg[c.lane] = ax;           // x marginal: lines (qx = lane, qy)
    smg[Q + c.lane] = by;       // y marginal: lines (qx, qy = lane)
    smg[2 * Q + c.lane] = cz;   // z marginal
    smg[3 * Q + c.lane] = 1.0 / cz;
  }
  wsync();
  double total = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) total += smg[q];
  const double itot = 1.0 / total;
  double lmin = 1e300, lmax = 0.0;
  if (c.line) {
    const double ihor = total / (smg[c.qa] * smg[Q + c.qb] * itot);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double r = co[q] * ihor * smg[3 * Q + q];
      lmin = fmin(lmin, r);
      lmax = fmax(lmax, r);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lmin = fmin(lmin, __shfl_xor(lmin, off, 64));
    lmax = fmax(lmax, __shfl_xor(lmax, off, 64));
  }
  if (c.lane < 3) {
    double Am[N1][N1];
    const double sc = c.lane == 0 ? 1.0 : itot;
    const double* wq = smg + c.lane * Q;
#pragma unroll
    for (int i = 0; i < N1; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) a += wq[q] * T.S[q][i] * T.S[q][j];
        Am[i][j] = sc * a;
        Am[j][i] = sc * a;
      }
    inv_spd_w3<N1>(Am);
#pragma unroll
    for (int i = 0; i < N1; ++i)
#pragma unroll
      for (int j = 0; j < N1; ++j) Mi[(c.lane * N1 + i) * N1 + j] = Am[i][j];
  }
  wsync();
  const double alpha = lmin, beta = lmax;
  const double theta = 0.5 * (alpha + beta);
  const double delta = fmax(0.5 * (beta - alpha), 1e-14 * theta);
  const double sigma1 = theta / delta;
  int nit = 1;
  {
    const float km1 = (float)((beta - alpha) / alpha);
    const float sk = sqrtf(1.0f + km1);
    const float sg = km1 / ((sk + 1.0f) * (sk + 1.0f));
    if (sg > 1e-30f) nit = (int)ceilf(logf(0.5f * (float)tol) / logf(sg));
    nit = nit < 1 ? 1 : (nit > maxit ? maxit : nit);
  }
  double* dl = wk + SW::MI;           // [NC][NPC] staged vector
  double* dt = dl + NC * NPC;         // [NC][NPC] second staging
  double* ew = dt + NC * NPC;         // sweeps
  const int ix = c.lane % N1, iy = (c.lane / N1) % N1, iz = c.lane / (N1 * N1);
  const bool node = c.lane < NPC;
  // z = P^{-1} r, node-distributed (three one-axis passes through LDS)
  auto precond = [&](const double (&r)[NC], double (&z)[NC]) {
    if (node)
#pragma unroll
      for (int a = 0; a < NC; ++a) dl[a * NPC + c.lane] = r[a];
    wsync();
    if (node)
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        double s0 = 0.0;
#pragma unroll
        for (int j = 0; j < N1; ++j) s0 += Mi[(0 * N1 + ix) * N1 + j] * dl[a * NPC + (iz * N1 + iy) * N1 + j];
        dt[a * NPC + c.lane] = s0;
      }
    wsync();
    if (node)
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        double s0 = 0.0;
#pragma unroll
        for (int j = 0; j < N1; ++j) s0 += Mi[(1 * N1 + iy) * N1 + j] * dt[a * NPC + (iz * N1 + j) * N1 + ix];
        dl[a * NPC + c.lane] = s0;
      }
    wsync();
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      double s0 = 0.0;
      if (node)
#pragma unroll
        for (int j = 0; j < N1; ++j) s0 += Mi[(2 * N1 + iz) * N1 + j] * dl[a * NPC + (j * N1 + iy) * N1 + ix];
      z[a] = s0;
    }
    wsync();
  };
  // y = A v, node-distributed
  auto apply = [&](const double (&v)[NC], double (&y)[NC]) {
    if (node)
#pragma unroll
      for (int a = 0; a < NC; ++a) dl[a * NPC + c.lane] = v[a];
    wsync();
    double vq[NC][Q];
    eval<K, Q, NC>(c, T, dl, NPC, ew, vq);
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
      for (int q = 0; q < Q; ++q) vq[a][q] *= co[q];
    const double G0[NC][3][Q] = {};
    integrate<K, Q, NC, true, false>(c, T, tb, ew, vq, G0, y);
  };
  double r[NC], z[NC], d[NC];
#pragma unroll
  for (int a = 0; a < NC; ++a) { x[a] = 0.0; r[a] = node ? b[a] : 0.0; }
  precond(r, z);
  double rho = 1.0 / sigma1;
#pragma unroll
  for (int a = 0; a < NC; ++a) d[a] = z[a] / theta;
  for (int k = 0; k < nit; ++k) {
#pragma unroll
    for (int a = 0; a < NC; ++a) x[a] += d[a];
    if (k == nit - 1) break;
    double ad[NC];
    apply(d, ad);
#pragma unroll
    for (int a = 0; a < NC; ++a) r[a] -= ad[a];
    precond(r, z);
    const double rn = 1.0 / (2.0 * sigma1 - rho);
#pragma unroll
    for (int a = 0; a < NC; ++a) d[a] = rn * rho * d[a] + (2.0 * rn / delta) * z[a];
    rho = rn;
  }
}

}  // namespace w3
}  // namespace afl
