// 3D FP64 operator kernels, degree k <= 3 on the operators' 2k + 1 Gauss points, built on the
// wave-per-cell blocks of afl_cell3d.h: one cell per wavefront, every own-cell field loaded at the
// kernel's start and staged in LDS, neighbour face nodes loaded one state ahead of their use, faces
// first (traces, fluxes, face integrals) and the volume terms second. The weak forms are those of the
// dimension-generic kernels in afl_ops.hip (each cites its assembler in euler_operator.h); only the
// work decomposition differs.
#include "afl_cell3d.h"
#include "afl_internal.h"

namespace afl {
namespace w3 {

constexpr int kWPB = 2;  // wavefronts (cells) per block

// Common prologue: block tables, the wave's cell (inactive waves leave after the table barrier).
#define W3_PROLOGUE(PER)                                   \
  constexpr int Q = 2 * K + 1;                             \
  using Z = Sz<K, Q>;                                      \
  constexpr int NPC = Z::NPC, NF2 = Z::NF2;                \
  (void)NF2;                                               \
  __shared__ Tabs<K, Q> tb;                                \
  __shared__ double shm[kWPB * (PER)];                     \
  const Tab& T = *Tp;                                      \
  load_tabs<K, Q>(T, tb);                                  \
  const int wv = threadIdx.x >> 6;                         \
  const int64_t cell_ = (int64_t)blockIdx.x * kWPB + wv;   \
  if (cell_ >= g.ncells) return;                           \
  Cell<K, Q> c(g, tb, cell_);                              \
  double* wbase = shm + wv * (PER);                        \
  (void)wbase;

template <int A, int B>
struct Max { static constexpr int v = A > B ? A : B; };
template <int A, int B, int C, int D = 0>
struct Max4 { static constexpr int v = Max<Max<A, B>::v, Max<C, D>::v>::v; };

// the z sweep of an integration, accumulated point by point: sa[i] += S[q][i] v + D[q][i] gz, ...
template <int K>
__device__ __attribute__((always_inline)) inline void zsweep_acc(const Tab& T, int q, double v, double gx, double gy,
                                                                  double gz, double (&sa)[K + 1], double (&sx)[K + 1],
                                                                  double (&sy)[K + 1]) {
#pragma unroll
  for (int i = 0; i <= K; ++i) {
    sa[i] += T.S[q][i] * v + T.D[q][i] * gz;
    sx[i] += T.S[q][i] * gx;
    sy[i] += T.S[q][i] * gy;
  }
}

// physical gradient contribution (g0, g1, g2) at quadrature point q of the lane's line -> reference
// coordinates times JxW
template <int K>
__device__ __attribute__((always_inline)) inline void to_ref(const LineGeo<3, K>& lg, const Geo& g, const Tab& T,
                                                              const int* ic, int q, double g0, double g1, double g2,
                                                              double& r0, double& r1, double& r2) {
  const double Gp[3] = {g0, g1, g2};
  double R[3];
  lg.gref(g, T, ic, q, Gp, R);
  r0 = R[0]; r1 = R[1]; r2 = R[2];
}

// ---------------------------------------------------------------------------------------
// Cell-local mass-type solve: M (rho null, scale 1), D (scale 1/(gamma-1)), A (rho_for_fixed),
// NC components (see k_mass_solve).
// b and x may be the same vector (every lane reads its cell's b before any x is written).
template <int K, int NC>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(2))) void k3_mass_solve(
    Geo g, const Tab* __restrict__ Tp, const double* __restrict__ rho, double scale, int64_t N,
    const double* b, double* x, double tol, int maxit) {
  constexpr int PER_ = Sz<K, 2 * K + 1>::NPC + Max<SolveWk<K, 2 * K + 1, NC>::SIZE, Sz<K, 2 * K + 1>::EV>::v;
  W3_PROLOGUE(PER_)
  double* nod = wbase;
  double* wk = wbase + NPC;
  double bb[NC];
#pragma unroll
  for (int a = 0; a < NC; ++a) bb[a] = c.own(b + a * N);
  const double rv = rho ? c.own(rho) : 0.0;
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  double co[Q];
  if (rho) {
    if (c.lane < NPC) nod[c.lane] = rv;
    wsync();
    double v[1][Q];
    eval<K, Q, 1>(c, T, nod, NPC, wk, v);
#pragma unroll
    for (int q = 0; q < Q; ++q) co[q] = v[0][q];
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) co[q] = 1.0;
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) co[q] *= scale * lg.jxw(g, T, q);
  double xx[NC];
  cheb_solve<K, Q, NC>(c, T, tb, co, wk, tol, maxit, bb, xx);
#pragma unroll
  for (int a = 0; a < NC; ++a) c.store(x + a * N, xx[a]);
}

// ---------------------------------------------------------------------------------------
// B: vmult_pressure (euler_operator.h:1510-1617) at node `lane`: volume -f p div v, faces +f {p} n.v
// (walls: p n.v), f = coeff dt / Ma^2. p staged at nod; its neighbour face nodes are loaded here,
// one direction ahead (nbf: Sz::NF2 values).
// Work-area values of integrate_B_volume: five z-swept line arrays (the four y-swept arrays reuse
// the same area).
template <int K, int Q>
struct BVol {
  static constexpr int N1 = K + 1, A1 = N1 * Q * Q, A2 = N1 * N1 * Q, SIZE = 5 * A1;
  static_assert(4 * A2 <= 5 * A1, "y-swept arrays fit the line arrays' area");
};

// The volume term of B, -f int p div(v) (euler_operator.h:1510-1617), for the three components at once.
// On the terrain-following map only the zeta row of J^{-1} is full, so with g = -f p JxW the reference
// integrands are g/hx d_xi (x), g/hy d_eta (y) and g J^{-1}_{zeta,a} d_zeta (a = x, y, z); after the
// vertical sweep on the lane's line they are five arrays of k + 1 sums (ps: S W p, pd: D W p, pdz:
// D W p (1 - zeta / H), each times the line's horizontal factors):
//   out_x = Sx Sy a0 + Dx Sy b0,  out_y = Sx Sy a1 + Sx Dy b1,  out_z = Sx Sy c2.
// The y sweep makes four arrays (Sy a0, Sy b0, Sy a1 + Dy b1, Sy c2), the x sweep the three results at
// node `lane`. The y-swept values stay in registers across a wavefront barrier and are then stored
// over the line arrays. wk: BVol::SIZE values.
template <int K, int Q>
__device__ __attribute__((always_inline)) inline void integrate_B_volume(const Cell<K, Q>& c, const Tabs<K, Q>& tb,
                                                                          double* wk, const double (&a0)[K + 1],
                                                                          const double (&b0)[K + 1], const double (&a1)[K + 1],
                                                                          const double (&b1)[K + 1], const double (&c2)[K + 1],
                                                                          double (&out)[3]) {
  using Z = Sz<K, Q>;
  using BV = BVol<K, Q>;
  constexpr int N1 = Z::N1, GQ = Z::GQ, NPC = Z::NPC, A1 = BV::A1, A2 = BV::A2;
  constexpr int NR = Z::cdiv(N1 * N1, GQ);  // rounds of the y sweep
  double* t = wk;  // [5][N1 iz][Q qy][Q qx]
  double* u = wk;  // [4][N1 iz][N1 iy][Q qx], written once every line array has been read
  if (c.line) {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const int o = (i * Q + c.qb) * Q + c.qa;
      t[o] = a0[i];
      t[A1 + o] = b0[i];
      t[2 * A1 + o] = a1[i];
      t[3 * A1 + o] = b1[i];
      t[4 * A1 + o] = c2[i];
    }
  }
  wsync();
  double sv[NR][4];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int pr = c.qb + GQ * r;  // (iz, iy)
    sv[r][0] = sv[r][1] = sv[r][2] = sv[r][3] = 0.0;
    if (c.grp_ok && pr < N1 * N1) {
      const int iz = pr / N1, iy = pr % N1;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int o = (iz * Q + q) * Q + c.qa;
        const double sy = tb.S[q * N1 + iy], dy = tb.D[q * N1 + iy];
        sv[r][0] += sy * t[o];
        sv[r][1] += sy * t[A1 + o];
        sv[r][2] += sy * t[2 * A1 + o] + dy * t[3 * A1 + o];
        sv[r][3] += sy * t[4 * A1 + o];
      }
    }
  }
  wsync();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int pr = c.qb + GQ * r;
    if (c.grp_ok && pr < N1 * N1) {
      const int o = pr * Q + c.qa;
#pragma unroll
      for (int a = 0; a < 4; ++a) u[a * A2 + o] = sv[r][a];
    }
  }
  wsync();
  out[0] = out[1] = out[2] = 0.0;
  if (c.lane < NPC) {
    const int ix = c.lane % N1, row = c.lane / N1;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double sx = tb.S[q * N1 + ix], dx = tb.D[q * N1 + ix];
      const int o = row * Q + q;
      out[0] += sx * u[o] + dx * u[A2 + o];
      out[1] += sx * u[2 * A2 + o];
      out[2] += sx * u[3 * A2 + o];
    }
  }
  wsync();
}

// B p at the cell's nodes (euler_operator.h:1510-1617): the face terms direction by direction (the x
// and y faces have the constant normals +-e_x, +-e_y, so only that component receives their
// integral), then the volume term from the vertical sums of p on the lane's line.
template <int K>
__device__ __attribute__((always_inline)) inline void apply_B_nodes(const Cell<K>& c, const Geo& g, const Tab& T,
                                                                     const Tabs<K, 2 * K + 1>& tb,
                                                                     const LineGeo<3, K>& lg, double f,
                                                                     const double* __restrict__ p, const double* nod,
                                                                     double* nbf, double* wk, double (&out)[3]) {
  constexpr int Q = 2 * K + 1;
  using Z = Sz<K, Q>;
  constexpr int N1 = K + 1;
  double fadd[3] = {0.0, 0.0, 0.0};
  double pn = c.nbr(p, 0);
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    c.put_nbr(nbf, pn);
    if (dir < 2) pn = c.nbr(p, dir + 1);
    wsync();
    double trp[1][2][2];
    traces<K, Q, 1>(c, tb, dir, nod, Z::NPC, nbf, Z::NF2, wk, trp);
    double n[2][3] = {}, jxw[2] = {};
    if (c.line) dir_geo<K>(g, T, c.ic, lg, dir, c.lane, n, jxw);
    if (dir < 2) {
      double Fa[1][2] = {};
      if (c.line)
#pragma unroll
        for (int side = 0; side < 2; ++side)
          Fa[0][side] = f * 0.5 * (trp[0][side][0] + trp[0][side][1]) * n[side][dir] * jxw[side];
      double add1[1] = {0.0};
      face_integrate<K, Q, 1>(c, tb, dir, wk, Fa, add1);
      fadd[dir] += add1[0];
    } else {
      double Fa[3][2] = {};
      if (c.line)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const double avg = 0.5 * (trp[0][side][0] + trp[0][side][1]);
#pragma unroll
          for (int a = 0; a < 3; ++a) Fa[a][side] = f * avg * n[side][a] * jxw[side];
        }
      face_integrate<K, Q, 3>(c, tb, dir, wk, Fa, fadd);
    }
  }
  double pq[1][Q];
  eval<K, Q, 1>(c, T, nod, Z::NPC, wk, pq);
  double a0[N1], b0[N1], a1[N1], b1[N1], c2[N1];
  {
    double ps[N1], pd[N1], pdz[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) ps[i] = pd[i] = pdz[i] = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double fz = 1.0 - (lg.zeta0 + T.QP[q] * lg.hz) * lg.iH;
      const double wp = T.W[q] * pq[0][q];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        ps[i] += T.S[q][i] * wp;
        pd[i] += T.D[q][i] * wp;
        pdz[i] += T.D[q][i] * (wp * fz);
      }
    }
    // the line's factors of gref (LineGeo) for the gradient fluxes -f p e_a
    const double fw = f * lg.wxy, hxy = lg.hx * lg.hy;
    const double ca0 = fw * hxy * lg.dhx, cb0 = -fw * lg.hy * lg.c, ca1 = fw * hxy * lg.dhy, cb1 = -fw * lg.hx * lg.c,
                 cc2 = -fw * hxy;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      a0[i] = ca0 * pdz[i];
      b0[i] = cb0 * ps[i];
      a1[i] = ca1 * pdz[i];
      b1[i] = cb1 * ps[i];
      c2[i] = cc2 * pd[i];
    }
  }
  double vol[3];
  integrate_B_volume<K, Q>(c, tb, wk, a0, b0, a1, b1, c2, vol);
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a] = vol[a] + fadd[a];
}

template <int K>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(2))) void k3_apply_B(
    Geo g, const Tab* __restrict__ Tp, double f, int64_t N, const double* __restrict__ p, double* __restrict__ dst) {
  constexpr int PER_ = Sz<K, 2 * K + 1>::NPC + Sz<K, 2 * K + 1>::NF2 +
                       Max4<BVol<K, 2 * K + 1>::SIZE, Sz<K, 2 * K + 1>::TR, 3 * Sz<K, 2 * K + 1>::FI, Sz<K, 2 * K + 1>::EV>::v;
  W3_PROLOGUE(PER_)
  double* nod = wbase;
  double* nbf = nod + NPC;
  double* wk = nbf + NF2;
  const double pv = c.own(p);
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  if (c.lane < NPC) nod[c.lane] = pv;
  wsync();
  double out[3];
  apply_B_nodes<K>(c, g, T, tb, lg, f, p, nod, nbf, wk, out);
#pragma unroll
  for (int a = 0; a < 3; ++a) c.store(dst + a * N, out[a]);
}

// ---------------------------------------------------------------------------------------
// C: vmult_enthalpy (euler_operator.h:1411-1507), nonlinear in w through lambda = max|w.n|;
// optionally fused with D pD (Schur second half): dst = scaleD M pD - C w.
template <int K, bool WITH_D>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(2))) void k3_apply_C(
    Geo g, const Tab* __restrict__ Tp, double cdt, int64_t N, const double* __restrict__ pfix,
    const double* __restrict__ w, const double* __restrict__ pD, double scaleD, double* __restrict__ dst) {
  constexpr int NO = WITH_D ? 5 : 4;  // staged own fields: pfix, w0, w1, w2 [, pD]
  constexpr int PER_ = NO * Sz<K, 2 * K + 1>::NPC + 4 * Sz<K, 2 * K + 1>::NF2 +
                       Max4<Sz<K, 2 * K + 1>::INGZ, 4 * Sz<K, 2 * K + 1>::TR, 2 * Sz<K, 2 * K + 1>::EV,
                            Sz<K, 2 * K + 1>::FI>::v;
  W3_PROLOGUE(PER_)
  double* nod = wbase;
  double* nbf = nod + NO * NPC;
  double* wk = nbf + 4 * NF2;
  const double* fl[4] = {pfix, w, w + N, w + 2 * N};
  double ov[NO], nv[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) ov[a] = c.own(fl[a]);
  if (WITH_D) ov[NO - 1] = c.own(pD);
#pragma unroll
  for (int a = 0; a < 4; ++a) nv[a] = c.nbr(fl[a], 0);
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  if (c.lane < NPC)
#pragma unroll
    for (int a = 0; a < NO; ++a) nod[a * NPC + c.lane] = ov[a];
  const double gr = kGamma / (kGamma - 1.0), igm1 = 1.0 / (kGamma - 1.0);
  const double sgn = WITH_D ? -1.0 : 1.0;
  double fadd[1] = {0.0};
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
#pragma unroll
    for (int a = 0; a < 4; ++a) c.put_nbr(nbf + a * NF2, nv[a]);
    if (dir < 2)
#pragma unroll
      for (int a = 0; a < 4; ++a) nv[a] = c.nbr(fl[a], dir + 1);
    wsync();
    double tr[4][2][2];
    traces<K, Q, 4>(c, tb, dir, nod, NPC, nbf, NF2, wk, tr);
    double Fo[1][2] = {};
    if (c.line) {
      double fn[2][3], fj[2];
      dir_geo<K>(g, T, c.ic, lg, dir, c.lane, fn, fj);
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int fc = 2 * dir + side;
        if (c.is_wall(fc)) continue;  // no boundary term (euler_operator.h:188-192)
        const double* n = fn[side];
        const double jxw = fj[side];
        const double pp = tr[0][side][0], pm = tr[0][side][1];
        double avg = 0.0, wnp = 0.0, wnm = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          avg += 0.5 * gr * (pp * tr[1 + d][side][0] + pm * tr[1 + d][side][1]) * n[d];
          wnp += tr[1 + d][side][0] * n[d];
          wnm += tr[1 + d][side][1] * n[d];
        }
        const double lam = fmax(fabs(wnp), fabs(wnm));
        Fo[0][side] = sgn * cdt * (avg + 0.5 * lam * igm1 * (pp - pm)) * jxw;
      }
    }
    face_integrate<K, Q, 1>(c, tb, dir, wk, Fo, fadd);
  }
  // volume: the fields on the lane's line after the horizontal sweeps (two at a time), the flux
  // G = s w with s = -cdt gamma/(gamma-1) p_fix mapped to reference coordinates as LineGeo::gref does
  // and swept vertically into the three line sums integrate_from_z takes
  constexpr int N1 = K + 1;
  double zp[2][N1], zw[2][N1];
  eval_lines<K, Q, 2>(c, nod, NPC, wk, zp);            // p_fix, w0
  eval_lines<K, Q, 2>(c, nod + 2 * NPC, NPC, wk, zw);  // w1, w2
  double zd[1][N1] = {};
  if (WITH_D) eval_lines<K, Q, 1>(c, nod + 4 * NPC, NPC, wk, zd);
  double sa[N1], sx[N1], sy[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) sa[i] = sx[i] = sy[i] = 0.0;
  if (c.line) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double wq = T.W[q];
      const double s = sgn * (-cdt * gr * line_value<K>(T, zp[0], q)) * wq;
      const double g0 = s * line_value<K>(T, zp[1], q), g1 = s * line_value<K>(T, zw[0], q),
                   g2 = s * line_value<K>(T, zw[1], q);
      const double fz = 1.0 - (lg.zeta0 + T.QP[q] * lg.hz) * lg.iH;
      const double gz = g2 - fz * (lg.dhx * g0 + lg.dhy * g1);
      const double vd = WITH_D ? scaleD * lg.c * wq * line_value<K>(T, zd[0], q) : 0.0;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        sx[i] += T.S[q][i] * g0;
        sy[i] += T.S[q][i] * g1;
        sa[i] += T.D[q][i] * gz + T.S[q][i] * vd;
      }
    }
    const double hxy = lg.hx * lg.hy, fx = lg.hy * lg.c * lg.wxy, fy = lg.hx * lg.c * lg.wxy, fa = hxy * lg.wxy;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      sx[i] *= fx;
      sy[i] *= fy;
      sa[i] *= fa;
    }
  }
  const double o = integrate_from_z<K, Q>(c, tb, wk, sa, sx, sy);
  c.store(dst, o + fadd[0]);
}

// ---------------------------------------------------------------------------------------
// E1: vmult_rhs_rho_update (euler_operator.h:342-473 cell, 476-692 faces, boundary none).
// src = {rho_j, u_j}, j < NS; stage weights wj.
template <int K, int NS>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(2))) void k3_rhs_rho(
    Geo g, const Tab* __restrict__ Tp, Srcs src, double dt, double w0, double w1, double w2, int64_t N,
    double* __restrict__ dst) {
  constexpr int PER_ = 4 * NS * Sz<K, 2 * K + 1>::NPC + 4 * NS * Sz<K, 2 * K + 1>::NF2 +
                       Max4<Sz<K, 2 * K + 1>::INGZ, 4 * Sz<K, 2 * K + 1>::TR, 2 * Sz<K, 2 * K + 1>::EV,
                            Sz<K, 2 * K + 1>::FI>::v;
  W3_PROLOGUE(PER_)
  double* nod = wbase;                 // [4 NS][NPC]: rho_j, u_j (3) per state
  double* nbf = nod + 4 * NS * NPC;    // [4 NS][NF2]: every state's neighbour face nodes of one direction
  double* wk = nbf + 4 * NS * NF2;
  const double wj[3] = {w0, w1, w2};
  auto fld = [&](int j, int m) -> const double* {
    return m == 0 ? src.p[2 * j] : src.p[2 * j + 1] + (m - 1) * N;
  };
  double nv[4 * NS];  // neighbour face nodes of the next direction, in flight during this one
  {
    double ov[4 * NS];
#pragma unroll
    for (int j = 0; j < NS; ++j)
#pragma unroll
      for (int m = 0; m < 4; ++m) ov[4 * j + m] = c.own(fld(j, m));
#pragma unroll
    for (int j = 0; j < NS; ++j)
#pragma unroll
      for (int m = 0; m < 4; ++m) nv[4 * j + m] = c.nbr(fld(j, m), 0);
    if (c.lane < NPC)
#pragma unroll
      for (int a = 0; a < 4 * NS; ++a) nod[a * NPC + c.lane] = ov[a];
  }
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  double fadd[1] = {0.0};
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    double F[2] = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4 * NS; ++a) c.put_nbr(nbf + a * NF2, nv[a]);
    if (dir < 2)
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int m = 0; m < 4; ++m) nv[4 * j + m] = c.nbr(fld(j, m), dir + 1);
    wsync();
    double fn[2][3], fj[2];
    if (c.line) dir_geo<K>(g, T, c.ic, lg, dir, c.lane, fn, fj);
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      double tr[4][2][2];
      traces<K, Q, 4>(c, tb, dir, nod + 4 * j * NPC, NPC, nbf + 4 * j * NF2, NF2, wk, tr);
      if (c.line)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int fc = 2 * dir + side;
          if (c.is_wall(fc)) continue;
          const double* n = fn[side];
          const double rp = tr[0][side][0], rm = tr[0][side][1];
          double avg = 0.0, unp = 0.0, unm = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            avg += 0.5 * (rp * tr[1 + d][side][0] + rm * tr[1 + d][side][1]) * n[d];
            unp += tr[1 + d][side][0] * n[d];
            unm += tr[1 + d][side][1] * n[d];
          }
          const double lam = fmax(fabs(unp), fabs(unm));
          F[side] += wj[j] * dt * (avg + 0.5 * lam * (rp - rm));
        }
    }
    double Fo[1][2] = {};
    if (c.line)
#pragma unroll
      for (int side = 0; side < 2; ++side) Fo[0][side] = -F[side] * fj[side];
    face_integrate<K, Q, 1>(c, tb, dir, wk, Fo, fadd);
  }
  // volume: rho^n psi + dt sum_j w_j (rho u)_j . grad psi. Each state's fields on the lane's line after
  // the horizontal sweeps; the flux mapped to reference coordinates as LineGeo::gref does and swept
  // vertically into the three line sums of integrate_from_z
  constexpr int N1 = K + 1;
  double sa[N1], sx[N1], sy[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) sa[i] = sx[i] = sy[i] = 0.0;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double zr[2][N1], zu[2][N1];
    eval_lines<K, Q, 2>(c, nod + 4 * j * NPC, NPC, wk, zr);         // rho_j, u_j0
    eval_lines<K, Q, 2>(c, nod + (4 * j + 2) * NPC, NPC, wk, zu);  // u_j1, u_j2
    if (c.line) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const double wq = T.W[q];
        const double r = line_value<K>(T, zr[0], q);
        const double s = wj[j] * dt * r * wq;
        const double g0 = s * line_value<K>(T, zr[1], q), g1 = s * line_value<K>(T, zu[0], q),
                     g2 = s * line_value<K>(T, zu[1], q);
        const double fz = 1.0 - (lg.zeta0 + T.QP[q] * lg.hz) * lg.iH;
        const double gz = g2 - fz * (lg.dhx * g0 + lg.dhy * g1);
        const double vm = j == 0 ? r * lg.c * wq : 0.0;  // rho^n JxW / (hx hy wxy)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          sx[i] += T.S[q][i] * g0;
          sy[i] += T.S[q][i] * g1;
          sa[i] += T.D[q][i] * gz + T.S[q][i] * vm;
        }
      }
    }
  }
  if (c.line) {
    const double hxy = lg.hx * lg.hy, fx = lg.hy * lg.c * lg.wxy, fy = lg.hx * lg.c * lg.wxy, fa = hxy * lg.wxy;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      sx[i] *= fx;
      sy[i] *= fy;
      sa[i] *= fa;
    }
  }
  c.store(dst, integrate_from_z<K, Q>(c, tb, wk, sa, sx, sy) + fadd[0]);
}

// ---------------------------------------------------------------------------------------
// E2: vmult_rhs_velocity_fixed (euler_operator.h:1620-1886 cell, 1889-2184 faces, 2187-2305 walls).
// Explicit states j < NE: (rho, u, p) = src[3j..3j+2]; gravity densities src[3j], j < NG (for j < NE
// the states' own densities, for j = NE the new density of stages 1-2).
template <int K, int NE, int NG>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(2))) void k3_rhs_u(
    Geo g, const Tab* __restrict__ Tp, Srcs src, E2Coef cf, double dt, double iMa2, double iFr2, int64_t N,
    double* __restrict__ dst) {
  constexpr bool XG = NG > NE;
  constexpr int NF = 5 * NE + (XG ? 1 : 0);
  constexpr int PER_ = NF * Sz<K, 2 * K + 1>::NPC + 5 * NE * Sz<K, 2 * K + 1>::NF2 +
                       Max4<Sz<K, 2 * K + 1>::ING, 5 * Sz<K, 2 * K + 1>::TR, 3 * Sz<K, 2 * K + 1>::EV,
                            3 * Sz<K, 2 * K + 1>::FI>::v;
  W3_PROLOGUE(PER_)
  double* nod = wbase;               // [NF][NPC]: (rho, u0, u1, u2, p) per state [, rho_new]
  double* nbf = nod + NF * NPC;      // [5 NE][NF2]: every state's neighbour face nodes of one direction
  double* wk = nbf + 5 * NE * NF2;
  auto fld = [&](int a) -> const double* {
    const int j = a / 5, m = a % 5;
    if (a == 5 * NE) return src.p[3 * NE];
    return m == 0 ? src.p[3 * j] : m == 4 ? src.p[3 * j + 2] : src.p[3 * j + 1] + (m - 1) * N;
  };
  double nv[5 * NE];  // neighbour face nodes of the next direction, in flight during this one
  {
    double ov[NF];
#pragma unroll
    for (int a = 0; a < NF; ++a) ov[a] = c.own(fld(a));
#pragma unroll
    for (int a = 0; a < 5 * NE; ++a) nv[a] = c.nbr(fld(a), 0);
    if (c.lane < NPC)
#pragma unroll
      for (int a = 0; a < NF; ++a) nod[a * NPC + c.lane] = ov[a];
  }
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  double fadd[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    double F[3][2] = {};
#pragma unroll
    for (int a = 0; a < 5 * NE; ++a) c.put_nbr(nbf + a * NF2, nv[a]);
    if (dir < 2)
#pragma unroll
      for (int a = 0; a < 5 * NE; ++a) nv[a] = c.nbr(fld(a), dir + 1);
    wsync();
    double fn[2][3], fj[2];
    if (c.line) dir_geo<K>(g, T, c.ic, lg, dir, c.lane, fn, fj);
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const double wj = cf.w[j], wpj = cf.wp[j];
      double tr[5][2][2];
      traces<K, Q, 5>(c, tb, dir, nod + 5 * j * NPC, NPC, nbf + 5 * j * NF2, NF2, wk, tr);
      if (c.line)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int fc = 2 * dir + side;
          const double* n = fn[side];
          const double pp = tr[4][side][0];
          if (!c.is_wall(fc)) {
            const double rp = tr[0][side][0], rm = tr[0][side][1], pm = tr[4][side][1];
            double unp = 0.0, unm = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) { unp += tr[1 + d][side][0] * n[d]; unm += tr[1 + d][side][1] * n[d]; }
            const double lam = fmax(fabs(unp), fabs(unm));
            const double avg_p = 0.5 * (pp + pm);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const double upa = tr[1 + a][side][0], uma = tr[1 + a][side][1];
              const double avg_tn = 0.5 * (rp * upa * unp + rm * uma * unm);
              const double jump = rp * upa - rm * uma;
              F[a][side] += wj * dt * avg_tn + wpj * dt * iMa2 * avg_p * n[a] + wj * dt * 0.5 * lam * jump;
            }
          } else {
            // wall: p_D = p, flux (p + p_D)/2 n (euler_operator.h:2213-2218)
#pragma unroll
            for (int a = 0; a < 3; ++a) F[a][side] += wpj * dt * iMa2 * (0.5 * (pp + pp)) * n[a];
          }
        }
    }
    if (c.line)
#pragma unroll
      for (int side = 0; side < 2; ++side)
#pragma unroll
        for (int a = 0; a < 3; ++a) F[a][side] *= -fj[side];
    face_integrate<K, Q, 3>(c, tb, dir, wk, F, fadd);
  }
  // cells second, state by state: the symmetric flux tensor rho u (x) u (with the pressure term on
  // its diagonal) of each state is mapped to reference coordinates and swept along the lane's line at
  // once, so only the z-swept sums of the three momentum components stay live
  double sa0[K + 1] = {}, sx0[K + 1] = {}, sy0[K + 1] = {}, sa1[K + 1] = {}, sx1[K + 1] = {}, sy1[K + 1] = {},
         sa2[K + 1] = {}, sx2[K + 1] = {}, sy2[K + 1] = {};

#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const double wj = cf.w[j], wpj = cf.wp[j];
    double zl[5][K + 1];
    {
      double z3[3][K + 1], z2[2][K + 1];
      eval_lines<K, Q, 3>(c, nod + 5 * j * NPC, NPC, wk, z3);        // rho, u0, u1
      eval_lines<K, Q, 2>(c, nod + (5 * j + 3) * NPC, NPC, wk, z2);  // u2, p
#pragma unroll
      for (int i = 0; i <= K; ++i) {
        zl[0][i] = z3[0][i]; zl[1][i] = z3[1][i]; zl[2][i] = z3[2][i];
        zl[3][i] = z2[0][i]; zl[4][i] = z2[1][i];
      }
    }
    // the flux tensor mapped to reference coordinates as LineGeo::gref does, with the line's
    // horizontal factors applied to the z-swept sums once, after the last state
#pragma unroll 1
    for (int q = 0; q < Q; ++q) {
      const double wq = T.W[q];
      const double r = line_value<K>(T, zl[0], q);
      const double u[3] = {line_value<K>(T, zl[1], q), line_value<K>(T, zl[2], q), line_value<K>(T, zl[3], q)};
      const double pw = wpj * dt * iMa2 * line_value<K>(T, zl[4], q) * wq;
      const double fz = 1.0 - (lg.zeta0 + T.QP[q] * lg.hz) * lg.iH;
      const double cw = lg.c * wq;  // JxW / (hx hy wxy)
      double V[3] = {0.0, 0.0, 0.0};
      if (j == 0)
#pragma unroll
        for (int d = 0; d < 3; ++d) V[d] = r * u[d] * cw;
      if (j < NG) V[2] -= cf.gw[j] * dt * iFr2 * r * cw;
      const double ru = wj * dt * r * wq;
      {
        const double g0 = ru * u[0] * u[0] + pw, g1 = ru * u[0] * u[1], g2 = ru * u[0] * u[2];
        zsweep_acc<K>(T, q, V[0], g0, g1, g2 - fz * (lg.dhx * g0 + lg.dhy * g1), sa0, sx0, sy0);
      }
      {
        const double g0 = ru * u[1] * u[0], g1 = ru * u[1] * u[1] + pw, g2 = ru * u[1] * u[2];
        zsweep_acc<K>(T, q, V[1], g0, g1, g2 - fz * (lg.dhx * g0 + lg.dhy * g1), sa1, sx1, sy1);
      }
      {
        const double g0 = ru * u[2] * u[0], g1 = ru * u[2] * u[1], g2 = ru * u[2] * u[2] + pw;
        zsweep_acc<K>(T, q, V[2], g0, g1, g2 - fz * (lg.dhx * g0 + lg.dhy * g1), sa2, sx2, sy2);
      }
    }
  }
  if (XG) {
    double zx[1][K + 1];
    eval_lines<K, Q, 1>(c, nod + 5 * NE * NPC, NPC, wk, zx);
#pragma unroll
    for (int q = 0; q < Q; ++q)
      zsweep_acc<K>(T, q, -cf.gw[NG - 1] * dt * iFr2 * line_value<K>(T, zx[0], q) * lg.c * T.W[q], 0.0, 0.0, 0.0,
                    sa2, sx2, sy2);
  }
  {
    const double fx = lg.hy * lg.c * lg.wxy, fy = lg.hx * lg.c * lg.wxy, fa = lg.hx * lg.hy * lg.wxy;
#pragma unroll
    for (int i = 0; i <= K; ++i) {
      sx0[i] *= fx; sy0[i] *= fy; sa0[i] *= fa;
      sx1[i] *= fx; sy1[i] *= fy; sa1[i] *= fa;
      sx2[i] *= fx; sy2[i] *= fy; sa2[i] *= fa;
    }
  }
  c.store(dst, integrate_from_z<K, Q>(c, tb, wk, sa0, sx0, sy0) + fadd[0]);
  c.store(dst + N, integrate_from_z<K, Q>(c, tb, wk, sa1, sx1, sy1) + fadd[1]);
  c.store(dst + 2 * N, integrate_from_z<K, Q>(c, tb, wk, sa2, sx2, sy2) + fadd[2]);
}

// ---------------------------------------------------------------------------------------
// E3: vmult_rhs_pressure (euler_operator.h:743-1007 cell, 1010-1361 faces, boundary none).
// Explicit states j < NE: (rho, u, p) = src[3j..]; kinetic correction with (src[3NE], src[3NE+1]);
// gravity work with (src[3j], src[3j+1]), j < NG (j = NE: the correction pair). FULL = false leaves
// out the two terms of stages 1-2 that read u_fixed (added by k_rhs_p_ufix).
// (the stage-2 form with every term holds the u_old traces as well: one wave per SIMD, no spills)
template <int K, int NE, int NG, int STAGE, bool FULL>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(STAGE == 2 && FULL ? 1 : 2))) void k3_rhs_p(
    Geo g, const Tab* __restrict__ Tp, Srcs src, E3Coef cf, double dt, double Ma2, double iFr2, int64_t N,
    double* __restrict__ dst) {
  constexpr bool KC = FULL || STAGE == 3;  // kinetic correction (and the extra gravity pair) present
  constexpr int NF = 5 * NE + (KC ? 4 : 0);  // staged states [, correction pair rho, u]
  constexpr int PER_ = NF * Sz<K, 2 * K + 1>::NPC + 5 * NE * Sz<K, 2 * K + 1>::NF2 +
                       Max4<Sz<K, 2 * K + 1>::INGZ, 5 * Sz<K, 2 * K + 1>::TR, 2 * Sz<K, 2 * K + 1>::EV,
                            Sz<K, 2 * K + 1>::FI>::v;
  W3_PROLOGUE(PER_)
  constexpr double gm1 = kGamma - 1.0;
  double* nod = wbase;                 // [NF][NPC]
  double* nbf = nod + NF * NPC;        // [5 NE][NF2]: every state's neighbour face nodes of one direction
  double* wk = nbf + 5 * NE * NF2;
  auto fld = [&](int a) -> const double* {
    if (a >= 5 * NE) return a == 5 * NE ? src.p[3 * NE] : src.p[3 * NE + 1] + (a - 5 * NE - 1) * N;
    const int j = a / 5, m = a % 5;
    return m == 0 ? src.p[3 * j] : m == 4 ? src.p[3 * j + 2] : src.p[3 * j + 1] + (m - 1) * N;
  };
  double nv[5 * NE];  // neighbour face nodes of the next direction, in flight during this one
  {
    double ov[NF];
#pragma unroll
    for (int a = 0; a < NF; ++a) ov[a] = c.own(fld(a));
#pragma unroll
    for (int a = 0; a < 5 * NE; ++a) nv[a] = c.nbr(fld(a), 0);
    if (c.lane < NPC)
#pragma unroll
      for (int a = 0; a < NF; ++a) nod[a * NPC + c.lane] = ov[a];
  }
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  double fadd[1] = {0.0};
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    double F[2] = {0.0, 0.0};
    double tu0[3][2][2];  // u_old traces (stage-2 kinetic jump as written, euler_operator.h:1182-1183)
#pragma unroll
    for (int a = 0; a < 5 * NE; ++a) c.put_nbr(nbf + a * NF2, nv[a]);
    if (dir < 2)
#pragma unroll
      for (int a = 0; a < 5 * NE; ++a) nv[a] = c.nbr(fld(a), dir + 1);
    wsync();
    double fn[2][3], fj[2];
    if (c.line) dir_geo<K>(g, T, c.ic, lg, dir, c.lane, fn, fj);
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const double wj = cf.w[j], wpj = cf.wt[j];
      double tr[5][2][2];
      traces<K, Q, 5>(c, tb, dir, nod + 5 * j * NPC, NPC, nbf + 5 * j * NF2, NF2, wk, tr);
      if (STAGE == 2 && j == 0)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int side = 0; side < 2; ++side) {
            tu0[d][side][0] = tr[1 + d][side][0];
            tu0[d][side][1] = tr[1 + d][side][1];
          }
      if (c.line)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int fc = 2 * dir + side;
          if (c.is_wall(fc)) continue;
          const double* n = fn[side];
          const double rp = tr[0][side][0], rm = tr[0][side][1], pp = tr[4][side][0], pm = tr[4][side][1];
          double uup = 0.0, uum = 0.0, unp = 0.0, unm = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double up = tr[1 + d][side][0], um = tr[1 + d][side][1];
            uup += up * up; uum += um * um;
            unp += up * n[d]; unm += um * n[d];
          }
          const double Ep = 1.0 / gm1 * pp / rp + 0.5 * Ma2 * uup;
          const double Em = 1.0 / gm1 * pm / rm + 0.5 * Ma2 * uum;
          double avg_kin = 0.0, avg_ent = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double up = tr[1 + d][side][0], um = tr[1 + d][side][1];
            avg_kin += 0.5 * (0.5 * uup * rp * up + 0.5 * uum * rm * um) * n[d];
            avg_ent += 0.5 * (((Ep - 0.5 * Ma2 * uup) * rp + pp) * up + ((Em - 0.5 * Ma2 * uum) * rm + pm) * um) * n[d];
          }
          const double lam = fmax(fabs(unp), fabs(unm));
          if (STAGE == 3) {
            const double jump = rp * Ep - rm * Em;
            F[side] += wj * dt * Ma2 * avg_kin + wj * dt * avg_ent + wj * dt * 0.5 * lam * jump;
          } else {
            double jk;
            if (STAGE == 2 && j == 1) {  // euler_operator.h:1182-1183, u_tmp_2 . u_old as written
              double d0p = 0.0, d0m = 0.0;
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                d0p += tr[1 + d][side][0] * tu0[d][side][0];
                d0m += tr[1 + d][side][1] * tu0[d][side][1];
              }
              jk = rp * 0.5 * d0p - rm * 0.5 * d0m;
            } else {
              jk = rp * 0.5 * uup - rm * 0.5 * uum;
            }
            const double je = rp * (Ep - 0.5 * Ma2 * uup) - rm * (Em - 0.5 * Ma2 * uum);
            F[side] += wj * dt * Ma2 * (avg_kin + 0.5 * lam * jk) + wpj * dt * (avg_ent + 0.5 * lam * je);
          }
        }
    }
    double Fo[1][2] = {};
    if (c.line)
#pragma unroll
      for (int side = 0; side < 2; ++side) Fo[0][side] = -F[side] * fj[side];
    face_integrate<K, Q, 1>(c, tb, dir, wk, Fo, fadd);
  }
  // volume: the terms of every state accumulated straight into the three vertical line sums of
  // integrate_from_z (fluxes mapped to reference coordinates as LineGeo::gref does; the line's
  // horizontal factors applied once at the end)
  constexpr int N1 = K + 1;
  double sa[N1], sx[N1], sy[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) sa[i] = sx[i] = sy[i] = 0.0;
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const double wj = cf.w[j], wpj = cf.wt[j];
    double zl[5][K + 1];
    {
      double za[2][K + 1], zb[2][K + 1], zc[1][K + 1];
      eval_lines<K, Q, 2>(c, nod + 5 * j * NPC, NPC, wk, za);        // rho, u0
      eval_lines<K, Q, 2>(c, nod + (5 * j + 2) * NPC, NPC, wk, zb);  // u1, u2
      eval_lines<K, Q, 1>(c, nod + (5 * j + 4) * NPC, NPC, wk, zc);  // p
#pragma unroll
      for (int i = 0; i <= K; ++i) {
        zl[0][i] = za[0][i]; zl[1][i] = za[1][i]; zl[2][i] = zb[0][i];
        zl[3][i] = zb[1][i]; zl[4][i] = zc[0][i];
      }
    }
    if (c.line) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const double wq = T.W[q];
        const double r = line_value<K>(T, zl[0], q), p = line_value<K>(T, zl[4], q);
        const double u[3] = {line_value<K>(T, zl[1], q), line_value<K>(T, zl[2], q), line_value<K>(T, zl[3], q)};
        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double E = 1.0 / gm1 * (p / r) + 0.5 * Ma2 * uu;
        double v = 0.0;
        if (j == 0) v += r * E;
        if (j < NG) v -= cf.gw[j] * dt * Ma2 * iFr2 * r * u[2];  // gravity work of state j
        v *= lg.c * wq;                                           // JxW / (hx hy wxy)
        const double kin = 0.5 * wj * dt * Ma2 * uu * r;
        const double ent = wpj * dt * (r * (E - 0.5 * Ma2 * uu) + p);
        const double f = (kin + ent) * wq;
        const double g0 = f * u[0], g1 = f * u[1], g2 = f * u[2];
        const double fz = 1.0 - (lg.zeta0 + T.QP[q] * lg.hz) * lg.iH;
        zsweep_acc<K>(T, q, v, g0, g1, g2 - fz * (lg.dhx * g0 + lg.dhy * g1), sa, sx, sy);
      }
    }
  }
  if (KC) {
    // kinetic correction, then (stages 1-2) the gravity work of the pair (rho_new, u_fix)
    double zk[4][K + 1];
    {
      double za[2][K + 1], zb[2][K + 1];
      eval_lines<K, Q, 2>(c, nod + 5 * NE * NPC, NPC, wk, za);
      eval_lines<K, Q, 2>(c, nod + (5 * NE + 2) * NPC, NPC, wk, zb);
#pragma unroll
      for (int i = 0; i <= K; ++i) { zk[0][i] = za[0][i]; zk[1][i] = za[1][i]; zk[2][i] = zb[0][i]; zk[3][i] = zb[1][i]; }
    }
    if (c.line) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const double r = line_value<K>(T, zk[0], q), w2 = line_value<K>(T, zk[3], q);
        const double u0 = line_value<K>(T, zk[1], q), u1 = line_value<K>(T, zk[2], q);
        double v = -0.5 * r * Ma2 * (u0 * u0 + u1 * u1 + w2 * w2);
        if (NG > NE) v -= cf.gw[NG - 1] * dt * Ma2 * iFr2 * r * w2;
        zsweep_acc<K>(T, q, v * lg.c * T.W[q], 0.0, 0.0, 0.0, sa, sx, sy);
      }
    }
  }
  if (c.line) {
    const double fx = lg.hy * lg.c * lg.wxy, fy = lg.hx * lg.c * lg.wxy, fa = lg.hx * lg.hy * lg.wxy;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      sx[i] *= fx;
      sy[i] *= fy;
      sa[i] *= fa;
    }
  }
  c.store(dst, integrate_from_z<K, Q>(c, tb, wk, sa, sx, sy) + fadd[0]);
}

// The u_fixed terms of E3 at an implicit stage (see k_rhs_p_ufix): dst = base + int psi V,
// V = -1/2 Ma^2 rho_new |u_fix|^2 - cg rho_new u_fix,z.
template <int K>
__global__ __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(2))) void k3_rhs_p_ufix(
    Geo g, const Tab* __restrict__ Tp, const double* __restrict__ rho, const double* __restrict__ u, double half_ma2,
    double cg, int64_t N, const double* __restrict__ base, double* __restrict__ dst) {
  constexpr int PER_ = 4 * Sz<K, 2 * K + 1>::NPC + Max<4 * Sz<K, 2 * K + 1>::EV, Sz<K, 2 * K + 1>::INV>::v;
  W3_PROLOGUE(PER_)
  double* nod = wbase;
  double* wk = nod + 4 * NPC;
  double ov[4];
  ov[0] = c.own(rho);
#pragma unroll
  for (int d = 0; d < 3; ++d) ov[1 + d] = c.own(u + d * N);
  const double bv = c.own(base);
  if (c.lane < NPC)
#pragma unroll
    for (int a = 0; a < 4; ++a) nod[a * NPC + c.lane] = ov[a];
  wsync();
  LineGeo<3, K> lg;
  lg.init(g, T, c.ic, c.lane);
  double v[4][Q];
  eval<K, Q, 4>(c, T, nod, NPC, wk, v);
  double V[1][Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) s += v[1 + d][q] * v[1 + d][q] * half_ma2 + (d == 2 ? cg * v[1 + d][q] : 0.0);
    V[0][q] = -v[0][q] * s * lg.jxw(g, T, q);
  }
  const double G0[1][3][Q] = {};
  double o[1];
  integrate<K, Q, 1, true, false>(c, T, tb, wk, V, G0, o);
  c.store(dst, o[0] + bv);
}

}  // namespace w3

// ---------------------------------------------------------------------------------------
// Launchers (3D, k <= 3); the callers in afl_ops.hip route those cases here.
#define W3_DISPATCH(KV, ...)                                     \
  do {                                                           \
    switch (KV) {                                                \
      case 1: { constexpr int K_ = 1; __VA_ARGS__; } break;      \
      case 2: { constexpr int K_ = 2; __VA_ARGS__; } break;      \
      case 3: { constexpr int K_ = 3; __VA_ARGS__; } break;      \
    }                                                            \
  } while (0)

static dim3 w3_grid(int64_t ncells) { return dim3((unsigned)((ncells + w3::kWPB - 1) / w3::kWPB)); }
static dim3 w3_block() { return dim3(64 * w3::kWPB); }

void launch3_mass_solve(const OpCtx& o, double by, const double* rho, double scale, int ncomp, const double* b,
                        double* x) {
  W3_DISPATCH(o.k, {
    if (ncomp == 1)
      launch_timed(o, kTimeMassSolve, (int)o.hyp, by, w3::k3_mass_solve<K_, 1>, w3_grid(o.geo.ncells), w3_block(),
                   o.geo, o.dT, rho, scale, o.N, b, x, o.cell_tol, o.cell_maxit);
    else
      launch_timed(o, kTimeASolve, (int)o.hyp, by, w3::k3_mass_solve<K_, 3>, w3_grid(o.geo.ncells), w3_block(),
                   o.geo, o.dT, rho, scale, o.N, b, x, o.cell_tol, o.cell_maxit);
  });
}

void launch3_apply_B(const OpCtx& o, double by, double f, const double* p, double* dst) {
  W3_DISPATCH(o.k, {
    launch_timed(o, kTimeApplyB, (int)o.hyp, by, w3::k3_apply_B<K_>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, f,
                 o.N, p, dst);
  });
}

// w = A^{-1} B p as two launches, B p into w and the cell solves in place: one kernel doing both needed
// 256 VGPRs (two waves per SIMD) against 188 and 120 for the halves, and the pair runs 15 % faster
// (96x96x40, k = 3: 2.52 + 2.24 ms against 5.61 ms). Timed as a B apply and a mass-type solve.
void launch3_schur_w(const OpCtx& o, double by, double f, const double* rho, const double* p, double* w) {
  (void)by;
  const double sz = 8.0 * (double)o.N;
  W3_DISPATCH(o.k, {
    launch_timed(o, kTimeApplyB, (int)o.hyp, sz * (1 + 3), w3::k3_apply_B<K_>, w3_grid(o.geo.ncells), w3_block(), o.geo,
                 o.dT, f, o.N, p, w);
    launch_timed(o, kTimeASolve, (int)o.hyp, sz * (1 + 3 + 3), w3::k3_mass_solve<K_, 3>, w3_grid(o.geo.ncells),
                 w3_block(), o.geo, o.dT, rho, 1.0, o.N, (const double*)w, w, o.cell_tol, o.cell_maxit);
  });
}

void launch3_apply_C(const OpCtx& o, double by, double cdt, const double* pfix, const double* w, const double* pD,
                     double* dst) {
  W3_DISPATCH(o.k, {
    if (pD)
      launch_timed(o, kTimeApplyC, (int)o.hyp, by, w3::k3_apply_C<K_, true>, w3_grid(o.geo.ncells), w3_block(), o.geo,
                   o.dT, cdt, o.N, pfix, w, pD, 1.0 / (kGamma - 1.0), dst);
    else
      launch_timed(o, kTimeApplyC, (int)o.hyp, by, w3::k3_apply_C<K_, false>, w3_grid(o.geo.ncells), w3_block(), o.geo,
                   o.dT, cdt, o.N, pfix, w, (const double*)nullptr, 0.0, dst);
  });
}

void launch3_rhs_rho(const OpCtx& o, double by, int stage, const Srcs& S, double w0, double w1, double w2,
                     double* dst) {
  W3_DISPATCH(o.k, {
    if (stage == 1)
      launch_timed(o, kTimeRhsRho, 1, by, w3::k3_rhs_rho<K_, 1>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, S,
                   o.dt, w0, w1, w2, o.N, dst);
    else if (stage == 2)
      launch_timed(o, kTimeRhsRho, 2, by, w3::k3_rhs_rho<K_, 2>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, S,
                   o.dt, w0, w1, w2, o.N, dst);
    else
      launch_timed(o, kTimeRhsRho, 3, by, w3::k3_rhs_rho<K_, 3>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, S,
                   o.dt, w0, w1, w2, o.N, dst);
  });
}

void launch3_rhs_u(const OpCtx& o, double by, int stage, const Srcs& S, const E2Coef& cf, double iMa2, double iFr2,
                   double* dst) {
  W3_DISPATCH(o.k, {
    if (stage == 1)
      launch_timed(o, kTimeRhsU, 1, by, w3::k3_rhs_u<K_, 1, 2>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, S, cf,
                   o.dt, iMa2, iFr2, o.N, dst);
    else if (stage == 2)
      launch_timed(o, kTimeRhsU, 2, by, w3::k3_rhs_u<K_, 2, 3>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, S, cf,
                   o.dt, iMa2, iFr2, o.N, dst);
    else
      launch_timed(o, kTimeRhsU, 3, by, w3::k3_rhs_u<K_, 3, 3>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT, S, cf,
                   o.dt, iMa2, iFr2, o.N, dst);
  });
}

void launch3_rhs_p(const OpCtx& o, double by, int stage, bool full, const Srcs& S, const E3Coef& cf, double Ma2,
                   double iFr2, double* dst) {
  W3_DISPATCH(o.k, {
    const dim3 gr = w3_grid(o.geo.ncells), bl = w3_block();
    if (stage == 1 && full)
      launch_timed(o, kTimeRhsP, 1, by, w3::k3_rhs_p<K_, 1, 2, 1, true>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N,
                   dst);
    else if (stage == 1)
      launch_timed(o, kTimeRhsP, 1, by, w3::k3_rhs_p<K_, 1, 2, 1, false>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2,
                   o.N, dst);
    else if (stage == 2 && full)
      launch_timed(o, kTimeRhsP, 2, by, w3::k3_rhs_p<K_, 2, 3, 2, true>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N,
                   dst);
    else if (stage == 2)
      launch_timed(o, kTimeRhsP, 2, by, w3::k3_rhs_p<K_, 2, 3, 2, false>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2,
                   o.N, dst);
    else
      launch_timed(o, kTimeRhsP, 3, by, w3::k3_rhs_p<K_, 3, 3, 3, true>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N,
                   dst);
  });
}

void launch3_rhs_p_ufix(const OpCtx& o, double by, int stage, const double* rho_new, const double* u_fix,
                        double half_ma2, double cg, const double* base, double* dst) {
  W3_DISPATCH(o.k, {
    launch_timed(o, kTimeRhsPUfix, stage, by, w3::k3_rhs_p_ufix<K_>, w3_grid(o.geo.ncells), w3_block(), o.geo, o.dT,
                 rho_new, u_fix, half_ma2, cg, o.N, base, dst);
  });
}

}  // namespace afl
