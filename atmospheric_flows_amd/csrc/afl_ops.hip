// Operator kernels of the IMEX-DG hot path for gfx950. Each kernel is a fused cell + face
// (+ wall) loop over one batch of cells per block (see afl_common.h for the decomposition).
// Weak forms follow euler_operator.h; every kernel cites the reference assembler it replaces.
#include "afl_internal.h"

namespace afl {

struct Srcs { const double* p[12]; };

template <int DIM, int K>
struct Lds {
  using C = Cfg<DIM, K>;
  static constexpr int EV = C::EV1 + C::EV2;
  static constexpr int IN = C::IN1 + C::IN2;
  static constexpr int W = EV > IN ? EV : IN;   // eval / integrate work area per cell
  static constexpr int F = 2 * DIM * C::NFQ;    // face flux block per cell and component
};

#define AFL_SLOT_PTRS                                                      \
  using C = Cfg<DIM, K>;                                                   \
  constexpr int NQ = C::NQ, N1 = C::N1;                                    \
  (void)N1;                                                                \
  Slot<DIM, K> s(g);                                                       \
  const int cbs = s.cb < C::CPB ? s.cb : 0;                                \
  double* t1 = shw + cbs * Lds<DIM, K>::W;                                 \
  double* t2 = t1 + C::EV1;                                                \
  double* i1 = t1;                                                         \
  double* i2 = t1 + C::IN1;                                                \
  (void)t2; (void)i2;

// ---------------------------------------------------------------------------------------
// E1: vmult_rhs_rho_update (euler_operator.h:342-473 cell, 476-692 faces, boundary none).
// src = {rho_j, u_j} for the NS explicit states; w = stage weights (a21 | a31,a32 | b1,b2,b3).
template <int DIM, int K, int NS>
__global__ __launch_bounds__(AFL_NT) void k_rhs_rho(Geo g, Tab T, Srcs src, double dt, double w0, double w1,
                                                     double w2, int64_t N, double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  const double wj[3] = {w0, w1, w2};
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double V[NQ], G[DIM][NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) { V[q] = 0.0; for (int d = 0; d < DIM; ++d) G[d][q] = 0.0; }
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double r[NQ];
    eval_line<DIM, K>(src.p[2 * j], s, t1, t2, T, r);
    if (j == 0) for (int q = 0; q < NQ; ++q) V[q] = r[q];
    for (int d = 0; d < DIM; ++d) {
      double u[NQ];
      eval_line<DIM, K>(src.p[2 * j + 1] + d * N, s, t1, t2, T, u);
#pragma unroll
      for (int q = 0; q < NQ; ++q) G[d][q] += wj[j] * dt * r[q] * u[q];
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double Gp[3] = {0, 0, 0}, R[3];
    for (int d = 0; d < DIM; ++d) Gp[d] = G[d][q];
    lg.gref(g, T, s.ic, q, Gp, R);
    for (int d = 0; d < DIM; ++d) G[d][q] = R[d];
    V[q] *= lg.jxw(g, T, q);
  }
  // faces: thread l handles face point l of every face of its cell
  double* Fq = shF + cbs * Lds<DIM, K>::F;
  if (s.act && s.l < C::NFQ) {
    for (int f = 0; f < 2 * DIM; ++f) {
      const int dir = f >> 1, side = f & 1;
      const int64_t nb = neighbour<DIM>(g, s.ic, f);
      double F = 0.0;
      if (nb >= 0) {
        double n[3], jxw;
        face_geo<DIM, K>(g, T, s.ic, f, s.l, n, jxw);
        for (int j = 0; j < NS; ++j) {
          const double rp = trace<DIM, K>(src.p[2 * j], s.cell, dir, side, s.l, T);
          const double rm = trace<DIM, K>(src.p[2 * j], nb, dir, 1 - side, s.l, T);
          double avg = 0.0, unp = 0.0, unm = 0.0;
          for (int d = 0; d < DIM; ++d) {
            const double up = trace<DIM, K>(src.p[2 * j + 1] + d * N, s.cell, dir, side, s.l, T);
            const double um = trace<DIM, K>(src.p[2 * j + 1] + d * N, nb, dir, 1 - side, s.l, T);
            avg += 0.5 * (rp * up + rm * um) * n[d];
            unp += up * n[d]; unm += um * n[d];
          }
          const double lam = fmax(fabs(unp), fabs(unm));
          F += wj[j] * dt * (avg + 0.5 * lam * (rp - rm));
        }
        F = -F * jxw;
      }
      Fq[f * C::NFQ + s.l] = F;
    }
  }
  __syncthreads();
  double out[N1];
  integrate_line<DIM, K, true, true>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst, s, out);
}

// ---------------------------------------------------------------------------------------
// E2: vmult_rhs_velocity_fixed (euler_operator.h:1620-1886 cell, 1889-2184 faces, 2187-2305 walls).
// Explicit states j < NE: (rho, u, p) = src[3j..3j+2], advective weight w[j], pressure weight wp[j].
// Gravity densities: src[3j] with weight gw[j], j < NG (rho_old, rho_tmp_2, rho_tmp_3 / rho_curr).
struct E2Coef { double w[3], wp[3], gw[3]; };

template <int DIM, int K, int NE, int NG>
__global__ __launch_bounds__(AFL_NT) void k_rhs_u(Geo g, Tab T, Srcs src, E2Coef cf, double dt, double iMa2,
                                                   double iFr2, int64_t N, double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F * DIM];
  AFL_SLOT_PTRS
  constexpr int NSYM = DIM * (DIM + 1) / 2;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double V[DIM][NQ], Sym[NSYM][NQ], P[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    for (int d = 0; d < DIM; ++d) V[d][q] = 0.0;
    for (int d = 0; d < NSYM; ++d) Sym[d][q] = 0.0;
    P[q] = 0.0;
  }
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    double r[NQ], u[DIM][NQ];
    eval_line<DIM, K>(src.p[3 * j], s, t1, t2, T, r);
    for (int d = 0; d < DIM; ++d) eval_line<DIM, K>(src.p[3 * j + 1] + d * N, s, t1, t2, T, u[d]);
    double pj[NQ];
    eval_line<DIM, K>(src.p[3 * j + 2], s, t1, t2, T, pj);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (j == 0) for (int d = 0; d < DIM; ++d) V[d][q] = r[q] * u[d][q];
      int m = 0;
      for (int a = 0; a < DIM; ++a)
        for (int b = a; b < DIM; ++b) Sym[m++][q] += cf.w[j] * dt * r[q] * u[a][q] * u[b][q];
      P[q] += cf.wp[j] * dt * iMa2 * pj[q];
    }
  }
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    double r[NQ];
    eval_line<DIM, K>(src.p[3 * j], s, t1, t2, T, r);  // gravity densities: rho_old, rho_tmp_2, rho_tmp_3/curr
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[DIM - 1][q] -= cf.gw[j] * dt * iFr2 * r[q];
  }
  // faces (all components)
  double* Fq = shF + cbs * Lds<DIM, K>::F * DIM;
  if (s.act && s.l < C::NFQ) {
    for (int f = 0; f < 2 * DIM; ++f) {
      const int dir = f >> 1, side = f & 1;
      const int64_t nb = neighbour<DIM>(g, s.ic, f);
      double n[3], jxw;
      face_geo<DIM, K>(g, T, s.ic, f, s.l, n, jxw);
      double F[DIM];
      for (int a = 0; a < DIM; ++a) F[a] = 0.0;
      for (int j = 0; j < NE; ++j) {
        const double pp = trace<DIM, K>(src.p[3 * j + 2], s.cell, dir, side, s.l, T);
        if (nb >= 0) {
          const double rp = trace<DIM, K>(src.p[3 * j], s.cell, dir, side, s.l, T);
          const double rm = trace<DIM, K>(src.p[3 * j], nb, dir, 1 - side, s.l, T);
          const double pm = trace<DIM, K>(src.p[3 * j + 2], nb, dir, 1 - side, s.l, T);
          double up[DIM], um[DIM], unp = 0.0, unm = 0.0;
          for (int d = 0; d < DIM; ++d) {
            up[d] = trace<DIM, K>(src.p[3 * j + 1] + d * N, s.cell, dir, side, s.l, T);
            um[d] = trace<DIM, K>(src.p[3 * j + 1] + d * N, nb, dir, 1 - side, s.l, T);
            unp += up[d] * n[d]; unm += um[d] * n[d];
          }
          const double lam = fmax(fabs(unp), fabs(unm));
          const double avg_p = 0.5 * (pp + pm);
          for (int a = 0; a < DIM; ++a) {
            const double avg_tn = 0.5 * (rp * up[a] * unp + rm * um[a] * unm);
            const double jump = rp * up[a] - rm * um[a];
            F[a] += cf.w[j] * dt * avg_tn + cf.wp[j] * dt * iMa2 * avg_p * n[a] + cf.w[j] * dt * 0.5 * lam * jump;
          }
        } else {
          // wall: p_D = p, flux (p + p_D)/2 n (euler_operator.h:2213-2218)
          for (int a = 0; a < DIM; ++a) F[a] += cf.wp[j] * dt * iMa2 * (0.5 * (pp + pp)) * n[a];
        }
      }
      for (int a = 0; a < DIM; ++a) Fq[(a * 2 * DIM + f) * C::NFQ + s.l] = -F[a] * jxw;
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    double Va[NQ], Ga[DIM][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double Gp[3] = {0, 0, 0}, R[3];
      for (int b = 0; b < DIM; ++b) {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const int m = lo * DIM - lo * (lo - 1) / 2 + (hi - lo);
        Gp[b] = Sym[m][q] + (a == b ? P[q] : 0.0);
      }
      lg.gref(g, T, s.ic, q, Gp, R);
      for (int b = 0; b < DIM; ++b) Ga[b][q] = R[b];
      Va[q] = V[a][q] * lg.jxw(g, T, q);
    }
    double out[N1];
    integrate_line<DIM, K, true, true>(s, i1, i2, T, Va, Ga, out);
    add_faces<DIM, K>(s, Fq + a * Lds<DIM, K>::F, T, out, 0xffu);
    store_row<DIM, K>(dst + a * N, s, out);
  }
}

// ---------------------------------------------------------------------------------------
// E3: vmult_rhs_pressure (euler_operator.h:743-1007 cell, 1010-1361 faces, boundary none).
// Explicit states j < NE: (rho,u,p) = src[3j..]; kinetic weight w[j], enthalpy weight wt[j].
// Kinetic correction with (src[3NE], src[3NE+1]); gravity work with (src[3j], src[3j+1]), j < NG.
struct E3Coef { double w[3], wt[3], gw[3]; };

template <int DIM, int K, int NE, int NG, int STAGE>
__global__ __launch_bounds__(AFL_NT) void k_rhs_p(Geo g, Tab T, Srcs src, E3Coef cf, double dt, double Ma2,
                                                   double iFr2, int64_t N, double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  constexpr double gm1 = kGamma - 1.0;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double V[NQ], G[DIM][NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) { V[q] = 0.0; for (int d = 0; d < DIM; ++d) G[d][q] = 0.0; }
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    double r[NQ], u[DIM][NQ], p[NQ];
    eval_line<DIM, K>(src.p[3 * j], s, t1, t2, T, r);
    for (int d = 0; d < DIM; ++d) eval_line<DIM, K>(src.p[3 * j + 1] + d * N, s, t1, t2, T, u[d]);
    eval_line<DIM, K>(src.p[3 * j + 2], s, t1, t2, T, p);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double uu = 0.0;
      for (int d = 0; d < DIM; ++d) uu += u[d][q] * u[d][q];
      const double E = 1.0 / gm1 * (p[q] / r[q]) + 0.5 * Ma2 * uu;
      if (j == 0) V[q] += r[q] * E;
      const double kin = 0.5 * cf.w[j] * dt * Ma2 * uu * r[q];
      const double ent = cf.wt[j] * dt * (r[q] * (E - 0.5 * Ma2 * uu) + p[q]);
      for (int d = 0; d < DIM; ++d) G[d][q] += (kin + ent) * u[d][q];
    }
  }
  {
    double r[NQ], u[NQ], uu[NQ];
    eval_line<DIM, K>(src.p[3 * NE], s, t1, t2, T, r);  // kinetic correction (rho, u) = src[3 NE], src[3 NE + 1]
#pragma unroll
    for (int q = 0; q < NQ; ++q) uu[q] = 0.0;
    for (int d = 0; d < DIM; ++d) {
      eval_line<DIM, K>(src.p[3 * NE + 1] + d * N, s, t1, t2, T, u);
#pragma unroll
      for (int q = 0; q < NQ; ++q) uu[q] += u[q] * u[q];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] -= 0.5 * r[q] * Ma2 * uu[q];
  }
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    double r[NQ], w[NQ];
    eval_line<DIM, K>(src.p[3 * j], s, t1, t2, T, r);  // gravity work: (rho, u) = src[3j], src[3j+1]
    eval_line<DIM, K>(src.p[3 * j + 1] + (DIM - 1) * N, s, t1, t2, T, w);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] -= cf.gw[j] * dt * Ma2 * iFr2 * r[q] * w[q];
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double Gp[3] = {0, 0, 0}, R[3];
    for (int d = 0; d < DIM; ++d) Gp[d] = G[d][q];
    lg.gref(g, T, s.ic, q, Gp, R);
    for (int d = 0; d < DIM; ++d) G[d][q] = R[d];
    V[q] *= lg.jxw(g, T, q);
  }
  double* Fq = shF + cbs * Lds<DIM, K>::F;
  if (s.act && s.l < C::NFQ) {
    for (int f = 0; f < 2 * DIM; ++f) {
      const int dir = f >> 1, side = f & 1;
      const int64_t nb = neighbour<DIM>(g, s.ic, f);
      double F = 0.0;
      if (nb >= 0) {
        double n[3], jxw;
        face_geo<DIM, K>(g, T, s.ic, f, s.l, n, jxw);
        double u0p[DIM], u0m[DIM];
        if (STAGE == 2)
          for (int d = 0; d < DIM; ++d) {
            u0p[d] = trace<DIM, K>(src.p[1] + d * N, s.cell, dir, side, s.l, T);
            u0m[d] = trace<DIM, K>(src.p[1] + d * N, nb, dir, 1 - side, s.l, T);
          }
        for (int j = 0; j < NE; ++j) {
          const double rp = trace<DIM, K>(src.p[3 * j], s.cell, dir, side, s.l, T);
          const double rm = trace<DIM, K>(src.p[3 * j], nb, dir, 1 - side, s.l, T);
          const double pp = trace<DIM, K>(src.p[3 * j + 2], s.cell, dir, side, s.l, T);
          const double pm = trace<DIM, K>(src.p[3 * j + 2], nb, dir, 1 - side, s.l, T);
          double up[DIM], um[DIM], uup = 0.0, uum = 0.0, unp = 0.0, unm = 0.0;
          for (int d = 0; d < DIM; ++d) {
            up[d] = trace<DIM, K>(src.p[3 * j + 1] + d * N, s.cell, dir, side, s.l, T);
            um[d] = trace<DIM, K>(src.p[3 * j + 1] + d * N, nb, dir, 1 - side, s.l, T);
            uup += up[d] * up[d]; uum += um[d] * um[d];
            unp += up[d] * n[d]; unm += um[d] * n[d];
          }
          const double Ep = 1.0 / gm1 * pp / rp + 0.5 * Ma2 * uup;
          const double Em = 1.0 / gm1 * pm / rm + 0.5 * Ma2 * uum;
          double avg_kin = 0.0, avg_ent = 0.0;
          for (int d = 0; d < DIM; ++d) {
            avg_kin += 0.5 * (0.5 * uup * rp * up[d] + 0.5 * uum * rm * um[d]) * n[d];
            avg_ent += 0.5 * (((Ep - 0.5 * Ma2 * uup) * rp + pp) * up[d] + ((Em - 0.5 * Ma2 * uum) * rm + pm) * um[d]) * n[d];
          }
          const double lam = fmax(fabs(unp), fabs(unm));
          if (STAGE == 3) {
            const double jump = rp * Ep - rm * Em;
            F += cf.w[j] * dt * Ma2 * avg_kin + cf.w[j] * dt * avg_ent + cf.w[j] * dt * 0.5 * lam * jump;
          } else {
            double jk;
            if (STAGE == 2 && j == 1) {  // euler_operator.h:1182-1183, u_tmp_2 . u_old as written
              double d0p = 0.0, d0m = 0.0;
              for (int d = 0; d < DIM; ++d) { d0p += up[d] * u0p[d]; d0m += um[d] * u0m[d]; }
              jk = rp * 0.5 * d0p - rm * 0.5 * d0m;
            } else {
              jk = rp * 0.5 * uup - rm * 0.5 * uum;
            }
            const double je = rp * (Ep - 0.5 * Ma2 * uup) - rm * (Em - 0.5 * Ma2 * uum);
            F += cf.w[j] * dt * Ma2 * (avg_kin + 0.5 * lam * jk) + cf.wt[j] * dt * (avg_ent + 0.5 * lam * je);
          }
        }
        F = -F * jxw;
      }
      Fq[f * C::NFQ + s.l] = F;
    }
  }
  __syncthreads();
  double out[N1];
  integrate_line<DIM, K, true, true>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst, s, out);
}

// ---------------------------------------------------------------------------------------
// Mass-type apply: dst = scale * M_rho src, M (euler_operator.h:715-740), D (:1384-1408),
// A (:2328-2357). rho == nullptr means rho = 1.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_mass_apply(Geo g, Tab T, const double* __restrict__ rho, double scale,
                                                        int ncomp, int64_t N, const double* __restrict__ src,
                                                        double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double co[NQ];
  if (rho) eval_line<DIM, K>(rho, s, t1, t2, T, co);
  else for (int q = 0; q < NQ; ++q) co[q] = 1.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) co[q] *= scale * lg.jxw(g, T, q);
  for (int a = 0; a < ncomp; ++a) {
    double v[NQ], G[DIM][NQ], out[N1];
    eval_line<DIM, K>(src + a * N, s, t1, t2, T, v);
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] *= co[q];
    integrate_line<DIM, K, true, false>(s, i1, i2, T, v, G, out);
    store_row<DIM, K>(dst + a * N, s, out);
  }
}

// Inverse diagonal, euler_operator.h:2456-2626: diag_i = sum_q JxW coef phi_i(q)^2.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_inv_diag(Geo g, Tab T, const double* __restrict__ rho, double scale,
                                                      int ncomp, int64_t N, double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double co[NQ];
  if (rho) eval_line<DIM, K>(rho, s, t1, t2, T, co);
  else for (int q = 0; q < NQ; ++q) co[q] = 1.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) co[q] *= scale * lg.jxw(g, T, q);
  double out[N1];
  if (DIM == 2) {
    if (s.act) {
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) a += T.S[q][i] * T.S[q][i] * co[q];
        i1[i * NQ + s.l] = a;
      }
    }
    __syncthreads();
    if (s.act && s.l < N1)
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) a += T.S[q][i] * T.S[q][i] * i1[s.l * NQ + q];
        out[i] = 1.0 / a;
      }
  } else {
    if (s.act) {
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) a += T.S[q][i] * T.S[q][i] * co[q];
        i1[i * NQ * NQ + s.l] = a;
      }
    }
    __syncthreads();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) a += T.S[q][i] * T.S[q][i] * i1[(iz * NQ + q) * NQ + qx];
        i2[(iz * N1 + i) * NQ + qx] = a;
      }
    }
    __syncthreads();
    if (s.act && s.l < N1 * N1)
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) a += T.S[q][i] * T.S[q][i] * i2[s.l * NQ + q];
        out[i] = 1.0 / a;
      }
  }
  for (int a = 0; a < ncomp; ++a) store_row<DIM, K>(dst + a * N, s, out);
}

// ---------------------------------------------------------------------------------------
// B: vmult_pressure (euler_operator.h:1510-1617). Volume -f p div v, faces +f {p} n . v
// (own side), walls +f p n . v, with f = coeff dt / Ma^2.
template <int DIM, int K>
__device__ inline void apply_B_rows(const Geo& g, const Tab& T, const Slot<DIM, K>& s, const LineGeo<DIM, K>& lg,
                                    const double* __restrict__ p, double f, double* t1, double* t2, double* i1,
                                    double* i2, double* Fq, double (&out)[DIM][K + 1]) {
  using C = Cfg<DIM, K>;
  constexpr int NQ = C::NQ;
  double pq[NQ];
  eval_line<DIM, K>(p, s, t1, t2, T, pq);
  if (s.act && s.l < C::NFQ) {
    for (int fc = 0; fc < 2 * DIM; ++fc) {
      const int dir = fc >> 1, side = fc & 1;
      const int64_t nb = neighbour<DIM>(g, s.ic, fc);
      double n[3], jxw;
      face_geo<DIM, K>(g, T, s.ic, fc, s.l, n, jxw);
      const double pp = trace<DIM, K>(p, s.cell, dir, side, s.l, T);
      const double pm = nb >= 0 ? trace<DIM, K>(p, nb, dir, 1 - side, s.l, T) : pp;
      const double avg = 0.5 * (pp + pm);
      for (int a = 0; a < DIM; ++a) Fq[(a * 2 * DIM + fc) * C::NFQ + s.l] = f * avg * n[a] * jxw;
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    double V[NQ], G[DIM][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double Gp[3] = {0, 0, 0}, R[3];
      Gp[a] = -f * pq[q];
      lg.gref(g, T, s.ic, q, Gp, R);
      for (int d = 0; d < DIM; ++d) G[d][q] = R[d];
    }
    integrate_line<DIM, K, false, true>(s, i1, i2, T, V, G, out[a]);
    add_faces<DIM, K>(s, Fq + a * 2 * DIM * C::NFQ, T, out[a], 0xffu);
  }
}

template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_apply_B(Geo g, Tab T, double f, int64_t N, const double* __restrict__ p,
                                                     double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F * DIM];
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double out[DIM][N1];
  apply_B_rows<DIM, K>(g, T, s, lg, p, f, t1, t2, i1, i2, shF + cbs * Lds<DIM, K>::F * DIM, out);
  for (int a = 0; a < DIM; ++a) store_row<DIM, K>(dst + a * N, s, out[a]);
}

// ---------------------------------------------------------------------------------------
// C: vmult_enthalpy (euler_operator.h:1411-1507); nonlinear in w through lambda = max|w.n|.
// Optionally fused with D p (Schur second half): dst = scaleD * M p - C w.
template <int DIM, int K, bool WITH_D>
__global__ __launch_bounds__(AFL_NT) void k_apply_C(Geo g, Tab T, double cdt, int64_t N,
                                                     const double* __restrict__ pfix, const double* __restrict__ w,
                                                     const double* __restrict__ pD, double scaleD,
                                                     double* __restrict__ dst) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  constexpr double gr = kGamma / (kGamma - 1.0), igm1 = 1.0 / (kGamma - 1.0);
  const double sgn = WITH_D ? -1.0 : 1.0;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double pq[NQ], G[DIM][NQ], V[NQ];
  eval_line<DIM, K>(pfix, s, t1, t2, T, pq);
  for (int d = 0; d < DIM; ++d) {
    double wq[NQ];
    eval_line<DIM, K>(w + d * N, s, t1, t2, T, wq);
#pragma unroll
    for (int q = 0; q < NQ; ++q) G[d][q] = sgn * (-cdt * gr * pq[q] * wq[q]);
  }
  if (WITH_D) {
    eval_line<DIM, K>(pD, s, t1, t2, T, V);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] *= scaleD * lg.jxw(g, T, q);
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double Gp[3] = {0, 0, 0}, R[3];
    for (int d = 0; d < DIM; ++d) Gp[d] = G[d][q];
    lg.gref(g, T, s.ic, q, Gp, R);
    for (int d = 0; d < DIM; ++d) G[d][q] = R[d];
  }
  double* Fq = shF + cbs * Lds<DIM, K>::F;
  if (s.act && s.l < C::NFQ) {
    for (int f = 0; f < 2 * DIM; ++f) {
      const int dir = f >> 1, side = f & 1;
      const int64_t nb = neighbour<DIM>(g, s.ic, f);
      double F = 0.0;
      if (nb >= 0) {
        double n[3], jxw;
        face_geo<DIM, K>(g, T, s.ic, f, s.l, n, jxw);
        const double pp = trace<DIM, K>(pfix, s.cell, dir, side, s.l, T);
        const double pm = trace<DIM, K>(pfix, nb, dir, 1 - side, s.l, T);
        double avg = 0.0, wnp = 0.0, wnm = 0.0;
        for (int d = 0; d < DIM; ++d) {
          const double wp = trace<DIM, K>(w + d * N, s.cell, dir, side, s.l, T);
          const double wm = trace<DIM, K>(w + d * N, nb, dir, 1 - side, s.l, T);
          avg += 0.5 * gr * (pp * wp + pm * wm) * n[d];
          wnp += wp * n[d]; wnm += wm * n[d];
        }
        const double lam = fmax(fabs(wnp), fabs(wnm));
        F = sgn * cdt * (avg + 0.5 * lam * igm1 * (pp - pm)) * jxw;
      }
      Fq[f * C::NFQ + s.l] = F;
    }
  }
  __syncthreads();
  double out[N1];
  integrate_line<DIM, K, WITH_D, true>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst, s, out);
}

// ---------------------------------------------------------------------------------------
// Cell-local preconditioned CG for the block-diagonal mass-type matrices (exact inverse up to
// the stopping tolerance). Matrix: sum_q co(q) phi_i phi_j with co = scale rho JxW on the
// thread's line; preconditioner: (MI x MI (x MI)) / mean(co / w_ref). b rows in, x rows out.
template <int DIM, int K, int NC>
__device__ inline void cell_pcg(const Geo& g, const Tab& T, const Slot<DIM, K>& s, const double (&co)[2 * K + 1],
                                double* cell_lds /* NPC */, double* t1, double* t2, double* i1, double* i2,
                                double* red /* CPB*ROWS */, int* flag, double tol2, int maxit,
                                const double (&b)[NC][K + 1], double (&x)[NC][K + 1]) {
  using C = Cfg<DIM, K>;
  constexpr int NQ = C::NQ, N1 = C::N1, ROWS = C::ROWS;
  const bool row_owner = s.act && s.l < ROWS;
  double* rd = red + (s.cb < C::CPB ? s.cb : 0) * ROWS;
  // mean coefficient relative to the reference weights: sum_q co / sum_q w_ref = sum_q co
  // (reference weights sum to 1); each line sums its own points, rows reduce.
  double lsum = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) lsum += co[q];
  // reduce over the LT lines of the cell through LDS
  __shared__ double red_lines[AFL_NT];
  red_lines[threadIdx.x] = lsum;
  __syncthreads();
  double mean = 1.0;
  if (s.act) {
    double a = 0.0;
    const int base = s.cb * C::LT;
    for (int t = 0; t < C::LT; ++t) a += red_lines[base + t];
    mean = a;
  }
  __syncthreads();
  const double pinv = 1.0 / mean;

  // per-cell reduction of row partial sums; result identical for all threads of the cell
  auto cell_sum = [&](double part) -> double {
    if (row_owner) rd[s.l] = part;
    __syncthreads();
    double a = 0.0;
    if (s.act) for (int t = 0; t < ROWS; ++t) a += rd[t];
    __syncthreads();
    return a;
  };
  // z = P^{-1} r for one component (rows in registers), tensor-product MI sweeps through LDS
  auto precond = [&](const double (&r)[N1], double (&z)[N1]) {
    double tmp[N1];
    if (row_owner) {
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int j = 0; j < N1; ++j) a += T.MI[i][j] * r[j];
        tmp[i] = a;
      }
      for (int i = 0; i < N1; ++i) cell_lds[s.l * N1 + i] = tmp[i];
    }
    __syncthreads();
    if (DIM == 2) {
      // along z: thread l < N1 handles column ix = l
      if (s.act && s.l < N1) {
        double c[N1];
        for (int j = 0; j < N1; ++j) c[j] = cell_lds[j * N1 + s.l];
        for (int i = 0; i < N1; ++i) {
          double a = 0.0;
          for (int j = 0; j < N1; ++j) a += T.MI[i][j] * c[j];
          t1[i * N1 + s.l] = a;
        }
      }
      __syncthreads();
      if (row_owner) for (int i = 0; i < N1; ++i) z[i] = pinv * t1[s.l * N1 + i];
      __syncthreads();
    } else {
      // along y: thread (ix, iz) = (l % N1, l / N1)
      if (s.act && s.l < N1 * N1) {
        const int ix = s.l % N1, iz = s.l / N1;
        double c[N1];
        for (int j = 0; j < N1; ++j) c[j] = cell_lds[(iz * N1 + j) * N1 + ix];
        for (int i = 0; i < N1; ++i) {
          double a = 0.0;
          for (int j = 0; j < N1; ++j) a += T.MI[i][j] * c[j];
          t1[(iz * N1 + i) * N1 + ix] = a;
        }
      }
      __syncthreads();
      // along z: thread (ix, iy)
      if (s.act && s.l < N1 * N1) {
        const int ix = s.l % N1, iy = s.l / N1;
        double c[N1];
        for (int j = 0; j < N1; ++j) c[j] = t1[(j * N1 + iy) * N1 + ix];
        for (int i = 0; i < N1; ++i) {
          double a = 0.0;
          for (int j = 0; j < N1; ++j) a += T.MI[i][j] * c[j];
          cell_lds[(i * N1 + iy) * N1 + ix] = a;
        }
      }
      __syncthreads();
      if (row_owner) for (int i = 0; i < N1; ++i) z[i] = pinv * cell_lds[s.l * N1 + i];
      __syncthreads();
    }
  };
  // y = A v for one component: v rows -> LDS -> evaluate -> * co -> integrate
  auto apply = [&](const double (&v)[N1], double (&y)[N1]) {
    if (row_owner) for (int i = 0; i < N1; ++i) cell_lds[s.l * N1 + i] = v[i];
    __syncthreads();
    double vq[NQ], G[DIM][NQ];
    eval_line_lds<DIM, K>(cell_lds, s, t1, t2, T, vq);
#pragma unroll
    for (int q = 0; q < NQ; ++q) vq[q] *= co[q];
    integrate_line<DIM, K, true, false>(s, i1, i2, T, vq, G, y);
  };

  double r[NC][N1], z[NC][N1], p[NC][N1], ap[N1];
  double rz = 0.0, bb = 0.0;
  for (int c = 0; c < NC; ++c) {
    for (int i = 0; i < N1; ++i) { x[c][i] = 0.0; r[c][i] = row_owner ? b[c][i] : 0.0; }
    precond(r[c], z[c]);
    double prz = 0.0, pbb = 0.0;
    if (row_owner) for (int i = 0; i < N1; ++i) { p[c][i] = z[c][i]; prz += r[c][i] * z[c][i]; pbb += b[c][i] * b[c][i]; }
    rz += cell_sum(prz);
    bb += cell_sum(pbb);
  }
  bool done = !(bb > 0.0);
  for (int it = 0; it < maxit; ++it) {
    // block-wide termination check
    if (threadIdx.x == 0) *flag = 0;
    __syncthreads();
    if (s.act && !done) *flag = 1;
    __syncthreads();
    if (*flag == 0) break;
    double pap = 0.0;
    double apv[NC][N1];
    for (int c = 0; c < NC; ++c) {
      apply(p[c], ap);
      double part = 0.0;
      if (row_owner) for (int i = 0; i < N1; ++i) { apv[c][i] = ap[i]; part += p[c][i] * ap[i]; }
      pap += cell_sum(part);
    }
    const double alpha = (done || pap == 0.0) ? 0.0 : rz / pap;
    double rr = 0.0;
    for (int c = 0; c < NC; ++c) {
      double part = 0.0;
      if (row_owner)
        for (int i = 0; i < N1; ++i) { x[c][i] += alpha * p[c][i]; r[c][i] -= alpha * apv[c][i]; part += r[c][i] * r[c][i]; }
      rr += cell_sum(part);
    }
    if (rr <= tol2 * bb) done = true;
    double rz_new = 0.0;
    for (int c = 0; c < NC; ++c) {
      precond(r[c], z[c]);
      double part = 0.0;
      if (row_owner) for (int i = 0; i < N1; ++i) part += r[c][i] * z[c][i];
      rz_new += cell_sum(part);
    }
    const double beta = (done || rz == 0.0) ? 0.0 : rz_new / rz;
    rz = rz_new;
    for (int c = 0; c < NC; ++c)
      if (row_owner) for (int i = 0; i < N1; ++i) p[c][i] = z[c][i] + beta * p[c][i];
  }
}

template <int DIM, int K, int NC>
__global__ __launch_bounds__(AFL_NT) void k_mass_solve(Geo g, Tab T, const double* __restrict__ rho, double scale,
                                                        int64_t N, const double* __restrict__ b,
                                                        double* __restrict__ x, double tol2, int maxit) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shc[Cfg<DIM, K>::CPB * Cfg<DIM, K>::NPC];
  __shared__ double shr[Cfg<DIM, K>::CPB * Cfg<DIM, K>::ROWS];
  __shared__ int flag;
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double co[NQ];
  if (rho) eval_line<DIM, K>(rho, s, t1, t2, T, co);
  else for (int q = 0; q < NQ; ++q) co[q] = 1.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) co[q] *= scale * lg.jxw(g, T, q);
  double bb[NC][N1], xx[NC][N1];
  for (int c = 0; c < NC; ++c)
    for (int i = 0; i < N1; ++i)
      bb[c][i] = (s.act && s.l < C::ROWS) ? b[c * N + s.cell * C::NPC + s.l * N1 + i] : 0.0;
  cell_pcg<DIM, K, NC>(g, T, s, co, shc + cbs * C::NPC, t1, t2, i1, i2, shr, &flag, tol2, maxit, bb, xx);
  for (int c = 0; c < NC; ++c) store_row<DIM, K>(x + c * N, s, xx[c]);
}

// Schur first half: w = A^{-1} B p (fused B apply + cell-local A solve).
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_schur_w(Geo g, Tab T, const double* __restrict__ rho, double f,
                                                     int64_t N, const double* __restrict__ p, double* __restrict__ w,
                                                     double tol2, int maxit) {
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::W];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F * DIM];
  __shared__ double shc[Cfg<DIM, K>::CPB * Cfg<DIM, K>::NPC];
  __shared__ double shr[Cfg<DIM, K>::CPB * Cfg<DIM, K>::ROWS];
  __shared__ int flag;
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double bp[DIM][N1];
  apply_B_rows<DIM, K>(g, T, s, lg, p, f, t1, t2, i1, i2, shF + cbs * Lds<DIM, K>::F * DIM, bp);
  double co[NQ];
  eval_line<DIM, K>(rho, s, t1, t2, T, co);
#pragma unroll
  for (int q = 0; q < NQ; ++q) co[q] *= lg.jxw(g, T, q);
  double xx[DIM][N1];
  cell_pcg<DIM, K, DIM>(g, T, s, co, shc + cbs * C::NPC, t1, t2, i1, i2, shr, &flag, tol2, maxit, bp, xx);
  for (int c = 0; c < DIM; ++c) store_row<DIM, K>(w + c * N, s, xx[c]);
}

// ---------------------------------------------------------------------------------------
// Launchers (dispatch on dim and degree).
template <int DIM, int K>
static dim3 grid_for(int ncells) {
  using C = Cfg<DIM, K>;
  return dim3((ncells + C::CPB - 1) / C::CPB);
}

#define AFL_DISPATCH(DIMV, KV, ...)                                   \
  do {                                                                 \
    if ((DIMV) == 2) {                                                 \
      switch (KV) {                                                    \
        case 1: { constexpr int D_ = 2, K_ = 1; __VA_ARGS__; } break;         \
        case 2: { constexpr int D_ = 2, K_ = 2; __VA_ARGS__; } break;         \
        case 3: { constexpr int D_ = 2, K_ = 3; __VA_ARGS__; } break;         \
        case 4: { constexpr int D_ = 2, K_ = 4; __VA_ARGS__; } break;         \
      }                                                                \
    } else {                                                           \
      switch (KV) {                                                    \
        case 1: { constexpr int D_ = 3, K_ = 1; __VA_ARGS__; } break;         \
        case 2: { constexpr int D_ = 3, K_ = 2; __VA_ARGS__; } break;         \
        case 3: { constexpr int D_ = 3, K_ = 3; __VA_ARGS__; } break;         \
        case 4: { constexpr int D_ = 3, K_ = 4; __VA_ARGS__; } break;         \
      }                                                                \
    }                                                                  \
  } while (0)

static Srcs make_srcs(const double* const* src, int n) {
  Srcs s{};
  for (int i = 0; i < n && i < 12; ++i) s.p[i] = src[i];
  return s;
}

void launch_rhs_rho(const OpCtx& o, int stage, const double* const* src, double* dst) {
  const Srcs S = make_srcs(src, 2 * stage);
  const TableauC& t = o.tab;
  AFL_DISPATCH(o.dim, o.k, {
    dim3 gr = grid_for<D_, K_>(o.geo.ncells);
    if (stage == 1)
      k_rhs_rho<D_, K_, 1><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, o.dt, t.a21, 0.0, 0.0, o.N, dst);
    else if (stage == 2)
      k_rhs_rho<D_, K_, 2><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, o.dt, t.a31, t.a32, 0.0, o.N, dst);
    else
      k_rhs_rho<D_, K_, 3><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, o.dt, t.b1, t.b2, t.b3, o.N, dst);
  });
}

void launch_rhs_u(const OpCtx& o, int stage, const double* const* src, double* dst) {
  const TableauC& t = o.tab;
  const Srcs S = make_srcs(src, stage == 1 ? 4 : (stage == 2 ? 7 : 9));
  E2Coef cf{};
  const double iMa2 = 1.0 / (o.Ma * o.Ma), iFr2 = 1.0 / (o.Fr * o.Fr);
  if (stage == 1) {
    cf.w[0] = t.a21; cf.wp[0] = t.at21;
    cf.gw[0] = t.at21; cf.gw[1] = t.at22;
  } else if (stage == 2) {
    cf.w[0] = t.a31; cf.w[1] = t.a32; cf.wp[0] = t.at31; cf.wp[1] = t.at32;
    cf.gw[0] = t.at31; cf.gw[1] = t.at32; cf.gw[2] = t.at33;
  } else {
    cf.w[0] = t.b1; cf.w[1] = t.b2; cf.w[2] = t.b3; cf.wp[0] = t.b1; cf.wp[1] = t.b2; cf.wp[2] = t.b3;
    cf.gw[0] = t.b1; cf.gw[1] = t.b2; cf.gw[2] = t.b3;
  }
  AFL_DISPATCH(o.dim, o.k, {
    dim3 gr = grid_for<D_, K_>(o.geo.ncells);
    if (stage == 1)
      k_rhs_u<D_, K_, 1, 2><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, cf, o.dt, iMa2, iFr2, o.N, dst);
    else if (stage == 2)
      k_rhs_u<D_, K_, 2, 3><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, cf, o.dt, iMa2, iFr2, o.N, dst);
    else
      k_rhs_u<D_, K_, 3, 3><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, cf, o.dt, iMa2, iFr2, o.N, dst);
  });
}

void launch_rhs_p(const OpCtx& o, int stage, const double* const* src, double* dst) {
  const TableauC& t = o.tab;
  const Srcs S = make_srcs(src, stage == 1 ? 5 : (stage == 2 ? 8 : 11));
  E3Coef cf{};
  const double Ma2 = o.Ma * o.Ma, iFr2 = 1.0 / (o.Fr * o.Fr);
  if (stage == 1) {
    cf.w[0] = t.a21; cf.wt[0] = t.at21;
    cf.gw[0] = t.at21; cf.gw[1] = t.at22;
  } else if (stage == 2) {
    cf.w[0] = t.a31; cf.wt[0] = t.at31; cf.w[1] = t.a32; cf.wt[1] = t.at32;
    cf.gw[0] = t.at31; cf.gw[1] = t.at32;
    cf.gw[2] = t.at33;
  } else {
    cf.w[0] = t.b1; cf.wt[0] = t.b1; cf.w[1] = t.b2; cf.wt[1] = t.b2; cf.w[2] = t.b3; cf.wt[2] = t.b3;
    cf.gw[0] = t.b1; cf.gw[1] = t.b2;
    cf.gw[2] = t.b3;
  }
  AFL_DISPATCH(o.dim, o.k, {
    dim3 gr = grid_for<D_, K_>(o.geo.ncells);
    if (stage == 1)
      k_rhs_p<D_, K_, 1, 2, 1><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, cf, o.dt, Ma2, iFr2, o.N, dst);
    else if (stage == 2)
      k_rhs_p<D_, K_, 2, 3, 2><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, cf, o.dt, Ma2, iFr2, o.N, dst);
    else
      k_rhs_p<D_, K_, 3, 3, 3><<<gr, dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, S, cf, o.dt, Ma2, iFr2, o.N, dst);
  });
}

void launch_mass_apply(const OpCtx& o, const double* rho, double scale, int ncomp, const double* src, double* dst) {
  AFL_DISPATCH(o.dim, o.k, {
    k_mass_apply<D_, K_><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, rho, scale, ncomp, o.N, src, dst);
  });
}

void launch_inv_diag(const OpCtx& o, const double* rho, double scale, int ncomp, double* dst) {
  AFL_DISPATCH(o.dim, o.k, {
    k_inv_diag<D_, K_><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, rho, scale, ncomp, o.N, dst);
  });
}

void launch_apply_B(const OpCtx& o, const double* p, double* dst) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double f = coeff * o.dt / (o.Ma * o.Ma);
  AFL_DISPATCH(o.dim, o.k, {
    k_apply_B<D_, K_><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, f, o.N, p, dst);
  });
}

void launch_apply_C(const OpCtx& o, const double* pfix, const double* w, const double* pD, double* dst) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double cdt = coeff * o.dt;
  AFL_DISPATCH(o.dim, o.k, {
    if (pD)
      k_apply_C<D_, K_, true><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, cdt, o.N, pfix, w, pD, 1.0 / (kGamma - 1.0), dst);
    else
      k_apply_C<D_, K_, false><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, cdt, o.N, pfix, w, nullptr, 0.0, dst);
  });
}

void launch_mass_solve(const OpCtx& o, const double* rho, double scale, int ncomp, const double* b, double* x) {
  const double tol2 = o.cell_tol * o.cell_tol;
  AFL_DISPATCH(o.dim, o.k, {
    if (ncomp == 1)
      k_mass_solve<D_, K_, 1><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, rho, scale, o.N, b, x, tol2, o.cell_maxit);
    else
      k_mass_solve<D_, K_, D_><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, rho, scale, o.N, b, x, tol2, o.cell_maxit);
  });
}

void launch_schur_w(const OpCtx& o, const double* rho, const double* p, double* w) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double f = coeff * o.dt / (o.Ma * o.Ma);
  const double tol2 = o.cell_tol * o.cell_tol;
  AFL_DISPATCH(o.dim, o.k, {
    k_schur_w<D_, K_><<<grid_for<D_, K_>(o.geo.ncells), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.T, rho, f, o.N, p, w, tol2, o.cell_maxit);
  });
}

}  // namespace afl
