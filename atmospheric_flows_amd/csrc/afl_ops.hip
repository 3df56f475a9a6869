// Operator kernels of the IMEX-DG hot path for gfx950. Each kernel is a fused cell + face
// (+ wall) loop over one batch of cells per block (see afl_common.h for the decomposition).
// Weak forms follow euler_operator.h; every kernel cites the reference assembler it replaces.
#include "afl_internal.h"

// Tuning of the FP32 multigrid level kernels (2D): minimum waves per SIMD the compiler must allow
// (register budget) and whether the C apply evaluates its fields in one batched pass (fewer
// wavefront syncs) or one at a time (half the LDS per cell). At 512x100, k = 4 the defaults (no
// register cap, batched) are within 2 % of every combination of 6 or 8 waves and unbatched
// evaluation: these kernels are not occupancy-bound.
#ifndef AFL_WLVL_WAVES
#define AFL_WLVL_WAVES 1
#endif
#ifndef AFL_CHEB_WAVES
#define AFL_CHEB_WAVES 1
#endif
#ifndef AFL_C_BATCH_F32
#define AFL_C_BATCH_F32 1
#endif
#ifndef AFL_E2_WAVES
#define AFL_E2_WAVES 2
#endif
// 3D FP32 level C apply + Chebyshev step at k <= 3: waves per SIMD the register budget must allow
// (4: 128 VGPRs with a 28-byte spill at k = 3, level C 521 -> 503 us at 96x96x40; at k = 4 the
// cap spills 68 bytes and gains nothing)
#ifndef AFL_CHEB_WAVES_3D
#define AFL_CHEB_WAVES_3D 4
#endif

namespace afl {


template <int DIM, int K, int Q = 2 * K + 1>
struct Lds {
  using C = Cfg<DIM, K, Q>;
  static constexpr int EV = C::EV1 + C::EV2;
  static constexpr int IN = C::IN1 + C::IN2;
  static constexpr int W = EV > IN ? EV : IN;   // eval / integrate work area per cell
  static constexpr int F = 2 * DIM * C::NFQ;    // face flux block per cell and component
  static constexpr int FN = FaceLds<DIM, K>::FN;
  static constexpr int WF = W + FN;             // work area + one field's face nodes
};

#define AFL_SLOT_PTRS                                                      \
  using C = Cfg<DIM, K>;                                                   \
  constexpr int NQ = C::NQ, N1 = C::N1;                                    \
  (void)N1;                                                                \
  Slot<DIM, K> s(g);                                                       \
  const int cbs = s.cb < C::CPB ? s.cb : 0;                                \
  double* t1 = shw + cbs * Lds<DIM, K>::WF;                                \
  double* t2 = t1 + C::EV1;                                                \
  double* i1 = t1;                                                         \
  double* i2 = t1 + C::IN1;                                                \
  double* fn = t1 + Lds<DIM, K>::W;                                        \
  (void)t2; (void)i2; (void)fn;

// Face geometry of the thread's face point on all 2*dim faces (outward normals, JxW).
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
struct FacePts {
  R n[2 * DIM][DIM];
  R jxw[2 * DIM];
  __device__ FacePts(const Geo& g, const TabT<R>& T, const Slot<DIM, K, Q>& s) {
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      R nn[3];
      face_geo<DIM, K, Q>(g, T, s.ic, f, s.l, nn, jxw[f]);
#pragma unroll
      for (int d = 0; d < DIM; ++d) n[f][d] = nn[d];
    }
  }
};

// ---------------------------------------------------------------------------------------
// E1: vmult_rhs_rho_update (euler_operator.h:342-473 cell, 476-692 faces, boundary none).
// src = {rho_j, u_j} for the NS explicit states; w = stage weights (a21 | a31,a32 | b1,b2,b3).
template <int DIM, int K, int NS>
__global__ __launch_bounds__(AFL_NT) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 2 : 1))) void k_rhs_rho(Geo g, const Tab* __restrict__ Tp, Srcs src, double dt, double w0, double w1,
                                                     double w2, int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  const double wj[3] = {w0, w1, w2};
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double V[NQ], G[DIM][NQ], F[2 * DIM];
#pragma unroll
  for (int q = 0; q < NQ; ++q) { V[q] = 0.0; for (int d = 0; d < DIM; ++d) G[d][q] = 0.0; }
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) F[f] = 0.0;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double r[NQ], trr[2 * DIM][2], tru[DIM][2 * DIM][2];
    eval_line_faces<DIM, K>(src.p[2 * j], s, nb, t1, t2, fn, T, r, trr);
    if (j == 0) for (int q = 0; q < NQ; ++q) V[q] = r[q];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double u[NQ];
      eval_line_faces<DIM, K>(src.p[2 * j + 1] + d * N, s, nb, t1, t2, fn, T, u, tru[d]);
#pragma unroll
      for (int q = 0; q < NQ; ++q) G[d][q] += wj[j] * dt * r[q] * u[q];
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      if (nb.c[f] < 0) continue;
      double nfv[3], jfv;  // face geometry recomputed here: fewer live registers
      face_geo<DIM, K>(g, T, s.ic, f, s.l, nfv, jfv);
      const double rp = trr[f][0], rm = trr[f][1];
      double avg = 0.0, unp = 0.0, unm = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        avg += 0.5 * (rp * tru[d][f][0] + rm * tru[d][f][1]) * nfv[d];
        unp += tru[d][f][0] * nfv[d];
        unm += tru[d][f][1] * nfv[d];
      }
      const double lam = fmax(fabs(unp), fabs(unm));
      F[f] += wj[j] * dt * (avg + 0.5 * lam * (rp - rm));
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
  double* Fq = shF + cbs * Lds<DIM, K>::F;
  if (s.act)
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double nfv[3], jfv;
      face_geo<DIM, K>(g, T, s.ic, f, s.l, nfv, jfv);
      Fq[f * C::NFQ + s.l] = -F[f] * jfv;
    }
  cell_sync<DIM, K>();
  double out[N1];
  integrate_line<DIM, K, true, true>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst, s, out);
}

// ---------------------------------------------------------------------------------------
// E2: vmult_rhs_velocity_fixed (euler_operator.h:1620-1886 cell, 1889-2184 faces, 2187-2305 walls).
// Explicit states j < NE: (rho, u, p) = src[3j..3j+2], advective weight w[j], pressure weight wp[j].
// Gravity densities: src[3j] with weight gw[j], j < NG (rho_old, rho_tmp_2, rho_tmp_3 / rho_curr).

template <int DIM, int K, int NE, int NG>
__global__ __launch_bounds__(AFL_NT) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? AFL_E2_WAVES : 1))) void k_rhs_u(Geo g, const Tab* __restrict__ Tp, Srcs src, E2Coef cf, double dt, double iMa2,
                                                   double iFr2, int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F * DIM];
  AFL_SLOT_PTRS
  constexpr int NSYM = DIM * (DIM + 1) / 2;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  // Sym: symmetric flux tensor rho u (x) u with the pressure term on its diagonal
  double V[DIM][NQ], Sym[NSYM][NQ], F[2 * DIM][DIM];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    for (int d = 0; d < DIM; ++d) V[d][q] = 0.0;
    for (int d = 0; d < NSYM; ++d) Sym[d][q] = 0.0;
  }
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f)
    for (int a = 0; a < DIM; ++a) F[f][a] = 0.0;
  // faces first, from the traces alone, then the cells from the values alone: the two sets of
  // a state's fields are never live together (2D: state loops kept rolled)
#pragma unroll(DIM == 2 ? 1 : NE)
  for (int j = 0; j < NE; ++j) {
    const double wj = pick<3>(cf.w, j), wpj = pick<3>(cf.wp, j);
    double trr[2 * DIM][2], tru[DIM][2 * DIM][2], trp[2 * DIM][2];
    trace_line_faces<DIM, K>(src_at(src, 3 * j), s, nb, fn, T, trr);
#pragma unroll
    for (int d = 0; d < DIM; ++d) trace_line_faces<DIM, K>(src_at(src, 3 * j + 1) + d * N, s, nb, fn, T, tru[d]);
    trace_line_faces<DIM, K>(src_at(src, 3 * j + 2), s, nb, fn, T, trp);
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double nfv[3], jfv;  // face geometry recomputed here: fewer live registers
      face_geo<DIM, K>(g, T, s.ic, f, s.l, nfv, jfv);
      const double pp = trp[f][0];
      if (nb.c[f] >= 0) {
        const double rp = trr[f][0], rm = trr[f][1], pm = trp[f][1];
        double unp = 0.0, unm = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) { unp += tru[d][f][0] * nfv[d]; unm += tru[d][f][1] * nfv[d]; }
        const double lam = fmax(fabs(unp), fabs(unm));
        const double avg_p = 0.5 * (pp + pm);
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          const double avg_tn = 0.5 * (rp * tru[a][f][0] * unp + rm * tru[a][f][1] * unm);
          const double jump = rp * tru[a][f][0] - rm * tru[a][f][1];
          F[f][a] += wj * dt * avg_tn + wpj * dt * iMa2 * avg_p * nfv[a] + wj * dt * 0.5 * lam * jump;
        }
      } else {
        // wall: p_D = p, flux (p + p_D)/2 n (euler_operator.h:2213-2218)
#pragma unroll
        for (int a = 0; a < DIM; ++a) F[f][a] += wpj * dt * iMa2 * (0.5 * (pp + pp)) * nfv[a];
      }
    }
  }
#pragma unroll(DIM == 2 ? 1 : NE)
  for (int j = 0; j < NE; ++j) {
    const double wj = pick<3>(cf.w, j), wpj = pick<3>(cf.wp, j);
    double r[NQ], u[DIM][NQ], pj[NQ];
    eval_line<DIM, K>(src_at(src, 3 * j), s, t1, t2, T, r);
#pragma unroll
    for (int d = 0; d < DIM; ++d) eval_line<DIM, K>(src_at(src, 3 * j + 1) + d * N, s, t1, t2, T, u[d]);
    eval_line<DIM, K>(src_at(src, 3 * j + 2), s, t1, t2, T, pj);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (j == 0) for (int d = 0; d < DIM; ++d) V[d][q] = r[q] * u[d][q];
      const double pw = wpj * dt * iMa2 * pj[q];
      int m = 0;
      for (int a = 0; a < DIM; ++a)
        for (int b = a; b < DIM; ++b) Sym[m++][q] += wj * dt * r[q] * u[a][q] * u[b][q] + (a == b ? pw : 0.0);
    }
  }
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    double r[NQ];
    eval_line<DIM, K>(src.p[3 * j], s, t1, t2, T, r);  // gravity densities: rho_old, rho_tmp_2, rho_tmp_3/curr
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[DIM - 1][q] -= cf.gw[j] * dt * iFr2 * r[q];
  }
  double* Fq = shF + cbs * Lds<DIM, K>::F * DIM;
  if (s.act)
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double nfv[3], jfv;
      face_geo<DIM, K>(g, T, s.ic, f, s.l, nfv, jfv);
#pragma unroll
      for (int a = 0; a < DIM; ++a) Fq[(a * 2 * DIM + f) * C::NFQ + s.l] = -F[f][a] * jfv;
    }
  cell_sync<DIM, K>();
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    double Va[NQ], Ga[DIM][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double Gp[3] = {0, 0, 0}, R[3];
      for (int b = 0; b < DIM; ++b) {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const int m = lo * DIM - lo * (lo - 1) / 2 + (hi - lo);
        Gp[b] = Sym[m][q];
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
// FULL = false (implicit stages only) leaves out the two terms that read u_fixed, the kinetic
// correction -1/2 Ma^2 rho_new |u_fix|^2 and the gravity work of (rho_new, u_fix): the rest does not
// change across the fixed-point iterations of a stage, and k_rhs_p_ufix adds those two terms.

template <int DIM, int K, int NE, int NG, int STAGE, bool FULL = true>
__global__ __launch_bounds__(AFL_NT) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 2 : 1))) void k_rhs_p(Geo g, const Tab* __restrict__ Tp, Srcs src, E3Coef cf, double dt, double Ma2,
                                                   double iFr2, int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  constexpr double gm1 = kGamma - 1.0;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double V[NQ], G[DIM][NQ], F[2 * DIM];
  double tu0[DIM][2 * DIM][2];  // u_old traces (stage-2 kinetic jump as written)
#pragma unroll
  for (int q = 0; q < NQ; ++q) { V[q] = 0.0; for (int d = 0; d < DIM; ++d) G[d][q] = 0.0; }
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) F[f] = 0.0;
  // faces first, from the traces alone, then the cells from the values alone: the two sets of
  // a state's fields are never live together (2D: state loops kept rolled)
#pragma unroll(DIM == 2 ? 1 : NE)
  for (int j = 0; j < NE; ++j) {
    const double wj = pick<3>(cf.w, j), wpj = pick<3>(cf.wt, j);
    double trr[2 * DIM][2], tru[DIM][2 * DIM][2], trp[2 * DIM][2];
    trace_line_faces<DIM, K>(src_at(src, 3 * j), s, nb, fn, T, trr);
#pragma unroll
    for (int d = 0; d < DIM; ++d) trace_line_faces<DIM, K>(src_at(src, 3 * j + 1) + d * N, s, nb, fn, T, tru[d]);
    trace_line_faces<DIM, K>(src_at(src, 3 * j + 2), s, nb, fn, T, trp);
    if (STAGE == 2 && j == 0)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) { tu0[d][f][0] = tru[d][f][0]; tu0[d][f][1] = tru[d][f][1]; }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      if (nb.c[f] < 0) continue;
      double nfv[3], jfv;  // face geometry recomputed here: fewer live registers
      face_geo<DIM, K>(g, T, s.ic, f, s.l, nfv, jfv);
      const double rp = trr[f][0], rm = trr[f][1], pp = trp[f][0], pm = trp[f][1];
      double uup = 0.0, uum = 0.0, unp = 0.0, unm = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        uup += tru[d][f][0] * tru[d][f][0]; uum += tru[d][f][1] * tru[d][f][1];
        unp += tru[d][f][0] * nfv[d]; unm += tru[d][f][1] * nfv[d];
      }
      const double Ep = 1.0 / gm1 * pp / rp + 0.5 * Ma2 * uup;
      const double Em = 1.0 / gm1 * pm / rm + 0.5 * Ma2 * uum;
      double avg_kin = 0.0, avg_ent = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        avg_kin += 0.5 * (0.5 * uup * rp * tru[d][f][0] + 0.5 * uum * rm * tru[d][f][1]) * nfv[d];
        avg_ent += 0.5 * (((Ep - 0.5 * Ma2 * uup) * rp + pp) * tru[d][f][0] +
                          ((Em - 0.5 * Ma2 * uum) * rm + pm) * tru[d][f][1]) * nfv[d];
      }
      const double lam = fmax(fabs(unp), fabs(unm));
      if (STAGE == 3) {
        const double jump = rp * Ep - rm * Em;
        F[f] += wj * dt * Ma2 * avg_kin + wj * dt * avg_ent + wj * dt * 0.5 * lam * jump;
      } else {
        double jk;
        if (STAGE == 2 && j == 1) {  // euler_operator.h:1182-1183, u_tmp_2 . u_old as written
          double d0p = 0.0, d0m = 0.0;
#pragma unroll
          for (int d = 0; d < DIM; ++d) { d0p += tru[d][f][0] * tu0[d][f][0]; d0m += tru[d][f][1] * tu0[d][f][1]; }
          jk = rp * 0.5 * d0p - rm * 0.5 * d0m;
        } else {
          jk = rp * 0.5 * uup - rm * 0.5 * uum;
        }
        const double je = rp * (Ep - 0.5 * Ma2 * uup) - rm * (Em - 0.5 * Ma2 * uum);
        F[f] += wj * dt * Ma2 * (avg_kin + 0.5 * lam * jk) + wpj * dt * (avg_ent + 0.5 * lam * je);
      }
    }
  }
#pragma unroll(DIM == 2 ? 1 : NE)
  for (int j = 0; j < NE; ++j) {
    const double wj = pick<3>(cf.w, j), wpj = pick<3>(cf.wt, j);
    double r[NQ], u[DIM][NQ], p[NQ];
    eval_line<DIM, K>(src_at(src, 3 * j), s, t1, t2, T, r);
#pragma unroll
    for (int d = 0; d < DIM; ++d) eval_line<DIM, K>(src_at(src, 3 * j + 1) + d * N, s, t1, t2, T, u[d]);
    eval_line<DIM, K>(src_at(src, 3 * j + 2), s, t1, t2, T, p);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double uu = 0.0;
      for (int d = 0; d < DIM; ++d) uu += u[d][q] * u[d][q];
      const double E = 1.0 / gm1 * (p[q] / r[q]) + 0.5 * Ma2 * uu;
      if (j == 0) V[q] += r[q] * E;
      const double kin = 0.5 * wj * dt * Ma2 * uu * r[q];
      const double ent = wpj * dt * (r[q] * (E - 0.5 * Ma2 * uu) + p[q]);
      for (int d = 0; d < DIM; ++d) G[d][q] += (kin + ent) * u[d][q];
    }
  }
  if (FULL || STAGE == 3) {
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
    if (!FULL && STAGE < 3 && j == NG - 1) continue;  // (rho_new, u_fix): k_rhs_p_ufix
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
  if (s.act)
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double nfv[3], jfv;
      face_geo<DIM, K>(g, T, s.ic, f, s.l, nfv, jfv);
      Fq[f * C::NFQ + s.l] = -F[f] * jfv;
    }
  cell_sync<DIM, K>();
  double out[N1];
  integrate_line<DIM, K, true, true>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst, s, out);
}

// The u_fixed terms of E3 at an implicit stage (euler_operator.h:743-1007, the kinetic-energy
// correction and the gravity work of the fixed-point velocity): dst = base + int psi V with
// V = -1/2 Ma^2 rho_new |u_fix|^2 - cg rho_new u_fix,vertical, cg = (gravity weight) dt Ma^2 / Fr^2.
// Cell-local (no faces); base = the u_fixed-free part of E3 from k_rhs_p<..., FULL = false>.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_rhs_p_ufix(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ rho,
                                                     const double* __restrict__ u, double half_ma2, double cg, int64_t N,
                                                     const double* __restrict__ base, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double r[NQ], V[1][NQ];
  eval_line<DIM, K>(rho, s, t1, t2, T, r);
#pragma unroll
  for (int q = 0; q < NQ; ++q) V[0][q] = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double w[NQ];
    eval_line<DIM, K>(u + d * N, s, t1, t2, T, w);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[0][q] += w[q] * w[q] * half_ma2 + (d == DIM - 1 ? cg * w[q] : 0.0);
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) V[0][q] = -r[q] * V[0][q] * lg.jxw(g, T, q);
  double out[1][N1], b[N1];
  load_row<DIM, K>(base, s, b);
  integrate_lines_v<DIM, K, 1>(s, i1, i2, T, V, out);
#pragma unroll
  for (int i = 0; i < N1; ++i) out[0][i] += b[i];
  store_row<DIM, K>(dst, s, out[0]);
}

// ---------------------------------------------------------------------------------------
// Mass-type apply: dst = scale * M_rho src, M (euler_operator.h:715-740), D (:1384-1408),
// A (:2328-2357). rho == nullptr means rho = 1.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_mass_apply(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ rho, double scale,
                                                        int ncomp, int64_t N, const double* __restrict__ src,
                                                        double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
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
__global__ __launch_bounds__(AFL_NT) void k_inv_diag(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ rho, double scale,
                                                      int ncomp, int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
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
    cell_sync<DIM, K>();
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
    cell_sync<DIM, K>();
    if (s.act && s.l < NQ * N1) {
      const int qx = s.l % NQ, iz = s.l / NQ;
      for (int i = 0; i < N1; ++i) {
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) a += T.S[q][i] * T.S[q][i] * i1[(iz * NQ + q) * NQ + qx];
        i2[(iz * N1 + i) * NQ + qx] = a;
      }
    }
    cell_sync<DIM, K>();
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
// (own side), walls +f p n . v, with f = coeff dt / Ma^2. fn: face-node LDS of this cell.
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void apply_B_rows(const Geo& g, const TabT<R>& T, const Slot<DIM, K, Q>& s, const LineGeo<DIM, K, R, Q>& lg,
                                    const Nbrs<DIM>& nb, const FacePts<DIM, K, R, Q>& fp, const R* __restrict__ p,
                                    R f, R* t1, R* t2, R* i1, R* i2, R* fn, R* Fq, R (&out)[DIM][K + 1]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int NQ = C::NQ;
  R pq[NQ], trp[2 * DIM][2];
  eval_line_faces<DIM, K, R>(p, s, nb, t1, t2, fn, T, pq, trp);
  // one face-flux block (Lds::F values per cell), refilled per component: the cell's LDS stays
  // small enough for one more block per CU in the level operators
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if (a > 0) cell_sync<DIM, K, Q>();  // the previous component's add_faces reads are done
    if (s.act) {
#pragma unroll
      for (int fc = 0; fc < 2 * DIM; ++fc) {
        const R avg = R(0.5) * (trp[fc][0] + trp[fc][1]);  // walls: trp[fc][1] == trp[fc][0] (p_D = p)
        Fq[fc * C::NFQ + s.l] = f * avg * fp.n[fc][a] * fp.jxw[fc];
      }
    }
    cell_sync<DIM, K, Q>();
    R V[NQ], G[DIM][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      R Gp[3] = {0, 0, 0}, Rg[3];
      Gp[a] = -f * pq[q];
      lg.gref(g, T, s.ic, q, Gp, Rg);
      for (int d = 0; d < DIM; ++d) G[d][q] = Rg[d];
    }
    integrate_line<DIM, K, false, true, R>(s, i1, i2, T, V, G, out[a]);
    add_faces<DIM, K, R>(s, Fq, T, out[a], 0xffu);
  }
}

template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_apply_B(Geo g, const Tab* __restrict__ Tp, double f, int64_t N, const double* __restrict__ p,
                                                     double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  const FacePts<DIM, K> fp(g, T, s);
  double out[DIM][N1];
  apply_B_rows<DIM, K, double>(g, T, s, lg, nb, fp, p, f, t1, t2, i1, i2, fn, shF + cbs * Lds<DIM, K>::F, out);
#pragma unroll
  for (int a = 0; a < DIM; ++a) store_row<DIM, K>(dst + a * N, s, out[a]);
}

// One component A of B p (3D, k = 4: a launch per component holds one face-flux block and one
// result row, two waves per SIMD instead of one).
template <int DIM, int K, int A>
__global__ __launch_bounds__(AFL_NT) void k_apply_B1(Geo g, const Tab* __restrict__ Tp, double f, int64_t N,
                                                      const double* __restrict__ p, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * Lds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * Lds<DIM, K>::F];
  AFL_SLOT_PTRS
  double* Fq = shF + cbs * Lds<DIM, K>::F;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  const FacePts<DIM, K> fp(g, T, s);
  double pq[NQ], trp[2 * DIM][2];
  eval_line_faces<DIM, K>(p, s, nb, t1, t2, fn, T, pq, trp);
  if (s.act) {
#pragma unroll
    for (int fc = 0; fc < 2 * DIM; ++fc)
      Fq[fc * C::NFQ + s.l] = f * 0.5 * (trp[fc][0] + trp[fc][1]) * fp.n[fc][A] * fp.jxw[fc];
  }
  cell_sync<DIM, K>();
  double V[NQ], G[DIM][NQ], out[N1];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double Gp[3] = {0, 0, 0}, Rg[3];
    Gp[A] = -f * pq[q];
    lg.gref(g, T, s.ic, q, Gp, Rg);
#pragma unroll
    for (int d = 0; d < DIM; ++d) G[d][q] = Rg[d];
  }
  integrate_line<DIM, K, false, true>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst + A * N, s, out);
}

// ---------------------------------------------------------------------------------------
// C: vmult_enthalpy (euler_operator.h:1411-1507); nonlinear in w through lambda = max|w.n|.
// LDS per cell of the C apply. FP32 (BATCH): all fields (p_fix, w, and p for the D term) are
// evaluated in one batched pass, so the work area holds every field's sweep and face nodes.
template <int DIM, int K, bool WITH_D, bool BATCH, int Q = 2 * K + 1>
struct LdsC {
  using C = Cfg<DIM, K, Q>;
  static constexpr int NF = 1 + DIM;                     // batched: p_fix, w_0..w_{d-1} (p apart)
  static constexpr int NFF = 1 + DIM;                    // fields needed on the faces
  static constexpr int EV = BATCH ? NF * (C::EV1 + C::EV2) : C::EV1 + C::EV2;
  static constexpr int IN = C::IN1 + C::IN2;
  static constexpr int W = EV > IN ? EV : IN;
  static constexpr int WF0 = W + (BATCH ? NFF : 1) * FaceLds<DIM, K>::FN;
  // the face-flux block is written once the evaluations are done: inside the work area after the
  // integration scratch when it fits there, else appended
  static constexpr int F = 2 * DIM * C::NFQ;
  static constexpr int FOFF = IN + F <= WF0 ? IN : WF0;
  static constexpr int WF = IN + F <= WF0 ? WF0 : WF0 + F;
};

// Result rows of C (optionally fused with D p) for the thread's node row: out = (WITH_D ?
// scaleD M pD - C w : C w). sh: the cell's work area (LdsC::WF), Fq: its face flux block (Lds::F).
template <int DIM, int K, bool WITH_D, typename R, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void apply_C_rows(const Geo& g, const TabT<R>& T, const Slot<DIM, K, Q>& s, R* sh, R* Fq, R cdt,
                                    const R* __restrict__ pfix, const R* __restrict__ w, int64_t N,
                                    const R* __restrict__ pD, R scaleD, R (&out)[K + 1],
                                    const R* rows_a = nullptr, const R* rows_b = nullptr, int sel = -1) {
  using C = Cfg<DIM, K, Q>;
  constexpr int NQ = C::NQ;
  constexpr bool BATCH = sizeof(R) == 4 && AFL_C_BATCH_F32 && DIM == 2;
  using LC = LdsC<DIM, K, WITH_D, BATCH, Q>;
  R* t1 = sh;
  R* i1 = t1;
  R* i2 = t1 + C::IN1;
  const R gr = R(kGamma / (kGamma - 1.0)), igm1 = R(1.0 / (kGamma - 1.0)), half = R(0.5);
  const R sgn = WITH_D ? R(-1) : R(1);
  LineGeo<DIM, K, R, Q> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  const FacePts<DIM, K, R, Q> fp(g, T, s);
  R pq[NQ], G[DIM][NQ], V[NQ];
  R trp[2 * DIM][2], trw[DIM][2 * DIM][2];
  // pD's node row: loaded ahead of the evaluations' barriers, or (sel = 0 / 1) the caller's
  // rows_a / rows_b, already loaded, picked where it is used
  R prow[K + 1];
  if (WITH_D && sel < 0) load_row<DIM, K, R>(pD, s, prow);
  auto pick_prow = [&]() {
    if (sel >= 0)
#pragma unroll
      for (int i = 0; i <= K; ++i) prow[i] = sel == 0 ? rows_a[i] : rows_b[i];
  };
  if constexpr (BATCH) {
    const R* fl[LC::NF];
    fl[0] = pfix;
#pragma unroll
    for (int d = 0; d < DIM; ++d) fl[1 + d] = w + d * N;
    R v[LC::NF][NQ], tr[LC::NFF][2 * DIM][2];
    eval_lines_faces<DIM, K, LC::NF, LC::NFF, R>(fl, s, nb, t1, t1 + LC::NF * C::EV1, sh + LC::W, T, v, tr);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      pq[q] = v[0][q];
#pragma unroll
      for (int d = 0; d < DIM; ++d) G[d][q] = sgn * (-cdt * gr * v[0][q] * v[1 + d][q]);
    }
    if (WITH_D) {
      pick_prow();
      eval_line_row<DIM, K, R>(prow, s, t1, t1 + C::EV1, T, V);
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] *= scaleD * lg.jxw(g, T, q);
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        trp[f][sd] = tr[0][f][sd];
#pragma unroll
        for (int d = 0; d < DIM; ++d) trw[d][f][sd] = tr[1 + d][f][sd];
      }
  } else {
    R* t2 = t1 + C::EV1;
    R* fn = t1 + LC::W;
    eval_line_faces<DIM, K, R>(pfix, s, nb, t1, t2, fn, T, pq, trp);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      R wq[NQ];
      eval_line_faces<DIM, K, R>(w + d * N, s, nb, t1, t2, fn, T, wq, trw[d]);
#pragma unroll
      for (int q = 0; q < NQ; ++q) G[d][q] = sgn * (-cdt * gr * pq[q] * wq[q]);
    }
    if (WITH_D) {
      pick_prow();
      eval_line_row<DIM, K, R>(prow, s, t1, t2, T, V);
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] *= scaleD * lg.jxw(g, T, q);
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    R Gp[3] = {0, 0, 0}, Rg[3];
    for (int d = 0; d < DIM; ++d) Gp[d] = G[d][q];
    lg.gref(g, T, s.ic, q, Gp, Rg);
    for (int d = 0; d < DIM; ++d) G[d][q] = Rg[d];
  }
  if (s.act) {
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      R F = R(0);
      if (nb.c[f] >= 0) {
        const R pp = trp[f][0], pm = trp[f][1];
        R avg = R(0), wnp = R(0), wnm = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          avg += half * gr * (pp * trw[d][f][0] + pm * trw[d][f][1]) * fp.n[f][d];
          wnp += trw[d][f][0] * fp.n[f][d];
          wnm += trw[d][f][1] * fp.n[f][d];
        }
        const R lam = fmax(fabs(wnp), fabs(wnm));
        F = sgn * cdt * (avg + half * lam * igm1 * (pp - pm)) * fp.jxw[f];
      }
      Fq[f * C::NFQ + s.l] = F;
    }
  }
  cell_sync<DIM, K, Q>();
  integrate_line<DIM, K, WITH_D, true, R>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K, R>(s, Fq, T, out, 0xffu);
}

// C: vmult_enthalpy (euler_operator.h:1411-1507); nonlinear in w through lambda = max|w.n|.
// Optionally fused with D p (Schur second half): dst = scaleD * M p - C w.
template <int DIM, int K, bool WITH_D, typename R = double, int Q = 2 * K + 1>
__global__ __launch_bounds__(AFL_NT) void k_apply_C(Geo g, const TabT<R>* __restrict__ Tp, R cdt, int64_t N, const R* __restrict__ pfix,
                                                     const R* __restrict__ w, const R* __restrict__ pD, R scaleD,
                                                     R* __restrict__ dst) {
  const auto& T = *Tp;
  using LC = LdsC<DIM, K, WITH_D, sizeof(R) == 4 && AFL_C_BATCH_F32 && DIM == 2, Q>;
  __shared__ R shw[Cfg<DIM, K, Q>::CPB * LC::WF];
  Slot<DIM, K, Q> s(g);
  const int cbs = s.cb < Cfg<DIM, K, Q>::CPB ? s.cb : 0;
  R out[K + 1];
  R* sh = shw + cbs * LC::WF;
  apply_C_rows<DIM, K, WITH_D, R>(g, T, s, sh, sh + LC::FOFF, cdt, pfix, w, N, pD, scaleD, out);
  store_row<DIM, K, R>(dst, s, out);
}

// Schur second half fused with one Chebyshev smoothing step of the multigrid levels:
//   t = S pD (= D pD - C w);  r' = rin - t -> rout;  z = (gamma-1)/Jbar (MI x MI [x MI]) r';
//   d = c1 d + c2 z;  x += d.
// The block preconditioner is cell-local (LDS transposes of the cell's rows); pD, d and x are
// read only inside the thread's own cell, so d and x (and pD == d or x) update in place.
template <int DIM, int K, typename R, int Q = 2 * K + 1>
__global__ __launch_bounds__(AFL_NT) __attribute__((amdgpu_waves_per_eu(DIM == 2 && sizeof(R) == 4 ? AFL_CHEB_WAVES : (DIM == 3 && K <= 3 && sizeof(R) == 4 ? AFL_CHEB_WAVES_3D : 1)))) void k_apply_C_cheb(Geo g, const TabT<R>* __restrict__ Tp, R cdt, int64_t N, const R* __restrict__ pfix,
                                                          const R* __restrict__ w, const R* pD, R scaleD,
                                                          const R* __restrict__ jbar, const R* __restrict__ rin,
                                                          R* __restrict__ rout, R c1, R c2, R* d, R* x) {
  const auto& T = *Tp;
  using LC = LdsC<DIM, K, true, sizeof(R) == 4 && AFL_C_BATCH_F32 && DIM == 2, Q>;
  __shared__ R shw[Cfg<DIM, K, Q>::CPB * LC::WF];
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NPC = C::NPC, ROWS = C::ROWS;
  Slot<DIM, K, Q> s(g);
  const int cbs = s.cb < C::CPB ? s.cb : 0;
  R* sh = shw + cbs * LC::WF;
  static_assert(2 * C::NPC <= LC::FOFF || LC::FOFF >= LC::W, "epilogue scratch must not overlap the face fluxes");
  const bool ro = s.act && s.l < ROWS;
  const int64_t base = (s.act ? s.cell : 0) * NPC + s.l * N1;
  // the epilogue's inputs are loaded first, so their latency overlaps the C apply (d and x are
  // read and written only by their row's thread, after every read of pD == d or x)
  const R jb = s.act ? jbar[s.cell] : R(1);
  R rr[N1], dr[N1], xr[N1];
  load_row<DIM, K, R>(rin, s, rr);
  load_row<DIM, K, R>(d, s, dr);
  load_row<DIM, K, R>(x, s, xr);
  // pD is d or x in the Chebyshev recursion: its rows are then already loaded
  const int sel = pD == d ? 0 : (pD == x ? 1 : -1);
  R t[N1];
  apply_C_rows<DIM, K, true, R>(g, T, s, sh, sh + LC::FOFF, cdt, pfix, w, N, pD, scaleD, t, dr, xr, sel);
  cell_sync<DIM, K, Q>();  // the face fluxes and work area are read until here; the epilogue reuses them
  R* tmp = sh;
  R* tmp2 = sh + NPC;
  if (ro) {
    R r[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      r[i] = rr[i] - t[i];
      rout[base + i] = r[i];
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      R a = R(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) a += T.MI[i][j] * r[j];
      tmp[s.l * N1 + i] = a;
    }
  }
  cell_sync<DIM, K, Q>();
  const R* zr;
  if (DIM == 2) {
    if (s.act && s.l < N1) {
      R v[N1];
#pragma unroll
      for (int j = 0; j < N1; ++j) v[j] = tmp[j * N1 + s.l];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = R(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) a += T.MI[i][j] * v[j];
        tmp2[i * N1 + s.l] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    zr = tmp2;
  } else {
    if (s.act && s.l < N1 * N1) {
      const int ix = s.l % N1, iz = s.l / N1;
      R v[N1];
#pragma unroll
      for (int j = 0; j < N1; ++j) v[j] = tmp[(iz * N1 + j) * N1 + ix];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = R(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) a += T.MI[i][j] * v[j];
        tmp2[(iz * N1 + i) * N1 + ix] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < N1 * N1) {
      const int ix = s.l % N1, iy = s.l / N1;
      R v[N1];
#pragma unroll
      for (int j = 0; j < N1; ++j) v[j] = tmp2[(j * N1 + iy) * N1 + ix];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = R(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) a += T.MI[i][j] * v[j];
        tmp[(i * N1 + iy) * N1 + ix] = a;
      }
    }
    cell_sync<DIM, K, Q>();
    zr = tmp;
  }
  if (ro) {
    const R zs = R(kGamma - 1.0) / jb;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const R dn = (c1 != R(0) ? c1 * dr[i] : R(0)) + c2 * zs * zr[s.l * N1 + i];
      d[base + i] = dn;
      x[base + i] = xr[i] + dn;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Cell-local solver for the block-diagonal mass-type matrices A = S^T diag(co) S (co = scale rho
// JxW at the quadrature points, S the tensor interpolation matrix): Chebyshev iteration (Saad,
// Iterative Methods, Alg. 12.1) preconditioned with the separable matrix P = S^T diag(c~) S,
// c~ = a (x) b [(x) c] / total^(dim-1) built from the marginal sums of co over the cell. P is a
// Kronecker product of dim weighted 1D mass matrices, so P^{-1} is the product of their 5x5
// inverses. The Rayleigh quotient of P^{-1} A lies in [min co/c~, max co/c~]: the interval and
// the iteration count for the requested A-norm tolerance (2 sigma^n <= tol) are known per cell
// up front, with no inner products and no convergence flags. On the terrain-following mesh the
// Jacobian factorises exactly, so mass and D solves finish in one step and density-weighted
// solves in a few. All NC components share every barrier. b rows in, x rows out.
template <int DIM, int K, int Q = 2 * K + 1>
struct SolvLds {
  using C = Cfg<DIM, K, Q>;
  static constexpr int EVF = C::EV1 + C::EV2;
  static constexpr int INF = IntV<DIM, K, Q>::I1 + IntV<DIM, K, Q>::I2;
  static constexpr int WF = EVF > INF ? EVF : INF;  // per component work area
  static constexpr int PER = WF + C::NPC;            // + the component's dofs
  // set-up scratch: co per line, line sums (3D), marginals, per-line ratio bounds
  // (3D: + the partial vertical marginals and partial bounds of the two-level reductions)
  static constexpr int SETUP = C::LT * C::NQ + (DIM == 3 ? C::LT : 0) + (DIM + 1) * C::NQ + 2 * C::LT +
                               (DIM == 3 ? C::NQ * C::NQ + 2 * C::NQ : 0);
  static constexpr int SEP = DIM * C::N1 * C::N1;    // the dim inverse 1D matrices (persist)
  static constexpr int sep_off(int nc) { return nc * PER > SETUP ? nc * PER : SETUP; }
  static constexpr int size(int nc) { return sep_off(nc) + SEP; }
};

// In-place inverse of a small SPD matrix (Gauss-Jordan without pivoting, fully unrolled).
template <int N, typename R>
__device__ __attribute__((always_inline)) inline void inv_spd(R (&A)[N][N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const R p = R(1) / A[k][k];
#pragma unroll
    for (int j = 0; j < N; ++j) A[k][j] = j == k ? p : A[k][j] * p;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == k) continue;
      const R f = A[i][k];
#pragma unroll
      for (int j = 0; j < N; ++j) A[i][j] = j == k ? -f * p : A[i][j] - f * A[k][j];
    }
  }
}

template <int DIM, int K, int NC, typename R = double, bool SEP = true, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void cell_cheb(const TabT<R>& T, const Slot<DIM, K, Q>& s, const LineGeo<DIM, K, R, Q>& lg,
                                 const R (&co)[Q], R* sl /* SolvLds::size(NC) per cell */,
                                 R tol, int maxit, const R (&b)[NC][K + 1], R (&x)[NC][K + 1]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int NQ = C::NQ, N1 = C::N1, ROWS = C::ROWS, NPC = C::NPC, LT = C::LT;
  constexpr int WF = SolvLds<DIM, K, Q>::WF;
  R* dl = sl;                 // NC*NPC dofs
  R* wk = sl + NC * NPC;      // NC*WF work
  R* t1 = wk;
  R* t2 = wk + NC * C::EV1;
  R* i1 = wk;
  R* i2 = wk + NC * IntV<DIM, K, Q>::I1;
  R* Mi = sl + SolvLds<DIM, K, Q>::sep_off(NC);  // [DIM][N1][N1]: x, (y,) z inverse 1D matrices
  const bool row_owner = s.act && s.l < ROWS;
  __shared__ int nit_blk;
  // ---- set-up: marginals of co, separable weights, ratio bounds, 1D inverses ----
  R* sco = sl;                                   // [LT][NQ]
  R* sL = sco + LT * NQ;                         // [LT] line sums (3D)
  R* smg = sL + (DIM == 3 ? LT : 0);             // [DIM][NQ] marginals a, (b,) c
  R* srm = smg + (DIM + 1) * NQ;                 // [2][LT] per-line min / max of co / c~
  R* part = srm + 2 * LT;                        // 3D: [NQ][NQ] partial vertical marginals
  R* smm = part + (DIM == 3 ? NQ * NQ : 0);      // 3D: [2][NQ] partial bounds
  if (threadIdx.x == 0) nit_blk = 1;
  R lsum = R(0);
#pragma unroll
  for (int q = 0; q < NQ; ++q) lsum += co[q];
  if (s.act) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sco[s.l * NQ + q] = co[q];
    if (DIM == 3) sL[s.l] = lsum;
  }
  cell_sync<DIM, K, Q>();
  if constexpr (DIM == 3) {
    // vertical marginal in two levels: thread l = qy NQ + q first sums the NQ lines of row qy at
    // height q (the LT-term sum of one thread is a long dependent chain)
    if (s.act) {
      const int q = s.l % NQ, qy = s.l / NQ;
      R c = R(0);
#pragma unroll
      for (int qx = 0; qx < NQ; ++qx) c += sco[(qy * NQ + qx) * NQ + q];
      part[s.l] = c;
    }
    cell_sync<DIM, K, Q>();
  }
  if (s.act && s.l < NQ) {
    // vertical marginal: sum over the lines of co at qz = l (and its reciprocal)
    R c = R(0);
    if constexpr (DIM == 3) {
#pragma unroll
      for (int qy = 0; qy < NQ; ++qy) c += part[qy * NQ + s.l];
    } else {
      for (int t = 0; t < LT; ++t) c += sco[t * NQ + s.l];
    }
    smg[(DIM - 1) * NQ + s.l] = c;
    smg[DIM * NQ + s.l] = R(1) / c;
    if (DIM == 2) {
      smg[s.l] = lsum;  // x marginal: the thread's own line sum
    } else {
      R ax = R(0), by = R(0);
#pragma unroll
      for (int t = 0; t < NQ; ++t) { ax += sL[t * NQ + s.l]; by += sL[s.l * NQ + t]; }
      smg[s.l] = ax;       // sum over qy of the lines (qx = l, qy)
      smg[NQ + s.l] = by;  // sum over qx of the lines (qx, qy = l)
    }
  }
  cell_sync<DIM, K, Q>();
  R total = R(0);
#pragma unroll
  for (int q = 0; q < NQ; ++q) total += smg[q];
  const R itot = R(1) / total;
  {
    // bounds of co / c~ on the thread's own line: c~ = prod_d marginal_d / total^(dim-1) (SEP),
    // or c~ = total * (reference weight) for the cheaper P = total (M1 x M1 [x M1])
    const int qx = s.l % NQ, qy = DIM == 3 ? s.l / NQ : 0;
    const R ihor = SEP ? total / (DIM == 2 ? smg[qx] : smg[qx] * smg[NQ + qy] * itot) : itot / lg.wxy;
    R lmin = R(1e30), lmax = R(0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      R r;
      if constexpr (SEP) r = co[q] * ihor * smg[DIM * NQ + q];
      else r = co[q] * ihor / T.W[q];
      lmin = fmin(lmin, r);
      lmax = fmax(lmax, r);
    }
    if (s.act) { srm[s.l] = lmin; srm[LT + s.l] = lmax; }
  }
  // 1D weighted mass matrices and their inverses: thread l < DIM of each cell builds axis l
  if (SEP && s.act && s.l < DIM) {
    R Am[N1][N1];
    const R sc = s.l == 0 ? R(1) : itot;
    const R* wq = smg + s.l * NQ;
#pragma unroll
    for (int i = 0; i < N1; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        R a = R(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) a += wq[q] * T.S[q][i] * T.S[q][j];
        Am[i][j] = sc * a;
        Am[j][i] = sc * a;
      }
    inv_spd<N1, R>(Am);
#pragma unroll
    for (int i = 0; i < N1; ++i)
#pragma unroll
      for (int j = 0; j < N1; ++j) Mi[(s.l * N1 + i) * N1 + j] = Am[i][j];
  }
  cell_sync<DIM, K, Q>();
  R cmin = R(1), cmax = R(1);
  if constexpr (DIM == 3) {
    // the cell's bounds in two levels: threads l < NQ reduce the lines l, l + NQ, ...
    if (s.act && s.l < NQ) {
      R mn = R(1e30), mx = R(0);
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        mn = fmin(mn, srm[r * NQ + s.l]);
        mx = fmax(mx, srm[LT + r * NQ + s.l]);
      }
      smm[s.l] = mn;
      smm[NQ + s.l] = mx;
    }
    cell_sync<DIM, K, Q>();
    if (s.act) {
      R mn = R(1e30), mx = R(0);
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        mn = fmin(mn, smm[r]);
        mx = fmax(mx, smm[NQ + r]);
      }
      cmin = mn; cmax = mx;
    }
  } else if (s.act) {
    R mn = R(1e30), mx = R(0);
    for (int t = 0; t < LT; ++t) {
      mn = fmin(mn, srm[t]);
      mx = fmax(mx, srm[LT + t]);
    }
    cmin = mn; cmax = mx;
  }
  const R alpha = cmin, beta = cmax;
  const R theta = R(0.5) * (alpha + beta);
  const R delta = fmax(R(0.5) * (beta - alpha), R(1e-14) * theta);
  const R sigma1 = theta / delta;
  int nit_lane = 1;
  {
    // sigma = (sqrt(kappa) - 1) / (sqrt(kappa) + 1) = (kappa - 1) / (sqrt(kappa) + 1)^2, with
    // kappa - 1 formed in R so that kappa close to one keeps its digits
    const float km1 = (float)((beta - alpha) / alpha);
    const float sk = sqrtf(1.0f + km1);
    const float sg = km1 / ((sk + 1.0f) * (sk + 1.0f));
    int n = 1;
    if (sg > 1e-30f) n = (int)ceilf(logf(0.5f * (float)tol) / logf(sg));
    n = n < 1 ? 1 : (n > maxit ? maxit : n);
    if constexpr (!C::WAVE_LOCAL)
      if (s.act && s.l == 0) atomicMax(&nit_blk, n);
    nit_lane = (s.act && s.l == 0) ? n : 1;
  }
  int nit;
  if constexpr (C::WAVE_LOCAL) {
    // the cells of one wavefront iterate together: maximum over its lanes
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nit_lane = max(nit_lane, __shfl_xor(nit_lane, off, 64));
    nit = nit_lane;
  } else {
    __syncthreads();
    nit = nit_blk;
  }
  const R zsc = SEP ? R(1) : itot;
  // entry (i, j) of the 1D inverse along axis ax: per-cell matrices in LDS (SEP) or the
  // reference table (uniform, no address taken of the kernel argument)
  auto minv = [&](int ax, int i, int j) -> R {
    if constexpr (SEP) return Mi[(ax * N1 + i) * N1 + j];
    else return T.MI[i][j];
  };

  // z = P^{-1} r for all components (rows in registers)
  auto precond = [&](const R (&r)[NC][N1], R (&z)[NC][N1]) {
    if (row_owner) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R m[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) m[j] = minv(0, i, j);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          R a = R(0);
#pragma unroll
          for (int j = 0; j < N1; ++j) a += m[j] * r[c][j];
          dl[c * NPC + s.l * N1 + i] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
    if (DIM == 2) {
      if (s.act && s.l < N1) {
        R v[NC][N1];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int j = 0; j < N1; ++j) v[c][j] = dl[c * NPC + j * N1 + s.l];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R m[N1];
#pragma unroll
          for (int j = 0; j < N1; ++j) m[j] = minv(DIM - 1, i, j);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            R a = R(0);
#pragma unroll
            for (int j = 0; j < N1; ++j) a += m[j] * v[c][j];
            t1[c * NPC + i * N1 + s.l] = a;
          }
        }
      }
      cell_sync<DIM, K, Q>();
      if (row_owner)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int i = 0; i < N1; ++i) z[c][i] = zsc * t1[c * NPC + s.l * N1 + i];
    } else {
      if (s.act && s.l < N1 * N1) {
        const int ix = s.l % N1, iz = s.l / N1;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          R v[N1];
#pragma unroll
          for (int j = 0; j < N1; ++j) v[j] = dl[c * NPC + (iz * N1 + j) * N1 + ix];
#pragma unroll
          for (int i = 0; i < N1; ++i) {
            R a = R(0);
#pragma unroll
            for (int j = 0; j < N1; ++j) a += minv(1, i, j) * v[j];
            t1[c * NPC + (iz * N1 + i) * N1 + ix] = a;
          }
        }
      }
      cell_sync<DIM, K, Q>();
      if (s.act && s.l < N1 * N1) {
        const int ix = s.l % N1, iy = s.l / N1;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          R v[N1];
#pragma unroll
          for (int j = 0; j < N1; ++j) v[j] = t1[c * NPC + (j * N1 + iy) * N1 + ix];
#pragma unroll
          for (int i = 0; i < N1; ++i) {
            R a = R(0);
#pragma unroll
            for (int j = 0; j < N1; ++j) a += minv(DIM - 1, i, j) * v[j];
            dl[c * NPC + (i * N1 + iy) * N1 + ix] = a;
          }
        }
      }
      cell_sync<DIM, K, Q>();
      if (row_owner)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int i = 0; i < N1; ++i) z[c][i] = zsc * dl[c * NPC + s.l * N1 + i];
    }
    cell_sync<DIM, K, Q>();
  };
  // y = A v for all components
  auto apply = [&](const R (&v)[NC][N1], R (&y)[NC][N1]) {
    if (row_owner)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < N1; ++i) dl[c * NPC + s.l * N1 + i] = v[c][i];
    cell_sync<DIM, K, Q>();
    R vq[NC][NQ];
    eval_lines_lds<DIM, K, NC, R>(dl, s, t1, t2, T, vq);
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int q = 0; q < NQ; ++q) vq[c][q] *= co[q];
    integrate_lines_v<DIM, K, NC, R>(s, i1, i2, T, vq, y);
  };

  R r[NC][N1], z[NC][N1], d[NC][N1];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < N1; ++i) { x[c][i] = R(0); r[c][i] = row_owner ? b[c][i] : R(0); }
  precond(r, z);
  R rho = R(1) / sigma1;
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < N1; ++i) d[c][i] = z[c][i] / theta;
  for (int k = 0; k < nit; ++k) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < N1; ++i) x[c][i] += d[c][i];
    if (k == nit - 1) break;
    R ad[NC][N1];
    apply(d, ad);
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < N1; ++i) r[c][i] -= ad[c][i];
    precond(r, z);
    const R rn = R(1) / (R(2) * sigma1 - rho);
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < N1; ++i) d[c][i] = rn * rho * d[c][i] + (R(2) * rn / delta) * z[c][i];
    rho = rn;
  }
}

// b and x may be the same vector: every row is read by its owner thread before the cell's solve
// and written by that thread after it. 3D k = 4, one component: held to 128 VGPRs (four waves per
// SIMD instead of three; no spills).
template <int DIM, int K, int NC>
__global__ __launch_bounds__(AFL_NT) __attribute__((amdgpu_waves_per_eu(DIM == 3 && K == 4 && NC == 1 ? 4 : 1))) void k_mass_solve(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ rho, double scale,
                                                        int64_t N, const double* b, double* x, double tol, int maxit) {
  const auto& T = *Tp;
  constexpr int PC = Lds<DIM, K>::W > SolvLds<DIM, K>::size(NC) ? Lds<DIM, K>::W : SolvLds<DIM, K>::size(NC);
  __shared__ double shw[Cfg<DIM, K>::CPB * PC];
  using C = Cfg<DIM, K>;
  constexpr int NQ = C::NQ, N1 = C::N1;
  Slot<DIM, K> s(g);
  const int cbs = s.cb < C::CPB ? s.cb : 0;
  double* base = shw + cbs * PC;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double co[NQ];
  if (rho) eval_line<DIM, K>(rho, s, base, base + C::EV1, T, co);
  else for (int q = 0; q < NQ; ++q) co[q] = 1.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) co[q] *= scale * lg.jxw(g, T, q);
  double bb[NC][N1], xx[NC][N1];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < N1; ++i)
      bb[c][i] = (s.act && s.l < C::ROWS) ? b[c * N + s.cell * C::NPC + s.l * N1 + i] : 0.0;
  cell_cheb<DIM, K, NC, double>(T, s, lg, co, base, tol, maxit, bb, xx);
#pragma unroll
  for (int c = 0; c < NC; ++c) store_row<DIM, K>(x + c * N, s, xx[c]);
}

// Schur first half: w = A^{-1} B p (fused B apply + cell-local A solve).
template <int DIM, int K, typename R = double, int Q = 2 * K + 1>
__global__ __launch_bounds__(AFL_NT) void k_schur_w(Geo g, const TabT<R>* __restrict__ Tp, const R* __restrict__ rho, R f, int64_t N,
                                                     const R* __restrict__ p, R* __restrict__ w, R tol, int maxit) {
  const auto& T = *Tp;
  constexpr int PB = Lds<DIM, K, Q>::WF + Lds<DIM, K, Q>::F;  // B phase: work + face nodes + face fluxes
  constexpr int PS = SolvLds<DIM, K, Q>::size(DIM);             // solver phase
  constexpr int PC = PB > PS ? PB : PS;
  __shared__ R shw[Cfg<DIM, K, Q>::CPB * PC];
  using C = Cfg<DIM, K, Q>;
  constexpr int NQ = C::NQ, N1 = C::N1;
  Slot<DIM, K, Q> s(g);
  const int cbs = s.cb < C::CPB ? s.cb : 0;
  R* base = shw + cbs * PC;
  R* t1 = base;
  R* t2 = base + C::EV1;
  R* i1 = base;
  R* i2 = base + C::IN1;
  R* fn = base + Lds<DIM, K, Q>::W;
  R* Fq = base + Lds<DIM, K, Q>::WF;
  LineGeo<DIM, K, R, Q> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  const FacePts<DIM, K, R, Q> fp(g, T, s);
  R bp[DIM][N1];
  apply_B_rows<DIM, K, R>(g, T, s, lg, nb, fp, p, f, t1, t2, i1, i2, fn, Fq, bp);
  R co[NQ];
  eval_line<DIM, K, R>(rho, s, t1, t2, T, co);
#pragma unroll
  for (int q = 0; q < NQ; ++q) co[q] *= lg.jxw(g, T, q);
  R xx[DIM][N1];
  // FP64 (exact solves): separable per-cell preconditioner; FP32 (multigrid level operator, one
  // Chebyshev step at preconditioner accuracy): the cheaper reference-mass preconditioner
  cell_cheb<DIM, K, DIM, R, sizeof(R) == 8>(T, s, lg, co, base, tol, maxit, bp, xx);
#pragma unroll
  for (int c = 0; c < DIM; ++c) store_row<DIM, K, R>(w + c * N, s, xx[c]);
}

// rows (held by the row owners) <- (MI x MI [x MI]) rows for NC components: the inverse of the
// reference-cell mass matrix, one LDS transpose per further axis. tmp: 2*NC*NPC values per cell.
template <int DIM, int K, int NC, typename R, int Q = 2 * K + 1>
__device__ __attribute__((always_inline)) inline void cell_mi_apply(const Slot<DIM, K, Q>& s, const TabT<R>& T, R* tmp, R (&rows)[NC][K + 1]) {
  using C = Cfg<DIM, K, Q>;
  constexpr int N1 = C::N1, NPC = C::NPC, ROWS = C::ROWS;
  R* t2 = tmp + NC * NPC;
  const bool ro = s.act && s.l < ROWS;
  if (ro) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        R a = R(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) a += T.MI[i][j] * rows[c][j];
        tmp[c * NPC + s.l * N1 + i] = a;
      }
  }
  cell_sync<DIM, K, Q>();
  const R* res;
  if (DIM == 2) {
    if (s.act && s.l < N1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R v[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = tmp[c * NPC + j * N1 + s.l];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = R(0);
#pragma unroll
          for (int j = 0; j < N1; ++j) a += T.MI[i][j] * v[j];
          t2[c * NPC + i * N1 + s.l] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
    res = t2;
  } else {
    if (s.act && s.l < N1 * N1) {
      const int ix = s.l % N1, iz = s.l / N1;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R v[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = tmp[c * NPC + (iz * N1 + j) * N1 + ix];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = R(0);
#pragma unroll
          for (int j = 0; j < N1; ++j) a += T.MI[i][j] * v[j];
          t2[c * NPC + (iz * N1 + i) * N1 + ix] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
    if (s.act && s.l < N1 * N1) {
      const int ix = s.l % N1, iy = s.l / N1;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R v[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = t2[c * NPC + (j * N1 + iy) * N1 + ix];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          R a = R(0);
#pragma unroll
          for (int j = 0; j < N1; ++j) a += T.MI[i][j] * v[j];
          tmp[c * NPC + (i * N1 + iy) * N1 + ix] = a;
        }
      }
    }
    cell_sync<DIM, K, Q>();
    res = tmp;
  }
  if (ro)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < N1; ++i) rows[c][i] = res[c * NPC + s.l * N1 + i];
}

// Multigrid level operator, first half, at preconditioner accuracy: w = P^{-1} B p with the
// mean-density block P = (sum_q rho JxW) (M1 x M1 [x M1]) in place of A (sinv[c] = 1 / sum_q rho
// JxW of cell c, refreshed with the stage coefficients). P^{-1} A has its spectrum within the
// cell's relative variation of rho J about the mean, a few percent.
template <int DIM, int K, typename R, int Q = 2 * K + 1>
__global__ __launch_bounds__(AFL_NT) __attribute__((amdgpu_waves_per_eu(DIM == 2 && sizeof(R) == 4 ? AFL_WLVL_WAVES : 1))) void k_schur_w_lvl(Geo g, const TabT<R>* __restrict__ Tp, const R* __restrict__ sinv, R f, int64_t N,
                                                         const R* __restrict__ p, R* __restrict__ w) {
  const auto& T = *Tp;
  using C = Cfg<DIM, K, Q>;
  constexpr int PB = Lds<DIM, K, Q>::WF + Lds<DIM, K, Q>::F;
  constexpr int PM = 2 * DIM * C::NPC;
  constexpr int PC = PB > PM ? PB : PM;
  // cells spread over several wavefronts need the mass-inverse scratch [0, PM) apart from the face
  // fluxes that slower waves' add_faces may still read; inside one wavefront LDS accesses stay in order
  static_assert(C::WAVE_LOCAL || PM <= Lds<DIM, K, Q>::WF, "the mass-inverse scratch must not overlap the face fluxes");
  __shared__ R shw[C::CPB * PC];
  constexpr int N1 = C::N1;
  Slot<DIM, K, Q> s(g);
  const int cbs = s.cb < C::CPB ? s.cb : 0;
  R* base = shw + cbs * PC;
  const R sc = s.act ? sinv[s.cell] : R(0);  // loaded first: not a round trip at the end
  LineGeo<DIM, K, R, Q> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  const FacePts<DIM, K, R, Q> fp(g, T, s);
  R bp[DIM][N1];
  apply_B_rows<DIM, K, R>(g, T, s, lg, nb, fp, p, f, base, base + C::EV1, base, base + C::IN1,
                          base + Lds<DIM, K, Q>::W, base + Lds<DIM, K, Q>::WF, bp);
  cell_mi_apply<DIM, K, DIM, R>(s, T, base, bp);
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
#pragma unroll
    for (int i = 0; i < N1; ++i) bp[c][i] *= sc;
    store_row<DIM, K, R>(w + c * N, s, bp[c]);
  }
}

// sinv[c] = 1 / sum_q rho(q) JxW(q) of cell c (one thread per cell). JxW(q) = F(qx, qy) W(qz) on the
// terrain-following map (F: the line's horizontal factor), so the sum factorises:
//   sum_(qx, qy) F(qx, qy) sum_(ix, iy) S[qx][ix] S[qy][iy] sum_iz wz[iz] rho(ix, iy, iz),
// wz[iz] = sum_qz W(qz) S[qz][iz] = the integral of the 1D basis function iz.
template <int DIM, int K, typename R, int Q = 2 * K + 1>
__global__ __launch_bounds__(AFL_NT) void k_cell_rho_mass_inv(Geo g, const TabT<R>* __restrict__ Tp, const R* __restrict__ rho,
                                                               R* __restrict__ sinv) {
  const auto& T = *Tp;
  using C = Cfg<DIM, K, Q>;
  constexpr int NQ = C::NQ, N1 = C::N1, NPC = C::NPC;
  constexpr int NH = DIM == 3 ? N1 * N1 : N1;  // horizontal node positions
  double wz[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) a += (double)T.W[q] * (double)T.S[q][i];
    wz[i] = a;
  }
  for (int64_t c = blockIdx.x * (int64_t)AFL_NT + threadIdx.x; c < g.ncells; c += (int64_t)gridDim.x * AFL_NT) {
    int ic[3];
    ic[0] = (int)(c % g.nc[0]);
    if (DIM == 2) { ic[1] = (int)(c / g.nc[0]); ic[2] = 0; }
    else { ic[1] = (int)((c / g.nc[0]) % g.nc[1]); ic[2] = (int)(c / ((int64_t)g.nc[0] * g.nc[1])); }
    const R* r = rho + c * NPC;
    double rz[NH];  // rho integrated along z at each horizontal node position
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      double a = 0.0;
#pragma unroll
      for (int iz = 0; iz < N1; ++iz) a += wz[iz] * (double)r[iz * NH + h];
      rz[h] = a;
    }
    double acc = 0.0;
    for (int l = 0; l < C::LT; ++l) {
      LineGeo<DIM, K, R, Q> lg;
      lg.init(g, T, ic, l);
      const int qx = l % NQ, qy = DIM == 3 ? l / NQ : 0;
      const double F = DIM == 3 ? (double)lg.hx * (double)lg.hy * (double)lg.c * (double)lg.wxy
                                : (double)lg.hx * (double)lg.c * (double)lg.wxy;
      double v = 0.0;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int ix = h % N1, iy = h / N1;
        double phi = (double)T.S[qx][ix];
        if (DIM == 3) phi *= (double)T.S[qy][iy];
        v += phi * rz[h];
      }
      acc += F * v;
    }
    sinv[c] = (R)(1.0 / acc);
  }
}

// ---------------------------------------------------------------------------------------
// Launchers (dispatch on dim and degree).
template <int DIM, int K, int Q = 2 * K + 1>
static dim3 grid_for(int ncells) {
  using C = Cfg<DIM, K, Q>;
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
  const double by = 8.0 * (double)o.N * (stage * (1 + o.dim) + 1);  // s (rho_j, u_j) in, rhs out
  if (use_w3(o)) {
    const double w0 = stage == 1 ? t.a21 : stage == 2 ? t.a31 : t.b1;
    const double w1 = stage == 1 ? 0.0 : stage == 2 ? t.a32 : t.b2;
    const double w2 = stage == 3 ? t.b3 : 0.0;
    launch3_rhs_rho(o, by, stage, S, w0, w1, w2, dst);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    dim3 gr = grid_for<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    if (stage == 1)
      launch_timed(o, kTimeRhsRho, 1, by, k_rhs_rho<D_, K_, 1>, gr, bl, o.geo, o.dT, S, o.dt, t.a21, 0.0, 0.0, o.N, dst);
    else if (stage == 2)
      launch_timed(o, kTimeRhsRho, 2, by, k_rhs_rho<D_, K_, 2>, gr, bl, o.geo, o.dT, S, o.dt, t.a31, t.a32, 0.0, o.N, dst);
    else
      launch_timed(o, kTimeRhsRho, 3, by, k_rhs_rho<D_, K_, 3>, gr, bl, o.geo, o.dT, S, o.dt, t.b1, t.b2, t.b3, o.N, dst);
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
  // sources: (rho, u, p) per explicit state plus rho_new at stages 1-2; rhs (dim fields) out
  const int d = o.dim;
  const double by = 8.0 * (double)o.N * (stage == 1 ? 3 + d + d : stage == 2 ? 2 * d + 5 + d : 3 * (2 + d) + d);
  if (use_w3(o)) {
    launch3_rhs_u(o, by, stage, S, cf, iMa2, iFr2, dst);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    dim3 gr = grid_for<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    if (stage == 1)
      launch_timed(o, kTimeRhsU, 1, by, k_rhs_u<D_, K_, 1, 2>, gr, bl, o.geo, o.dT, S, cf, o.dt, iMa2, iFr2, o.N, dst);
    else if (stage == 2)
      launch_timed(o, kTimeRhsU, 2, by, k_rhs_u<D_, K_, 2, 3>, gr, bl, o.geo, o.dT, S, cf, o.dt, iMa2, iFr2, o.N, dst);
    else
      launch_timed(o, kTimeRhsU, 3, by, k_rhs_u<D_, K_, 3, 3>, gr, bl, o.geo, o.dT, S, cf, o.dt, iMa2, iFr2, o.N, dst);
  });
}

void launch_rhs_p(const OpCtx& o, int stage, const double* const* src, double* dst, bool full) {
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
  // sources as ordered for the stage (without u_fixed when !full); rhs out
  const int d = o.dim;
  const int nin = stage == 1 ? (full ? 2 * d + 3 : d + 3) : stage == 2 ? (full ? 3 * d + 5 : 2 * d + 5) : 4 * d + 7;
  const double by = 8.0 * (double)o.N * (nin + 1);
  if (use_w3(o)) {
    launch3_rhs_p(o, by, stage, full, S, cf, Ma2, iFr2, dst);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    dim3 gr = grid_for<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    if (stage == 1 && full)
      launch_timed(o, kTimeRhsP, 1, by, k_rhs_p<D_, K_, 1, 2, 1>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N, dst);
    else if (stage == 1)
      launch_timed(o, kTimeRhsP, 1, by, k_rhs_p<D_, K_, 1, 2, 1, false>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N,
                   dst);
    else if (stage == 2 && full)
      launch_timed(o, kTimeRhsP, 2, by, k_rhs_p<D_, K_, 2, 3, 2>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N, dst);
    else if (stage == 2)
      launch_timed(o, kTimeRhsP, 2, by, k_rhs_p<D_, K_, 2, 3, 2, false>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N,
                   dst);
    else
      launch_timed(o, kTimeRhsP, 3, by, k_rhs_p<D_, K_, 3, 3, 3>, gr, bl, o.geo, o.dT, S, cf, o.dt, Ma2, iFr2, o.N, dst);
  });
}

void launch_rhs_p_ufix(const OpCtx& o, int stage, const double* rho_new, const double* u_fix, const double* base,
                       double* dst) {
  const TableauC& t = o.tab;
  const double Ma2 = o.Ma * o.Ma, iFr2 = 1.0 / (o.Fr * o.Fr);
  const double cg = (stage == 1 ? t.at22 : t.at33) * o.dt * Ma2 * iFr2;
  const double by = 8.0 * (double)o.N * (o.dim + 3);  // rho_new, u_fix, base in; rhs out
  if (use_w3(o)) {
    launch3_rhs_p_ufix(o, by, stage, rho_new, u_fix, 0.5 * Ma2, cg, base, dst);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    launch_timed(o, kTimeRhsPUfix, stage, by, k_rhs_p_ufix<D_, K_>, grid_for<D_, K_>(o.geo.ncells),
                 dim3(Cfg<D_, K_>::NT), o.geo, o.dT, rho_new, u_fix, 0.5 * Ma2, cg, o.N, base, dst);
  });
}

void launch_mass_apply(const OpCtx& o, const double* rho, double scale, int ncomp, const double* src, double* dst) {
  AFL_DISPATCH(o.dim, o.k, {
    k_mass_apply<D_, K_><<<grid_for<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(o.geo, o.dT, rho, scale, ncomp, o.N, src, dst);
  });
}

void launch_inv_diag(const OpCtx& o, const double* rho, double scale, int ncomp, double* dst) {
  AFL_DISPATCH(o.dim, o.k, {
    k_inv_diag<D_, K_><<<grid_for<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(o.geo, o.dT, rho, scale, ncomp, o.N, dst);
  });
}

void launch_apply_B(const OpCtx& o, const double* p, double* dst) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double f = coeff * o.dt / (o.Ma * o.Ma);
  const double by = 8.0 * (double)o.N * (1 + o.dim);
  if (use_w3(o)) {
    launch3_apply_B(o, by, f, p, dst);
    return;
  }
  if (o.dim == 3 && o.k == 4) {
    const dim3 gr = grid_for<3, 4>(o.geo.ncells), bl(Cfg<3, 4>::NT);
    const double byc = 8.0 * (double)o.N * 2;
    launch_timed(o, kTimeApplyB, (int)o.hyp, byc, k_apply_B1<3, 4, 0>, gr, bl, o.geo, o.dT, f, o.N, p, dst);
    launch_timed(o, kTimeApplyB, (int)o.hyp, byc, k_apply_B1<3, 4, 1>, gr, bl, o.geo, o.dT, f, o.N, p, dst);
    launch_timed(o, kTimeApplyB, (int)o.hyp, byc, k_apply_B1<3, 4, 2>, gr, bl, o.geo, o.dT, f, o.N, p, dst);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    launch_timed(o, kTimeApplyB, (int)o.hyp, by, k_apply_B<D_, K_>, grid_for<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT),
                 o.geo, o.dT, f, o.N, p, dst);
  });
}

void launch_apply_C(const OpCtx& o, const double* pfix, const double* w, const double* pD, double* dst) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double cdt = coeff * o.dt;
  const double by = 8.0 * (double)o.N * (o.dim + (pD ? 3 : 2));  // pfix, w [, pD] in; out
  if (use_w3(o)) {
    launch3_apply_C(o, by, cdt, pfix, w, pD, dst);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    const dim3 gr = grid_for<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    if (pD)
      launch_timed(o, kTimeApplyC, (int)o.hyp, by, k_apply_C<D_, K_, true, double>, gr, bl, o.geo, o.dT, cdt, o.N, pfix,
                   w, pD, 1.0 / (kGamma - 1.0), dst);
    else
      launch_timed(o, kTimeApplyC, (int)o.hyp, by, k_apply_C<D_, K_, false, double>, gr, bl, o.geo, o.dT, cdt, o.N, pfix,
                   w, (const double*)nullptr, 0.0, dst);
  });
}

void launch_mass_solve(const OpCtx& o, const double* rho, double scale, int ncomp, const double* b, double* x) {
  const double tol2 = o.cell_tol;
  const double by = 8.0 * (double)o.N * (2 * ncomp + (rho ? 1 : 0));
  if (use_w3(o)) {
    launch3_mass_solve(o, by, rho, scale, ncomp, b, x);
    return;
  }
  if (o.dim == 3 && o.k == 4 && ncomp > 1) {
    // one launch per component: the one-component kernel runs at three waves per SIMD, the
    // three-component one at one (its per-thread iterates triple), which more than pays for
    // reading rho and forming the cell's preconditioner three times
    const dim3 gr = grid_for<3, 4>(o.geo.ncells), bl(Cfg<3, 4>::NT);
    const double byc = 8.0 * (double)o.N * (2 + (rho ? 1 : 0));
    for (int c = 0; c < ncomp; ++c)
      launch_timed(o, kTimeASolve, (int)o.hyp, byc, k_mass_solve<3, 4, 1>, gr, bl, o.geo, o.dT, rho, scale, o.N,
                   b + c * o.N, x + c * o.N, tol2, o.cell_maxit);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    const dim3 gr = grid_for<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    if (ncomp == 1)
      launch_timed(o, kTimeMassSolve, (int)o.hyp, by, k_mass_solve<D_, K_, 1>, gr, bl, o.geo, o.dT, rho, scale, o.N, b, x,
                   tol2, o.cell_maxit);
    else
      launch_timed(o, kTimeASolve, (int)o.hyp, by, k_mass_solve<D_, K_, D_>, gr, bl, o.geo, o.dT, rho, scale, o.N, b,
                   x, tol2, o.cell_maxit);
  });
}

void launch_schur_w(const OpCtx& o, const double* rho, const double* p, double* w) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double f = coeff * o.dt / (o.Ma * o.Ma);
  const double tol2 = o.cell_tol;
  const double by = 8.0 * (double)o.N * (2 + o.dim);  // p, rho in; w out
  if (use_w3(o)) {
    launch3_schur_w(o, by, f, rho, p, w);
    return;
  }
  if (o.dim == 3 && o.k == 4) {
    // the fused kernel holds the B apply's and the cell solve's registers at once (one wave per
    // SIMD); as separate launches each phase runs at its own occupancy, for one extra pass over w
    launch_apply_B(o, p, w);
    launch_mass_solve(o, rho, 1.0, 3, w, w);
    return;
  }
  AFL_DISPATCH(o.dim, o.k, {
    launch_timed(o, kTimeSchurW, (int)o.hyp, by, k_schur_w<D_, K_, double>, grid_for<D_, K_>(o.geo.ncells),
                 dim3(Cfg<D_, K_>::NT), o.geo, o.dT, rho, f, o.N, p, w, tol2, o.cell_maxit);
  });
}

// Multigrid level operators: besides dim and degree, the quadrature rule of the level, k + 1
// (rq) or 2k + 1 Gauss points per direction.
#define AFL_DISPATCH_LQ(DIMV, KV, RQ, ...)                          \
  AFL_DISPATCH(DIMV, KV, {                                           \
    if (RQ) { constexpr int Q_ = K_ + 1; __VA_ARGS__; }              \
    else { constexpr int Q_ = 2 * K_ + 1; __VA_ARGS__; }             \
  })

// Level Schur apply y = D p - C A^{-1} B p in precision R (multigrid levels).
template <typename R>
void launch_schur_pair(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st, const R* rho,
                       const R* sinv, const R* pfix, R f, R cdt, R tol, int maxit, int64_t N, const R* p, R* w, R* y) {
  AFL_DISPATCH_LQ(dim, k, rq, {
    const dim3 gr = grid_for<D_, K_, Q_>(geo.ncells);
    const dim3 bl(Cfg<D_, K_, Q_>::NT);
    if (sinv) k_schur_w_lvl<D_, K_, R, Q_><<<gr, bl, 0, st>>>(geo, T, sinv, f, N, p, w);
    else k_schur_w<D_, K_, R, Q_><<<gr, bl, 0, st>>>(geo, T, rho, f, N, p, w, tol, maxit);
    k_apply_C<D_, K_, true, R, Q_><<<gr, bl, 0, st>>>(geo, T, cdt, N, pfix, w, p, (R)(1.0 / (kGamma - 1.0)), y);
  });
}
// w = A^{-1} B p alone in precision R (the first half of the level operator)
template <typename R>
void launch_schur_w_r(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st, const R* rho,
                      const R* sinv, R f, R tol, int maxit, int64_t N, const R* p, R* w, hipEvent_t* ev) {
  AFL_DISPATCH_LQ(dim, k, rq, {
    const dim3 gr = grid_for<D_, K_, Q_>(geo.ncells);
    const dim3 bl(Cfg<D_, K_, Q_>::NT);
    if (sinv && ev)
      hipExtLaunchKernelGGL(k_schur_w_lvl<D_, K_, R, Q_>, gr, bl, 0, st, ev[0], ev[1], 0, geo, T, sinv, f, N, p, w);
    else if (sinv)
      k_schur_w_lvl<D_, K_, R, Q_><<<gr, bl, 0, st>>>(geo, T, sinv, f, N, p, w);
    else
      k_schur_w<D_, K_, R, Q_><<<gr, bl, 0, st>>>(geo, T, rho, f, N, p, w, tol, maxit);
  });
}
template <typename R>
void launch_cell_rho_mass_inv(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st,
                              const R* rho, R* sinv) {
  const int nb = (int)((geo.ncells + AFL_NT - 1) / AFL_NT);
  AFL_DISPATCH_LQ(dim, k, rq, { k_cell_rho_mass_inv<D_, K_, R, Q_><<<dim3(nb), dim3(AFL_NT), 0, st>>>(geo, T, rho, sinv); });
}
// S p (second half) fused with one Chebyshev step (see k_apply_C_cheb)
template <typename R>
void launch_apply_C_cheb(int dim, int k, bool rq, const Geo& geo, const TabT<R>* T, hipStream_t st, const R* pfix,
                         R cdt, int64_t N, const R* w, const R* p, const R* jbar, const R* rin, R* rout, R c1, R c2,
                         R* d, R* x, hipEvent_t* ev) {
  const R sD = (R)(1.0 / (kGamma - 1.0));
  AFL_DISPATCH_LQ(dim, k, rq, {
    const dim3 gr = grid_for<D_, K_, Q_>(geo.ncells);
    const dim3 bl(Cfg<D_, K_, Q_>::NT);
    if (ev)
      hipExtLaunchKernelGGL(k_apply_C_cheb<D_, K_, R, Q_>, gr, bl, 0, st, ev[0], ev[1], 0, geo, T, cdt, N, pfix, w, p,
                            sD, jbar, rin, rout, c1, c2, d, x);
    else
      k_apply_C_cheb<D_, K_, R, Q_><<<gr, bl, 0, st>>>(geo, T, cdt, N, pfix, w, p, sD, jbar, rin, rout, c1, c2, d, x);
  });
}

#define AFL_LEVEL_INSTANCES(R)                                                                                       \
  template void launch_schur_pair<R>(int, int, bool, const Geo&, const TabT<R>*, hipStream_t, const R*, const R*,    \
                                     const R*, R, R, R, int, int64_t, const R*, R*, R*);                               \
  template void launch_schur_w_r<R>(int, int, bool, const Geo&, const TabT<R>*, hipStream_t, const R*, const R*, R, \
                                    R, int, int64_t, const R*, R*, hipEvent_t*);                                       \
  template void launch_cell_rho_mass_inv<R>(int, int, bool, const Geo&, const TabT<R>*, hipStream_t, const R*, R*); \
  template void launch_apply_C_cheb<R>(int, int, bool, const Geo&, const TabT<R>*, hipStream_t, const R*, R, int64_t, \
                                       const R*, const R*, const R*, const R*, R*, R, R, R*, R*, hipEvent_t*);
AFL_LEVEL_INSTANCES(float)
AFL_LEVEL_INSTANCES(double)
#undef AFL_LEVEL_INSTANCES

}  // namespace afl
