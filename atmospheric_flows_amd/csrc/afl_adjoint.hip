// Discrete-adjoint building blocks: transposes of the linearised forward operators.
// The reference has no adjoint; these kernels differentiate its forward weak forms exactly:
//   E1 (vmult_rhs_rho_update, euler_operator.h:342-692): vector-Jacobian product with respect
//      to every source (rho_j, u_j);
//   B  (vmult_pressure, :1510-1617): B^T;
//   C  (vmult_enthalpy, :1411-1507): the tangent C'(w0) (C is nonlinear in w through the Rusanov
//      speed max(|w+.n|, |w-.n|)) and its transpose.
// Same decomposition as the forward kernels (one thread per quadrature line, cells packed into
// wavefronts). Faces are handled cell-centrically from both sides: a cell's adjoint gathers the
// derivative of every face integral its own traces enter, i.e. the integral tested against its
// own functions and the one tested against the neighbour's, so no cell writes another's dofs.
#include "afl_internal.h"

namespace afl {

namespace {

template <int DIM, int K>
struct AdjLds {
  using C = Cfg<DIM, K>;
  static constexpr int EV = C::EV1 + C::EV2;
  static constexpr int IN = C::IN1 + C::IN2;
  static constexpr int W = EV > IN ? EV : IN;  // eval / gradient / integrate work area per cell
  static constexpr int F = 2 * DIM * C::NFQ;   // face values per cell
  static constexpr int FN = FaceLds<DIM, K>::FN;
  static constexpr int WF = W + FN;
};

#define AFL_ADJ_SLOT                                                       \
  using C = Cfg<DIM, K>;                                                   \
  constexpr int NQ = C::NQ;                                                \
  Slot<DIM, K> s(g);                                                       \
  const int cbs = s.cb < C::CPB ? s.cb : 0;                                \
  double* t1 = shw + cbs * AdjLds<DIM, K>::WF;                             \
  double* t2 = t1 + C::EV1;                                                \
  double* i1 = t1;                                                         \
  double* i2 = t1 + C::IN1;                                                \
  double* fn = t1 + AdjLds<DIM, K>::W;                                     \
  double* Fq = shF + cbs * AdjLds<DIM, K>::F;                              \
  (void)t2; (void)i2;

// Rusanov speed max(|a|, |b|): d lambda = wa da + wb db; a tie splits the subgradient evenly so
// that both cells of a face use the same one.
__device__ __attribute__((always_inline)) inline void lam_weights(double a, double b, double& wa, double& wb) {
  const double fa = fabs(a), fb = fabs(b);
  const double sa = a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0), sb = b > 0.0 ? 1.0 : (b < 0.0 ? -1.0 : 0.0);
  wa = fa > fb ? sa : (fa < fb ? 0.0 : 0.5 * sa);
  wb = fb > fa ? sb : (fb < fa ? 0.0 : 0.5 * sb);
}

// Integrate one output field from its line values V (times JxW) and optional reference-gradient
// terms G (gref form), add the face values Fv (own face point, times JxW_f) and store the rows.
template <int DIM, int K, bool HAS_V, bool HAS_G>
__device__ __attribute__((always_inline)) inline void finish_field(const Slot<DIM, K>& s, double* i1, double* i2, double* Fq,
                                                                   const Tab& T, const double (&V)[2 * K + 1],
                                                                   const double (&G)[DIM][2 * K + 1],
                                                                   const double (&Fv)[2 * DIM], double* dst) {
  using C = Cfg<DIM, K>;
  cell_sync<DIM, K>();  // the previous field's face reads are done
  if (s.act)
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fq[f * C::NFQ + s.l] = Fv[f];
  double out[K + 1];
  integrate_line<DIM, K, HAS_V, HAS_G>(s, i1, i2, T, V, G, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  store_row<DIM, K>(dst, s, out);
}

// Physical gradient times JxW of a scalar field at the line's points.
template <int DIM, int K>
__device__ __attribute__((always_inline)) inline void grad_jxw(const double* f, const Slot<DIM, K>& s, double* t1, const Tab& T,
                                                               const Geo& g, const LineGeo<DIM, K>& lg,
                                                               double (&P)[DIM][2 * K + 1]) {
  double gr[DIM][2 * K + 1];
  eval_line_grad<DIM, K>(f, s, t1, T, gr);
#pragma unroll
  for (int q = 0; q < 2 * K + 1; ++q) {
    double a[3], b[3];
    for (int d = 0; d < DIM; ++d) a[d] = gr[d][q];
    lg.gref_t(g, T, q, a, b);
    for (int d = 0; d < DIM; ++d) P[d][q] = b[d];
  }
}

template <int DIM, int K>
__device__ __attribute__((always_inline)) inline void face_normals(const Geo& g, const Tab& T, const Slot<DIM, K>& s,
                                                                   double (&nv)[2 * DIM][DIM], double (&jf)[2 * DIM]) {
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    double n3[3];
    face_geo<DIM, K>(g, T, s.ic, f, s.l, n3, jf[f]);
    for (int d = 0; d < DIM; ++d) nv[f][d] = n3[d];
  }
}

// Source pointers and their adjoint outputs, positional (up to the 11 of E3).
struct AdjSrcs { const double* p[12]; double* o[12]; };

// io.p[i] / io.o[i] for a runtime i without indexing the kernel argument through memory
__device__ __attribute__((always_inline)) inline const double* in_at(const AdjSrcs& io, int i) {
  const double* v = io.p[0];
#pragma unroll
  for (int j = 1; j < 12; ++j) v = (i == j) ? io.p[j] : v;
  return v;
}
__device__ __attribute__((always_inline)) inline double* out_at(const AdjSrcs& io, int i) {
  double* v = io.o[0];
#pragma unroll
  for (int j = 1; j < 12; ++j) v = (i == j) ? io.o[j] : v;
  return v;
}

// E1 adjoint: out[i] = (d rhs_rho / d src[i])^T rbar, src = (rho_j, u_j), j < NS; one launch per
// state j = jsel.
template <int DIM, int K, int NS>
__global__ __launch_bounds__(AFL_NT) void k_vjp_rhs_rho(Geo g, const Tab* __restrict__ Tp, AdjSrcs io,
                                                        const double* __restrict__ rbar, double dt, double w0,
                                                        double w1, double w2, int64_t N, int jsel) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double R[NQ], trR[2 * DIM][2], P[DIM][NQ];
  eval_line_faces<DIM, K>(rbar, s, nb, t1, t2, fn, T, R, trR);
  grad_jxw<DIM, K>(rbar, s, t1, T, g, lg, P);
  const double G0[DIM][NQ] = {};
  {
    const int j = jsel;
    const double cj = (j == 0 ? w0 : (j == 1 ? w1 : w2)) * dt;
    const double* rs = in_at(io, 2 * j);
    const double* us = in_at(io, 2 * j + 1);
    double* ro = out_at(io, 2 * j);
    double* uo = out_at(io, 2 * j + 1);
    double r[NQ], trr[2 * DIM][2], u[DIM][NQ], tru[DIM][2 * DIM][2];
    eval_line_faces<DIM, K>(rs, s, nb, t1, t2, fn, T, r, trr);
#pragma unroll
    for (int d = 0; d < DIM; ++d) eval_line_faces<DIM, K>(us + d * N, s, nb, t1, t2, fn, T, u[d], tru[d]);
    // face weights W = c (R_nb - R_own) JxW_f, own normal velocity and Rusanov weight
    double W[2 * DIM], uno[2 * DIM], lam[2 * DIM], la[2 * DIM], nv[2 * DIM][DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double n3[3], jf;
      face_geo<DIM, K>(g, T, s.ic, f, s.l, n3, jf);
      double a = 0.0, b = 0.0;
      for (int d = 0; d < DIM; ++d) { nv[f][d] = n3[d]; a += tru[d][f][0] * n3[d]; b += tru[d][f][1] * n3[d]; }
      double lb;
      lam_weights(a, b, la[f], lb);
      uno[f] = a;
      lam[f] = fmax(fabs(a), fabs(b));
      W[f] = nb.c[f] >= 0 ? cj * (trR[f][1] - trR[f][0]) * jf : 0.0;
    }
    double V[NQ], Fv[2 * DIM];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double v = j == 0 ? R[q] * lg.jxw(g, T, q) : 0.0;
      for (int d = 0; d < DIM; ++d) v += cj * u[d][q] * P[d][q];
      V[q] = v;
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = W[f] * 0.5 * (uno[f] + lam[f]);
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, ro);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] = cj * r[q] * P[d][q];
#pragma unroll
      for (int f = 0; f < 2 * DIM; ++f)
        Fv[f] = W[f] * 0.5 * (trr[f][0] + (trr[f][0] - trr[f][1]) * la[f]) * nv[f][d];
      finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, uo + d * N);
    }
  }
}

// E2 adjoint: out[i] = (d rhs_u / d src[i])^T rbar. Explicit states (rho, u, p) = src[3j..3j+2],
// j < NE, advective weight w[j], pressure weight wp[j]; gravity densities src[3g] with weight
// gw[g], g < NG (the densities of states beyond NE enter through gravity only).
struct AdjE2Coef { double w[3], wp[3], gw[3]; };

// One launch per explicit state (jsel < NE), and one for the gravity-only densities (jsel == NE).
template <int DIM, int K, int NE, int NG>
__global__ __launch_bounds__(AFL_NT) void k_vjp_rhs_u(Geo g, const Tab* __restrict__ Tp, AdjSrcs io,
                                                      const double* __restrict__ rbar, AdjE2Coef cf, double dt,
                                                      double iMa2, double iFr2, int64_t N, int jsel) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double nv[2 * DIM][DIM], jf[2 * DIM];
  face_normals<DIM, K>(g, T, s, nv, jf);
  // rbar (DIM fields): values, face weights W = (R_nb - R_own) JxW_f, wall normal traces, gradients
  double R[DIM][NQ], W[2 * DIM][DIM], Rw[2 * DIM], P[DIM][DIM][NQ];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) Rw[f] = 0.0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    double tr[2 * DIM][2];
    eval_line_faces<DIM, K>(rbar + a * N, s, nb, t1, t2, fn, T, R[a], tr);
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      W[f][a] = nb.c[f] >= 0 ? (tr[f][1] - tr[f][0]) * jf[f] : 0.0;
      if (nb.c[f] < 0) Rw[f] += tr[f][0] * nv[f][a] * jf[f];
    }
    grad_jxw<DIM, K>(rbar + a * N, s, t1, T, g, lg, P[a]);
  }
  double jx[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) jx[q] = lg.jxw(g, T, q);
  const double G0[DIM][NQ] = {};
  if (jsel < NE) {
    const int j = jsel;
    const double cj = pick<3>(cf.w, j) * dt, cpj = pick<3>(cf.wp, j) * dt * iMa2;
    const double cg = (j < NG ? pick<3>(cf.gw, j) : 0.0) * dt * iFr2;
    double r[NQ], trr[2 * DIM][2], u[DIM][NQ], tru[DIM][2 * DIM][2];
    eval_line_faces<DIM, K>(in_at(io, 3 * j), s, nb, t1, t2, fn, T, r, trr);
#pragma unroll
    for (int d = 0; d < DIM; ++d) eval_line_faces<DIM, K>(in_at(io, 3 * j + 1) + d * N, s, nb, t1, t2, fn, T, u[d], tru[d]);
    double una[2 * DIM], lam[2 * DIM], la[2 * DIM], Wua[2 * DIM], Wj[2 * DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double a = 0.0, b = 0.0, x = 0.0, y = 0.0, z = 0.0;
      for (int d = 0; d < DIM; ++d) {
        a += tru[d][f][0] * nv[f][d];
        b += tru[d][f][1] * nv[f][d];
        x += W[f][d] * tru[d][f][0];
        y += W[f][d] * tru[d][f][1];
        z += W[f][d] * (trr[f][0] * tru[d][f][0] - trr[f][1] * tru[d][f][1]);
      }
      double lb;
      lam_weights(a, b, la[f], lb);
      una[f] = a;
      lam[f] = fmax(fabs(a), fabs(b));
      Wua[f] = x;
      Wj[f] = z;
      (void)y;
    }
    double V[NQ], Fv[2 * DIM];
    // density of state j (with its gravity term)
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double v = -cg * R[DIM - 1][q] * jx[q];
      for (int a = 0; a < DIM; ++a) {
        if (j == 0) v += u[a][q] * R[a][q] * jx[q];
        for (int b = 0; b < DIM; ++b) v += cj * u[a][q] * u[b][q] * P[a][b][q];
      }
      V[q] = v;
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = cj * 0.5 * (una[f] + lam[f]) * Wua[f];
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * j));
    // velocity components
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double v = j == 0 ? r[q] * R[c][q] * jx[q] : 0.0;
        for (int b = 0; b < DIM; ++b) v += cj * r[q] * u[b][q] * (P[c][b][q] + P[b][c][q]);
        V[q] = v;
      }
#pragma unroll
      for (int f = 0; f < 2 * DIM; ++f) {
        const double ra = trr[f][0];
        Fv[f] = cj * 0.5 * (ra * (W[f][c] * una[f] + Wua[f] * nv[f][c]) + la[f] * nv[f][c] * Wj[f] +
                            lam[f] * ra * W[f][c]);
      }
      finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * j + 1) + c * N);
    }
    // pressure of state j: divergence of rbar, average flux and walls (p_D = p)
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double v = 0.0;
      for (int a = 0; a < DIM; ++a) v += P[a][a][q];
      V[q] = cpj * v;
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double wn = 0.0;
      for (int d = 0; d < DIM; ++d) wn += W[f][d] * nv[f][d];
      Fv[f] = nb.c[f] >= 0 ? cpj * 0.5 * wn : -cpj * Rw[f];
    }
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * j + 2));
    return;
  }
  // densities that enter through gravity only
#pragma unroll
  for (int gi = NE; gi < NG; ++gi) {
    const double cg = pick<3>(cf.gw, gi) * dt * iFr2;
    double V[NQ], Fv[2 * DIM];
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] = -cg * R[DIM - 1][q] * jx[q];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = 0.0;
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * gi));
  }
}

// E3 adjoint: out[i] = (d rhs_p / d src[i])^T rbar. Explicit states (rho, u, p) = src[3j..3j+2],
// j < NE, weights w[j] (kinetic) and wt[j] (enthalpy); the kinetic-energy correction reads
// (src[3 NE], src[3 NE + 1]); gravity work reads (src[3g], src[3g+1]), g < NG (for stages 1 and 2
// the last gravity pair is the correction's pair). With rho (E - Ma^2 |u|^2 / 2) = p / (gamma - 1)
// the volume flux of state j is (w dt Ma^2 rho |u|^2 / 2 + wt dt gamma/(gamma-1) p) u and every face
// flux is c1 (avg_kin.n + lam/2 jump_kin) + c2 (avg_ent.n + lam/2 jump_rhoe), c1 = w dt Ma^2,
// c2 = wt dt. STAGE 2 keeps the reference's u_tmp_2 . u_old kinetic jump (euler_operator.h:1182-1183).
struct AdjE3Coef { double w[3], wt[3], gw[3]; };

// One launch per explicit state: jsel < NE differentiates state jsel, jsel == NE the kinetic-energy
// correction (one state's inputs live at a time keeps the 3D kernels free of register spills).
template <int DIM, int K, int NE, int NG, int STAGE>
__global__ __launch_bounds__(AFL_NT) void k_vjp_rhs_p(Geo g, const Tab* __restrict__ Tp, AdjSrcs io,
                                                      const double* __restrict__ rbar, AdjE3Coef cf, double dt,
                                                      double Ma2, double iFr2, int64_t N, int jsel) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  constexpr double gm1 = kGamma - 1.0, gr = kGamma / (kGamma - 1.0);
  constexpr int KI = 3 * NE;
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double nv[2 * DIM][DIM], jf[2 * DIM];
  face_normals<DIM, K>(g, T, s, nv, jf);
  double R[NQ], trR[2 * DIM][2], P[DIM][NQ], W[2 * DIM], jx[NQ];
  eval_line_faces<DIM, K>(rbar, s, nb, t1, t2, fn, T, R, trR);
  grad_jxw<DIM, K>(rbar, s, t1, T, g, lg, P);
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) W[f] = nb.c[f] >= 0 ? (trR[f][1] - trR[f][0]) * jf[f] : 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) jx[q] = lg.jxw(g, T, q);
  const double G0[DIM][NQ] = {};
  if (jsel < NE) {
    const int j = jsel;
    // stage 2: state 1's kinetic jump reads the traces of u_old (state 0's velocity), and that jump's
    // derivative with respect to u_old on the own faces goes to u_old's adjoint (state 0)
    double tru0[DIM][2 * DIM][2], Q0[2 * DIM][DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        Q0[f][d] = 0.0;
        tru0[d][f][0] = tru0[d][f][1] = 0.0;
      }
    if (STAGE == 2 && j == 1) {
      double v[NQ];
#pragma unroll
      for (int d = 0; d < DIM; ++d) eval_line_faces<DIM, K>(in_at(io, 1) + d * N, s, nb, t1, t2, fn, T, v, tru0[d]);
    }
    if (STAGE == 2 && j == 0) {
      double v[NQ], r1[NQ], trr1[2 * DIM][2], tru1[DIM][2 * DIM][2];
      eval_line_faces<DIM, K>(in_at(io, 3), s, nb, t1, t2, fn, T, r1, trr1);
#pragma unroll
      for (int d = 0; d < DIM; ++d) eval_line_faces<DIM, K>(in_at(io, 4) + d * N, s, nb, t1, t2, fn, T, v, tru1[d]);
      const double c1 = pick<3>(cf.w, 1) * dt * Ma2;
#pragma unroll
      for (int f = 0; f < 2 * DIM; ++f) {
        double a = 0.0, b = 0.0;
        for (int d = 0; d < DIM; ++d) { a += tru1[d][f][0] * nv[f][d]; b += tru1[d][f][1] * nv[f][d]; }
        const double lam = fmax(fabs(a), fabs(b));
        for (int d = 0; d < DIM; ++d) Q0[f][d] = W[f] * c1 * 0.5 * lam * trr1[f][0] * 0.5 * tru1[d][f][0];
      }
    }
    const double c1 = pick<3>(cf.w, j) * dt * Ma2, c2 = pick<3>(cf.wt, j) * dt;
    const double cg = (j < NG ? pick<3>(cf.gw, j) : 0.0) * dt * Ma2 * iFr2;
    const bool quirk = STAGE == 2 && j == 1;
    double r[NQ], trr[2 * DIM][2], u[DIM][NQ], tru[DIM][2 * DIM][2], pj[NQ], trp[2 * DIM][2];
    eval_line_faces<DIM, K>(in_at(io, 3 * j), s, nb, t1, t2, fn, T, r, trr);
#pragma unroll
    for (int d = 0; d < DIM; ++d) eval_line_faces<DIM, K>(in_at(io, 3 * j + 1) + d * N, s, nb, t1, t2, fn, T, u[d], tru[d]);
    eval_line_faces<DIM, K>(in_at(io, 3 * j + 2), s, nb, t1, t2, fn, T, pj, trp);
    // own-side face factors
    double una[2 * DIM], lam[2 * DIM], la[2 * DIM], uua[2 * DIM], ka[2 * DIM], jk[2 * DIM], jre[2 * DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double a = 0.0, b = 0.0, aa = 0.0, bb = 0.0, za = 0.0, zb = 0.0;
      for (int d = 0; d < DIM; ++d) {
        a += tru[d][f][0] * nv[f][d];
        b += tru[d][f][1] * nv[f][d];
        aa += tru[d][f][0] * tru[d][f][0];
        bb += tru[d][f][1] * tru[d][f][1];
        if (STAGE == 2) { za += tru[d][f][0] * tru0[d][f][0]; zb += tru[d][f][1] * tru0[d][f][1]; }
      }
      double lb;
      lam_weights(a, b, la[f], lb);
      una[f] = a;
      lam[f] = fmax(fabs(a), fabs(b));
      uua[f] = aa;
      ka[f] = quirk ? 0.5 * za : 0.5 * aa;
      const double kb = quirk ? 0.5 * zb : 0.5 * bb;
      jk[f] = trr[f][0] * ka[f] - trr[f][1] * kb;
      jre[f] = (trp[f][0] - trp[f][1]) / gm1;
    }
    double V[NQ], Fv[2 * DIM];
    // density
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double uu = 0.0, Gu = 0.0;
      for (int d = 0; d < DIM; ++d) { uu += u[d][q] * u[d][q]; Gu += u[d][q] * P[d][q]; }
      V[q] = 0.5 * c1 * uu * Gu + (j == 0 ? 0.5 * Ma2 * uu * R[q] * jx[q] : 0.0) - cg * u[DIM - 1][q] * R[q] * jx[q];
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = W[f] * c1 * (0.25 * uua[f] * una[f] + 0.5 * lam[f] * ka[f]);
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * j));
    // velocity components
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double uu = 0.0, Gu = 0.0;
        for (int d = 0; d < DIM; ++d) { uu += u[d][q] * u[d][q]; Gu += u[d][q] * P[d][q]; }
        const double aj = 0.5 * c1 * r[q] * uu + c2 * gr * pj[q];
        double v = aj * P[c][q] + c1 * r[q] * u[c][q] * Gu;
        if (j == 0) v += Ma2 * r[q] * u[c][q] * R[q] * jx[q];
        if (c == DIM - 1) v -= cg * r[q] * R[q] * jx[q];
        V[q] = v;
      }
#pragma unroll
      for (int f = 0; f < 2 * DIM; ++f) {
        const double ra = trr[f][0], uc = tru[c][f][0], n = nv[f][c];
        const double dka = quirk ? 0.5 * tru0[c][f][0] : uc;
        Fv[f] = W[f] * (c1 * (0.25 * (2.0 * uc * ra * una[f] + uua[f] * ra * n) + 0.5 * la[f] * n * jk[f] +
                              0.5 * lam[f] * ra * dka) +
                        c2 * 0.5 * (gr * trp[f][0] * n + la[f] * n * jre[f]));
        if (j == 0) Fv[f] += Q0[f][c];
      }
      finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * j + 1) + c * N);
    }
    // pressure
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double Gu = 0.0;
      for (int d = 0; d < DIM; ++d) Gu += u[d][q] * P[d][q];
      V[q] = c2 * gr * Gu + (j == 0 ? R[q] * jx[q] / gm1 : 0.0);
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = W[f] * c2 * 0.5 * (gr * una[f] + lam[f] / gm1);
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, 3 * j + 2));
  } else {
    // kinetic-energy correction (and, for stages 1 and 2, the last gravity pair, the same sources)
    const double cg = (NG > NE ? pick<3>(cf.gw, NE) : 0.0) * dt * Ma2 * iFr2;
    double r[NQ], u[DIM][NQ];
    eval_line<DIM, K>(in_at(io, KI), s, t1, t2, T, r);
#pragma unroll
    for (int d = 0; d < DIM; ++d) eval_line<DIM, K>(in_at(io, KI + 1) + d * N, s, t1, t2, T, u[d]);
    double V[NQ], Fv[2 * DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double uu = 0.0;
      for (int d = 0; d < DIM; ++d) uu += u[d][q] * u[d][q];
      V[q] = (-0.5 * Ma2 * uu - cg * u[DIM - 1][q]) * R[q] * jx[q];
    }
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, KI));
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] = (-Ma2 * r[q] * u[c][q] - (c == DIM - 1 ? cg * r[q] : 0.0)) * R[q] * jx[q];
      finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, out_at(io, KI + 1) + c * N);
    }
  }
}

// B^T y (y: DIM fields, out: one field); f = c dt / Ma^2 as in the forward B.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_apply_Bt(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ y, double fB,
                                                     int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double V[NQ], Fv[2 * DIM];
#pragma unroll
  for (int q = 0; q < NQ; ++q) V[q] = 0.0;
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) Fv[f] = 0.0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    double ya[NQ], tr[2 * DIM][2], P[DIM][NQ];
    eval_line_faces<DIM, K>(y + a * N, s, nb, t1, t2, fn, T, ya, tr);
    grad_jxw<DIM, K>(y + a * N, s, t1, T, g, lg, P);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] -= fB * (a == 0 ? P[0][q] : (a == 1 ? P[1][q] : P[DIM - 1][q]));
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      double n3[3], jf;
      face_geo<DIM, K>(g, T, s.ic, f, s.l, n3, jf);
      const double na = pick<3>(n3, a);
      // interior: 1/2 f n.(y_own - y_nb); wall (p_D = p): f n.y
      Fv[f] += (nb.c[f] >= 0 ? 0.5 * fB * (tr[f][0] - tr[f][1]) : fB * tr[f][0]) * na * jf;
    }
  }
  const double G0[DIM][NQ] = {};
  finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, dst);
}

// Normal components of a vector field's traces on every face: own [0] and neighbour [1].
template <int DIM, int K>
__device__ __attribute__((always_inline)) inline void normal_traces(const double* w, const Slot<DIM, K>& s, const Nbrs<DIM>& nb,
                                                                    double* t1, double* t2, double* fn, const Tab& T,
                                                                    const double (&nv)[2 * DIM][DIM], int64_t N,
                                                                    double (&wn)[2 * DIM][2]) {
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) wn[f][0] = wn[f][1] = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double v[2 * K + 1], tr[2 * DIM][2];
    eval_line_faces<DIM, K>(w + d * N, s, nb, t1, t2, fn, T, v, tr);
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      const double nd = pick<DIM>(nv[f], d);
      wn[f][0] += tr[f][0] * nd;
      wn[f][1] += tr[f][1] * nd;
    }
  }
}


// C'(w0) dw with the fixed pressure pf: the forward C with lambda replaced by its derivative.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_apply_C_tangent(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ pf,
                                                            const double* __restrict__ dw, const double* __restrict__ w0,
                                                            double cdt, int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  constexpr double gr = kGamma / (kGamma - 1.0), igm1 = 1.0 / (kGamma - 1.0);
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double nv[2 * DIM][DIM], jf[2 * DIM];
  face_normals<DIM, K>(g, T, s, nv, jf);
  double pv[NQ], trp[2 * DIM][2];
  eval_line_faces<DIM, K>(pf, s, nb, t1, t2, fn, T, pv, trp);
  double G[DIM][NQ], avg[2 * DIM], dn[2 * DIM][2];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) { avg[f] = 0.0; dn[f][0] = dn[f][1] = 0.0; }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double v[NQ], tr[2 * DIM][2];
    eval_line_faces<DIM, K>(dw + d * N, s, nb, t1, t2, fn, T, v, tr);
#pragma unroll
    for (int q = 0; q < NQ; ++q) G[d][q] = -cdt * gr * pv[q] * v[q];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      avg[f] += 0.5 * gr * (trp[f][0] * tr[f][0] + trp[f][1] * tr[f][1]) * nv[f][d];
      dn[f][0] += tr[f][0] * nv[f][d];
      dn[f][1] += tr[f][1] * nv[f][d];
    }
  }
  double wn[2 * DIM][2];
  normal_traces<DIM, K>(w0, s, nb, t1, t2, fn, T, nv, N, wn);
  double Fv[2 * DIM];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    double la, lb;
    lam_weights(wn[f][0], wn[f][1], la, lb);
    const double F = cdt * (avg[f] + 0.5 * (la * dn[f][0] + lb * dn[f][1]) * (trp[f][0] - trp[f][1]) * igm1);
    Fv[f] = nb.c[f] >= 0 ? F * jf[f] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double Gp[3] = {0, 0, 0}, Rr[3];
    for (int d = 0; d < DIM; ++d) Gp[d] = G[d][q];
    lg.gref(g, T, s.ic, q, Gp, Rr);
    for (int d = 0; d < DIM; ++d) G[d][q] = Rr[d];
  }
  const double V0[NQ] = {};
  finish_field<DIM, K, false, true>(s, i1, i2, Fq, T, V0, G, Fv, dst);
}

// C'(w0)^T cb -> DIM fields. SVM: minus the scaled velocity mass of cb in the same pass,
// dst_d -= int phi cb r (alpha u_d + beta [d = z]) (r, u: rs, us), as k_cell_products MODE 1.
template <int DIM, int K, bool SVM = false>
__global__ __launch_bounds__(AFL_NT) void k_apply_C_tangent_t(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ pf,
                                                              const double* __restrict__ cb, const double* __restrict__ w0,
                                                              double cdt, int64_t N, double* __restrict__ dst,
                                                              const double* __restrict__ rs = nullptr,
                                                              const double* __restrict__ us = nullptr, double alpha = 0.0,
                                                              double beta = 0.0) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  constexpr double gr = kGamma / (kGamma - 1.0), igm1 = 1.0 / (kGamma - 1.0);
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double nv[2 * DIM][DIM], jf[2 * DIM];
  face_normals<DIM, K>(g, T, s, nv, jf);
  double pv[NQ], trp[2 * DIM][2], cv[NQ], trc[2 * DIM][2], P[DIM][NQ], wn[2 * DIM][2];
  eval_line_faces<DIM, K>(pf, s, nb, t1, t2, fn, T, pv, trp);
  eval_line_faces<DIM, K>(cb, s, nb, t1, t2, fn, T, cv, trc);
  grad_jxw<DIM, K>(cb, s, t1, T, g, lg, P);
  normal_traces<DIM, K>(w0, s, nb, t1, t2, fn, T, nv, N, wn);
  double Wf[2 * DIM];  // c dt (C_own - C_nb) JxW_f 1/2 (gr p_own + la [p] / (gamma - 1))
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    double la, lb;
    lam_weights(wn[f][0], wn[f][1], la, lb);
    Wf[f] = nb.c[f] >= 0 ? cdt * (trc[f][0] - trc[f][1]) * jf[f] * 0.5 *
                               (gr * trp[f][0] + la * (trp[f][0] - trp[f][1]) * igm1)
                         : 0.0;
  }
  const double G0[DIM][NQ] = {};
  double cr[NQ];  // SVM: cb r JxW at the line's points
  if constexpr (SVM) {
    double r[NQ];
    eval_line<DIM, K>(rs, s, t1, t2, T, r);
#pragma unroll
    for (int q = 0; q < NQ; ++q) cr[q] = cv[q] * r[q] * lg.jxw(g, T, q);
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double V[NQ], Fv[2 * DIM];
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] = -cdt * gr * pv[q] * (d == 0 ? P[0][q] : (d == 1 ? P[1][q] : P[DIM - 1][q]));
    if constexpr (SVM) {
      double u[NQ];
      eval_line<DIM, K>(us + d * N, s, t1, t2, T, u);
      const double bz = d == DIM - 1 ? beta : 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] -= cr[q] * (alpha * u[q] + bz);
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fv[f] = Wf[f] * pick<DIM>(nv[f], d);
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, dst + d * N);
  }
}

// (d C(w; pf) / d pf)^T cb: C is linear in its coefficient pf for a fixed w (vmult_enthalpy,
// euler_operator.h:1411-1507); N -> N.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_apply_C_coef_t(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ w,
                                                           const double* __restrict__ cb, double cdt, int64_t N,
                                                           double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  constexpr double gr = kGamma / (kGamma - 1.0), igm1 = 1.0 / (kGamma - 1.0);
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double nv[2 * DIM][DIM], jf[2 * DIM];
  face_normals<DIM, K>(g, T, s, nv, jf);
  double cv[NQ], trc[2 * DIM][2], P[DIM][NQ], V[NQ];
  eval_line_faces<DIM, K>(cb, s, nb, t1, t2, fn, T, cv, trc);
  grad_jxw<DIM, K>(cb, s, t1, T, g, lg, P);
#pragma unroll
  for (int q = 0; q < NQ; ++q) V[q] = 0.0;
  double wn[2 * DIM][2];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) wn[f][0] = wn[f][1] = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double v[NQ], tr[2 * DIM][2];
    eval_line_faces<DIM, K>(w + d * N, s, nb, t1, t2, fn, T, v, tr);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] -= cdt * gr * v[q] * (d == 0 ? P[0][q] : (d == 1 ? P[1][q] : P[DIM - 1][q]));
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      const double nd = pick<DIM>(nv[f], d);
      wn[f][0] += tr[f][0] * nd;
      wn[f][1] += tr[f][1] * nd;
    }
  }
  double Fv[2 * DIM];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    const double lam = fmax(fabs(wn[f][0]), fabs(wn[f][1]));
    Fv[f] = nb.c[f] >= 0 ? cdt * (trc[f][0] - trc[f][1]) * jf[f] * 0.5 * (gr * wn[f][0] + lam * igm1) : 0.0;
  }
  const double G0[DIM][NQ] = {};
  finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, dst);
}

// The coefficient part of the adjoint stage operator in one pass:
//   dst = scaleD M cb + (dC(wE; pf)/dpf)^T cb - (dC(wB; pf)/dpf)^T cb - acc,
// with cb evaluated once for both velocities. acc may be dst: each row is read by its owner thread
// just before that thread writes it.
template <int DIM, int K>
__global__ __launch_bounds__(AFL_NT) void k_apply_C_coef_pair_t(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ wE,
                                                                const double* __restrict__ wB, const double* __restrict__ cb,
                                                                double cdt, double scaleD, const double* acc, int64_t N,
                                                                double* dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  constexpr double gr = kGamma / (kGamma - 1.0), igm1 = 1.0 / (kGamma - 1.0);
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  const Nbrs<DIM> nb(g, s.ic);
  double nv[2 * DIM][DIM], jf[2 * DIM];
  face_normals<DIM, K>(g, T, s, nv, jf);
  double cv[NQ], trc[2 * DIM][2], P[DIM][NQ], V[NQ];
  eval_line_faces<DIM, K>(cb, s, nb, t1, t2, fn, T, cv, trc);
  grad_jxw<DIM, K>(cb, s, t1, T, g, lg, P);
#pragma unroll
  for (int q = 0; q < NQ; ++q) V[q] = scaleD * cv[q] * lg.jxw(g, T, q);
  double wnE[2 * DIM][2], wnB[2 * DIM][2];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) wnE[f][0] = wnE[f][1] = wnB[f][0] = wnB[f][1] = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double v[NQ], tr[2 * DIM][2];
    eval_line_faces<DIM, K>(wE + d * N, s, nb, t1, t2, fn, T, v, tr);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] -= cdt * gr * v[q] * P[d][q];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      wnE[f][0] += tr[f][0] * nv[f][d];
      wnE[f][1] += tr[f][1] * nv[f][d];
    }
    eval_line_faces<DIM, K>(wB + d * N, s, nb, t1, t2, fn, T, v, tr);
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] += cdt * gr * v[q] * P[d][q];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      wnB[f][0] += tr[f][0] * nv[f][d];
      wnB[f][1] += tr[f][1] * nv[f][d];
    }
  }
  double Fv[2 * DIM];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    const double lamE = fmax(fabs(wnE[f][0]), fabs(wnE[f][1]));
    const double lamB = fmax(fabs(wnB[f][0]), fabs(wnB[f][1]));
    const double jc = cdt * (trc[f][0] - trc[f][1]) * jf[f] * 0.5;
    Fv[f] = nb.c[f] >= 0 ? jc * (gr * wnE[f][0] + lamE * igm1) - jc * (gr * wnB[f][0] + lamB * igm1) : 0.0;
  }
  const double G0[DIM][NQ] = {};
  cell_sync<DIM, K>();  // the last field's face reads are done
  if (s.act)
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) Fq[f * C::NFQ + s.l] = Fv[f];
  double out[K + 1];
  integrate_line<DIM, K, true, false>(s, i1, i2, T, V, G0, out);
  add_faces<DIM, K>(s, Fq, T, out, 0xffu);
  if (s.act && s.l < C::ROWS) {
    const int64_t r0 = s.cell * C::NPC + s.l * C::N1;
#pragma unroll
    for (int i = 0; i < C::N1; ++i) dst[r0 + i] = out[i] - acc[r0 + i];
  }
}

// Cell-local products integrated against the test functions:
//   MODE 0: dst = int phi (a . b)                     (a, b: DIM fields; d(A(rho) b)/d rho transposed)
//   MODE 1: dst_d = int phi a r (alpha u_d + beta [d = z])  (a, r: one field, u: DIM fields)
template <int DIM, int K, int MODE>
__global__ __launch_bounds__(AFL_NT) void k_cell_products(Geo g, const Tab* __restrict__ Tp, const double* __restrict__ a,
                                                          const double* __restrict__ b, const double* __restrict__ c,
                                                          double alpha, double beta, int64_t N, double* __restrict__ dst) {
  const auto& T = *Tp;
  __shared__ double shw[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::WF];
  __shared__ double shF[Cfg<DIM, K>::CPB * AdjLds<DIM, K>::F];
  AFL_ADJ_SLOT
  LineGeo<DIM, K> lg;
  lg.init(g, T, s.ic, s.l);
  double jx[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) jx[q] = lg.jxw(g, T, q);
  const double G0[DIM][NQ] = {};
  double Fv[2 * DIM];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) Fv[f] = 0.0;
  double V[NQ];
  if constexpr (MODE == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) V[q] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double x[NQ], y[NQ];
      eval_line<DIM, K>(a + d * N, s, t1, t2, T, x);
      eval_line<DIM, K>(b + d * N, s, t1, t2, T, y);
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] += x[q] * y[q] * jx[q];
    }
    finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, dst);
  } else {
    double x[NQ], r[NQ];
    eval_line<DIM, K>(a, s, t1, t2, T, x);
    eval_line<DIM, K>(b, s, t1, t2, T, r);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double u[NQ];
      eval_line<DIM, K>(c + d * N, s, t1, t2, T, u);
      const double bz = d == DIM - 1 ? beta : 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) V[q] = x[q] * r[q] * (alpha * u[q] + bz) * jx[q];
      finish_field<DIM, K, true, false>(s, i1, i2, Fq, T, V, G0, Fv, dst + d * N);
    }
  }
}

template <int DIM, int K>
dim3 adj_grid(int ncells) {
  using C = Cfg<DIM, K>;
  return dim3((ncells + C::CPB - 1) / C::CPB);
}

#define AFL_ADJ_DISPATCH(DIMV, KV, ...)                                \
  do {                                                                 \
    switch ((DIMV) * 10 + (KV)) {                                      \
      case 21: { constexpr int D_ = 2, K_ = 1; __VA_ARGS__; } break;   \
      case 22: { constexpr int D_ = 2, K_ = 2; __VA_ARGS__; } break;   \
      case 23: { constexpr int D_ = 2, K_ = 3; __VA_ARGS__; } break;   \
      case 24: { constexpr int D_ = 2, K_ = 4; __VA_ARGS__; } break;   \
      case 31: { constexpr int D_ = 3, K_ = 1; __VA_ARGS__; } break;   \
      case 32: { constexpr int D_ = 3, K_ = 2; __VA_ARGS__; } break;   \
      case 33: { constexpr int D_ = 3, K_ = 3; __VA_ARGS__; } break;   \
      case 34: { constexpr int D_ = 3, K_ = 4; __VA_ARGS__; } break;   \
      default: break;                                                  \
    }                                                                  \
  } while (0)

}  // namespace

void launch_vjp_rhs_rho(const OpCtx& o, int stage, const double* rbar, const double* const* src, double* const* out) {
  AdjSrcs io{};
  for (int i = 0; i < 2 * stage; ++i) { io.p[i] = src[i]; io.o[i] = out[i]; }
  const TableauC& t = o.tab;
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    const dim3 gr = adj_grid<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    for (int j = 0; j < stage; ++j) {
      if (stage == 1) k_vjp_rhs_rho<D_, K_, 1><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, o.dt, t.a21, 0.0, 0.0, o.N, j);
      else if (stage == 2) k_vjp_rhs_rho<D_, K_, 2><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, o.dt, t.a31, t.a32, 0.0, o.N, j);
      else k_vjp_rhs_rho<D_, K_, 3><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, o.dt, t.b1, t.b2, t.b3, o.N, j);
    }
  });
}

void launch_vjp_rhs_u(const OpCtx& o, int stage, const double* rbar, const double* const* src, double* const* out) {
  const int n = stage == 1 ? 4 : (stage == 2 ? 7 : 9);
  AdjSrcs io{};
  for (int i = 0; i < n; ++i) { io.p[i] = src[i]; io.o[i] = out[i]; }
  const TableauC& t = o.tab;
  AdjE2Coef cf{};
  if (stage == 1) {
    cf.w[0] = t.a21; cf.wp[0] = t.at21; cf.gw[0] = t.at21; cf.gw[1] = t.at22;
  } else if (stage == 2) {
    cf.w[0] = t.a31; cf.w[1] = t.a32; cf.wp[0] = t.at31; cf.wp[1] = t.at32;
    cf.gw[0] = t.at31; cf.gw[1] = t.at32; cf.gw[2] = t.at33;
  } else {
    cf.w[0] = t.b1; cf.w[1] = t.b2; cf.w[2] = t.b3; cf.wp[0] = t.b1; cf.wp[1] = t.b2; cf.wp[2] = t.b3;
    cf.gw[0] = t.b1; cf.gw[1] = t.b2; cf.gw[2] = t.b3;
  }
  const double iMa2 = 1.0 / (o.Ma * o.Ma), iFr2 = 1.0 / (o.Fr * o.Fr);
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    const dim3 gr = adj_grid<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    // states 0..NE-1, then (stages 1 and 2) the gravity-only density
    const int nl = stage == 3 ? 3 : stage + 1;
    for (int j = 0; j < nl; ++j) {
      if (stage == 1) k_vjp_rhs_u<D_, K_, 1, 2><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, cf, o.dt, iMa2, iFr2, o.N, j);
      else if (stage == 2) k_vjp_rhs_u<D_, K_, 2, 3><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, cf, o.dt, iMa2, iFr2, o.N, j);
      else k_vjp_rhs_u<D_, K_, 3, 3><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, cf, o.dt, iMa2, iFr2, o.N, j);
    }
  });
}

void launch_vjp_rhs_p(const OpCtx& o, int stage, const double* rbar, const double* const* src, double* const* out) {
  const int n = stage == 1 ? 5 : (stage == 2 ? 8 : 11);
  AdjSrcs io{};
  for (int i = 0; i < n; ++i) { io.p[i] = src[i]; io.o[i] = out[i]; }
  const TableauC& t = o.tab;
  AdjE3Coef cf{};
  if (stage == 1) {
    cf.w[0] = t.a21; cf.wt[0] = t.at21; cf.gw[0] = t.at21; cf.gw[1] = t.at22;
  } else if (stage == 2) {
    cf.w[0] = t.a31; cf.wt[0] = t.at31; cf.w[1] = t.a32; cf.wt[1] = t.at32;
    cf.gw[0] = t.at31; cf.gw[1] = t.at32; cf.gw[2] = t.at33;
  } else {
    cf.w[0] = t.b1; cf.wt[0] = t.b1; cf.w[1] = t.b2; cf.wt[1] = t.b2; cf.w[2] = t.b3; cf.wt[2] = t.b3;
    cf.gw[0] = t.b1; cf.gw[1] = t.b2; cf.gw[2] = t.b3;
  }
  const double Ma2 = o.Ma * o.Ma, iFr2 = 1.0 / (o.Fr * o.Fr);
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    const dim3 gr = adj_grid<D_, K_>(o.geo.ncells), bl(Cfg<D_, K_>::NT);
    for (int j = 0; j <= stage; ++j) {
      if (stage == 1) k_vjp_rhs_p<D_, K_, 1, 2, 1><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, cf, o.dt, Ma2, iFr2, o.N, j);
      else if (stage == 2) k_vjp_rhs_p<D_, K_, 2, 3, 2><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, cf, o.dt, Ma2, iFr2, o.N, j);
      else k_vjp_rhs_p<D_, K_, 3, 3, 3><<<gr, bl, 0, o.stream>>>(o.geo, o.dT, io, rbar, cf, o.dt, Ma2, iFr2, o.N, j);
    }
  });
}

void launch_apply_Bt(const OpCtx& o, const double* y, double* dst) {
  const double coeff = o.hyp == 1 ? o.tab.at22 : o.tab.at33;
  const double fB = coeff * o.dt / (o.Ma * o.Ma);
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_apply_Bt<D_, K_><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(o.geo, o.dT, y, fB, o.N, dst);
  });
}

void launch_apply_C_coef_t(const OpCtx& o, const double* w, const double* cb, double* dst) {
  const double cdt = (o.hyp == 1 ? o.tab.at22 : o.tab.at33) * o.dt;
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_apply_C_coef_t<D_, K_><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(o.geo, o.dT, w, cb, cdt,
                                                                                                     o.N, dst);
  });
}

void launch_apply_C_coef_pair_t(const OpCtx& o, const double* wE, const double* wB, const double* cb, double scaleD,
                                const double* acc, double* dst) {
  const double cdt = (o.hyp == 1 ? o.tab.at22 : o.tab.at33) * o.dt;
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_apply_C_coef_pair_t<D_, K_><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(
        o.geo, o.dT, wE, wB, cb, cdt, scaleD, acc, o.N, dst);
  });
}

void launch_dot_mass(const OpCtx& o, const double* a, const double* b, double* dst) {
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_cell_products<D_, K_, 0><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(
        o.geo, o.dT, a, b, nullptr, 0.0, 0.0, o.N, dst);
  });
}

void launch_scaled_velocity_mass(const OpCtx& o, const double* lam, const double* rho, const double* u, double alpha,
                                 double beta, double* dst) {
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_cell_products<D_, K_, 1><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(
        o.geo, o.dT, lam, rho, u, alpha, beta, o.N, dst);
  });
}

void launch_apply_C_tangent(const OpCtx& o, const double* pf, const double* dw, const double* w0, double* dst) {
  const double cdt = (o.hyp == 1 ? o.tab.at22 : o.tab.at33) * o.dt;
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_apply_C_tangent<D_, K_><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(o.geo, o.dT, pf, dw, w0,
                                                                                                      cdt, o.N, dst);
  });
}

void launch_apply_C_tangent_t(const OpCtx& o, const double* pf, const double* cb, const double* w0, double* dst) {
  const double cdt = (o.hyp == 1 ? o.tab.at22 : o.tab.at33) * o.dt;
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_apply_C_tangent_t<D_, K_><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(o.geo, o.dT, pf, cb, w0,
                                                                                                        cdt, o.N, dst);
  });
}

void launch_apply_C_tangent_t_svm(const OpCtx& o, const double* pf, const double* cb, const double* w0, const double* rho,
                                  const double* u, double alpha, double beta, double* dst) {
  const double cdt = (o.hyp == 1 ? o.tab.at22 : o.tab.at33) * o.dt;
  AFL_ADJ_DISPATCH(o.dim, o.k, {
    k_apply_C_tangent_t<D_, K_, true><<<adj_grid<D_, K_>(o.geo.ncells), dim3(Cfg<D_, K_>::NT), 0, o.stream>>>(
        o.geo, o.dT, pf, cb, w0, cdt, o.N, dst, rho, u, alpha, beta);
  });
}

}  // namespace afl
