// Observation operator: the potential-temperature perturbation theta' = theta(rho, p) - theta_bar(z)
// at fixed physical points, and its transpose (the misfit of the adjoint configurations; the
// reference has no point evaluation). rho and p are evaluated at a point from the cell's nodal
// values with the tensor-product Lagrange basis on the support points;
// theta = p / (rho R (p / p0)^(R/cp)) with the background's reference pressure p0,
// theta_bar = theta0 exp(N^2 z / g).
#include "afl_internal.h"

#include <algorithm>
#include <cmath>

namespace afl {

namespace {

__device__ __host__ inline double lagrange_1d(const double* nd, int n1, int j, double x) {
  double v = 1.0;
  for (int m = 0; m < n1; ++m)
    if (m != j) v *= (x - nd[m]) / (nd[j] - nd[m]);
  return v;
}

// values of rho and p at point k (cell c, reference coordinates r)
__device__ inline void point_values(const Tab& T, int dim, int n1, int64_t c, const double* r, const double* rho,
                                    const double* p, double& rv, double& pv) {
  double bx[5], by[5], bz[5];
  for (int j = 0; j < n1; ++j) {
    bx[j] = lagrange_1d(T.ND, n1, j, r[0]);
    by[j] = dim == 3 ? lagrange_1d(T.ND, n1, j, r[1]) : 1.0;
    bz[j] = lagrange_1d(T.ND, n1, j, r[dim - 1]);
  }
  const int npc = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const double* rc = rho + c * npc;
  const double* pc = p + c * npc;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < npc; ++i) {
    const int ix = i % n1, iy = dim == 3 ? (i / n1) % n1 : 0, iz = dim == 2 ? i / n1 : i / (n1 * n1);
    const double w = bx[ix] * (dim == 3 ? by[iy] : 1.0) * bz[iz];
    a += w * rc[i];
    b += w * pc[i];
  }
  rv = a;
  pv = b;
}

constexpr double kRd = 287.0, kG = 9.80616;

__global__ void k_observe(const Tab* __restrict__ Tp, int dim, int n1, int npts, const int64_t* __restrict__ cell,
                          const double* __restrict__ ref, const double* __restrict__ zpt, const double* __restrict__ rho,
                          const double* __restrict__ p, double theta0, double n2, double p0, double* __restrict__ out,
                          const double* __restrict__ wbar, double* __restrict__ fr, double* __restrict__ fp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  const double kap = (kGamma - 1.0) / kGamma;
  double rv, pv;
  point_values(*Tp, dim, n1, cell[k], ref + 3 * k, rho, p, rv, pv);
  const double th = pv / (rv * kRd * pow(pv / p0, kap));
  if (out) out[k] = th - theta0 * exp(n2 * zpt[k] / kG);
  if (wbar) {  // derivative factors of theta in rho and p, weighted by the point adjoint
    fr[k] = -wbar[k] * th / rv;
    fp[k] = wbar[k] * th * (1.0 - kap) / pv;
  }
}

// transpose scatter: one thread per (cell with points, node); the cell's points are contiguous
__global__ void k_observe_t(const Tab* __restrict__ Tp, int dim, int n1, int ncp, const int64_t* __restrict__ ccell,
                            const int* __restrict__ cstart, const int* __restrict__ order,
                            const double* __restrict__ ref, const double* __restrict__ fr,
                            const double* __restrict__ fp, double* __restrict__ rbar, double* __restrict__ pbar) {
  const int npc = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)ncp * npc) return;
  const int j = (int)(t / npc), i = (int)(t % npc);
  const int ix = i % n1, iy = dim == 3 ? (i / n1) % n1 : 0, iz = dim == 2 ? i / n1 : i / (n1 * n1);
  double a = 0.0, b = 0.0;
  for (int q = cstart[j]; q < cstart[j + 1]; ++q) {
    const int k = order[q];
    const double* r = ref + 3 * k;
    const double w = lagrange_1d(Tp->ND, n1, ix, r[0]) * (dim == 3 ? lagrange_1d(Tp->ND, n1, iy, r[1]) : 1.0) *
                     lagrange_1d(Tp->ND, n1, iz, r[dim - 1]);
    a += w * fr[k];
    b += w * fp[k];
  }
  rbar[ccell[j] * npc + i] = a;
  pbar[ccell[j] * npc + i] = b;
}

}  // namespace

// Host-side location of the points: cell and reference coordinates on the terrain-following map
// z = zeta + h(x, y) (1 - zeta / H). Returns the index of the first point outside the domain, or -1.
int locate_points(const Geo& g, int dim, int npts, const double* xyz, std::vector<int64_t>& cell,
                  std::vector<double>& ref, std::vector<double>& zpt) {
  cell.resize(npts);
  ref.assign((size_t)3 * npts, 0.0);
  zpt.resize(npts);
  for (int k = 0; k < npts; ++k) {
    const double* X = xyz + (size_t)dim * k;
    int ic[3] = {0, 0, 0};
    double r2 = 0.0;
    for (int d = 0; d < dim - 1; ++d) {
      const double s = (X[d] - g.lo[d]) / g.hc[d];
      if (!(s >= 0.0 && s <= g.nc[d])) return k;
      ic[d] = std::min((int)std::floor(s), g.nc[d] - 1);
      ref[3 * k + d] = s - ic[d];
      r2 += X[d] * X[d];
    }
    const double h = g.h0 / (1.0 + r2 * g.inv_a2);
    const double zeta = (X[dim - 1] - h) / (1.0 - h / g.H);
    const double s = zeta / g.hc[dim - 1];
    if (!(s >= 0.0 && s <= g.nc[dim - 1])) return k;
    ic[dim - 1] = std::min((int)std::floor(s), g.nc[dim - 1] - 1);
    ref[3 * k + dim - 1] = s - ic[dim - 1];
    cell[k] = dim == 2 ? ic[0] + (int64_t)g.nc[0] * ic[1]
                       : ic[0] + (int64_t)g.nc[0] * (ic[1] + (int64_t)g.nc[1] * ic[2]);
    zpt[k] = X[dim - 1];
  }
  return -1;
}

void launch_observe(const OpCtx& o, const ObsDev& ob, const double* rho, const double* p, double theta0, double n2,
                    double p0, double* out, const double* wbar, double* fr, double* fp) {
  const int nb = (ob.npts + 255) / 256;
  k_observe<<<dim3(nb), dim3(256), 0, o.stream>>>(o.dT, o.dim, o.k + 1, ob.npts, ob.cell, ob.ref, ob.z, rho, p, theta0,
                                                 n2, p0, out, wbar, fr, fp);
}

void launch_observe_t(const OpCtx& o, const ObsDev& ob, const double* fr, const double* fp, double* rbar,
                      double* pbar) {
  const int n1 = o.k + 1, npc = o.dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const int64_t work = (int64_t)ob.ncp * npc;
  if (work == 0) return;
  k_observe_t<<<dim3((unsigned)((work + 255) / 256)), dim3(256), 0, o.stream>>>(o.dT, o.dim, n1, ob.ncp, ob.ccell,
                                                                               ob.cstart, ob.order, ob.ref, fr, fp,
                                                                               rbar, pbar);
}

}  // namespace afl
