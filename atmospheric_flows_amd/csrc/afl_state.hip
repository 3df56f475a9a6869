// Initial state of the hill problems and the Rayleigh sponge, evaluated per node on the
// device. Background: constant Brunt-Vaisala frequency N, hydrostatic, closed form
//   theta = theta0 exp(N^2 z / g),  pi = 1 + g^2/(cp theta0 N^2) (exp(-N^2 z/g) - 1),
//   p = p0 pi^(cp/R),  rho = p / (R pi theta);
// perturbation: theta' = sigma * N(0,1) per node from splitmix64 + Box-Muller, keyed by
// (seed, global node index) so the field does not depend on the launch shape.
#include "afl_internal.h"

namespace afl {

static constexpr double kR = 287.0;      // equation_data.h:25
static constexpr double kGrav = 9.80616; // equation_data.h:28

__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline double unit_open(uint64_t r) { return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// physical coordinates of node i of cell c
template <int DIM>
__device__ inline void node_xyz(const Geo& g, const Tab& T, int n1, int64_t c, int i, double* X) {
  int ic[3];
  ic[0] = (int)(c % g.nc[0]);
  if (DIM == 2) { ic[1] = (int)(c / g.nc[0]); ic[2] = 0; }
  else { ic[1] = (int)((c / g.nc[0]) % g.nc[1]); ic[2] = (int)(c / ((int64_t)g.nc[0] * g.nc[1])); }
  const int ii[3] = {i % n1, (i / n1) % n1, i / (n1 * n1)};
  double x = g.lo[0] + (ic[0] + T.ND[ii[0]]) * g.hc[0];
  double y = DIM == 3 ? g.lo[1] + (ic[1] + T.ND[ii[1]]) * g.hc[1] : 0.0;
  const double zeta = (ic[DIM - 1] + T.ND[ii[DIM - 1]]) * g.hc[DIM - 1];
  double h, hx, hy;
  hill(g, x, y, h, hx, hy);
  X[0] = x;
  if (DIM == 3) X[1] = y;
  X[DIM - 1] = zeta + h * (1.0 - zeta / g.H);
}

template <int DIM>
__global__ void k_init_state(Geo g, const Tab* __restrict__ Tp, int n1, int npc, int64_t N, HillParams hp, double* __restrict__ rho,
                             double* __restrict__ u, double* __restrict__ p, double* __restrict__ th_out, bool noise) {
  const auto& T = *Tp;
  const double cp = kGamma * kR / (kGamma - 1.0);
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
    double X[3];
    node_xyz<DIM>(g, T, n1, idx / npc, (int)(idx % npc), X);
    const double z = X[DIM - 1];
    const double th = hp.theta0 * exp(hp.N2 * z / kGrav);
    const double pi = 1.0 + kGrav * kGrav / (cp * hp.theta0 * hp.N2) * (exp(-hp.N2 * z / kGrav) - 1.0);
    const double pb = hp.p0 * pow(pi, cp / kR);
    double theta = th;
    if (noise) {
      const uint64_t r1 = mix64(hp.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(2 * idx + 1)));
      const uint64_t r2 = mix64(r1);
      theta += hp.sigma * sqrt(-2.0 * log(unit_open(r1))) * cos(2.0 * M_PI * unit_open(r2));
    }
    p[idx] = pb;
    rho[idx] = pb / (kR * pi * theta);
    if (u) {
      u[idx] = hp.u0;
      for (int d = 1; d < DIM; ++d) u[d * N + idx] = 0.0;
    }
    if (th_out) th_out[idx] = th;
  }
}

template <int DIM>
__global__ void k_sponge(Geo g, const Tab* __restrict__ Tp, int n1, int npc, int64_t N, double zs, double adt, double u0, double u1,
                         double u2, double* __restrict__ u) {
  const auto& T = *Tp;
  const double ur[3] = {u0, u1, u2};
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
    double X[3];
    node_xyz<DIM>(g, T, n1, idx / npc, (int)(idx % npc), X);
    const double z = X[DIM - 1];
    if (z <= zs) continue;
    const double sn = sin(0.5 * M_PI * (z - zs) / (g.H - zs));
    const double a = adt * sn * sn;
    for (int d = 0; d < DIM; ++d) u[d * N + idx] = (u[d * N + idx] + a * ur[d]) / (1.0 + a);
  }
}

static int grid_nodes(int64_t n) {
  int64_t b = (n + AFL_NT - 1) / AFL_NT;
  return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

void launch_init_state(const OpCtx& o, const HillParams& hp, double* rho, double* u, double* p) {
  const int n1 = o.k + 1, npc = o.dim == 2 ? n1 * n1 : n1 * n1 * n1;
  if (o.dim == 2)
    k_init_state<2><<<dim3(grid_nodes(o.N)), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.dT, n1, npc, o.N, hp, rho, u, p, nullptr, true);
  else
    k_init_state<3><<<dim3(grid_nodes(o.N)), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.dT, n1, npc, o.N, hp, rho, u, p, nullptr, true);
}

void launch_background(const OpCtx& o, const HillParams& hp, double* rho, double* p, double* th) {
  const int n1 = o.k + 1, npc = o.dim == 2 ? n1 * n1 : n1 * n1 * n1;
  if (o.dim == 2)
    k_init_state<2><<<dim3(grid_nodes(o.N)), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.dT, n1, npc, o.N, hp, rho, nullptr, p, th, false);
  else
    k_init_state<3><<<dim3(grid_nodes(o.N)), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.dT, n1, npc, o.N, hp, rho, nullptr, p, th, false);
}

void launch_sponge(const OpCtx& o, double z_start, double alpha_dt, const double* uref, double* u) {
  const int n1 = o.k + 1, npc = o.dim == 2 ? n1 * n1 : n1 * n1 * n1;
  if (o.dim == 2)
    k_sponge<2><<<dim3(grid_nodes(o.N)), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.dT, n1, npc, o.N, z_start, alpha_dt, uref[0], uref[1], uref[2], u);
  else
    k_sponge<3><<<dim3(grid_nodes(o.N)), dim3(AFL_NT), 0, o.stream>>>(o.geo, o.dT, n1, npc, o.N, z_start, alpha_dt, uref[0], uref[1], uref[2], u);
}

}  // namespace afl
