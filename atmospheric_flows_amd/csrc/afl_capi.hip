// C-ABI of libafl_hip.so (declarations and reference citations in include/afl.h).
// Host side of the drop-in: EULEROperator state, NS_stage dispatch, the device-resident
// Schur GMRES and the IMEX-ARK2 step driver. All vectors live on the device; the host only
// sees reduction scalars (one small copy per Krylov iteration / fixed-point iteration).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/afl.h"
#include "afl_internal.h"
#include "afl_mg.h"

using namespace afl;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_device(const char* where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(AFL_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
  return AFL_OK;
}

// ---- 1D tables (host) -------------------------------------------------------------------
void legendre(int n, double x, double& p, double& dp) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { p = 1.0; dp = 0.0; return; }
  for (int j = 2; j <= n; ++j) {
    const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
    p0 = p1;
    p1 = p2;
  }
  p = p1;
  dp = n * (x * p1 - p0) / (x * x - 1.0);
}

// Gauss-Legendre points/weights on [0,1] (Newton on P_n from Chebyshev-like guesses)
void gauss01(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double t = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(n, t, p, dp);
      const double d = p / dp;
      t -= d;
      if (std::fabs(d) < 1e-16) break;
    }
    double p, dp;
    legendre(n, t, p, dp);
    x[i] = 0.5 * (t + 1.0);
    w[i] = 1.0 / ((1.0 - t * t) * dp * dp);
  }
}

// support points: Gauss-Lobatto (end points + roots of P'_k) or equidistant, on [0,1]
void nodes01(int k, int type, double* x) {
  if (type == AFL_NODES_EQUIDISTANT) {
    for (int i = 0; i <= k; ++i) x[i] = double(i) / k;
    return;
  }
  x[0] = 0.0;
  x[k] = 1.0;
  for (int i = 1; i < k; ++i) {
    double t = -std::cos(M_PI * i / k);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(k, t, p, dp);
      const double d2p = (2.0 * t * dp - k * (k + 1.0) * p) / (1.0 - t * t);
      const double d = dp / d2p;
      t -= d;
      if (std::fabs(d) < 1e-16) break;
    }
    x[i] = 0.5 * (t + 1.0);
  }
}

double lag(const double* xn, int n1, int i, double x) {
  double v = 1.0;
  for (int j = 0; j < n1; ++j)
    if (j != i) v *= (x - xn[j]) / (xn[i] - xn[j]);
  return v;
}

double lag_d(const double* xn, int n1, int i, double x) {
  double s = 0.0;
  for (int m = 0; m < n1; ++m) {
    if (m == i) continue;
    double v = 1.0 / (xn[i] - xn[m]);
    for (int j = 0; j < n1; ++j)
      if (j != i && j != m) v *= (x - xn[j]) / (xn[i] - xn[j]);
    s += v;
  }
  return s;
}

// 1D tables of degree k on nq Gauss points (default 2k + 1, the operator's rule; the reference
// mass matrix is exact for nq >= k + 1)
void build_tab(int k, int node_type, Tab& T, int nq = 0) {
  std::memset(&T, 0, sizeof(T));
  const int n1 = k + 1;
  if (nq <= 0) nq = 2 * k + 1;
  nodes01(k, node_type, T.ND);
  gauss01(nq, T.QP, T.W);
  for (int q = 0; q < nq; ++q)
    for (int i = 0; i < n1; ++i) {
      T.S[q][i] = lag(T.ND, n1, i, T.QP[q]);
      T.D[q][i] = lag_d(T.ND, n1, i, T.QP[q]);
    }
  // inverse of the reference 1D mass matrix by Gauss-Jordan
  double A[5][10];
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) {
      double m = 0.0;
      for (int q = 0; q < nq; ++q) m += T.W[q] * T.S[q][i] * T.S[q][j];
      A[i][j] = m;
      A[i][n1 + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < n1; ++c) {
    int piv = c;
    for (int r = c + 1; r < n1; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    for (int j = 0; j < 2 * n1; ++j) std::swap(A[c][j], A[piv][j]);
    const double d = A[c][c];
    for (int j = 0; j < 2 * n1; ++j) A[c][j] /= d;
    for (int r = 0; r < n1; ++r)
      if (r != c) {
        const double f = A[r][c];
        for (int j = 0; j < 2 * n1; ++j) A[r][j] -= f * A[c][j];
      }
  }
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) T.MI[i][j] = A[i][n1 + j];
}

TableauC make_tableau() {
  TableauC t{};
  const double g = 2.0 - std::sqrt(2.0);
  t.a21 = g; t.a22 = 0.0; t.a31 = 0.5; t.a32 = 0.5; t.a33 = 0.0;
  t.at21 = 0.5 * g; t.at22 = 0.5 * g;
  t.at31 = std::sqrt(2.0) / 4.0; t.at32 = std::sqrt(2.0) / 4.0; t.at33 = 1.0 - std::sqrt(2.0) / 2.0;
  t.b1 = t.at31; t.b2 = t.at32; t.b3 = t.at33;
  return t;
}

// Device buffer owned by the handle (or by a scope): released by its destructor, so an early
// return after a failed allocation leaks nothing. Move-only.
struct DevBuf {
  double* p = nullptr;
  int64_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  int alloc(int64_t count) {
    if (p && n >= count) return AFL_OK;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, (size_t)(count > 0 ? count : 1) * sizeof(double)) != hipSuccess)
      return fail(AFL_ERR_DEVICE, "hipMalloc failed");
    n = count;
    return AFL_OK;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct afl_operator {
  OpCtx o;
  unsigned ns = 1;
  afl_solver_params sp{1e-8, 30000, 28, 1e-10, 100, 0, AFL_SCHUR_JACOBI, 3, 0, 1, 15.0, 0.05};
  // p-multigrid preconditioner state
  MgT<float> mg32;
  MgT<double> mg64;
  int mg_built = 0;          // precision of the built hierarchy (0: none)
  bool mg_eig_stale = true;  // rho_for_fixed changed since the last eigenvalue estimate
  bool mg_coef_stale = true; // rho_for_fixed / pres_fixed changed since the levels copied them
  Tab tabs_by_k[5]{};        // 1D tables per degree (same node family)
  Tab tabs_rq_by_k[5]{};     // the same on k + 1 Gauss points (multigrid level quadrature)
  DevBuf gZ, mgw, diagw;
  double min_diam = 0.0;  // GridTools::minimal_cell_diameter of the mapped mesh
  afl_sponge sponge{15000.0, 0.0, {0.0, 0.0, 0.0}, 0, 0};
  DevBuf rho_fix, p_fix, u_fix, inv_diag, dinv_D, red, dsc, wS, gV, gW, ydev;
  int diag_comps = 0;
  bool dinv_ready = false;
  // IMEX work vectors
  DevBuf rho2, rho3, rhoc, u2, u3, uc, p2, p3, pc, rhs_r, rhs_u, rhs_p, tmp1, tmp2, Bp, pfix, pfix_old, ufix;
  DevBuf rhs_u0, rhs_p0;  // fixed-point loop invariants: E2 and the u_fixed-free part of E3
  // adjoint work vectors (Schur tangent and its transpose; step adjoint)
  DevBuf adjW0, adjT1, adjT2, adjC;
  static constexpr int kAdjBufs = 34;
  DevBuf aj[kAdjBufs];
  // last pressure multiplier of each implicit stage: the initial guess of the next adjoint stage
  // solve (adjoint variables change little from one step to the previous)
  DevBuf adj_lam_prev[2];
  bool adj_lam_valid[2] = {false, false};
  int64_t adj_krylov_its = 0;  // Krylov iterations of the adjoint stage solves since the last reset
  bool adj_keep_stages = true;  // afl_adjoint_run keeps stage sets (afl_set_adjoint_stage_sets)
  // observation points (device copies of their location) and point scratch
  ObsDev obs;
  DevBuf obs_f;
  void release_obs() {
    hipFree(obs.cell); hipFree(obs.ref); hipFree(obs.z); hipFree(obs.ccell); hipFree(obs.cstart); hipFree(obs.order);
    obs = ObsDev{};
    obs_f.release();
  }
  double* host_sc = nullptr;  // pinned scalars
  Tab* dtab = nullptr;        // device copy of the basis tables (o.dT)
  int verbose = 0;            // progress lines on stderr (afl_set_verbose)
  KernelTimer timer;          // per-kernel HIP-event timing (afl_kernel_timer_*)
  // phase timing of afl_imex_step (afl_section_timer_*): segment s spans events 2s, 2s+1
  struct SectionTimer {
    static constexpr size_t kMaxSegments = 1 << 16;
    bool on = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> sec;
    size_t nseg = 0;
    // opens a segment of `section`: returns its index, or -1 when off or full
    int open(hipStream_t st, int section) {
      if (!on || nseg >= kMaxSegments) return -1;
      while (ev.size() < 2 * (nseg + 1)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ev.push_back(e);
      }
      if (sec.size() < nseg + 1) sec.resize(nseg + 1);
      sec[nseg] = section;
      hipEventRecord(ev[2 * nseg], st);
      return (int)nseg++;
    }
    void close(hipStream_t st, int s) {
      if (s >= 0) hipEventRecord(ev[2 * (size_t)s + 1], st);
    }
  } sect;
  std::vector<double> timer_ms;  // durations of the timed launches of the last timing window
  ~afl_operator() {
    DevBuf* all[] = {&rho_fix, &p_fix, &u_fix, &inv_diag, &dinv_D, &red, &dsc, &wS, &gV, &gW, &ydev, &gZ, &mgw, &diagw,
                     &rho2, &rho3, &rhoc, &u2, &u3, &uc, &p2, &p3, &pc, &rhs_r, &rhs_u, &rhs_p,
                     &tmp1, &tmp2, &Bp, &pfix, &pfix_old, &ufix, &rhs_u0, &rhs_p0, &adjW0, &adjT1, &adjT2, &adjC};
    for (DevBuf* b : all) b->release();
    for (DevBuf& b : aj) b.release();
    for (DevBuf& b : adj_lam_prev) b.release();
    release_obs();
    if (host_sc) hipHostFree(host_sc);
    if (dtab) hipFree(dtab);
    for (hipEvent_t e : timer.ev) hipEventDestroy(e);
    for (hipEvent_t e : sect.ev) hipEventDestroy(e);
  }
  int64_t N() const { return o.N; }
  int dim() const { return o.dim; }
};

namespace {

// minimal_cell_diameter on the terrain-following map: min over cells of the longest diagonal
// between opposite vertices (deal.II's CellAccessor::diameter for quadrilaterals / hexahedra)
double min_cell_diameter(const afl_mesh_desc* md) {
  const int d3 = md->dim == 3;
  const int nx = md->n_cells[0], ny = d3 ? md->n_cells[1] : 1, nz = md->n_cells[2];
  const double hx = (md->x_max - md->x_min) / nx, hy = d3 ? (md->y_max - md->y_min) / ny : 0.0,
               hz = md->z_top / nz;
  const double ia2 = 1.0 / (md->hill_width * md->hill_width);
  auto zmap = [&](double x, double y, double zeta) {
    const double h = md->hill_height / (1.0 + (x * x + y * y) * ia2);
    return zeta + h * (1.0 - zeta / md->z_top);
  };
  double best = 1e300;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const double x0 = md->x_min + ix * hx, y0 = d3 ? md->y_min + iy * hy : 0.0, z0 = iz * hz;
        double diam = 0.0;
        // opposite vertex pairs: (a, b, 0) <-> (1-a, 1-b, 1)
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < (d3 ? 2 : 1); ++b) {
            const double xa = x0 + a * hx, ya = y0 + b * hy, xb = x0 + (1 - a) * hx, yb = y0 + (d3 ? (1 - b) * hy : 0.0);
            const double za = zmap(xa, ya, z0), zb = zmap(xb, yb, z0 + hz);
            const double dx = xb - xa, dy = yb - ya, dz = zb - za;
            diam = std::max(diam, std::sqrt(dx * dx + dy * dy + dz * dz));
          }
        best = std::min(best, diam);
      }
  return best;
}

// ---- device scratch of the allocation-free entry point ---------------------------------------
// afl_l2_difference_noalloc keeps its partial sums here instead of in operator buffers: at most
// kL2ScratchBlocks per-block pairs (the partial pass strides over the cell blocks, so any mesh fits)
// and the two finished sums, 32.8 KB. Each GPU holds its own copy until the process ends. Host code
// never passes the array's name to a kernel or a copy (in host code it is a stand-in, not a device
// address): the address of this device's copy is looked up once per device and kept in
// g_scratch_dev. Calls that overlap on one GPU share the array and race.
constexpr int kL2ScratchBlocks = 2048;
constexpr int kL2ScratchSize = 2 * kL2ScratchBlocks + 2;
static __device__ __align__(256) double scratch[kL2ScratchSize];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// the current device's copy of scratch (looked up on the first call for that device)
int current_device_scratch(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return fail(AFL_ERR_DEVICE, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return fail(AFL_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(AFL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return AFL_OK;
}

constexpr int kMaxRestart = 60;
// device scalar slots: [0,64) first CGS pass, [64,128) second pass, 128 norm, 129.. misc
// device/pinned scalar slots: CGS2 columns, norms, reduction results, and the Krylov coefficients y
// (written by the host into pinned memory; the stream reads them before any later fetch returns)
constexpr int kScH1 = 0, kScH2 = 64, kScNrm = 128, kScA = 129, kScB = 130, kScY = 160, kScSize = 224;

int ensure_common(afl_operator* op) {
  int rc;
  if ((rc = op->red.alloc(red_work_size(kMaxRestart + 2)))) return rc;
  if ((rc = op->dsc.alloc(kScSize))) return rc;
  if ((rc = op->wS.alloc(op->N() * op->dim()))) return rc;
  if (!op->host_sc) {
    if (hipHostMalloc(&op->host_sc, kScSize * sizeof(double), hipHostMallocDefault) != hipSuccess)
      return fail(AFL_ERR_DEVICE, "hipHostMalloc failed");
  }
  return AFL_OK;
}

// fetch device scalars [off, off+cnt) to host_sc[off..]
int fetch(afl_operator* op, int off, int cnt) {
  if (hipMemcpyAsync(op->host_sc + off, op->dsc.p + off, cnt * sizeof(double), hipMemcpyDeviceToHost, op->o.stream) !=
      hipSuccess)
    return fail(AFL_ERR_DEVICE, "hipMemcpyAsync failed");
  if (hipStreamSynchronize(op->o.stream) != hipSuccess) return fail(AFL_ERR_DEVICE, "hipStreamSynchronize failed");
  return AFL_OK;
}

// S p = D p - C A^{-1} B p, euler_operator.h:2374-2408 (A^{-1}: exact cell-local solve)
void schur_apply(afl_operator* op, const double* p, double* y) {
  launch_schur_w(op->o, op->rho_fix.p, p, op->wS.p);
  launch_apply_C(op->o, op->p_fix.p, op->wS.p, p, y);
}

int ensure_dinv(afl_operator* op) {
  int rc;
  if ((rc = op->dinv_D.alloc(op->N()))) return rc;
  if (!op->dinv_ready) {
    launch_inv_diag(op->o, nullptr, 1.0 / (kGamma - 1.0), 1, op->dinv_D.p);
    op->dinv_ready = true;
  }
  return AFL_OK;
}

// Right-preconditioned restarted GMRES, Jacobi(D), classical Gram-Schmidt applied twice.
// (The reference's left-preconditioned deal.II GMRES compares the preconditioned residual
// with eps*||rhs||; the Arnoldi estimate here is the true residual ||b - S x||.)
// The operator is a callable apply(in, out); what names the solve in errors.
// Sx0 (optional): apply(x) already computed by the caller, used for the first residual.
template <class Apply>
int gmres_op(afl_operator* op, Apply&& apply, double* x, const double* b, double tol, int max_it, int m, int* iters,
             double* res_out, const char* what, const double* Sx0 = nullptr, bool tol_rel = false) {
  int rc;
  const int64_t N = op->N();
  hipStream_t st = op->o.stream;
  if (m < 1 || m > kMaxRestart) return fail(AFL_ERR_INVALID_ARGUMENT, "gmres_restart must be in 1..60");
  if ((rc = ensure_common(op)) || (rc = ensure_dinv(op))) return rc;
  if ((rc = op->gV.alloc((int64_t)(m + 1) * N)) || (rc = op->gW.alloc(2 * N)) || (rc = op->ydev.alloc(kMaxRestart)))
    return rc;
  double* V = op->gV.p;
  double* w = op->gW.p;
  double* z = op->gW.p + N;
  double* sc = op->dsc.p;
  double* hs = op->host_sc;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  int it = 0;
  double res = 0.0;
  while (true) {
    const double* Ax = w;
    if (Sx0) Ax = Sx0;
    else apply(x, w);
    Sx0 = nullptr;
    vec_sub(st, N, b, Ax, V);
    vec_norm2(st, N, V, op->red.p, sc + kScNrm);
    // tol_rel: tol is relative to the norm the caller left in slot kScA, fetched with this one
    if ((rc = fetch(op, kScNrm, tol_rel ? 2 : 1))) return rc;
    if (tol_rel) {
      tol *= hs[kScA];
      tol_rel = false;
    }
    const double beta = hs[kScNrm];
    res = beta;
    if (!(beta > tol)) break;
    if (it >= max_it) {
      *iters = it; *res_out = res;
      return fail(AFL_ERR_NOT_CONVERGED, what);
    }
    vec_scale_by_inv(st, N, sc + kScNrm, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    bool conv = false;
    for (; j < m && it < max_it; ++j) {
      ++it;
      const double* vj = V + (int64_t)j * N;
      double* vn = V + (int64_t)(j + 1) * N;
      vec_mul(st, N, vj, op->dinv_D.p, z);
      apply(z, vn);
      vec_mdot(st, N, V, N, j + 1, vn, op->red.p, sc + kScH1);
      vec_maxpy(st, N, V, N, j + 1, sc + kScH1, -1.0, vn);
      vec_mdot(st, N, V, N, j + 1, vn, op->red.p, sc + kScH2);
      vec_maxpy(st, N, V, N, j + 1, sc + kScH2, -1.0, vn);
      vec_norm2(st, N, vn, op->red.p, sc + kScNrm);
      vec_scale_by_inv(st, N, sc + kScNrm, vn);
      if ((rc = fetch(op, 0, kScNrm + 1))) return rc;
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hs[kScH1 + i] + hs[kScH2 + i];
      H[(size_t)(j + 1) * m + j] = hs[kScNrm];
      for (int i = 0; i < j; ++i) {
        const double a0 = H[(size_t)i * m + j], a1 = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a0 + sn[i] * a1;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
      }
      const double a0 = H[(size_t)j * m + j], a1 = H[(size_t)(j + 1) * m + j];
      const double r = std::hypot(a0, a1);
      cs[j] = a0 / r;
      sn[j] = a1 / r;
      H[(size_t)j * m + j] = r;
      H[(size_t)(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      res = std::fabs(g[j + 1]);
      if (res <= tol) { ++j; conv = true; break; }
    }
    for (int i = j - 1; i >= 0; --i) {
      double v = g[i];
      for (int l = i + 1; l < j; ++l) v -= H[(size_t)i * m + l] * y[l];
      y[i] = v / H[(size_t)i * m + i];
    }
    if (j > 0) {
      // y goes through pinned memory, which the host rewrites only after a later fetch has waited
      // for the stream, so the copy needs no wait of its own
      std::copy(y.begin(), y.begin() + j, hs + kScY);
      if (hipMemcpyAsync(op->ydev.p, hs + kScY, j * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(AFL_ERR_DEVICE, "hipMemcpyAsync failed");
      vec_fill(st, N, 0.0, z);
      vec_maxpy(st, N, V, N, j, op->ydev.p, 1.0, z);
      vec_addmul(st, N, op->dinv_D.p, z, x);
    }
    if (conv) break;
    if (it >= max_it) {
      *iters = it; *res_out = res;
      return fail(AFL_ERR_NOT_CONVERGED, what);
    }
  }
  *iters = it;
  *res_out = res;
  return check_device("gmres");
}

int gmres(afl_operator* op, double* x, const double* b, double tol, int max_it, int m, int* iters, double* res_out,
          const double* Sx0, bool tol_rel) {
  return gmres_op(
      op, [op](const double* in, double* out) { schur_apply(op, in, out); }, x, b, tol, max_it, m, iters, res_out,
      "Schur GMRES did not converge", Sx0, tol_rel);
}

// p-multigrid hierarchy for the current solver parameters (built once, levels share the mesh)
// (coarsest-level Chebyshev degree unless set: in 3D 5 at k <= 3, where the hill solves at 48x48x20,
// k = 3 take one FGMRES iteration from degree 5 up, and 12 at k = 4 (24x48x64-cell slab of the
// 192x192x64 hill: forward step 330 ms at 5, 285 at 8, 255 at 12, 272 at 16; the store-all adjoint
// 2.53 s at 5, 1.89 s at 12, 1.95 s at 16); in 2D 16 at k = 4 (512x100 hill: 16 fastest of 10..32) and
// 8 at k <= 3 (256x60, k = 3 hill: forward steps within 4 % of the fastest degree, 12, and the
// adjoint stage solves, which reduce their residual by ten orders from the previous step's
// multiplier, cheapest from 6 to 10))
MgConfig mg_config(const afl_solver_params& sp, int dim, int k) {
  MgConfig c;
  if (sp.mg_smoother_degree > 0) c.smoother_degree = sp.mg_smoother_degree;
  c.coarse_degree = sp.mg_coarse_degree > 0 ? sp.mg_coarse_degree : (dim == 2 ? (k >= 4 ? 16 : 8) : (k >= 4 ? 12 : 5));
  if (sp.mg_levels > 0) c.max_levels = sp.mg_levels;
  if (sp.mg_smoothing_range > 1.0) c.smoothing_range = sp.mg_smoothing_range;
  if (sp.mg_cell_tol > 0.0) c.cell_tol = sp.mg_cell_tol;
  // k + 1 points per direction on every level unless the operator's own 2k + 1 are asked for
  c.reduced_quadrature = sp.mg_quad_points != 2 * k + 1;
  return c;
}

int ensure_mg(afl_operator* op) {
  const int want = op->sp.schur_precond;
  int rc;
  if (op->mg_built != want) {
    const MgConfig cfg = mg_config(op->sp, op->o.dim, op->o.k);
    rc = want == AFL_SCHUR_MG_FP32 ? op->mg32.setup_levels(op->o, cfg, op->tabs_by_k, op->tabs_rq_by_k)
                                   : op->mg64.setup_levels(op->o, cfg, op->tabs_by_k, op->tabs_rq_by_k);
    if (rc) return fail(AFL_ERR_DEVICE, "multigrid set-up: hipMalloc failed");
    op->mg_built = want;
    op->mg_eig_stale = true;
    op->mg_coef_stale = true;
  }
  {  // smoothing degrees and range take effect without a rebuild
    const MgConfig cfg = mg_config(op->sp, op->o.dim, op->o.k);
    auto upd = [&](auto& mg) {
      mg.smoother_degree = cfg.smoother_degree;
      mg.coarse_degree = cfg.coarse_degree;
      mg.smoothing_range = cfg.smoothing_range;
    };
    if (want == AFL_SCHUR_MG_FP32) upd(op->mg32);
    else upd(op->mg64);
  }
  if ((rc = op->mgw.alloc(op->N() + red_work_size(1) + 16))) return rc;
  // stage coefficients (rho_for_fixed, pres_fixed) on every level
  if (op->mg_coef_stale) {
    // rho_for_fixed changed whenever the eigenvalue estimate is stale
    if (want == AFL_SCHUR_MG_FP32) op->mg32.set_coefficients(op->rho_fix.p, op->p_fix.p, op->mg_eig_stale);
    else op->mg64.set_coefficients(op->rho_fix.p, op->p_fix.p, op->mg_eig_stale);
    op->mg_coef_stale = false;
  }
  if (op->mg_eig_stale) {
    const int it = mg_config(op->sp, op->o.dim, op->o.k).eig_iters;
    rc = want == AFL_SCHUR_MG_FP32
             ? op->mg32.estimate_eigenvalues(op->o, it, op->mgw.p, op->dsc.p + kScA + 4, op->host_sc + kScA + 4)
             : op->mg64.estimate_eigenvalues(op->o, it, op->mgw.p, op->dsc.p + kScA + 4, op->host_sc + kScA + 4);
    if (rc) return fail(AFL_ERR_DEVICE, "multigrid eigenvalue estimate failed");
    op->mg_eig_stale = false;
  }
  return check_device("multigrid set-up");
}

// Flexible GMRES (Saad 1993) with the p-multigrid V-cycle as variable right preconditioner:
// z_j = MG(v_j) is stored, x += Z y. Outer arithmetic, Arnoldi basis and residual in FP64.
// The operator is a callable apply(in, out); the preconditioner is the multigrid V-cycle of the
// current solver parameters.
// Sx0 (optional): apply(x) already computed by the caller, used for the first residual.
template <class Apply>
int fgmres_op(afl_operator* op, Apply&& apply, double* x, const double* b, double tol, int max_it, int m, int* iters,
              double* res_out, const char* what, const double* Sx0 = nullptr, bool tol_rel = false) {
  int rc;
  const int64_t N = op->N();
  hipStream_t st = op->o.stream;
  if (m < 1 || m > kMaxRestart) return fail(AFL_ERR_INVALID_ARGUMENT, "gmres_restart must be in 1..60");
  if ((rc = ensure_common(op)) || (rc = ensure_mg(op))) return rc;
  if ((rc = op->gV.alloc((int64_t)(m + 1) * N)) || (rc = op->gZ.alloc((int64_t)m * N)) || (rc = op->gW.alloc(2 * N)) ||
      (rc = op->ydev.alloc(kMaxRestart)))
    return rc;
  double* V = op->gV.p;
  double* Z = op->gZ.p;
  double* w = op->gW.p;
  double* sc = op->dsc.p;
  double* hs = op->host_sc;
  const bool f32 = op->sp.schur_precond == AFL_SCHUR_MG_FP32;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  int it = 0;
  double res = 0.0;
  while (true) {
    const double* Ax = w;
    if (Sx0) Ax = Sx0;
    else apply(x, w);
    Sx0 = nullptr;
    vec_sub(st, N, b, Ax, V);
    vec_norm2(st, N, V, op->red.p, sc + kScNrm);
    // tol_rel: tol is relative to the norm the caller left in slot kScA, fetched with this one
    if ((rc = fetch(op, kScNrm, tol_rel ? 2 : 1))) return rc;
    if (tol_rel) {
      tol *= hs[kScA];
      tol_rel = false;
    }
    const double beta = hs[kScNrm];
    res = beta;
    if (!(beta > tol)) break;
    if (it >= max_it) {
      *iters = it; *res_out = res;
      return fail(AFL_ERR_NOT_CONVERGED, what);
    }
    vec_scale_by_inv(st, N, sc + kScNrm, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    bool conv = false;
    for (; j < m && it < max_it; ++j) {
      ++it;
      const double* vj = V + (int64_t)j * N;
      double* zj = Z + (int64_t)j * N;
      double* vn = V + (int64_t)(j + 1) * N;
      if (f32) op->mg32.apply(op->o, vj, zj);
      else op->mg64.apply(op->o, vj, zj);
      apply(zj, vn);
      vec_mdot(st, N, V, N, j + 1, vn, op->red.p, sc + kScH1);
      vec_maxpy(st, N, V, N, j + 1, sc + kScH1, -1.0, vn);
      vec_mdot(st, N, V, N, j + 1, vn, op->red.p, sc + kScH2);
      vec_maxpy(st, N, V, N, j + 1, sc + kScH2, -1.0, vn);
      vec_norm2(st, N, vn, op->red.p, sc + kScNrm);
      vec_scale_by_inv(st, N, sc + kScNrm, vn);
      if ((rc = fetch(op, 0, kScNrm + 1))) return rc;
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hs[kScH1 + i] + hs[kScH2 + i];
      H[(size_t)(j + 1) * m + j] = hs[kScNrm];
      for (int i = 0; i < j; ++i) {
        const double a0 = H[(size_t)i * m + j], a1 = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a0 + sn[i] * a1;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
      }
      const double a0 = H[(size_t)j * m + j], a1 = H[(size_t)(j + 1) * m + j];
      const double r = std::hypot(a0, a1);
      cs[j] = a0 / r;
      sn[j] = a1 / r;
      H[(size_t)j * m + j] = r;
      H[(size_t)(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      res = std::fabs(g[j + 1]);
      if (res <= tol) { ++j; conv = true; break; }
    }
    for (int i = j - 1; i >= 0; --i) {
      double v = g[i];
      for (int l = i + 1; l < j; ++l) v -= H[(size_t)i * m + l] * y[l];
      y[i] = v / H[(size_t)i * m + i];
    }
    if (j > 0) {
      // y through pinned memory, as in gmres_op
      std::copy(y.begin(), y.begin() + j, hs + kScY);
      if (hipMemcpyAsync(op->ydev.p, hs + kScY, j * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(AFL_ERR_DEVICE, "hipMemcpyAsync failed");
      vec_maxpy(st, N, Z, N, j, op->ydev.p, 1.0, x);
    }
    if (conv) break;
    if (it >= max_it) {
      *iters = it; *res_out = res;
      return fail(AFL_ERR_NOT_CONVERGED, what);
    }
  }
  *iters = it;
  *res_out = res;
  return check_device("fgmres");
}

int fgmres(afl_operator* op, double* x, const double* b, double tol, int max_it, int m, int* iters, double* res_out,
           const double* Sx0, bool tol_rel) {
  return fgmres_op(
      op, [op](const double* in, double* out) { schur_apply(op, in, out); }, x, b, tol, max_it, m, iters, res_out,
      "Schur FGMRES did not converge", Sx0, tol_rel);
}

bool valid_op(afl_operator* op) { return op != nullptr; }

}  // namespace

// =========================================================================================
extern "C" {

const char* afl_last_error(void) { return g_err.c_str(); }

int afl_create(const afl_mesh_desc* md, double mach, double froude, double dt, afl_operator** out) {
  if (!out) return fail(AFL_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!md) return fail(AFL_ERR_INVALID_ARGUMENT, "mesh is NULL");
  if (md->dim != 2 && md->dim != 3) return fail(AFL_ERR_INVALID_ARGUMENT, "dim must be 2 or 3");
  if (md->degree < 1 || md->degree > 4) return fail(AFL_ERR_INVALID_ARGUMENT, "degree must be 1..4");
  if (md->node_type != AFL_NODES_GAUSS_LOBATTO && md->node_type != AFL_NODES_EQUIDISTANT)
    return fail(AFL_ERR_INVALID_ARGUMENT, "unknown node_type");
  const int nx = md->n_cells[0], ny = md->dim == 3 ? md->n_cells[1] : 1, nz = md->n_cells[2];
  if (nx < 1 || ny < 1 || nz < 1) return fail(AFL_ERR_INVALID_ARGUMENT, "cell counts must be positive");
  if (!(md->x_max > md->x_min) || (md->dim == 3 && !(md->y_max > md->y_min)) || !(md->z_top > 0.0))
    return fail(AFL_ERR_INVALID_ARGUMENT, "empty domain");
  if (!(md->hill_width > 0.0) || !(md->hill_height < md->z_top) || md->hill_height < 0.0)
    return fail(AFL_ERR_INVALID_ARGUMENT, "hill must satisfy 0 <= h0 < z_top and a > 0");
  if (!(dt > 0.0) || !(mach > 0.0) || !(froude > 0.0)) return fail(AFL_ERR_INVALID_ARGUMENT, "dt, Ma, Fr must be > 0");
  const int64_t ncells = (int64_t)nx * ny * nz;
  if (ncells > 2000000000LL) return fail(AFL_ERR_INVALID_ARGUMENT, "too many cells");
  auto* op = new afl_operator;
  OpCtx& o = op->o;
  o.dim = md->dim;
  o.k = md->degree;
  build_tab(o.k, md->node_type, o.T);
  for (int kk = 1; kk <= 4; ++kk) {
    build_tab(kk, md->node_type, op->tabs_by_k[kk]);
    build_tab(kk, md->node_type, op->tabs_rq_by_k[kk], kk + 1);
  }
  if (hipMalloc((void**)&op->dtab, sizeof(Tab)) != hipSuccess ||
      hipMemcpy(op->dtab, &o.T, sizeof(Tab), hipMemcpyHostToDevice) != hipSuccess) {
    delete op;
    return fail(AFL_ERR_DEVICE, "afl_create: table upload failed");
  }
  o.dT = op->dtab;
  Geo& g = o.geo;
  g.nc[0] = nx; g.nc[1] = md->dim == 3 ? ny : nz; g.nc[2] = md->dim == 3 ? nz : 1;
  g.ncells = (int)ncells;
  g.per[0] = md->periodic[0] ? 1 : 0;
  g.per[1] = (md->dim == 3 && md->periodic[1]) ? 1 : 0;
  g.lo[0] = md->x_min; g.lo[1] = md->y_min;
  g.hc[0] = (md->x_max - md->x_min) / nx;
  if (md->dim == 3) { g.hc[1] = (md->y_max - md->y_min) / ny; g.hc[2] = md->z_top / nz; }
  else { g.hc[1] = md->z_top / nz; g.hc[2] = 0.0; }
  g.H = md->z_top; g.h0 = md->hill_height; g.inv_a2 = 1.0 / (md->hill_width * md->hill_width);
  const int n1 = o.k + 1;
  o.N = ncells * (md->dim == 2 ? n1 * n1 : n1 * n1 * n1);
  o.Ma = mach; o.Fr = froude; o.dt = dt;
  o.tab = make_tableau();
  int rc;
  if ((rc = op->rho_fix.alloc(o.N)) || (rc = op->p_fix.alloc(o.N)) || (rc = op->u_fix.alloc(o.N * o.dim)) ||
      (rc = ensure_common(op))) {
    delete op;
    return rc;
  }
  double* scr = nullptr;  // the device's copy of the no-alloc scratch, resolved at set-up
  if ((rc = current_device_scratch(&scr))) {
    delete op;
    return rc;
  }
  op->min_diam = min_cell_diameter(md);
  vec_fill(o.stream, o.N, 1.0, op->rho_fix.p);
  vec_fill(o.stream, o.N, 0.0, op->p_fix.p);
  vec_fill(o.stream, o.N * o.dim, 0.0, op->u_fix.p);
  if ((rc = check_device("afl_create"))) { delete op; return rc; }
  *out = op;
  return AFL_OK;
}

int afl_destroy(afl_operator* op) {
  if (op) {
    hipStreamSynchronize(op->o.stream);
    delete op;
  }
  return AFL_OK;
}

int afl_set_stream(afl_operator* op, void* s) {
  if (!valid_op(op)) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  op->o.stream = (hipStream_t)s;
  // the multigrid levels launch on their own copy of the stream: every level follows
  for (auto& L : op->mg32.levels) L.stream = op->o.stream;
  for (auto& L : op->mg64.levels) L.stream = op->o.stream;
  return AFL_OK;
}

int afl_n_dofs(const afl_operator* op, int idx, int64_t* n) {
  if (!op || !n) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (idx < 0 || idx > 2) return fail(AFL_ERR_INVALID_ARGUMENT, "dof_index must be 0, 1 or 2");
  *n = idx == 0 ? op->o.N * op->o.dim : op->o.N;
  return AFL_OK;
}

int afl_m(const afl_operator* op, int64_t* m) {
  if (!op || !m) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  *m = (op->ns == 3 || op->ns == 5) ? op->o.N * op->o.dim : op->o.N;
  return AFL_OK;
}

// dt, Ma and Fr scale the Schur operator, so the multigrid's Chebyshev interval (estimated for the
// old values) is re-estimated before the next preconditioner apply
int afl_set_dt(afl_operator* op, double v) {
  if (!op || !(v > 0.0)) return fail(AFL_ERR_INVALID_ARGUMENT, "dt must be > 0");
  if (v != op->o.dt) op->mg_eig_stale = true;
  op->o.dt = v;
  return AFL_OK;
}
int afl_set_mach(afl_operator* op, double v) {
  if (!op || !(v > 0.0)) return fail(AFL_ERR_INVALID_ARGUMENT, "Ma must be > 0");
  if (v != op->o.Ma) op->mg_eig_stale = true;
  op->o.Ma = v;
  return AFL_OK;
}
int afl_set_froude(afl_operator* op, double v) {
  if (!op || !(v > 0.0)) return fail(AFL_ERR_INVALID_ARGUMENT, "Fr must be > 0");
  if (v != op->o.Fr) op->mg_eig_stale = true;
  op->o.Fr = v;
  return AFL_OK;
}
int afl_set_hyperbolic_stage(afl_operator* op, unsigned s) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  if (s < 1 || s > 3) return fail(AFL_ERR_INVALID_ARGUMENT, "HYPERBOLIC_stage must be 1..3");
  op->o.hyp = s;
  return AFL_OK;
}
int afl_set_ns_stage(afl_operator* op, unsigned s) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  if (s < 1 || s > 6) return fail(AFL_ERR_INVALID_ARGUMENT, "NS_stage must be 1..6");
  op->ns = s;
  return AFL_OK;
}
int afl_set_rho_for_fixed(afl_operator* op, const double* v) {
  if (!op || !v) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  vec_copy(op->o.stream, op->o.N, v, op->rho_fix.p);
  op->mg_coef_stale = true;
  op->mg_eig_stale = true;
  return check_device("afl_set_rho_for_fixed");
}
int afl_set_pres_fixed(afl_operator* op, const double* v) {
  if (!op || !v) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  vec_copy(op->o.stream, op->o.N, v, op->p_fix.p);
  op->mg_coef_stale = true;
  return check_device("afl_set_pres_fixed");
}
int afl_set_u_fixed(afl_operator* op, const double* v) {
  if (!op || !v) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  vec_copy(op->o.stream, op->o.N * op->o.dim, v, op->u_fix.p);
  return check_device("afl_set_u_fixed");
}
int afl_set_solver_params(afl_operator* op, const afl_solver_params* sp) {
  if (!op || !sp) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (sp->gmres_restart < 1 || sp->gmres_restart > kMaxRestart) return fail(AFL_ERR_INVALID_ARGUMENT, "gmres_restart must be 1..60");
  if (!(sp->gmres_eps > 0.0) || sp->gmres_max_iter < 1 || sp->fp_max_iter < 1 || !(sp->fp_tol > 0.0))
    return fail(AFL_ERR_INVALID_ARGUMENT, "invalid solver parameters");
  if (sp->schur_precond < AFL_SCHUR_JACOBI || sp->schur_precond > AFL_SCHUR_MG_FP64)
    return fail(AFL_ERR_INVALID_ARGUMENT, "unknown schur_precond");
  if (sp->mg_smoother_degree > 12 || sp->mg_coarse_degree > 64 || sp->mg_levels > 4)
    return fail(AFL_ERR_INVALID_ARGUMENT, "multigrid parameters out of range");
  if (sp->mg_quad_points != 0 && sp->mg_quad_points != op->o.k + 1 && sp->mg_quad_points != 2 * op->o.k + 1)
    return fail(AFL_ERR_INVALID_ARGUMENT, "mg_quad_points must be 0, k + 1 or 2k + 1");
  const bool rebuild = sp->schur_precond != op->sp.schur_precond || sp->mg_levels != op->sp.mg_levels ||
                       sp->mg_cell_tol != op->sp.mg_cell_tol || sp->mg_quad_points != op->sp.mg_quad_points;
  op->sp = *sp;
  if (rebuild) op->mg_built = 0;
  return AFL_OK;
}
int afl_set_sponge(afl_operator* op, const afl_sponge* s) {
  if (!op || !s) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  op->sponge = *s;
  return AFL_OK;
}

static int check_src(const double* const* src, int n_src, int need, const char* what) {
  if (!src) return fail(AFL_ERR_INVALID_ARGUMENT, std::string(what) + ": src is NULL");
  if (n_src != need)
    return fail(AFL_ERR_INVALID_ARGUMENT, std::string(what) + ": expected " + std::to_string(need) +
                                              " source vectors for this HYPERBOLIC_stage, got " + std::to_string(n_src));
  for (int i = 0; i < n_src; ++i)
    if (!src[i]) return fail(AFL_ERR_INVALID_ARGUMENT, std::string(what) + ": NULL source vector");
  return AFL_OK;
}

int afl_vmult_rhs_rho_update(afl_operator* op, double* dst, const double* const* src, int n_src) {
  if (!op || !dst) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int st = (int)op->o.hyp;
  int rc;
  if ((rc = check_src(src, n_src, 2 * st, "vmult_rhs_rho_update"))) return rc;
  launch_rhs_rho(op->o, st, src, dst);
  return check_device("vmult_rhs_rho_update");
}

int afl_vmult_rhs_velocity_fixed(afl_operator* op, double* dst, const double* const* src, int n_src) {
  if (!op || !dst) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int st = (int)op->o.hyp;
  int rc;
  if ((rc = check_src(src, n_src, st == 1 ? 4 : (st == 2 ? 7 : 9), "vmult_rhs_velocity_fixed"))) return rc;
  launch_rhs_u(op->o, st, src, dst);
  return check_device("vmult_rhs_velocity_fixed");
}

int afl_vmult_rhs_pressure(afl_operator* op, double* dst, const double* const* src, int n_src) {
  if (!op || !dst) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int st = (int)op->o.hyp;
  int rc;
  if ((rc = check_src(src, n_src, st == 1 ? 5 : (st == 2 ? 8 : 11), "vmult_rhs_pressure"))) return rc;
  launch_rhs_p(op->o, st, src, dst);
  return check_device("vmult_rhs_pressure");
}

int afl_vmult(afl_operator* op, double* dst, const double* src) {
  if (!op || !dst || !src) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  switch (op->ns) {
    case 1: case 4: launch_mass_apply(op->o, nullptr, 1.0, 1, src, dst); break;
    case 2: {
      int rc;
      if ((rc = ensure_common(op))) return rc;
      schur_apply(op, src, dst);
      break;
    }
    case 3: case 5: launch_mass_apply(op->o, op->rho_fix.p, 1.0, op->o.dim, src, dst); break;
    default: launch_mass_apply(op->o, nullptr, 1.0 / (kGamma - 1.0), 1, src, dst); break;
  }
  return check_device("vmult");
}

int afl_vmult_pressure(afl_operator* op, double* dst, const double* src) {
  if (!op || !dst || !src) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_apply_B(op->o, src, dst);
  return check_device("vmult_pressure");
}

int afl_vmult_enthalpy(afl_operator* op, double* dst, const double* src) {
  if (!op || !dst || !src) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_apply_C(op->o, op->p_fix.p, src, nullptr, dst);
  return check_device("vmult_enthalpy");
}

int afl_vmult_schur_velocity(afl_operator* op, double* w, const double* p) {
  if (!op || !w || !p) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_schur_w(op->o, op->rho_fix.p, p, w);
  return check_device("vmult_schur_velocity");
}

// ---- discrete-adjoint building blocks ----------------------------------------------------

int afl_vjp_rhs_rho_update(afl_operator* op, const double* rbar, double* const* src_bar, const double* const* src,
                           int n_src) {
  if (!op || !rbar || !src_bar) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int st = (int)op->o.hyp;
  int rc;
  if ((rc = check_src(src, n_src, 2 * st, "vjp_rhs_rho_update"))) return rc;
  for (int i = 0; i < n_src; ++i)
    if (!src_bar[i]) return fail(AFL_ERR_INVALID_ARGUMENT, "vjp_rhs_rho_update: null output vector");
  launch_vjp_rhs_rho(op->o, st, rbar, src, src_bar);
  return check_device("vjp_rhs_rho_update");
}

int afl_vjp_rhs_velocity_fixed(afl_operator* op, const double* rbar, double* const* src_bar, const double* const* src,
                               int n_src) {
  if (!op || !rbar || !src_bar) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int st = (int)op->o.hyp;
  int rc;
  if ((rc = check_src(src, n_src, st == 1 ? 4 : (st == 2 ? 7 : 9), "vjp_rhs_velocity_fixed"))) return rc;
  for (int i = 0; i < n_src; ++i)
    if (!src_bar[i]) return fail(AFL_ERR_INVALID_ARGUMENT, "vjp_rhs_velocity_fixed: null output vector");
  launch_vjp_rhs_u(op->o, st, rbar, src, src_bar);
  return check_device("vjp_rhs_velocity_fixed");
}

int afl_vjp_rhs_pressure(afl_operator* op, const double* rbar, double* const* src_bar, const double* const* src,
                         int n_src) {
  if (!op || !rbar || !src_bar) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int st = (int)op->o.hyp;
  int rc;
  if ((rc = check_src(src, n_src, st == 1 ? 5 : (st == 2 ? 8 : 11), "vjp_rhs_pressure"))) return rc;
  for (int i = 0; i < n_src; ++i)
    if (!src_bar[i]) return fail(AFL_ERR_INVALID_ARGUMENT, "vjp_rhs_pressure: null output vector");
  launch_vjp_rhs_p(op->o, st, rbar, src, src_bar);
  return check_device("vjp_rhs_pressure");
}

int afl_vmult_pressure_transpose(afl_operator* op, double* dst, const double* src) {
  if (!op || !dst || !src) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_apply_Bt(op->o, src, dst);
  return check_device("vmult_pressure_transpose");
}

int afl_vmult_enthalpy_tangent(afl_operator* op, double* dst, const double* dw, const double* w0) {
  if (!op || !dst || !dw || !w0) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_apply_C_tangent(op->o, op->p_fix.p, dw, w0, dst);
  return check_device("vmult_enthalpy_tangent");
}

int afl_vmult_enthalpy_tangent_transpose(afl_operator* op, double* dst, const double* cbar, const double* w0) {
  if (!op || !dst || !cbar || !w0) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_apply_C_tangent_t(op->o, op->p_fix.p, cbar, w0, dst);
  return check_device("vmult_enthalpy_tangent_transpose");
}

int afl_vmult_enthalpy_coefficient_transpose(afl_operator* op, double* dst, const double* cbar, const double* w) {
  if (!op || !dst || !cbar || !w) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_apply_C_coef_t(op->o, w, cbar, dst);
  return check_device("vmult_enthalpy_coefficient_transpose");
}

int afl_vjp_rho_for_fixed(afl_operator* op, double* dst, const double* ybar, const double* u) {
  if (!op || !dst || !ybar || !u) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  launch_dot_mass(op->o, ybar, u, dst);
  return check_device("vjp_rho_for_fixed");
}

int afl_vjp_rhs_pressure_u_fixed(afl_operator* op, double* dst, const double* rbar, const double* rho_new,
                                 const double* u_fixed) {
  if (!op || !dst || !rbar || !rho_new || !u_fixed) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const OpCtx& o = op->o;
  if (o.hyp == 3) return fail(AFL_ERR_STATE, "vjp_rhs_pressure_u_fixed: stages 1 and 2 read u_fixed");
  const double Ma2 = o.Ma * o.Ma, iFr2 = 1.0 / (o.Fr * o.Fr);
  const double cgz = (o.hyp == 1 ? o.tab.at22 : o.tab.at33) * o.dt * Ma2 * iFr2;
  launch_scaled_velocity_mass(o, rbar, rho_new, u_fixed, -Ma2, -cgz, dst);
  return check_device("vjp_rhs_pressure_u_fixed");
}

namespace {
// S'(p0) x = D x - C'(w0) A^{-1} B x, or its transpose D x - B^T A^{-1} C'(w0)^T x, w0 = A^{-1} B p0
int schur_tangent(afl_operator* op, double* dst, const double* src, const double* p0, bool transpose) {
  if (!op || !dst || !src || !p0) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int64_t N = op->N(), NU = N * op->dim();
  int rc;
  if ((rc = op->adjW0.alloc(NU)) || (rc = op->adjT1.alloc(NU)) || (rc = op->adjT2.alloc(NU)) ||
      (rc = op->adjC.alloc(N)))
    return rc;
  const OpCtx& o = op->o;
  launch_schur_w(o, op->rho_fix.p, p0, op->adjW0.p);
  if (!transpose) {
    launch_schur_w(o, op->rho_fix.p, src, op->adjT2.p);
    launch_apply_C_tangent(o, op->p_fix.p, op->adjT2.p, op->adjW0.p, op->adjC.p);
  } else {
    launch_apply_C_tangent_t(o, op->p_fix.p, src, op->adjW0.p, op->adjT1.p);
    launch_mass_solve(o, op->rho_fix.p, 1.0, op->dim(), op->adjT1.p, op->adjT2.p);
    launch_apply_Bt(o, op->adjT2.p, op->adjC.p);
  }
  launch_mass_apply(o, nullptr, 1.0 / (kGamma - 1.0), 1, src, dst);
  vec_axpby(o.stream, N, -1.0, op->adjC.p, 1.0, dst);
  return AFL_OK;
}
}  // namespace

int afl_vmult_schur_tangent(afl_operator* op, double* dst, const double* src, const double* p0) {
  int rc;
  if ((rc = schur_tangent(op, dst, src, p0, false))) return rc;
  return check_device("vmult_schur_tangent");
}

int afl_vmult_schur_tangent_transpose(afl_operator* op, double* dst, const double* src, const double* p0) {
  int rc;
  if ((rc = schur_tangent(op, dst, src, p0, true))) return rc;
  return check_device("vmult_schur_tangent_transpose");
}

int afl_mg_vmult(afl_operator* op, double* dst, const double* src) {
  if (!op || !dst || !src) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int pc = op->sp.schur_precond;
  if (pc != AFL_SCHUR_MG_FP32 && pc != AFL_SCHUR_MG_FP64)
    return fail(AFL_ERR_STATE, "mg_vmult needs solver_params.schur_precond = AFL_SCHUR_MG_FP32 or AFL_SCHUR_MG_FP64");
  int rc;
  if ((rc = ensure_common(op)) || (rc = ensure_mg(op))) return rc;
  if (pc == AFL_SCHUR_MG_FP32) op->mg32.apply(op->o, src, dst);
  else op->mg64.apply(op->o, src, dst);
  return check_device("mg_vmult");
}

int afl_mg_level_vmult(afl_operator* op, double* dst, const double* src) {
  if (!op || !dst || !src) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int pc = op->sp.schur_precond;
  if (pc != AFL_SCHUR_MG_FP32 && pc != AFL_SCHUR_MG_FP64)
    return fail(AFL_ERR_STATE, "mg_level_vmult needs solver_params.schur_precond = AFL_SCHUR_MG_FP32 or AFL_SCHUR_MG_FP64");
  int rc;
  if ((rc = ensure_common(op)) || (rc = ensure_mg(op))) return rc;
  if (pc == AFL_SCHUR_MG_FP32) op->mg32.level_vmult(op->o, src, dst);
  else op->mg64.level_vmult(op->o, src, dst);
  return check_device("mg_level_vmult");
}

int afl_vmult_schur_velocity_f32(afl_operator* op, float* w, const float* p) {
  if (!op || !w || !p) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (op->sp.schur_precond != AFL_SCHUR_MG_FP32)
    return fail(AFL_ERR_STATE, "vmult_schur_velocity_f32 needs solver_params.schur_precond = AFL_SCHUR_MG_FP32");
  int rc;
  if ((rc = ensure_common(op)) || (rc = ensure_mg(op))) return rc;
  op->mg32.schur_velocity(op->o, p, w);
  return check_device("vmult_schur_velocity_f32");
}

int afl_set_verbose(afl_operator* op, int level) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  op->verbose = level;
  return AFL_OK;
}

int afl_section_timer_start(afl_operator* op) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  op->sect.nseg = 0;
  op->sect.on = true;
  return AFL_OK;
}

int afl_section_timer_stop(afl_operator* op) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  op->sect.on = false;
  return AFL_OK;
}

int afl_section_timer_result(afl_operator* op, int section, int* calls, double* total_ms) {
  if (!op || !calls || !total_ms) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (section < 0 || section >= AFL_N_SECTIONS) return fail(AFL_ERR_INVALID_ARGUMENT, "unknown section");
  if (hipStreamSynchronize(op->o.stream) != hipSuccess) return fail(AFL_ERR_DEVICE, "hipStreamSynchronize failed");
  int n = 0;
  double ms = 0.0;
  for (size_t s = 0; s < op->sect.nseg; ++s) {
    if (op->sect.sec[s] != section) continue;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, op->sect.ev[2 * s], op->sect.ev[2 * s + 1]) != hipSuccess)
      return fail(AFL_ERR_DEVICE, "hipEventElapsedTime failed");
    ++n;
    ms += t;
  }
  *calls = n;
  *total_ms = ms;
  return AFL_OK;
}

int afl_kernel_timer_start(afl_operator* op, int kernels, int max_launches) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  constexpr int kAll = AFL_TIME_LEVEL_W | AFL_TIME_LEVEL_C_CHEB | AFL_TIME_RHS_RHO | AFL_TIME_RHS_U | AFL_TIME_RHS_P |
                       AFL_TIME_RHS_P_UFIX | AFL_TIME_SCHUR_W | AFL_TIME_APPLY_C | AFL_TIME_MASS_SOLVE |
                       AFL_TIME_APPLY_B | AFL_TIME_A_SOLVE;
  if (kernels <= 0 || (kernels & ~kAll)) return fail(AFL_ERR_INVALID_ARGUMENT, "unknown timed kernel");
  if (max_launches < 1 || max_launches > (1 << 20)) return fail(AFL_ERR_INVALID_ARGUMENT, "max_launches out of range");
  KernelTimer& t = op->timer;
  while ((int)t.ev.size() < 2 * max_launches) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return fail(AFL_ERR_DEVICE, "hipEventCreate failed");
    t.ev.push_back(e);
  }
  t.flag.assign(t.ev.size() / 2, 0);
  t.stage.assign(t.ev.size() / 2, 0);
  t.bytes.assign(t.ev.size() / 2, 0.0);
  t.mask = (unsigned)kernels;
  t.used = 0;
  t.dropped = 0;
  op->o.timer = &t;
  op->mg32.timer = &t;
  op->mg64.timer = &t;
  return AFL_OK;
}
int afl_kernel_timer_stop(afl_operator* op, int* launches, double* total_ms) {
  if (!op || !launches || !total_ms) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  KernelTimer& t = op->timer;
  t.mask = 0;
  op->o.timer = nullptr;
  if (hipStreamSynchronize(op->o.stream) != hipSuccess) return fail(AFL_ERR_DEVICE, "hipStreamSynchronize failed");
  double sum = 0.0;
  op->timer_ms.assign(t.used, 0.0);
  for (int i = 0; i < t.used; ++i) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]) != hipSuccess)
      return fail(AFL_ERR_DEVICE, "hipEventElapsedTime failed");
    op->timer_ms[i] = ms;
    sum += ms;
  }
  *launches = t.used;
  *total_ms = sum;
  if (t.dropped) return fail(AFL_ERR_STATE, "kernel timer: more launches than max_launches");
  return AFL_OK;
}
int afl_kernel_timer_result(afl_operator* op, int kernel, int stage, int* launches, double* total_ms,
                            double* bytes) {
  if (!op || !launches || !total_ms || !bytes) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (stage < 0 || stage > 3) return fail(AFL_ERR_INVALID_ARGUMENT, "stage must be 0..3");
  const KernelTimer& t = op->timer;
  if (t.mask) return fail(AFL_ERR_STATE, "kernel timer still running: call afl_kernel_timer_stop first");
  int n = 0;
  double ms = 0.0, b = 0.0;
  for (int i = 0; i < t.used && i < (int)op->timer_ms.size(); ++i)
    if (t.flag[i] == kernel && (stage == 0 || t.stage[i] == stage)) {
      ++n;
      ms += op->timer_ms[i];
      b += t.bytes[i];
    }
  *launches = n;
  *total_ms = ms;
  *bytes = b;
  return AFL_OK;
}

// ---- VTU output (EulerSolver::output_results, Atmospheric_Flow.cc:886-918) ------------------
extern "C++" {
namespace {
void b64_append(std::string& out, const unsigned char* d, size_t n) {
  static const char* tb = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    const unsigned v = (d[i] << 16) | (d[i + 1] << 8) | d[i + 2];
    out += tb[(v >> 18) & 63]; out += tb[(v >> 12) & 63]; out += tb[(v >> 6) & 63]; out += tb[v & 63];
  }
  if (i < n) {
    const unsigned v = (d[i] << 16) | ((i + 1 < n ? d[i + 1] : 0) << 8);
    out += tb[(v >> 18) & 63]; out += tb[(v >> 12) & 63];
    out += i + 1 < n ? tb[(v >> 6) & 63] : '=';
    out += '=';
  }
}
// one binary DataArray payload: base64 of (UInt64 byte count, raw little-endian data)
template <typename T>
std::string vtu_binary(const std::vector<T>& v) {
  std::vector<unsigned char> buf(8 + v.size() * sizeof(T));
  const uint64_t nb = v.size() * sizeof(T);
  std::memcpy(buf.data(), &nb, 8);
  if (nb) std::memcpy(buf.data() + 8, v.data(), nb);
  std::string o;
  o.reserve(buf.size() * 4 / 3 + 8);
  b64_append(o, buf.data(), buf.size());
  return o;
}
}  // namespace
}  // extern "C++"

int afl_write_vtu(afl_operator* op, const char* path, const double* rho, const double* u, const double* p,
                  int high_order) {
  if (!op || !path || !rho || !u || !p) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const OpCtx& o = op->o;
  const int dim = o.dim, k = o.k, n1 = k + 1;
  const int64_t N = o.N, ncells = o.geo.ncells;
  const int npc = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  std::vector<double> hr(N), hu(N * dim), hp(N);
  if (hipStreamSynchronize(o.stream) != hipSuccess || hipMemcpy(hr.data(), rho, N * 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hu.data(), u, N * dim * 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hp.data(), p, N * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(AFL_ERR_DEVICE, "write_vtu: device to host copy failed");
  // patch points: k+1 equidistant points per direction of every cell (DataOut::build_patches
  // with n_subdivisions = degree); E[i][j] = phi_j(i/k) on the support points
  double E[5][5];
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) E[i][j] = lag(o.T.ND, n1, j, (double)i / k);
  const Geo& g = o.geo;
  const int nx = g.nc[0], ny = dim == 3 ? g.nc[1] : 1;
  auto zmap = [&](double x, double y, double zeta) {
    const double h = g.h0 / (1.0 + (x * x + y * y) * g.inv_a2);
    return zeta + h * (1.0 - zeta / g.H);
  };
  const int64_t npts = ncells * npc;
  std::vector<double> X(npts * 3), R(npts), U(npts * 3, 0.0), P(npts);
  std::vector<double> tmp(npc);
  auto interp = [&](const double* f, double* out) {  // nodal -> patch values of one cell
    for (int q = 0; q < npc; ++q) {
      const int qx = q % n1, qy = (q / n1) % n1, qz = dim == 3 ? q / (n1 * n1) : 0;
      double a = 0.0;
      for (int i = 0; i < npc; ++i) {
        const int ix = i % n1, iy = (i / n1) % n1, iz = dim == 3 ? i / (n1 * n1) : 0;
        double w = E[qx][ix] * E[qy][iy];
        if (dim == 3) w *= E[qz][iz];
        a += w * f[i];
      }
      out[q] = a;
    }
  };
  for (int64_t c = 0; c < ncells; ++c) {
    const int ix = (int)(c % nx), iy = dim == 3 ? (int)((c / nx) % ny) : 0;
    const int iz = dim == 3 ? (int)(c / ((int64_t)nx * ny)) : (int)(c / nx);
    const int64_t base = c * npc;
    interp(&hr[base], &R[base]);
    interp(&hp[base], &P[base]);
    for (int a = 0; a < dim; ++a) {
      interp(&hu[a * N + base], tmp.data());
      for (int q = 0; q < npc; ++q) U[(base + q) * 3 + a] = tmp[q];
    }
    for (int q = 0; q < npc; ++q) {
      const int qx = q % n1, qy = (q / n1) % n1, qz = dim == 3 ? q / (n1 * n1) : qy;
      const double x = g.lo[0] + (ix + (double)qx / k) * g.hc[0];
      const double y = dim == 3 ? g.lo[1] + (iy + (double)qy / k) * g.hc[1] : 0.0;
      const double zeta = (iz + (double)qz / k) * g.hc[dim - 1];
      double* xp = &X[(base + q) * 3];
      xp[0] = x;
      if (dim == 2) { xp[1] = zmap(x, 0.0, zeta); xp[2] = 0.0; }
      else { xp[1] = y; xp[2] = zmap(x, y, zeta); }
    }
  }
  std::vector<int64_t> conn, offs;
  std::vector<uint8_t> types;
  auto pid = [&](int64_t c, int a, int b, int cc) { return c * npc + (int64_t)(cc * n1 + b) * n1 + a; };
  for (int64_t c = 0; c < ncells; ++c) {
    if (high_order && dim == 3) {
      // VTK_LAGRANGE_HEXAHEDRON (file version 2.2 ordering): corners (0,0,0) (k,0,0) (k,k,0) (0,k,0), the
      // same four at z = k; the bottom edges y = 0, x = k, y = k, x = 0 (each in increasing x or y), the
      // same four at z = k, the vertical edges at (x, y) = (0,0) (k,0) (k,k) (0,k); the faces x = 0, x = k
      // (y fastest), y = 0, y = k (x fastest), z = 0, z = k (x fastest); the interior, x fastest
      const int c0[8][3] = {{0, 0, 0}, {k, 0, 0}, {k, k, 0}, {0, k, 0}, {0, 0, k}, {k, 0, k}, {k, k, k}, {0, k, k}};
      for (const auto& v : c0) conn.push_back(pid(c, v[0], v[1], v[2]));
      for (int zc : {0, k}) {
        for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, 0, zc));
        for (int j = 1; j < k; ++j) conn.push_back(pid(c, k, j, zc));
        for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, k, zc));
        for (int j = 1; j < k; ++j) conn.push_back(pid(c, 0, j, zc));
      }
      const int vx[4][2] = {{0, 0}, {k, 0}, {k, k}, {0, k}};
      for (const auto& v : vx)
        for (int l = 1; l < k; ++l) conn.push_back(pid(c, v[0], v[1], l));
      for (int xf : {0, k})
        for (int l = 1; l < k; ++l)
          for (int j = 1; j < k; ++j) conn.push_back(pid(c, xf, j, l));
      for (int yf : {0, k})
        for (int l = 1; l < k; ++l)
          for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, yf, l));
      for (int zf : {0, k})
        for (int j = 1; j < k; ++j)
          for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, j, zf));
      for (int l = 1; l < k; ++l)
        for (int j = 1; j < k; ++j)
          for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, j, l));
      offs.push_back((int64_t)conn.size());
      types.push_back(72);
      continue;
    }
    if (high_order) {
      // VTK_LAGRANGE_QUADRILATERAL: corners, then the edges (0,1), (1,2), (3,2), (0,3), then the interior
      conn.push_back(pid(c, 0, 0, 0)); conn.push_back(pid(c, k, 0, 0));
      conn.push_back(pid(c, k, k, 0)); conn.push_back(pid(c, 0, k, 0));
      for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, 0, 0));
      for (int j = 1; j < k; ++j) conn.push_back(pid(c, k, j, 0));
      for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, k, 0));
      for (int j = 1; j < k; ++j) conn.push_back(pid(c, 0, j, 0));
      for (int j = 1; j < k; ++j)
        for (int i = 1; i < k; ++i) conn.push_back(pid(c, i, j, 0));
      offs.push_back((int64_t)conn.size());
      types.push_back(70);
      continue;
    }
    for (int zc = 0; zc < (dim == 3 ? k : 1); ++zc)
      for (int b = 0; b < k; ++b)
        for (int a = 0; a < k; ++a) {
          if (dim == 2) {
            conn.push_back(pid(c, a, b, 0)); conn.push_back(pid(c, a + 1, b, 0));
            conn.push_back(pid(c, a + 1, b + 1, 0)); conn.push_back(pid(c, a, b + 1, 0));
            types.push_back(9);
          } else {
            for (int lz = 0; lz < 2; ++lz) {
              conn.push_back(pid(c, a, b, zc + lz)); conn.push_back(pid(c, a + 1, b, zc + lz));
              conn.push_back(pid(c, a + 1, b + 1, zc + lz)); conn.push_back(pid(c, a, b + 1, zc + lz));
            }
            types.push_back(12);
          }
          offs.push_back((int64_t)conn.size());
        }
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(AFL_ERR_INVALID_ARGUMENT, std::string("write_vtu: cannot open ") + path);
  // version 2.2: the Lagrange hexahedron's vertical edges in corner order (VTK >= 9.1)
  std::string hdr = "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"2.2\" "
                    "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n<UnstructuredGrid>\n<Piece NumberOfPoints=\"" +
                    std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(types.size()) + "\">\n";
  auto arr = [&](const char* type, const char* name, int ncomp, const std::string& data) {
    std::string a = std::string("<DataArray type=\"") + type + "\"";
    if (name) a += std::string(" Name=\"") + name + "\"";
    if (ncomp > 1) a += " NumberOfComponents=\"" + std::to_string(ncomp) + "\"";
    a += " format=\"binary\">\n" + data + "\n</DataArray>\n";
    return a;
  };
  std::string body = hdr + "<PointData Scalars=\"rho\" Vectors=\"u\">\n" + arr("Float64", "rho", 1, vtu_binary(R)) +
                     arr("Float64", "u", 3, vtu_binary(U)) + arr("Float64", "p", 1, vtu_binary(P)) +
                     "</PointData>\n<Points>\n" + arr("Float64", "Points", 3, vtu_binary(X)) +
                     "</Points>\n<Cells>\n" + arr("Int64", "connectivity", 1, vtu_binary(conn)) +
                     arr("Int64", "offsets", 1, vtu_binary(offs)) + arr("UInt8", "types", 1, vtu_binary(types)) +
                     "</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
  const size_t wr = std::fwrite(body.data(), 1, body.size(), f);
  std::fclose(f);
  if (wr != body.size()) return fail(AFL_ERR_DEVICE, "write_vtu: short write");
  return AFL_OK;
}

int afl_compute_diagnostics(afl_operator* op, const double* rho, const double* u, const double* p,
                            afl_diagnostics* out) {
  if (!op || !rho || !out) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  int rc;
  if ((rc = ensure_common(op)) || (rc = op->diagw.alloc(diag_work_size(op->o)))) return rc;
  launch_nodal_stats(op->o, rho, u, p, op->diagw.p, op->dsc.p + kScA + 4);
  if ((rc = fetch(op, kScA + 4, 4))) return rc;
  const double* h = op->host_sc + kScA + 4;
  const OpCtx& o = op->o;
  out->min_density = -h[0];
  out->max_density = h[1];
  out->max_celerity = p ? h[2] : 0.0;
  out->max_velocity = u ? h[3] : 0.0;
  out->min_cell_diameter = op->min_diam;
  const double f = o.dt * o.k * std::sqrt((double)o.dim) / op->min_diam;
  out->cfl_c = f * out->max_celerity / o.Ma;
  out->cfl_u = f * out->max_velocity;
  return check_device("compute_diagnostics");
}

int afl_l2_difference(afl_operator* op, int ncomp, const double* a, const double* b, double* err_l2,
                      double* ref_l2) {
  if (!op || !a || !err_l2 || !ref_l2) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (ncomp != 1 && ncomp != op->o.dim) return fail(AFL_ERR_INVALID_ARGUMENT, "ncomp must be 1 or dim");
  int rc;
  if ((rc = ensure_common(op)) || (rc = op->diagw.alloc(diag_work_size(op->o)))) return rc;
  launch_l2_difference(op->o, ncomp, a, b, op->diagw.p, op->dsc.p + kScA + 4);
  if ((rc = fetch(op, kScA + 4, 2))) return rc;
  *err_l2 = std::sqrt(std::max(0.0, op->host_sc[kScA + 4]));
  *ref_l2 = std::sqrt(std::max(0.0, op->host_sc[kScA + 5]));
  return check_device("l2_difference");
}

int afl_l2_difference_noalloc(afl_operator* op, int ncomp, const double* a, const double* b, double* err_l2,
                              double* ref_l2) {
  if (!op || !a || !err_l2 || !ref_l2) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (ncomp != 1 && ncomp != op->o.dim) return fail(AFL_ERR_INVALID_ARGUMENT, "ncomp must be 1 or dim");
  int rc;
  double* scr = nullptr;
  if ((rc = current_device_scratch(&scr))) return rc;
  const int64_t nb = l2_blocks(op->o, kL2ScratchBlocks);
  if (2 * nb + 2 > kL2ScratchSize)
    return fail(AFL_ERR_INVALID_ARGUMENT, "l2_difference_noalloc: " + std::to_string(nb) +
                                              " partial blocks exceed the scratch of " +
                                              std::to_string(kL2ScratchBlocks));
  // every partial slot the finishing pass reads is written by the partial pass first: nothing to clear
  double* sums = scr + 2 * kL2ScratchBlocks;
  launch_l2_difference(op->o, ncomp, a, b, scr, sums, kL2ScratchBlocks);
  double h[2];
  if (hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, op->o.stream) != hipSuccess)
    return fail(AFL_ERR_DEVICE, "hipMemcpyAsync failed");
  if (hipStreamSynchronize(op->o.stream) != hipSuccess) return fail(AFL_ERR_DEVICE, "hipStreamSynchronize failed");
  *err_l2 = std::sqrt(std::max(0.0, h[0]));
  *ref_l2 = std::sqrt(std::max(0.0, h[1]));
  return check_device("l2_difference_noalloc");
}

int afl_compute_diagonal(afl_operator* op) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null handle");
  const unsigned ns = op->ns;
  const int comps = (ns == 3 || ns == 5) ? op->o.dim : 1;
  int rc;
  if ((rc = op->inv_diag.alloc(op->o.N * comps))) return rc;
  if (ns == 1 || ns == 4) launch_inv_diag(op->o, nullptr, 1.0, 1, op->inv_diag.p);
  else if (ns == 2 || ns == 6) launch_inv_diag(op->o, nullptr, 1.0 / (kGamma - 1.0), 1, op->inv_diag.p);
  else launch_inv_diag(op->o, op->rho_fix.p, 1.0, comps, op->inv_diag.p);
  op->diag_comps = comps;
  return check_device("compute_diagonal");
}

int afl_get_matrix_diagonal_inverse(afl_operator* op, double* dst) {
  if (!op || !dst) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (op->diag_comps == 0) return fail(AFL_ERR_STATE, "compute_diagonal has not been called");
  vec_copy(op->o.stream, op->o.N * op->diag_comps, op->inv_diag.p, dst);
  return check_device("get_matrix_diagonal_inverse");
}

int afl_solve_mass(afl_operator* op, double* x, const double* b) {
  if (!op || !x || !b) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  switch (op->ns) {
    case 1: case 4: launch_mass_solve(op->o, nullptr, 1.0, 1, b, x); break;
    case 6: launch_mass_solve(op->o, nullptr, 1.0 / (kGamma - 1.0), 1, b, x); break;
    case 3: case 5: launch_mass_solve(op->o, op->rho_fix.p, 1.0, op->o.dim, b, x); break;
    default: return fail(AFL_ERR_STATE, "solve_mass: NS_stage 2 is the Schur system, use afl_schur_solve");
  }
  return check_device("solve_mass");
}

// Schur solve S p = rhs from the initial guess in p; Sp0 (optional) holds S p already.
static int schur_solve(afl_operator* op, double* p, const double* rhs, const double* Sp0, int* iters,
                       double* residual) {
  int rc;
  if ((rc = ensure_common(op))) return rc;
  // the tolerance eps * |rhs|: the norm is fetched together with the first residual norm
  vec_norm2(op->o.stream, op->o.N, rhs, op->red.p, op->dsc.p + kScA);
  int it = 0;
  double res = 0.0;
  if (op->sp.schur_precond == AFL_SCHUR_JACOBI)
    rc = gmres(op, p, rhs, op->sp.gmres_eps, op->sp.gmres_max_iter, op->sp.gmres_restart, &it, &res, Sp0, true);
  else
    rc = fgmres(op, p, rhs, op->sp.gmres_eps, op->sp.gmres_max_iter, op->sp.gmres_restart, &it, &res, Sp0, true);
  if (iters) *iters = it;
  if (residual) *residual = res;
  return rc;
}

int afl_schur_solve(afl_operator* op, double* p, const double* rhs, int* iters, double* residual) {
  if (!op || !p || !rhs) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  return schur_solve(op, p, rhs, nullptr, iters, residual);
}

int afl_imex_step(afl_operator* op, double* rho, double* u, double* p, afl_step_info* info) {
  if (!op || !rho || !u || !p) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  const int64_t N = op->o.N, NU = N * op->o.dim;
  hipStream_t st = op->o.stream;
  int rc;
  DevBuf* sc[] = {&op->rho2, &op->rho3, &op->rhoc, &op->p2, &op->p3, &op->pc, &op->rhs_r,
                  &op->rhs_p, &op->tmp2, &op->pfix, &op->pfix_old, &op->rhs_p0};
  for (DevBuf* b : sc)
    if ((rc = b->alloc(N))) return rc;
  DevBuf* vc[] = {&op->u2, &op->u3, &op->uc, &op->rhs_u, &op->tmp1, &op->Bp, &op->ufix, &op->rhs_u0};
  for (DevBuf* b : vc)
    if ((rc = b->alloc(NU))) return rc;
  if ((rc = ensure_common(op))) return rc;
  afl_step_info inf{};
  OpCtx& o = op->o;

  auto fixed_point = [&](int stage, const double* rho_new, const double* p_start, const double* u_start,
                         double* p_out, double* u_out) -> int {
    vec_copy(st, N, p_start, op->pfix_old.p);
    vec_copy(st, NU, u_start, op->ufix.p);
    // Loop invariants of pressure_fixed_point (Atmospheric_Flow.cc:545-673), which recomputes them on
    // every iteration from the same inputs: the momentum right-hand side E2 (no source changes inside
    // the loop), tmp1 = A(rho_new)^{-1} E2, and the part of E3 that does not read u_fixed. Each
    // iteration adds E3's two u_fixed terms to that part (launch_rhs_p_ufix).
    if (stage == 1) {
      const double* sp_[5] = {rho, u, p, rho_new, op->ufix.p};
      launch_rhs_p(o, 1, sp_, op->rhs_p0.p, false);
      const double* su_[4] = {rho, u, p, rho_new};
      launch_rhs_u(o, 1, su_, op->rhs_u0.p);
    } else {
      const double* sp_[8] = {rho, u, p, op->rho2.p, op->u2.p, op->p2.p, rho_new, op->ufix.p};
      launch_rhs_p(o, 2, sp_, op->rhs_p0.p, false);
      const double* su_[7] = {rho, u, p, op->rho2.p, op->u2.p, op->p2.p, rho_new};
      launch_rhs_u(o, 2, su_, op->rhs_u0.p);
    }
    launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, op->rhs_u0.p, op->tmp1.p);
    // Bp holds w = A^{-1} B pfix_old from the second iteration on: the velocity update computes it
    // (u_fix = A^{-1} E2 - w, A^{-1} being linear), and the next Schur solve starts from
    // S pfix_old = D pfix_old - C(p_fix; w) without a second B apply and cell solve (C's coefficient
    // p_fix changes between iterations, w does not).
    bool have_w = false;
    int it = 0, conv = 0, gm_total = 0;
    for (; it < op->sp.fp_max_iter; ++it) {
      // pressure_fixed_point, Atmospheric_Flow.cc:545-673
      vec_copy(st, N, op->pfix_old.p, op->p_fix.p);
      op->mg_coef_stale = true;
      vec_copy(st, NU, op->ufix.p, op->u_fix.p);
      launch_rhs_p_ufix(o, stage, rho_new, op->ufix.p, op->rhs_p0.p, op->rhs_p.p);
      launch_apply_C(o, op->p_fix.p, op->tmp1.p, nullptr, op->tmp2.p);
      vec_axpby(st, N, -1.0, op->tmp2.p, 1.0, op->rhs_p.p);
      vec_copy(st, N, op->pfix_old.p, op->pfix.p);
      if (have_w) launch_apply_C(o, op->p_fix.p, op->Bp.p, op->pfix.p, op->tmp2.p);  // S pfix_old
      int gi = 0;
      double gr = 0.0;
      const int r = schur_solve(op, op->pfix.p, op->rhs_p.p, have_w ? op->tmp2.p : nullptr, &gi, &gr);
      gm_total += gi;
      inf.schur_residual = gr;
      if (r) return r;
      if (have_w && gi == 0 && op->sp.fp_tol > 0.0) {
        // the solve kept pfix = pfix_old: the velocity update would reproduce u_fix and the
        // pressure change is exactly zero, so the fixed point has converged
        conv = 1;
        ++it;
        break;
      }
      // update_velocity, Atmospheric_Flow.cc:681-781: u_fix = A^{-1} (E2 - B p)
      launch_schur_w(o, op->rho_fix.p, op->pfix.p, op->Bp.p);
      vec_sub(st, NU, op->tmp1.p, op->Bp.p, op->ufix.p);
      have_w = true;
      // relative inf-norm change of the nodal pressure, Atmospheric_Flow.cc:1140-1150
      vec_linf(st, N, op->pfix.p, op->red.p, op->dsc.p + kScA);
      vec_linf_diff(st, N, op->pfix.p, op->pfix_old.p, op->red.p, op->dsc.p + kScB);
      int rr;
      if ((rr = fetch(op, kScA, 2))) return rr;
      const double den = op->host_sc[kScA], num = op->host_sc[kScB];
      if (num / (den + 1e-10) < op->sp.fp_tol) { conv = 1; ++it; break; }
      vec_copy(st, N, op->pfix.p, op->pfix_old.p);
    }
    inf.fp_iters[stage - 1] = it;
    inf.fp_converged[stage - 1] = conv;
    inf.gmres_iters[stage - 1] = gm_total;
    vec_copy(st, N, op->pfix.p, p_out);
    vec_copy(st, NU, op->ufix.p, u_out);
    return AFL_OK;
  };

  auto& sect = op->sect;
  int seg;
  // stage 1
  o.hyp = 1;
  seg = sect.open(st, AFL_SECTION_UPDATE_DENSITY);
  { const double* s[2] = {rho, u}; launch_rhs_rho(o, 1, s, op->rhs_r.p); }
  launch_mass_solve(o, nullptr, 1.0, 1, op->rhs_r.p, op->rho2.p);
  sect.close(st, seg);
  vec_copy(st, N, op->rho2.p, op->rho_fix.p);
  op->mg_coef_stale = true;
  op->mg_eig_stale = true;
  seg = sect.open(st, AFL_SECTION_FIXED_POINT_PRESSURE);
  rc = fixed_point(1, op->rho2.p, p, u, op->p2.p, op->u2.p);
  sect.close(st, seg);
  if (rc) return rc;
  // stage 2
  o.hyp = 2;
  seg = sect.open(st, AFL_SECTION_UPDATE_DENSITY);
  { const double* s[4] = {rho, u, op->rho2.p, op->u2.p}; launch_rhs_rho(o, 2, s, op->rhs_r.p); }
  launch_mass_solve(o, nullptr, 1.0, 1, op->rhs_r.p, op->rho3.p);
  sect.close(st, seg);
  vec_copy(st, N, op->rho3.p, op->rho_fix.p);
  op->mg_coef_stale = true;
  op->mg_eig_stale = true;
  seg = sect.open(st, AFL_SECTION_FIXED_POINT_PRESSURE);
  rc = fixed_point(2, op->rho3.p, op->p2.p, op->u2.p, op->p3.p, op->u3.p);
  sect.close(st, seg);
  if (rc) return rc;
  // stage 3
  o.hyp = 3;
  seg = sect.open(st, AFL_SECTION_UPDATE_DENSITY);
  { const double* s[6] = {rho, u, op->rho2.p, op->u2.p, op->rho3.p, op->u3.p}; launch_rhs_rho(o, 3, s, op->rhs_r.p); }
  launch_mass_solve(o, nullptr, 1.0, 1, op->rhs_r.p, op->rhoc.p);
  sect.close(st, seg);
  vec_copy(st, N, op->rhoc.p, op->rho_fix.p);
  op->mg_coef_stale = true;
  op->mg_eig_stale = true;
  seg = sect.open(st, AFL_SECTION_UPDATE_VELOCITY);
  {
    const double* s[9] = {rho, u, p, op->rho2.p, op->u2.p, op->p2.p, op->rho3.p, op->u3.p, op->p3.p};
    launch_rhs_u(o, 3, s, op->rhs_u.p);
  }
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, op->rhs_u.p, op->uc.p);
  sect.close(st, seg);
  seg = sect.open(st, AFL_SECTION_UPDATE_PRESSURE);
  {
    const double* s[11] = {rho, u, p, op->rho2.p, op->u2.p, op->p2.p, op->rho3.p, op->u3.p, op->p3.p, op->rhoc.p, op->uc.p};
    launch_rhs_p(o, 3, s, op->rhs_p.p);
  }
  launch_mass_solve(o, nullptr, 1.0 / (kGamma - 1.0), 1, op->rhs_p.p, op->pc.p);
  sect.close(st, seg);
  vec_copy(st, N, op->rhoc.p, rho);
  vec_copy(st, NU, op->uc.p, u);
  vec_copy(st, N, op->pc.p, p);
  if (op->sponge.enabled)
    launch_sponge(o, op->sponge.z_start, op->sponge.alpha_max * o.dt, op->sponge.u_ref, u);
  if (info) *info = inf;
  return check_device("imex_step");
}

// ---- adjoint of one IMEX step ----------------------------------------------------------------
// Reverse of afl_imex_step: the forward step is recomputed from (rho, u, p) to restore the stage
// vectors, then every stage is differentiated backwards. The implicit stages are differentiated at
// their converged fixed point (p*, u*) by the implicit-function theorem applied to
//   R_u = A(rho_new) u* + B p* - E2(x) = 0,
//   R_p = D p* - C(A^{-1} B p*; p*) + C(A^{-1} E2(x); p*) - E3(x, rho_new, u*) = 0,
// the equations the fixed-point loop of Atmospheric_Flow.cc:545-781 converges to (C is applied to the
// two velocities separately, as the loop does). The transposed system is reduced to the pressure
// multiplier, Z lam_p = pbar - B^T A^{-1} ubar with
//   Z = D - B^T A^{-1} C'(w_B)^T - C_pf(w_B)^T + C_pf(w_E)^T + B^T A^{-1} E3_u^T,
// solved by GMRES with the Jacobi(D) preconditioner, and lam_u = A^{-1} (ubar + E3_u^T lam_p).
extern "C++" {
namespace {

struct StepAdj {
  afl_operator* op;
  int64_t N, NU;
  double* b[afl_operator::kAdjBufs];
  hipStream_t st;
  void acc(double* dst, const double* src, int64_t n) { vec_axpby(st, n, 1.0, src, 1.0, dst); }
};

// adjoint of one converged implicit stage; tgt[i] receives the adjoint of E3 source i (i < n3 - 1,
// E2 source i for i < n2); the E3 source n3 - 1 is u_fix = u*. rnew_pos: position of rho_new.
int fp_adjoint(StepAdj& S, int stage, const double* rho_new, const double* ps, const double* us, const double* pbar,
               const double* ubar, const double* const* e2s, int n2, const double* const* e3s, int n3,
               double* const* tgt, int rnew_pos) {
  afl_operator* op = S.op;
  OpCtx& o = op->o;
  const int64_t N = S.N, NU = S.NU;
  hipStream_t st = S.st;
  double* const* B = S.b;
  double *wB = B[22], *wE = B[23], *lamp = B[24], *lamu = B[25], *t1 = B[26], *t2 = B[27], *t3 = B[28], *r = B[30];
  double* outs[11];
  for (int i = 0; i < 11; ++i) outs[i] = B[9 + i];
  o.hyp = (unsigned)stage;
  vec_copy(st, N, rho_new, op->rho_fix.p);
  vec_copy(st, N, ps, op->p_fix.p);
  op->mg_coef_stale = true;
  op->mg_eig_stale = true;
  const TableauC& t = o.tab;
  const double Ma2 = o.Ma * o.Ma, iFr2 = 1.0 / (o.Fr * o.Fr);
  const double cgz = (stage == 1 ? t.at22 : t.at33) * o.dt * Ma2 * iFr2;  // gravity weight of (rho_new, u_fix)
  launch_schur_w(o, op->rho_fix.p, ps, wB);
  launch_rhs_u(o, stage, e2s, t1);
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, t1, wE);
  // r = pbar - B^T A^{-1} ubar
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, ubar, t1);
  launch_apply_Bt(o, t1, r);
  vec_axpby(st, N, 1.0, pbar, -1.0, r);
  auto Z = [&](const double* in, double* out) {
    launch_apply_C_tangent_t_svm(o, op->p_fix.p, in, wB, rho_new, us, -Ma2, -cgz, t1);
    launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, t1, t2);
    launch_apply_Bt(o, t2, out);
    // out = D in - out - C_pf(w_B)^T in + C_pf(w_E)^T in
    launch_apply_C_coef_pair_t(o, wE, wB, in, 1.0 / (kGamma - 1.0), out, out);
  };
  int rc;
  if ((rc = ensure_common(op))) return rc;
  vec_norm2(st, N, r, op->red.p, op->dsc.p + kScA);
  if ((rc = fetch(op, kScA, 1))) return rc;
  const double tol = op->sp.gmres_eps * op->host_sc[kScA];
  if ((rc = op->adj_lam_prev[stage - 1].alloc(N))) return rc;
  if (op->adj_lam_valid[stage - 1] && op->host_sc[kScA] > 0.0) vec_copy(st, N, op->adj_lam_prev[stage - 1].p, lamp);
  else vec_fill(st, N, 0.0, lamp);
  int it = 0;
  double res = 0.0;
  // Z differs from the Schur operator's transpose by coefficient terms, and the Schur operator is
  // nearly symmetric (C = -kappa B^T up to the Rusanov term), so the Schur preconditioner serves Z
  if (op->host_sc[kScA] > 0.0) {
    if (op->sp.schur_precond == AFL_SCHUR_JACOBI)
      rc = gmres_op(op, Z, lamp, r, tol, op->sp.gmres_max_iter, op->sp.gmres_restart, &it, &res,
                    "adjoint stage GMRES did not converge");
    else
      rc = fgmres_op(op, Z, lamp, r, tol, op->sp.gmres_max_iter, op->sp.gmres_restart, &it, &res,
                     "adjoint stage FGMRES did not converge");
    if (rc) return rc;
  }
  op->adj_krylov_its += it;
  vec_copy(st, N, lamp, op->adj_lam_prev[stage - 1].p);
  op->adj_lam_valid[stage - 1] = true;
  // lam_u = A^{-1} (ubar + E3_u^T lam_p)
  launch_scaled_velocity_mass(o, lamp, rho_new, us, -Ma2, -cgz, t1);
  vec_axpby(st, NU, 1.0, ubar, 1.0, t1);
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, t1, lamu);
  // E2 sources: VJP of lam_u - A^{-1} C'(w_E)^T lam_p
  launch_apply_C_tangent_t(o, op->p_fix.p, lamp, wE, t1);
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, t1, t2);  // t2 = A^{-1} C'(w_E)^T lam_p (kept below)
  vec_copy(st, NU, lamu, B[29]);
  vec_axpby(st, NU, -1.0, t2, 1.0, B[29]);
  launch_vjp_rhs_u(o, stage, B[29], e2s, outs);
  {
    AccList L;
    for (int i = 0; i < n2; ++i) L.add(tgt[i], outs[i], i % 3 == 1 ? NU : N);
    vec_acc_multi(st, L);
  }
  // rho_new through A(rho_new): + int phi (A^{-1} C'(w_E)^T lam_p) . w_E - int phi lam_u . u*
  launch_dot_mass(o, t2, wE, t3);
  S.acc(tgt[rnew_pos], t3, N);
  launch_dot_mass(o, lamu, us, t3);
  vec_axpby(st, N, -1.0, t3, 1.0, tgt[rnew_pos]);
  //                     - int phi (A^{-1} C'(w_B)^T lam_p) . w_B
  launch_apply_C_tangent_t(o, op->p_fix.p, lamp, wB, t1);
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, t1, t2);
  launch_dot_mass(o, t2, wB, t3);
  vec_axpby(st, N, -1.0, t3, 1.0, tgt[rnew_pos]);
  // E3 sources except u_fix: VJP of lam_p
  launch_vjp_rhs_p(o, stage, lamp, e3s, outs);
  {
    AccList L;
    for (int i = 0; i < n3 - 1; ++i) {
      const bool vel = i < 3 * stage ? (i % 3 == 1) : false;
      L.add(tgt[i], outs[i], vel ? NU : N);
    }
    vec_acc_multi(st, L);
  }
  return AFL_OK;
}

}  // namespace
}  // extern "C++"

// Reverse part of the step adjoint, given the step's input state and its stage vectors
// stg = {rho2, u2, p2, rho3, u3, p3, rho_c, u_c (before the sponge)}.
int step_vjp_reverse(afl_operator* op, const double* rho, const double* u, const double* p, const double* const* stg,
                     const double* rho_bar_out, const double* u_bar_out, const double* p_bar_out, double* rho_bar,
                     double* u_bar, double* p_bar) {
  OpCtx& o = op->o;
  const int64_t N = o.N, NU = N * o.dim;
  hipStream_t st = o.stream;
  int rc;
  for (DevBuf& b : op->aj)
    if ((rc = b.alloc(NU))) return rc;
  StepAdj S{op, N, NU, {}, st};
  for (int i = 0; i < afl_operator::kAdjBufs; ++i) S.b[i] = op->aj[i].p;
  double* const* B = S.b;
  const double *r2 = stg[0], *u2 = stg[1], *p2 = stg[2], *r3 = stg[3], *u3 = stg[4], *p3 = stg[5];
  const double *rc_ = stg[6], *uc = stg[7];
  double *rb2 = B[0], *ub2 = B[1], *pb2 = B[2], *rb3 = B[3], *ub3 = B[4], *pb3 = B[5], *rbc = B[6], *ubc = B[7],
         *pbc = B[8], *e = B[20], *eU = B[21], *t3 = B[28];
  double* outs[11];
  for (int i = 0; i < 11; ++i) outs[i] = B[9 + i];
  vec_fill(st, N, 0.0, rho_bar);
  vec_fill(st, NU, 0.0, u_bar);
  vec_fill(st, N, 0.0, p_bar);
  for (double* z : {rb2, pb2, rb3, pb3}) vec_fill(st, N, 0.0, z);
  for (double* z : {ub2, ub3}) vec_fill(st, NU, 0.0, z);
  vec_copy(st, N, rho_bar_out, rbc);
  vec_copy(st, NU, u_bar_out, ubc);
  vec_copy(st, N, p_bar_out, pbc);
  if (op->sponge.enabled) {  // u <- (u + a u_ref) / (1 + a): the transpose divides by 1 + a
    const double zero[3] = {0.0, 0.0, 0.0};
    launch_sponge(o, op->sponge.z_start, op->sponge.alpha_max * o.dt, zero, ubc);
  }
  // ---- stage 3: pc = D^{-1} E3, uc = A(rho_c)^{-1} E2, rho_c = M^{-1} E1
  o.hyp = 3;
  vec_copy(st, N, rc_, op->rho_fix.p);
  op->mg_coef_stale = true;
  op->mg_eig_stale = true;
  launch_mass_solve(o, nullptr, 1.0 / (kGamma - 1.0), 1, pbc, e);
  {
    const double* s11[11] = {rho, u, p, r2, u2, p2, r3, u3, p3, rc_, uc};
    launch_vjp_rhs_p(o, 3, e, s11, outs);
    double* tg[11] = {rho_bar, u_bar, p_bar, rb2, ub2, pb2, rb3, ub3, pb3, rbc, ubc};
    AccList L;
    for (int i = 0; i < 11; ++i) L.add(tg[i], outs[i], (i % 3 == 1 && i < 9) || i == 10 ? NU : N);
    vec_acc_multi(st, L);
  }
  launch_mass_solve(o, op->rho_fix.p, 1.0, o.dim, ubc, eU);
  launch_dot_mass(o, eU, uc, t3);
  vec_axpby(st, N, -1.0, t3, 1.0, rbc);
  {
    const double* s9[9] = {rho, u, p, r2, u2, p2, r3, u3, p3};
    launch_vjp_rhs_u(o, 3, eU, s9, outs);
    double* tg[9] = {rho_bar, u_bar, p_bar, rb2, ub2, pb2, rb3, ub3, pb3};
    AccList L;
    for (int i = 0; i < 9; ++i) L.add(tg[i], outs[i], i % 3 == 1 ? NU : N);
    vec_acc_multi(st, L);
  }
  launch_mass_solve(o, nullptr, 1.0, 1, rbc, e);
  {
    const double* s6[6] = {rho, u, r2, u2, r3, u3};
    launch_vjp_rhs_rho(o, 3, e, s6, outs);
    double* tg[6] = {rho_bar, u_bar, rb2, ub2, rb3, ub3};
    AccList L;
    for (int i = 0; i < 6; ++i) L.add(tg[i], outs[i], i % 2 == 1 ? NU : N);
    vec_acc_multi(st, L);
  }
  // ---- stage 2: fixed point (p3, u3) with rho_new = rho3, then rho3 = M^{-1} E1
  {
    const double* e2s[7] = {rho, u, p, r2, u2, p2, r3};
    const double* e3s[8] = {rho, u, p, r2, u2, p2, r3, u3};
    double* tg[7] = {rho_bar, u_bar, p_bar, rb2, ub2, pb2, rb3};
    if ((rc = fp_adjoint(S, 2, r3, p3, u3, pb3, ub3, e2s, 7, e3s, 8, tg, 6))) return rc;
  }
  o.hyp = 2;
  launch_mass_solve(o, nullptr, 1.0, 1, rb3, e);
  {
    const double* s4[4] = {rho, u, r2, u2};
    launch_vjp_rhs_rho(o, 2, e, s4, outs);
    double* tg[4] = {rho_bar, u_bar, rb2, ub2};
    AccList L;
    for (int i = 0; i < 4; ++i) L.add(tg[i], outs[i], i % 2 == 1 ? NU : N);
    vec_acc_multi(st, L);
  }
  // ---- stage 1: fixed point (p2, u2) with rho_new = rho2, then rho2 = M^{-1} E1
  {
    const double* e2s[4] = {rho, u, p, r2};
    const double* e3s[5] = {rho, u, p, r2, u2};
    double* tg[4] = {rho_bar, u_bar, p_bar, rb2};
    if ((rc = fp_adjoint(S, 1, r2, p2, u2, pb2, ub2, e2s, 4, e3s, 5, tg, 3))) return rc;
  }
  o.hyp = 1;
  launch_mass_solve(o, nullptr, 1.0, 1, rb2, e);
  {
    const double* s2[2] = {rho, u};
    launch_vjp_rhs_rho(o, 1, e, s2, outs);
    AccList L;
    L.add(rho_bar, outs[0], N);
    L.add(u_bar, outs[1], NU);
    vec_acc_multi(st, L);
  }
  return AFL_OK;
}

int afl_imex_step_vjp(afl_operator* op, const double* rho, const double* u, const double* p, const double* rho_bar_out,
                      const double* u_bar_out, const double* p_bar_out, double* rho_bar, double* u_bar, double* p_bar,
                      afl_step_info* info) {
  if (!op || !rho || !u || !p || !rho_bar_out || !u_bar_out || !p_bar_out || !rho_bar || !u_bar || !p_bar)
    return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  OpCtx& o = op->o;
  const int64_t N = o.N, NU = N * o.dim;
  hipStream_t st = o.stream;
  int rc;
  for (DevBuf& b : op->aj)
    if ((rc = b.alloc(NU))) return rc;
  // forward recompute from copies of the input state: restores the stage vectors
  vec_copy(st, N, rho, op->aj[31].p);
  vec_copy(st, NU, u, op->aj[32].p);
  vec_copy(st, N, p, op->aj[33].p);
  if ((rc = afl_imex_step(op, op->aj[31].p, op->aj[32].p, op->aj[33].p, info))) return rc;
  const double* stg[8] = {op->rho2.p, op->u2.p, op->p2.p, op->rho3.p, op->u3.p, op->p3.p, op->rhoc.p, op->uc.p};
  if ((rc = step_vjp_reverse(op, rho, u, p, stg, rho_bar_out, u_bar_out, p_bar_out, rho_bar, u_bar, p_bar))) return rc;
  return check_device("imex_step_vjp");
}

// ---- gradient over many steps with binomial checkpointing ------------------------------------
// The reverse sweep over steps [a, b) starts from the state at a (held in a snapshot) and s free
// snapshots: one step is reversed directly; otherwise the state is advanced to a split point m, kept
// in a snapshot, [m, b) is reversed with s - 1 snapshots and then [a, m) with s. The split minimises
// the forward steps, T(t, s) = min_m [m + T(t - m, s - 1) + T(m, s)], T(1, s) = 0, T(t, 0) =
// t (t - 1) / 2 (Griewank's binomial checkpointing; the optimum is found by dynamic programming).
extern "C++" {
namespace {

struct Schedule {
  int T = 0, S = 0;
  std::vector<int64_t> cost;
  std::vector<int> split;
  int64_t& c(int t, int s) { return cost[(size_t)t * (S + 1) + s]; }
  int& m(int t, int s) { return split[(size_t)t * (S + 1) + s]; }
  void build(int t_max, int s_max) {
    T = t_max; S = s_max;
    cost.assign((size_t)(T + 1) * (S + 1), 0);
    split.assign((size_t)(T + 1) * (S + 1), 0);
    for (int t = 1; t <= T; ++t)
      for (int s = 0; s <= S; ++s) {
        if (t == 1) { c(t, s) = 0; continue; }
        if (s == 0) { c(t, s) = (int64_t)t * (t - 1) / 2; continue; }
        int64_t best = -1;
        int bm = 1;
        for (int mm = 1; mm < t; ++mm) {
          const int64_t v = mm + c(t - mm, s - 1) + c(mm, s);
          if (best < 0 || v < best) { best = v; bm = mm; }
        }
        c(t, s) = best;
        m(t, s) = bm;
      }
  }
};

// Stage sets: with keep_stages, an advance that ends at a split point keeps the stage vectors of its
// last step (rho2, u2, p2, rho3, u3, p3, rho_c, u_c). Every range of the reversal ends at the step
// count or at a split point, so the reversal of a step whose set is held restores nothing: it runs
// step_vjp_reverse on the set directly (a set lives from its advance to that step's reversal; at most
// snapshots + 1 are held). Steps without a set (the last step, and the ranges reversed without
// snapshots) recompute their forward step as afl_imex_step_vjp does.
struct AdjRun {
  afl_operator* op;
  int64_t N, NU;
  std::vector<DevBuf> snap;  // 3 buffers per slot: rho, u, p
  DevBuf wr, wu, wp;         // working state
  DevBuf lam[2][3];          // adjoint ping-pong
  int cur = 0, used = 0, max_used = 0;
  std::vector<int> free_slots;
  Schedule sch;
  afl_adjoint_stats st{};
  bool keep_stages = true;
  static constexpr int kStg = 8;
  std::vector<DevBuf> sets;     // kStg buffers per stage set
  std::vector<int> set_step;    // step whose stages the set holds, -1 when free
  int sets_held = 0;
  // keeps the stage vectors of the step just taken (step index j) in a free set
  int keep(int j) {
    int k = -1;
    for (int i = 0; i < (int)set_step.size(); ++i)
      if (set_step[i] < 0) { k = i; break; }
    if (k < 0) {
      k = (int)set_step.size();
      set_step.push_back(-1);
      sets.resize(sets.size() + kStg);
      const bool vec[kStg] = {false, true, false, false, true, false, false, true};
      for (int f = 0; f < kStg; ++f) {
        int rc;
        if ((rc = sets[(size_t)kStg * k + f].alloc(vec[f] ? NU : N))) return rc;
      }
    }
    const bool vec[kStg] = {false, true, false, false, true, false, false, true};
    const double* sv[kStg] = {op->rho2.p, op->u2.p, op->p2.p, op->rho3.p, op->u3.p, op->p3.p, op->rhoc.p, op->uc.p};
    for (int f = 0; f < kStg; ++f) vec_copy(stream(), vec[f] ? NU : N, sv[f], sets[(size_t)kStg * k + f].p);
    set_step[k] = j;
    ++sets_held;
    st.stage_sets_used = std::max(st.stage_sets_used, sets_held);
    return AFL_OK;
  }
  int find_set(int j) const {
    for (int i = 0; i < (int)set_step.size(); ++i)
      if (set_step[i] == j) return i;
    return -1;
  }
  hipStream_t stream() const { return op->o.stream; }
  double* sr(int k) { return snap[3 * k].p; }
  double* su(int k) { return snap[3 * k + 1].p; }
  double* sp(int k) { return snap[3 * k + 2].p; }
  int take() {
    const int k = free_slots.back();
    free_slots.pop_back();
    max_used = std::max(max_used, ++used);
    return k;
  }
  void give(int k) { free_slots.push_back(k); --used; }
  // working state <- state at slot k (step a) advanced by n steps; with keep_set the stage vectors
  // of the last of them (step a + n - 1) are kept
  int advance(int k, int a, int n, bool keep_set) {
    vec_copy(stream(), N, sr(k), wr.p);
    vec_copy(stream(), NU, su(k), wu.p);
    vec_copy(stream(), N, sp(k), wp.p);
    for (int i = 0; i < n; ++i) {
      int rc;
      if ((rc = afl_imex_step(op, wr.p, wu.p, wp.p, nullptr))) return rc;
      ++st.forward_steps;
      if (keep_set && keep_stages && i == n - 1 && (rc = keep(a + i))) return rc;
    }
    return AFL_OK;
  }
  // reverses step j from the working state (the state at step j)
  int vjp_from_working(int j) {
    const int nx = 1 - cur;
    const int k = find_set(j);
    int rc;
    if (k >= 0) {
      const double* stg[kStg];
      for (int f = 0; f < kStg; ++f) stg[f] = sets[(size_t)kStg * k + f].p;
      rc = step_vjp_reverse(op, wr.p, wu.p, wp.p, stg, lam[cur][0].p, lam[cur][1].p, lam[cur][2].p, lam[nx][0].p,
                            lam[nx][1].p, lam[nx][2].p);
      set_step[k] = -1;
      --sets_held;
    } else {
      rc = afl_imex_step_vjp(op, wr.p, wu.p, wp.p, lam[cur][0].p, lam[cur][1].p, lam[cur][2].p, lam[nx][0].p,
                             lam[nx][1].p, lam[nx][2].p, nullptr);
      ++st.recomputed_steps;
    }
    if (rc) return rc;
    cur = nx;
    ++st.adjoint_steps;
    if (op->verbose) {
      std::fprintf(stderr, "[afl] adjoint_run: %d steps reversed, %d forward steps\n", st.adjoint_steps,
                   st.forward_steps);
      std::fflush(stderr);
    }
    return AFL_OK;
  }
  int reverse(int a, int b, int slot, int s) {
    int rc;
    const int t = b - a;
    if (t == 1) {
      if ((rc = advance(slot, a, 0, false))) return rc;
      return vjp_from_working(a);
    }
    if (s == 0) {
      for (int k = b - 1; k >= a; --k) {
        if ((rc = advance(slot, a, k - a, false))) return rc;
        if ((rc = vjp_from_working(k))) return rc;
      }
      return AFL_OK;
    }
    const int mm = sch.m(t, s);
    if ((rc = advance(slot, a, mm, true))) return rc;
    const int k = take();
    vec_copy(stream(), N, wr.p, sr(k));
    vec_copy(stream(), NU, wu.p, su(k));
    vec_copy(stream(), N, wp.p, sp(k));
    if ((rc = reverse(a + mm, b, k, s - 1))) return rc;
    give(k);
    return reverse(a, a + mm, slot, s);
  }
};

}  // namespace
}  // extern "C++"

int afl_set_adjoint_stage_sets(afl_operator* op, int keep) {
  if (!op) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  op->adj_keep_stages = keep != 0;
  return AFL_OK;
}

int afl_adjoint_run(afl_operator* op, const double* rho0, const double* u0, const double* p0, int nsteps,
                    int snapshots, const double* rho_bar_T, const double* u_bar_T, const double* p_bar_T,
                    double* rho_bar_0, double* u_bar_0, double* p_bar_0, afl_adjoint_stats* stats) {
  if (!op || !rho0 || !u0 || !p0 || !rho_bar_T || !u_bar_T || !p_bar_T || !rho_bar_0 || !u_bar_0 || !p_bar_0)
    return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (nsteps < 1 || snapshots < AFL_ADJOINT_STORE_ALL || snapshots > 4096)
    return fail(AFL_ERR_INVALID_ARGUMENT, "adjoint_run: nsteps >= 1 and AFL_ADJOINT_STORE_ALL <= snapshots <= 4096");
  const int64_t N = op->N(), NU = N * op->dim();
  op->adj_krylov_its = 0;
  if (snapshots == AFL_ADJOINT_STORE_ALL) {
    // one forward sweep keeping every step's input state and stage vectors, then the reverse sweep
    // with no recomputation
    constexpr int kF = 11;  // rho, u, p, rho2, u2, p2, rho3, u3, p3, rho_c, u_c
    const bool vec[kF] = {false, true, false, false, true, false, false, true, false, false, true};
    // one allocation for every step's fields (thousands of separate allocations cost a visible
    // fraction of a forward run on small meshes); field f of step k at slot(k, f)
    int64_t per_step = 0;
    int64_t off[kF];
    for (int f = 0; f < kF; ++f) {
      off[f] = per_step;
      per_step += vec[f] ? NU : N;
    }
    DevBuf all;
    DevBuf x[3], lam[2][3];
    int rc;
    auto release = [&]() {
      all.release();
      for (DevBuf& b : x) b.release();
      for (auto& l : lam)
        for (DevBuf& b : l) b.release();
    };
    if ((rc = all.alloc(per_step * nsteps))) {
      release();
      return fail(AFL_ERR_DEVICE, "adjoint_run: not enough device memory to store every step; use snapshots");
    }
    auto slot = [&](int k, int f) { return all.p + per_step * k + off[f]; };
    for (int f = 0; f < 3; ++f)
      if ((rc = x[f].alloc(f == 1 ? NU : N)) || (rc = lam[0][f].alloc(f == 1 ? NU : N)) ||
          (rc = lam[1][f].alloc(f == 1 ? NU : N))) {
        release();
        return rc;
      }
    hipStream_t st = op->o.stream;
    afl_adjoint_stats as{};
    vec_copy(st, N, rho0, x[0].p);
    vec_copy(st, NU, u0, x[1].p);
    vec_copy(st, N, p0, x[2].p);
    for (int k = 0; k < nsteps && !rc; ++k) {
      for (int f = 0; f < 3; ++f) vec_copy(st, f == 1 ? NU : N, x[f].p, slot(k, f));
      if ((rc = afl_imex_step(op, x[0].p, x[1].p, x[2].p, nullptr))) break;
      const double* sv[8] = {op->rho2.p, op->u2.p, op->p2.p, op->rho3.p, op->u3.p, op->p3.p, op->rhoc.p, op->uc.p};
      for (int f = 0; f < 8; ++f) vec_copy(st, vec[3 + f] ? NU : N, sv[f], slot(k, 3 + f));
      ++as.forward_steps;
    }
    int cur = 0;
    if (!rc) {
      vec_copy(st, N, rho_bar_T, lam[0][0].p);
      vec_copy(st, NU, u_bar_T, lam[0][1].p);
      vec_copy(st, N, p_bar_T, lam[0][2].p);
    }
    for (int k = nsteps - 1; k >= 0 && !rc; --k) {
      const double* stg[8];
      for (int f = 0; f < 8; ++f) stg[f] = slot(k, 3 + f);
      const int nx = 1 - cur;
      rc = step_vjp_reverse(op, slot(k, 0), slot(k, 1), slot(k, 2), stg, lam[cur][0].p, lam[cur][1].p, lam[cur][2].p,
                            lam[nx][0].p, lam[nx][1].p, lam[nx][2].p);
      cur = nx;
      ++as.adjoint_steps;
      if (op->verbose) {
        std::fprintf(stderr, "[afl] adjoint_run: %d steps reversed\n", as.adjoint_steps);
        std::fflush(stderr);
      }
    }
    if (!rc) {
      vec_copy(st, N, lam[cur][0].p, rho_bar_0);
      vec_copy(st, NU, lam[cur][1].p, u_bar_0);
      vec_copy(st, N, lam[cur][2].p, p_bar_0);
      rc = check_device("adjoint_run");
    }
    hipStreamSynchronize(st);
    as.snapshots_used = nsteps;
    as.stage_solve_iterations = (int32_t)std::min<int64_t>(op->adj_krylov_its, INT32_MAX);
    if (stats) *stats = as;
    release();
    return rc;
  }
  const int s = std::min(snapshots, nsteps - 1);
  AdjRun R;
  R.op = op; R.N = N; R.NU = NU;
  R.keep_stages = op->adj_keep_stages;
  R.snap.resize((size_t)3 * (s + 1));
  int rc;
  for (int k = 0; k <= s; ++k)
    if ((rc = R.snap[3 * k].alloc(N)) || (rc = R.snap[3 * k + 1].alloc(NU)) || (rc = R.snap[3 * k + 2].alloc(N)))
      return rc;
  if ((rc = R.wr.alloc(N)) || (rc = R.wu.alloc(NU)) || (rc = R.wp.alloc(N))) return rc;
  for (int i = 0; i < 2; ++i)
    if ((rc = R.lam[i][0].alloc(N)) || (rc = R.lam[i][1].alloc(NU)) || (rc = R.lam[i][2].alloc(N))) return rc;
  for (int k = s; k >= 1; --k) R.free_slots.push_back(k);
  R.used = R.max_used = 1;  // slot 0: the initial state
  hipStream_t st = op->o.stream;
  vec_copy(st, N, rho0, R.sr(0));
  vec_copy(st, NU, u0, R.su(0));
  vec_copy(st, N, p0, R.sp(0));
  vec_copy(st, N, rho_bar_T, R.lam[0][0].p);
  vec_copy(st, NU, u_bar_T, R.lam[0][1].p);
  vec_copy(st, N, p_bar_T, R.lam[0][2].p);
  R.sch.build(nsteps, s);
  rc = R.reverse(0, nsteps, 0, s);
  if (!rc) {
    vec_copy(st, N, R.lam[R.cur][0].p, rho_bar_0);
    vec_copy(st, NU, R.lam[R.cur][1].p, u_bar_0);
    vec_copy(st, N, R.lam[R.cur][2].p, p_bar_0);
    rc = check_device("adjoint_run");
  }
  hipStreamSynchronize(st);
  R.st.snapshots_used = R.max_used;
  R.st.stage_solve_iterations = (int32_t)std::min<int64_t>(op->adj_krylov_its, INT32_MAX);
  if (stats) *stats = R.st;
  for (DevBuf& b : R.snap) b.release();
  for (DevBuf& b : R.sets) b.release();
  for (DevBuf* b : {&R.wr, &R.wu, &R.wp}) b->release();
  for (auto& l : R.lam)
    for (DevBuf& b : l) b.release();
  return rc;
}

// ---- observation operator ----------------------------------------------------------------------
int afl_set_observation_points(afl_operator* op, int npts, const double* xyz) {
  if (!op || npts < 0 || (npts > 0 && !xyz)) return fail(AFL_ERR_INVALID_ARGUMENT, "bad observation points");
  std::vector<int64_t> cell;
  std::vector<double> ref, zpt;
  const int bad = locate_points(op->o.geo, op->o.dim, npts, xyz, cell, ref, zpt);
  if (bad >= 0) return fail(AFL_ERR_INVALID_ARGUMENT, "observation point outside the domain");
  // group the points by cell (stable), for the deterministic transpose
  std::vector<int> order(npts);
  for (int k = 0; k < npts; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cell[a] < cell[b]; });
  std::vector<int64_t> ccell;
  std::vector<int> cstart;
  for (int q = 0; q < npts; ++q)
    if (q == 0 || cell[order[q]] != cell[order[q - 1]]) { ccell.push_back(cell[order[q]]); cstart.push_back(q); }
  cstart.push_back(npts);
  op->release_obs();
  ObsDev& ob = op->obs;
  ob.npts = npts;
  ob.ncp = (int)ccell.size();
  auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
    if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return false;
    return !bytes || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up((void**)&ob.cell, cell.data(), cell.size() * sizeof(int64_t)) ||
      !up((void**)&ob.ref, ref.data(), ref.size() * sizeof(double)) ||
      !up((void**)&ob.z, zpt.data(), zpt.size() * sizeof(double)) ||
      !up((void**)&ob.ccell, ccell.data(), ccell.size() * sizeof(int64_t)) ||
      !up((void**)&ob.cstart, cstart.data(), cstart.size() * sizeof(int)) ||
      !up((void**)&ob.order, order.data(), order.size() * sizeof(int)))
    return fail(AFL_ERR_DEVICE, "observation points: device copy failed");
  int rc;
  if ((rc = op->obs_f.alloc(2 * (int64_t)npts))) return rc;
  return AFL_OK;
}

int afl_observe_theta(afl_operator* op, const afl_hill_state* bg, const double* rho, const double* p,
                      double* theta_prime) {
  if (!op || !bg || !rho || !p || !theta_prime) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (op->obs.npts == 0) return AFL_OK;
  launch_observe(op->o, op->obs, rho, p, bg->theta0, bg->brunt * bg->brunt, bg->p0, theta_prime, nullptr, nullptr,
                 nullptr);
  return check_device("observe_theta");
}

int afl_observe_theta_transpose(afl_operator* op, const afl_hill_state* bg, const double* rho, const double* p,
                                const double* wbar, double* rho_bar, double* p_bar) {
  if (!op || !bg || !rho || !p || !wbar || !rho_bar || !p_bar) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  vec_fill(op->o.stream, op->N(), 0.0, rho_bar);
  vec_fill(op->o.stream, op->N(), 0.0, p_bar);
  if (op->obs.npts == 0) return check_device("observe_theta_transpose");
  double* fr = op->obs_f.p;
  double* fp = op->obs_f.p + op->obs.npts;
  launch_observe(op->o, op->obs, rho, p, bg->theta0, bg->brunt * bg->brunt, bg->p0, nullptr, wbar, fr, fp);
  launch_observe_t(op->o, op->obs, fr, fp, rho_bar, p_bar);
  return check_device("observe_theta_transpose");
}

int afl_init_hill_state(afl_operator* op, const afl_hill_state* s, double* rho, double* u, double* p) {
  if (!op || !s || !rho || !u || !p) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (!(s->brunt > 0.0) || !(s->theta0 > 0.0) || !(s->p0 > 0.0)) return fail(AFL_ERR_INVALID_ARGUMENT, "bad background");
  HillParams hp{s->theta0, s->brunt * s->brunt, s->p0, s->u0, s->theta_sigma, s->seed};
  launch_init_state(op->o, hp, rho, u, p);
  return check_device("init_hill_state");
}

int afl_background(afl_operator* op, const afl_hill_state* s, double* rho, double* p, double* th) {
  if (!op || !s) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (!(s->brunt > 0.0) || !(s->theta0 > 0.0) || !(s->p0 > 0.0)) return fail(AFL_ERR_INVALID_ARGUMENT, "bad background");
  int rc;
  DevBuf tr, tp, tt;
  double* r = rho; double* pp = p; double* t = th;
  if (!r) { if ((rc = tr.alloc(op->o.N))) return rc; r = tr.p; }
  if (!pp) { if ((rc = tp.alloc(op->o.N))) return rc; pp = tp.p; }
  if (!t) { if ((rc = tt.alloc(op->o.N))) return rc; t = tt.p; }
  HillParams hp{s->theta0, s->brunt * s->brunt, s->p0, s->u0, 0.0, 0};
  launch_background(op->o, hp, r, pp, t);
  rc = check_device("background");
  hipStreamSynchronize(op->o.stream);
  tr.release(); tp.release(); tt.release();
  return rc;
}

}  // extern "C"

// ---- layout conversion -------------------------------------------------------------------
namespace {
__global__ void k_pack(int64_t N, int npc, int ncomp, const double* __restrict__ src, const int64_t* __restrict__ perm,
                       int64_t cs, int64_t as, int64_t ns, double* __restrict__ dst, bool to_afl) {
  const int64_t total = N * ncomp;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(idx / N);
    const int64_t r = idx % N;
    const int64_t c = r / npc;
    const int i = (int)(r % npc);
    const int64_t pc = perm ? perm[c] : c;
    const int64_t ci = pc * cs + a * as + i * ns;
    if (to_afl) dst[idx] = src[ci];
    else dst[ci] = src[idx];
  }
}
}  // namespace

extern "C" int afl_mg_transfer(afl_operator* op, int k_coarse, int k_fine, int transpose, const double* src,
                               double* dst) {
  if (!op || !src || !dst) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (k_coarse < 1 || k_fine > 4 || k_coarse >= k_fine)
    return fail(AFL_ERR_INVALID_ARGUMENT, "mg_transfer: degrees must satisfy 1 <= k_coarse < k_fine <= 4");
  if (mg_transfer(op->o, op->tabs_by_k, k_coarse, k_fine, transpose != 0, src, dst))
    return fail(AFL_ERR_INVALID_ARGUMENT, "mg_transfer: unsupported degree pair");
  return check_device("mg_transfer");
}

extern "C" int afl_pack_dofs(afl_operator* op, int ncomp, const double* caller, const int64_t* perm, int64_t cs,
                             int64_t as, int64_t ns, double* out) {
  if (!op || !caller || !out || ncomp < 1) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (ncomp != 1 && ncomp != op->o.dim) return fail(AFL_ERR_INVALID_ARGUMENT, "ncomp must be 1 or dim");
  if (cs < 0 || as < 0 || ns < 0) return fail(AFL_ERR_INVALID_ARGUMENT, "strides must be non-negative");
  const int n1 = op->o.k + 1, npc = op->o.dim == 2 ? n1 * n1 : n1 * n1 * n1;
  int64_t b = (op->o.N * ncomp + AFL_NT - 1) / AFL_NT;
  k_pack<<<dim3((unsigned)(b > 65536 ? 65536 : b)), dim3(AFL_NT), 0, op->o.stream>>>(op->o.N, npc, ncomp, caller, perm, cs, as, ns, out, true);
  return check_device("pack_dofs");
}

extern "C" int afl_unpack_dofs(afl_operator* op, int ncomp, const double* in, const int64_t* perm, int64_t cs,
                               int64_t as, int64_t ns, double* caller) {
  if (!op || !caller || !in || ncomp < 1) return fail(AFL_ERR_INVALID_ARGUMENT, "null argument");
  if (ncomp != 1 && ncomp != op->o.dim) return fail(AFL_ERR_INVALID_ARGUMENT, "ncomp must be 1 or dim");
  if (cs < 0 || as < 0 || ns < 0) return fail(AFL_ERR_INVALID_ARGUMENT, "strides must be non-negative");
  const int n1 = op->o.k + 1, npc = op->o.dim == 2 ? n1 * n1 : n1 * n1 * n1;
  int64_t b = (op->o.N * ncomp + AFL_NT - 1) / AFL_NT;
  k_pack<<<dim3((unsigned)(b > 65536 ? 65536 : b)), dim3(AFL_NT), 0, op->o.stream>>>(op->o.N, npc, ncomp, in, perm, cs, as, ns, caller, false);
  return check_device("unpack_dofs");
}
