// Coupled elevation / sediment-flux solves: the two-field systems goSPL
// assembles as PETSc nested matrices for the transport-limited stream power
// (SPL._coupledEDSystem, gospl/eroder/SPL.py:116-258) and for marine
// deposition (SEAMesh._depMarineSystem, gospl/sed/seaplex.py:240-402).
//
// Both read
//     [E - diag(f)     -diag(f dt / a)] [h]   [r0]
//     [diag(a / dt)     F             ] [q] = [r1]
// with F = (I - W)^T the flow matrix of a receiver set (goSPL's matrixFlow,
// flowplex.py:429-478) and E the stream-power matrix of SPL._eroMats
// (SPL.py:45-114; the identity for the marine system).  Eliminating
// q = F^-1 (r1 - (a / dt) h) leaves the N x N system
//     S h = (E - diag(f) + diag(d) F^-1 diag(s)) h = r0 + d (F^-1 r1),
//     d = f dt / a,  s = a / dt,
// solved here by restarted GMRES right-preconditioned with E^-1.  One
// iteration applies E^-1 (one exact erosion-direction sweep of the receiver
// DAG), F^-1 (one exact flow-direction sweep) and the per-vertex rows -- the
// two triangular solves goSPL's Schur fieldsplit preconditioner applies per
// block (GASM/ILU, SPL.py:181-206), exact here, inside a Krylov method as
// goSPL's TFQMR.  The reductions are per-workgroup partial sums added in a
// fixed order, so a solve is bit-reproducible.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace gospl {
namespace {

constexpr int KRY_M = 20;       // restart length
constexpr int DOT_BLOCKS = 512;  // partial sums per reduction
constexpr int DOT_THREADS = 256;

// rows of the workspace (each N + 64 doubles): V (KRY_M + 1), Z (KRY_M), then
// w, t, y, b, x and six per-vertex vectors of the system
constexpr int ROW_W = 2 * KRY_M + 1;
constexpr int KRY_ROWS = ROW_W + 11;

__global__ __launch_bounds__(DOT_THREADS) void k_mdot(int64_t N, const double* __restrict__ V, int64_t ld,
                                                       const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[DOT_THREADS];
  const double* v = V + (int64_t)blockIdx.y * ld;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * DOT_THREADS + threadIdx.x; i < N; i += (int64_t)DOT_BLOCKS * DOT_THREADS)
    acc += v[i] * w[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = DOT_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * DOT_BLOCKS + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(DOT_THREADS) void k_msum(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[DOT_THREADS];
  const double* p = part + (int64_t)blockIdx.x * DOT_BLOCKS;
  double acc = 0.0;
  for (int i = threadIdx.x; i < DOT_BLOCKS; i += DOT_THREADS) acc += p[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = DOT_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// out = out + sign * sum_j c[j] V_j
__global__ void k_lincomb(int64_t N, const double* __restrict__ V, int64_t ld, int k, const double* __restrict__ c,
                          double sign, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  for (int j = 0; j < k; ++j) acc += c[j] * V[(int64_t)j * ld + i];
  out[i] = out[i] + sign * acc;
}

__global__ void k_scale(int64_t N, const double* __restrict__ w, double alpha, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = alpha * w[i];
}

__global__ void k_mul(int64_t N, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = a[i] * b[i];
}

// w = E z - f z + d y (E given by off/diag on the receivers, or the identity)
__global__ void k_coupled_apply(int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ off,
                                const double* __restrict__ diag, const double* __restrict__ f,
                                const double* __restrict__ d, const double* __restrict__ z,
                                const double* __restrict__ y, double* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double zi = z[i];
  double e;
  if (rcv) {
    e = diag[i] * zi;
    for (int k = 0; k < NR; ++k) e += off[(int64_t)k * N + i] * z[rcv[(int64_t)k * N + i]];
  } else {
    e = zi;
  }
  if (f) e = e - f[i] * zi;
  w[i] = e + d[i] * y[i];
}

// r = b - w
__global__ void k_sub(int64_t N, const double* __restrict__ b, const double* __restrict__ w, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) r[i] = b[i] - w[i];
}

// Transport-limited stream power (SPL.py:306-333, :155-174): fDep = G area /
// max(PA, G area / 0.99) (0 where PA = 0, at sea and at flat-mesh outlets,
// capped at 0.99), d = fDep dt / area, s = area / dt, sh = (hOld area) / dt.
__global__ void k_ed_prep(int64_t N, double G, const double* __restrict__ area, const double* __restrict__ fal,
                          double sl, const double* __restrict__ lfill, const int8_t* __restrict__ outlet, int flat,
                          double dt, const double* __restrict__ h, double* __restrict__ f, double* __restrict__ d,
                          double* __restrict__ s, double* __restrict__ sh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double a = area[i];
  const double num = G * a;
  const double floor = num / 0.99;
  const double pa = fal[i];
  double fd = 0.0;
  if (pa != 0.0) fd = num / (pa > floor ? pa : floor);
  if (lfill[i] <= sl) fd = 0.0;
  if (fd > 0.99) fd = 0.99;
  if (flat && outlet[i]) fd = 0.0;
  f[i] = fd;
  d[i] = fd * dt / a;
  s[i] = a / dt;
  sh[i] = (h[i] * a) * (1.0 / dt);
}

// b = (1 - f) hOld + d y0
__global__ void k_ed_rhs(int64_t N, const double* __restrict__ f, const double* __restrict__ d,
                         const double* __restrict__ h, const double* __restrict__ y0, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) b[i] = (1.0 - f[i]) * h[i] + d[i] * y0[i];
}

// t = sh - s x  (the right-hand side of the flux block at the solution)
__global__ void k_flux_rhs(int64_t N, const double* __restrict__ sh, const double* __restrict__ s,
                           const double* __restrict__ x, double* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) t[i] = sh[i] - s[i] * x[i];
}

// t = ((hOld - x) area) / dt: the eroded volume rate whose accumulation is the
// total sediment flux Qt of the non-linear residual (nlSPL.py:104-107)
__global__ void k_nled_qrhs(int64_t N, const double* __restrict__ hold, const double* __restrict__ x,
                            const double* __restrict__ area, double dt, double* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) t[i] = ((hold[i] - x[i]) * area[i]) * (1.0 / dt);
}

// One Newton linearisation of the non-linear transport-limited residual
// (_form_residual_ed, nlSPL.py:82-125) at x:
//   F = (x - hOld)(1 - f) + Kbr S^n - f dt max(Qt, 0) / area,
// S the weighted positive slope to the lower receivers with w > 0 (local_spl,
// functions.F90:1469-1502).  Writes the rows of the triangular part
// (dd = n Kbr S^(n-1); off -dd c_k, diagonal 1 - f + dd sum c), rhs = -F and
// the coupling d' = f dt / area where Qt > 0 (the derivative of the clipped
// flux): J = T + diag(d') F^-1 diag(area / dt).  The preconditioning sweep
// takes T without the -f of the coupled rows (fcoup = f there, 0 elsewhere),
// applied beside it: J = (T + diag(fcoup)) - diag(fcoup) + diag(d') F^-1 ...,
// as the stream-power rows precondition the linear solve (a sweep of 1 - f
// on a coupled row would divide by as little as 0.01).
__global__ void k_nled_jac(int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ wgt,
                           const double* __restrict__ dist, const double* __restrict__ x,
                           const double* __restrict__ hold, const double* __restrict__ kbr, double n,
                           const double* __restrict__ f, const double* __restrict__ qt,
                           const double* __restrict__ area, double dt, double* __restrict__ off,
                           double* __restrict__ diag, double* __restrict__ rhs, double* __restrict__ dprime,
                           double* __restrict__ fcoup) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double xi = x[i];
  double S = 0.0, c0 = 0.0;
  double ck[GOSPL_MAXW];
  for (int k = 0; k < NR; ++k) {
    ck[k] = 0.0;
    const double w = wgt[(int64_t)k * N + i];
    if (w > 0.0) {
      const double xr = x[rcv[(int64_t)k * N + i]];
      if (xi > xr) {
        const double d = dist[(int64_t)k * N + i];
        S = S + w * (xi - xr) / d;
        ck[k] = -w / d;
        c0 = c0 + w / d;
      }
    }
  }
  const double kb = kbr[i];
  const double fi = f[i];
  const double q = qt[i];
  const double qp = q < 0.0 ? 0.0 : q;
  double F = (xi - hold[i]) * (1.0 - fi);
  F = F + kb * (S > 0.0 ? pow(S, n) : 0.0);
  F = F - (fi * dt) * qp / area[i];
  const double dd = S > 0.0 ? n * kb * pow(S, n - 1.0) : 0.0;
  for (int k = 0; k < NR; ++k) off[(int64_t)k * N + i] = dd * ck[k];
  // where the flux couples (Qt > 0) the -f of J's diagonal is applied beside
  // the sweep; elsewhere the row is triangular and the sweep takes it whole
  const bool coupled = q > 0.0;
  diag[i] = (coupled ? 1.0 : 1.0 - fi) + dd * c0;
  rhs[i] = -F;
  dprime[i] = coupled ? fi * dt / area[i] : 0.0;
  fcoup[i] = coupled ? fi : 0.0;
}

// x += d; keys of max |d| and max |x| (one atomic per workgroup)
__global__ __launch_bounds__(GOSPL_BLOCK) void k_newton_step(int64_t N, const double* __restrict__ d,
                                                            double* __restrict__ x,
                                                            unsigned long long* __restrict__ keys) {
  __shared__ unsigned long long sh[2][GOSPL_BLOCK / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long a = 0ull, b = 0ull;
  if (i < N) {
    const double dv = d[i];
    const double xv = x[i] + dv;
    x[i] = xv;
    a = dkey(fabs(dv));
    b = dkey(fabs(xv));
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long ca = __shfl_xor(a, o), cb = __shfl_xor(b, o);
    a = ca > a ? ca : a;
    b = cb > b ? cb : b;
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
      a = sh[0][w] > a ? sh[0][w] : a;
      b = sh[1][w] > b ? sh[1][w] : b;
    }
    atomicMax(keys, a);
    atomicMax(keys + 1, b);
  }
}

// Marine deposition coefficients (seaplex.py:262-293): fDepm = min(Gmar,
// 0.99), 0 where h > sea level; d = fDepm dt / area, s = area / dt;
// r0 = (1 - fDepm) h, r1 = vol / dt + (h area) / dt (kept in b and t).
__global__ void k_mdep_prep(int64_t N, double Gm, double sl, double dt, const double* __restrict__ area,
                            const double* __restrict__ h, const double* __restrict__ vol, double* __restrict__ f,
                            double* __restrict__ d, double* __restrict__ s, double* __restrict__ r0,
                            double* __restrict__ r1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double fd = Gm > 0.99 ? 0.99 : Gm;
  const double hi = h[i];
  if (hi > sl) fd = 0.0;
  const double a = area[i];
  f[i] = fd;
  d[i] = fd * dt / a;
  s[i] = a / dt;
  r0[i] = (1.0 - fd) * hi;
  r1[i] = vol[i] / dt + (hi * a) * (1.0 / dt);
}

// b = r0 + d y (y = F^-1 r1), in place of r0
__global__ void k_mdep_rhs(int64_t N, const double* __restrict__ d, const double* __restrict__ y,
                           double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) b[i] = b[i] + d[i] * y[i];
}

// volDep = max((hnew - h) area, 0) (seaplex.py:394-399)
__global__ void k_mdep_vol(int64_t N, const double* __restrict__ hn, const double* __restrict__ h,
                           const double* __restrict__ area, double* __restrict__ vol) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double v = (hn[i] - h[i]) * area[i];
  vol[i] = v < 0.0 ? 0.0 : v;
}

struct Coupled {
  int nE;
  const int32_t* rcvE;  // nullptr: E = I
  const double* offE;
  const double* diagE;
  int nF;
  const int32_t* rcvF;
  const double* wgtF;
  const double* f;
  const double* d;
  const double* s;
};

struct Work {
  int64_t N, ld;
  double* V;  // KRY_M + 1 rows
  double* Z;  // KRY_M rows
  double* w;
  double* t;
  double* y;
  double* b;
  double* x;
  double* vec[6];
  double* part;  // DOT_BLOCKS * (KRY_M + 1) partial sums (device)
  double* dsum;  // KRY_M + 1 sums (device)
  double* dcoef; // KRY_M + 1 coefficients (device)
  double* hsum;  // pinned: sums read back
  double* hcoef; // pinned: coefficients sent
  unsigned long long* keys;  // two max keys (device)
};

int ensure_work(gospl_mesh* m, Work* wk) {
  const int64_t ld = m->N + 64;
  if (!m->kry_buf) GOSPL_CHECK(hipMalloc(&m->kry_buf, sizeof(double) * KRY_ROWS * ld));
  if (!m->kry_d) GOSPL_CHECK(hipMalloc(&m->kry_d, sizeof(double) * ((DOT_BLOCKS + 2) * (KRY_M + 2) + 4)));
  if (!m->kry_h) GOSPL_CHECK(hipHostMalloc((void**)&m->kry_h, sizeof(double) * 2 * (KRY_M + 2), hipHostMallocDefault));
  wk->N = m->N;
  wk->ld = ld;
  wk->V = m->kry_buf;
  wk->Z = m->kry_buf + (int64_t)(KRY_M + 1) * ld;
  wk->w = m->kry_buf + (int64_t)ROW_W * ld;
  wk->t = wk->w + ld;
  wk->y = wk->t + ld;
  wk->b = wk->y + ld;
  wk->x = wk->b + ld;
  for (int k = 0; k < 6; ++k) wk->vec[k] = wk->x + (int64_t)(k + 1) * ld;
  wk->part = m->kry_d;
  wk->dsum = m->kry_d + (int64_t)DOT_BLOCKS * (KRY_M + 2);
  wk->dcoef = wk->dsum + (KRY_M + 2);
  wk->hsum = m->kry_h;
  wk->hcoef = m->kry_h + (KRY_M + 2);
  wk->keys = reinterpret_cast<unsigned long long*>(wk->dcoef + (KRY_M + 2));
  return 0;
}

// out[j] = V_j . w for j < k (synchronises s)
int mdot(const Work& wk, const double* V, int k, const double* w, double* out, hipStream_t s) {
  k_mdot<<<dim3(DOT_BLOCKS, k), DOT_THREADS, 0, s>>>(wk.N, V, wk.ld, w, wk.part);
  k_msum<<<k, DOT_THREADS, 0, s>>>(wk.part, wk.dsum);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(wk.hsum, wk.dsum, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  std::memcpy(out, wk.hsum, sizeof(double) * k);
  return 0;
}

// out += sign * sum_j c[j] V_j (the coefficients go through the pinned block;
// every caller synchronises before that block is written again)
int lincomb(const Work& wk, const double* V, int k, const double* c, double sign, double* out, hipStream_t s) {
  std::memcpy(wk.hcoef, c, sizeof(double) * k);
  GOSPL_CHECK(hipMemcpyAsync(wk.dcoef, wk.hcoef, sizeof(double) * k, hipMemcpyHostToDevice, s));
  k_lincomb<<<nblocks(wk.N), GOSPL_BLOCK, 0, s>>>(wk.N, V, wk.ld, k, wk.dcoef, sign, out);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

// w = S z
int apply(gospl_mesh* m, const Coupled& c, const Work& wk, const double* z, double* w, hipStream_t s) {
  const int g = nblocks(wk.N);
  k_mul<<<g, GOSPL_BLOCK, 0, s>>>(wk.N, c.s, z, wk.t);
  GOSPL_CHECK(hipGetLastError());
  int64_t bad = 0;
  if (dag_solve_ext(m, 0, c.nF, c.rcvF, c.wgtF, nullptr, wk.t, wk.y, &bad, s)) return 1;
  if (bad) {
    set_error("coupled solve: the flow sweep left %lld vertices on an undrained receiver cycle", (long long)bad);
    return 1;
  }
  k_coupled_apply<<<g, GOSPL_BLOCK, 0, s>>>(wk.N, c.nE, c.rcvE, c.offE, c.diagE, c.f, c.d, z, wk.y, w);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

// z = E^-1 v
int precond(gospl_mesh* m, const Coupled& c, const Work& wk, const double* v, double* z, hipStream_t s) {
  if (!c.rcvE) {
    GOSPL_CHECK(hipMemcpyAsync(z, v, sizeof(double) * wk.N, hipMemcpyDeviceToDevice, s));
    return 0;
  }
  return dag_solve_ext(m, 1, c.nE, c.rcvE, c.offE, c.diagE, v, z, nullptr, s);
}

// GMRES(KRY_M) on S x = b, right-preconditioned by E^-1; x holds the initial
// guess.  Classical Gram-Schmidt applied twice; the true residual is taken
// at every restart.  Stops once ||b - S x||_2 <= rtol ||b||_2 or after maxit
// iterations (then *relres reports what was reached).
int gmres(gospl_mesh* m, const Coupled& c, const Work& wk, const double* b, double* x, double rtol, int maxit,
          int* iters, double* relres, hipStream_t s, bool zero_guess = false, double atol = 0.0,
          double* bnorm = nullptr) {
  const int64_t N = wk.N;
  double nb = 0.0;
  if (mdot(wk, b, 1, b, &nb, s)) return 1;
  nb = std::sqrt(nb);
  if (bnorm) *bnorm = nb;
  // stop at ||r|| <= tol_abs: rtol relative to ||b||, or the absolute floor atol
  const double tol_abs = rtol * nb > atol ? rtol * nb : atol;
  *iters = 0;
  *relres = 0.0;
  if (nb == 0.0) {
    GOSPL_CHECK(hipMemsetAsync(x, 0, sizeof(double) * N, s));
    return 0;
  }
  std::vector<double> H((KRY_M + 1) * KRY_M), cs(KRY_M), sn(KRY_M), g(KRY_M + 1), yv(KRY_M), hcol(KRY_M + 1),
      h2(KRY_M + 1);
  int it = 0;
  while (true) {
    // r = b - S x into V_0
    double* v0 = wk.V;
    if (zero_guess && it == 0) {
      GOSPL_CHECK(hipMemsetAsync(x, 0, sizeof(double) * N, s));
      GOSPL_CHECK(hipMemcpyAsync(v0, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    } else {
      if (apply(m, c, wk, x, wk.w, s)) return 1;
      k_sub<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, b, wk.w, v0);
      GOSPL_CHECK(hipGetLastError());
    }
    double beta = 0.0;
    if (mdot(wk, v0, 1, v0, &beta, s)) return 1;
    beta = std::sqrt(beta);
    *relres = beta / nb;
    if (!(beta > tol_abs) || it >= maxit) break;
    k_scale<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, v0, 1.0 / beta, v0);
    GOSPL_CHECK(hipGetLastError());
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < KRY_M && it < maxit; ++j) {
      double* vj = wk.V + (int64_t)j * wk.ld;
      double* zj = wk.Z + (int64_t)j * wk.ld;
      if (precond(m, c, wk, vj, zj, s)) return 1;
      if (apply(m, c, wk, zj, wk.w, s)) return 1;
      ++it;
      // two passes of classical Gram-Schmidt against V_0..V_j
      if (mdot(wk, wk.V, j + 1, wk.w, hcol.data(), s)) return 1;
      if (lincomb(wk, wk.V, j + 1, hcol.data(), -1.0, wk.w, s)) return 1;
      if (mdot(wk, wk.V, j + 1, wk.w, h2.data(), s)) return 1;
      if (lincomb(wk, wk.V, j + 1, h2.data(), -1.0, wk.w, s)) return 1;
      for (int i = 0; i <= j; ++i) hcol[i] += h2[i];
      double hn = 0.0;
      if (mdot(wk, wk.w, 1, wk.w, &hn, s)) return 1;
      hn = std::sqrt(hn);
      hcol[j + 1] = hn;
      // earlier rotations, then the one that zeroes hcol[j + 1]
      for (int i = 0; i < j; ++i) {
        const double t0 = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
        hcol[i] = t0;
      }
      const double rr = std::hypot(hcol[j], hcol[j + 1]);
      cs[j] = rr > 0.0 ? hcol[j] / rr : 1.0;
      sn[j] = rr > 0.0 ? hcol[j + 1] / rr : 0.0;
      hcol[j] = rr;
      hcol[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; ++i) H[i * KRY_M + j] = hcol[i];
      k = j + 1;
      if (!(std::fabs(g[j + 1]) > tol_abs) || hn == 0.0) break;
      k_scale<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, wk.w, 1.0 / hn, wk.V + (int64_t)(j + 1) * wk.ld);
      GOSPL_CHECK(hipGetLastError());
    }
    // y = H^-1 g (upper triangular), x += Z y
    for (int i = k - 1; i >= 0; --i) {
      double acc = g[i];
      for (int l = i + 1; l < k; ++l) acc -= H[i * KRY_M + l] * yv[l];
      yv[i] = H[i * KRY_M + i] != 0.0 ? acc / H[i * KRY_M + i] : 0.0;
    }
    if (lincomb(wk, wk.Z, k, yv.data(), 1.0, x, s)) return 1;
    GOSPL_CHECK(hipStreamSynchronize(s));  // the pinned coefficients are free again
  }
  *iters = it;
  return 0;
}

}  // namespace
}  // namespace gospl

using namespace gospl;

extern "C" {

int gospl_spl_dep(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti, const double* disti,
                  const double* fal, const double* kfac, double K, double mexp, double G, double dt,
                  double sealevel, const double* lfill, const int8_t* lsink, double* h, double* cumed, double* eb,
                  double* eblocal, double rtol, int32_t maxit, int32_t* iters, double* relres, double* qs,
                  gospl_stream_t stream) {
  if (!m || nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_spl_dep: null handle or nrcv outside 1..12");
    return 1;
  }
  if (!(G > 0.0) || !(dt > 0.0) || maxit < 1) {
    set_error("gospl_spl_dep: G and dt must be positive and maxit >= 1");
    return 1;
  }
  hipStream_t s = stream;
  const int64_t N = m->N;
  if (gospl_ensure_coef(m, nrcv)) return 1;
  double* off = m->wcoef;
  double* diag = m->wd[3];
  if (spl_coef_launch(m, nrcv, rcvi, wgti, disti, fal, kfac, K, mexp, dt, sealevel, lfill, h, off, diag, s))
    return 1;
  Work wk;
  if (ensure_work(m, &wk)) return 1;
  double *f = wk.vec[0], *d = wk.vec[1], *sv = wk.vec[2], *sh = wk.vec[3];
  const int g = nblocks(N);
  k_ed_prep<<<g, GOSPL_BLOCK, 0, s>>>(N, G, m->area, fal, sealevel, lfill, m->outlet, m->flat, dt, h, f, d, sv, sh);
  GOSPL_CHECK(hipGetLastError());
  Coupled c{nrcv, rcvi, off, diag, nrcv, rcvi, wgti, f, d, sv};
  // b = (1 - f) hOld + d F^-1 (hOld area / dt)
  int64_t bad = 0;
  if (dag_solve_ext(m, 0, nrcv, rcvi, wgti, nullptr, sh, wk.y, &bad, s)) return 1;
  if (bad) {
    set_error("gospl_spl_dep: the flow sweep left %lld vertices on an undrained receiver cycle", (long long)bad);
    return 1;
  }
  k_ed_rhs<<<g, GOSPL_BLOCK, 0, s>>>(N, f, d, h, wk.y, wk.b);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(wk.x, h, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  int it = 0;
  double rr = 0.0;
  if (gmres(m, c, wk, wk.b, wk.x, rtol, maxit, &it, &rr, s)) return 1;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  if (qs) {
    k_flux_rhs<<<g, GOSPL_BLOCK, 0, s>>>(N, sh, sv, wk.x, wk.t);
    GOSPL_CHECK(hipGetLastError());
    if (dag_solve_ext(m, 0, nrcv, rcvi, wgti, nullptr, wk.t, qs, &bad, s)) return 1;
  }
  return spl_update_launch(m, wk.x, lsink, dt, h, cumed, eb, eblocal, s);
}

int gospl_spl_nl_dep(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti, const double* disti,
                     const double* fal, const double* kfac, double K, double mexp, double nexp, double G, double dt,
                     double sealevel, const double* lfill, const int8_t* lsink, double* h, double* cumed,
                     double* eb, double* eblocal, double tol, int32_t maxit, int32_t* iters, int32_t* inner,
                     double* step_rel, gospl_stream_t stream) {
  if (!m || nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_spl_nl_dep: null handle or nrcv outside 1..12");
    return 1;
  }
  if (!(G > 0.0) || !(nexp > 0.0) || !(dt > 0.0) || maxit < 1) {
    set_error("gospl_spl_nl_dep: G, n and dt must be positive and maxit >= 1");
    return 1;
  }
  hipStream_t s = stream;
  const int64_t N = m->N;
  const int g = nblocks(N);
  if (gospl_ensure_coef(m, nrcv)) return 1;
  double* off = m->wcoef;
  double* diag = m->wd[3];
  Work wk;
  if (ensure_work(m, &wk)) return 1;
  double *f = wk.vec[0], *d = wk.vec[1], *sv = wk.vec[2], *sh = wk.vec[3], *kbr = wk.vec[4], *fcoup = wk.vec[5];
  double* rhs = wk.b;
  double* dprime = d;  // the coupling of each Newton step (d = f dt / area is not used here)
  double* qt = wk.y;   // overwritten by the GMRES sweeps after k_nled_jac has read it
  double* step = wk.t;
  double* xs = wk.x;   // the Newton iterate
  k_ed_prep<<<g, GOSPL_BLOCK, 0, s>>>(N, G, m->area, fal, sealevel, lfill, m->outlet, m->flat, dt, h, f, d, sv, sh);
  GOSPL_CHECK(hipGetLastError());
  if (nl_kbr_launch(m, nrcv, rcvi, h, fal, kfac, K, mexp, dt, sealevel, lfill, kbr, s)) return 1;
  GOSPL_CHECK(hipMemcpyAsync(xs, h, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  // the Newton step: GMRES's solution vector (sh, the linear solve's right-hand side, is not used here)
  double* dx = sh;
  int it = 0, tot = 0;
  double rel = INFINITY, fnorm0 = 0.0;
  while (it < maxit) {
    k_nled_qrhs<<<g, GOSPL_BLOCK, 0, s>>>(N, h, xs, m->area, dt, step);
    GOSPL_CHECK(hipGetLastError());
    int64_t bad = 0;
    if (dag_solve_ext(m, 0, nrcv, rcvi, wgti, nullptr, step, qt, &bad, s)) return 1;
    if (bad) {
      set_error("gospl_spl_nl_dep: the flow sweep left %lld vertices on an undrained receiver cycle", (long long)bad);
      return 1;
    }
    k_nled_jac<<<g, GOSPL_BLOCK, 0, s>>>(N, nrcv, rcvi, wgti, disti, xs, h, kbr, nexp, f, qt, m->area, dt, off, diag,
                                         rhs, dprime, fcoup);
    GOSPL_CHECK(hipGetLastError());
    Coupled c{nrcv, rcvi, off, diag, nrcv, rcvi, wgti, fcoup, dprime, sv};
    int git = 0;
    double grr = 0.0, fn = 0.0;
    // each linear solve to 1e-13 of its residual or of the first Newton
    // residual, whichever is larger (near the root -F is round-off)
    if (gmres(m, c, wk, rhs, dx, 1.0e-13, 400, &git, &grr, s, true, 1.0e-13 * fnorm0, &fn)) return 1;
    if (it == 0) fnorm0 = fn;
    tot += git;
    GOSPL_CHECK(hipMemsetAsync(wk.keys, 0, sizeof(unsigned long long) * 2, s));
    k_newton_step<<<g, GOSPL_BLOCK, 0, s>>>(N, dx, xs, wk.keys);
    GOSPL_CHECK(hipGetLastError());
    GOSPL_CHECK(hipMemcpyAsync(wk.hsum, wk.keys, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    unsigned long long hk[2];
    std::memcpy(hk, wk.hsum, sizeof(hk));
    ++it;
    const double dmax = dkey_inv_host(hk[0]), xmax = dkey_inv_host(hk[1]);
    rel = dmax / (xmax > 1.0 ? xmax : 1.0);
    if (!(rel == rel)) break;
    if (rel <= tol) break;
  }
  if (iters) *iters = it;
  if (inner) *inner = tot;
  if (step_rel) *step_rel = rel;
  if (!(rel <= tol)) {
    set_error("gospl_spl_nl_dep: Newton did not converge in %d steps (last step %.3g of max|h|)", it, rel);
    return 1;
  }
  return spl_update_launch(m, xs, lsink, dt, h, cumed, eb, eblocal, s);
}

int gospl_marine_dep(gospl_mesh_t* m, const int32_t* rcv, const double* wgt, const double* h, const double* vol,
                     double Gmar, double sealevel, double dt, double* voldep, double rtol, int32_t maxit,
                     int32_t* iters, double* relres, gospl_stream_t stream) {
  if (!m || !rcv || !wgt || !h || !vol || !voldep) {
    set_error("gospl_marine_dep: null argument");
    return 1;
  }
  if (!(Gmar > 0.0) || !(dt > 0.0) || maxit < 1) {
    set_error("gospl_marine_dep: Gmar and dt must be positive and maxit >= 1");
    return 1;
  }
  hipStream_t s = stream;
  const int64_t N = m->N;
  const int g = nblocks(N);
  Work wk;
  if (ensure_work(m, &wk)) return 1;
  double *f = wk.vec[0], *d = wk.vec[1], *sv = wk.vec[2], *r1 = wk.vec[3];
  k_mdep_prep<<<g, GOSPL_BLOCK, 0, s>>>(N, Gmar, sealevel, dt, m->area, h, vol, f, d, sv, wk.b, r1);
  GOSPL_CHECK(hipGetLastError());
  int64_t bad = 0;
  if (dag_solve_ext(m, 0, GOSPL_MAXW, rcv, wgt, nullptr, r1, wk.y, &bad, s)) return 1;
  if (bad) {
    set_error("gospl_marine_dep: the routing sweep left %lld vertices on an undrained receiver cycle",
              (long long)bad);
    return 1;
  }
  k_mdep_rhs<<<g, GOSPL_BLOCK, 0, s>>>(N, d, wk.y, wk.b);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(wk.x, h, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  Coupled c{0, nullptr, nullptr, nullptr, GOSPL_MAXW, rcv, wgt, f, d, sv};
  int it = 0;
  double rr = 0.0;
  if (gmres(m, c, wk, wk.b, wk.x, rtol, maxit, &it, &rr, s)) return 1;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  k_mdep_vol<<<g, GOSPL_BLOCK, 0, s>>>(N, wk.x, h, m->area, voldep);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
