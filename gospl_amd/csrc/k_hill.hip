// Linear hillslope diffusion (gospl/sed/hillslope.py:171-233, 614-649).
//
// Operator rows from `sethillslopecoeff` (functions.F90:901-938):
// v_kn = (0.5 * (Kd_k + Kd_n) / A_k) * vDist_kn / eLgt_kn, diag = 1 + sum v,
// off = -v; flat-mesh outlets are identity rows (hillslope.py:253-255).
// M h' = h is solved by Jacobi sweeps: M is strictly diagonally dominant with
// contraction factor max_k sum v / (1 + sum v) (~1e-5 at the benchmark
// parameters), so a handful of sweeps reach a fixed point in FP64.  Row k
// depends on the elevation only through the sea flags (h <= sealevel) of k and
// its neighbours, so the rows are kept in the handle with the flags they were
// built from, and a call rebuilds only the rows where one of those flags
// changed (goSPL keeps its operator until the coastline moves,
// hillslope.py:184-200); new diffusivities or dt rebuild every row.
#include <cmath>
#include <cstring>
#include <vector>

#include <utility>

#include "common.h"

namespace gospl {

// One operator row; the sea flags (is_sea(v)) decide each vertex's
// diffusivity.  The FP32 copies are written when given.
template <class Sea>
__device__ __forceinline__ void hill_row(int64_t N, int W, int64_t i, const int32_t* __restrict__ nbr,
                                         const int32_t* __restrict__ deg, const double* __restrict__ area,
                                         const double* __restrict__ elgt, const double* __restrict__ vdist,
                                         const int8_t* __restrict__ outlet, int flat, const Sea& is_sea,
                                         double cda, double cdm, double dt, double* __restrict__ coef,
                                         double* __restrict__ diag, float* __restrict__ coef32,
                                         float* __restrict__ diag32) {
  const double a = area[i];
  const bool ident = flat && outlet[i];
  const double ck = (is_sea(i) ? cdm : cda) * dt;
  double s1 = 0.0;
  const int dg = deg[i];
  for (int p = 0; p < W; ++p) {
    double c_out = 0.0;
    if (!ident && a > 0.0 && p < dg) {
      const double vd = vdist[(int64_t)p * N + i];
      if (vd > 0.0) {
        const int n = nbr[(int64_t)p * N + i];
        const double cn = (is_sea(n) ? cdm : cda) * dt;
        const double c = 0.5 * (ck + cn) / a;
        const double v = c * vd / elgt[(int64_t)p * N + i];
        s1 = s1 + v;
        c_out = -v;
      }
    }
    coef[(int64_t)p * N + i] = c_out;
    if (coef32) coef32[(int64_t)p * N + i] = (float)c_out;
  }
  const double d = ident ? 1.0 : (a > 0.0 ? 1.0 + s1 : 0.0);
  diag[i] = d;
  if (diag32) diag32[i] = (float)d;
}

struct SeaByLevel {
  const double* h;
  double sl;
  __device__ __forceinline__ bool operator()(int64_t v) const { return h[v] <= sl; }
};
struct SeaByFlag {
  const int8_t* f;
  __device__ __forceinline__ bool operator()(int64_t v) const { return f[v] != 0; }
};

// Every row, and the sea flags they were built from.
__global__ void k_hill_coef(int64_t N, int W, const int32_t* __restrict__ nbr,
                            const int32_t* __restrict__ deg, const double* __restrict__ area,
                            const double* __restrict__ elgt, const double* __restrict__ vdist,
                            const int8_t* __restrict__ outlet, int flat, const double* __restrict__ h,
                            double sl, double cda, double cdm, double dt, double* __restrict__ coef,
                            double* __restrict__ diag, float* __restrict__ coef32, float* __restrict__ diag32,
                            int8_t* __restrict__ sea) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  hill_row(N, W, i, nbr, deg, area, elgt, vdist, outlet, flat, SeaByLevel{h, sl}, cda, cdm, dt, coef, diag, coef32,
           diag32);
  sea[i] = h[i] <= sl;
}

// The sea flags of the current elevation against the kept ones: chg = 1
// where a vertex changed side; the kept flags are updated.
__global__ void k_hill_sea(int64_t N, const double* __restrict__ h, double sl, int8_t* __restrict__ sea,
                           int8_t* __restrict__ chg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int8_t f = h[i] <= sl;
  chg[i] = f != sea[i];
  sea[i] = f;
}

// The rows whose vertex or a neighbour changed side, rebuilt.
__global__ void k_hill_coef_changed(int64_t N, int W, const int32_t* __restrict__ nbr,
                                    const int32_t* __restrict__ deg, const double* __restrict__ area,
                                    const double* __restrict__ elgt, const double* __restrict__ vdist,
                                    const int8_t* __restrict__ outlet, int flat, const double* __restrict__ h,
                                    double sl, double cda, double cdm, double dt,
                                    const int8_t* __restrict__ chg, double* __restrict__ coef,
                                    double* __restrict__ diag, float* __restrict__ coef32,
                                    float* __restrict__ diag32) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  bool need = chg[i] != 0;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg && !need; ++p) {
    const int n = nbr[(int64_t)p * N + i];
    if (n >= 0 && chg[n]) need = true;
  }
  if (need)
    hill_row(N, W, i, nbr, deg, area, elgt, vdist, outlet, flat, SeaByLevel{h, sl}, cda, cdm, dt, coef, diag, coef32,
             diag32);
}

__global__ void k_hill_sweep(int64_t N, int W, const int32_t* __restrict__ nbr,
                             const int32_t* __restrict__ deg, const double* __restrict__ coef,
                             const double* __restrict__ diag, const double* __restrict__ rhs,
                             const double* __restrict__ xin, double* __restrict__ xout) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double d = diag[i];
  const double r = rhs[i];
  if (d == 0.0) { xout[i] = r; return; }
  double s = r;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    const double c = coef[(int64_t)p * N + i];
    if (c != 0.0) s = s - c * xin[nbr[(int64_t)p * N + i]];
  }
  xout[i] = s / d;
}

// One step of the Chebyshev semi-iteration over that sweep:
// xout = omega (sweep(xcur) - xprev) + xprev, the sweep in k_hill_sweep's
// operation order.
__global__ void k_hill_cheb(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                            const double* __restrict__ coef, const double* __restrict__ diag,
                            const double* __restrict__ rhs, const double* __restrict__ xcur,
                            const double* __restrict__ xprev, double omega, double* __restrict__ xout) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double d = diag[i];
  const double r = rhs[i];
  double v = r;
  if (d != 0.0) {
    double s = r;
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      const double c = coef[(int64_t)p * N + i];
      if (c != 0.0) s = s - c * xcur[nbr[(int64_t)p * N + i]];
    }
    v = s / d;
  }
  const double xp = xprev[i];
  xout[i] = omega * (v - xp) + xp;
}

// Two maxima (order-preserving keys; a and b valid in thread 0) of every
// workgroup, as per-workgroup partials; k_max2_finish folds them into
// keys[0..1] in the next launch.  (A last-workgroup reduction through one
// completion counter took ~25 ns per workgroup on the counter's same-address
// atomic, ~250 us over 2.6M rows.)
__device__ __forceinline__ void max2_part(unsigned long long a, unsigned long long b, unsigned long long* part) {
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

__global__ __launch_bounds__(1024) void k_max2_finish(int nblk, const unsigned long long* __restrict__ part,
                                                      unsigned long long* __restrict__ keys) {
  __shared__ unsigned long long sm[2][1024 / 64];
  unsigned long long ma = 0ull, mb = 0ull;
  for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
    const unsigned long long x = part[2 * k], y = part[2 * k + 1];
    ma = x > ma ? x : ma;
    mb = y > mb ? y : mb;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(ma, o), y = __shfl_xor(mb, o);
    ma = x > ma ? x : ma;
    mb = y > mb ? y : mb;
  }
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = ma;
    sm[1][threadIdx.x >> 6] = mb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      ma = sm[0][w] > ma ? sm[0][w] : ma;
      mb = sm[1][w] > mb ? sm[1][w] : mb;
    }
    keys[0] = ma > keys[0] ? ma : keys[0];
    keys[1] = mb > keys[1] ? mb : keys[1];
  }
}

// The sweep of k_hill_sweep that also records max |xout - xin| and max
// |xout| (order-preserving keys, reduced through per-workgroup partials):
// the convergence test of the iterate-to-convergence mode.
__global__ __launch_bounds__(GOSPL_BLOCK) void k_hill_sweep_chk(int64_t N, int W, const int32_t* __restrict__ nbr,
                                 const int32_t* __restrict__ deg, const double* __restrict__ coef,
                                 const double* __restrict__ diag, const double* __restrict__ rhs,
                                 const double* __restrict__ xin, double* __restrict__ xout,
                                 unsigned long long* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double chg = 0.0, mag = 0.0;
  if (i < N) {
    const double d = diag[i];
    const double r = rhs[i];
    double x = r;
    if (d != 0.0) {
      double s = r;
      const int dg = deg[i];
      for (int p = 0; p < W && p < dg; ++p) {
        const double c = coef[(int64_t)p * N + i];
        if (c != 0.0) s = s - c * xin[nbr[(int64_t)p * N + i]];
      }
      x = s / d;
    }
    xout[i] = x;
    chg = fabs(x - xin[i]);
    mag = fabs(x);
  }
  for (int o = 32; o > 0; o >>= 1) {
    chg = fmax(chg, __shfl_xor(chg, o, 64));
    mag = fmax(mag, __shfl_xor(mag, o, 64));
  }
  __shared__ double sh[2][GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = chg;
    sh[1][threadIdx.x >> 6] = mag;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
      chg = fmax(chg, sh[0][w]);
      mag = fmax(mag, sh[1][w]);
    }
  }
  max2_part(dkey(chg), dkey(mag), part);
}

// max over the rows of the Jacobi contraction (diag - 1) / diag = s / (1 + s)
// into key[0] (left as it is when no row contracts: 0 after the caller's clear)
__global__ __launch_bounds__(GOSPL_BLOCK) void k_hill_rho(int64_t N, const double* __restrict__ diag,
                                                         unsigned long long* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (i < N) {
    const double d = diag[i];
    if (d > 0.0) r = (d - 1.0) / d;
  }
  for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
  __shared__ double sh[GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) r = fmax(r, sh[w]);
  // a key of 0 for a workgroup without a contracting row: the key keeps the
  // caller's 0 (the second key is only compared with 0 and kept)
  max2_part(r > 0.0 ? dkey(r) : 0ull, 0ull, part);
}

// Rows from given sea flags (the frozen operator of the adjoint and its
// forward replay), in buffers of their own.
__global__ void k_hill_rows_flags(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                                  const double* __restrict__ area, const double* __restrict__ elgt,
                                  const double* __restrict__ vdist, const int8_t* __restrict__ outlet, int flat,
                                  const int8_t* __restrict__ sea, double cda, double cdm, double dt,
                                  double* __restrict__ coef, double* __restrict__ diag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  hill_row(N, W, i, nbr, deg, area, elgt, vdist, outlet, flat, SeaByFlag{sea}, cda, cdm, dt, coef, diag,
           (float*)nullptr, (float*)nullptr);
}

// One Jacobi sweep of the transposed operator M^T y = rhs: the coefficient
// of y_j in row i of M^T is row j's coefficient at i, found through the
// reverse neighbour positions (rev[q * N + i] = position of i in the list of
// its neighbour q).  Rows with a zero diagonal are identity rows of M.
__global__ void k_hill_sweepT(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                              const int8_t* __restrict__ rev, const double* __restrict__ coef,
                              const double* __restrict__ diag, const double* __restrict__ rhs,
                              const double* __restrict__ xin, double* __restrict__ xout) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = rhs[i];
  const int dg = deg[i];
  for (int q = 0; q < W && q < dg; ++q) {
    const int j = nbr[(int64_t)q * N + i];
    const int p = rev[(int64_t)q * N + i];
    if (j < 0 || p < 0) continue;
    const double c = coef[(int64_t)p * N + j];
    if (c != 0.0) s = s - c * xin[j];
  }
  const double d = diag[i];
  xout[i] = d != 0.0 ? s / d : s;
}

// Refined solve (FP32 corrections of an FP64 residual).  Control words in
// device memory: [0] done, [1] passes, [2] relative residual, [3] its first value.
struct HillCtl {
  double done;
  double pass;
  double relres;
  double relres0;
};

// r = h - M x in FP64 (rows with a zero diagonal hold x = h), stored as FP32
// for the correction sweeps; per-workgroup sum r^2 and sum h^2.
__global__ void k_hill_resid(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                             const double* __restrict__ coef, const double* __restrict__ diag,
                             const double* __restrict__ h, const double* __restrict__ x, float* __restrict__ r32,
                             double* __restrict__ part, const HillCtl* __restrict__ ctl) {
  if (ctl->done != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, hh = 0.0;
  if (i < N) {
    const double d = diag[i];
    const double hi = h[i];
    double r = hi - x[i];
    if (d != 0.0) {
      double s = d * x[i];
      const int dg = deg[i];
      for (int p = 0; p < W && p < dg; ++p) {
        const double c = coef[(int64_t)p * N + i];
        if (c != 0.0) s = s + c * x[nbr[(int64_t)p * N + i]];
      }
      r = hi - s;
    }
    r32[i] = (float)r;
    rr = r * r;
    hh = hi * hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    rr += __shfl_xor(rr, o, 64);
    hh += __shfl_xor(hh, o, 64);
  }
  __shared__ double sh[2][GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = hh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0][0], b = sh[1][0];
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
      a += sh[0][w];
      b += sh[1][w];
    }
    part[2 * (int64_t)blockIdx.x] = a;
    part[2 * (int64_t)blockIdx.x + 1] = b;
  }
}

// Relative residual (one workgroup, fixed order) and the stop test.
__global__ void k_hill_norm(int nblk, const double* __restrict__ part, double rtol, int pass, int maxpass,
                            HillCtl* __restrict__ ctl) {
  if (ctl->done != 0.0) return;
  double rr = 0.0, hh = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    rr += part[2 * (int64_t)b];
    hh += part[2 * (int64_t)b + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    rr += __shfl_xor(rr, o, 64);
    hh += __shfl_xor(hh, o, 64);
  }
  __shared__ double sh[2][GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = hh;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
    sh[0][0] += sh[0][w];
    sh[1][0] += sh[1][w];
  }
  rr = sh[0][0];
  hh = sh[1][0];
  const double rel = hh > 0.0 ? sqrt(rr / hh) : 0.0;
  ctl->relres = rel;
  if (pass == 0) ctl->relres0 = rel;
  if (!(rel > rtol) || pass >= maxpass) ctl->done = 1.0;  // a NaN residual stops too
}

// One FP32 Jacobi sweep of M d = r (first: from d = 0).
__global__ void k_hill_sweep32(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                               const float* __restrict__ coef32, const float* __restrict__ diag32,
                               const float* __restrict__ r32, const float* __restrict__ din, int first,
                               float* __restrict__ dout, const HillCtl* __restrict__ ctl) {
  if (ctl->done != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float d = diag32[i];
  const float r = r32[i];
  if (d == 0.0f) {
    dout[i] = r;
    return;
  }
  float s = r;
  if (!first) {
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      const float c = coef32[(int64_t)p * N + i];
      if (c != 0.0f) s = s - c * din[nbr[(int64_t)p * N + i]];
    }
  }
  dout[i] = s / d;
}

// Segmented refined solve (ensembles: one union mesh of independent
// members, each segment a run of whole workgroups).  Every segment keeps
// its own control words, residual norm (summed over its own workgroups in
// the order of the single-mesh solve), sweep count and stop test, so each
// member's result equals its own solve's bit for bit.  aux (int32):
// [0, nseg) first workgroup, [nseg, 2 nseg) workgroup count, [2 nseg,
// 3 nseg) sweeps, [3 nseg, 4 nseg) first vertex.  Padding vertices
// (segment < 0) keep x = h.
__global__ void k_hill_resid_seg(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                                 const double* __restrict__ coef, const double* __restrict__ diag,
                                 const double* __restrict__ h, const double* __restrict__ x, float* __restrict__ r32,
                                 double* __restrict__ part, const int32_t* __restrict__ seg,
                                 const HillCtl* __restrict__ ctl) {
  const int64_t b0 = (int64_t)blockIdx.x * blockDim.x;
  const int bs = b0 < N ? seg[b0] : -1;  // a workgroup lies in one segment and starts on a real vertex
  if (bs < 0 || ctl[bs].done != 0.0) return;
  const int64_t i = b0 + threadIdx.x;
  double rr = 0.0, hh = 0.0;
  if (i < N && seg[i] >= 0) {
    const double d = diag[i];
    const double hi = h[i];
    double r = hi - x[i];
    if (d != 0.0) {
      double s = d * x[i];
      const int dg = deg[i];
      for (int p = 0; p < W && p < dg; ++p) {
        const double c = coef[(int64_t)p * N + i];
        if (c != 0.0) s = s + c * x[nbr[(int64_t)p * N + i]];
      }
      r = hi - s;
    }
    r32[i] = (float)r;
    rr = r * r;
    hh = hi * hi;
  } else if (i < N) {
    r32[i] = 0.0f;
  }
  for (int o = 32; o > 0; o >>= 1) {
    rr += __shfl_xor(rr, o, 64);
    hh += __shfl_xor(hh, o, 64);
  }
  __shared__ double sh[2][GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = hh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0][0], b = sh[1][0];
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
      a += sh[0][w];
      b += sh[1][w];
    }
    part[2 * (int64_t)blockIdx.x] = a;
    part[2 * (int64_t)blockIdx.x + 1] = b;
  }
}

// One workgroup per segment; the partials of the segment's workgroups are
// summed in k_hill_norm's order.
__global__ void k_hill_norm_seg(int nseg, const double* __restrict__ part, const int32_t* __restrict__ aux,
                                double rtol, int pass, int maxpass, HillCtl* __restrict__ ctl) {
  const int sg = blockIdx.x;
  if (sg >= nseg || ctl[sg].done != 0.0) return;
  const int64_t blk0 = aux[sg];
  const int nblk = aux[nseg + sg];
  double rr = 0.0, hh = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    rr += part[2 * (blk0 + b)];
    hh += part[2 * (blk0 + b) + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    rr += __shfl_xor(rr, o, 64);
    hh += __shfl_xor(hh, o, 64);
  }
  __shared__ double sh[2][GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = hh;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
    sh[0][0] += sh[0][w];
    sh[1][0] += sh[1][w];
  }
  rr = sh[0][0];
  hh = sh[1][0];
  const double rel = hh > 0.0 ? sqrt(rr / hh) : 0.0;
  ctl[sg].relres = rel;
  if (pass == 0) ctl[sg].relres0 = rel;
  if (!(rel > rtol) || pass >= maxpass) ctl[sg].done = 1.0;
}

// Sweep k of the segmented correction: segments whose count is reached copy
// their values through.
__global__ void k_hill_sweep32_seg(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                                   const float* __restrict__ coef32, const float* __restrict__ diag32,
                                   const float* __restrict__ r32, const float* __restrict__ din, int k,
                                   float* __restrict__ dout, const int32_t* __restrict__ seg, int nseg,
                                   const int32_t* __restrict__ aux, const HillCtl* __restrict__ ctl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int sg = seg[i];
  if (sg < 0) {
    dout[i] = 0.0f;
    return;
  }
  if (ctl[sg].done != 0.0) return;
  if (k >= aux[2 * nseg + sg]) {
    dout[i] = din[i];
    return;
  }
  const float d = diag32[i];
  const float r = r32[i];
  if (d == 0.0f) {
    dout[i] = r;
    return;
  }
  float s = r;
  if (k != 0) {
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      const float c = coef32[(int64_t)p * N + i];
      if (c != 0.0f) s = s - c * din[nbr[(int64_t)p * N + i]];
    }
  }
  dout[i] = s / d;
}

__global__ void k_hill_addcorr_seg(int64_t N, const float* __restrict__ d, double* __restrict__ x,
                                   const int32_t* __restrict__ seg, int nseg, const int32_t* __restrict__ aux,
                                   HillCtl* __restrict__ ctl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int sg = seg[i];
  if (sg < 0 || ctl[sg].done != 0.0) return;
  if (i == aux[3 * nseg + sg]) ctl[sg].pass = ctl[sg].pass + 1.0;
  x[i] = x[i] + (double)d[i];
}

// x += d (FP64 iterate) and count the pass.
__global__ void k_hill_addcorr(int64_t N, const float* __restrict__ d, double* __restrict__ x,
                               HillCtl* __restrict__ ctl) {
  if (ctl->done != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) ctl->pass = ctl->pass + 1.0;
  if (i >= N) return;
  x[i] = x[i] + (double)d[i];
}

__global__ void k_hill_update(int64_t N, const double* __restrict__ x, double dt, double* __restrict__ h,
                              double* __restrict__ cumed, double* __restrict__ eblocal) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double t = x[i] - h[i];
  h[i] = x[i];
  cumed[i] = cumed[i] + t;
  eblocal[i] = eblocal[i] + t / dt;
}

// sethillslopecoeff (functions.F90:901-938) for a per-vertex Kd: out is
// (13, N), row 0 the diagonal and row 1 + p the coefficient of neighbour p.
__global__ void k_hill_coef_kd(int64_t N, int W, const int32_t* __restrict__ nbr,
                               const int32_t* __restrict__ deg, const double* __restrict__ area,
                               const double* __restrict__ elgt, const double* __restrict__ vdist,
                               const double* __restrict__ kd, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double a = area[i];
  const double ck = kd[i];
  double s1 = 0.0;
  const int dg = deg[i];
  for (int p = 0; p < GOSPL_MAXW; ++p) {
    double c_out = 0.0;
    if (a > 0.0 && p < dg && p < W) {
      const double vd = vdist[(int64_t)p * N + i];
      if (vd > 0.0) {
        const int n = nbr[(int64_t)p * N + i];
        const double c = 0.5 * (ck + kd[n]) / a;
        const double v = c * vd / elgt[(int64_t)p * N + i];
        s1 = s1 + v;
        c_out = -v;
      }
    }
    out[(int64_t)(p + 1) * N + i] = c_out;
  }
  out[i] = a > 0.0 ? 1.0 + s1 : 0.0;
}

// flat-mesh outlets: identity rows (hillslope.py:253-255)
__global__ void k_hill_outlet_rows(int64_t N, const int8_t* __restrict__ outlet, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !outlet[i]) return;
  out[i] = 1.0;
  for (int p = 1; p <= GOSPL_MAXW; ++p) out[(int64_t)p * N + i] = 0.0;
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_diffuse_fixed(gospl_mesh_t* m, const double* kd, const double* rhs, double* x, int32_t maxit,
                                   int32_t* sweeps, gospl_stream_t s) {
  using namespace gospl;
  if (!m || maxit < 1) {
    set_error("gospl_diffuse_fixed: null handle or maxit < 1");
    return 1;
  }
  const int64_t N = m->N;
  const int g = nblocks(N);
  if (gospl_ensure_coef(m, 13)) return 1;
  double* op = m->wcoef;  // (13, N): row 0 the diagonal, row 1 + p neighbour p
  k_hill_coef_kd<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->area, m->elgt, m->vdist, kd, op);
  if (m->flat) k_hill_outlet_rows<<<g, GOSPL_BLOCK, 0, s>>>(N, m->outlet, op);
  GOSPL_CHECK(hipGetLastError());
  // Chebyshev semi-iteration over the Jacobi sweep from x0 = rhs.  rho = max
  // over rows of (d - 1) / d bounds the spectrum of the sweep's iteration
  // matrix, which is real (the operator is symmetric up to the row scaling
  // by the cell areas).  The count is the least k with 2 sigma^k <= 2^-49
  // (1 - rho), sigma = rho / (1 + sqrt(1 - rho^2)): the error bound in the
  // norm that symmetrises the sweep, with 2^-16 of room for that norm against
  // the max norm, under the 2^-33 of the plain sweeps' bound rho^k / (1 -
  // rho) (goSPL: Richardson to rtol 1e-8).  Weights omega_1 = 1 / (1 - rho^2
  // / 2), omega_k = 1 / (1 - rho^2 omega_{k-1} / 4).
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(m->dflag + 48);
  unsigned long long* hk = reinterpret_cast<unsigned long long*>(m->hflag + 48);
  GOSPL_CHECK(hipMemsetAsync(keys + 2, 0, sizeof(unsigned long long), s));
  k_hill_rho<<<g, GOSPL_BLOCK, 0, s>>>(N, op, m->red_part);
  k_max2_finish<<<1, 1024, 0, s>>>(g, m->red_part, keys + 2);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(hk + 2, keys + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  const double rho = hk[2] ? dkey_inv_host(hk[2]) : 0.0;
  if (!(rho < 1.0)) {
    set_error("gospl_diffuse_fixed: the Jacobi sweep does not contract (rho %.6g)", rho);
    return 1;
  }
  const double r2 = rho * rho;
  const double sig = rho / (1.0 + std::sqrt(1.0 - r2));
  const double tgt = 0x1p-49 * (1.0 - rho);
  double pw = 2.0;
  int nsw = 0;
  while (pw > tgt && nsw < maxit) {
    pw *= sig;
    ++nsw;
  }
  if (pw > tgt) {
    set_error("gospl_diffuse_fixed: %d sweeps do not reach the error bound (contraction %.6g)", maxit, rho);
    return 1;
  }
  // x1 = sweep(rhs); then three buffers in turn (rhs is never written)
  double* buf[3] = {m->wd[1], m->wd[2], x};
  const double* prev = rhs;
  int cur = 0;
  k_hill_sweep<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, op + N, op, rhs, rhs, buf[0]);
  double omega = 1.0;
  int prev_i = -1;
  for (int k = 1; k < nsw; ++k) {
    omega = k == 1 ? 1.0 / (1.0 - 0.5 * r2) : 1.0 / (1.0 - 0.25 * r2 * omega);
    int nxt = 0;
    while (nxt == cur || nxt == prev_i) ++nxt;
    k_hill_cheb<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, op + N, op, rhs, buf[cur], prev, omega,
                                          buf[nxt]);
    prev = buf[cur];
    prev_i = cur;
    cur = nxt;
  }
  GOSPL_CHECK(hipGetLastError());
  if (buf[cur] != x) GOSPL_CHECK(hipMemcpyAsync(x, buf[cur], sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  if (sweeps) *sweeps = nsw;
  return 0;
}

extern "C" int gospl_hillslope_coeff(gospl_mesh_t* m, const double* kd, double* coeff,
                                     gospl_stream_t s) {
  k_hill_coef_kd<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, m->W, m->nbr, m->deg, m->area, m->elgt,
                                                       m->vdist, kd, coeff);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

// The operator rows kept in the handle, brought up to date for h: all rows
// when the parameters changed, else the rows next to a vertex that changed
// side of sea level.
static int hill_operator(gospl_mesh* m, double cda, double cdm, double dt, double sealevel, const double* h,
                         hipStream_t s) {
  const int64_t N = m->N;
  const int g = nblocks(N);
  if (!m->hcoef) {
    GOSPL_CHECK(hipMalloc(&m->hcoef, sizeof(double) * (int64_t)m->W * (N + 64)));
    GOSPL_CHECK(hipMalloc(&m->hdiag, sizeof(double) * (N + 64)));
    GOSPL_CHECK(hipMalloc(&m->hcoef32, sizeof(float) * (int64_t)m->W * (N + 64)));
    GOSPL_CHECK(hipMalloc(&m->hdiag32, sizeof(float) * (N + 64)));
    GOSPL_CHECK(hipMalloc(&m->hsea, N + 64));
    m->hvalid = 0;
  }
  if (!m->hvalid || m->hpar[0] != cda || m->hpar[1] != cdm || m->hpar[2] != dt) {
    k_hill_coef<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->area, m->elgt, m->vdist, m->outlet,
                                           m->flat, h, sealevel, cda, cdm, dt, m->hcoef, m->hdiag, m->hcoef32,
                                           m->hdiag32, m->hsea);
    GOSPL_CHECK(hipGetLastError());
    m->hpar[0] = cda;
    m->hpar[1] = cdm;
    m->hpar[2] = dt;
    m->hvalid = 1;
  } else {
    int8_t* chg = reinterpret_cast<int8_t*>(m->wi[0]);
    k_hill_sea<<<g, GOSPL_BLOCK, 0, s>>>(N, h, sealevel, m->hsea, chg);
    k_hill_coef_changed<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->area, m->elgt, m->vdist,
                                                   m->outlet, m->flat, h, sealevel, cda, cdm, dt, chg, m->hcoef,
                                                   m->hdiag, m->hcoef32, m->hdiag32);
    GOSPL_CHECK(hipGetLastError());
  }
  return 0;
}

extern "C" int gospl_hillslope(gospl_mesh_t* m, double cda, double cdm, double dt, double sealevel,
                               int32_t sweeps, double* h, double* cumed, double* eblocal,
                               gospl_stream_t s) {

  const int64_t N = m->N;
  const int g = nblocks(N);
  if (hill_operator(m, cda, cdm, dt, sealevel, h, s)) return 1;
  double* coef = m->hcoef;
  double* diag = m->hdiag;
  double* xa = m->wd[1];
  double* xb = m->wd[2];
  const double* xin = h;
  double* xout = xa;
  if (sweeps > 0) {
    // the count the caller derived from the contraction bound
    for (int k = 0; k < sweeps; ++k) {
      k_hill_sweep<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, coef, diag, h, xin, xout);
      xin = xout;
      xout = (xout == xa) ? xb : xa;
    }
  } else {
    // iterate to convergence: with contraction rho the error after a sweep
    // is at most rho / (1 - rho) times its change, so stop once that bound
    // is below 2^-52 of the largest value; chunks of 32 sweeps, the last one
    // measuring (one read-back per chunk)
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(m->dflag + 48);
    GOSPL_CHECK(hipMemsetAsync(keys, 0, sizeof(unsigned long long) * 3, s));
    k_hill_rho<<<g, GOSPL_BLOCK, 0, s>>>(N, diag, m->red_part);
    k_max2_finish<<<1, 1024, 0, s>>>(g, m->red_part, keys + 2);
    GOSPL_CHECK(hipGetLastError());
    unsigned long long* hk = reinterpret_cast<unsigned long long*>(m->hflag + 48);
    GOSPL_CHECK(hipMemcpyAsync(hk + 2, keys + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    const double rho = hk[2] ? dkey_inv_host(hk[2]) : 0.0;
    const double factor = rho < 1.0 ? rho / (1.0 - rho) : 1.0e300;
    int64_t done = 0;
    bool conv = false;
    while (!conv && done < GOSPL_HILL_MAX_SWEEPS) {
      for (int k = 0; k < 31; ++k) {
        k_hill_sweep<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, coef, diag, h, xin, xout);
        xin = xout;
        xout = (xout == xa) ? xb : xa;
      }
      GOSPL_CHECK(hipMemsetAsync(keys, 0, sizeof(unsigned long long) * 2, s));
      k_hill_sweep_chk<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, coef, diag, h, xin, xout, m->red_part);
      k_max2_finish<<<1, 1024, 0, s>>>(g, m->red_part, keys);
      GOSPL_CHECK(hipGetLastError());
      xin = xout;
      xout = (xout == xa) ? xb : xa;
      done += 32;
      GOSPL_CHECK(hipMemcpyAsync(hk, keys, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, s));
      GOSPL_CHECK(hipStreamSynchronize(s));
      const double chg = dkey_inv_host(hk[0]), mag = dkey_inv_host(hk[1]);
      if (!(chg == chg) || !(mag == mag)) break;  // NaN: not converging
      conv = chg * factor <= 0x1p-52 * mag;
    }
    if (!conv) {
      set_error("gospl_hillslope: Jacobi sweeps did not converge within %lld sweeps (contraction %.6g)",
                (long long)GOSPL_HILL_MAX_SWEEPS, rho);
      return 1;
    }
  }
  k_hill_update<<<g, GOSPL_BLOCK, 0, s>>>(N, xin, dt, h, cumed, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

extern "C" int gospl_hillslope_refine(gospl_mesh_t* m, double cda, double cdm, double dt, double sealevel,
                                      int32_t sweeps32, double rtol, int32_t maxpass, double* h, double* cumed,
                                      double* eblocal, int32_t* passes, double* relres, gospl_stream_t s) {
  if (sweeps32 <= 0) {
    // too stiff for a fixed count of FP32 sweeps: the converged FP64 solve
    if (passes) *passes = -1;
    if (relres) *relres = 0.0;
    return gospl_hillslope(m, cda, cdm, dt, sealevel, 0, h, cumed, eblocal, s);
  }
  if (maxpass < 1) maxpass = 1;
  const int64_t N = m->N;
  const int g = nblocks(N);
  if (hill_operator(m, cda, cdm, dt, sealevel, h, s)) return 1;
  if (!m->rx_ctl) {
    GOSPL_CHECK(hipMalloc(&m->rx_ctl, sizeof(double) * 64));
    GOSPL_CHECK(hipHostMalloc((void**)&m->rx_hctl, sizeof(double) * 64, hipHostMallocDefault));
  }
  HillCtl* ctl = reinterpret_cast<HillCtl*>(m->rx_ctl + 32);
  double* x = m->wd[1];
  double* part = m->wd[0];  // 2 per workgroup <= N + 64
  float* r32 = reinterpret_cast<float*>(m->wi[5]);
  float* da = reinterpret_cast<float*>(m->wi[6]);
  float* db = reinterpret_cast<float*>(m->wi[7]);
  GOSPL_CHECK(hipMemsetAsync(ctl, 0, sizeof(HillCtl), s));
  GOSPL_CHECK(hipMemcpyAsync(x, h, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  for (int pass = 0; pass <= maxpass; ++pass) {
    k_hill_resid<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->hcoef, m->hdiag, h, x, r32, part, ctl);
    k_hill_norm<<<1, GOSPL_BLOCK, 0, s>>>(g, part, rtol, pass, maxpass, ctl);
    if (pass == maxpass) break;
    const float* din = da;
    float* dout = da;
    for (int k = 0; k < sweeps32; ++k) {
      k_hill_sweep32<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->hcoef32, m->hdiag32, r32, din, k == 0,
                                               dout, ctl);
      din = dout;
      dout = (dout == da) ? db : da;
    }
    k_hill_addcorr<<<g, GOSPL_BLOCK, 0, s>>>(N, din, x, ctl);
    GOSPL_CHECK(hipGetLastError());
  }
  k_hill_update<<<g, GOSPL_BLOCK, 0, s>>>(N, x, dt, h, cumed, eblocal);
  GOSPL_CHECK(hipGetLastError());
  if (passes || relres) {
    GOSPL_CHECK(hipMemcpyAsync(m->rx_hctl + 32, ctl, sizeof(HillCtl), hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    const HillCtl* hc = reinterpret_cast<const HillCtl*>(m->rx_hctl + 32);
    if (passes) *passes = (int32_t)hc->pass;
    if (relres) *relres = hc->relres;
  }
  return 0;
}

// The frozen operator's rows (sea flags given), in the adjoint's buffers.
static int hill_frozen_rows(gospl_mesh* m, double cda, double cdm, double dt, const int8_t* sea, hipStream_t s) {
  const int64_t N = m->N;
  if (!m->adj_hcoef) {
    GOSPL_CHECK(hipMalloc(&m->adj_hcoef, sizeof(double) * (int64_t)m->W * (N + 64)));
    GOSPL_CHECK(hipMalloc(&m->adj_hdiag, sizeof(double) * (N + 64)));
  }
  k_hill_rows_flags<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->area, m->elgt, m->vdist,
                                                       m->outlet, m->flat, sea, cda, cdm, dt, m->adj_hcoef,
                                                       m->adj_hdiag);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

extern "C" int gospl_hillslope_frozen(gospl_mesh_t* m, double cda, double cdm, double dt, const int8_t* sea,
                                      int32_t sweeps, const double* h_in, double* h_out, gospl_stream_t s) {
  if (sweeps < 1) {
    set_error("gospl_hillslope_frozen: sweeps must be >= 1 (a count from the contraction bound)");
    return 1;
  }
  if (hill_frozen_rows(m, cda, cdm, dt, sea, s)) return 1;
  const int64_t N = m->N;
  const int g = nblocks(N);
  double* xa = m->wd[1];
  double* xb = m->wd[2];
  const double* xin = h_in;
  double* xout = xa;
  for (int k = 0; k < sweeps; ++k) {
    k_hill_sweep<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->adj_hcoef, m->adj_hdiag, h_in, xin, xout);
    xin = xout;
    xout = (xout == xa) ? xb : xa;
  }
  GOSPL_CHECK(hipMemcpyAsync(h_out, xin, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

extern "C" int gospl_hillslope_adjoint(gospl_mesh_t* m, double cda, double cdm, double dt, const int8_t* sea,
                                       int32_t sweeps, const double* lam, double* out, gospl_stream_t s) {
  if (sweeps < 1) {
    set_error("gospl_hillslope_adjoint: sweeps must be >= 1 (a count from the contraction bound)");
    return 1;
  }
  if (ensure_records(m, s)) return 1;  // the reverse neighbour positions
  if (hill_frozen_rows(m, cda, cdm, dt, sea, s)) return 1;
  const int64_t N = m->N;
  const int g = nblocks(N);
  double* xa = m->wd[1];
  double* xb = m->wd[2];
  const double* xin = lam;
  double* xout = xa;
  for (int k = 0; k < sweeps; ++k) {
    k_hill_sweepT<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->rev, m->adj_hcoef, m->adj_hdiag, lam, xin,
                                           xout);
    xin = xout;
    xout = (xout == xa) ? xb : xa;
  }
  GOSPL_CHECK(hipMemcpyAsync(out, xin, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

extern "C" int gospl_hillslope_refine_seg(gospl_mesh_t* m, double cda, double cdm, double dt, double sealevel,
                                          const int32_t* seg, int32_t nseg, const int64_t* offsets,
                                          const int32_t* sweeps32, double rtol, int32_t maxpass, double* h,
                                          double* cumed, double* eblocal, int32_t* passes, gospl_stream_t s) {
  const int64_t N = m->N;
  if (nseg < 1 || !seg || !offsets || !sweeps32) {
    set_error("gospl_hillslope_refine_seg: need seg, offsets and sweeps32, nseg >= 1");
    return 1;
  }
  if (offsets[0] != 0 || offsets[nseg] != N) {
    set_error("gospl_hillslope_refine_seg: offsets must run from 0 to N (%lld)", (long long)N);
    return 1;
  }
  int maxsw = 0;
  for (int k = 0; k < nseg; ++k) {
    if (offsets[k + 1] <= offsets[k] || offsets[k] % GOSPL_BLOCK != 0 || offsets[k + 1] - offsets[k] > INT32_MAX) {
      set_error("gospl_hillslope_refine_seg: segment %d must be a non-empty run of whole %d-vertex workgroups", k,
                GOSPL_BLOCK);
      return 1;
    }
    if (sweeps32[k] < 1) {
      set_error("gospl_hillslope_refine_seg: segment %d has %d FP32 sweeps (>= 1 needed)", k, (int)sweeps32[k]);
      return 1;
    }
    maxsw = sweeps32[k] > maxsw ? sweeps32[k] : maxsw;
  }
  if (maxpass < 1) maxpass = 1;
  if (ensure_seg(m, nseg)) return 1;
  if (!m->seg_haux || m->seg_haux_n < 4 * nseg) {
    free(m->seg_haux);
    m->seg_haux = static_cast<int32_t*>(calloc(4 * (size_t)nseg, sizeof(int32_t)));
    if (!m->seg_haux) {
      set_error("gospl_hillslope_refine_seg: out of host memory");
      return 1;
    }
    m->seg_haux_n = -1;  // forces the upload below
  }
  // segment layout and sweep counts: uploaded when they change
  std::vector<int32_t> aux(4 * (size_t)nseg);
  for (int k = 0; k < nseg; ++k) {
    aux[k] = (int32_t)(offsets[k] / GOSPL_BLOCK);
    aux[nseg + k] = (int32_t)((offsets[k + 1] - offsets[k]) / GOSPL_BLOCK);
    aux[2 * nseg + k] = sweeps32[k];
    aux[3 * nseg + k] = (int32_t)offsets[k];
  }
  if (m->seg_haux_n != 4 * nseg || memcmp(m->seg_haux, aux.data(), sizeof(int32_t) * 4 * nseg) != 0) {
    GOSPL_CHECK(hipMemcpyAsync(m->seg_aux, aux.data(), sizeof(int32_t) * 4 * nseg, hipMemcpyHostToDevice, s));
    GOSPL_CHECK(hipStreamSynchronize(s));  // the host vector goes out of scope
    memcpy(m->seg_haux, aux.data(), sizeof(int32_t) * 4 * nseg);
    m->seg_haux_n = 4 * nseg;
  }
  const int g = nblocks(N);
  if (hill_operator(m, cda, cdm, dt, sealevel, h, s)) return 1;
  HillCtl* ctl = reinterpret_cast<HillCtl*>(m->seg_ctl);
  double* x = m->wd[1];
  double* part = m->wd[0];  // 2 per workgroup <= N + 64
  float* r32 = reinterpret_cast<float*>(m->wi[5]);
  float* da = reinterpret_cast<float*>(m->wi[6]);
  float* db = reinterpret_cast<float*>(m->wi[7]);
  GOSPL_CHECK(hipMemsetAsync(ctl, 0, sizeof(HillCtl) * nseg, s));
  GOSPL_CHECK(hipMemcpyAsync(x, h, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  for (int pass = 0; pass <= maxpass; ++pass) {
    k_hill_resid_seg<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->hcoef, m->hdiag, h, x, r32, part, seg,
                                               ctl);
    k_hill_norm_seg<<<nseg, GOSPL_BLOCK, 0, s>>>(nseg, part, m->seg_aux, rtol, pass, maxpass, ctl);
    if (pass == maxpass) break;
    const float* din = da;
    float* dout = da;
    for (int k = 0; k < maxsw; ++k) {
      k_hill_sweep32_seg<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->hcoef32, m->hdiag32, r32, din, k,
                                                   dout, seg, nseg, m->seg_aux, ctl);
      din = dout;
      dout = (dout == da) ? db : da;
    }
    k_hill_addcorr_seg<<<g, GOSPL_BLOCK, 0, s>>>(N, din, x, seg, nseg, m->seg_aux, ctl);
    GOSPL_CHECK(hipGetLastError());
  }
  k_hill_update<<<g, GOSPL_BLOCK, 0, s>>>(N, x, dt, h, cumed, eblocal);
  GOSPL_CHECK(hipGetLastError());
  if (passes) {
    std::vector<HillCtl> hc(nseg);
    GOSPL_CHECK(hipMemcpyAsync(hc.data(), ctl, sizeof(HillCtl) * nseg, hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < nseg; ++k) passes[k] = (int32_t)hc[k].pass;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Non-linear thickness diffusion dh/dt = div(Cd(h) grad h) over one step by
// goSPL's lagged-diffusivity scheme (hillSLP._diffuseImplicitPicard,
// gospl/sed/hillslope.py:925-1076): `nsub` backward-Euler sub-steps of dt /
// nsub, each with `npic` Picard updates that freeze Cd at the current iterate
// and solve (I + dt_sub L(Cd)) h' = h_start, L the finite-volume operator of
// jacobiancoeff with a zero derivative term (functions.F90:1421-1467).
// Cd = minDiff + Cd0 (1 - exp(-dexp dh)) on the masked cells (0 elsewhere),
// dh = h - bed with thicknesses below 0.1 m taken as 0 (hillslope.py:
// 1030-1034); a face carries flux only when both of its cells diffuse.  Each
// linear system is solved to convergence by Jacobi sweeps from the current
// iterate (the contraction bound rho = max (d - 1) / d of the rows: stop once
// the change times rho / (1 - rho) is below 2^-52 of the largest value);
// goSPL stops its Richardson solve at rtol 1e-8.
// ---------------------------------------------------------------------------

namespace gospl {

__device__ __forceinline__ double nld_cd(int64_t i, const double* __restrict__ h, const double* __restrict__ zb,
                                         const int8_t* __restrict__ mask, double cd0, double mind,
                                         double dexp) {
  if (!mask[i]) return 0.0;
  double dh = h[i] - zb[i];
  if (dh < 0.1) dh = 0.0;
  return mind + cd0 * (1.0 - exp(-dexp * dh));
}

// The jacobiancoeff rows (Kp = 0) of the current iterate h, with what the
// linear solve needs first: the row's contraction (d - 1) / d and the first Jacobi sweep from x0 = h with the
// row's own coefficients (k_hill_sweep_chk's operation order), xout = x1;
// per-workgroup maxima of |x1 - x0|, |x1| and the contraction into
// part[3 b + 0..2] for k_max3_finish.
__global__ __launch_bounds__(GOSPL_BLOCK) void k_nld_rows_first(
    int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
    const double* __restrict__ area, const double* __restrict__ elgt, const double* __restrict__ vdist,
    const double* __restrict__ h, const double* __restrict__ zb, const int8_t* __restrict__ mask, double cd0,
    double mind, double dexp, double dts, const double* __restrict__ rhs, double* __restrict__ coef,
    double* __restrict__ diag, double* __restrict__ xout, unsigned long long* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double chg = 0.0, mag = 0.0, rho = 0.0;
  if (i < N) {
    double s = 0.0;
    const double a = area[i];
    const double ck = nld_cd(i, h, zb, mask, cd0, mind, dexp);
    const int dg = deg[i];
    double off[GOSPL_MAXW];
    int nb[GOSPL_MAXW];
#pragma unroll
    for (int p = 0; p < GOSPL_MAXW; ++p) {
      off[p] = 0.0;
      nb[p] = 0;
      if (p >= W) continue;
      if (a > 0.0 && p < dg) {
        const int64_t e = (int64_t)p * N + i;
        const double vd = vdist[e];
        const int n = nbr[e];
        nb[p] = n;
        if (vd > 0.0) {
          const double cn = nld_cd(n, h, zb, mask, cd0, mind, dexp);
          if (ck > 0.0 && cn > 0.0) {
            const double c = 0.5 * (ck + cn) / a;
            const double v = c * vd / elgt[e];
            s = s + v;
            off[p] = dts * (-v);
          }
        }
      }
      coef[(int64_t)p * N + i] = off[p];
    }
    const double d = dts * s + 1.0;
    diag[i] = d;
    if (d > 0.0) rho = (d - 1.0) / d;
    const double r = rhs[i];
    double x = r;
    if (d != 0.0) {
      double t = r;
#pragma unroll
      for (int p = 0; p < GOSPL_MAXW; ++p)
        if (p < W && p < dg && off[p] != 0.0) t = t - off[p] * h[nb[p]];
      x = t / d;
    }
    xout[i] = x;
    chg = fabs(x - h[i]);
    mag = fabs(x);
  }
  for (int o = 32; o > 0; o >>= 1) {
    chg = fmax(chg, __shfl_xor(chg, o, 64));
    mag = fmax(mag, __shfl_xor(mag, o, 64));
    rho = fmax(rho, __shfl_xor(rho, o, 64));
  }
  __shared__ double sh[3][GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = chg;
    sh[1][threadIdx.x >> 6] = mag;
    sh[2][threadIdx.x >> 6] = rho;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) {
      chg = fmax(chg, sh[0][w]);
      mag = fmax(mag, sh[1][w]);
      rho = fmax(rho, sh[2][w]);
    }
    part[3 * blockIdx.x] = dkey(chg);
    part[3 * blockIdx.x + 1] = dkey(mag);
    part[3 * blockIdx.x + 2] = rho > 0.0 ? dkey(rho) : 0ull;
  }
}

// The three maxima of k_nld_rows_first's workgroups into keys[0..2].
__global__ __launch_bounds__(1024) void k_max3_finish(int nblk, const unsigned long long* __restrict__ part,
                                                      unsigned long long* __restrict__ keys) {
  __shared__ unsigned long long sm[3][1024 / 64];
  unsigned long long m3[3] = {0ull, 0ull, 0ull};
  for (int k = threadIdx.x; k < nblk; k += blockDim.x)
    for (int j = 0; j < 3; ++j) {
      const unsigned long long x = part[3 * k + j];
      m3[j] = x > m3[j] ? x : m3[j];
    }
  for (int o = 32; o > 0; o >>= 1)
    for (int j = 0; j < 3; ++j) {
      const unsigned long long x = __shfl_xor(m3[j], o);
      m3[j] = x > m3[j] ? x : m3[j];
    }
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 3; ++j) sm[j][threadIdx.x >> 6] = m3[j];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int j = 0; j < 3; ++j) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m3[j] = sm[j][w] > m3[j] ? sm[j][w] : m3[j];
      keys[j] = m3[j] > keys[j] ? m3[j] : keys[j];
    }
}

__global__ void k_nld_init(int64_t N, const double* __restrict__ dh, const double* __restrict__ zb,
                           double* __restrict__ h) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  h[i] = dh[i] + zb[i];
}

__global__ void k_nld_out(int64_t N, const double* __restrict__ h, const double* __restrict__ zb,
                          double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double d = h[i] - zb[i];
  out[i] = d < 0.0 ? 0.0 : d;
}

}  // namespace gospl

extern "C" int gospl_nl_diffusion(gospl_mesh_t* m, const double* dh, const int8_t* mask, double cd0,
                                  double min_diff, double dexp, double dt, int32_t nsub, int32_t npic,
                                  const double* zb, double* ndepo, int64_t* sweeps, gospl_stream_t s) {
  if (nsub < 1 || npic < 1 || !(dt > 0.0)) {
    set_error("gospl_nl_diffusion: bad arguments (nsub=%d, npic=%d, dt=%g)", (int)nsub, (int)npic, dt);
    return 1;
  }
  const int64_t N = m->N;
  const int64_t ld = N + 64;
  const int g = nblocks(N);
  if (!m->nld_buf) GOSPL_CHECK(hipMalloc(&m->nld_buf, sizeof(double) * ld * (m->W + 5)));
  double* coef = m->nld_buf;
  double* diag = coef + ld * m->W;
  double* h = diag + ld;       // current iterate
  double* hs = h + ld;         // the sub-step's start (right-hand side)
  double* xa = hs + ld;
  double* xb = xa + ld;
  // the coefficient kernel reads rows column-major with stride N
  const double dts = dt / nsub;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(m->dflag + 48);
  unsigned long long* hk = reinterpret_cast<unsigned long long*>(m->hflag + 48);
  int64_t total = 0;
  k_nld_init<<<g, GOSPL_BLOCK, 0, s>>>(N, dh, zb, h);
  GOSPL_CHECK(hipGetLastError());
  for (int sub = 0; sub < nsub; ++sub) {
    GOSPL_CHECK(hipMemcpyAsync(hs, h, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    for (int it = 0; it < npic; ++it) {
      // the rows, their contraction and the first Jacobi sweep from the
      // current iterate (its largest change and magnitude) in one pass
      GOSPL_CHECK(hipMemsetAsync(keys, 0, sizeof(unsigned long long) * 3, s));
      k_nld_rows_first<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->area, m->elgt, m->vdist, h, zb, mask,
                                                 cd0, min_diff, dexp, dts, hs, coef, diag, xa, m->red_part);
      k_max3_finish<<<1, 1024, 0, s>>>(g, m->red_part, keys);
      GOSPL_CHECK(hipGetLastError());
      GOSPL_CHECK(hipMemcpyAsync(hk, keys, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost, s));
      GOSPL_CHECK(hipStreamSynchronize(s));
      const double rho = hk[2] ? dkey_inv_host(hk[2]) : 0.0;
      const double chg = dkey_inv_host(hk[0]), mag = dkey_inv_host(hk[1]);
      if (!(rho < 1.0) || !(chg == chg) || !(mag == mag)) {
        set_error("gospl_nl_diffusion: Jacobi sweeps do not converge (contraction %.6g)", rho);
        return 1;
      }
      // Chebyshev semi-iteration over the sweep (the operator is symmetric up
      // to the row scaling by the cell areas, so the sweep's spectrum is real
      // and within [-rho, rho]): x_{k+1} = w_k (sweep(x_k) - x_{k-1}) +
      // x_{k-1}, w_1 = 1 / (1 - rho^2 / 2), w_k = 1 / (1 - rho^2 w_{k-1} / 4),
      // for the least count whose error bound 2 sigma^k ||x0 - x*||, sigma =
      // rho / (1 + sqrt(1 - rho^2)), ||x0 - x*|| <= chg / (1 - rho), with
      // 2^16 of room for the norm that symmetrises the sweep against the max
      // norm, is at most 2^-52 of the solution's magnitude.
      const double r2 = rho * rho;
      const double sig = rho / (1.0 + std::sqrt(1.0 - r2));
      const double tgt = 0x1p-52 * mag;
      double pw = 0x1p17 * chg / (1.0 - rho) * sig;
      int nsw = 1;
      while (pw > tgt && nsw < GOSPL_HILL_MAX_SWEEPS) {
        pw *= sig;
        ++nsw;
      }
      if (pw > tgt) {
        set_error("gospl_nl_diffusion: %d sweeps do not reach the error bound (contraction %.6g)",
                  (int)GOSPL_HILL_MAX_SWEEPS, rho);
        return 1;
      }
      double* buf[3] = {h, xa, xb};
      int prev = 0, cur = 1;
      double omega = 1.0;
      for (int k = 1; k < nsw; ++k) {
        omega = k == 1 ? 1.0 / (1.0 - 0.5 * r2) : 1.0 / (1.0 - 0.25 * r2 * omega);
        const int nxt = 3 - prev - cur;
        k_hill_cheb<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, coef, diag, hs, buf[cur], buf[prev], omega,
                                              buf[nxt]);
        prev = cur;
        cur = nxt;
      }
      GOSPL_CHECK(hipGetLastError());
      total += nsw;
      if (cur != 0) GOSPL_CHECK(hipMemcpyAsync(h, buf[cur], sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    }
  }
  k_nld_out<<<g, GOSPL_BLOCK, 0, s>>>(N, h, zb, ndepo);
  GOSPL_CHECK(hipGetLastError());
  if (sweeps) *sweeps = total;
  return 0;
}
