// Marine deposition of river sediment: the per-vertex parts of goSPL's
// SEAMesh.seaChange (gospl/sed/seaplex.py:738-861).
//
//  * gospl_coast_distance: _distanceCoasts / _globalCoastsTree (:51-135).
//    The coastline is VTK's contour of the elevation at sea level on the
//    triangle mesh: one point on every edge whose ends straddle the level
//    (one end below, the other at or above it), linearly interpolated from
//    the lower end, x = x_lo + t (x_hi - x_lo), t = (level - z_lo) /
//    (z_hi - z_lo).  Each selected vertex gets the Euclidean distance to the
//    nearest coastline point (goSPL queries a k-d tree): here an exact search
//    of a bounding-box tree over the points (counting-sorted by the Morton
//    code of a uniform grid's cells, buckets of 8 consecutive points as
//    leaves, a complete binary tree above them), nearer child first,
//    pruning boxes that cannot hold a closer point.
//  * gospl_marine_receivers: mfdrcvrs (functions.F90:1673-1785) as _matOcean
//    calls it (12 receivers, sea level at the -1e6 sentinel): every lower
//    neighbour on the filled surface whose drop to the power 1.1 over the
//    edge length does not underflow, in neighbour order, each with
//    weight 1 / (their count) as the Fortran computes it (1. / kk, single
//    precision); no lower neighbour: a sink (flat-mesh outlets: no
//    receivers, not sinks).
//  * gospl_marine_route: _distOcean (:404-619): the incoming volumes pushed
//    down the receivers pass by pass, each vertex keeping what its shelf
//    capacity (clinoH - h) area still holds, sinks keeping everything,
//    outlets discarding it, until the volume still moving is at most
//    max(1, 1e-6 of the input) or 5000 passes; what still moves is then
//    deposited where it is.  A pass is three kernels that return at once
//    when a device flag says the loop has ended, so passes are issued in
//    batches and the host reads the flag once per batch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "common.h"

namespace gospl {
namespace {

constexpr int ROUTE_BATCH = 16;
constexpr int ROUTE_MAX_PASSES = 5000;
constexpr int SUM_BLOCKS = 1024;

// ---------------------------------------------------------------- coastline

// crossing edges counted from their lower end
__global__ void k_cst_count(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                            const double* __restrict__ z, double v, int32_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;
  int c = 0;
  if (i < N && z[i] < v) {
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      const int n = nbr[(int64_t)p * N + i];
      if (n >= 0 && z[n] >= v) ++c;
    }
  }
  cnt[i] = c;
}

__global__ void k_cst_points(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                             const double* __restrict__ z, double v, const double* __restrict__ xyz,
                             const int32_t* __restrict__ off, double* __restrict__ pts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double zi = z[i];
  if (!(zi < v)) return;
  int o = off[i];
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    const int n = nbr[(int64_t)p * N + i];
    if (n < 0) continue;
    const double zn = z[n];
    if (!(zn >= v)) continue;
    const double t = (v - zi) / (zn - zi);
    for (int j = 0; j < 3; ++j) {
      const double x1 = xyz[3 * i + j], x2 = xyz[3 * (int64_t)n + j];
      pts[3 * (int64_t)o + j] = x1 + t * (x2 - x1);
    }
    ++o;
  }
}

// bounding box of the points: keys of min x, y, z, then of -max x, y, z
__global__ void k_cst_bbox(int64_t M, const double* __restrict__ pts, unsigned long long* __restrict__ keys) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long lo[3], hi[3];
  for (int j = 0; j < 3; ++j) {
    lo[j] = ~0ull;
    hi[j] = ~0ull;
  }
  if (k < M)
    for (int j = 0; j < 3; ++j) {
      lo[j] = dkey(pts[3 * k + j]);
      hi[j] = dkey(-pts[3 * k + j]);
    }
  for (int j = 0; j < 3; ++j)
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(lo[j], o), b = __shfl_xor(hi[j], o);
      lo[j] = a < lo[j] ? a : lo[j];
      hi[j] = b < hi[j] ? b : hi[j];
    }
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 3; ++j) {
      atomicMin(keys + j, lo[j]);
      atomicMin(keys + 3 + j, hi[j]);
    }
}

struct Grid {
  double lo[3];
  double cs;
  int n[3];
};

__device__ __forceinline__ int grid_axis(double x, double lo, double cs, int n) {
  int c = (int)floor((x - lo) / cs);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// bits of x spread to every third position
__device__ __forceinline__ unsigned int spread3(unsigned int x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Sort key of each point: the Morton code of its grid cell, the cell
// coordinates shifted right by `sh` so that the keys stay below 2^(3 bits).
__global__ void k_cst_cell(int64_t M, const double* __restrict__ pts, Grid g, int sh, int32_t* __restrict__ cell,
                           int32_t* __restrict__ cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const int cx = grid_axis(pts[3 * k], g.lo[0], g.cs, g.n[0]) >> sh;
  const int cy = grid_axis(pts[3 * k + 1], g.lo[1], g.cs, g.n[1]) >> sh;
  const int cz = grid_axis(pts[3 * k + 2], g.lo[2], g.cs, g.n[2]) >> sh;
  const int c = (int)(spread3((unsigned)cx) | (spread3((unsigned)cy) << 1) | (spread3((unsigned)cz) << 2));
  cell[k] = c;
  atomicAdd(cnt + c, 1);
}

__global__ void k_cst_scatter(int64_t M, const double* __restrict__ pts, const int32_t* __restrict__ cell,
                              const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                              double* __restrict__ sorted) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const int c = cell[k];
  const int slot = start[c] + atomicAdd(fill + c, 1);
  for (int j = 0; j < 3; ++j) sorted[3 * (int64_t)slot + j] = pts[3 * k + j];
}

#define CST_LEAF 8  // sorted points per leaf box

// Leaf boxes: leaf j holds sorted points [8 j, 8 j + 8) (fewer in the last);
// leaves past the points are empty boxes (lo = +inf, hi = -inf).  Tree in
// heap order: node 1 the root, children 2 i and 2 i + 1, leaves at P .. 2 P - 1
// (P a power of two); each node 6 doubles, lo xyz then hi xyz.
__global__ void k_cst_leaves(int64_t M, int64_t P, const double* __restrict__ sorted, double* __restrict__ box) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P) return;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const int64_t a = j * CST_LEAF, b = a + CST_LEAF < M ? a + CST_LEAF : M;
  for (int64_t k = a; k < b; ++k)
    for (int d = 0; d < 3; ++d) {
      const double x = sorted[3 * k + d];
      lo[d] = x < lo[d] ? x : lo[d];
      hi[d] = x > hi[d] ? x : hi[d];
    }
  double* o = box + 6 * (P + j);
  for (int d = 0; d < 3; ++d) {
    o[d] = lo[d];
    o[3 + d] = hi[d];
  }
}

// Inner boxes, level by level up to the root, in one workgroup.
__global__ __launch_bounds__(1024) void k_cst_inner(int64_t P, double* __restrict__ box) {
  for (int64_t w = P >> 1; w >= 1; w >>= 1) {
    for (int64_t i = w + threadIdx.x; i < 2 * w; i += blockDim.x) {
      const double* l = box + 6 * (2 * i);
      const double* r = box + 6 * (2 * i + 1);
      double* o = box + 6 * i;
      for (int d = 0; d < 3; ++d) {
        o[d] = l[d] < r[d] ? l[d] : r[d];
        o[3 + d] = l[3 + d] > r[3 + d] ? l[3 + d] : r[3 + d];
      }
    }
    __syncthreads();
  }
}

// Lower bound of the squared distance from q to a box: each axis's gap, in
// the distance's own order (rounding is monotone, so the bound never
// exceeds a computed distance to a point inside); +inf for an empty box.
__device__ __forceinline__ double box_lb(const double* b, double qx, double qy, double qz) {
  if (!(b[0] <= b[3])) return INFINITY;
  const double gx = qx < b[0] ? b[0] - qx : (qx > b[3] ? qx - b[3] : 0.0);
  const double gy = qy < b[1] ? b[1] - qy : (qy > b[4] ? qy - b[4] : 0.0);
  const double gz = qz < b[2] ? b[2] - qz : (qz > b[5] ? qz - b[5] : 0.0);
  return (gx * gx + gy * gy) + gz * gz;
}

// nearest point for each selected vertex (0 elsewhere): depth-first, the
// nearer child first, boxes no nearer than the best point so far skipped
__global__ void k_cst_query(int64_t N, const double* __restrict__ xyz, const int8_t* __restrict__ mask, int64_t M,
                            int64_t P, const double* __restrict__ box, const double* __restrict__ sorted,
                            double* __restrict__ dist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (mask && !mask[i]) {
    dist[i] = 0.0;
    return;
  }
  const double qx = xyz[3 * i], qy = xyz[3 * i + 1], qz = xyz[3 * i + 2];
  double best = INFINITY;
  int64_t stack[64];
  int sp = 0;
  stack[sp++] = 1;
  while (sp > 0) {
    const int64_t nd = stack[--sp];
    if (box_lb(box + 6 * nd, qx, qy, qz) >= best) continue;
    if (nd >= P) {
      const int64_t a = (nd - P) * CST_LEAF, b = a + CST_LEAF < M ? a + CST_LEAF : M;
      for (int64_t k = a; k < b; ++k) {
        const double ex = qx - sorted[3 * k];
        const double ey = qy - sorted[3 * k + 1];
        const double ez = qz - sorted[3 * k + 2];
        const double d2 = (ex * ex + ey * ey) + ez * ez;
        best = d2 < best ? d2 : best;
      }
      continue;
    }
    const double la = box_lb(box + 6 * (2 * nd), qx, qy, qz);
    const double lb = box_lb(box + 6 * (2 * nd + 1), qx, qy, qz);
    // the farther child below the nearer one (depth <= 62: two per level)
    const int64_t nf = la <= lb ? 2 * nd + 1 : 2 * nd, nn = la <= lb ? 2 * nd : 2 * nd + 1;
    const double lf = la <= lb ? lb : la, ln = la <= lb ? la : lb;
    if (lf < best && sp < 64) stack[sp++] = nf;
    if (ln < best && sp < 64) stack[sp++] = nn;
  }
  dist[i] = sqrt(best);
}

// ------------------------------------------------------------ receivers

__global__ void k_mar_rcv(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                          const double* __restrict__ elgt, const double* __restrict__ fz,
                          const int8_t* __restrict__ outlet, int flat, int32_t* __restrict__ rcv,
                          double* __restrict__ wgt, int8_t* __restrict__ sink) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double zi = fz[i];
  // val = (z_k - z_n)^fexp / L > 0, fexp 1.1 above the -1e6 sentinel and 0.01
  // at it (functions.F90:1709-1729): a drop whose power underflows (or is
  // negative: NaN) does not count
  const double fexp = zi > -1.0e6 ? 1.1 : 0.01;
  int ids[GOSPL_MAXW];
  int kk = 0;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    const int n = nbr[(int64_t)p * N + i];
    if (n < 0) continue;
    const double L = elgt[(int64_t)p * N + i];
    if (!(L > 0.0)) continue;
    const double val = pow(zi - fz[n], fexp) / L;
    if (val > 0.0) ids[kk++] = n;
  }
  const bool out = flat && outlet[i];
  if (out) kk = 0;
  // the Fortran's 1. / kk: a single-precision quotient stored in a double
  const double w = kk > 0 ? (double)__fdiv_rn(1.0f, (float)kk) : 0.0;
  for (int k = 0; k < GOSPL_MAXW; ++k) {
    rcv[(int64_t)k * N + i] = k < kk ? ids[k] : (int32_t)i;
    wgt[(int64_t)k * N + i] = k < kk ? w : 0.0;
  }
  sink[i] = (kk == 0 && !out) ? 1 : 0;
}

// -------------------------------------------------------------- routing

// Shelf capacity and incoming volume (seaplex.py:761-776): clinoH = sl -
// 1e-3 - dist clinSlp (hl where hl >= sl or where clinoH < hl); cap =
// (clinoH - hl) area; vin = vsed dt on cells with lfill <= sl, 0 elsewhere
// or where negative.
__global__ void k_mar_prep(int64_t N, const double* __restrict__ hl, const double* __restrict__ cdist,
                           const double* __restrict__ lfill, const double* __restrict__ vsed,
                           const double* __restrict__ area, double sl, double clin, double dt,
                           double* __restrict__ cap, double* __restrict__ vin) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double h = hl[i];
  double c = sl - 1.0e-3 - cdist[i] * clin;
  if (h >= sl) c = h;
  if (c < h) c = h;
  cap[i] = (c - h) * area[i];
  double v = vsed[i] * dt;
  if (!(lfill[i] <= sl)) v = 0.0;
  if (v < 0.0) v = 0.0;
  vin[i] = v;
}

// start of _distOcean: capacities copied, sinks keep their input, the sum of
// what moves (per-workgroup partials)
__global__ void k_route_init(int64_t N, const double* __restrict__ cap0, const double* __restrict__ vin,
                             const int8_t* __restrict__ sink, double* __restrict__ cap, double* __restrict__ mv,
                             double* __restrict__ dep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  cap[i] = cap0[i];
  const double v = vin[i];
  if (sink[i]) {
    dep[i] = v;
    mv[i] = 0.0;
  } else {
    dep[i] = 0.0;
    mv[i] = v;
  }
}

// per-workgroup partial sums of x
__global__ __launch_bounds__(256) void k_psum(int64_t N, const double* __restrict__ x, double* __restrict__ part,
                                              const int32_t* __restrict__ ctl) {
  if (ctl && ctl[0]) return;
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (int64_t)SUM_BLOCKS * 256) acc += x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ctl[0] done, ctl[1] passes run; tot[0] the first sum (the threshold is set
// from it), tot[1] the last
__global__ __launch_bounds__(256) void k_route_check(const double* __restrict__ part, int32_t* __restrict__ ctl,
                                                     double* __restrict__ tot, int first) {
  if (!first && ctl[0]) return;
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < SUM_BLOCKS; i += 256) acc += part[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double s = sh[0];
  if (first) {
    tot[0] = s;
    tot[2] = 1.0e-6 * s > 1.0 ? 1.0e-6 * s : 1.0;  // max(1, 1e-6 initial total)
    ctl[0] = 0;
    ctl[1] = 0;
  } else {
    ctl[1] += 1;
  }
  tot[1] = s;
  if (!(s > tot[2]) || ctl[1] >= ROUTE_MAX_PASSES) ctl[0] = 1;
}

// The donors of each vertex r -- the neighbours that have r among their
// receivers (gospl_marine_receivers: a vertex's receivers come first, then
// itself) -- sorted by input id, and each vertex's uniform weight: built once
// per routing call, read by every pass.
__global__ void k_route_donors(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                               const int32_t* __restrict__ gid, const int32_t* __restrict__ rcv,
                               const double* __restrict__ wgt, int32_t* __restrict__ don, int8_t* __restrict__ nd,
                               double* __restrict__ wv) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  int dl[GOSPL_MAXW];
  int n_d = 0;
  const int dg = deg[r];
  for (int p = 0; p < W && p < dg; ++p) {
    const int n = nbr[(int64_t)p * N + r];
    if (n < 0) continue;
    for (int k = 0; k < GOSPL_MAXW; ++k) {
      const int q = rcv[(int64_t)k * N + n];
      if (q == (int)n) break;  // past the last receiver
      if (q == (int)r) {
        // insertion by input id
        int j = n_d++;
        const int gn = gid[n];
        while (j > 0 && gid[dl[j - 1]] > gn) {
          dl[j] = dl[j - 1];
          --j;
        }
        dl[j] = n;
        break;
      }
    }
  }
  for (int j = 0; j < n_d; ++j) don[(int64_t)j * N + r] = dl[j];
  nd[r] = (int8_t)n_d;
  wv[r] = wgt[r];  // slot 0: the weight of every receiver of r (0 without one)
}

// One pass of _distOcean (seaplex.py:514-570): moved[r] = sum over r's donors
// (in increasing input id) of mv_in[i] w_i (dMat1.mult), then the capacity
// test; what keeps moving goes to mv_out.
__global__ void k_route_pass(int64_t N, const int32_t* __restrict__ don, const int8_t* __restrict__ nd,
                             const double* __restrict__ wv, const double* __restrict__ mv_in,
                             double* __restrict__ mv_out, const int8_t* __restrict__ sink,
                             const int8_t* __restrict__ outlet, int flat, double* __restrict__ cap,
                             double* __restrict__ dep, const int32_t* __restrict__ ctl) {
  if (ctl[0]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int n_d = nd[i];
  double s = 0.0;
  for (int j = 0; j < n_d; ++j) {
    const int d = don[(int64_t)j * N + i];
    s += mv_in[d] * wv[d];
  }
  double c = cap[i];
  double dp = dep[i];
  if (s >= c) {
    s = s - c;
    dp = dp + c;
    c = 0.0;
  } else {
    c = c - s;
    dp = dp + s;
    s = 0.0;
  }
  if (flat && outlet[i]) {
    dp = 0.0;
    s = 0.0;
  }
  if (sink[i]) {
    dp = dp + s;
    s = 0.0;
  }
  cap[i] = c;
  dep[i] = dp;
  mv_out[i] = s;
}

// what still moves is deposited in place (seaplex.py:596-606)
__global__ void k_route_end(int64_t N, const double* __restrict__ mv, const int8_t* __restrict__ outlet, int flat,
                            double* __restrict__ dep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double d = dep[i] + mv[i];
  if (flat && outlet[i]) d = 0.0;
  dep[i] = d;
}

// dh = vdep / area (0 where area is 0), below the 1e-3 m floor set to 0
// (seaplex.py:814-818)
__global__ void k_mar_thick(int64_t N, const double* __restrict__ vdep, const double* __restrict__ area,
                            double floor_m, double* __restrict__ dh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double a = area[i];
  double t = a != 0.0 ? vdep[i] / a : 0.0;
  if (t < floor_m) t = 0.0;
  dh[i] = t;
}

// cumed += t; h += t; eblocal += t / dt (seaplex.py:836-847)
__global__ void k_mar_apply(int64_t N, const double* __restrict__ t, double dt, double* __restrict__ h,
                            double* __restrict__ cumed, double* __restrict__ eblocal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double v = t[i];
  cumed[i] = cumed[i] + v;
  h[i] = h[i] + v;
  eblocal[i] = eblocal[i] + v / dt;
}

int ensure_sea(gospl_mesh* m, int64_t npts) {
  // sea_d: 6 (N + 64) doubles of routing vectors; sea_i: 2 (N + 64) ints of
  // coastline counts / offsets; sea_cell and sea_pts: the points' cells and
  // the points with their grid-sorted copy, grown to npts
  const int64_t N = m->N;
  if (!m->sea_d) {
    GOSPL_CHECK(hipMalloc(&m->sea_d, sizeof(double) * 6 * (N + 64)));
    GOSPL_CHECK(hipMalloc(&m->sea_ctl, sizeof(double) * (SUM_BLOCKS + 64)));
    GOSPL_CHECK(hipMalloc(&m->sea_i, sizeof(int32_t) * 2 * (N + 64)));
    GOSPL_CHECK(hipHostMalloc((void**)&m->sea_h, sizeof(unsigned long long) * 8, hipHostMallocDefault));
  }
  if (npts > m->sea_pcap) {
    if (m->sea_pts) GOSPL_CHECK(hipFree(m->sea_pts));
    if (m->sea_cell) GOSPL_CHECK(hipFree(m->sea_cell));
    m->sea_pts = nullptr;
    m->sea_cell = nullptr;
    m->sea_pcap = 0;
    GOSPL_CHECK(hipMalloc(&m->sea_pts, sizeof(double) * 6 * (npts + 64)));
    GOSPL_CHECK(hipMalloc(&m->sea_cell, sizeof(int32_t) * (npts + 64)));
    m->sea_pcap = npts;
  }
  return 0;
}

}  // namespace
}  // namespace gospl

using namespace gospl;

extern "C" {

int gospl_coast_distance(gospl_mesh_t* m, const double* xyz, const double* z, double level, const int8_t* mask,
                         double* dist, int64_t* npoints, gospl_stream_t stream) {
  if (!m || !xyz || !z || !dist) {
    set_error("gospl_coast_distance: null argument");
    return 1;
  }
  hipStream_t s = stream;
  const int64_t N = m->N;
  if (ensure_sea(m, 0)) return 1;
  int32_t* cnt = m->sea_i;  // N + 1 counts, then offsets in place of the next N + 1
  int32_t* off = m->sea_i + (N + 64);
  k_cst_count<<<nblocks(N + 1), GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, z, level, cnt);
  GOSPL_CHECK(hipGetLastError());
  size_t bytes = m->cub_tmp_bytes;
  GOSPL_CHECK(rocprim::exclusive_scan(m->cub_tmp, bytes, cnt, off, 0, (size_t)(N + 1), rocprim::plus<int32_t>(), s));
  int32_t* hs = reinterpret_cast<int32_t*>(m->sea_h);
  GOSPL_CHECK(hipMemcpyAsync(hs, off + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  const int64_t M = hs[0];
  if (npoints) *npoints = M;
  if (M == 0) {
    // no coastline: goSPL leaves every distance at 0
    GOSPL_CHECK(hipMemsetAsync(dist, 0, sizeof(double) * N, s));
    return 0;
  }
  if (ensure_sea(m, M)) return 1;
  int32_t* cellid = m->sea_cell;
  double* pts = m->sea_pts;
  double* sorted = m->sea_pts + 3 * (m->sea_pcap + 32);
  k_cst_points<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, z, level, xyz, off, pts);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(m->sea_ctl);
  GOSPL_CHECK(hipMemsetAsync(keys, 0xff, sizeof(unsigned long long) * 6, s));
  k_cst_bbox<<<nblocks(M), GOSPL_BLOCK, 0, s>>>(M, pts, keys);
  GOSPL_CHECK(hipGetLastError());
  unsigned long long hk[6];
  GOSPL_CHECK(hipMemcpyAsync(m->sea_h, keys, sizeof(hk), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  std::memcpy(hk, m->sea_h, sizeof(hk));
  Grid g;
  double ext = 0.0;
  for (int j = 0; j < 3; ++j) {
    g.lo[j] = dkey_inv_host(hk[j]);
    const double hi = -dkey_inv_host(hk[3 + j]);
    ext = hi - g.lo[j] > ext ? hi - g.lo[j] : ext;
  }
  // about two cells per cube root of the point count along the widest axis,
  // with (G + 1)^3 + 1 <= N + 1 (the scan workspace holds N + 1 entries)
  int G = (int)std::ceil(2.0 * std::cbrt((double)M));
  const int gmax = (int)std::floor(std::cbrt((double)N)) - 1;
  G = G > gmax ? gmax : G;
  G = G < 1 ? 1 : G;
  g.cs = ext > 0.0 ? ext / G : 1.0;
  for (int j = 0; j < 3; ++j) {
    const double hi = -dkey_inv_host(hk[3 + j]);
    int n = (int)std::floor((hi - g.lo[j]) / g.cs) + 1;
    g.n[j] = n < 1 ? 1 : (n > G + 1 ? G + 1 : n);
  }
  // sort keys: Morton codes of the cells, the cell coordinates cut to b bits
  // each so that the 2^(3 b) keys fit the N + 64 counts
  int gbits = 0;
  while ((1 << gbits) < g.n[0] || (1 << gbits) < g.n[1] || (1 << gbits) < g.n[2]) ++gbits;
  int b = gbits < 10 ? gbits : 10;
  while (b > 0 && (int64_t(1) << (3 * b)) + 1 > N + 64) --b;
  const int sh = gbits - b;
  const int64_t nkey = int64_t(1) << (3 * b);
  GOSPL_CHECK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nkey + 1), s));
  k_cst_cell<<<nblocks(M), GOSPL_BLOCK, 0, s>>>(M, pts, g, sh, cellid, cnt);
  GOSPL_CHECK(hipGetLastError());
  bytes = m->cub_tmp_bytes;
  GOSPL_CHECK(rocprim::exclusive_scan(m->cub_tmp, bytes, cnt, off, 0, (size_t)(nkey + 1), rocprim::plus<int32_t>(),
                                      s));
  GOSPL_CHECK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nkey + 1), s));
  k_cst_scatter<<<nblocks(M), GOSPL_BLOCK, 0, s>>>(M, pts, cellid, off, cnt, sorted);
  // the tree: leaves of CST_LEAF consecutive sorted points, P leaves (a power
  // of two), nodes 1 .. 2 P - 1
  const int64_t nleaf = (M + CST_LEAF - 1) / CST_LEAF;
  int64_t P = 1;
  while (P < nleaf) P <<= 1;
  if (2 * P > m->sea_bvh_cap) {
    if (m->sea_bvh) GOSPL_CHECK(hipFree(m->sea_bvh));
    m->sea_bvh = nullptr;
    m->sea_bvh_cap = 0;
    GOSPL_CHECK(hipMalloc(&m->sea_bvh, sizeof(double) * 6 * 2 * P));
    m->sea_bvh_cap = 2 * P;
  }
  k_cst_leaves<<<nblocks(P), GOSPL_BLOCK, 0, s>>>(M, P, sorted, m->sea_bvh);
  if (P > 1) k_cst_inner<<<1, 1024, 0, s>>>(P, m->sea_bvh);
  k_cst_query<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, xyz, mask, M, P, m->sea_bvh, sorted, dist);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_marine_receivers(gospl_mesh_t* m, const double* fillz, int32_t* rcv, double* wgt, int8_t* sink,
                           gospl_stream_t stream) {
  if (!m || !fillz || !rcv || !wgt || !sink) {
    set_error("gospl_marine_receivers: null argument");
    return 1;
  }
  k_mar_rcv<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, m->W, m->nbr, m->deg, m->elgt, fillz, m->outlet, m->flat,
                                                        rcv, wgt, sink);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_marine_prep(gospl_mesh_t* m, const double* hl, const double* cdist, const double* lfill,
                      const double* vsed, double sealevel, double clinslp, double dt, double* cap, double* vin,
                      gospl_stream_t stream) {
  if (!m) {
    set_error("gospl_marine_prep: null handle");
    return 1;
  }
  k_mar_prep<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, hl, cdist, lfill, vsed, m->area, sealevel, clinslp, dt,
                                                         cap, vin);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_marine_route(gospl_mesh_t* m, const int32_t* rcv, const double* wgt, const int8_t* sink,
                       const double* cap0, const double* vin, double* vdep, int32_t* passes, gospl_stream_t stream) {
  if (!m || !rcv || !wgt || !sink || !cap0 || !vin || !vdep) {
    set_error("gospl_marine_route: null argument");
    return 1;
  }
  hipStream_t s = stream;
  const int64_t N = m->N;
  const int g = nblocks(N);
  if (ensure_sea(m, 0)) return 1;
  const int64_t ld = N + 64;
  double* cap = m->sea_d;
  double* mvb[2] = {m->sea_d + ld, m->sea_d + 2 * ld};  // what moves, before / after a pass
  double* wv = m->sea_d + 3 * ld;
  double* part = m->sea_ctl;
  int32_t* ctl = m->dflag + 24;        // done, passes
  double* tot = m->sea_ctl + SUM_BLOCKS;  // first sum, last sum, threshold
  if (!m->sea_don) {
    GOSPL_CHECK(hipMalloc(&m->sea_don, sizeof(int32_t) * (int64_t)m->W * ld));
    GOSPL_CHECK(hipMalloc(&m->sea_nd, ld));
  }
  k_route_donors<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->gid, rcv, wgt, m->sea_don, m->sea_nd, wv);
  k_route_init<<<g, GOSPL_BLOCK, 0, s>>>(N, cap0, vin, sink, cap, mvb[0], vdep);
  k_psum<<<SUM_BLOCKS, 256, 0, s>>>(N, mvb[0], part, nullptr);
  k_route_check<<<1, 256, 0, s>>>(part, ctl, tot, 1);
  GOSPL_CHECK(hipGetLastError());
  int32_t hc[2] = {0, 0};
  int launched = 0;  // pass j reads mvb[j % 2]; passes past the end return at once
  for (;;) {
    GOSPL_CHECK(hipMemcpyAsync(m->sea_h, ctl, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    std::memcpy(hc, m->sea_h, sizeof(hc));
    if (hc[0]) break;
    for (int b = 0; b < ROUTE_BATCH; ++b, ++launched) {
      double* in = mvb[launched & 1];
      double* out = mvb[(launched + 1) & 1];
      k_route_pass<<<g, GOSPL_BLOCK, 0, s>>>(N, m->sea_don, m->sea_nd, wv, in, out, sink, m->outlet, m->flat, cap,
                                             vdep, ctl);
      k_psum<<<SUM_BLOCKS, 256, 0, s>>>(N, out, part, ctl);
      k_route_check<<<1, 256, 0, s>>>(part, ctl, tot, 0);
    }
    GOSPL_CHECK(hipGetLastError());
  }
  // the passes run: hc[1]; the last one wrote mvb[hc[1] % 2]
  k_route_end<<<g, GOSPL_BLOCK, 0, s>>>(N, mvb[hc[1] & 1], m->outlet, m->flat, vdep);
  GOSPL_CHECK(hipGetLastError());
  if (passes) *passes = hc[1];
  return 0;
}

int gospl_marine_thickness(gospl_mesh_t* m, const double* vdep, double floor_m, double* dh, gospl_stream_t stream) {
  if (!m) {
    set_error("gospl_marine_thickness: null handle");
    return 1;
  }
  k_mar_thick<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, vdep, m->area, floor_m, dh);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_marine_apply(gospl_mesh_t* m, const double* t, double dt, double* h, double* cumed, double* eblocal,
                       gospl_stream_t stream) {
  if (!m) {
    set_error("gospl_marine_apply: null handle");
    return 1;
  }
  k_mar_apply<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, t, dt, h, cumed, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
