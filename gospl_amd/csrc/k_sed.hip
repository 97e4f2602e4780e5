// Continental sediment routing and deposition (goSPL's SEDMesh,
// gospl/sed/sedplex.py): the per-vertex pieces of _getSedFlux,
// _moveDownstream, _distributeSediment and _updateSinks.  The solves reuse
// the flow-accumulation sweep, the pit tables come from gospl_pit_info_sed,
// the per-pit sums from gospl_pit_inflow's fixed-order grouped reduction;
// the per-pit decisions (a few numbers per depression) are the caller's.
#include <cmath>

#include "common.h"

namespace gospl {

// vSed source: the eroded volume rate -Eb * area (sedplex.py:125-131)
__global__ void k_sed_source(int64_t N, const double* __restrict__ eb, const double* __restrict__ area,
                             double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  b[i] = (-eb[i]) * area[i];
}

// Flow-terminal cells that are not a depression (lsink & pitID < 0): their
// sediment leaves the routed volume; on a domain without outlets it is kept
// for deposition in place (sedplex.py:236-251).
__global__ void k_sed_stray(int64_t N, const int8_t* __restrict__ lsink, const int32_t* __restrict__ pit,
                            int closed, double* __restrict__ vsed, double* __restrict__ depo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (lsink[i] && pit[i] < 0) {
    if (closed) depo[i] = depo[i] + vsed[i];
    vsed[i] = 0.0;
  }
}

// One routed pass (sedplex.py:424-431): the next volumes q (zero at draining
// borders) and the load rate vSed += q / dt where the cell is not under a
// sediment-filled surface.
__global__ void k_sed_add(int64_t N, const double* __restrict__ q, const double* __restrict__ hl,
                          const double* __restrict__ sedf, const int8_t* __restrict__ outlet, int flat,
                          double dt, double* __restrict__ vnext, double* __restrict__ vsedl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double v = q[i];
  const double nv = hl[i] < sedf[i] ? 0.0 : v;
  vsedl[i] = vsedl[i] + nv / dt;
  vnext[i] = (flat && outlet[i]) ? 0.0 : v;
}

// Deposit thickness from a volume (closed sinks, sedplex.py:491-500)
__global__ void k_sed_closed(int64_t N, const double* __restrict__ depo, const double* __restrict__ area,
                             double* __restrict__ delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double v = depo[i];
  delta[i] = v > 0.0 ? v / area[i] : 0.0;
}

// Bottom-up lake fill (sedplex.py:525-560): cells of the selected pits rise
// to min(target level, filled level); target -inf elsewhere.
__global__ void k_sed_bottomup(int64_t N, const int32_t* __restrict__ pit, const double* __restrict__ target,
                               const double* __restrict__ lfill, const double* __restrict__ hl,
                               double* __restrict__ delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int p = pit[i];
  double d = 0.0;
  if (p >= 0) {
    const double t = target[p];
    if (t > -INFINITY) {
      const double lf = lfill[i];
      const double sfc = t < lf ? t : lf;
      const double dd = sfc - hl[i];
      d = dd > 0.0 ? dd : 0.0;
    }
  }
  delta[i] = d;
}

// Micro-tilt (sedplex.py:584-655): cells of the selected pits that receive a
// deposit, their distance to the pit's spill point.
__global__ void k_sed_tilt_dist(int64_t N, const double* __restrict__ xyz, const int32_t* __restrict__ pit,
                                const int8_t* __restrict__ psel, const double* __restrict__ spill_xyz,
                                const double* __restrict__ delta, int8_t* __restrict__ sel,
                                double* __restrict__ dist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int p = pit[i];
  const int ps = p >= 0 ? p : 0;
  const double dx = xyz[3 * i] - spill_xyz[3 * ps];
  const double dy = xyz[3 * i + 1] - spill_xyz[3 * ps + 1];
  const double dz = xyz[3 * i + 2] - spill_xyz[3 * ps + 2];
  dist[i] = sqrt((dx * dx + dy * dy) + dz * dz);
  sel[i] = (p >= 0 && psel[p] && delta[i] > 0.0) ? 1 : 0;
}

__global__ void k_sed_tilt_apply(int64_t N, const int32_t* __restrict__ pit, const int8_t* __restrict__ sel,
                                 const double* __restrict__ dist, const double* __restrict__ mean,
                                 double* __restrict__ delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !sel[i]) return;
  const double t = 1.0e-6 * (dist[i] - mean[pit[i]]);
  const double d = delta[i] + t;
  delta[i] = d < 0.0 ? 0.0 : d;
}

// Pit inlets (sedplex.py:662-697): cells of a selected pit that receive flow
// from a donor outside that pit.
__global__ void k_sed_inlets(int64_t N, int nrcv, const int32_t* __restrict__ rcv, const int32_t* __restrict__ pit,
                             const int8_t* __restrict__ psel, int8_t* __restrict__ inlet) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int pi = pit[i];
  for (int k = 0; k < nrcv; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    if (r == (int)i || r < 0) continue;
    const int pr = pit[r];
    if (pr >= 0 && pr != pi && psel[pr]) inlet[r] = 1;
  }
}

// Large-pit initial deposit (sedplex.py:833-871): water depth below the
// target level on the selected cells, and wd * area for its per-pit sums.
__global__ void k_sed_lake_depth(int64_t N, const int32_t* __restrict__ pit, const int8_t* __restrict__ psel,
                                 const double* __restrict__ target, const double* __restrict__ hl,
                                 const double* __restrict__ area, int8_t* __restrict__ sel,
                                 double* __restrict__ wd, double* __restrict__ wda) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int p = pit[i];
  const bool s = p >= 0 && psel[p];
  double w = 0.0;
  if (s) {
    const double d = target[p] - hl[i];
    w = d > 0.0 ? d : 0.0;
  }
  sel[i] = s ? 1 : 0;
  wd[i] = w;
  wda[i] = w * area[i];
}

// delta_init = (1 - bias) depo * wd / denom (denom > 0), plus at the inlets
// bias_at * depo / (count * area) (sedplex.py:873-898); ca[p] = (1 - bias)
// depo[p], cb[p] = bias_at[p] depo[p].
__global__ void k_sed_lake_init(int64_t N, const int32_t* __restrict__ pit, const int8_t* __restrict__ sel,
                                const int8_t* __restrict__ inlet, const double* __restrict__ ca,
                                const double* __restrict__ cb, const double* __restrict__ denom,
                                const double* __restrict__ count, const double* __restrict__ wd,
                                const double* __restrict__ area, double* __restrict__ delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double d = 0.0;
  const int p = pit[i];
  if (sel[i]) {
    const double dn = denom[p];
    if (dn > 0.0) d = ca[p] * wd[i] / (dn > 1.0e-30 ? dn : 1.0e-30);
    if (inlet[i] && count[p] > 0.0) d = d + cb[p] / (count[p] * area[i]);
  }
  delta[i] = d;
}

// x * area on the selected cells (per-pit volume sums)
__global__ void k_sed_vol(int64_t N, const double* __restrict__ x, const double* __restrict__ area,
                          double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  out[i] = x[i] * area[i];
}

// delta *= scale[pit] on the selected cells, 0 on the others (sedplex.py:
// 907-926)
__global__ void k_sed_scale(int64_t N, const int32_t* __restrict__ pit, const int8_t* __restrict__ sel,
                            const double* __restrict__ scale, double* __restrict__ delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  delta[i] = sel[i] ? delta[i] * scale[pit[i]] : delta[i] * 0.0;
}

// cumED += delta; h += delta (sedplex.py:1041-1046)
__global__ void k_sed_apply(int64_t N, const double* __restrict__ delta, double* __restrict__ h,
                            double* __restrict__ cumed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double d = delta[i];
  cumed[i] = cumed[i] + d;
  h[i] = h[i] + d;
}

__global__ void k_sed_const(int64_t n, double v, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// mean = sum / count where count > 0, else 0 (sedplex.py:637-641)
__global__ void k_sed_mean(int64_t n, const double* __restrict__ cnt, double* __restrict__ sum) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double c = cnt[p];
  sum[p] = c > 0.0 ? sum[p] / (c > 1.0 ? c : 1.0) : 0.0;
}

static int ensure_sed(gospl_mesh_t* m) {
  if (!m->sed_mask) GOSPL_CHECK(hipMalloc(&m->sed_mask, m->N + 64));
  if (!m->sed_buf) GOSPL_CHECK(hipMalloc(&m->sed_buf, sizeof(double) * 3 * (m->N + 64)));
  return 0;
}

}  // namespace gospl

using namespace gospl;

extern "C" {

int gospl_sed_source(gospl_mesh_t* m, const double* eb, double* b, gospl_stream_t s) {
  k_sed_source<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, eb, m->area, b);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_stray(gospl_mesh_t* m, const int8_t* lsink, const int32_t* pitids, int32_t closed, double* vsed,
                    double* closed_depo, gospl_stream_t s) {
  k_sed_stray<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, lsink, pitids, closed ? 1 : 0, vsed, closed_depo);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_add(gospl_mesh_t* m, const double* q, const double* hl, const double* sedfilled, double dt,
                  double* vnext, double* vsedl, gospl_stream_t s) {
  k_sed_add<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, q, hl, sedfilled, m->outlet, m->flat, dt, vnext, vsedl);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_closed(gospl_mesh_t* m, const double* closed_depo, double* h, double* cumed, gospl_stream_t s) {
  if (ensure_sed(m)) return 1;
  const int g = nblocks(m->N);
  k_sed_closed<<<g, GOSPL_BLOCK, 0, s>>>(m->N, closed_depo, m->area, m->sed_buf);
  k_sed_apply<<<g, GOSPL_BLOCK, 0, s>>>(m->N, m->sed_buf, h, cumed);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_bottomup(gospl_mesh_t* m, const int32_t* pitids, const double* target, const double* lfill,
                       const double* hl, double* delta, gospl_stream_t s) {
  k_sed_bottomup<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, pitids, target, lfill, hl, delta);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_tilt(gospl_mesh_t* m, const double* xyz, const int32_t* pitids, int64_t npit, const int8_t* psel,
                   const double* spill_xyz, double* delta, double* mean, gospl_stream_t s) {
  if (ensure_sed(m)) return 1;
  const int64_t N = m->N;
  const int g = nblocks(N);
  double* dist = m->sed_buf;
  double* ones = m->sed_buf + (N + 64);
  double* cnt = m->sed_buf + 2 * (N + 64);
  k_sed_tilt_dist<<<g, GOSPL_BLOCK, 0, s>>>(N, xyz, pitids, psel, spill_xyz, delta, m->sed_mask, dist);
  k_sed_const<<<g, GOSPL_BLOCK, 0, s>>>(N, 1.0, ones);
  GOSPL_CHECK(hipGetLastError());
  // per-pit sum of the distances and count of the tilted cells, then the mean
  if (gospl_pit_inflow(m, dist, m->sed_mask, pitids, npit, mean, s)) return 1;
  if (gospl_pit_inflow(m, ones, m->sed_mask, pitids, npit, cnt, s)) return 1;
  k_sed_mean<<<nblocks(npit > 0 ? npit : 1), GOSPL_BLOCK, 0, s>>>(npit, cnt, mean);
  k_sed_tilt_apply<<<g, GOSPL_BLOCK, 0, s>>>(N, pitids, m->sed_mask, dist, mean, delta);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_inlets(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcv, const int32_t* pitids, int64_t npit,
                     const int8_t* psel, int8_t* inlet, double* count, gospl_stream_t s) {
  if (ensure_sed(m)) return 1;
  const int64_t N = m->N;
  const int g = nblocks(N);
  double* ones = m->sed_buf + (N + 64);
  GOSPL_CHECK(hipMemsetAsync(inlet, 0, N, s));
  k_sed_inlets<<<g, GOSPL_BLOCK, 0, s>>>(N, nrcv, rcv, pitids, psel, inlet);
  k_sed_const<<<g, GOSPL_BLOCK, 0, s>>>(N, 1.0, ones);
  GOSPL_CHECK(hipGetLastError());
  return gospl_pit_inflow(m, ones, inlet, pitids, npit, count, s);
}

int gospl_sed_lake_depth(gospl_mesh_t* m, const int32_t* pitids, int64_t npit, const int8_t* psel,
                         const double* target, const double* hl, int8_t* sel, double* wd, double* denom,
                         gospl_stream_t s) {
  if (ensure_sed(m)) return 1;
  const int64_t N = m->N;
  double* wda = m->sed_buf;
  k_sed_lake_depth<<<nblocks(N), GOSPL_BLOCK, 0, s>>>(N, pitids, psel, target, hl, m->area, sel, wd, wda);
  GOSPL_CHECK(hipGetLastError());
  return gospl_pit_inflow(m, wda, sel, pitids, npit, denom, s);
}

int gospl_sed_lake_init(gospl_mesh_t* m, const int32_t* pitids, const int8_t* sel, const int8_t* inlet,
                        const double* ca, const double* cb, const double* denom, const double* count,
                        const double* wd, double* delta, gospl_stream_t s) {
  k_sed_lake_init<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, pitids, sel, inlet, ca, cb, denom, count, wd,
                                                        m->area, delta);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_pit_volume(gospl_mesh_t* m, const double* delta, const int8_t* sel, const int32_t* pitids,
                         int64_t npit, double* vol, gospl_stream_t s) {
  if (ensure_sed(m)) return 1;
  double* va = m->sed_buf;
  k_sed_vol<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, delta, m->area, va);
  GOSPL_CHECK(hipGetLastError());
  return gospl_pit_inflow(m, va, sel, pitids, npit, vol, s);
}

int gospl_sed_scale(gospl_mesh_t* m, const int32_t* pitids, const int8_t* sel, const double* scale, double* delta,
                    gospl_stream_t s) {
  k_sed_scale<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, pitids, sel, scale, delta);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_sed_apply(gospl_mesh_t* m, const double* delta, double* h, double* cumed, gospl_stream_t s) {
  k_sed_apply<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, delta, h, cumed);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
