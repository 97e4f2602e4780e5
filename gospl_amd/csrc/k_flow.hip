// Receivers (MFD) and the two receiver-DAG solves (flow accumulation, top
// down; stream-power erosion, bottom up).
//
// Receivers: one thread per vertex; the <= 12 neighbour candidates live in
// registers (loops fully unrolled over the 12-slot stencil), the top-nrcv
// selection ranks candidates by pairwise comparison.  goSPL orders them with
// a Fortran quicksort (functions.F90:683-762) whose result is unique when
// the perturbed slopes are distinct; when two candidates carry exactly equal
// values the kernel replays that quicksort on a private copy so the tie order
// is the reference's.
//
// DAG solves: a persistent cooperative kernel walks the DAG level by level.
// Per level every thread takes frontier vertices, computes their value by
// pulling from already-solved neighbours in the FV neighbour order (so the
// sums are deterministic, no float atomics), then decrements the pending
// counters of the vertices that depend on it; a counter reaching zero appends
// that vertex to the next frontier.  One grid barrier per level.
#include <hip/hip_cooperative_groups.h>

#include "common.h"

namespace cg = cooperative_groups;

namespace gospl {

// ---------------------------------------------------------------------------
// goSPL's quicksort/split on 1-based arrays of 13 (functions.F90:683-762),
// iterative form; used only when perturbed slopes tie exactly.
// ---------------------------------------------------------------------------
__device__ void f_split(double* a, int low, int high, int* mid, int* idx) {
  int left = low, right = high;
  double pivot = a[low];
  int ipivot = idx[low];
  for (;;) {
    if (left >= right) break;
    for (;;) {
      if (left >= right || a[right] < pivot) break;
      right--;
    }
    for (;;) {
      if (left > 12) break;
      if (a[left] > pivot) break;
      left++;
    }
    if (left < right) {
      double s = a[left]; a[left] = a[right]; a[right] = s;
      int t = idx[left]; idx[left] = idx[right]; idx[right] = t;
    }
  }
  a[low] = a[right];
  a[right] = pivot;
  *mid = right;
  idx[low] = idx[right];
  idx[right] = ipivot;
}

__device__ void f_quicksort(double* a, int first, int last, int* idx) {
  int stk[64];
  int sp = 0;
  stk[sp++] = first;
  stk[sp++] = last;
  while (sp > 0) {
    int l = stk[--sp];
    int f = stk[--sp];
    if (f < l) {
      int mid;
      f_split(a, f, l, &mid, idx);
      // recursion order: (f, mid-1) then (mid+1, l); push second first
      stk[sp++] = mid + 1;
      stk[sp++] = l;
      stk[sp++] = f;
      stk[sp++] = mid - 1;
    }
  }
}

// ---------------------------------------------------------------------------
// Receivers (functions.F90:1544-1671) + _buildFlowDirection post-processing
// (flowplex.py:493-540).
// ---------------------------------------------------------------------------
template <int NR>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_receivers(
    int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
    const double* __restrict__ elgt, const int32_t* __restrict__ gid,
    const int8_t* __restrict__ outlet, int flat, double fexp0, const double* __restrict__ h,
    double sl, const double* __restrict__ lfill, int post, int down,
    const int32_t* __restrict__ pitids, const int32_t* __restrict__ flatdirs,
    int32_t* __restrict__ rcv, double* __restrict__ dist, double* __restrict__ wgt,
    int8_t* __restrict__ lsink) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int me = (int)i;
  const double hk = h[i];
  int r_out[NR];
  double d_out[NR], w_out[NR];
#pragma unroll
  for (int p = 0; p < NR; ++p) { r_out[p] = me; d_out[p] = 0.0; w_out[p] = 0.0; }

  if (!(hk <= sl)) {
    const int dg = deg[i];
    int nb[GOSPL_MAXW];
    double L[GOSPL_MAXW], hn[GOSPL_MAXW];
    double e = 0.0;
#pragma unroll
    for (int p = 0; p < GOSPL_MAXW; ++p) {
      nb[p] = -1; L[p] = 0.0; hn[p] = 0.0;
      if (p < W && p < dg) {
        nb[p] = nbr[(int64_t)p * N + i];
        L[p] = elgt[(int64_t)p * N + i];
        if (nb[p] >= 0) hn[p] = h[nb[p]];
        if (nb[p] >= 0 && L[p] > 0.0) {
          double v = (hk - hn[p]) / L[p];
          e = (v > e) ? v : e;
        }
      }
    }
    const double fexp = (e > 0.0) ? 5.0 * (e < 1.0 ? e : 1.0) + fexp0 : 1.0;
    // candidates in neighbour order
    double slp[GOSPL_MAXW], dst[GOSPL_MAXW];
    int id[GOSPL_MAXW];
    int kk = 0;
#pragma unroll
    for (int p = 0; p < GOSPL_MAXW; ++p) {
      slp[p] = 0.0; dst[p] = 0.0; id[p] = 0;
    }
#pragma unroll
    for (int p = 0; p < GOSPL_MAXW; ++p) {
      if (nb[p] >= 0 && L[p] > 0.0) {
        double val = pow(hk - hn[p], fexp) / L[p];
        if (val > 0.0) {
          double s = val * (1.0 - 1.0e-15 * (double)gid[nb[p]]);
          // compact into slot kk (kk <= p) without dynamic register indexing
#pragma unroll
          for (int q = 0; q <= p; ++q) {
            if (q == kk) { slp[q] = s; id[q] = nb[p]; dst[q] = L[p]; }
          }
          kk++;
        }
      }
    }
    if (kk > 0 && kk <= NR) {
      double val = 0.0;
#pragma unroll
      for (int p = 0; p < NR; ++p) {
        if (p < kk) { r_out[p] = id[p]; d_out[p] = dst[p]; val = val + slp[p]; }
      }
#pragma unroll
      for (int p = 0; p < NR; ++p) w_out[p] = slp[p] / val;
    } else if (kk > NR) {
      // distinct values? (ties need the reference's quicksort order)
      bool tie = false;
#pragma unroll
      for (int p = 0; p < GOSPL_MAXW; ++p)
#pragma unroll
        for (int q = p + 1; q < GOSPL_MAXW; ++q)
          if (q < kk && slp[p] == slp[q]) tie = true;
      // dist takes the unsorted candidate lengths at sorted positions kk..kk-NR+1
      double dd[NR];
#pragma unroll
      for (int n = 0; n < NR; ++n) {
        dd[n] = 0.0;
#pragma unroll
        for (int q = 0; q < GOSPL_MAXW; ++q)
          if (q == kk - 1 - n) dd[n] = dst[q];
      }
      double sl_sorted[NR];
      int id_sorted[NR];
      if (!tie) {
#pragma unroll
        for (int n = 0; n < NR; ++n) { sl_sorted[n] = 0.0; id_sorted[n] = me; }
#pragma unroll
        for (int p = 0; p < GOSPL_MAXW; ++p) {
          if (p < kk) {
            int rank = 0;
#pragma unroll
            for (int q = 0; q < GOSPL_MAXW; ++q)
              if (q < kk && slp[q] > slp[p]) rank++;
#pragma unroll
            for (int n = 0; n < NR; ++n)
              if (rank == n) { sl_sorted[n] = slp[p]; id_sorted[n] = id[p]; }
          }
        }
      } else {
        double a[13];
        int ix[13];
        a[0] = 0.0; ix[0] = 0;
        for (int p = 0; p < 12; ++p) { a[p + 1] = slp[p]; ix[p + 1] = id[p]; }
        f_quicksort(a, 1, kk, ix);
        for (int n = 0; n < NR; ++n) { sl_sorted[n] = a[kk - n]; id_sorted[n] = ix[kk - n]; }
      }
      double val = 0.0;
#pragma unroll
      for (int n = 0; n < NR; ++n) {
        r_out[n] = id_sorted[n];
        d_out[n] = dd[n];
        val = val + sl_sorted[n];
      }
#pragma unroll
      for (int n = 0; n < NR; ++n) w_out[n] = sl_sorted[n] / val;
    }
  }

  if (post) {
    double sw = w_out[0];
#pragma unroll
    for (int p = 1; p < NR; ++p) sw = sw + w_out[p];
    if (down && sw == 0.0 && hk == lfill[i] && pitids[i] > -1 && flatdirs[i] > -1) {
#pragma unroll
      for (int p = 0; p < NR; ++p) { r_out[p] = me; w_out[p] = 0.0; }
      r_out[0] = flatdirs[i];
      w_out[0] = 1.0;
    }
    const bool is_out = flat && outlet[i];
    if (is_out) {
#pragma unroll
      for (int p = 0; p < NR; ++p) { r_out[p] = me; d_out[p] = 0.0; w_out[p] = 0.0; }
    }
    sw = w_out[0];
#pragma unroll
    for (int p = 1; p < NR; ++p) sw = sw + w_out[p];
    lsink[i] = (sw == 0.0 && !is_out && !(lfill[i] <= sl)) ? 1 : 0;
  }
#pragma unroll
  for (int p = 0; p < NR; ++p) {
    rcv[(int64_t)p * N + i] = r_out[p];
    dist[(int64_t)p * N + i] = d_out[p];
    wgt[(int64_t)p * N + i] = w_out[p];
  }
}

template <int NR>
static void launch_receivers(gospl_mesh* m, double fexp, const double* h, double sl,
                             const double* lfill, int post, int down, const int32_t* pitids,
                             const int32_t* flatdirs, int32_t* rcv, double* dist, double* wgt,
                             int8_t* lsink, hipStream_t s) {
  k_receivers<NR><<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(
      m->N, m->W, m->nbr, m->deg, m->elgt, m->gid, m->outlet, m->flat, fexp, h, sl, lfill, post,
      down, pitids, flatdirs, rcv, dist, wgt, lsink);
}

static int receivers_dispatch(gospl_mesh* m, int nrcv, double fexp, const double* h, double sl,
                              const double* lfill, int post, int down, const int32_t* pitids,
                              const int32_t* flatdirs, int32_t* rcv, double* dist, double* wgt,
                              int8_t* lsink, hipStream_t s) {
  switch (nrcv) {
#define GOSPL_RCASE(K) \
  case K: launch_receivers<K>(m, fexp, h, sl, lfill, post, down, pitids, flatdirs, rcv, dist, wgt, lsink, s); break;
    GOSPL_RCASE(1) GOSPL_RCASE(2) GOSPL_RCASE(3) GOSPL_RCASE(4) GOSPL_RCASE(5) GOSPL_RCASE(6)
    GOSPL_RCASE(7) GOSPL_RCASE(8) GOSPL_RCASE(9) GOSPL_RCASE(10) GOSPL_RCASE(11) GOSPL_RCASE(12)
#undef GOSPL_RCASE
    default:
      set_error("nrcv must be in 1..12 (got %d)", nrcv);
      return 1;
  }
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// DAG solves
// ---------------------------------------------------------------------------

// Flow accumulation: pending[i] = number of donors j (neighbours of i that
// list i as a receiver with non-zero weight).
__global__ void k_fa_pending(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                             const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
                             const double* __restrict__ wgt, int32_t* __restrict__ pend,
                             int32_t* __restrict__ front, int32_t* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c = 0;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    int j = nbr[(int64_t)p * N + i];
    if (j < 0) continue;
    for (int k = 0; k < NR; ++k)
      if (rcv[(int64_t)k * N + j] == (int)i && wgt[(int64_t)k * N + j] != 0.0) c++;
  }
  pend[i] = c;
  if (c == 0) front[atomicAdd(&counts[0], 1)] = (int)i;
}

// SPL: pending[i] = number of receivers k with a non-zero coefficient.
__global__ void k_spl_pending(int64_t N, int NR, const int32_t* __restrict__ rcv,
                              const double* __restrict__ off, int32_t* __restrict__ pend,
                              int32_t* __restrict__ front, int32_t* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c = 0;
  for (int k = 0; k < NR; ++k) {
    int r = rcv[(int64_t)k * N + i];
    if (r != (int)i && off[(int64_t)k * N + i] != 0.0) c++;
  }
  pend[i] = c;
  if (c == 0) front[atomicAdd(&counts[0], 1)] = (int)i;
}

__device__ __forceinline__ int load_count(int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// mode 0: flow accumulation  q_i = b_i + sum_{donors j} w_j q_j, then release receivers
// mode 1: SPL                x_i = (rhs_i - sum_k off_k x_rk) / diag_i, then release donors
__global__ __launch_bounds__(GOSPL_BLOCK) void k_dag_sweep(
    int mode, int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
    const double* __restrict__ coef, const double* __restrict__ diag,
    const double* __restrict__ rhs, double* __restrict__ x, int32_t* __restrict__ pend,
    int32_t* __restrict__ fa, int32_t* __restrict__ fb, int32_t* __restrict__ counts,
    int32_t* __restrict__ nlev) {
  cg::grid_group grid = cg::this_grid();
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nth = (int64_t)gridDim.x * blockDim.x;
  int level = 0;
  for (;;) {
    const int n = load_count(&counts[level]);
    if (n == 0) break;
    int32_t* cur = (level & 1) ? fb : fa;
    int32_t* nxt = (level & 1) ? fa : fb;
    for (int64_t t = tid; t < n; t += nth) {
      const int i = cur[t];
      const int dg = deg[i];
      if (mode == 0) {
        double s = rhs[i];
        for (int p = 0; p < W && p < dg; ++p) {
          int j = nbr[(int64_t)p * N + i];
          if (j < 0) continue;
          for (int k = 0; k < NR; ++k) {
            if (rcv[(int64_t)k * N + j] == i) {
              double w = coef[(int64_t)k * N + j];
              if (w != 0.0) s = s + w * x[j];
            }
          }
        }
        x[i] = s;
        for (int k = 0; k < NR; ++k) {
          int r = rcv[(int64_t)k * N + i];
          if (r != i && r >= 0 && coef[(int64_t)k * N + i] != 0.0) {
            if (atomicSub(&pend[r], 1) == 1) nxt[atomicAdd(&counts[level + 1], 1)] = r;
          }
        }
      } else {
        double s = rhs[i];
        for (int k = 0; k < NR; ++k) {
          int r = rcv[(int64_t)k * N + i];
          double o = coef[(int64_t)k * N + i];
          if (r != i && o != 0.0) s = s - o * x[r];
        }
        x[i] = s / diag[i];
        for (int p = 0; p < W && p < dg; ++p) {
          int j = nbr[(int64_t)p * N + i];
          if (j < 0) continue;
          for (int k = 0; k < NR; ++k) {
            if (rcv[(int64_t)k * N + j] == i && coef[(int64_t)k * N + j] != 0.0) {
              if (atomicSub(&pend[j], 1) == 1) nxt[atomicAdd(&counts[level + 1], 1)] = j;
            }
          }
        }
      }
    }
    grid.sync();
    level++;
  }
  if (tid == 0) *nlev = level;
}

// Vertices never reached (a cycle in the receivers) -> 0, counted.
__global__ void k_dag_unsolved(int64_t N, const int32_t* __restrict__ pend, double* __restrict__ x,
                               int32_t* __restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (pend[i] > 0) {
    x[i] = 0.0;
    atomicAdd(cnt, 1);
  }
}

// SPL coefficients (SPL.py:56-108): off[k] = -(Kbr*limiter/dist)*w (0 for
// self receivers), diag = ((1 - off0) - off1) - ...
__global__ void k_spl_coef(int64_t N, int NR, const int32_t* __restrict__ rcv,
                           const double* __restrict__ wgt, const double* __restrict__ dist,
                           const double* __restrict__ fal, const double* __restrict__ kfac,
                           double K, double m, double dt, double sl,
                           const double* __restrict__ lfill, const double* __restrict__ h,
                           double* __restrict__ off, double* __restrict__ diag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool sea = lfill[i] <= sl;
  double kbr = K * 1.0;
  if (kfac) kbr = kbr * kfac[i];
  kbr = kbr * (dt * pow(fal[i], m));
  if (sea) kbr = 0.0;
  const double hi = h[i];
  double d = 1.0;
  for (int k = 0; k < NR; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    const double w = sea ? 0.0 : wgt[(int64_t)k * N + i];
    const double dh = hi - h[r];
    const double lim = (dh != 0.0) ? dh / (dh + 1.0e-2) : 0.0;
    const double dk = dist[(int64_t)k * N + i];
    double data = (dk != 0.0) ? (kbr * lim) / dk : 0.0;
    data = data * -w;
    if (r == (int)i) data = 0.0;
    off[(int64_t)k * N + i] = data;
    d = d - data;
  }
  diag[i] = d;
}

// Eb and the erodepSPL state update (SPL.py:345-353, 422-438).
__global__ void k_spl_update(int64_t N, const double* __restrict__ x, const int8_t* __restrict__ outlet,
                             int flat, const int8_t* __restrict__ lsink, double dt,
                             double* __restrict__ h, double* __restrict__ cumed,
                             double* __restrict__ eb, double* __restrict__ eblocal) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double hold = h[i];
  double e = (x[i] - hold) / dt;
  if ((flat && outlet[i]) || lsink[i]) e = 0.0;
  eb[i] = e;
  const double t = e * dt;
  cumed[i] = cumed[i] + t;
  h[i] = hold + t;
  eblocal[i] = t / dt;
}

static int dag_solve(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                     const double* diag, const double* rhs, double* x, int32_t* nlevels,
                     int64_t* unsolved, hipStream_t s) {
  int32_t* pend = m->wi[0];
  int32_t* fa = m->wi[1];
  int32_t* fb = m->wi[2];
  int32_t* counts = m->counts;
  int32_t* dev = m->dflag;
  GOSPL_CHECK(hipMemsetAsync(counts, 0, sizeof(int32_t) * (m->N + 2), s));
  GOSPL_CHECK(hipMemsetAsync(dev, 0, sizeof(int32_t) * 4, s));
  if (mode == 0)
    k_fa_pending<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, m->W, nrcv, m->nbr, m->deg, rcv, coef,
                                                       pend, fa, counts);
  else
    k_spl_pending<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, nrcv, rcv, coef, pend, fa, counts);
  GOSPL_CHECK(hipGetLastError());
  int64_t N = m->N;
  int W = m->W;
  int NR = nrcv;
  int32_t* nlev = dev;
  void* args[] = {&mode, &N, &W, &NR, (void*)&m->nbr, (void*)&m->deg, (void*)&rcv, (void*)&coef,
                  (void*)&diag, (void*)&rhs, &x, &pend, &fa, &fb, &counts, &nlev};
  GOSPL_CHECK(hipLaunchCooperativeKernel((const void*)k_dag_sweep, dim3(m->coop_blocks),
                                         dim3(GOSPL_BLOCK), args, 0, s));
  k_dag_unsolved<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, pend, x, dev + 1);
  GOSPL_CHECK(hipGetLastError());
  if (nlevels || unsolved) {
    GOSPL_CHECK(hipMemcpyAsync(m->hflag, dev, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    if (nlevels) *nlevels = m->hflag[0];
    if (unsolved) *unsolved = m->hflag[1];
  }
  return 0;
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_dag_init_occupancy(gospl_mesh* m) {
  int per_cu = 0;
  GOSPL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_dag_sweep,
                                                           GOSPL_BLOCK, 0));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 4) per_cu = 4;
  m->coop_blocks = per_cu * m->num_cu;
  return 0;
}

extern "C" {

int gospl_mfdreceivers(gospl_mesh_t* m, int32_t nrcv, double flowexp, const double* elev,
                       double sealevel, int32_t* rcv, double* dist, double* wgt,
                       gospl_stream_t stream) {
  return receivers_dispatch(m, nrcv, flowexp, elev, sealevel, nullptr, 0, 0, nullptr, nullptr, rcv,
                            dist, wgt, nullptr, stream);
}

int gospl_flow_direction(gospl_mesh_t* m, int32_t nrcv, double flowexp, const double* h,
                         double sealevel, const double* lfill, int32_t down,
                         const int32_t* pitids, const int32_t* flatdirs, int32_t* rcv,
                         double* dist, double* wgt, int8_t* lsink, gospl_stream_t stream) {
  if (down && (!pitids || !flatdirs)) {
    set_error("gospl_flow_direction: down=1 needs pitids and flatdirs");
    return 1;
  }
  return receivers_dispatch(m, nrcv, flowexp, h, sealevel, lfill, 1, down, pitids, flatdirs, rcv,
                            dist, wgt, lsink, stream);
}

int gospl_flowacc(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcv, const double* wgt,
                  const double* b, double* q, int32_t* nlevels, int64_t* unsolved,
                  gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_flowacc: nrcv must be in 1..12");
    return 1;
  }
  return dag_solve(m, 0, nrcv, rcv, wgt, nullptr, b, q, nlevels, unsolved, stream);
}

int gospl_spl(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti,
              const double* disti, const double* fal, const double* kfac, double K, double mexp,
              double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
              double* cumed, double* eb, double* eblocal, int32_t* nlevels,
              gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_spl: nrcv must be in 1..12");
    return 1;
  }

  if (gospl_ensure_coef(m, nrcv)) return 1;
  double* offc = m->wcoef;
  double* diag = m->wd[3];
  double* x = m->wd[2];
  k_spl_coef<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, nrcv, rcvi, wgti, disti, fal, kfac, K,
                                                        mexp, dt, sealevel, lfill, h, offc, diag);
  GOSPL_CHECK(hipGetLastError());
  int rc = dag_solve(m, 1, nrcv, rcvi, offc, diag, h, x, nlevels, nullptr, stream);
  if (rc) return rc;
  k_spl_update<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, x, m->outlet, m->flat, lsink, dt, h,
                                                          cumed, eb, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
