// Solves on the receiver DAG: flow accumulation (I - W)^T q = b (top down,
// flowplex.py:429-478 matrix) and the implicit stream-power system
// (bottom up, SPL.py:45-114 matrix).
//
// Both matrices are triangular in a topological order of the receiver graph
// (receivers are strictly lower, or closer to a spill point on flats), so
// they are solved exactly by one sweep.  The sweep is level-synchronous in a
// persistent cooperative kernel:
//   * a donor list (reverse receiver graph, in FV neighbour order) is built
//     first so each vertex pulls its dependencies from a contiguous row;
//   * pending counters hold the number of unsolved dependencies; a vertex
//     whose counter reaches zero joins the next frontier (wavefront-
//     aggregated appends: one atomic per wavefront, not per vertex);
//   * one light grid barrier (atomic arrival counter + generation word)
//     separates levels; the grid is small because the sweep is bound by the
//     latency of each level, not by bandwidth.
// Every value is a fixed-order sum over its own row, so results do not depend
// on the schedule (no floating-point atomics).
#include <cstdlib>

#include <hipcub/hipcub.hpp>

#include "common.h"

namespace gospl {

__device__ __forceinline__ int load_count(int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// donors of i: neighbours j with rcv[k][j] == i and coef[k][j] != 0
__global__ void k_dep_count(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                            const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
                            const double* __restrict__ coef, int32_t* __restrict__ dcnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;
  if (i == N) { dcnt[N] = 0; return; }
  int c = 0;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    int j = nbr[(int64_t)p * N + i];
    if (j < 0) continue;
    for (int k = 0; k < NR; ++k)
      if (rcv[(int64_t)k * N + j] == (int)i && coef[(int64_t)k * N + j] != 0.0) c++;
  }
  dcnt[i] = c;
}

__global__ void k_dep_fill(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
                           const double* __restrict__ coef, const int32_t* __restrict__ doff,
                           int32_t* __restrict__ don, double* __restrict__ dval) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c = doff[i];
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    int j = nbr[(int64_t)p * N + i];
    if (j < 0) continue;
    for (int k = 0; k < NR; ++k) {
      if (rcv[(int64_t)k * N + j] == (int)i) {
        double w = coef[(int64_t)k * N + j];
        if (w != 0.0) { don[c] = j; dval[c] = w; c++; }
      }
    }
  }
}

// pending counters + frontier 0
__global__ void k_pending(int mode, int64_t N, int NR, const int32_t* __restrict__ rcv,
                          const double* __restrict__ coef, const int32_t* __restrict__ doff,
                          int32_t* __restrict__ pend, int32_t* __restrict__ front,
                          int32_t* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < N;
  int c = 0;
  if (valid) {
    if (mode == 0) {
      c = doff[i + 1] - doff[i];
    } else {
      for (int k = 0; k < NR; ++k) {
        int r = rcv[(int64_t)k * N + i];
        if (r != (int)i && coef[(int64_t)k * N + i] != 0.0) c++;
      }
    }
    pend[i] = c;
  }
  wave_append(valid && c == 0, (int)i, &counts[0], front);
}

// Grid barrier (all blocks co-resident: cooperative launch).  bar[0]
// arrivals, bar[1] generation, bar[2] abort.  A waiter that spins too long
// raises the abort flag instead of hanging; all blocks then leave the level
// loop and the host reports the failure.
__device__ __forceinline__ bool grid_barrier(unsigned int* bar, unsigned int nblocks) {
  __syncthreads();
  __shared__ int s_abort;
  if (threadIdx.x == 0) {
    const unsigned int gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned int arrived =
        __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    int ab = 0;
    if (arrived == nblocks - 1) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar[1], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1ull << 26) ||
            __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ab = 1;
          break;
        }
      }
    }
    __threadfence();
    s_abort = ab;
  }
  __syncthreads();
  return s_abort == 0;
}

template <int MODE>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_dag_level(
    int64_t N, int W, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ coef,
    const double* __restrict__ diag, const double* __restrict__ rhs, double* x,
    const int32_t* __restrict__ doff, const int32_t* __restrict__ don,
    const double* __restrict__ dval, int32_t* __restrict__ pend, int32_t* __restrict__ fa,
    int32_t* __restrict__ fb, int32_t* __restrict__ counts, unsigned int* __restrict__ bar,
    int32_t* __restrict__ nlev) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nth = (int64_t)gridDim.x * blockDim.x;
  int level = 0;
  for (;;) {
    const int n = load_count(&counts[level]);
    if (n == 0) break;
    int32_t* cur = (level & 1) ? fb : fa;
    int32_t* nxt = (level & 1) ? fa : fb;
    int32_t* cnt = &counts[level + 1];
    // uniform trip count per wavefront so the aggregated appends see all lanes
    for (int64_t t0 = tid - (tid & 63); t0 < n; t0 += nth) {
      const int64_t t = t0 + (tid & 63);
      const bool valid = t < n;
      const int v = valid ? cur[t] : 0;
      if (MODE == 0) {
        int r[GOSPL_MAXW];
        double w[GOSPL_MAXW];
        if (valid) {
          for (int k = 0; k < NR; ++k) {
            r[k] = rcv[(int64_t)k * N + v];
            w[k] = coef[(int64_t)k * N + v];
          }
          double s = rhs[v];
          const int e = doff[v + 1];
          for (int c = doff[v]; c < e; ++c) s = s + dval[c] * x[don[c]];
          x[v] = s;
        }
        for (int k = 0; k < NR; ++k) {
          bool rel = false;
          if (valid && r[k] != v && r[k] >= 0 && w[k] != 0.0) rel = atomicSub(&pend[r[k]], 1) == 1;
          wave_append(rel, valid ? r[k] : 0, cnt, nxt);
        }
      } else {
        int b = 0, e = 0;
        if (valid) {
          double s = rhs[v];
          for (int k = 0; k < NR; ++k) {
            int rr = rcv[(int64_t)k * N + v];
            double o = coef[(int64_t)k * N + v];
            if (rr != v && o != 0.0) s = s - o * x[rr];
          }
          x[v] = s / diag[v];
          b = doff[v];
          e = doff[v + 1];
        }
        for (int c = 0; c < W; ++c) {
          bool rel = false;
          int j = 0;
          if (valid && b + c < e) {
            j = don[b + c];
            rel = atomicSub(&pend[j], 1) == 1;
          }
          wave_append(rel, j, cnt, nxt);
        }
      }
    }
    if (!grid_barrier(bar, gridDim.x)) break;
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

// ---------------------------------------------------------------------------
// Tile-resident asynchronous sweep.
//
// Block b owns the Morton tile [b*S, min((b+1)*S, N)).  Dependencies inside
// the tile are counted in LDS and released with LDS atomics; a vertex whose
// in-tile counter reaches zero is queued in LDS if its out-of-tile counter
// (global, `ext`) is already zero.  Dependencies on other tiles are released
// with a global atomic carrying release semantics; the owning block finds
// such vertices by polling its boundary list (vertices that had out-of-tile
// dependencies) with acquire loads.  No grid barrier: a hop inside a tile
// costs a block barrier and a few L2 hits.  All blocks are co-resident
// (cooperative launch); a block that sees no global progress for a long
// time raises `abort` (then every block leaves and the host falls back to the
// level-synchronous sweep, which also handles cycles).
// ---------------------------------------------------------------------------

template <int MODE>
__global__ void k_tile_ext(int64_t N, int NR, int S, const int32_t* __restrict__ rcv,
                           const double* __restrict__ coef, const int32_t* __restrict__ doff,
                           const int32_t* __restrict__ don, int32_t* __restrict__ ext) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const int tv = (int)(v / S);
  int c = 0;
  if (MODE == 0) {
    for (int e = doff[v]; e < doff[v + 1]; ++e)
      if (don[e] / S != tv) c++;
  } else {
    for (int k = 0; k < NR; ++k) {
      int r = rcv[(int64_t)k * N + v];
      if (r != (int)v && coef[(int64_t)k * N + v] != 0.0 && r / S != tv) c++;
    }
  }
  ext[v] = c;
}

template <int MODE>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_dag_tile(
    int64_t N, int NR, int S, const int32_t* __restrict__ rcv, const double* __restrict__ coef,
    const double* __restrict__ diag, const double* __restrict__ rhs, double* x,
    const int32_t* __restrict__ doff, const int32_t* __restrict__ don,
    const double* __restrict__ dval, int32_t* ext, unsigned int* __restrict__ ctl) {
  // ctl[0] global progress (vertices done), ctl[1] abort
  extern __shared__ int smem[];
  int* pend = smem;                                           // S ints
  short* q0 = reinterpret_cast<short*>(pend + S);             // S shorts
  short* q1 = q0 + S;                                         // S shorts
  short* bnd = q1 + S;                                        // S shorts (boundary list)
  __shared__ int s_qn[2], s_nb, s_done, s_stop;
  const int t0 = blockIdx.x * S;
  const int n = (int)((int64_t)t0 + S <= N ? S : N - t0);
  if (threadIdx.x == 0) { s_qn[0] = 0; s_qn[1] = 0; s_nb = 0; s_done = 0; s_stop = 0; }
  __syncthreads();
  // in-tile dependency counts, initial queue, boundary list
  for (int l = threadIdx.x; l < n; l += blockDim.x) {
    const int v = t0 + l;
    int c = 0;
    if (MODE == 0) {
      for (int e = doff[v]; e < doff[v + 1]; ++e) {
        int j = don[e];
        if (j >= t0 && j < t0 + n) c++;
      }
    } else {
      for (int k = 0; k < NR; ++k) {
        int r = rcv[(int64_t)k * N + v];
        if (r != v && coef[(int64_t)k * N + v] != 0.0 && r >= t0 && r < t0 + n) c++;
      }
    }
    const int e = ext[v];
    pend[l] = c;
    if (e > 0) bnd[atomicAdd(&s_nb, 1)] = (short)l;
    if (c == 0 && e == 0) {
      pend[l] = -1000000;
      q0[atomicAdd(&s_qn[0], 1)] = (short)l;
    }
  }
  __syncthreads();
  int cur = 0;
  unsigned int last_global = 0;
  int idle = 0;
  while (true) {
    __syncthreads();
    if (s_done >= n || s_stop) break;
    short* qc = cur ? q1 : q0;
    short* qn = cur ? q0 : q1;
    if (s_qn[cur] == 0) {
      // poll the boundary list for vertices released by other tiles
      for (int b = threadIdx.x; b < s_nb; b += blockDim.x) {
        const int l = bnd[b];
        if (l < 0) continue;
        if (pend[l] == 0 &&
            __hip_atomic_load(&ext[t0 + l], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
          if (atomicCAS(&pend[l], 0, -1000000) == 0) {
            qc[atomicAdd(&s_qn[cur], 1)] = (short)l;
            bnd[b] = -1;
          }
        }
      }
      __syncthreads();
      if (s_qn[cur] == 0) {
        if (threadIdx.x == 0) {
          const unsigned int gp = __hip_atomic_load(&ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (gp != last_global) { last_global = gp; idle = 0; } else { ++idle; }
          if (idle > (1 << 18) || __hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_store(&ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_stop = 1;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        continue;
      }
    }
    const int m = s_qn[cur];
    for (int it = threadIdx.x; it < m; it += blockDim.x) {
      const int l = qc[it];
      const int v = t0 + l;
      double s = rhs[v];
      if (MODE == 0) {
        const int e = doff[v + 1];
        for (int c = doff[v]; c < e; ++c) s = s + dval[c] * x[don[c]];
        x[v] = s;
        for (int k = 0; k < NR; ++k) {
          const int r = rcv[(int64_t)k * N + v];
          if (r == v || r < 0 || coef[(int64_t)k * N + v] == 0.0) continue;
          if (r >= t0 && r < t0 + n) {
            if (atomicSub(&pend[r - t0], 1) == 1 &&
                __hip_atomic_load(&ext[r], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                atomicCAS(&pend[r - t0], 0, -1000000) == 0)
              qn[atomicAdd(&s_qn[cur ^ 1], 1)] = (short)(r - t0);
          } else {
            __hip_atomic_fetch_add(&ext[r], -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      } else {
        for (int k = 0; k < NR; ++k) {
          const int r = rcv[(int64_t)k * N + v];
          const double o = coef[(int64_t)k * N + v];
          if (r != v && o != 0.0) s = s - o * x[r];
        }
        x[v] = s / diag[v];
        const int e = doff[v + 1];
        for (int c = doff[v]; c < e; ++c) {
          const int j = don[c];
          if (j >= t0 && j < t0 + n) {
            if (atomicSub(&pend[j - t0], 1) == 1 &&
                __hip_atomic_load(&ext[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                atomicCAS(&pend[j - t0], 0, -1000000) == 0)
              qn[atomicAdd(&s_qn[cur ^ 1], 1)] = (short)(j - t0);
          } else {
            __hip_atomic_fetch_add(&ext[j], -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_done += m;
      s_qn[cur] = 0;
      __hip_atomic_fetch_add(&ctl[0], (unsigned int)m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      idle = 0;
    }
    cur ^= 1;
  }
}

static int ensure_donors(gospl_mesh* m, int64_t cap) {
  if (m->wdon_cap >= cap) return 0;
  if (m->wdon) GOSPL_CHECK(hipFree(m->wdon));
  if (m->wdval) GOSPL_CHECK(hipFree(m->wdval));
  m->wdon = nullptr;
  m->wdval = nullptr;
  GOSPL_CHECK(hipMalloc(&m->wdon, sizeof(int32_t) * (cap + 64)));
  GOSPL_CHECK(hipMalloc(&m->wdval, sizeof(double) * (cap + 64)));
  m->wdon_cap = cap;
  return 0;
}

static int dag_solve(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                     const double* diag, const double* rhs, double* x, int32_t* nlevels,
                     int64_t* unsolved, hipStream_t s) {
  const int64_t N = m->N;
  const int g = nblocks(N);
  int32_t* pend = m->wi[0];
  int32_t* fa = m->wi[1];
  int32_t* fb = m->wi[2];
  int32_t* dcnt = m->wi[4];
  int32_t* doff = m->wi[5];
  int32_t* counts = m->counts;
  int32_t* dev = m->dflag;
  if (ensure_donors(m, (int64_t)nrcv * N)) return 1;
  GOSPL_CHECK(hipMemsetAsync(counts, 0, sizeof(int32_t) * (N + 2), s));
  GOSPL_CHECK(hipMemsetAsync(dev, 0, sizeof(int32_t) * 4, s));
  unsigned int* bar = reinterpret_cast<unsigned int*>(dev + 32);
  GOSPL_CHECK(hipMemsetAsync(bar, 0, sizeof(unsigned int) * 4, s));
  // donor rows
  k_dep_count<<<nblocks(N + 1), GOSPL_BLOCK, 0, s>>>(N, m->W, nrcv, m->nbr, m->deg, rcv, coef, dcnt);
  size_t bytes = m->cub_tmp_bytes;
  GOSPL_CHECK(hipcub::DeviceScan::ExclusiveSum(m->cub_tmp, bytes, dcnt, doff, (int)(N + 1), s));
  k_dep_fill<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, nrcv, m->nbr, m->deg, rcv, coef, doff, m->wdon, m->wdval);
  int64_t NN = N;
  int W = m->W;
  int NR = nrcv;
  int32_t* nlev = dev;
  int32_t* don = m->wdon;
  double* dval = m->wdval;
  if (m->tile_S > 0) {
    // tile-resident asynchronous sweep
    int S = m->tile_S;
    int32_t* ext = pend;
    unsigned int* ctl = reinterpret_cast<unsigned int*>(dev + 48);
    GOSPL_CHECK(hipMemsetAsync(ctl, 0, sizeof(unsigned int) * 4, s));
    if (mode == 0)
      k_tile_ext<0><<<g, GOSPL_BLOCK, 0, s>>>(N, NR, S, rcv, coef, doff, don, ext);
    else
      k_tile_ext<1><<<g, GOSPL_BLOCK, 0, s>>>(N, NR, S, rcv, coef, doff, don, ext);
    GOSPL_CHECK(hipGetLastError());
    void* targs[] = {&NN, &NR, &S, (void*)&rcv, (void*)&coef, (void*)&diag, (void*)&rhs, &x, &doff,
                     &don, &dval, &ext, &ctl};
    const void* tk = mode == 0 ? (const void*)k_dag_tile<0> : (const void*)k_dag_tile<1>;
    const int nt = (int)((N + S - 1) / S);
    if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[0], s));
    GOSPL_CHECK(hipLaunchCooperativeKernel(tk, dim3(nt), dim3(GOSPL_BLOCK), targs,
                                           (size_t)10 * S, s));
    if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[1], s));
    GOSPL_CHECK(hipMemcpyAsync(m->hflag + 40, ctl, sizeof(unsigned int) * 2, hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    if (m->hflag[41] == 0 && m->hflag[40] == (int32_t)N) {
      if (m->timing) {
        float ms = 0.0f;
        GOSPL_CHECK(hipEventElapsedTime(&ms, m->tev[0], m->tev[1]));
        m->t_ms[mode] += ms;
        m->t_calls[mode] += 1;
      }
      if (nlevels) *nlevels = -1;
      if (unsolved) *unsolved = 0;
      return 0;
    }
    // no progress (a cycle in the receivers): fall through to the level sweep
  }
  k_pending<<<g, GOSPL_BLOCK, 0, s>>>(mode, N, nrcv, rcv, coef, doff, pend, fa, counts);
  GOSPL_CHECK(hipGetLastError());
  void* args[] = {&NN, &W, &NR, (void*)&rcv, (void*)&coef, (void*)&diag, (void*)&rhs, &x, &doff,
                  &don, &dval, &pend, &fa, &fb, &counts, &bar, &nlev};
  const void* kern = mode == 0 ? (const void*)k_dag_level<0> : (const void*)k_dag_level<1>;
  if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[0], s));
  GOSPL_CHECK(hipLaunchCooperativeKernel(kern, dim3(m->coop_blocks), dim3(GOSPL_BLOCK), args, 0, s));
  if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[1], s));
  k_dag_unsolved<<<g, GOSPL_BLOCK, 0, s>>>(N, pend, x, dev + 1);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(m->hflag, dev, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipMemcpyAsync(m->hflag + 34, bar + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  if (m->hflag[34]) {
    set_error("receiver-DAG sweep aborted: a grid barrier did not complete");
    return 1;
  }
  if (m->timing) {
    float ms = 0.0f;
    GOSPL_CHECK(hipEventElapsedTime(&ms, m->tev[0], m->tev[1]));
    m->t_ms[mode] += ms;
    m->t_calls[mode] += 1;
  }
  if (nlevels) *nlevels = m->hflag[0];
  if (unsolved) *unsolved = m->hflag[1];
  return 0;
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_dag_init_occupancy(gospl_mesh* m) {
  int per_cu = 1 << 30;
  const void* ks[2] = {(const void*)k_dag_level<0>, (const void*)k_dag_level<1>};
  for (int i = 0; i < 2; ++i) {
    int b = 0;
    GOSPL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, ks[i], GOSPL_BLOCK, 0));
    per_cu = b < per_cu ? b : per_cu;
  }
  if (per_cu < 1) per_cu = 1;
  const int maxb = per_cu * m->num_cu;
  // Few blocks: the sweep is bound by per-level latency (dependent loads and
  // the grid barrier, ~3 us at 32 blocks vs ~18 us at 256 on MI355X), not by
  // bandwidth.  Measured best on the 2.6M-vertex sphere: 32.
  int want = 32;
  const char* env = getenv("GOSPL_DAG_BLOCKS");
  if (env && atoi(env) > 0) want = atoi(env);
  m->coop_blocks = want < maxb ? want : maxb;

  // tile size for the tile-resident sweep: one tile per co-resident block,
  // LDS = 10 bytes per tile vertex, tile-local indices in 16 bits
  m->tile_S = 0;
  const char* te = getenv("GOSPL_DAG_TILE");
  if (!(te && atoi(te) == 0)) {
    int64_t S = (m->N + m->num_cu - 1) / m->num_cu;
    if (S < 1024) S = 1024;
    if (S <= 16000) {
      const size_t lds = (size_t)10 * S;
      GOSPL_CHECK(hipFuncSetAttribute((const void*)k_dag_tile<0>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      GOSPL_CHECK(hipFuncSetAttribute((const void*)k_dag_tile<1>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      int b0 = 0, b1 = 0;
      GOSPL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, (const void*)k_dag_tile<0>,
                                                               GOSPL_BLOCK, lds));
      GOSPL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, (const void*)k_dag_tile<1>,
                                                               GOSPL_BLOCK, lds));
      const int64_t nt = (m->N + S - 1) / S;
      const int64_t cap = (int64_t)(b0 < b1 ? b0 : b1) * m->num_cu;
      if (nt <= cap) m->tile_S = (int)S;
    }
  }
  return 0;
}

extern "C" {

int gospl_sweep_timing(gospl_mesh_t* m, int32_t enable, double* fa_ms, int64_t* fa_calls,
                       double* spl_ms, int64_t* spl_calls) {
  if (enable && !m->tev[0]) {
    GOSPL_CHECK(hipEventCreate(&m->tev[0]));
    GOSPL_CHECK(hipEventCreate(&m->tev[1]));
  }
  if (fa_ms) *fa_ms = m->t_ms[0];
  if (fa_calls) *fa_calls = m->t_calls[0];
  if (spl_ms) *spl_ms = m->t_ms[1];
  if (spl_calls) *spl_calls = m->t_calls[1];
  m->t_ms[0] = m->t_ms[1] = 0.0;
  m->t_calls[0] = m->t_calls[1] = 0;
  m->timing = enable ? 1 : 0;
  return 0;
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
