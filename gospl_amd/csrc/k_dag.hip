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
//     whose counter reaches zero joins the next frontier;
//   * one light grid barrier (atomic arrival counter + generation word)
//     separates levels;
//   * optional chaining (GOSPL_DAG_CHAIN=1): the thread that releases a
//     vertex whose other dependencies were published in earlier levels
//     computes it at once.
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
                          int32_t* __restrict__ pend, int32_t* __restrict__ lv,
                          int32_t* __restrict__ front, int32_t* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c = 0;
  if (mode == 0) {
    c = doff[i + 1] - doff[i];
  } else {
    for (int k = 0; k < NR; ++k) {
      int r = rcv[(int64_t)k * N + i];
      if (r != (int)i && coef[(int64_t)k * N + i] != 0.0) c++;
    }
  }
  pend[i] = c;
  lv[i] = -1;
  if (c == 0) front[atomicAdd(&counts[0], 1)] = (int)i;
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
__device__ __forceinline__ double solve_one(int v, int64_t N, int NR, const int32_t* __restrict__ rcv,
                                            const double* __restrict__ coef,
                                            const double* __restrict__ diag,
                                            const double* __restrict__ rhs, const double* x,
                                            const int32_t* __restrict__ doff,
                                            const int32_t* __restrict__ don,
                                            const double* __restrict__ dval) {
  double s = rhs[v];
  if (MODE == 0) {
    const int e = doff[v + 1];
    for (int c = doff[v]; c < e; ++c) s = s + dval[c] * x[don[c]];
    return s;
  } else {
    for (int k = 0; k < NR; ++k) {
      int r = rcv[(int64_t)k * N + v];
      double o = coef[(int64_t)k * N + v];
      if (r != v && o != 0.0) s = s - o * x[r];
    }
    return s / diag[v];
  }
}

// Chaining test: every dependency of r other than v was solved in an earlier
// level (so published by a barrier).
template <int MODE>
__device__ __forceinline__ bool ready_now(int r, int v, int level, int64_t N, int NR,
                                          const int32_t* __restrict__ rcv,
                                          const double* __restrict__ coef,
                                          const int32_t* __restrict__ doff,
                                          const int32_t* __restrict__ don, const int32_t* lv) {
  if (MODE == 0) {
    const int e = doff[r + 1];
    for (int c = doff[r]; c < e; ++c) {
      int j = don[c];
      if (j == v) continue;
      int l = __hip_atomic_load(&lv[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (l < 0 || l >= level) return false;
    }
  } else {
    for (int k = 0; k < NR; ++k) {
      int q = rcv[(int64_t)k * N + r];
      if (q == r || q == v || coef[(int64_t)k * N + r] == 0.0) continue;
      int l = __hip_atomic_load(&lv[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (l < 0 || l >= level) return false;
    }
  }
  return true;
}

template <int MODE, int CHAIN>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_dag_level(
    int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ coef,
    const double* __restrict__ diag, const double* __restrict__ rhs, double* x,
    const int32_t* __restrict__ doff, const int32_t* __restrict__ don,
    const double* __restrict__ dval, int32_t* __restrict__ pend, int32_t* lv,
    int32_t* __restrict__ fa, int32_t* __restrict__ fb, int32_t* __restrict__ counts,
    unsigned int* __restrict__ bar, int32_t* __restrict__ nlev, int small) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nth = (int64_t)gridDim.x * blockDim.x;
  int level = 0;
  __shared__ int s_n;
  for (;;) {
    const int n = load_count(&counts[level]);
    if (n == 0) break;
    if (n <= small) {
      // Narrow frontier: block 0 runs consecutive narrow levels on its own,
      // separated by block barriers only; the other blocks wait at the next
      // grid barrier and resume at the level block 0 publishes.
      if (blockIdx.x == 0) {
        int ln = n;
        for (;;) {
          int32_t* cur = (level & 1) ? fb : fa;
          int32_t* nxt = (level & 1) ? fa : fb;
          for (int t = threadIdx.x; t < ln; t += blockDim.x) {
            const int v = cur[t];
            x[v] = solve_one<MODE>(v, N, NR, rcv, coef, diag, rhs, x, doff, don, dval);
            if (MODE == 0) {
              for (int k = 0; k < NR; ++k) {
                int r = rcv[(int64_t)k * N + v];
                if (r != v && r >= 0 && coef[(int64_t)k * N + v] != 0.0)
                  if (atomicSub(&pend[r], 1) == 1) nxt[atomicAdd(&counts[level + 1], 1)] = r;
              }
            } else {
              const int e = doff[v + 1];
              for (int c = doff[v]; c < e; ++c) {
                int j = don[c];
                if (atomicSub(&pend[j], 1) == 1) nxt[atomicAdd(&counts[level + 1], 1)] = j;
              }
            }
          }
          __threadfence_block();
          __syncthreads();
          level++;
          if (threadIdx.x == 0) s_n = load_count(&counts[level]);
          __syncthreads();
          ln = s_n;
          if (ln == 0 || ln > small) break;
        }
        if (threadIdx.x == 0) __hip_atomic_store(&bar[3], (unsigned int)level, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
      }
      if (!grid_barrier(bar, gridDim.x)) break;
      level = (int)__hip_atomic_load(&bar[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      continue;
    }
    int32_t* cur = (level & 1) ? fb : fa;
    int32_t* nxt = (level & 1) ? fa : fb;
    for (int64_t t = tid; t < n; t += nth) {
      int v = cur[t];
      while (v >= 0) {
        x[v] = solve_one<MODE>(v, N, NR, rcv, coef, diag, rhs, x, doff, don, dval);
        if (CHAIN) __hip_atomic_store(&lv[v], level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int next = -1;
        if (MODE == 0) {
          for (int k = 0; k < NR; ++k) {
            int r = rcv[(int64_t)k * N + v];
            if (r != v && r >= 0 && coef[(int64_t)k * N + v] != 0.0) {
              if (atomicSub(&pend[r], 1) == 1) {
                if (CHAIN && next < 0 && ready_now<0>(r, v, level, N, NR, rcv, coef, doff, don, lv))
                  next = r;
                else
                  nxt[atomicAdd(&counts[level + 1], 1)] = r;
              }
            }
          }
        } else {
          const int e = doff[v + 1];
          for (int c = doff[v]; c < e; ++c) {
            int j = don[c];
            if (atomicSub(&pend[j], 1) == 1) {
              if (CHAIN && next < 0 && ready_now<1>(j, v, level, N, NR, rcv, coef, doff, don, lv))
                next = j;
              else
                nxt[atomicAdd(&counts[level + 1], 1)] = j;
            }
          }
        }
        v = next;
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

template <int MODE>
static const void* level_kernel(int chain) {
  return chain ? (const void*)k_dag_level<MODE, 1> : (const void*)k_dag_level<MODE, 0>;
}

static int dag_solve(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                     const double* diag, const double* rhs, double* x, int32_t* nlevels,
                     int64_t* unsolved, hipStream_t s) {
  const int64_t N = m->N;
  const int g = nblocks(N);
  int32_t* pend = m->wi[0];
  int32_t* fa = m->wi[1];
  int32_t* fb = m->wi[2];
  int32_t* lv = m->wi[3];
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
  k_pending<<<g, GOSPL_BLOCK, 0, s>>>(mode, N, nrcv, rcv, coef, doff, pend, lv, fa, counts);
  GOSPL_CHECK(hipGetLastError());
  int64_t NN = N;
  int NR = nrcv;
  int32_t* nlev = dev;
  int32_t* don = m->wdon;
  double* dval = m->wdval;
  int small = m->dag_small;
  void* args[] = {&NN, &NR, (void*)&rcv, (void*)&coef, (void*)&diag, (void*)&rhs, &x, &doff, &don,
                  &dval, &pend, &lv, &fa, &fb, &counts, &bar, &nlev, &small};
  const void* kern = mode == 0 ? level_kernel<0>(m->dag_chain) : level_kernel<1>(m->dag_chain);
  GOSPL_CHECK(hipLaunchCooperativeKernel(kern, dim3(m->coop_blocks), dim3(GOSPL_BLOCK), args, 0, s));
  k_dag_unsolved<<<g, GOSPL_BLOCK, 0, s>>>(N, pend, x, dev + 1);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(m->hflag, dev, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipMemcpyAsync(m->hflag + 34, bar + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  if (m->hflag[34]) {
    set_error("receiver-DAG sweep aborted: a grid barrier did not complete");
    return 1;
  }
  if (nlevels) *nlevels = m->hflag[0];
  if (unsolved) *unsolved = m->hflag[1];
  return 0;
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_dag_init_occupancy(gospl_mesh* m) {
  int per_cu = 1 << 30;
  const void* ks[4] = {(const void*)k_dag_level<0, 0>, (const void*)k_dag_level<0, 1>,
                       (const void*)k_dag_level<1, 0>, (const void*)k_dag_level<1, 1>};
  for (int i = 0; i < 4; ++i) {
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
  const char* ch = getenv("GOSPL_DAG_CHAIN");
  m->dag_chain = (ch && atoi(ch) > 0) ? 1 : 0;
  const char* sm = getenv("GOSPL_DAG_SMALL");
  m->dag_small = (sm && atoi(sm) >= 0) ? atoi(sm) : 0;
  return 0;
}

extern "C" {

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
