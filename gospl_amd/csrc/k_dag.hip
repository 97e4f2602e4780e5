// Solves on the receiver DAG: flow accumulation (I - W)^T q = b (top down,
// flowplex.py:429-478 matrix) and the implicit stream-power system
// (bottom up, SPL.py:45-114 matrix).
//
// Both matrices are triangular in a topological order of the receiver graph
// (receivers are strictly lower, or closer to a spill point on flats), so
// they are solved exactly by one sweep over its levels:
//   * a donor list (reverse receiver graph, in FV neighbour order) is built
//     first; for flow accumulation every donor entry has a slot, and a solved
//     donor writes its weighted value w*q into its slot of each receiver's
//     row, so a vertex finds all its inputs in one contiguous row;
//   * pending counters hold the number of unsolved dependencies; a vertex
//     whose counter reaches zero joins the next frontier (wavefront-
//     aggregated appends: one atomic per wavefront, not per vertex);
//   * level 0 (no dependencies) runs on the whole GPU; the following levels
//     run in a small persistent cooperative kernel separated by a light grid
//     barrier, and once the frontier is small, in one workgroup with LDS
//     frontiers and workgroup barriers.
// Past level 0 the sweep is bound by the latency of each level: the loads of a
// vertex are issued together (unrolled over the stencil width) so one level
// costs about three dependent memory round trips plus the barrier.
// Every value is a fixed-order sum over its own row, so results do not depend
// on the schedule (no floating-point atomics).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "dag_common.h"

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

// Donor rows in neighbour order; slot[k][j] = position of (donor j, its
// receiver k) in the receiver's row (-1 where there is none).
__global__ void k_dep_fill(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
                           const double* __restrict__ coef, const int32_t* __restrict__ doff,
                           int32_t* __restrict__ don, int32_t* __restrict__ slot) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c = doff[i];
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    int j = nbr[(int64_t)p * N + i];
    if (j < 0) continue;
    for (int k = 0; k < NR; ++k) {
      if (rcv[(int64_t)k * N + j] == (int)i && coef[(int64_t)k * N + j] != 0.0) {
        don[c] = j;
        if (slot) slot[(int64_t)k * N + j] = c;
        c++;
      }
    }
  }
}

// Pending counters.  Level 0 (pend == 0) is not listed: k_dag_level0 scans
// all vertices, which avoids a million appends on one counter.
__global__ void k_pending(int mode, int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                          const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
                          const double* __restrict__ coef, const int32_t* __restrict__ doff,
                          int32_t* __restrict__ pend, int32_t* __restrict__ nonnbr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  // a dependency on a vertex outside the neighbour list has no row entry
  const int dg = deg[i];
  for (int k = 0; k < NR; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    if (r == (int)i || r < 0 || coef[(int64_t)k * N + i] == 0.0) continue;
    bool found = false;
    for (int p = 0; p < W && p < dg; ++p) found = found || nbr[(int64_t)p * N + i] == r;
    if (!found) atomicAdd(nonnbr, 1);
  }
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
}

// Grid barrier that also sums the next frontier's size (all blocks
// co-resident: cooperative launch).  bar[0] (64-bit) collects arrivals in its
// high word and the blocks' appended counts in its low word; the last block
// to arrive publishes (level << 32 | total) in bar[1]; bar[2] (low word) is
// the abort flag.  A waiter that spins too long raises the abort flag
// instead of hanging; all blocks then leave the level loop and the host
// reports the failure.  Returns the next frontier's size, or -1 on abort.
__device__ __forceinline__ int grid_barrier_count(unsigned long long* bar, unsigned int nblocks,
                                                  int level, int appended) {
  __syncthreads();
  __shared__ int s_next;
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned long long old = __hip_atomic_fetch_add(
        &bar[0], (1ull << 32) + (unsigned long long)(unsigned int)appended, __ATOMIC_ACQ_REL,
        __HIP_MEMORY_SCOPE_AGENT);
    int next;
    if ((unsigned int)(old >> 32) == nblocks - 1) {
      next = (int)((unsigned int)old + (unsigned int)appended);
      __hip_atomic_store(&bar[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar[1], ((unsigned long long)(unsigned int)level << 32) | (unsigned int)next,
                         __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long spins = 0;
      for (;;) {
        const unsigned long long w = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if ((unsigned int)(w >> 32) == (unsigned int)level) {
          next = (int)(unsigned int)w;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1ull << 26) ||
            __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          __hip_atomic_store(&bar[2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          next = -1;
          break;
        }
      }
    }
    __threadfence();
    s_next = next;
  }
  __syncthreads();
  return s_next;
}

#define DAG_BLOCK 1024

// Operands of one sweep.
struct DagArgs {
  int64_t N;
  int NR;
  const int32_t* rcv;   // (NR, N)
  const double* coef;   // (NR, N) weights (flow) or off-diagonal coefficients (SPL)
  const double* diag;   // (N) SPL diagonal
  const double* rhs;    // (N)
  double* x;            // (N) solution
  const int32_t* doff;  // (N + 1) donor row offsets
  const int32_t* don;   // donor rows
  double* acc;          // flow: weighted donor values, one per donor entry
  const int32_t* slot;  // flow: (NR, N) slot of (vertex, receiver k) in acc
  int32_t* pend;        // (N) unsolved dependencies
  double* rec;          // packed vertex records (nrcv <= 3), 16 words per vertex
  double* oacc;         // flow: row entries past the record, (c - 8) * N + i
  const int32_t* odon;  // stream power: row entries past the record, (c - 14) * N + i
};

// Decrement of a pending counter, returning the old value.  Device scope
// while several workgroups sweep, workgroup scope in the one-workgroup tail
// (there the counters are touched by one workgroup only, so the atomic can
// stay in this XCD's L2).
template <bool WG>
__device__ __forceinline__ int dec(int32_t* p) {
  if (WG) return __hip_atomic_fetch_add(p, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return atomicSub(p, 1);
}

// Appends the K candidate values of every lane whose flag is set with one
// reservation per wavefront.  Must be reached by all lanes of the wavefront.
template <int K, class Append>
__device__ __forceinline__ void append_all(const bool (&pred)[K], const int (&val)[K], int kmax,
                                           const Append& app) {
  unsigned long long m[K];
  int total = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    m[k] = (k < kmax) ? __ballot(pred[k]) : 0ull;
    total += __popcll(m[k]);
  }
  if (total == 0) return;
  const int lane = __lane_id();
  int base = 0;
  if (lane == 0) base = app.reserve(total);
  base = __shfl(base, 0);
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < kmax && pred[k]) app.write(base + __popcll(m[k] & below), val[k]);
    base += __popcll(m[k]);
  }
}

// Solve vertex v (if valid) and release its dependents into the next
// frontier.  Reached by all lanes of a wavefront.  All loads of a vertex,
// then all its counter decrements, are issued together (unrolled over the
// stencil width, results used only after the last one is issued), so a vertex
// costs three dependent memory round trips and one frontier reservation.
template <int MODE, class Append, bool WG = false>
__device__ __forceinline__ void process_vertex(bool valid, int v, const DagArgs& a,
                                               const Append& app) {
  const int64_t N = a.N;
  const int NR = a.NR;
  if (MODE == 0) {
    // q_v = b_v + sum over donor entries (row order) of w * q_donor
    int r[GOSPL_MAXW];
    int old[GOSPL_MAXW];
    bool dep[GOSPL_MAXW];
#pragma unroll
    for (int k = 0; k < GOSPL_MAXW; ++k) {
      r[k] = 0;
      old[k] = 0;
      dep[k] = false;
    }
    if (valid) {
      double w[GOSPL_MAXW];
      int sl[GOSPL_MAXW];
      const int b = a.doff[v];
      const int e = a.doff[v + 1];
      double s = a.rhs[v];
#pragma unroll
      for (int k = 0; k < GOSPL_MAXW; ++k) {
        if (k < NR) {
          r[k] = a.rcv[(int64_t)k * N + v];
          w[k] = a.coef[(int64_t)k * N + v];
          sl[k] = a.slot[(int64_t)k * N + v];
        }
      }
      double cv[GOSPL_MAXW];
#pragma unroll
      for (int c = 0; c < GOSPL_MAXW; ++c)
        if (b + c < e) cv[c] = a.acc[b + c];
#pragma unroll
      for (int c = 0; c < GOSPL_MAXW; ++c)
        if (b + c < e) s = s + cv[c];
      for (int c = b + GOSPL_MAXW; c < e; ++c) s = s + a.acc[c];
      a.x[v] = s;
#pragma unroll
      for (int k = 0; k < GOSPL_MAXW; ++k) {
        if (k < NR && r[k] != v && r[k] >= 0 && w[k] != 0.0) {
          dep[k] = true;
          if (sl[k] >= 0) a.acc[sl[k]] = w[k] * s;
          old[k] = dec<WG>(&a.pend[r[k]]);
        }
      }
    }
    bool rel[GOSPL_MAXW];
#pragma unroll
    for (int k = 0; k < GOSPL_MAXW; ++k) rel[k] = dep[k] && old[k] == 1;
    append_all<GOSPL_MAXW>(rel, r, NR, app);
  } else {
    // x_v = (h_v - sum_k off_k x_{rcv_k}) / diag_v, then release the donors
    int j[GOSPL_MAXW];
    int old[GOSPL_MAXW];
#pragma unroll
    for (int c = 0; c < GOSPL_MAXW; ++c) {
      j[c] = 0;
      old[c] = 0;
    }
    int b = 0, e = 0;
    if (valid) {
      int rr[GOSPL_MAXW];
      double o[GOSPL_MAXW], xr[GOSPL_MAXW];
      b = a.doff[v];
      e = a.doff[v + 1];
      double s = a.rhs[v];
      const double dg = a.diag[v];
#pragma unroll
      for (int k = 0; k < GOSPL_MAXW; ++k) {
        if (k < NR) {
          rr[k] = a.rcv[(int64_t)k * N + v];
          o[k] = a.coef[(int64_t)k * N + v];
        }
      }
#pragma unroll
      for (int c = 0; c < GOSPL_MAXW; ++c)
        if (b + c < e) j[c] = a.don[b + c];
#pragma unroll
      for (int k = 0; k < GOSPL_MAXW; ++k)
        if (k < NR && rr[k] != v && o[k] != 0.0) xr[k] = a.x[rr[k]];
#pragma unroll
      for (int k = 0; k < GOSPL_MAXW; ++k)
        if (k < NR && rr[k] != v && o[k] != 0.0) s = s - o[k] * xr[k];
      a.x[v] = s / dg;
#pragma unroll
      for (int c = 0; c < GOSPL_MAXW; ++c)
        if (b + c < e) old[c] = dec<WG>(&a.pend[j[c]]);
    }
    bool rel[GOSPL_MAXW];
#pragma unroll
    for (int c = 0; c < GOSPL_MAXW; ++c) rel[c] = b + c < e && old[c] == 1;
    append_all<GOSPL_MAXW>(rel, j, GOSPL_MAXW, app);
    // rows longer than the stencil width (not produced by an FV mesh)
    for (int c = GOSPL_MAXW; __ballot(b + c < e) != 0ull; ++c) {
      bool rl[1] = {false};
      int jj[1] = {0};
      if (b + c < e) {
        jj[0] = a.don[b + c];
        rl[0] = dec<WG>(&a.pend[jj[0]]) == 1;
      }
      append_all<1>(rl, jj, 1, app);
    }
  }
}

// ---------------------------------------------------------------------------
// Packed vertex records (up to three receivers).
//
// Past level 0 a level costs as much as the distinct cache lines a vertex
// touches on its dependent path, so with up to three receivers each vertex
// keeps what the sweep reads about it in one 128-byte record (16 words), and
// the values its dependencies produce are pushed into that record:
//   flow accumulation: w0 = row length, w1 = b, w2..4 = weights, w5..7 =
//     receivers r[3] then their slots sl[3] (int32), w8..15 = the first 8 row
//     entries (w * q of each donor, in row order; later entries in `oacc`).
//     sl[k] = word index of this vertex's entry in receiver k's record, or
//     -2 - its index in `oacc`, or -1 (none).
//   stream power: w0 = row length | dependency mask << 32, w1 = h,
//     w2 = diagonal, w3..5 = off-diagonal coefficients, w6..8 = the receivers'
//     values (pushed by them), w9..15 = the first 14 row entries as int32 word
//     indices of the donors' receiver-value slots (later entries in `odon`).
// A vertex then reads one line, and writes its pushes and counter decrements.
// ---------------------------------------------------------------------------
#define REC_W 16
#define REC_LOG 4
#define REC_ACC 8
#define REC_NACC 8
#define REC_XR 6
#define REC_TGT 9
#define REC_NTGT 14

__device__ __forceinline__ int lo32(double d) {
  return (int)(unsigned int)((unsigned long long)__double_as_longlong(d) & 0xffffffffull);
}
__device__ __forceinline__ int hi32(double d) {
  return (int)(unsigned int)((unsigned long long)__double_as_longlong(d) >> 32);
}

__device__ __forceinline__ double pack32(int lo, int hi) {
  return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) |
                                          (unsigned long long)(unsigned int)lo));
}

// Reverse neighbour positions (static per mesh): rev[q * N + i] = position of
// i in the neighbour list of its neighbour q, -1 if absent.
__global__ void k_rev_nbr(int64_t N, int W, const int32_t* __restrict__ nbr,
                          const int32_t* __restrict__ deg, int8_t* __restrict__ rev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int dg = deg[i];
  for (int q = 0; q < W; ++q) {
    int pos = -1;
    if (q < dg) {
      const int j = nbr[(int64_t)q * N + i];
      if (j >= 0) {
        const int dj = deg[j];
        for (int p = 0; p < W && p < dj; ++p)
          if (nbr[(int64_t)p * N + j] == (int)i) {
            pos = p;
            break;
          }
      }
    }
    rev[(int64_t)q * N + i] = (int8_t)pos;
  }
}

// Row masks: receiver r's row is the bitmask over (neighbour position p,
// receiver slot k), bit p * NR + k, of the donors that send to r through
// their slot k; bit order is row order (FV neighbour order, then slot).  Set
// from the donor side, one atomicOr per dependency (the mask is zeroed
// before).
// The position of i in the list of its receiver k is kept in rpos[k * N + i]
// (-1: no dependency) for the record builds, and, when dmask is given, the
// receivers i depends on in the stream-power sweep (bit k: a receiver other
// than i with a non-zero coefficient).
// The first workgroup also clears the sweep's small counters: zero[0..nz).
__global__ void k_rec_mask(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ deg, const int8_t* __restrict__ rev,
                           const int32_t* __restrict__ rcv, const double* __restrict__ coef,
                           unsigned long long* __restrict__ rowmask, int8_t* __restrict__ rpos,
                           int8_t* __restrict__ dmask, ZeroList zl) {
  if (blockIdx.x == 0)
    for (int k = 0; k < zl.n; ++k)
      for (int t = threadIdx.x; t < zl.len[k]; t += blockDim.x) zl.p[k][t] = 0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  // the neighbour list once, searched for every receiver
  int nb[GOSPL_MAXW];
  const int dg = deg[i];
#pragma unroll
  for (int q = 0; q < GOSPL_MAXW; ++q) nb[q] = (q < W && q < dg) ? nbr[(int64_t)q * N + i] : -1;
  int dm = 0;
  for (int k = 0; k < NR; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    const double ck = coef[(int64_t)k * N + i];
    int p = -1;
    if (k < 3 && r != (int)i && ck != 0.0) dm |= 1 << k;
    if (r != (int)i && r >= 0 && ck != 0.0) {
      int q = -1;
#pragma unroll
      for (int t = GOSPL_MAXW - 1; t >= 0; --t)
        if (nb[t] == r) q = t;
      if (q >= 0) p = rev[(int64_t)q * N + i];
      if (p >= 0) atomicOr(&rowmask[r], 1ull << (p * NR + k));
    }
    rpos[(int64_t)k * N + i] = (int8_t)p;
  }
  if (dmask) dmask[i] = (int8_t)dm;
}

// Appends i to the start list of the sweep when `start` (one reservation per
// wavefront).  Reached by all lanes of the wavefront.
__device__ __forceinline__ void start_append(bool start, int i, int32_t* __restrict__ fb,
                                             int32_t* __restrict__ cnt) {
  const unsigned long long m = __ballot(start);
  if (m == 0ull) return;
  const int lane = __lane_id();
  int base = 0;
  if (lane == 0) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, 0);
  if (start) fb[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

// Records of a workgroup staged in LDS (9 double2 per record: the padding
// keeps the lanes' 16-byte writes on different banks) and written out as
// one contiguous run, so each store instruction covers whole lines.
#define REC_STAGE 9
__device__ __forceinline__ void rec_store_staged(double2 (*stage)[REC_STAGE], const double2 (&out)[8], bool valid,
                                                 double* __restrict__ rec, int64_t N) {
  if (valid) {
#pragma unroll
    for (int k = 0; k < 8; ++k) stage[threadIdx.x][k] = out[k];
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * GOSPL_BLOCK;
  const int nrec = (int)(N - base < GOSPL_BLOCK ? N - base : GOSPL_BLOCK);
  double2* dst = reinterpret_cast<double2*>(rec + base * REC_W);
  for (int t = threadIdx.x; t < nrec * 8; t += GOSPL_BLOCK) dst[t] = stage[t >> 3][t & 7];
}

// An empty row entry: a signalling NaN, which no arithmetic produces, so a
// reader can tell an entry its donor has not written yet.
#define REC_EMPTY_BITS 0x7ff4000000000001ll
__device__ __forceinline__ double rec_empty() { return __longlong_as_double(REC_EMPTY_BITS); }
__device__ __forceinline__ bool is_empty(double d) { return __double_as_longlong(d) == REC_EMPTY_BITS; }

// Whole flow-accumulation record of vertex i, written at once.  The slot of
// (i, receiver k) is the rank of its bit in the receiver's row mask.  Level 0
// of the sweep is done here: the entry of a donor without donors of its own
// (its value is its b) is its push w * b, made by the row's owner, and the
// other entries start empty; pend[i] = the donors still to come.  A vertex
// without donors is solved (x = b); one whose donors all are is appended to
// the start list fb (count *cnt).
__global__ void k_rec_build0(int64_t N, int NR, const int32_t* __restrict__ nbr,
                             const int8_t* __restrict__ rpos, const int32_t* __restrict__ rcv,
                             const double* __restrict__ coef, const double* __restrict__ rhs,
                             const unsigned long long* __restrict__ rowmask, double* __restrict__ rec,
                             double* __restrict__ oacc, int32_t* __restrict__ pend,
                             int32_t* __restrict__ nonnbr, double* __restrict__ x, int32_t* __restrict__ fb,
                             int32_t* __restrict__ cnt) {
  __shared__ double2 stage[GOSPL_BLOCK][REC_STAGE];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((int64_t)blockIdx.x * blockDim.x >= N) return;  // whole workgroup past the mesh
  const bool valid = i < N;
  double2 out[8];
  bool start = false;
  if (valid) {
  int r[3], sl[3];
  double w[3];
  for (int k = 0; k < 3; ++k) {
    r[k] = (int)i;
    w[k] = 0.0;
    sl[k] = -1;
    if (k >= NR) continue;
    r[k] = rcv[(int64_t)k * N + i];
    w[k] = coef[(int64_t)k * N + i];
    const int rk = r[k];
    const int p = rpos[(int64_t)k * N + i];
    if (p < 0) {
      // a dependency with no row entry: its receiver is not a neighbour
      if (rk != (int)i && rk >= 0 && w[k] != 0.0) atomicAdd(nonnbr, 1);
      continue;
    }
    const int c = __popcll(rowmask[rk] & ((1ull << (p * NR + k)) - 1ull));
    sl[k] = c < REC_NACC ? (int)(rk * (int64_t)REC_W + REC_ACC + c)
                         : (int)(-2 - ((int64_t)(c - REC_NACC) * N + rk));
  }
  const unsigned long long mi = rowmask[i];
  const int nd = __popcll(mi);
  const double e = rec_empty();
  double ent[REC_NACC];
#pragma unroll
  for (int c = 0; c < REC_NACC; ++c) ent[c] = e;
  int npend = 0;
  unsigned long long mk = mi;
  for (int c = 0; mk != 0ull; ++c) {
    const int b = __ffsll((long long)mk) - 1;
    mk &= mk - 1ull;
    const int p = b / NR, k = b - p * NR;
    const int j = nbr[(int64_t)p * N + i];
    double val = e;
    if (rowmask[j] == 0ull) val = coef[(int64_t)k * N + j] * rhs[j];
    else ++npend;
    if (c < REC_NACC) {
#pragma unroll
      for (int q = 0; q < REC_NACC; ++q)
        if (q == c) ent[q] = val;
    } else {
      oacc[(int64_t)(c - REC_NACC) * N + i] = val;
    }
  }
  pend[i] = npend;
  if (nd == 0) x[i] = rhs[i];
  start = nd > 0 && npend == 0;
  out[0] = make_double2(pack32(nd, 0), rhs[i]);
  out[1] = make_double2(w[0], w[1]);
  out[2] = make_double2(w[2], pack32(r[0], r[1]));
  out[3] = make_double2(pack32(r[2], sl[0]), pack32(sl[1], sl[2]));
  out[4] = make_double2(ent[0], ent[1]);
  out[5] = make_double2(ent[2], ent[3]);
  out[6] = make_double2(ent[4], ent[5]);
  out[7] = make_double2(ent[6], ent[7]);
  }
  start_append(start, (int)i, fb, cnt);
  rec_store_staged(stage, out, valid, rec, N);
}

// Whole stream-power record of vertex i, written at once: its dependency
// mask, coefficients, its receivers' values and its row (the donors'
// receiver-value slots, in row order, from its row mask; entries past the
// record go to `odon`).  Level 0 of the sweep is done here: the value of a
// receiver that depends on nothing (dmask 0: h / diagonal) is made by i, the
// others start empty until pushed; pend[i] = the receivers still to come.  A
// vertex that depends on nothing is solved; one whose receivers all are is
// appended to the start list fb (count *cnt).
__global__ void k_rec_build1(int64_t N, int NR, const int32_t* __restrict__ nbr,
                             const int8_t* __restrict__ rpos, const int32_t* __restrict__ rcv,
                             const double* __restrict__ coef, const double* __restrict__ diag,
                             const double* __restrict__ rhs, const unsigned long long* __restrict__ rowmask,
                             double* __restrict__ rec, int32_t* __restrict__ odon,
                             int32_t* __restrict__ pend, int32_t* __restrict__ nonnbr,
                             const int8_t* __restrict__ dmask, double* __restrict__ x, int32_t* __restrict__ fb,
                             int32_t* __restrict__ cnt) {
  __shared__ double2 stage[GOSPL_BLOCK][REC_STAGE];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((int64_t)blockIdx.x * blockDim.x >= N) return;  // whole workgroup past the mesh
  const bool valid = i < N;
  double2 out[8];
  bool start = false;
  if (valid) {
  int mask = 0;
  double o[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < NR && k < 3; ++k) {
    o[k] = coef[(int64_t)k * N + i];
    const int rk = rcv[(int64_t)k * N + i];
    if (rk != (int)i && o[k] != 0.0) {
      mask |= 1 << k;
      // its value would never be pushed: the receiver is not a neighbour
      if (rpos[(int64_t)k * N + i] < 0) atomicAdd(nonnbr, 1);
    }
  }
  int t[REC_NTGT];
#pragma unroll
  for (int c = 0; c < REC_NTGT; ++c) t[c] = 0;
  unsigned long long mk = rowmask[i];
  const int nd = __popcll(mk);
  for (int c = 0; mk != 0ull; ++c) {
    const int b = __ffsll((long long)mk) - 1;
    mk &= mk - 1ull;
    const int p = b / NR, k = b - p * NR;
    const int j = nbr[(int64_t)p * N + i];
    const int tj = (int)((int64_t)j * REC_W + REC_XR + k);
    if (c < REC_NTGT) {
#pragma unroll
      for (int q = 0; q < REC_NTGT; ++q)
        if (q == c) t[q] = tj;
    } else {
      odon[(int64_t)(c - REC_NTGT) * N + i] = tj;
    }
  }
  const double e = rec_empty();
  double xr[3] = {e, e, e};
  int npend = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if ((mask >> k) & 1) {
      const int r = rcv[(int64_t)k * N + i];
      if (r >= 0 && r < N && dmask[r] == 0) xr[k] = rhs[r] / diag[r];
      else ++npend;
    }
  }
  pend[i] = npend;
  if (mask == 0) x[i] = rhs[i] / diag[i];
  start = mask != 0 && npend == 0;
  out[0] = make_double2(pack32(nd, mask), rhs[i]);
  out[1] = make_double2(diag[i], o[0]);
  out[2] = make_double2(o[1], o[2]);
  out[3] = make_double2(xr[0], xr[1]);
  out[4] = make_double2(xr[2], pack32(t[0], t[1]));
  out[5] = make_double2(pack32(t[2], t[3]), pack32(t[4], t[5]));
  out[6] = make_double2(pack32(t[6], t[7]), pack32(t[8], t[9]));
  out[7] = make_double2(pack32(t[10], t[11]), pack32(t[12], t[13]));
  }
  start_append(start, (int)i, fb, cnt);
  rec_store_staged(stage, out, valid, rec, N);
}

template <int MODE, class Append, bool WG>
__device__ __forceinline__ void process_rec(bool valid, int v, const DagArgs& a, const Append& app) {
  const int64_t N = a.N;
  if (MODE == 0) {
    int r[3] = {0, 0, 0};
    int old[3] = {0, 0, 0};
    bool dep[3] = {false, false, false};
    if (valid) {
      const double2* rp = reinterpret_cast<const double2*>(a.rec + (int64_t)v * REC_W);
      const double2 q0 = rp[0], q1 = rp[1], q2 = rp[2], q3 = rp[3];
      const double2 q4 = rp[4], q5 = rp[5], q6 = rp[6], q7 = rp[7];
      const int nd = lo32(q0.x);
      double s = q0.y;
      const double w[3] = {q1.x, q1.y, q2.x};
      r[0] = lo32(q2.y);
      r[1] = hi32(q2.y);
      r[2] = lo32(q3.x);
      const int sl[3] = {hi32(q3.x), lo32(q3.y), hi32(q3.y)};
      const double acc[REC_NACC] = {q4.x, q4.y, q5.x, q5.y, q6.x, q6.y, q7.x, q7.y};
#pragma unroll
      for (int c = 0; c < REC_NACC; ++c)
        if (c < nd) s = s + acc[c];
      for (int c = REC_NACC; c < nd; ++c) s = s + a.oacc[(int64_t)(c - REC_NACC) * N + v];
      a.x[v] = s;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (r[k] != v && r[k] >= 0 && w[k] != 0.0) {
          dep[k] = true;
          const double val = w[k] * s;
          if (sl[k] >= 0) a.rec[sl[k]] = val;
          else if (sl[k] <= -2) a.oacc[-2 - (int64_t)sl[k]] = val;
          old[k] = dec<WG>(&a.pend[r[k]]);
        }
      }
    }
    bool rel[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) rel[k] = dep[k] && old[k] == 1;
    append_all<3>(rel, r, 3, app);
  } else {
    int j[REC_NTGT];
    int old[REC_NTGT];
#pragma unroll
    for (int c = 0; c < REC_NTGT; ++c) {
      j[c] = 0;
      old[c] = 0;
    }
    int nd = 0;
    double xv = 0.0;
    if (valid) {
      const double2* rp = reinterpret_cast<const double2*>(a.rec + (int64_t)v * REC_W);
      const double2 q0 = rp[0], q1 = rp[1], q2 = rp[2], q3 = rp[3];
      const double2 q4 = rp[4], q5 = rp[5], q6 = rp[6], q7 = rp[7];
      nd = lo32(q0.x);
      const int mask = hi32(q0.x);
      double s = q0.y;
      const double dg = q1.x;
      const double o[3] = {q1.y, q2.x, q2.y};
      const double xr[3] = {q3.x, q3.y, q4.x};
      const int t[REC_NTGT] = {lo32(q4.y), hi32(q4.y), lo32(q5.x), hi32(q5.x), lo32(q5.y),
                               hi32(q5.y), lo32(q6.x), hi32(q6.x), lo32(q6.y), hi32(q6.y),
                               lo32(q7.x), hi32(q7.x), lo32(q7.y), hi32(q7.y)};
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if ((mask >> k) & 1) s = s - o[k] * xr[k];
      xv = s / dg;
      a.x[v] = xv;
#pragma unroll
      for (int c = 0; c < REC_NTGT; ++c) {
        if (c < nd) {
          a.rec[t[c]] = xv;
          j[c] = t[c] >> REC_LOG;
          old[c] = dec<WG>(&a.pend[j[c]]);
        }
      }
    }
    bool rel[REC_NTGT];
#pragma unroll
    for (int c = 0; c < REC_NTGT; ++c) rel[c] = c < nd && old[c] == 1;
    append_all<REC_NTGT>(rel, j, REC_NTGT, app);
    // rows longer than the record (more than 14 donor entries)
    for (int c = REC_NTGT; __ballot(c < nd) != 0ull; ++c) {
      bool rl[1] = {false};
      int jj[1] = {0};
      if (c < nd) {
        const int tc = a.odon[(int64_t)(c - REC_NTGT) * N + v];
        a.rec[tc] = xv;
        jj[0] = tc >> REC_LOG;
        rl[0] = dec<WG>(&a.pend[jj[0]]) == 1;
      }
      append_all<1>(rl, jj, 1, app);
    }
  }
}

template <int MODE, bool REC, bool WG, class Append>
__device__ __forceinline__ void solve_vertex(bool valid, int v, const DagArgs& a, const Append& app) {
  if (REC) process_rec<MODE, Append, WG>(valid, v, a, app);
  else process_vertex<MODE, Append, WG>(valid, v, a, app);
}

// Level 0 (vertices without dependencies, typically a third of the mesh) on
// the whole GPU before the persistent sweep takes over at level 1.  The
// dependency count is recomputed from the rows (pend[] is being decremented
// by other threads).  Same-address atomics serialise (~25 ns each on MI355X),
// so the level-0 size is not counted.
template <int MODE, bool REC>
__global__ void k_dag_level0(DagArgs a, int32_t* __restrict__ fb, int32_t* __restrict__ counts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t - (t & 63) >= a.N) return;  // whole wavefront beyond the mesh
  bool valid = t < a.N;
  const int v = (int)t;
  if (valid) {
    if (REC) {
      const double w0 = a.rec[(int64_t)v * REC_W];
      valid = MODE == 0 ? lo32(w0) == 0 : hi32(w0) == 0;
    } else if (MODE == 0) {
      valid = a.doff[v + 1] == a.doff[v];
    } else {
      for (int k = 0; k < a.NR && valid; ++k) {
        const int r = a.rcv[(int64_t)k * a.N + v];
        if (r != v && a.coef[(int64_t)k * a.N + v] != 0.0) valid = false;
      }
    }
  }
  solve_vertex<MODE, REC, false>(valid, v, a, GlobalAppend{&counts[1], fb, nullptr});
}

// Levels 1.. on a small co-resident grid.  Stops at the first level whose
// frontier is empty or smaller than tail_n (k_dag_tail continues there);
// *nlev = that level.
template <int MODE, bool REC>
__global__ __launch_bounds__(DAG_BLOCK) void k_dag_level(
    DagArgs a, int32_t* __restrict__ fa, int32_t* __restrict__ fb, int32_t* __restrict__ counts,
    unsigned long long* __restrict__ bar, int32_t* __restrict__ nlev, long long* __restrict__ trace,
    int interleave, int tail_n) {
  __shared__ int32_t s_app[2];  // appended by this workgroup, by level parity
  // Wavefront w of block b takes frontier slice (w * gridDim.x + b) when
  // interleaving, so a small frontier spreads over all CUs instead of filling
  // the first blocks.
  const int64_t wave = interleave
                           ? (int64_t)(threadIdx.x >> 6) * gridDim.x + blockIdx.x
                           : (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t tid = wave * 64 + (threadIdx.x & 63);
  const int64_t nth = (int64_t)gridDim.x * blockDim.x;
  int level = 1;  // level 0 ran in k_dag_level0
  int n = load_count(&counts[1]);
  if (threadIdx.x == 0) s_app[1] = 0;
  __syncthreads();
  for (;;) {
    if (trace && tid == 0 && level < 8192) {
      trace[2 * level] = (long long)wall_clock64();
      trace[2 * level + 1] = n;
    }
    if (n < tail_n || n == 0) break;
    int32_t* cur = (level & 1) ? fb : fa;
    int32_t* nxt = (level & 1) ? fa : fb;
    int32_t* cnt = &counts[level + 1];
    // this level appends to s_app[level & 1]; the other slot was read at the
    // previous barrier, before its closing workgroup barrier
    if (threadIdx.x == 0) s_app[(level + 1) & 1] = 0;
    const GlobalAppend app{cnt, nxt, &s_app[level & 1]};
    // uniform trip count per wavefront so the aggregated appends see all lanes
    for (int64_t t0 = tid - (tid & 63); t0 < n; t0 += nth) {
      const int64_t t = t0 + (tid & 63);
      const bool valid = t < n;
      const int v = valid ? cur[t] : 0;
      solve_vertex<MODE, REC, false>(valid, v, a, app);
    }
    __syncthreads();
    n = grid_barrier_count(bar, gridDim.x, level, s_app[level & 1]);
    if (n < 0) break;
    level++;
  }
  if (tid == 0) *nlev = level;
}

#define TAIL_THREADS 1024
#define TAIL_Q 8192

// Appends into the tail's next frontier: slots below TAIL_Q in LDS, the rest
// in the global list at the same index.
struct TailAppend {
  int32_t* cnt;  // LDS counter
  int32_t* lq;   // LDS list
  int32_t* gq;   // global list (capacity N)
  __device__ __forceinline__ int reserve(int n) const { return atomicAdd(cnt, n); }
  __device__ __forceinline__ void write(int idx, int value) const {
    if (idx < TAIL_Q) lq[idx] = value;
    else gq[idx] = value;
  }
};

// The remaining levels of the sweep on one workgroup, from the level where
// k_dag_level stopped (*nlev).  Frontiers live in LDS and levels are separated
// by workgroup barriers instead of a device-wide barrier with its cache
// write-back/invalidate.
template <int MODE, bool REC>
__global__ __launch_bounds__(TAIL_THREADS) void k_dag_tail(
    DagArgs a, int32_t* __restrict__ fa, int32_t* __restrict__ fb,
    const int32_t* __restrict__ counts, const unsigned long long* __restrict__ bar,
    int32_t* __restrict__ nlev, long long* __restrict__ trace) {
  __shared__ int32_t sq[2][TAIL_Q];
  __shared__ int32_t sc[3];
  if (bar[2]) return;  // the multi-block phase aborted
  int level = *nlev;
  int n = counts[level];
  if (threadIdx.x < 3) sc[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  bool first = true;  // the first frontier is entirely in the global list
  while (n > 0) {
    if (trace && threadIdx.x == 0 && level < 8192) {
      trace[2 * level] = (long long)wall_clock64();
      trace[2 * level + 1] = n;
    }
    const int p = level & 1;
    const int32_t* gcur = p ? fb : fa;
    int32_t* gnxt = p ? fa : fb;
    // counters rotate over three slots: this level appends to (level+1)%3 and
    // clears (level+2)%3, which every thread read before the previous barrier
    if (threadIdx.x == 0) sc[(level + 2) % 3] = 0;
    const TailAppend app{&sc[(level + 1) % 3], sq[p ^ 1], gnxt};
    for (int t0 = threadIdx.x - lane; t0 < n; t0 += blockDim.x) {
      const int t = t0 + lane;
      const bool valid = t < n;
      int v = 0;
      if (valid) v = (first || t >= TAIL_Q) ? gcur[t] : sq[p][t];
      solve_vertex<MODE, REC, true>(valid, v, a, app);
    }
    __syncthreads();
    n = sc[(level + 1) % 3];
    first = false;
    level++;
  }
  if (threadIdx.x == 0) {
    if (trace && level < 8192) {
      trace[2 * level] = (long long)wall_clock64();
      trace[2 * level + 1] = 0;
    }
    *nlev = level;
  }
}

// ---------------------------------------------------------------------------
// Dataflow sweeps (records only).  After level 0, the thread whose decrement
// releases a vertex goes on with it at once, so a dependency chain costs two
// dependent memory round trips per vertex (its record, then its pushes and
// counter decrements) and no grid barrier.  Each wavefront owns a contiguous
// share of the start list (the level-0 releases) and an LDS pool: a vertex
// that releases several vertices keeps one and pools the others, and idle
// lanes take work from the pool first, then from the share.  Flow
// accumulation mostly merges paths; the stream-power sweep runs upstream and
// fans out, so a wavefront whose pool fills hands batches of 64 vertices to a
// global queue, from which idle wavefronts claim them.
//
// Nothing waits on another thread except for data already on its way: pushes
// are coherent stores into row entries that start empty (a reader that still
// sees an empty entry, its donor's store not landed yet, reads it again), and
// a queue batch is claimed only once reserved, its entries starting at -1.
// These waits are bounded: past CHAIN_SPIN_MAX re-reads the abort flag is
// raised, every wavefront leaves and the host reports the failure.  An idle
// wavefront leaves when the queue is empty and no wavefront holds work (or
// after CHAIN_SPIN_MAX idle passes); work is only ever held by a running
// wavefront, so leaving early costs parallelism, never vertices.
// ---------------------------------------------------------------------------
// One wavefront's work state.  Wave-uniform except v, ci, next.
struct ChainWave {
  int32_t* pool;
  int np;           // pooled vertices
  int pool_cap;     // pool capacity (<= CHAIN_POOL, >= 128)
  int share_at;     // pool size kept when a batch goes to the global queue
  int32_t* gq;      // global queue, 64 entries per batch
  int64_t gq_cap;   // in batches
  int32_t* ctl;
  int lane;
  unsigned long long below;
  int next;         // the vertex this lane goes on with (-1: none)

  // Move the newest 64 pooled vertices to a global batch.
  __device__ __forceinline__ void spill() {
    int b = 0;
    if (lane == 0)
      b = (int)(atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[CQ_QUEUE]), 1ull << 32) >> 32);
    b = __shfl(b, 0);
    const int u = pool[np - 64 + lane];
    if ((int64_t)b < gq_cap) st_agent_i(&gq[(int64_t)b * 64 + lane], u);
    else if (lane == 0) st_agent_i(&ctl[CQ_ABORT], 1);
    np -= 64;
    wave_sync();
  }
  // A released vertex `val` of lanes with `rl`: kept as the lane's next
  // vertex if it has none, pooled otherwise.  Reached by all lanes.
  __device__ __forceinline__ void offer(bool rl, int val) {
    const bool keep = rl && next < 0;
    if (keep) next = val;
    const bool ex = rl && !keep;
    const unsigned long long em = __ballot(ex);
    if (em == 0ull) return;
    const int e = __popcll(em);
    while (np + e > pool_cap) spill();  // pool_cap >= 128 and e <= 64: np >= 64 here
    if (ex) pool[np + __popcll(em & below)] = val;
    np += e;
    wave_sync();
  }
};

#define CHAIN_AHEAD 3   // dependents' record lines loaded ahead per solved vertex (flow)
#define CHAIN_AHEAD1 6  // (stream power: the first donors of the row)

template <int MODE>
__global__ __launch_bounds__(CHAIN_THREADS) void k_chain(DagArgs a, const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ nlist,
                                                       int32_t* __restrict__ gq, int64_t gq_cap,
                                                       int32_t* __restrict__ ctl, int pool_cap,
                                                       int share_at, int pollers, int spin_max,
                                                       long long* __restrict__ trace) {
  __shared__ int32_t s_pool[CHAIN_THREADS / 64][CHAIN_POOL];
  const int64_t N = a.N;
  const int lane = __lane_id();
  ChainWave W{s_pool[threadIdx.x >> 6], 0, pool_cap, share_at, gq, gq_cap, ctl, lane,
              (1ull << lane) - 1ull, -1};
  const int64_t nw = (int64_t)gridDim.x * (CHAIN_THREADS / 64);
  const int64_t w = (int64_t)blockIdx.x * (CHAIN_THREADS / 64) + (threadIdx.x >> 6);
  const int64_t n0 = *nlist;
  int64_t cur = w * n0 / nw;  // this wavefront's share of the start list
  const int64_t end = (w + 1) * n0 / nw;
  int v = -1;                 // vertex this lane works on
  int64_t ci = -1;            // global-queue entry this lane waits for
  int retry = 0;              // passes this lane waited on a row or a queue entry
  int idle = 0;               // consecutive idle passes of the wavefront
  int pf[CHAIN_AHEAD1] = {};  // first words of dependents' records, loaded ahead
  bool act = true;            // counted in ctl[CQ_ACTIVE]
  bool poller = false;        // counted in ctl[CQ_IDLERS]
  unsigned long long* qword = reinterpret_cast<unsigned long long*>(&ctl[CQ_QUEUE]);
  if (lane == 0) atomicAdd(&ctl[CQ_ACTIVE], 1);
  for (;;) {
    const long long t_top = trace ? (long long)wall_clock64() : 0;
    // lanes waiting on a claimed queue entry
    if (ci >= 0) {
      const int q = ld_agent_i(&gq[ci]);
      if (q >= 0) {
        st_agent_i(&gq[ci], -1);  // the queue is empty again when the sweep ends
        v = q;
        ci = -1;
        retry = 0;
      }
    }
    // idle lanes take pooled vertices, then the next ones of the share
    const bool need = v < 0 && ci < 0;
    const unsigned long long nm = __ballot(need);
    if (nm != 0ull && (W.np > 0 || cur < end)) {
      const int rank = __popcll(nm & W.below);
      const int nn = __popcll(nm);
      const int fp = nn < W.np ? nn : W.np;
      if (need && rank < fp) v = W.pool[W.np - 1 - rank];
      const int64_t left = end - cur;
      const int fs = (int)((int64_t)(nn - fp) < left ? (int64_t)(nn - fp) : left);
      if (need && rank >= fp && rank - fp < fs) v = list[cur + (rank - fp)];
      W.np -= fp;
      cur += fs;
      wave_sync();
    }
    if (__ballot(v >= 0 || ci >= 0) == 0ull) {
      // pool and share are empty: claim a global batch, or leave.  A
      // wavefront leaves only after it has seen the queue empty, which
      // includes its own batches (same word), so no batch is left behind.
      // Up to CHAIN_POLLERS idle wavefronts stay while others hold work.
      int claimed = -1, stop = 0;
      if (lane == 0) {
        if (act) {
          atomicSub(&ctl[CQ_ACTIVE], 1);
          act = false;
        }
        const unsigned long long qw = __hip_atomic_load(qword, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int h = (int)(unsigned int)qw, t = (int)(qw >> 32);
        if (h < t) {
          if (atomicCAS(qword, qw, qw + 1ull) == qw) {
            atomicAdd(&ctl[CQ_ACTIVE], 1);
            act = true;
            claimed = h;
            if (poller) atomicSub(&ctl[CQ_IDLERS], 1);
            poller = false;
          }
        } else if (ld_agent_i(&ctl[CQ_ABORT]) != 0) {
          stop = 1;
        } else if (!poller) {
          if (atomicAdd(&ctl[CQ_IDLERS], 1) < pollers) poller = true;
          else stop = 1;
        } else if (ld_agent_i(&ctl[CQ_ACTIVE]) == 0) {
          stop = 1;
        }
      }
      claimed = __shfl(claimed, 0);
      stop = __shfl(stop, 0);
      act = __shfl((int)act, 0) != 0;
      poller = __shfl((int)poller, 0) != 0;
      if (claimed >= 0) {
        if ((int64_t)claimed >= gq_cap) {
          if (lane == 0) st_agent_i(&ctl[CQ_ABORT], 1);
          break;
        }
        ci = (int64_t)claimed * 64 + lane;
        idle = 0;
        continue;
      }
      if (stop) break;
      if (++idle > spin_max) {
        if (lane == 0) st_agent_i(&ctl[CQ_ABORT], 1);
        break;
      }
      __builtin_amdgcn_s_sleep(16);
      continue;
    }
    idle = 0;
    // solve the lanes' vertices whose rows are complete; push, decrement
    W.next = -1;
    bool solved = false;
    if (MODE == 0) {
      int r[3] = {0, 0, 0};
      int old[3] = {0, 0, 0};
      bool dep[3] = {false, false, false};
      if (v >= 0) {
        const double* rv = a.rec + (int64_t)v * REC_W;
        const double2* rp = reinterpret_cast<const double2*>(rv);
        const double2 q0 = rp[0], q1 = rp[1], q2 = rp[2], q3 = rp[3];
        double e[REC_NACC];
#pragma unroll
        for (int c = 0; c < REC_NACC; ++c) e[c] = ld_agent(rv + REC_ACC + c);  // with the rest of the line
        const int nd = lo32(q0.x);
        double s = q0.y;
        // branch-free over the record's entries (selects, not branches)
#pragma unroll
        for (int c = 0; c < REC_NACC; ++c) {
          const double t = s + e[c];
          s = c < nd ? t : s;
        }
        for (int c = REC_NACC; c < nd; ++c) s = s + ld_agent(a.oacc + (int64_t)(c - REC_NACC) * N + v);
        // an empty entry makes the sum NaN; only then look at the entries
        // (a NaN can also have arrived with the data)
        bool ready = !isnan(s);
        if (!ready) {
          ready = true;
#pragma unroll
          for (int c = 0; c < REC_NACC; ++c) ready = ready && (c >= nd || !is_empty(e[c]));
          for (int c = REC_NACC; c < nd; ++c)
            ready = ready && !is_empty(ld_agent(a.oacc + (int64_t)(c - REC_NACC) * N + v));
        }
        if (ready) {
          solved = true;
          a.x[v] = s;
          if (trace) {
            trace[v] = (long long)wall_clock64();
            trace[N + v] = t_top;
          }
          const double wk[3] = {q1.x, q1.y, q2.x};
          r[0] = lo32(q2.y);
          r[1] = hi32(q2.y);
          r[2] = lo32(q3.x);
          const int sl[3] = {hi32(q3.x), lo32(q3.y), hi32(q3.y)};
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            if (r[k] != v && r[k] >= 0 && wk[k] != 0.0) {
              dep[k] = true;
              const double val = wk[k] * s;
              if (sl[k] >= 0) st_agent(a.rec + sl[k], val);
              else if (sl[k] <= -2) st_agent(a.oacc + (-2 - (int64_t)sl[k]), val);
              old[k] = atomicSub(&a.pend[r[k]], 1);
            }
          }
        }
      }
      // One word of each dependent's record, loaded beside the decrements
      // (unconditionally, the lane's own line standing in): the lane whose
      // decrement releases a dependent reads that record next and finds the
      // line on its way to this XCD's L2.  Consumed at the end of the pass.
      {
        const int64_t own = (int64_t)(v >= 0 ? v : 0) * REC_W;
#pragma unroll
        for (int k = 0; k < CHAIN_AHEAD; ++k)
          pf[k] = *reinterpret_cast<const int*>(a.rec + (k < 3 && dep[k] ? (int64_t)r[k] * REC_W : own));
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) W.offer(dep[k] && old[k] == 1, r[k]);
    } else {
      int j[REC_NTGT];
      int old[REC_NTGT];
#pragma unroll
      for (int c = 0; c < REC_NTGT; ++c) {
        j[c] = 0;
        old[c] = 0;
      }
      int nd = 0;
      double xv = 0.0;
      if (v >= 0) {
        const double* rv = a.rec + (int64_t)v * REC_W;
        const double2* rp = reinterpret_cast<const double2*>(rv);
        const double2 q0 = rp[0], q1 = rp[1], q2 = rp[2];
        const double2 q5 = rp[5], q6 = rp[6], q7 = rp[7];
        const double q4y = rv[9];
        const double xr[3] = {ld_agent(rv + REC_XR), ld_agent(rv + REC_XR + 1), ld_agent(rv + REC_XR + 2)};
        const int mask = hi32(q0.x);
        double s = q0.y;
        const double dg = q1.x;
        const double o[3] = {q1.y, q2.x, q2.y};
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if ((mask >> k) & 1) s = s - o[k] * xr[k];
        const double xs = s / dg;
        // an empty receiver value makes the result NaN; only then look at them
        bool ready = !isnan(xs);
        if (!ready) {
          ready = true;
#pragma unroll
          for (int k = 0; k < 3; ++k) ready = ready && (((mask >> k) & 1) == 0 || !is_empty(xr[k]));
        }
        if (ready) {
          solved = true;
          nd = lo32(q0.x);
          xv = xs;
          a.x[v] = xv;
          if (trace) {
            trace[v] = (long long)wall_clock64();
            trace[N + v] = t_top;
          }
          const int t[REC_NTGT] = {lo32(q4y), hi32(q4y), lo32(q5.x), hi32(q5.x), lo32(q5.y),
                                   hi32(q5.y), lo32(q6.x), hi32(q6.x), lo32(q6.y), hi32(q6.y),
                                   lo32(q7.x), hi32(q7.x), lo32(q7.y), hi32(q7.y)};
#pragma unroll
          for (int c = 0; c < REC_NTGT; ++c) {
            if (c < nd) {
              st_agent(a.rec + t[c], xv);
              j[c] = t[c] >> REC_LOG;
              old[c] = atomicSub(&a.pend[j[c]], 1);
            }
          }
        }
      }
      {
        const int64_t own = (int64_t)(v >= 0 ? v : 0) * REC_W;
#pragma unroll
        for (int c = 0; c < CHAIN_AHEAD1; ++c)
          pf[c] = *reinterpret_cast<const int*>(a.rec + (c < nd ? (int64_t)j[c] * REC_W : own));
      }
#pragma unroll
      for (int c = 0; c < REC_NTGT; ++c) W.offer(c < nd && old[c] == 1, j[c]);
      // rows longer than the record (more than 14 donor entries)
      for (int c = REC_NTGT; __ballot(c < nd) != 0ull; ++c) {
        bool rl = false;
        int jj = 0;
        if (c < nd) {
          const int tc = a.odon[(int64_t)(c - REC_NTGT) * N + v];
          st_agent(a.rec + tc, xv);
          jj = tc >> REC_LOG;
          rl = atomicSub(&a.pend[jj], 1) == 1;
        }
        W.offer(rl, jj);
      }
    }
    // the words loaded ahead: a record's first word is its row length, far
    // below 2^30
    {
      int any = 0;
#pragma unroll
      for (int k = 0; k < (MODE == 0 ? CHAIN_AHEAD : CHAIN_AHEAD1); ++k) any |= pf[k];
      if (any & 0x40000000) st_agent_i(&ctl[CQ_RETRIES], any);
    }
    // share work with idle wavefronts once the pool is long
    if (W.np >= W.share_at + 64) W.spill();
    const bool waiting = v >= 0 && !solved;
    if (trace && waiting) atomicAdd(&ctl[CQ_RETRIES], 1);
    if (solved) {
      v = W.next;
      retry = 0;
    }
    if ((waiting || ci >= 0) && ++retry > spin_max) st_agent_i(&ctl[CQ_ABORT], 1);
    // the abort flag is read only by wavefronts that wait on something
    if (__ballot(waiting || ci >= 0) != 0ull) {
      int ab = 0;
      if (lane == 0) ab = ld_agent_i(&ctl[CQ_ABORT]);
      if (__shfl(ab, 0) != 0) break;
    }
  }
}

// Vertices never reached (a cycle in the receivers and everything below it):
// x = 0, listed.
__global__ void k_dag_unsolved(int64_t N, const int32_t* __restrict__ pend, double* __restrict__ x,
                               int32_t* __restrict__ cnt, int32_t* __restrict__ list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool left = i < N && pend[i] > 0;
  if (left) x[i] = 0.0;
  wave_append(left, (int)i, cnt, list);
}

#define CYCLE_THREADS 1024
#define CYCLE_MAXIT 100000

__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, sh[w]);
  return r;
}

// Flow accumulation on the left-over vertices U of a sweep (a receiver cycle
// and everything below it).  goSPL's iterative solve converges there when
// the cycle drains and gives up otherwise (flowplex.py:178-322): here the
// fixed-point iteration q <- b + W^T q over U in Jacobi form, from q = 0 with
// the solved vertices fixed, in one workgroup, until the largest change is at
// most 1e-14 of the largest value.  Each pass sums a row in its fixed order,
// so the result does not depend on scheduling.  res[0] = 1 when converged.
// Pushes w_k * s of vertex v into its receivers' rows.
template <bool REC>
__device__ __forceinline__ void cycle_push(const DagArgs& a, int v, double s) {
  const int64_t N = a.N;
  if (REC) {
    const double* r = a.rec + (int64_t)v * REC_W;
    const int32_t* ri = reinterpret_cast<const int32_t*>(r);
    for (int k = 0; k < 3; ++k) {
      const int rk = ri[10 + k];
      const int sl = ri[13 + k];
      const double w = r[2 + k];
      if (rk != v && rk >= 0 && w != 0.0) {
        if (sl >= 0) a.rec[sl] = w * s;
        else if (sl <= -2) a.oacc[-2 - (int64_t)sl] = w * s;
      }
    }
  } else {
    for (int k = 0; k < a.NR; ++k) {
      const int r = a.rcv[(int64_t)k * N + v];
      const double w = a.coef[(int64_t)k * N + v];
      const int sl = a.slot[(int64_t)k * N + v];
      if (r != v && r >= 0 && w != 0.0 && sl >= 0) a.acc[sl] = w * s;
    }
  }
}

// b_v plus the row of v, in row order.
template <bool REC>
__device__ __forceinline__ double cycle_row(const DagArgs& a, int v) {
  const int64_t N = a.N;
  if (REC) {
    const double* r = a.rec + (int64_t)v * REC_W;
    const int nd = reinterpret_cast<const int32_t*>(r)[0];
    double s = r[1];
    for (int c = 0; c < nd; ++c)
      s = s + (c < REC_NACC ? r[REC_ACC + c] : a.oacc[(int64_t)(c - REC_NACC) * N + v]);
    return s;
  }
  double s = a.rhs[v];
  const int e = a.doff[v + 1];
  for (int c = a.doff[v]; c < e; ++c) s = s + a.acc[c];
  return s;
}

template <bool REC>
__global__ __launch_bounds__(CYCLE_THREADS) void k_flow_cycles(DagArgs a, const int32_t* __restrict__ U,
                                                             const int32_t* __restrict__ nU,
                                                             double* __restrict__ xn,
                                                             int32_t* __restrict__ res) {
  __shared__ double sh[3][CYCLE_THREADS / 64];
  const int n = *nU;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int v = U[t];
    a.x[v] = 0.0;
    cycle_push<REC>(a, v, 0.0);
  }
  __syncthreads();
  int conv = 0;
  for (int it = 0; it < CYCLE_MAXIT; ++it) {
    double dmax = 0.0, xmax = 0.0, bad = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int v = U[t];
      const double s = cycle_row<REC>(a, v);
      xn[t] = s;
      if (!isfinite(s)) bad = 1.0;
      dmax = fmax(dmax, fabs(s - a.x[v]));
      xmax = fmax(xmax, fabs(s));
    }
    dmax = block_max(dmax, sh[0]);
    xmax = block_max(xmax, sh[1]);
    bad = block_max(bad, sh[2]);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int v = U[t];
      const double s = xn[t];
      a.x[v] = s;
      cycle_push<REC>(a, v, s);
    }
    __syncthreads();
    if (bad != 0.0) break;
    if (dmax <= 1.0e-14 * xmax) {
      conv = 1;
      break;
    }
  }
  if (threadIdx.x == 0) res[0] = conv;
}

// SPL coefficients (SPL.py:56-108): off[k] = -(Kbr*limiter/dist)*w (0 for
// self receivers), diag = ((1 - off0) - off1) - ...
__global__ void k_spl_coef(int64_t N, int NR, const int32_t* __restrict__ rcv,
                           const double* __restrict__ wgt, const double* __restrict__ dist,
                           const double* __restrict__ fal, const double* __restrict__ kfac,
                           double K, double m, const double* __restrict__ mvec, double dt, double sl,
                           const double* __restrict__ lfill, const double* __restrict__ h,
                           double* __restrict__ off, double* __restrict__ diag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool sea = lfill[i] <= sl;
  double kbr = K * 1.0;
  if (kfac) kbr = kbr * kfac[i];
  kbr = kbr * (dt * pow(fal[i], mvec ? mvec[i] : m));
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
// Non-linear stream power (n != 1, detachment-limited; nlSPL._solveNL,
// gospl/eroder/nlSPL.py:321-396).  F(x)_i = x_i - hOld_i + Kbr_i S_i(x)^n
// with S_i = sum over receivers with w > 0 and x_i > x_r of w (x_i - x_r)/d
// (local_spl, functions.F90:1469-1502) is lower-triangular in the receiver
// order, and so is its Jacobian (local_spl_coeff, :1504-1542): each Newton
// step J d = -F is one exact erosion-direction sweep.
// ---------------------------------------------------------------------------

// Kbr = K kfac dt FAL^m elimiter, 0 at sea, with elimiter = dh / (dh + 0.01)
// of the largest drop to a receiver dh (0 when dh = 0), nlSPL.py:330-346.
__global__ void k_nl_kbr(int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ h,
                         const double* __restrict__ fal, const double* __restrict__ kfac, double K, double m,
                         double dt, double sl, const double* __restrict__ lfill, double* __restrict__ kbr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double hi = h[i];
  double dh = -INFINITY;
  for (int k = 0; k < NR; ++k) {
    const double d = hi - h[rcv[(int64_t)k * N + i]];
    dh = d > dh ? d : dh;
  }
  const double el = dh != 0.0 ? dh / (dh + 1.0e-2) : 0.0;
  double kb = K * 1.0;
  if (kfac) kb = kb * kfac[i];
  kb = kb * ((dt * pow(fal[i], m)) * el);
  kbr[i] = lfill[i] <= sl ? 0.0 : kb;
}

// One Newton linearisation at x: off/diag of J and rhs = -F.
__global__ void k_nl_jac(int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ wgt,
                         const double* __restrict__ dist, const double* __restrict__ x,
                         const double* __restrict__ hold, const double* __restrict__ kbr, double n,
                         double* __restrict__ off, double* __restrict__ diag, double* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double xi = x[i];
  double S = 0.0, c0 = 0.0;
  double ck[GOSPL_MAXW];
  for (int k = 0; k < NR; ++k) {
    ck[k] = 0.0;
    const double w = wgt[(int64_t)k * N + i];
    if (w > 0.0) {
      const int r = rcv[(int64_t)k * N + i];
      const double xr = x[r];
      if (xi > xr) {
        const double d = dist[(int64_t)k * N + i];
        S = S + w * (xi - xr) / d;
        ck[k] = -w / d;
        c0 = c0 + w / d;
      }
    }
  }
  const double kb = kbr[i];
  const double F = (xi - hold[i]) + kb * (S > 0.0 ? pow(S, n) : 0.0);
  const double dd = S > 0.0 ? n * kb * pow(S, n - 1.0) : 0.0;
  for (int k = 0; k < NR; ++k) off[(int64_t)k * N + i] = dd * ck[k];
  diag[i] = 1.0 + dd * c0;
  rhs[i] = -F;
}

// x += d; keys of max |d| and max |x| (one atomic per workgroup).
__global__ __launch_bounds__(GOSPL_BLOCK) void k_nl_step(int64_t N, const double* __restrict__ d,
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

// ---------------------------------------------------------------------------
// Adjoints of the receiver-DAG solves (gradients with receivers, weights,
// discharge and sea/sink flags held at their forward values).
// ---------------------------------------------------------------------------

// (I - W) y = lam, the transpose of the flow system: off = -w (0 for self
// receivers), diag = 1, solved bottom-up like the stream-power system.
__global__ void k_fa_adj_prep(int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ wgt,
                              double* __restrict__ off, double* __restrict__ diag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int k = 0; k < NR; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    const double w = wgt[(int64_t)k * N + i];
    off[(int64_t)k * N + i] = (r != (int)i && r >= 0) ? -w : 0.0;
  }
  diag[i] = 1.0;
}

// Stream-power adjoint, step 1: the right-hand side of A^T mu = lam_x
// (lam_x = 0 where the forward kept h: flat outlets and sinks, SPL.py:345-352)
// and the transposed system's weights c_ik / diag_i, so that nu = diag * mu
// solves nu_i = lam_x_i + sum over donors j of (c_jk / diag_j) nu_j, a
// top-down sweep like flow accumulation.
__global__ void k_spl_adj_prep(int64_t N, int NR, const double* __restrict__ off, const double* __restrict__ diag,
                               const double* __restrict__ lam, const int8_t* __restrict__ outlet, int flat,
                               const int8_t* __restrict__ lsink, double* __restrict__ wt, double* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double d = diag[i];
  for (int k = 0; k < NR; ++k) wt[(int64_t)k * N + i] = -off[(int64_t)k * N + i] / d;
  const bool reset = (flat && outlet[i]) || lsink[i];
  rhs[i] = reset ? 0.0 : lam[i];
}

// Step 2: mu = nu / diag, the local terms of each dependency (i, k) and the
// erodibility gradient.  With c_ik = Kbr_i l(h_i - h_rk) w_ik / d_ik,
// l(s) = s / (s + 0.01), Kbr_i = K kfac_i dt FAL_i^m (0 at sea), and the
// forward residual (A x)_i - h_i = x_i + sum_k c_ik (x_i - x_rk) - h_i:
//   e_ik = mu_i Kbr_i w_ik / d_ik l'(h_i - h_rk) (x_i - x_rk)
//   dL/dlog10(kfac_i) = -ln(10) mu_i sum_k c_ik (x_i - x_rk)
//   dL/dFAL_i        = -m / FAL_i mu_i sum_k c_ik (x_i - x_rk)
__global__ void k_spl_adj_terms(int64_t N, int NR, const int32_t* __restrict__ rcv, const double* __restrict__ wgt,
                                const double* __restrict__ dist, const double* __restrict__ fal,
                                const double* __restrict__ kfac, double K, double m, double dt, double sl,
                                const double* __restrict__ lfill, const double* __restrict__ h,
                                const double* __restrict__ x, const double* __restrict__ off,
                                const double* __restrict__ diag, double* __restrict__ mu, double* __restrict__ e,
                                double* __restrict__ glogk, double* __restrict__ gfal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double mi = mu[i] / diag[i];
  mu[i] = mi;
  const bool sea = lfill[i] <= sl;
  double kbr = K * 1.0;
  if (kfac) kbr = kbr * kfac[i];
  kbr = kbr * (dt * pow(fal[i], m));
  if (sea) kbr = 0.0;
  double S = 0.0;
  for (int k = 0; k < NR; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    const double o = off[(int64_t)k * N + i];
    const double dk = dist[(int64_t)k * N + i];
    double ek = 0.0;
    if (r != (int)i && r >= 0 && kbr != 0.0 && dk != 0.0) {
      // l'(s) = 0.01 / (s + 0.01)^2, also at s = 0 where the forward's
      // l(0) = 0 is taken without the division
      const double dx = x[i] - x[r];
      S = S + (-o) * dx;
      const double dh = h[i] - h[r];
      const double w = wgt[(int64_t)k * N + i];
      const double t = dh + 1.0e-2;
      ek = mi * (kbr * w / dk) * (1.0e-2 / (t * t)) * dx;
    }
    e[(int64_t)k * N + i] = ek;
  }
  if (glogk) glogk[i] = -2.302585092994046 * mi * S;
  if (gfal) gfal[i] = fal[i] > 0.0 ? -m / fal[i] * mi * S : 0.0;
}

// Step 3: dL/dh_j = (lam_j where the forward kept h) + mu_j - sum_k e_jk +
// sum over donors i of j (through their slot k) of e_ik, gathered through the
// neighbour list in a fixed order (no floating-point atomics).
__global__ void k_spl_adj_gather(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                                 const int32_t* __restrict__ deg, const int32_t* __restrict__ rcv,
                                 const double* __restrict__ e, const double* __restrict__ mu,
                                 const double* __restrict__ lam, const int8_t* __restrict__ outlet, int flat,
                                 const int8_t* __restrict__ lsink, double* __restrict__ gh) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double g = mu[j];
  for (int k = 0; k < NR; ++k) g = g - e[(int64_t)k * N + j];
  const int dg = deg[j];
  for (int p = 0; p < W && p < dg; ++p) {
    const int i = nbr[(int64_t)p * N + j];
    if (i < 0) continue;
    for (int k = 0; k < NR; ++k)
      if (rcv[(int64_t)k * N + i] == (int)j) g = g + e[(int64_t)k * N + i];
  }
  const bool reset = (flat && outlet[j]) || lsink[j];
  gh[j] = reset ? g + lam[j] : g;
}

int ensure_records(gospl_mesh* m, hipStream_t s) {
  if (m->wrec) return 0;
  GOSPL_CHECK(hipMalloc(&m->wrec, sizeof(double) * REC_W * (m->N + 64)));
  GOSPL_CHECK(hipMalloc(&m->rev, sizeof(int8_t) * m->W * (m->N + 64)));
  k_rev_nbr<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, m->W, m->nbr, m->deg, m->rev);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int ensure_donors(gospl_mesh* m, int64_t cap) {
  if (m->wdon_cap >= cap) return 0;
  if (m->wdon) GOSPL_CHECK(hipFree(m->wdon));
  if (m->wacc) GOSPL_CHECK(hipFree(m->wacc));
  if (m->wslot) GOSPL_CHECK(hipFree(m->wslot));
  m->wdon = nullptr;
  m->wacc = nullptr;
  m->wslot = nullptr;
  GOSPL_CHECK(hipMalloc(&m->wdon, sizeof(int32_t) * (cap + 64)));
  GOSPL_CHECK(hipMalloc(&m->wacc, sizeof(double) * (cap + 64)));
  GOSPL_CHECK(hipMalloc(&m->wslot, sizeof(int32_t) * (cap + 64)));
  m->wdon_cap = cap;
  return 0;
}

// A sweep in two halves: dag_issue enqueues everything, the status words
// included (their copies to the host flags too), without waiting;
// dag_finish waits for the stream and reads them.  The sweeps use their
// own integer workspace (wdag), so pit information may run beside them on
// another stream.
static int dag_finish(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                      int32_t* nlevels, int64_t* unsolved, hipStream_t s);

static int dag_issue(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                     const double* diag, const double* rhs, double* x, hipStream_t s) {
  const int64_t N = m->N;
  const int g = nblocks(N);
  int32_t* pend = m->wdag[0];
  int32_t* fa = m->wdag[1];
  int32_t* fb = m->wdag[2];
  int32_t* dcnt = m->wdag[3];
  int32_t* doff = m->wdag[4];
  int32_t* counts = m->counts;
  int32_t* dev = m->dflag;
  // packed records: up to 3 receivers, int32 word indices, row masks of W * nrcv <= 64 bits
  const bool rec = nrcv <= 3 && !m->dag_csr && N * REC_W < (int64_t)INT32_MAX && m->W * nrcv <= 64;
  unsigned long long* bar = reinterpret_cast<unsigned long long*>(dev + 32);
  const bool chain = rec && m->dag_chain;
  // dataflow sweep: global batch queue (entries at -1 between sweeps: each
  // taken entry is reset) and control words
  const int64_t gq_cap = N / 64 + 1;
  if (chain && !m->chain_ctl) {
    GOSPL_CHECK(hipMalloc(&m->chain_ctl, sizeof(int32_t) * CQ_WORDS));
    GOSPL_CHECK(hipMalloc(&m->chain_gq, sizeof(int32_t) * gq_cap * 64));
    GOSPL_CHECK(hipMemsetAsync(m->chain_gq, 0xff, sizeof(int32_t) * gq_cap * 64, s));
  }
  // the small counters (status words, barrier words, dataflow control words,
  // the dataflow's frontier count) are cleared by the record-mask kernel;
  // per-level frontier counts of the level sweeps by memset
  ZeroList zl{{dev, dev + 32, counts, m->chain_ctl}, {4, 6, 2, CQ_WORDS}, chain ? 4 : 3};
  if (!chain) GOSPL_CHECK(hipMemsetAsync(counts, 0, sizeof(int32_t) * (N + 2), s));
  if (!rec) {
    GOSPL_CHECK(hipMemsetAsync(dev, 0, sizeof(int32_t) * 4, s));
    GOSPL_CHECK(hipMemsetAsync(bar, 0, sizeof(unsigned long long) * 3, s));
  }
  int32_t* slot = nullptr;
  if (rec) {
    // packed records; rows hold at most nrcv entries per neighbour, the ones
    // past the record go to the overflow arrays
    if (ensure_records(m, s)) return 1;
    if (ensure_donors(m, (int64_t)nrcv * m->W * N)) return 1;
    unsigned long long* rowmask = m->wu[0];
    GOSPL_CHECK(hipMemsetAsync(rowmask, 0, sizeof(unsigned long long) * N, s));
    int8_t* rpos = reinterpret_cast<int8_t*>(m->wdag[4]);  // 3 N bytes
    int8_t* dmask = mode == 1 ? rpos + 3 * N : nullptr;     // the fourth N bytes
    k_rec_mask<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, nrcv, m->nbr, m->deg, m->rev, rcv, coef, rowmask, rpos,
                                         dmask, zl);
    // the sweep proper: the builds do level 0, their releases are the start list
    if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[0], s));
    if (mode == 0)
      k_rec_build0<<<g, GOSPL_BLOCK, 0, s>>>(N, nrcv, m->nbr, rpos, rcv, coef, rhs, rowmask, m->wrec, m->wacc,
                                              pend, dev + 3, x, fb, counts + 1);
    else
      k_rec_build1<<<g, GOSPL_BLOCK, 0, s>>>(N, nrcv, m->nbr, rpos, rcv, coef, diag, rhs, rowmask, m->wrec,
                                              m->wdon, pend, dev + 3, dmask, x, fb, counts + 1);
  } else {
    // CSR donor rows
    if (ensure_donors(m, (int64_t)nrcv * N)) return 1;
    slot = mode == 0 ? m->wslot : nullptr;
    if (slot) GOSPL_CHECK(hipMemsetAsync(slot, 0xff, sizeof(int32_t) * nrcv * N, s));
    k_dep_count<<<nblocks(N + 1), GOSPL_BLOCK, 0, s>>>(N, m->W, nrcv, m->nbr, m->deg, rcv, coef, dcnt);
    size_t bytes = m->cub_tmp_bytes;
    GOSPL_CHECK(rocprim::exclusive_scan(m->dag_tmp, bytes, dcnt, doff, 0, (size_t)(N + 1),
                                        rocprim::plus<int32_t>(), s));
    k_dep_fill<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, nrcv, m->nbr, m->deg, rcv, coef, doff, m->wdon, slot);
    k_pending<<<g, GOSPL_BLOCK, 0, s>>>(mode, N, m->W, nrcv, m->nbr, m->deg, rcv, coef, doff, pend, dev + 3);
  }
  GOSPL_CHECK(hipGetLastError());
  DagArgs a{N, nrcv, rcv, coef, diag, rhs, x, doff, m->wdon, m->wacc, slot, pend,
            rec ? m->wrec : nullptr, m->wacc, m->wdon};
  int32_t* nlev = dev;
  // GOSPL_DAG_TRACE=1: per-level wall-clock stamps and frontier sizes, summarised on stderr
  static const bool do_trace = getenv("GOSPL_DAG_TRACE") && atoi(getenv("GOSPL_DAG_TRACE")) > 0;
  long long* trace = nullptr;
  if (do_trace) {
    if (!m->trace_buf) GOSPL_CHECK(hipMalloc(&m->trace_buf, sizeof(long long) * 2 * 8192));
    trace = m->trace_buf;
    GOSPL_CHECK(hipMemsetAsync(trace, 0, sizeof(long long) * 2 * 8192, s));
  }
  // dataflow sweeps trace one solve time per vertex instead
  long long* vtrace = nullptr;
  if (do_trace && rec && m->dag_chain) {
    if (!m->trace_v) GOSPL_CHECK(hipMalloc(&m->trace_v, sizeof(long long) * 2 * N));
    vtrace = m->trace_v;
    GOSPL_CHECK(hipMemsetAsync(vtrace, 0, sizeof(long long) * 2 * N, s));
  }
  void* args[] = {&a, &fa, &fb, &counts, &bar, &nlev, &trace, &m->dag_interleave, &m->dag_tail};
  const void* kern = rec ? (mode == 0 ? (const void*)k_dag_level<0, true> : (const void*)k_dag_level<1, true>)
                        : (mode == 0 ? (const void*)k_dag_level<0, false> : (const void*)k_dag_level<1, false>);
  int32_t* gq = m->chain_gq;
  if (m->timing && !rec) GOSPL_CHECK(hipEventRecord(m->tev[0], s));
  if (chain) {
    if (mode == 0)
      k_chain<0><<<m->chain_blocks[0], CHAIN_THREADS, 0, s>>>(a, fb, counts + 1, gq, gq_cap, m->chain_ctl,
                                                           m->chain_pool, m->chain_share, m->chain_pollers, m->chain_spin, vtrace);
    else
      k_chain<1><<<m->chain_blocks[1], CHAIN_THREADS, 0, s>>>(a, fb, counts + 1, gq, gq_cap, m->chain_ctl,
                                                           m->chain_pool, m->chain_share, m->chain_pollers, m->chain_spin, vtrace);
    GOSPL_CHECK(hipGetLastError());
  } else {
    // level by level: level 0 (done by the record builds), the persistent
    // level kernel, then the one-workgroup tail
    if (!rec) {
      if (mode == 0) k_dag_level0<0, false><<<g, GOSPL_BLOCK, 0, s>>>(a, fb, counts);
      else k_dag_level0<1, false><<<g, GOSPL_BLOCK, 0, s>>>(a, fb, counts);
    }
    GOSPL_CHECK(hipGetLastError());
    if (m->dag_coop)
      GOSPL_CHECK(hipLaunchCooperativeKernel(kern, dim3(m->coop_blocks), dim3(m->dag_threads), args, 0, s));
    else  // plain launch (counter collection); the grid fits the GPU many times over
      GOSPL_CHECK(hipLaunchKernel(kern, dim3(m->coop_blocks), dim3(m->dag_threads), args, 0, s));
    if (m->dag_tail > 0) {
      if (rec) {
        if (mode == 0) k_dag_tail<0, true><<<1, TAIL_THREADS, 0, s>>>(a, fa, fb, counts, bar, nlev, trace);
        else k_dag_tail<1, true><<<1, TAIL_THREADS, 0, s>>>(a, fa, fb, counts, bar, nlev, trace);
      } else {
        if (mode == 0) k_dag_tail<0, false><<<1, TAIL_THREADS, 0, s>>>(a, fa, fb, counts, bar, nlev, trace);
        else k_dag_tail<1, false><<<1, TAIL_THREADS, 0, s>>>(a, fa, fb, counts, bar, nlev, trace);
      }
      GOSPL_CHECK(hipGetLastError());
    }
  }
  if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[1], s));
  int32_t* left = fa;  // the frontier lists are free again
  k_dag_unsolved<<<g, GOSPL_BLOCK, 0, s>>>(N, pend, x, dev + 1, left);
  GOSPL_CHECK(hipGetLastError());
  if (mode == 0) {
    // a receiver cycle: iterate on what the sweep left over (nothing to do,
    // and converged at once, when the sweep solved every vertex)
    if (rec) k_flow_cycles<true><<<1, CYCLE_THREADS, 0, s>>>(a, left, dev + 1, m->wd[1], dev + 2);
    else k_flow_cycles<false><<<1, CYCLE_THREADS, 0, s>>>(a, left, dev + 1, m->wd[1], dev + 2);
    GOSPL_CHECK(hipGetLastError());
  }
  GOSPL_CHECK(hipMemcpyAsync(m->hflag, dev, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, s));
  // abort word
  GOSPL_CHECK(hipMemcpyAsync(m->hflag + 34, chain ? m->chain_ctl + CQ_ABORT : (int32_t*)(bar + 2),
                             sizeof(int32_t), hipMemcpyDeviceToHost, s));
  m->dag_pending_chain = chain ? 1 : 0;
  m->dag_pending_trace = vtrace != nullptr;
  m->dag_pending_ltrace = trace != nullptr;
  return 0;
}

static int dag_finish(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                      int32_t* nlevels, int64_t* unsolved, hipStream_t s) {
  const int64_t N = m->N;
  const bool chain = m->dag_pending_chain != 0;
  int32_t* gq = m->chain_gq;
  const int64_t gq_cap = N / 64 + 1;
  long long* vtrace = m->dag_pending_trace ? m->trace_v : nullptr;
  long long* trace = m->dag_pending_ltrace ? m->trace_buf : nullptr;
  GOSPL_CHECK(hipStreamSynchronize(s));
  if (m->hflag[34]) {
    // an aborted dataflow sweep can leave queue entries behind
    if (chain) GOSPL_CHECK(hipMemsetAsync(gq, 0xff, sizeof(int32_t) * gq_cap * 64, s));
    set_error(chain ? "dataflow sweep aborted: a bounded wait ran out"
                    : "receiver-DAG sweep aborted: a grid barrier did not complete");
    return 1;
  }
  if (m->hflag[3] > 0) {
    set_error("%d receivers are not neighbours of their donor (each receiver must be in the donor's "
              "neighbour list)", m->hflag[3]);
    return 1;
  }
  if (mode == 0 && m->hflag[1] > 0 && m->hflag[2]) m->hflag[1] = 0;  // the cycle iteration converged
  if (m->timing) {
    float ms = 0.0f;
    GOSPL_CHECK(hipEventElapsedTime(&ms, m->tev[0], m->tev[1]));
    m->t_ms[mode] += ms;
    m->t_calls[mode] += 1;
  }
  if (vtrace) {
    // solve times relative to the first one: quantiles, and what the last 10 % of the time solved
    std::vector<long long> t(N);
    GOSPL_CHECK(hipMemcpy(t.data(), vtrace, sizeof(long long) * N, hipMemcpyDeviceToHost));
    std::vector<long long> u;
    for (long long x : t)
      if (x) u.push_back(x);
    std::sort(u.begin(), u.end());
    if (!u.empty()) {
      const double span = (u.back() - u.front()) * 0.01;  // us, 100 MHz counter
      auto q = [&](double f) { return (u[(size_t)(f * (u.size() - 1))] - u.front()) * 0.01; };
      size_t late = 0;
      for (long long x : u)
        if ((x - u.front()) * 0.01 > 0.9 * span) late++;
      int retries = 0;
      GOSPL_CHECK(hipMemcpy(&retries, m->chain_ctl + CQ_RETRIES, sizeof(int), hipMemcpyDeviceToHost));
      fprintf(stderr, "[dataflow mode %d] %zu solves over %.0f us; quantiles (us) 50%% %.0f 90%% %.0f "
              "99%% %.0f 99.9%% %.0f; last 10%% of the time: %zu solves; %d incomplete-row passes\n",
              mode, u.size(), span, q(0.5), q(0.9), q(0.99), q(0.999), late, retries);
      // hop latency: a vertex's solve time after the last of its dependencies
      std::vector<int32_t> hr((size_t)nrcv * N);
      std::vector<double> hc((size_t)nrcv * N);
      GOSPL_CHECK(hipMemcpy(hr.data(), rcv, sizeof(int32_t) * nrcv * N, hipMemcpyDeviceToHost));
      GOSPL_CHECK(hipMemcpy(hc.data(), coef, sizeof(double) * nrcv * N, hipMemcpyDeviceToHost));
      std::vector<long long> last(N, 0);
      for (int64_t i = 0; i < N; ++i)
        for (int k = 0; k < nrcv; ++k) {
          const int r = hr[(size_t)k * N + i];
          if (r == (int)i || r < 0 || hc[(size_t)k * N + i] == 0.0) continue;
          // flow: i before r; stream power: r before i
          const int64_t from = mode == 0 ? i : r, to = mode == 0 ? r : i;
          if (t[from] > last[to]) last[to] = t[from];
        }
      std::vector<long long> hop, pre, post;
      std::vector<long long> top(N);
      GOSPL_CHECK(hipMemcpy(top.data(), vtrace + N, sizeof(long long) * N, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < N; ++i)
        if (t[i] && last[i]) {
          hop.push_back(t[i] - last[i]);
          pre.push_back(top[i] - last[i]);  // last dependency solved -> pass that solves i starts
          post.push_back(t[i] - top[i]);    // that pass: start -> i solved (record read)
        }
      std::sort(pre.begin(), pre.end());
      std::sort(post.begin(), post.end());
      if (!pre.empty()) {
        auto pq = [&](std::vector<long long>& h, double f) { return h[(size_t)(f * (h.size() - 1))] * 10.0; };
        fprintf(stderr, "[dataflow mode %d] dependency solved -> pass start (ns): 10%% %.0f 50%% %.0f 90%% %.0f; "
                "pass start -> solved: 10%% %.0f 50%% %.0f 90%% %.0f\n", mode, pq(pre, 0.1), pq(pre, 0.5),
                pq(pre, 0.9), pq(post, 0.1), pq(post, 0.5), pq(post, 0.9));
      }
      std::sort(hop.begin(), hop.end());
      if (!hop.empty()) {
        auto hq = [&](double f) { return hop[(size_t)(f * (hop.size() - 1))] * 10.0; };
        fprintf(stderr, "[dataflow mode %d] hop after last dependency (ns): 10%% %.0f 50%% %.0f 90%% %.0f "
                "99%% %.0f\n", mode, hq(0.1), hq(0.5), hq(0.9), hq(0.99));
      }
    }
  }
  if (trace) {
    static long long h[2 * 8192];
    GOSPL_CHECK(hipMemcpy(h, trace, sizeof(h), hipMemcpyDeviceToHost));
    const int L = m->hflag[0] < 8191 ? m->hflag[0] : 8191;
    double tb[4] = {0, 0, 0, 0};
    long long nb[4] = {0, 0, 0, 0};
    for (int l = 1; l < L; ++l) {  // level 0 runs in its own kernel
      const double us = (h[2 * (l + 1)] - h[2 * l]) * 0.01;  // 100 MHz counter
      const long long n = h[2 * l + 1];
      const int b = n >= 100000 ? 0 : n >= 10000 ? 1 : n >= 1000 ? 2 : 3;
      tb[b] += us;
      nb[b] += 1;
    }
    fprintf(stderr, "[dag mode %d] levels %d: >=1e5: %lld lv %.0f us | 1e4-1e5: %lld lv %.0f us | "
            "1e3-1e4: %lld lv %.0f us | <1e3: %lld lv %.0f us\n", mode, L, nb[0], tb[0], nb[1], tb[1],
            nb[2], tb[2], nb[3], tb[3]);
  }
  if (nlevels) *nlevels = m->hflag[0];
  if (unsolved) *unsolved = m->hflag[1];
  return 0;
}

static int dag_solve(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                     const double* diag, const double* rhs, double* x, int32_t* nlevels,
                     int64_t* unsolved, hipStream_t s) {
  if (dag_issue(m, mode, nrcv, rcv, coef, diag, rhs, x, s)) return 1;
  return dag_finish(m, mode, nrcv, rcv, coef, nlevels, unsolved, s);
}

int spl_coef_launch(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* wgt, const double* dist,
                    const double* fal, const double* kfac, double K, double mexp, double dt, double sl,
                    const double* lfill, const double* h, double* off, double* diag, hipStream_t s) {
  k_spl_coef<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, nrcv, rcv, wgt, dist, fal, kfac, K, mexp, nullptr, dt, sl,
                                                   lfill, h, off, diag);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int spl_update_launch(gospl_mesh* m, const double* x, const int8_t* lsink, double dt, double* h, double* cumed,
                      double* eb, double* eblocal, hipStream_t s) {
  k_spl_update<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, x, m->outlet, m->flat, lsink, dt, h, cumed, eb, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int nl_kbr_launch(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* h, const double* fal, const double* kfac,
                  double K, double mexp, double dt, double sl, const double* lfill, double* kbr, hipStream_t s) {
  k_nl_kbr<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, nrcv, rcv, h, fal, kfac, K, mexp, dt, sl, lfill, kbr);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int dag_solve_ext(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef, const double* diag,
                  const double* rhs, double* x, int64_t* unsolved, hipStream_t s) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("receiver-DAG solve: nrcv must be in 1..12");
    return 1;
  }
  return dag_solve(m, mode, nrcv, rcv, coef, diag, rhs, x, nullptr, unsolved, s);
}

int flowacc_issue(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* wgt, const double* b, double* q,
                  hipStream_t s) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_flowacc: nrcv must be in 1..12");
    return 1;
  }
  return dag_issue(m, 0, nrcv, rcv, wgt, nullptr, b, q, s);
}

int flowacc_finish(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* wgt, int32_t* nlevels,
                   int64_t* unsolved, hipStream_t s) {
  return dag_finish(m, 0, nrcv, rcv, wgt, nlevels, unsolved, s);
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_dag_init_occupancy(gospl_mesh* m) {
  // Launch shape of the sweep past level 0 (env overrides for tuning):
  // GOSPL_DAG_BLOCKS x GOSPL_DAG_THREADS co-resident threads, frontier slices
  // interleaved over blocks, frontiers below GOSPL_DAG_TAIL on one workgroup.
  m->dag_threads = 256;
  const char* te_threads = getenv("GOSPL_DAG_THREADS");
  if (te_threads) {
    const int t = atoi(te_threads);
    if (t >= 64 && t <= DAG_BLOCK && t % 64 == 0) m->dag_threads = t;
  }
  m->dag_interleave = 1;
  const char* te_il = getenv("GOSPL_DAG_INTERLEAVE");
  if (te_il) m->dag_interleave = atoi(te_il) > 0;
  m->dag_tail = 256;
  m->dag_coop = 1;
  const char* te_coop = getenv("GOSPL_DAG_COOP");
  if (te_coop) m->dag_coop = atoi(te_coop) != 0;
  const char* te_tail = getenv("GOSPL_DAG_TAIL");
  if (te_tail && atoi(te_tail) >= 0) m->dag_tail = atoi(te_tail);
  int per_cu = 1 << 30;
  // GOSPL_DAG_CSR=1 keeps the CSR rows even for three receivers or fewer
  const char* te_csr = getenv("GOSPL_DAG_CSR");
  m->dag_csr = te_csr && atoi(te_csr) > 0;
  const void* ks[4] = {(const void*)k_dag_level<0, false>, (const void*)k_dag_level<1, false>,
                       (const void*)k_dag_level<0, true>, (const void*)k_dag_level<1, true>};
  for (int i = 0; i < 4; ++i) {
    int b = 0;
    GOSPL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, ks[i], m->dag_threads, 0));
    per_cu = b < per_cu ? b : per_cu;
  }
  if (per_cu < 1) per_cu = 1;
  const int maxb = per_cu * m->num_cu;
  // dataflow sweeps (GOSPL_DAG_CHAIN=0: level by level); one workgroup per
  // CU (GOSPL_CHAIN_BLOCKS overrides).  More resident wavefronts do not
  // shorten the dependency chains: they share the SIMDs and the memory
  // pipeline with the chain that is on the longest path (two per CU: flow
  // 1.27 / 0.99 ms, erosion 1.37 / 1.65 ms on early / late steps of the
  // 2.6M-vertex sphere, against 1.19 / 0.86 and 1.27 / 1.60 with one).
  const char* te_chain = getenv("GOSPL_DAG_CHAIN");
  m->dag_chain = !(te_chain && atoi(te_chain) == 0);
  const void* kc[2] = {(const void*)k_chain<0>, (const void*)k_chain<1>};
  const char* te_cb = getenv("GOSPL_CHAIN_BLOCKS");
  for (int i = 0; i < 2; ++i) {
    int b = 0;
    GOSPL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kc[i], CHAIN_THREADS, 0));
    if (b < 1) {
      set_error("the dataflow kernel does not fit on a compute unit");
      return 1;
    }
    m->chain_resident[i] = b * m->num_cu;
    m->chain_blocks[i] = m->num_cu;
    if (te_cb && atoi(te_cb) > 0) m->chain_blocks[i] = atoi(te_cb);
    // every workgroup resident at once
    if (m->chain_blocks[i] > m->chain_resident[i]) m->chain_blocks[i] = m->chain_resident[i];
  }
  // pool capacity per wavefront (GOSPL_CHAIN_POOL, 128..512) and the pool
  // size kept when batches go to the global queue (GOSPL_CHAIN_SHARE,
  // 0..capacity - 64: a batch of 64 leaves once the pool holds share + 64)
  m->chain_pool = CHAIN_POOL;
  const char* te_pool = getenv("GOSPL_CHAIN_POOL");
  if (te_pool && atoi(te_pool) >= 128 && atoi(te_pool) <= CHAIN_POOL) m->chain_pool = atoi(te_pool);
  // 0: a batch leaves as soon as the pool holds one.  Erosion fans out
  // upstream, and a vertex kept in a busy wavefront's pool waits there; on
  // the 2.6M-vertex sphere the erosion sweep takes 1.37 ms (early steps) and
  // 1.64 ms (steps 100-160) against 1.46 and 2.29 ms when 128 are kept.
  m->chain_share = 0;
  const char* te_share = getenv("GOSPL_CHAIN_SHARE");
  if (te_share && atoi(te_share) >= 0) m->chain_share = atoi(te_share);
  if (m->chain_share > m->chain_pool - 64) m->chain_share = m->chain_pool - 64;
  // bound of every wait in the dataflow kernel (GOSPL_CHAIN_SPIN_MAX: small
  // values force the abort path, for testing it)
  m->chain_spin = CHAIN_SPIN_MAX;
  const char* te_spin = getenv("GOSPL_CHAIN_SPIN_MAX");
  if (te_spin && atoi(te_spin) >= 0) m->chain_spin = atoi(te_spin);
  // idle wavefronts that stay to poll the global queue (GOSPL_CHAIN_POLLERS)
  m->chain_pollers = CHAIN_POLLERS;
  const char* te_pol = getenv("GOSPL_CHAIN_POLLERS");
  if (te_pol && atoi(te_pol) >= 0) m->chain_pollers = atoi(te_pol);

  // Few blocks: the grid barrier costs ~3 us at 32 blocks and ~18 us at 256 on
  // MI355X, and the frontiers past level 0 hold a few thousand vertices.
  int want = 32;
  const char* env = getenv("GOSPL_DAG_BLOCKS");
  if (env && atoi(env) > 0) want = atoi(env);
  m->coop_blocks = want < maxb ? want : maxb;
  return 0;
}

extern "C" {

int gospl_sweep_timing(gospl_mesh_t* m, int32_t enable, double* fa_ms, int64_t* fa_calls,
                       double* spl_ms, int64_t* spl_calls) {
  if (enable && !m->tev[0]) {
    GOSPL_CHECK(hipEventCreate(&m->tev[0]));
    GOSPL_CHECK(hipEventCreate(&m->tev[1]));
  }
  // exact sweeps and refined solves of the same system are reported together
  if (fa_ms) *fa_ms = m->t_ms[0] + m->t_ms[2];
  if (fa_calls) *fa_calls = m->t_calls[0] + m->t_calls[2];
  if (spl_ms) *spl_ms = m->t_ms[1] + m->t_ms[3];
  if (spl_calls) *spl_calls = m->t_calls[1] + m->t_calls[3];
  for (int k = 0; k < 4; ++k) {
    m->t_ms[k] = 0.0;
    m->t_calls[k] = 0;
  }
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

int gospl_flowacc_refine(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcv, const double* wgt,
                         const double* b, double* q, double rtol, int32_t maxpass, int32_t* passes,
                         double* relres, int64_t* unsolved, gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_flowacc_refine: nrcv must be in 1..12");
    return 1;
  }
  if (passes) *passes = 0;
  if (relres) *relres = 0.0;
  if (unsolved) *unsolved = 0;
  if (nrcv <= 3 && m->W * nrcv <= 64) {
    RefineStatus st{};
    if (refine_solve(m, 0, nrcv, rcv, wgt, nullptr, b, q, rtol, maxpass, &st, stream)) return 1;
    if (st.aborted) {
      set_error("refined flow solve aborted: a bounded wait ran out");
      return 1;
    }
    if (!st.fallback) {
      if (passes) *passes = st.passes;
      if (relres) *relres = st.relres;
      return 0;
    }
  }
  // more than three receivers, a receiver cycle or an overflow: the exact sweep
  if (passes) *passes = -1;
  return dag_solve(m, 0, nrcv, rcv, wgt, nullptr, b, q, nullptr, unsolved, stream);
}

int gospl_spl_refine(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                     const double* disti, const double* fal, const double* kfac, double K, double mexp,
                     double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
                     double* cumed, double* eb, double* eblocal, double rtol, int32_t maxpass,
                     int32_t* passes, double* relres, gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_spl_refine: nrcv must be in 1..12");
    return 1;
  }
  if (gospl_ensure_coef(m, nrcv)) return 1;
  double* offc = m->wcoef;
  double* diag = m->wd[3];
  double* x = m->wd[2];
  if (passes) *passes = 0;
  if (relres) *relres = 0.0;
  k_spl_coef<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, nrcv, rcvi, wgti, disti, fal, kfac, K,
                                                        mexp, nullptr, dt, sealevel, lfill, h, offc, diag);
  GOSPL_CHECK(hipGetLastError());
  // the iterate starts at the old surface
  GOSPL_CHECK(hipMemcpyAsync(x, h, sizeof(double) * m->N, hipMemcpyDeviceToDevice, stream));
  bool exact = true;
  if (nrcv <= 3 && m->W * nrcv <= 64) {
    RefineStatus st{};
    if (refine_solve(m, 1, nrcv, rcvi, offc, diag, h, x, rtol, maxpass, &st, stream)) return 1;
    if (st.aborted) {
      set_error("refined stream-power solve aborted: a bounded wait ran out");
      return 1;
    }
    exact = st.fallback != 0;
    if (!exact) {
      if (passes) *passes = st.passes;
      if (relres) *relres = st.relres;
    }
  }
  if (exact) {
    if (passes) *passes = -1;
    if (dag_solve(m, 1, nrcv, rcvi, offc, diag, h, x, nullptr, nullptr, stream)) return 1;
  }
  k_spl_update<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, x, m->outlet, m->flat, lsink, dt, h,
                                                          cumed, eb, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

static int spl_step(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                    const double* disti, const double* fal, const double* kfac, double K, double mexp,
                    const double* mvec, double dt, double sealevel, const double* lfill, const int8_t* lsink,
                    double* h, double* cumed, double* eb, double* eblocal, int32_t* nlevels,
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
                                                        mexp, mvec, dt, sealevel, lfill, h, offc, diag);
  GOSPL_CHECK(hipGetLastError());
  int rc = dag_solve(m, 1, nrcv, rcvi, offc, diag, h, x, nlevels, nullptr, stream);
  if (rc) return rc;
  k_spl_update<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, x, m->outlet, m->flat, lsink, dt, h,
                                                          cumed, eb, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_spl(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti,
              const double* disti, const double* fal, const double* kfac, double K, double mexp,
              double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
              double* cumed, double* eb, double* eblocal, int32_t* nlevels,
              gospl_stream_t stream) {
  return spl_step(m, nrcv, rcvi, wgti, disti, fal, kfac, K, mexp, nullptr, dt, sealevel, lfill, lsink, h, cumed,
                  eb, eblocal, nlevels, stream);
}

int gospl_spl_batch(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                    const double* disti, const double* fal, const double* kfac, double K, const double* mvec,
                    double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
                    double* cumed, double* eb, double* eblocal, int32_t* nlevels, gospl_stream_t stream) {
  if (!mvec) {
    set_error("gospl_spl_batch: mvec (the per-vertex exponent) is required");
    return 1;
  }
  return spl_step(m, nrcv, rcvi, wgti, disti, fal, kfac, K, 0.0, mvec, dt, sealevel, lfill, lsink, h, cumed, eb,
                  eblocal, nlevels, stream);
}

int gospl_dag_set_shape(gospl_mesh_t* m, int32_t chain_blocks, int32_t chain_share, int32_t chain_pollers) {
  if (!m) {
    set_error("gospl_dag_set_shape: null handle");
    return 1;
  }
  const int lim = m->chain_resident[0] < m->chain_resident[1] ? m->chain_resident[0] : m->chain_resident[1];
  if (chain_blocks > lim) {
    set_error("gospl_dag_set_shape: %d workgroups exceed the %d that fit on the device at once", (int)chain_blocks,
              lim);
    return 1;
  }
  if (chain_share > m->chain_pool - 64 || chain_pollers > 4096) {
    set_error("gospl_dag_set_shape: share must be <= %d and pollers <= 4096", m->chain_pool - 64);
    return 1;
  }
  if (chain_blocks > 0) m->chain_blocks[0] = m->chain_blocks[1] = chain_blocks;
  if (chain_share >= 0) m->chain_share = chain_share;
  if (chain_pollers >= 0) m->chain_pollers = chain_pollers;
  return 0;
}

int gospl_dag_get_shape(gospl_mesh_t* m, int32_t* out) {
  if (!m || !out) {
    set_error("gospl_dag_get_shape: null argument");
    return 1;
  }
  out[0] = m->chain_blocks[0];
  out[1] = m->chain_share;
  out[2] = m->chain_pollers;
  out[3] = m->chain_resident[0] < m->chain_resident[1] ? m->chain_resident[0] : m->chain_resident[1];
  out[4] = m->num_cu;
  return 0;
}

int gospl_spl_nl(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti, const double* disti,
                 const double* fal, const double* kfac, double K, double mexp, double nexp, double dt,
                 double sealevel, const double* lfill, const int8_t* lsink, double* h, double* cumed, double* eb,
                 double* eblocal, double tol, int32_t maxit, int32_t* iters, double* step_rel,
                 gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_spl_nl: nrcv must be in 1..12");
    return 1;
  }
  if (!(nexp > 0.0)) {
    set_error("gospl_spl_nl: n must be positive");
    return 1;
  }
  // coefficient rows: the Jacobian's nrcv off-diagonal rows, then Kbr
  if (gospl_ensure_coef(m, nrcv + 1)) return 1;
  const int64_t N = m->N;
  const int g = nblocks(N);
  double* off = m->wcoef;
  double* kbr = m->wcoef + (int64_t)nrcv * (N + 64);
  double* diag = m->wd[3];
  double* rhs = m->wd[2];
  double* d = m->wd[0];
  double* x = m->wd[1];  // the iterate (the sweeps' cycle scratch is not used in erosion mode)
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(m->dflag + 56);
  k_nl_kbr<<<g, GOSPL_BLOCK, 0, stream>>>(N, nrcv, rcvi, h, fal, kfac, K, mexp, dt, sealevel, lfill, kbr);
  GOSPL_CHECK(hipMemcpyAsync(x, h, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
  GOSPL_CHECK(hipGetLastError());
  int it = 0;
  double rel = INFINITY;
  while (it < maxit) {
    k_nl_jac<<<g, GOSPL_BLOCK, 0, stream>>>(N, nrcv, rcvi, wgti, disti, x, h, kbr, nexp, off, diag, rhs);
    GOSPL_CHECK(hipGetLastError());
    if (dag_solve(m, 1, nrcv, rcvi, off, diag, rhs, d, nullptr, nullptr, stream)) return 1;
    GOSPL_CHECK(hipMemsetAsync(keys, 0, sizeof(unsigned long long) * 2, stream));
    k_nl_step<<<g, GOSPL_BLOCK, 0, stream>>>(N, d, x, keys);
    GOSPL_CHECK(hipGetLastError());
    unsigned long long hk[2];
    GOSPL_CHECK(hipMemcpyAsync(m->hflag + 56, keys, sizeof(hk), hipMemcpyDeviceToHost, stream));
    GOSPL_CHECK(hipStreamSynchronize(stream));
    std::memcpy(hk, m->hflag + 56, sizeof(hk));
    ++it;
    const double dmax = dkey_inv_host(hk[0]), xmax = dkey_inv_host(hk[1]);
    rel = dmax / (xmax > 1.0 ? xmax : 1.0);
    if (!(rel == rel)) break;  // NaN
    if (rel <= tol) break;
  }
  if (iters) *iters = it;
  if (step_rel) *step_rel = rel;
  if (!(rel <= tol)) {
    set_error("gospl_spl_nl: Newton did not converge in %d iterations (last step %.3g of max|h|)", it, rel);
    return 1;
  }
  k_spl_update<<<g, GOSPL_BLOCK, 0, stream>>>(N, x, m->outlet, m->flat, lsink, dt, h, cumed, eb, eblocal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_flowacc_adjoint(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcv, const double* wgt,
                          const double* lam, double* y, gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_flowacc_adjoint: nrcv must be in 1..12");
    return 1;
  }
  if (gospl_ensure_coef(m, nrcv)) return 1;
  k_fa_adj_prep<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, nrcv, rcv, wgt, m->wcoef, m->wd[3]);
  GOSPL_CHECK(hipGetLastError());
  int64_t bad = 0;
  if (dag_solve(m, 1, nrcv, rcv, m->wcoef, m->wd[3], lam, y, nullptr, &bad, stream)) return 1;
  if (bad) {
    set_error("gospl_flowacc_adjoint: the receivers hold %lld vertices on cycles", (long long)bad);
    return 1;
  }
  return 0;
}

int gospl_spl_adjoint(gospl_mesh_t* m, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                      const double* disti, const double* fal, const double* kfac, double K, double mexp,
                      double dt, double sealevel, const double* lfill, const int8_t* lsink, const double* h,
                      const double* x, const double* lam, double* grad_h, double* grad_logk, double* grad_fal,
                      gospl_stream_t stream) {
  if (nrcv < 1 || nrcv > GOSPL_MAXW) {
    set_error("gospl_spl_adjoint: nrcv must be in 1..12");
    return 1;
  }
  // coefficient rows: off (nrcv rows), transposed weights (nrcv rows), local
  // terms (nrcv rows), mu (one row; wd[1] is the sweep's cycle scratch)
  if (gospl_ensure_coef(m, 3 * nrcv + 1)) return 1;
  const int64_t N = m->N;
  double* off = m->wcoef;
  double* wt = m->wcoef + (int64_t)nrcv * (N + 64);
  double* e = m->wcoef + (int64_t)2 * nrcv * (N + 64);
  double* mu = m->wcoef + (int64_t)3 * nrcv * (N + 64);
  double* diag = m->wd[3];
  double* rhs = m->wd[2];
  const int g = nblocks(N);
  k_spl_coef<<<g, GOSPL_BLOCK, 0, stream>>>(N, nrcv, rcvi, wgti, disti, fal, kfac, K, mexp, nullptr, dt, sealevel,
                                           lfill, h, off, diag);
  GOSPL_CHECK(hipGetLastError());
  int64_t bad = 0;
  if (!x) {
    // the forward solution A x = h, solved again (the forward kept h at
    // flat outlets and sinks, whose x still enters their donors' rows)
    double* xs = m->wd[0];
    if (dag_solve(m, 1, nrcv, rcvi, off, diag, h, xs, nullptr, &bad, stream)) return 1;
    if (bad) {
      set_error("gospl_spl_adjoint: the receivers hold %lld vertices on cycles", (long long)bad);
      return 1;
    }
    x = xs;
  }
  k_spl_adj_prep<<<g, GOSPL_BLOCK, 0, stream>>>(N, nrcv, off, diag, lam, m->outlet, m->flat, lsink, wt, rhs);
  GOSPL_CHECK(hipGetLastError());
  if (dag_solve(m, 0, nrcv, rcvi, wt, nullptr, rhs, mu, nullptr, &bad, stream)) return 1;
  if (bad) {
    set_error("gospl_spl_adjoint: the receivers hold %lld vertices on cycles", (long long)bad);
    return 1;
  }
  k_spl_adj_terms<<<g, GOSPL_BLOCK, 0, stream>>>(N, nrcv, rcvi, wgti, disti, fal, kfac, K, mexp, dt, sealevel, lfill,
                                                h, x, off, diag, mu, e, grad_logk, grad_fal);
  k_spl_adj_gather<<<g, GOSPL_BLOCK, 0, stream>>>(N, m->W, nrcv, m->nbr, m->deg, rcvi, e, mu, lam, m->outlet,
                                                 m->flat, lsink, grad_h);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
