// Refined solves on the receiver DAG: the flow system (I - W^T) q = b
// (flowplex.py:323-478) and the implicit stream-power system (SPL.py:45-114)
// solved by iterative refinement.  Each pass computes the FP64 residual
// r = rhs - A x of the current FP64 iterate x, then a correction A d = r by one
// dataflow sweep in 64-bit fixed point, and x += d; passes stop once
// ||r||_2 <= rtol ||rhs||_2.
//
// The correction sweep accumulates each vertex's inputs in one 64-bit word:
// (unsolved dependencies << 56) | (2^55 + sum), the sum in units of
// u = bound / 2^48, where bound is an a-priori bound on every partial sum of
// the sweep (sum |r| for flow, whose solution operator has entries in [0, 1];
// (max diag - 1) * max |r| for stream power, whose matrix has diag - sum |off|
// = 1 in every row).  A dependency pushes its quantised contribution and
// decrements the count with ONE integer atomic add; the atomic that takes the
// count to zero returns the complete sum to the thread that made it, which
// goes on with that vertex at once.  Its static record was loaded while the
// atomic was in flight, so a hop of the sweep costs one memory round trip
// instead of a record read followed by a push and a decrement.  Integer
// addition is associative, so the sums and the results do not depend on the
// schedule (bit-reproducible runs).
//
// Records (64 bytes, 8 words per vertex):
//   flow:         w0 accumulator, w1 rhs, w2..4 weights (f64),
//                 w5 receivers 0 | 1 << 32, w6 receiver 2
//   stream power: w0 accumulator, w1 rhs, w2 diag, w3 entries | dependent 0
//                 << 32, w4 its coefficient, w5 dependents 1 | 2 << 32, w6, w7
//                 their coefficients (f64); later ones in (odon, ocoef).
// A sweep writes its correction d to dx; the leftover check adds it to x
// (no read of x on the dependent path).
#include <cmath>

#include "dag_common.h"

namespace gospl {

#define RX_W 8                        // words per record
#define RX_SHIFT 56
#define RX_ONE (1ll << RX_SHIFT)      // one dependency in the count field
#define RX_BIAS (1ll << 55)           // offset of the sum field
#define RX_MASK ((1ll << RX_SHIFT) - 1)
#define RX_UNIT_BITS 48               // unit = bound / 2^48
#define RX_QMAX 562949953421312.0     // 2^49: a larger push is an overflow
#define RX_SD 3                       // dependents held in a stream-power record
#define RX_PF 3                       // dependents whose records a stream-power hop prefetches
#define RX_PARTS 4                    // partial sums per workgroup of the residual kernel

// Control block of a refined solve (device memory, 512 bytes).
struct RxCtl {
  long long done;      // the iteration has stopped (converged, passes used up, or error)
  long long fallback;  // overflow, unsolved vertices or a non-finite residual
  long long pass;      // correction sweeps run
  long long nonnbr;    // dependencies on a vertex outside the neighbour list
  double unit;         // fixed-point unit of the current pass
  double inv_unit;
  double relres;       // ||r|| / ||rhs|| of the last residual
  double relres0;      // of the first one
  long long unsolved;
  long long pad[7];
};

struct RxArgs {
  int64_t N;
  long long* rec;        // (N, RX_W) records
  double* x;             // FP64 iterate
  double* dx;            // the sweep's correction (added to x by k_rx_unsolved)
  const int32_t* odon;   // stream power: dependents past the record, (c - RX_SD) * N + v
  const double* ocoef;   // and their coefficients
  const int32_t* cnt0;   // dependencies of each vertex at the start of the sweep
  RxCtl* ctl;
};

__device__ __forceinline__ int rx_cnt(long long acc) { return (int)((unsigned long long)acc >> RX_SHIFT); }
__device__ __forceinline__ long long rx_sum(long long acc) { return (acc & RX_MASK) - RX_BIAS; }
__device__ __forceinline__ double rx_f64(long long w) { return __longlong_as_double(w); }
__device__ __forceinline__ long long rx_bits(double d) { return __double_as_longlong(d); }
__device__ __forceinline__ int rx_lo(long long w) { return (int)(unsigned int)((unsigned long long)w & 0xffffffffull); }
__device__ __forceinline__ int rx_hi(long long w) { return (int)(unsigned int)((unsigned long long)w >> 32); }
__device__ __forceinline__ long long rx_pack(int lo, int hi) {
  return (long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned long long)(unsigned int)lo);
}

__device__ __forceinline__ long long rx_atomic_add(long long* p, long long v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long rx_ld_agent(const long long* p) {
  return __hip_atomic_load(const_cast<long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A contribution in fixed point; out-of-range or non-finite values raise the
// fallback flag (the bound of the unit was violated) and push nothing.
__device__ __forceinline__ long long rx_quant(double v, double inv, RxCtl* ctl) {
  const double t = v * inv;
  if (!(fabs(t) <= RX_QMAX)) {
    ctl->fallback = 1;
    return 0;
  }
  return __double2ll_rn(t);
}

// The static words 1..7 of a record (plain loads: they do not change during
// a sweep; the accumulator, word 0, is read only through atomics or coherent
// loads).
struct RxRec {
  long long w[7];
};
__device__ __forceinline__ void rx_load(const long long* __restrict__ rec, int v, RxRec& r) {
  const longlong2* p = reinterpret_cast<const longlong2*>(rec + (int64_t)v * RX_W);
  const longlong2 a = p[0], b = p[1], c = p[2], d = p[3];
  r.w[0] = a.y;
  r.w[1] = b.x;
  r.w[2] = b.y;
  r.w[3] = c.x;
  r.w[4] = c.y;
  r.w[5] = d.x;
  r.w[6] = d.y;
}

// Row masks of the donors (as k_rec_mask): bit p * NR + k of rowmask[r] is set
// when neighbour p of r sends to r through its receiver slot k; dependencies
// on a vertex that is not a neighbour are counted in *nonnbr.
__global__ void k_rx_mask(int64_t N, int W, int NR, const int32_t* __restrict__ nbr,
                          const int32_t* __restrict__ deg, const int8_t* __restrict__ rev,
                          const int32_t* __restrict__ rcv, const double* __restrict__ coef,
                          unsigned long long* __restrict__ rowmask, long long* __restrict__ nonnbr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int nb[GOSPL_MAXW];
  const int dg = deg[i];
#pragma unroll
  for (int q = 0; q < GOSPL_MAXW; ++q) nb[q] = (q < W && q < dg) ? nbr[(int64_t)q * N + i] : -1;
  for (int k = 0; k < NR; ++k) {
    const int r = rcv[(int64_t)k * N + i];
    if (r == (int)i || r < 0 || coef[(int64_t)k * N + i] == 0.0) continue;
    int q = -1;
#pragma unroll
    for (int t = GOSPL_MAXW - 1; t >= 0; --t)
      if (nb[t] == r) q = t;
    const int p = q >= 0 ? rev[(int64_t)q * N + i] : -1;
    if (p >= 0) atomicOr(&rowmask[r], 1ull << (p * NR + k));
    else atomicAdd(reinterpret_cast<unsigned long long*>(nonnbr), 1ull);
  }
}

// Fixed-order reduction of RX_PARTS values over a workgroup of GOSPL_BLOCK
// threads: sums for slots with max_slot[t] == 0, maxima otherwise.
__device__ __forceinline__ void rx_block_reduce(double (&v)[RX_PARTS], const bool (&is_max)[RX_PARTS],
                                                double* __restrict__ out) {
  __shared__ double sh[RX_PARTS][GOSPL_BLOCK / 64];
#pragma unroll
  for (int t = 0; t < RX_PARTS; ++t)
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_xor(v[t], o, 64);
      v[t] = is_max[t] ? fmax(v[t], w) : v[t] + w;
    }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int t = 0; t < RX_PARTS; ++t) sh[t][threadIdx.x >> 6] = v[t];
  __syncthreads();
  if (threadIdx.x < RX_PARTS) {
    const int t = threadIdx.x;
    double a = sh[t][0];
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) a = is_max[t] ? fmax(a, sh[t][w]) : a + sh[t][w];
    out[t] = a;
  }
}

// Residual r = rhs - A x (FP64) of every vertex, its records for the next
// correction sweep (build != 0), and per-workgroup partials:
//   flow:         sum |r|, sum r^2, sum rhs^2, 0
//   stream power: max |r|, sum r^2, sum rhs^2, max diag
// The first workgroup clears the sweep's small counters.
template <int MODE>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_rx_resid(
    int64_t N, int NR, const int32_t* __restrict__ nbr, const unsigned long long* __restrict__ rowmask,
    const int32_t* __restrict__ rcv, const double* __restrict__ coef, const double* __restrict__ diag,
    const double* __restrict__ rhs, const double* __restrict__ x, int build, long long* __restrict__ rec,
    int32_t* __restrict__ odon, double* __restrict__ ocoef, int32_t* __restrict__ cnt0, double* __restrict__ part,
    const RxCtl* __restrict__ ctl, ZeroList zl) {
  if (ctl->done) return;
  if (blockIdx.x == 0)
    for (int k = 0; k < zl.n; ++k)
      for (int t = threadIdx.x; t < zl.len[k]; t += blockDim.x) zl.p[k][t] = 0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double pv[RX_PARTS] = {0.0, 0.0, 0.0, 0.0};
  const bool is_max[RX_PARTS] = {MODE == 1, false, false, MODE == 1};
  if (i < N) {
    const double b = rhs[i];
    const unsigned long long mk = rowmask[i];
    double r;
    if (MODE == 0) {
      // b - x_i + sum over donors j (slot k) of w_jk x_j
      r = b - x[i];
      unsigned long long m = mk;
      while (m) {
        const int bit = __ffsll((long long)m) - 1;
        m &= m - 1ull;
        const int p = bit / NR, k = bit - p * NR;
        const int j = nbr[(int64_t)p * N + i];
        r = r + coef[(int64_t)k * N + j] * x[j];
      }
    } else {
      const double dg = diag[i];
      r = b - dg * x[i];
      for (int k = 0; k < NR; ++k) {
        const int rk = rcv[(int64_t)k * N + i];
        const double o = coef[(int64_t)k * N + i];
        if (rk != (int)i && rk >= 0 && o != 0.0) r = r - o * x[rk];
      }
      pv[3] = dg;
    }
    pv[0] = fabs(r);
    pv[1] = r * r;
    pv[2] = b * b;
    if (build) {
      long long* rp = rec + i * RX_W;
      if (MODE == 0) {
        int rr[3] = {(int)i, (int)i, (int)i};
        double w[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < NR && k < 3; ++k) {
          rr[k] = rcv[(int64_t)k * N + i];
          w[k] = coef[(int64_t)k * N + i];
        }
        longlong2* q = reinterpret_cast<longlong2*>(rp);
        cnt0[i] = __popcll(mk);
        q[0] = make_longlong2(((long long)__popcll(mk) << RX_SHIFT) + RX_BIAS, rx_bits(r));
        q[1] = make_longlong2(rx_bits(w[0]), rx_bits(w[1]));
        q[2] = make_longlong2(rx_bits(w[2]), rx_pack(rr[0], rr[1]));
        q[3] = make_longlong2(rx_pack(rr[2], 0), 0ll);
      } else {
        int ndep = 0;
        for (int k = 0; k < NR; ++k) {
          const int rk = rcv[(int64_t)k * N + i];
          if (rk != (int)i && rk >= 0 && coef[(int64_t)k * N + i] != 0.0) ndep++;
        }
        // dependents: donors j (slot k) with coefficient -off_jk >= 0
        int dj[RX_SD] = {0, 0, 0};
        double dc[RX_SD] = {0.0, 0.0, 0.0};
        unsigned long long m = mk;
        const int nd = __popcll(mk);
        for (int c = 0; m; ++c) {
          const int bit = __ffsll((long long)m) - 1;
          m &= m - 1ull;
          const int p = bit / NR, k = bit - p * NR;
          const int j = nbr[(int64_t)p * N + i];
          const double cf = -coef[(int64_t)k * N + j];
          if (c < RX_SD) {
#pragma unroll
            for (int t = 0; t < RX_SD; ++t)
              if (t == c) {
                dj[t] = j;
                dc[t] = cf;
              }
          } else {
            odon[(int64_t)(c - RX_SD) * N + i] = j;
            ocoef[(int64_t)(c - RX_SD) * N + i] = cf;
          }
        }
        longlong2* q = reinterpret_cast<longlong2*>(rp);
        cnt0[i] = ndep;
        q[0] = make_longlong2(((long long)ndep << RX_SHIFT) + RX_BIAS, rx_bits(r));
        q[1] = make_longlong2(rx_bits(diag[i]), rx_pack(nd, dj[0]));
        q[2] = make_longlong2(rx_bits(dc[0]), rx_pack(dj[1], dj[2]));
        q[3] = make_longlong2(rx_bits(dc[1]), rx_bits(dc[2]));
      }
    }
  }
  rx_block_reduce(pv, is_max, part + (int64_t)blockIdx.x * RX_PARTS);
}

// Norms of the residual (one workgroup, fixed reduction order), the stop test
// and the next pass's fixed-point unit.
template <int MODE>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_rx_norm(int nblk, const double* __restrict__ part, double rtol,
                                                         int pass, int maxpass, RxCtl* __restrict__ ctl) {
  if (ctl->done) return;
  double pv[RX_PARTS] = {0.0, 0.0, 0.0, 0.0};
  const bool is_max[RX_PARTS] = {MODE == 1, false, false, MODE == 1};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
    for (int t = 0; t < RX_PARTS; ++t) {
      const double w = part[(int64_t)b * RX_PARTS + t];
      pv[t] = is_max[t] ? fmax(pv[t], w) : pv[t] + w;
    }
  __shared__ double tot[RX_PARTS];
  rx_block_reduce(pv, is_max, tot);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double rr = tot[1], bb = tot[2];
  const double relres = bb > 0.0 ? sqrt(rr / bb) : (rr > 0.0 ? INFINITY : 0.0);
  ctl->relres = relres;
  if (pass == 0) ctl->relres0 = relres;
  if (!isfinite(relres) || !isfinite(tot[0]) || !isfinite(tot[3])) {
    ctl->fallback = 1;
    ctl->done = 1;
    return;
  }
  if (relres <= rtol || pass >= maxpass) {
    ctl->done = 1;
    return;
  }
  const double bound = MODE == 0 ? tot[0] : fmax(tot[3] - 1.0, 1.0) * tot[0];
  if (!(bound > 0.0)) {
    ctl->done = 1;
    return;
  }
  ctl->unit = ldexp(bound, -RX_UNIT_BITS);
  ctl->inv_unit = ldexp(1.0 / bound, RX_UNIT_BITS);
}

// The solved correction of a vertex from its record and input sum; x += d.
template <int MODE>
__device__ __forceinline__ double rx_value(const RxRec& rc, long long S, double unit) {
  const double s = rx_f64(rc.w[0]) + (double)S * unit;
  return MODE == 0 ? s : s / rx_f64(rc.w[1]);
}

// Dependent c of a record: its vertex (-1 if none) and push coefficient.
template <int MODE>
__device__ __forceinline__ void rx_dep(const RxRec& rc, int v, int c, int& d, double& cf) {
  d = -1;
  cf = 0.0;
  if (MODE == 0) {
    const int r = c == 0 ? rx_lo(rc.w[4]) : c == 1 ? rx_hi(rc.w[4]) : rx_lo(rc.w[5]);
    const double w = rx_f64(c == 0 ? rc.w[1] : c == 1 ? rc.w[2] : rc.w[3]);
    if (r != v && r >= 0 && w != 0.0) {
      d = r;
      cf = w;
    }
  } else {
    const int nd = rx_lo(rc.w[2]);
    if (c < nd && c < RX_SD) {
      // constant indices only (a runtime index would put the record in scratch)
      d = c == 0 ? rx_hi(rc.w[2]) : c == 1 ? rx_lo(rc.w[4]) : rx_hi(rc.w[4]);
      cf = rx_f64(c == 0 ? rc.w[3] : c == 1 ? rc.w[5] : rc.w[6]);
    }
  }
}

// Level 0: vertices without dependencies solve on the whole grid; the ones
// they release form the start list of the dataflow kernel.
template <int MODE>
__global__ __launch_bounds__(GOSPL_BLOCK) void k_rx_level0(RxArgs a, int32_t* __restrict__ list,
                                                           int32_t* __restrict__ count) {
  RxCtl* ctl = a.ctl;
  if (ctl->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->pass = ctl->pass + 1;
  const int64_t N = a.N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t - (t & 63) >= N) return;  // whole wavefront past the mesh
  const double unit = ctl->unit, inv = ctl->inv_unit;
  const int v = (int)t;
  bool src = false;
  RxRec rc;
  double d = 0.0;
  if (t < N) {
    // sources by their initial count: the live accumulator of a vertex that
    // another source releases here reaches zero too, and that vertex goes to
    // the start list, not through this kernel
    src = a.cnt0[v] == 0;
    if (src) {
      rx_load(a.rec, v, rc);
      d = rx_value<MODE>(rc, 0, unit);
      a.dx[v] = d;
    }
  }
  // dependents: the three receivers (flow) or the record's dependents and
  // the overflow entries (stream power), one ballot round each
  const int nd_lane = !src ? 0 : (MODE == 0 ? 3 : rx_lo(rc.w[2]));
  int ndmax = nd_lane;
  for (int o = 32; o > 0; o >>= 1) ndmax = max(ndmax, __shfl_xor(ndmax, o, 64));
  for (int c = 0; c < ndmax; ++c) {
    bool rel = false;
    int dv = -1;
    if (c < nd_lane) {
      double cf = 0.0;
      if (MODE == 0 || c < RX_SD) {
        rx_dep<MODE>(rc, v, c, dv, cf);
      } else {
        dv = a.odon[(int64_t)(c - RX_SD) * N + v];
        cf = a.ocoef[(int64_t)(c - RX_SD) * N + v];
      }
      if (dv >= 0) {
        const long long q = rx_quant(cf * d, inv, ctl);
        const long long old = rx_atomic_add(&a.rec[(int64_t)dv * RX_W], q - RX_ONE);
        rel = rx_cnt(old) == 1;
      }
    }
    wave_append(rel, dv, count, list);
  }
}

// Dataflow correction sweep (after level 0).  A lane holds a vertex v with
// its input sum S and static record; it solves v, pushes to v's dependents
// (one atomic each; the records of the first RX_PF / three are loaded
// meanwhile) and goes on with the first dependent it released.  Other
// releases go to the wavefront's LDS pool with their sums; pooled work is
// shared through the global queue of 64-vertex batches (as the FP64 sweep),
// where a claimer reads the sum back from the accumulator.  The scheduling
// (share of the start list, pool, queue, pollers, bounded waits, leaving) is
// that of k_chain.
template <int MODE>
__global__ __launch_bounds__(CHAIN_THREADS) void k_rx_chain(RxArgs a, const int32_t* __restrict__ list,
                                                          const int32_t* __restrict__ nlist,
                                                          int32_t* __restrict__ gq, int64_t gq_cap,
                                                          int32_t* __restrict__ qctl, int pool_cap,
                                                          int pollers, int spin_max) {
  __shared__ int32_t s_pv[CHAIN_THREADS / 64][CHAIN_POOL];
  __shared__ long long s_ps[CHAIN_THREADS / 64][CHAIN_POOL];
  RxCtl* ctl = a.ctl;
  if (ctl->done) return;
  const int64_t N = a.N;
  const double unit = ctl->unit, inv = ctl->inv_unit;
  const int lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  int32_t* pv = s_pv[threadIdx.x >> 6];
  long long* ps = s_ps[threadIdx.x >> 6];
  int np = 0;
  const int64_t nw = (int64_t)gridDim.x * (CHAIN_THREADS / 64);
  const int64_t w = (int64_t)blockIdx.x * (CHAIN_THREADS / 64) + (threadIdx.x >> 6);
  const int64_t n0 = *nlist;
  int64_t cur = w * n0 / nw;
  const int64_t end = (w + 1) * n0 / nw;
  int v = -1;            // vertex of this lane
  bool have_sum = false; // S holds v's complete input sum
  bool have_rec = false; // rc holds v's static record
  long long S = 0;
  RxRec rc;
  int64_t ci = -1;
  int retry = 0, idle = 0;
  bool act = true, poller = false;
  unsigned long long* qword = reinterpret_cast<unsigned long long*>(&qctl[CQ_QUEUE]);
  if (lane == 0) atomicAdd(&qctl[CQ_ACTIVE], 1);
  // move the newest 64 pooled vertices to a global batch (their sums are
  // read back from the accumulators by the claimer)
  auto spill = [&]() {
    int b = 0;
    if (lane == 0) b = (int)(atomicAdd(qword, 1ull << 32) >> 32);
    b = __shfl(b, 0);
    const int u = pv[np - 64 + lane];
    if ((int64_t)b < gq_cap) st_agent_i(&gq[(int64_t)b * 64 + lane], u);
    else if (lane == 0) st_agent_i(&qctl[CQ_ABORT], 1);
    np -= 64;
    wave_sync();
  };
  for (;;) {
    if (ci >= 0) {
      const int q = ld_agent_i(&gq[ci]);
      if (q >= 0) {
        st_agent_i(&gq[ci], -1);
        v = q;
        ci = -1;
        retry = 0;
        have_sum = false;
        have_rec = false;
      }
    }
    const bool need = v < 0 && ci < 0;
    const unsigned long long nm = __ballot(need);
    if (nm != 0ull && (np > 0 || cur < end)) {
      const int rank = __popcll(nm & below);
      const int nn = __popcll(nm);
      const int fp = nn < np ? nn : np;
      if (need && rank < fp) {
        v = pv[np - 1 - rank];
        S = ps[np - 1 - rank];
        have_sum = true;
        have_rec = false;
      }
      const int64_t left = end - cur;
      const int fs = (int)((int64_t)(nn - fp) < left ? (int64_t)(nn - fp) : left);
      if (need && rank >= fp && rank - fp < fs) {
        v = list[cur + (rank - fp)];
        have_sum = false;
        have_rec = false;
      }
      np -= fp;
      cur += fs;
      wave_sync();
    }
    if (__ballot(v >= 0 || ci >= 0) == 0ull) {
      int claimed = -1, stop = 0;
      if (lane == 0) {
        if (act) {
          atomicSub(&qctl[CQ_ACTIVE], 1);
          act = false;
        }
        const unsigned long long qw = __hip_atomic_load(qword, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int h = (int)(unsigned int)qw, t = (int)(qw >> 32);
        if (h < t) {
          if (atomicCAS(qword, qw, qw + 1ull) == qw) {
            atomicAdd(&qctl[CQ_ACTIVE], 1);
            act = true;
            claimed = h;
            if (poller) atomicSub(&qctl[CQ_IDLERS], 1);
            poller = false;
          }
        } else if (ld_agent_i(&qctl[CQ_ABORT]) != 0) {
          stop = 1;
        } else if (!poller) {
          if (atomicAdd(&qctl[CQ_IDLERS], 1) < pollers) poller = true;
          else stop = 1;
        } else if (ld_agent_i(&qctl[CQ_ACTIVE]) == 0) {
          stop = 1;
        }
      }
      claimed = __shfl(claimed, 0);
      stop = __shfl(stop, 0);
      act = __shfl((int)act, 0) != 0;
      poller = __shfl((int)poller, 0) != 0;
      if (claimed >= 0) {
        if ((int64_t)claimed >= gq_cap) {
          if (lane == 0) st_agent_i(&qctl[CQ_ABORT], 1);
          break;
        }
        ci = (int64_t)claimed * 64 + lane;
        idle = 0;
        continue;
      }
      if (stop) break;
      if (++idle > spin_max) {
        if (lane == 0) st_agent_i(&qctl[CQ_ABORT], 1);
        break;
      }
      __builtin_amdgcn_s_sleep(16);
      continue;
    }
    idle = 0;
    // lanes that took a vertex from the start list, the pool or the queue
    // load what they miss: the static record; the sum (a coherent read of the
    // accumulator, complete once its count is zero)
    if (v >= 0 && !have_rec) {
      rx_load(a.rec, v, rc);
      have_rec = true;
    }
    if (v >= 0 && !have_sum) {
      const long long acc = rx_ld_agent(&a.rec[(int64_t)v * RX_W]);
      if (rx_cnt(acc) == 0) {
        S = rx_sum(acc);
        have_sum = true;
      }
    }
    const bool solved = v >= 0 && have_sum;
    int nxt = -1;
    long long nS = 0;
    RxRec nrc;
    bool nrec = false;
    // dependents handled with prefetch: 3 (flow) or RX_PF (stream power)
    constexpr int PF = MODE == 0 ? 3 : RX_PF;
    int dv[PF];
    long long dq[PF], old[PF];
    RxRec prc[PF];
    double d = 0.0;
    int nd = 0;
    if (solved) {
      d = rx_value<MODE>(rc, S, unit);
      a.dx[v] = d;
      nd = MODE == 0 ? 3 : rx_lo(rc.w[2]);
#pragma unroll
      for (int c = 0; c < PF; ++c) {
        double cf = 0.0;
        rx_dep<MODE>(rc, v, c, dv[c], cf);
        dq[c] = dv[c] >= 0 ? rx_quant(cf * d, inv, ctl) - RX_ONE : 0;
      }
      // all atomics and prefetches in flight together
#pragma unroll
      for (int c = 0; c < PF; ++c)
        if (dv[c] >= 0) old[c] = rx_atomic_add(&a.rec[(int64_t)dv[c] * RX_W], dq[c]);
#pragma unroll
      for (int c = 0; c < PF; ++c)
        if (dv[c] >= 0) rx_load(a.rec, dv[c], prc[c]);
    } else {
#pragma unroll
      for (int c = 0; c < PF; ++c) dv[c] = -1;
    }
    // releases: the first one is this lane's next vertex, the others pooled
#pragma unroll
    for (int c = 0; c < PF; ++c) {
      const bool rl = dv[c] >= 0 && rx_cnt(old[c]) == 1;
      const long long sc = rl ? rx_sum(old[c] + dq[c]) : 0;
      const bool keep = rl && nxt < 0;
      if (keep) {
        nxt = dv[c];
        nS = sc;
        nrc = prc[c];
        nrec = true;
      }
      const bool ex = rl && !keep;
      const unsigned long long em = __ballot(ex);
      if (em != 0ull) {
        const int e = __popcll(em);
        while (np + e > pool_cap) spill();
        if (ex) {
          pv[np + __popcll(em & below)] = dv[c];
          ps[np + __popcll(em & below)] = sc;
        }
        np += e;
        wave_sync();
      }
    }
    // stream power: dependents past the prefetched ones (record entries, then
    // the overflow arrays), one ballot round each
    if (MODE == 1) {
      int ndmax = solved ? nd : 0;
      for (int o = 32; o > 0; o >>= 1) ndmax = max(ndmax, __shfl_xor(ndmax, o, 64));
      for (int c = PF; c < ndmax; ++c) {
        bool rl = false;
        int dc = -1;
        long long sc = 0;
        if (solved && c < nd) {
          double cf = 0.0;
          if (c < RX_SD) {
            rx_dep<MODE>(rc, v, c, dc, cf);
          } else {
            dc = a.odon[(int64_t)(c - RX_SD) * N + v];
            cf = a.ocoef[(int64_t)(c - RX_SD) * N + v];
          }
          if (dc >= 0) {
            const long long q = rx_quant(cf * d, inv, ctl) - RX_ONE;
            const long long o = rx_atomic_add(&a.rec[(int64_t)dc * RX_W], q);
            rl = rx_cnt(o) == 1;
            sc = rl ? rx_sum(o + q) : 0;
          }
        }
        const bool keep = rl && nxt < 0;
        if (keep) {
          nxt = dc;
          nS = sc;
          nrec = false;
        }
        const bool ex = rl && !keep;
        const unsigned long long em = __ballot(ex);
        if (em != 0ull) {
          const int e = __popcll(em);
          while (np + e > pool_cap) spill();
          if (ex) {
            pv[np + __popcll(em & below)] = dc;
            ps[np + __popcll(em & below)] = sc;
          }
          np += e;
          wave_sync();
        }
      }
    }
    if (np >= 64) spill();
    const bool waiting = v >= 0 && !solved;
    if (solved) {
      v = nxt;
      S = nS;
      have_sum = nxt >= 0;
      have_rec = nrec;
      if (nrec) rc = nrc;
      retry = 0;
    }
    if ((waiting || ci >= 0) && ++retry > spin_max) st_agent_i(&qctl[CQ_ABORT], 1);
    if (__ballot(waiting || ci >= 0) != 0ull) {
      int ab = 0;
      if (lane == 0) ab = ld_agent_i(&qctl[CQ_ABORT]);
      if (__shfl(ab, 0) != 0) break;
    }
  }
}

// x += dx at the vertices the sweep solved; the ones it did not reach (a
// receiver cycle and what lies below it) are counted, and any makes the solve
// fall back to the exact FP64 sweep.
__global__ void k_rx_unsolved(int64_t N, const long long* __restrict__ rec, const double* __restrict__ dx,
                              double* __restrict__ x, RxCtl* __restrict__ ctl) {
  if (ctl->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool left = i < N && rx_cnt(rec[i * RX_W]) != 0;
  if (i < N && !left) x[i] = x[i] + dx[i];
  const unsigned long long m = __ballot(left);
  if (m != 0ull && __lane_id() == 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->unsolved), (unsigned long long)__popcll(m));
    ctl->fallback = 1;
  }
}

// Stops the iteration after a sweep that left vertices or overflowed.
__global__ void k_rx_check(RxCtl* __restrict__ ctl) {
  if (threadIdx.x == 0 && !ctl->done && ctl->fallback) ctl->done = 1;
}

int refine_solve(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                 const double* diag, const double* rhs, double* x, double rtol, int maxpass,
                 RefineStatus* st, hipStream_t s) {
  const int64_t N = m->N;
  if (nrcv < 1 || nrcv > 3 || m->W * nrcv > 64) {
    set_error("refined solve: needs 1..3 receivers (got %d)", nrcv);
    return 1;
  }
  if (maxpass < 1) maxpass = 1;
  if (ensure_records(m, s)) return 1;
  if (ensure_donors(m, (int64_t)nrcv * m->W * N)) return 1;
  if (!m->rx_ctl) {
    GOSPL_CHECK(hipMalloc(&m->rx_ctl, sizeof(double) * 64));
    GOSPL_CHECK(hipHostMalloc((void**)&m->rx_hctl, sizeof(double) * 64, hipHostMallocDefault));
  }
  const int64_t gq_cap = N / 64 + 1;
  if (!m->chain_ctl) {
    GOSPL_CHECK(hipMalloc(&m->chain_ctl, sizeof(int32_t) * CQ_WORDS));
    GOSPL_CHECK(hipMalloc(&m->chain_gq, sizeof(int32_t) * gq_cap * 64));
    GOSPL_CHECK(hipMemsetAsync(m->chain_gq, 0xff, sizeof(int32_t) * gq_cap * 64, s));
  }
  const int g = nblocks(N);
  RxCtl* ctl = reinterpret_cast<RxCtl*>(m->rx_ctl);
  unsigned long long* rowmask = m->wu[0];
  double* part = m->wd[0];  // RX_PARTS per workgroup <= N + 64 doubles
  int32_t* list = m->wi[2];
  int32_t* counts = m->counts;
  long long* rec = reinterpret_cast<long long*>(m->wrec);
  int32_t* odon = m->wdon;
  double* ocoef = m->wacc;
  double* dx = m->wd[1];
  GOSPL_CHECK(hipMemsetAsync(ctl, 0, sizeof(RxCtl), s));
  GOSPL_CHECK(hipMemsetAsync(rowmask, 0, sizeof(unsigned long long) * N, s));
  k_rx_mask<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, nrcv, m->nbr, m->deg, m->rev, rcv, coef, rowmask, &ctl->nonnbr);
  GOSPL_CHECK(hipGetLastError());
  ZeroList zl{{counts, m->chain_ctl, nullptr, nullptr}, {2, CQ_WORDS, 0, 0}, 2};
  int32_t* cnt0 = m->wi[0];
  RxArgs a{N, rec, x, dx, odon, ocoef, cnt0, ctl};
  auto residual = [&](int pass) {
    const int build = pass < maxpass;
    if (mode == 0) {
      k_rx_resid<0><<<g, GOSPL_BLOCK, 0, s>>>(N, nrcv, m->nbr, rowmask, rcv, coef, diag, rhs, x, build, rec, odon,
                                              ocoef, cnt0, part, ctl, zl);
      k_rx_norm<0><<<1, GOSPL_BLOCK, 0, s>>>(g, part, rtol, pass, maxpass, ctl);
    } else {
      k_rx_resid<1><<<g, GOSPL_BLOCK, 0, s>>>(N, nrcv, m->nbr, rowmask, rcv, coef, diag, rhs, x, build, rec, odon,
                                              ocoef, cnt0, part, ctl, zl);
      k_rx_norm<1><<<1, GOSPL_BLOCK, 0, s>>>(g, part, rtol, pass, maxpass, ctl);
    }
  };
  // pass 0's residual decides whether the first correction runs; after each
  // correction the next residual is read back (one synchronisation per pass
  // run), so only the passes that run are launched and timed
  residual(0);
  GOSPL_CHECK(hipGetLastError());
  for (int pass = 0; pass < maxpass; ++pass) {
    if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[0], s));
    if (mode == 0) {
      k_rx_level0<0><<<g, GOSPL_BLOCK, 0, s>>>(a, list, counts + 1);
      k_rx_chain<0><<<m->chain_blocks[0], CHAIN_THREADS, 0, s>>>(a, list, counts + 1, m->chain_gq, gq_cap,
                                                                m->chain_ctl, m->chain_pool, m->chain_pollers,
                                                                m->chain_spin);
    } else {
      k_rx_level0<1><<<g, GOSPL_BLOCK, 0, s>>>(a, list, counts + 1);
      k_rx_chain<1><<<m->chain_blocks[1], CHAIN_THREADS, 0, s>>>(a, list, counts + 1, m->chain_gq, gq_cap,
                                                                m->chain_ctl, m->chain_pool, m->chain_pollers,
                                                                m->chain_spin);
    }
    if (m->timing) GOSPL_CHECK(hipEventRecord(m->tev[1], s));
    k_rx_unsolved<<<g, GOSPL_BLOCK, 0, s>>>(N, rec, dx, x, ctl);
    k_rx_check<<<1, 64, 0, s>>>(ctl);
    residual(pass + 1);
    GOSPL_CHECK(hipGetLastError());
    GOSPL_CHECK(hipMemcpyAsync(m->rx_hctl, ctl, sizeof(RxCtl), hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipMemcpyAsync(m->hflag + 34, m->chain_ctl + CQ_ABORT, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    const RxCtl* hc = reinterpret_cast<const RxCtl*>(m->rx_hctl);
    if (m->timing && hc->pass > pass) {
      float ms = 0.0f;
      GOSPL_CHECK(hipEventElapsedTime(&ms, m->tev[0], m->tev[1]));
      m->t_ms[2 + mode] += ms;
      m->t_calls[2 + mode] += 1;
    }
    if (hc->done || m->hflag[34]) break;
  }
  const RxCtl* h = reinterpret_cast<const RxCtl*>(m->rx_hctl);
  if (h->nonnbr > 0) {
    set_error("%lld receivers are not neighbours of their donor (each receiver must be in the donor's "
              "neighbour list)", (long long)h->nonnbr);
    return 1;
  }
  st->passes = (int)h->pass;
  st->relres = h->relres;
  st->fallback = h->fallback != 0;
  st->aborted = m->hflag[34] != 0;
  if (st->aborted) GOSPL_CHECK(hipMemsetAsync(m->chain_gq, 0xff, sizeof(int32_t) * gq_cap * 64, s));
  return 0;
}

}  // namespace gospl
