// Multiple-flow-direction receivers and _buildFlowDirection.
//
// Receivers: one thread per vertex; the <= 12 neighbour candidates live in
// registers (loops fully unrolled over the 12-slot stencil), the top-nrcv
// selection ranks candidates by pairwise comparison.  goSPL orders them with
// a Fortran quicksort (functions.F90:683-762) whose result is unique when
// the perturbed slopes are distinct; when two candidates carry exactly equal
// values the kernel replays that quicksort on a private copy so the tie order
// is the reference's.
//

#include "common.h"

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
      m->N, m->W, m->nbr, m->deg, m->elgt, m->tgid, m->outlet, m->flat, fexp, h, sl, lfill, post,
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

}  // namespace gospl

using namespace gospl;

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

}  // extern "C"
