// goSPL's FAMesh.flowAccumulation with its pit-overflow cascade
// (gospl/flow/flowplex.py:576-686, 752-945) driven from native code: the
// device entry points of this library plus the cascade's host decisions,
// so that the step's orchestration adds no interpreter time between its
// synchronisations.  The decisions are goSPL's (flowplex.py:596-686,
// 849-904): per pass, pit inflow = sum of FA over the pit's sink cells;
// eV = inflow - remaining capacity; a pit with eV > 0 fills, spills eV at its
// spill cell (a cell shared by several pits keeps the value of the last of
// them) and empties its capacity; a pit with eV < 0 and inflow keeps eV less
// capacity and takes the listed water level whose volume is closest; the
// receivers are rebuilt on the water surface (on the filled surface at pass
// 100) when some eV > 1e-3, and the spilled volumes are routed when their
// sum / dt exceeds the largest cell area; passes stop when nothing is
// routed, after 100, or after 3 passes without a 0.1 % improvement of
// that sum.  The sum is numpy's pairwise summation, in its order, so every
// decision matches the Python driver bit for bit.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace gospl {

__global__ void k_scale_copy(int64_t N, const double* __restrict__ a, double s, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = a[i] * s;
}

__global__ void k_sub_inplace(int64_t N, double* __restrict__ a, const double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) a[i] = a[i] - b[i];
}

// numpy's pairwise summation of a contiguous float64 run (add.reduce's
// inner loop: blocks of at most 128 summed with eight accumulators,
// halves split at multiples of 8 above that).
static double np_pairwise_sum(const double* a, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return np_pairwise_sum(a, n2) + np_pairwise_sum(a + n2, n - n2);
}

// pinned host tables / staging and device staging of at least n doubles
static int ensure_flow_bufs(gospl_mesh* m, int64_t n, hipStream_t s) {
  if (m->flow_cap >= n) return 0;
  GOSPL_CHECK(hipStreamSynchronize(s));  // the old buffers may be in flight
  if (m->flow_htab) GOSPL_CHECK(hipHostFree(m->flow_htab));
  if (m->flow_hstage) GOSPL_CHECK(hipHostFree(m->flow_hstage));
  if (m->flow_dstage) GOSPL_CHECK(hipFree(m->flow_dstage));
  m->flow_htab = m->flow_hstage = m->flow_dstage = nullptr;
  m->flow_cap = 0;
  const int64_t cap = n < 4096 ? 4096 : 2 * n;
  GOSPL_CHECK(hipHostMalloc((void**)&m->flow_htab, sizeof(double) * cap, hipHostMallocDefault));
  GOSPL_CHECK(hipHostMalloc((void**)&m->flow_hstage, sizeof(double) * cap, hipHostMallocDefault));
  GOSPL_CHECK(hipMalloc((void**)&m->flow_dstage, sizeof(double) * cap));
  m->flow_cap = cap;
  return 0;
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_flow_tables(gospl_mesh_t* m, const double** tables, int64_t* n1) {
  if (!m || !tables || !n1) {
    set_error("gospl_flow_tables: null argument");
    return 1;
  }
  *tables = m->flow_htab;
  *n1 = m->flow_n1;
  return 0;
}

// np.sum of a contiguous float64 array: add.reduce works through its
// buffer of 8192 elements, adding each chunk's pairwise sum in turn.
static double np_sum(const double* a, int64_t n) {
  double r = 0.0;
  for (int64_t k = 0; k < n; k += 8192) r += np_pairwise_sum(a + k, n - k < 8192 ? n - k : 8192);
  return r;
}

extern "C" double gospl_np_sum(const double* a, int64_t n) { return np_sum(a, n); }

extern "C" int gospl_flow_accumulation(gospl_mesh_t* m, const gospl_flow_params* p, const gospl_flow_buffers* b,
                                       gospl_flow_result* res, gospl_stream_t s) {
  if (!m || !p || !b || !res) {
    set_error("gospl_flow_accumulation: null argument");
    return 1;
  }
  const int64_t N = m->N;
  const int g = nblocks(N);
  std::memset(res, 0, sizeof(*res));
  if (!m->side) {
    GOSPL_CHECK(hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking));
    for (int k = 0; k < 3; ++k) GOSPL_CHECK(hipEventCreateWithFlags(&m->side_ev[k], hipEventDisableTiming));
  }
  hipStream_t s2 = m->side;
  // fillElevation (pitfilling.py:849-903)
  if (gospl_fill_elevation(m, b->h, p->add_eps, p->floor_level, p->sealevel, b->lfill, s)) return 1;
  // The step's receivers (the main set, down = False: they read the filled
  // surface only through the sea test lfill <= sealevel, which the marine
  // restore of the pit information, fill <= sealevel -> h <= fill, leaves
  // unchanged), the runoff source and the main flow solve run on the side
  // stream, from a copy of the filled surface taken before the pit
  // information rewrites it, while the pit information runs here.  The flow
  // sweeps have their own workspace.  Same inputs, same kernels: the results
  // equal the one-stream order bit for bit.
  double* lsnap = b->fa;  // free until the cascade
  GOSPL_CHECK(hipEventRecord(m->side_ev[0], s));
  GOSPL_CHECK(hipStreamWaitEvent(s2, m->side_ev[0], 0));
  GOSPL_CHECK(hipMemcpyAsync(lsnap, b->lfill, sizeof(double) * N, hipMemcpyDeviceToDevice, s2));
  GOSPL_CHECK(hipEventRecord(m->side_ev[1], s2));
  GOSPL_CHECK(hipStreamWaitEvent(s, m->side_ev[1], 0));
  GOSPL_CHECK(hipMemcpyAsync(b->hl, b->h, sizeof(double) * N, hipMemcpyDeviceToDevice, s2));
  if (gospl_flow_direction(m, p->nrcv, p->flowexp, b->hl, p->sealevel, lsnap, 0, b->pitids, b->flatdirs, b->rcv0,
                           b->dist0, b->wgt0, b->lsink0, s2))
    return 1;
  if (gospl_rain_source(m, b->rain, lsnap, p->sealevel, b->b, s2)) return 1;
  if (flowacc_issue(m, p->nrcv, b->rcv0, b->wgt0, b->b, b->fal, s2)) return 1;
  int64_t npits = 0;
  const int rc_pit = gospl_pit_info(m, b->h, b->lfill, p->sealevel, b->pitids, b->flatdirs, b->ptdir, b->lspill,
                                    b->cap, b->spill, b->pitparams, b->flvl, b->fvol, &npits, s);
  int32_t lev = 0;
  int64_t bad = 0;
  // the side stream's status words (waits for the main solve)
  const int rc_main = flowacc_finish(m, p->nrcv, b->rcv0, b->wgt0, &lev, &bad, s2);
  GOSPL_CHECK(hipEventRecord(m->side_ev[2], s2));
  GOSPL_CHECK(hipStreamWaitEvent(s, m->side_ev[2], 0));
  if (rc_pit || rc_main) return 1;
  res->flow_solves += 1;
  res->undrained = bad;
  if (bad) {
    set_error("flow accumulation did not converge (%lld undrained vertices)", (long long)bad);
    return 2;
  }
  const int64_t n1 = npits + 1;
  // tables (14 n1), and a pass's staging (<= 4 n1: full, level, spill cells, volumes)
  if (ensure_flow_bufs(m, 14 * n1, s)) return 1;
  if (gospl_pit_tables(m, npits, b->cap, b->spill, b->pitparams, b->flvl, b->fvol, m->flow_htab, s)) return 1;
  m->flow_n1 = n1;
  res->npits = npits;
  const double* T = m->flow_htab;
  auto spill_of = [&](int64_t q) { return (int64_t)T[q]; };
  auto fvol = [&](int64_t q, int k) { return T[9 * n1 + k * n1 + q]; };
  auto flvl = [&](int64_t q, int k) { return T[4 * n1 + k * n1 + q]; };
  std::vector<double> pitVol(n1);
  for (int64_t q = 0; q < n1; ++q) pitVol[q] = T[n1 + q];
  res->last_set = 0;
  const int8_t* lsink_cur = b->lsink0;
  GOSPL_CHECK(hipMemcpyAsync(b->water, b->hl, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  bool has_pits = false;
  for (int64_t q = 0; q < n1; ++q) has_pits = has_pits || pitVol[q] > 0.0;
  int passes = 0;
  if (has_pits) {
    k_scale_copy<<<g, GOSPL_BLOCK, 0, s>>>(N, b->fal, p->dt, b->fa);
    GOSPL_CHECK(hipGetLastError());
    const double* FA = b->fa;
    bool excess = true;
    int step = 0;
    bool have_best = false;
    double best = 0.0;
    int stall = 0;
    std::vector<double> inV(n1), eV(n1), full(n1), level(n1), vv, vdt;
    std::vector<int64_t> u;
    std::vector<std::pair<int64_t, int64_t>> pts;  // (spill cell, pit) in pit order
    while (excess && step <= p->max_passes) {
      // _distributeDownstream (flowplex.py:576-686)
      if (gospl_pit_inflow(m, FA, lsink_cur, b->pitids, n1, b->inv, s)) return 1;
      GOSPL_CHECK(hipMemcpyAsync(m->flow_hstage, b->inv, sizeof(double) * n1, hipMemcpyDeviceToHost, s));
      GOSPL_CHECK(hipStreamSynchronize(s));
      std::memcpy(inV.data(), m->flow_hstage, sizeof(double) * n1);
      bool anypos = false, anyneg = false, go = false, anyfull = false, anylevel = false;
      for (int64_t q = 0; q < n1; ++q) {
        eV[q] = inV[q] - pitVol[q];
        full[q] = 0.0;
        level[q] = NAN;
        anypos = anypos || eV[q] > 0.0;
        anyneg = anyneg || eV[q] < 0.0;
        go = go || eV[q] > 1.0e-3;
      }
      u.clear();
      vv.clear();
      double nfa_sum = 0.0;
      if (anypos) {
        pts.clear();
        for (int64_t q = 0; q < n1; ++q)
          if (eV[q] > 0.0) {
            pitVol[q] = 0.0;
            full[q] = 1.0;
            anyfull = true;
            if (spill_of(q) >= 0) pts.emplace_back(spill_of(q), q);
          }
        if (!pts.empty()) {
          // sorted unique spill cells, each with the value of its last pit
          std::vector<std::pair<int64_t, int64_t>> srt(pts);
          std::stable_sort(srt.begin(), srt.end(),
                           [](const std::pair<int64_t, int64_t>& x, const std::pair<int64_t, int64_t>& y) {
                             return x.first < y.first;
                           });
          for (size_t k = 0; k < srt.size(); ++k) {
            if (k + 1 < srt.size() && srt[k + 1].first == srt[k].first) continue;  // a later pit wins
            u.push_back(srt[k].first);
            vv.push_back(eV[srt[k].second]);
          }
          vdt.resize(vv.size());
          for (size_t k = 0; k < vv.size(); ++k) vdt[k] = vv[k] / p->dt;
          nfa_sum = np_sum(vdt.data(), (int64_t)vdt.size());
        }
      }
      if (anyneg) {
        for (int64_t q = 0; q < n1; ++q)
          if (eV[q] < 0.0 && inV[q] > 0.0) {
            pitVol[q] += eV[q];
            int nid = 0;
            double dmin = std::fabs(fvol(q, 0) - pitVol[q]);
            for (int k = 1; k < 5; ++k) {
              const double d = std::fabs(fvol(q, k) - pitVol[q]);
              if (d < dmin) {
                dmin = d;
                nid = k;
              }
            }
            level[q] = flvl(q, nid);
            anylevel = true;
          }
      }
      bool ex = false;
      if (go || anyfull || anylevel) {
        const int64_t k = (int64_t)u.size();
        const int64_t n = 2 * n1 + 2 * k;  // k <= n1: within the 14 n1 ensured above
        double* st = m->flow_hstage;
        std::memcpy(st, full.data(), sizeof(double) * n1);
        std::memcpy(st + n1, level.data(), sizeof(double) * n1);
        for (int64_t j = 0; j < k; ++j) {
          st[2 * n1 + j] = (double)u[j];
          st[2 * n1 + k + j] = vv[j];
        }
        GOSPL_CHECK(hipMemcpyAsync(m->flow_dstage, st, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (gospl_cascade_seed(m, m->flow_dstage, n1, k, b->pitids, b->lfill, b->water, p->dt, b->rhs, s)) return 1;
      }
      if (go) {
        const double* src = step == 100 ? b->lfill : b->water;
        if (gospl_flow_direction(m, p->nrcv, p->flowexp, src, p->sealevel, b->lfill, 1, b->pitids, b->flatdirs,
                                 b->rcv1, b->dist1, b->wgt1, b->lsink1, s))
          return 1;
        res->last_set = 1;
        lsink_cur = b->lsink1;
        if (nfa_sum > p->maxarea) {
          ex = true;
          int64_t bad2 = 0;
          if (gospl_flowacc(m, p->nrcv, b->rcv1, b->wgt1, b->rhs, b->q, &lev, &bad2, s)) return 1;
          res->flow_solves += 1;
          res->undrained = bad2;
          if (bad2) GOSPL_CHECK(hipMemsetAsync(b->q, 0, sizeof(double) * N, s));  // goSPL zeroes it
          if (gospl_cascade_add(m, b->q, b->hl, b->water, p->dt, b->nfa, b->fal, s)) return 1;
          FA = b->nfa;
        }
      }
      excess = ex;
      step += 1;
      if (excess) {
        if (!have_best || nfa_sum < best * (1.0 - p->rel_improve)) {
          best = nfa_sum;
          have_best = true;
          stall = 0;
        } else {
          stall += 1;
          if (stall >= p->patience) break;
        }
      }
    }
    passes = step;
  }
  res->cascade_passes = passes;
  if (gospl_flow_finalize(m, b->hl, b->water, b->pitids, has_pits ? 1 : 0, b->fal, b->fillfal, s)) return 1;
  k_sub_inplace<<<g, GOSPL_BLOCK, 0, s>>>(N, b->water, b->hl);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}
