// Depression filling and depression information.
//
// Fill.  goSPL floods from seed cells with a priority queue (Barnes'
// priority-flood, functions.F90:2082-2178); the filled value of a reached
// cell above the level is its minimax ("spill") elevation over paths from a
// seed.  Here it is computed without a queue:
//   1. every included cell points at its lowest included neighbour if that
//      is lower in (elevation, index) order, else at itself (a pit bottom);
//      seeds point at themselves;
//   2. pointer jumping gives each cell the root of its descent tree (its
//      basin); along a descent path elevations never increase, so inside a
//      basin the minimax value of a cell is its own elevation joined with the
//      basin's spill level;
//   3. basin spill levels S (seed basins -inf) relax over the edges that
//      cross basins: S[ru] = min(S[ru], max(z_u, z_v, S[rv])) until nothing
//      changes (atomicMin on order-preserving 64-bit keys: exact, no
//      rounding);
//   4. fill = max(z, S[basin]) for reached cells above the level, z
//      elsewhere.
// Only comparisons and min/max touch the elevations, so the result is
// bit-identical to the queue-based flood.
//
// Depression information (pitfilling.py:704-847): equal-fill connected
// components with a local-minimum cell (union-find with atomic CAS hooking,
// ECL-CC style), numbered by the smallest input id among their bottom cells
// (label_pits' visit order) via a scan over input ids; spill points by
// atomicMin of input ids; flat-routing distances by relaxation of the
// breadth-first distance; per-pit sums in a fixed order (k_pits.hip).
#include <climits>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "common.h"

namespace gospl {

constexpr unsigned long long KEY_POS_INF = 0xFFF0000000000000ull;  // dkey(+inf)
constexpr unsigned long long KEY_NEG_INF = 0x000FFFFFFFFFFFFFull;  // dkey(-inf)

// The fill level: the given one, or, when `lv` is set, fillElevation's
// choice from the minimum elevation the device found (pitfilling.py:860-866):
// minh = min(h) (+ 1e-3 when lv->add_eps), level = max(minh, lv->floor) with
// Python's max (the first argument unless the second is greater).
struct FillLevel {
  const unsigned long long* minkey;  // dkey of min(h), one per segment
  int add_eps;
  double floor;
  const int32_t* seg;  // segment of each vertex (NULL: one segment); < 0: padding at `floor`
};
__device__ __forceinline__ double fill_level(double level, const FillLevel& lv, int64_t i) {
  if (!lv.minkey) return level;
  int sg = 0;
  if (lv.seg) {
    sg = lv.seg[i];
    if (sg < 0) return lv.floor;
  }
  double mh = dkey_inv(lv.minkey[sg]);
  if (lv.add_eps) mh += 1.0e-3;
  return lv.floor > mh ? lv.floor : mh;
}

// dkey of min(h): a fixed grid strides over the mesh, one atomic per
// workgroup (exact and independent of order); *key starts at ~0.
#define FILL_MIN_BLOCKS 512
__global__ __launch_bounds__(GOSPL_BLOCK) void k_min_key(int64_t N, const double* __restrict__ h,
                                                        unsigned long long* __restrict__ key) {
  __shared__ unsigned long long sh[GOSPL_BLOCK / 64];
  unsigned long long k = ~0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = dkey(h[i]);
    k = c < k ? c : k;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long c = __shfl_xor(k, o);
    k = c < k ? c : k;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) k = sh[w] < k ? sh[w] : k;
    atomicMin(key, k);
  }
}

// dkey of min(h) per segment: one atomic per workgroup whose vertices share
// a segment (the ensemble layout aligns segments to workgroups), one per
// vertex otherwise; padding vertices (segment < 0) take no part.
__global__ __launch_bounds__(GOSPL_BLOCK) void k_min_key_seg(int64_t N, const double* __restrict__ h,
                                                            const int32_t* __restrict__ seg,
                                                            unsigned long long* __restrict__ keys) {
  __shared__ unsigned long long sh[GOSPL_BLOCK / 64];
  __shared__ int s_lo, s_hi;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = i < N ? seg[i] : -1;
  unsigned long long k = (sg >= 0) ? dkey(h[i]) : ~0ull;
  if (threadIdx.x == 0) {
    s_lo = 0x7fffffff;
    s_hi = -1;
  }
  __syncthreads();
  if (sg >= 0) {
    atomicMin(&s_lo, sg);
    atomicMax(&s_hi, sg);
  }
  __syncthreads();
  const int lo = s_lo, hi = s_hi;
  if (hi < 0) return;
  if (lo != hi) {  // mixed workgroup
    if (sg >= 0) atomicMin(keys + sg, k);
    return;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long c = __shfl_xor(k, o);
    k = c < k ? c : k;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < GOSPL_BLOCK / 64; ++w) k = sh[w] < k ? sh[w] : k;
    atomicMin(keys + lo, k);
  }
}

// state: 1 seed, 0 included, -1 excluded (edge_tile, functions.F90:2036-2080).
// raw (epsfill's base, functions.F90:2633-2697): elevations unshifted, seeds
// the cells below the level next to one at or above it, no border seeds.
__global__ void k_fill_init(int64_t N, int W, const int32_t* __restrict__ nbr,
                            const int32_t* __restrict__ deg, const int8_t* __restrict__ outlet,
                            int flat, const double* __restrict__ h, double level_in, FillLevel lv,
                            double* __restrict__ dem, int32_t* __restrict__ state, int raw) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double level = fill_level(level_in, lv, i);
  const double z = raw ? h[i] : h[i] - level;
  dem[i] = z;
  const bool border = !raw && flat && outlet[i];
  int st = 0;
  if (raw ? z < level : z < 0.0) {
    bool up = false;
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      int n = nbr[(int64_t)p * N + i];
      if (n >= 0 && (raw ? h[n] >= level : h[n] - level >= 0.0)) { up = true; break; }
    }
    st = up ? 1 : -1;
  } else if (border) {
    st = 1;
  }
  state[i] = st;
}

__global__ void k_fill_descent(int64_t N, int W, const int32_t* __restrict__ nbr,
                               const int32_t* __restrict__ deg, const double* __restrict__ dem,
                               const int32_t* __restrict__ state, int32_t* __restrict__ root,
                               unsigned long long* __restrict__ S) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int st = state[i];
  int best = (int)i;
  if (st == 0) {
    double bz = dem[i];
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      int n = nbr[(int64_t)p * N + i];
      if (n < 0 || state[n] < 0) continue;
      double zn = dem[n];
      if (zn < bz || (zn == bz && n < best)) { bz = zn; best = n; }
    }
  }
  root[i] = best;
  S[i] = (st == 1) ? KEY_NEG_INF : KEY_POS_INF;
}

// In-place pointer jumping; a few hops per launch.
__global__ void k_jump(int64_t N, int32_t* __restrict__ root, int32_t* __restrict__ flag, int pass) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int r = root[i];
  int r0 = r;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int rr = root[r];
    if (rr == r) break;
    r = rr;
  }
  if (r != r0) {
    root[i] = r;
    if (root[r] != r) *flag = pass;
  }
}

// Cells that touch another basin (the only ones the relaxation reads
// from), flagged one cell per thread; the list is then selected by flag.
__global__ void k_basin_boundary(int64_t N, int W, const int32_t* __restrict__ nbr,
                                 const int32_t* __restrict__ deg, const int32_t* __restrict__ state,
                                 const int32_t* __restrict__ root, int8_t* __restrict__ flag) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= N) return;
  int8_t f = 0;
  if (state[u] >= 0) {
    const int ru = root[u];
    const int dg = deg[u];
    for (int p = 0; p < W && p < dg; ++p) {
      const int v = nbr[(int64_t)p * N + u];
      if (v >= 0 && state[v] >= 0 && root[v] != ru) f = 1;
    }
  }
  flag[u] = f;
}

// Step 3 over the list of basin-boundary cells: S[ru] = min(S[ru],
// max(key(max(z_u, z_v)), S[rv])) over included neighbours v in another
// basin; a pass that lowers some S writes its number into *flag.
#define FILL_REPEAT 4
__global__ __launch_bounds__(GOSPL_BLOCK) void k_fill_relax_list(
    int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
    const double* __restrict__ dem, const int32_t* __restrict__ state, const int32_t* __restrict__ root,
    unsigned long long* __restrict__ S, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
    int32_t* __restrict__ flag, int pass) {
  const int n = *count;
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;  // whole workgroup past the list
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool own = t < n;
  const int u = own ? list[t] : 0;
  const int ru = own ? root[u] : 0;
  // the neighbours in other basins, found once: their basin roots and the
  // key of the edge's crossing height max(z_u, z_v) (position p: root or -1)
  int rn[GOSPL_MAXW];
  unsigned long long kn[GOSPL_MAXW];
  {
    const double zu = own ? dem[u] : 0.0;
    const int dg = own ? deg[u] : 0;
#pragma unroll
    for (int p = 0; p < GOSPL_MAXW; ++p) {
      int rv = -1;
      unsigned long long kw = 0;
      if (p < W && p < dg) {
        const int v = nbr[(int64_t)p * N + u];
        if (v >= 0 && state[v] >= 0) {
          rv = root[v];
          if (rv == ru) rv = -1;
          const double w = dem[v] > zu ? dem[v] : zu;
          kw = dkey(w);
        }
      }
      rn[p] = rv;
      kn[p] = kw;
    }
  }
  // a few rounds per launch, the workgroup's cells (consecutive boundary
  // cells, a compact patch) seeing each other's lowered levels between
  // rounds; a round is one load of this basin's level and of each listed
  // neighbour basin's
  bool changed = false;
  for (int round = 0; round < FILL_REPEAT; ++round) {
    if (own) {
      const unsigned long long cur =
          __hip_atomic_load(&S[ru], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur != KEY_NEG_INF) {
        unsigned long long best = cur;
#pragma unroll
        for (int p = 0; p < GOSPL_MAXW; ++p) {
          if (rn[p] < 0) continue;
          const unsigned long long sv =
              __hip_atomic_load(&S[rn[p]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (sv == KEY_POS_INF) continue;
          const unsigned long long c = kn[p] > sv ? kn[p] : sv;
          if (c < best) best = c;
        }
        if (best < cur) {
          const unsigned long long old = atomicMin(&S[ru], best);
          if (best < old) changed = true;
        }
      }
    }
    __syncthreads();
  }
  if (changed) *flag = pass;
}

__global__ void k_fill_final(int64_t N, const double* __restrict__ dem,
                             const int32_t* __restrict__ state, const int32_t* __restrict__ root,
                             const unsigned long long* __restrict__ S, double level_in, FillLevel lv,
                             double sl, double* __restrict__ lfill, int8_t* __restrict__ reach) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double level = fill_level(level_in, lv, i);
  const double z = dem[i];
  double f = z;
  bool reached = false;
  if (state[i] >= 0) {
    unsigned long long s = S[root[i]];
    reached = s != KEY_POS_INF;
    if (reached && (reach ? z > level : z > 0.0)) {
      double sv = dkey_inv(s);
      f = sv > z ? sv : z;
    }
  }
  if (reach) {
    // raw mode (epsfill's base): unshifted, no marine reset
    reach[i] = reached ? 1 : 0;
    lfill[i] = f;
    return;
  }
  if (f < sl - level) f = z;
  lfill[i] = f + level;
}

// epsfill (functions.F90:2633-2697): the priority flood raises a reached
// cell c at or above the level to the next double above the fill of the
// first neighbour popped before it, the one with the least fill:
//   F(c) = max(z_c, nextafter(min_n F(n), +inf)),
// seeds and excluded cells keep z.  F is the unique fixed point of that map
// (the queue pops fills in non-decreasing order); the minimax fill is below
// it everywhere, so raising from it until no cell changes reaches it.
__global__ void k_eps_relax(int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
                            const double* __restrict__ h, const int32_t* __restrict__ state,
                            const int8_t* __restrict__ reach, double* __restrict__ F, int32_t* __restrict__ flag,
                            int pass) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (state[i] != 0 || !reach[i]) return;
  double mn = INFINITY;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    const int n = nbr[(int64_t)p * N + i];
    if (n >= 0) {
      const double v = F[n];
      mn = v < mn ? v : mn;
    }
  }
  if (mn == INFINITY) return;
  const double c = nextafter(mn, INFINITY);
  const double z = h[i];
  const double nf = z > c ? z : c;
  if (nf > F[i]) {
    F[i] = nf;
    *flag = pass;
  }
}

// ---------------------------------------------------------------------------
// depression information
// ---------------------------------------------------------------------------

__global__ void k_pit_init(int64_t N, int W, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ deg, const double* __restrict__ fill,
                           double lvl, int32_t* __restrict__ parent, int8_t* __restrict__ noflow,
                           int32_t* __restrict__ cmin, int32_t* __restrict__ cnon) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double f = fill[i];
  cmin[i] = INT_MAX;
  cnon[i] = 0;
  if (!(f > lvl)) {
    parent[i] = -1;
    noflow[i] = 0;
    return;
  }
  parent[i] = (int)i;
  bool nf = true;
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    int n = nbr[(int64_t)p * N + i];
    if (n >= 0 && f > fill[n]) { nf = false; break; }
  }
  noflow[i] = nf ? 1 : 0;
}

__device__ __forceinline__ int uf_find(int32_t* parent, int x) {
  int p = parent[x];
  while (p != x) {
    int gp = parent[p];
    if (gp != p) parent[x] = gp;  // path halving (benign race: moves toward the root)
    x = p;
    p = gp;
  }
  return x;
}

__global__ void k_cc_hook(int64_t N, int W, const int32_t* __restrict__ nbr,
                          const int32_t* __restrict__ deg, const double* __restrict__ fill,
                          int32_t* parent) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (parent[i] < 0) return;
  const double f = fill[i];
  const int dg = deg[i];
  for (int p = 0; p < W && p < dg; ++p) {
    int j = nbr[(int64_t)p * N + i];
    if (j <= (int)i || fill[j] != f) continue;
    int a = (int)i, b = j;
    for (;;) {
      a = uf_find(parent, a);
      b = uf_find(parent, b);
      if (a == b) break;
      int hi = a > b ? a : b, lo = a > b ? b : a;
      int old = atomicCAS(&parent[hi], hi, lo);
      if (old == hi) break;
      a = lo;
      b = old;
    }
  }
}

// Read-only find: after hooking no pointer may move any more, and a halving
// store here could overwrite a neighbour's final root with an intermediate
// ancestor.
__device__ __forceinline__ int uf_root(const int32_t* parent, int x) {
  int p = parent[x];
  while (p != x) {
    x = p;
    p = parent[x];
  }
  return x;
}

// Same-address atomics serialise, and a lake is one component of many
// cells: the lanes of a wavefront that share a root reduce their minimum
// first, one atomic per root and wavefront.
__global__ void k_cc_compress(int64_t N, int32_t* parent, const int8_t* __restrict__ noflow,
                              const int32_t* __restrict__ gid, const int8_t* __restrict__ outlet,
                              int flat, int32_t* __restrict__ cmin, int32_t* __restrict__ cnon) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i - (i & 63) >= N) return;  // whole wavefront past the mesh
  const bool in = i < N && parent[i] >= 0;
  int r = -1;
  if (in) {
    r = uf_root(parent, (int)i);
    parent[i] = r;
    if (!(flat && outlet[i])) cnon[r] = 1;
  }
  const bool want = in && noflow[i];
  const int g = want ? gid[i] : INT_MAX;
  unsigned long long todo = __ballot(want);
  const int lane = __lane_id();
  while (todo != 0ull) {
    const int leader = __ffsll((long long)todo) - 1;
    const int rl = __shfl(r, leader);
    const bool mine = want && r == rl;
    int m = mine ? g : INT_MAX;
    for (int o = 32; o > 0; o >>= 1) {
      const int t = __shfl_xor(m, o);
      m = t < m ? t : m;
    }
    if (lane == leader) atomicMin(&cmin[rl], m);
    todo &= ~__ballot(mine);
  }
}

__global__ void k_pit_mark(int64_t N, const int32_t* __restrict__ parent,
                           const int32_t* __restrict__ cmin, const int32_t* __restrict__ cnon,
                           int32_t* __restrict__ mark) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (parent[i] == (int)i && cmin[i] != INT_MAX && cnon[i]) mark[cmin[i]] = 1;
}

__global__ void k_pit_ids(int64_t N, const int32_t* __restrict__ parent,
                          const int32_t* __restrict__ cmin, const int32_t* __restrict__ cnon,
                          const int32_t* __restrict__ rank, const int8_t* __restrict__ outlet,
                          int flat, int32_t* __restrict__ pitids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int id = -1;
  const int r = parent[i];
  if (r >= 0 && cmin[r] != INT_MAX && cnon[r] && !(flat && outlet[i])) id = rank[cmin[r]] + 1;
  pitids[i] = id;
}

// spill candidates (pitfilling.py:741-801)
__global__ void k_spill_best(int64_t N, int W, const int32_t* __restrict__ nbr,
                             const int32_t* __restrict__ deg, const double* __restrict__ fill,
                             const int32_t* __restrict__ pitids, const int8_t* __restrict__ outlet,
                             int flat, const int32_t* __restrict__ gid, int32_t* __restrict__ best,
                             int8_t* __restrict__ qual) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int8_t q = 0;
  const int pid = pitids[i];
  if (pid > -1) {
    const double f = fill[i];
    bool lower = false, eq = false;
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      int n = nbr[(int64_t)p * N + i];
      if (n < 0) continue;
      double fn = fill[n];
      if (fn < f) lower = true;
      else if (fn == f && ((flat && outlet[n]) || pitids[n] == -1)) eq = true;
    }
    if (lower) q = 2;
    else if (eq) q = 1;
    if (q) atomicMin(&best[pid], gid[i]);
  }
  qual[i] = q;
}

__global__ void k_spill_set(int64_t N, int W, const int32_t* __restrict__ nbr,
                            const int32_t* __restrict__ deg, const double* __restrict__ fill,
                            const int32_t* __restrict__ pitids, const int8_t* __restrict__ outlet,
                            int flat, const int32_t* __restrict__ gid,
                            const int32_t* __restrict__ best, const int8_t* __restrict__ qual,
                            int32_t* __restrict__ spill, int8_t* __restrict__ lspill) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int8_t q = qual[i];
  if (!q) return;
  const int pid = pitids[i];
  if (best[pid] != gid[i]) return;
  if (q == 2) {
    spill[pid] = (int)i;
    lspill[i] = 1;
  } else {
    const double f = fill[i];
    int s = -1, sg = INT_MAX;
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      int n = nbr[(int64_t)p * N + i];
      if (n < 0) continue;
      if (fill[n] == f && ((flat && outlet[n]) || pitids[n] == -1) && gid[n] < sg) {
        sg = gid[n];
        s = n;
      }
    }
    spill[pid] = s;
    lspill[s] = 1;
  }
}

__global__ void k_ptdir_init(int64_t N, const int8_t* __restrict__ ls, int32_t* __restrict__ pd) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) pd[i] = ls[i] ? 1 : -1;
}

// Breadth-first distance from the spill cells over pit cells along
// not-descending edges (fill_dir + nghb_dir, functions.F90:1845-1969).
#define PTDIR_REPEAT 8
#define PTDIR_BLOCK 256
__global__ __launch_bounds__(PTDIR_BLOCK) void k_ptdir_relax(
    int64_t N, int W, const int32_t* __restrict__ nbr, const int32_t* __restrict__ deg,
    const double* __restrict__ fill, const int32_t* __restrict__ ncells, const int32_t* __restrict__ cells,
    int32_t* ptdir, int32_t* __restrict__ flag, int pass) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = *ncells;
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;  // whole workgroup past the list
  const bool own = t < n;
  const int c = own ? cells[t] : 0;
  // only this thread writes ptdir[c]: its current value stays in a register
  int cur = own ? ptdir[c] : 1;
  const bool active = own && cur != 1;
  // the neighbours a distance can come from (fill not above this cell's),
  // found once: position p holds the neighbour or -1 (fixed positions keep
  // the list in registers)
  int nb[GOSPL_MAXW];
  {
    const double fc = active ? fill[c] : 0.0;
    const int dg = active ? deg[c] : 0;
#pragma unroll
    for (int p = 0; p < GOSPL_MAXW; ++p) {
      int i = -1;
      if (p < W && p < dg) {
        i = nbr[(int64_t)p * N + c];
        if (i >= 0 && !(fc >= fill[i])) i = -1;
      }
      nb[p] = i;
    }
  }
  // a few rounds per launch, the workgroup's cells (consecutive pit cells, a
  // compact patch) seeing each other's updates between rounds; a round is
  // one load of each listed neighbour's distance
  bool changed = false;
  for (int round = 0; round < PTDIR_REPEAT; ++round) {
    if (active) {
      int best = cur;
#pragma unroll
      for (int p = 0; p < GOSPL_MAXW; ++p) {
        if (nb[p] >= 0) {
          const int di = ptdir[nb[p]];
          if (di > -1) {
            const int cand = di + 1;
            if (best < 0 || cand < best) best = cand;
          }
        }
      }
      if (best != cur) {
        ptdir[c] = best;
        cur = best;
        changed = true;
      }
    }
    __syncthreads();
  }
  if (changed) *flag = pass;
}

// Flat receivers (pitfilling.py:479-517); spill cells and non-pit cells -1.
__global__ void k_flatdirs(int64_t N, int W, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ deg, const double* __restrict__ fill,
                           const int32_t* __restrict__ pitids, const int32_t* __restrict__ ptdir,
                           const int32_t* __restrict__ gid, const int8_t* __restrict__ lspill,
                           int32_t* __restrict__ flatdirs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int out = -1;
  if (pitids[i] > -1) {
    const int di = ptdir[i];
    const double fi = fill[i];
    long long bestk = LLONG_MAX;
    int bestn = (int)i;
    const int dg = deg[i];
    for (int p = 0; p < W && p < dg; ++p) {
      int c = nbr[(int64_t)p * N + i];
      if (c < 0) continue;
      int dc = ptdir[c];
      if (dc > -1 && dc < di && fill[c] <= fi) {
        long long k = ((long long)dc << 32) | (unsigned int)gid[c];
        if (k < bestk) { bestk = k; bestn = c; }
      }
    }
    out = bestn;
  }
  if (lspill[i]) out = -1;
  flatdirs[i] = out;
}

// Marine restore (pitfilling.py:822-832): cells filled to sea level or below
// take their unfilled elevation and have no flat receiver.
__global__ void k_pit_marine(int64_t N, const double* __restrict__ hl, double* __restrict__ fill,
                             int32_t* __restrict__ flatdirs, double sl) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (fill[i] <= sl) {
    fill[i] = hl[i];
    flatdirs[i] = -1;
  }
}

// Host helper: run relaxation passes until one changes nothing.  A pass
// that changes something writes its (1-based) number into the device flag,
// so after a batch the flag holds the last changing pass; if a later pass of
// the batch exists, the fixed point has been reached and confirmed.  The
// first batch is as long as the previous call needed (`hint`, kept in the
// mesh handle), so a call usually checks once.
template <typename F>
static int until_stable(gospl_mesh* m, hipStream_t s, int max_iter, int check_every, int* hint,
                        F launch, int* iters) {
  int32_t* flag = m->dflag + 8;
  GOSPL_CHECK(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
  int done = 0;
  int batch = *hint > check_every ? *hint : check_every;
  for (;;) {
    for (int k = 0; k < batch; ++k) launch(flag, done + k + 1);
    done += batch;
    GOSPL_CHECK(hipGetLastError());
    GOSPL_CHECK(hipMemcpyAsync(m->hflag + 8, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    const int last = m->hflag[8];
    if (last < done) {
      *hint = last + 2;  // passes to reach and confirm the fixed point, plus one
      break;
    }
    if (done >= max_iter) {
      set_error("iteration did not converge within %d passes", max_iter);
      return 1;
    }
    batch = check_every;
  }
  if (iters) *iters = done;
  return 0;
}

int ensure_seg(gospl_mesh* m, int nseg) {
  if (nseg <= m->seg_cap) return 0;
  if (m->seg_keys) GOSPL_CHECK(hipFree(m->seg_keys));
  if (m->seg_ctl) GOSPL_CHECK(hipFree(m->seg_ctl));
  if (m->seg_aux) GOSPL_CHECK(hipFree(m->seg_aux));
  m->seg_keys = nullptr;
  m->seg_ctl = nullptr;
  m->seg_aux = nullptr;
  m->seg_cap = 0;
  const int cap = nseg < 256 ? 256 : nseg;
  GOSPL_CHECK(hipMalloc(&m->seg_keys, sizeof(unsigned long long) * cap));
  GOSPL_CHECK(hipMalloc(&m->seg_ctl, sizeof(double) * 4 * cap));
  GOSPL_CHECK(hipMalloc(&m->seg_aux, sizeof(int32_t) * 4 * (cap + 1)));
  m->seg_cap = cap;
  return 0;
}

}  // namespace gospl

using namespace gospl;

extern "C" {

static int fill_body(gospl_mesh_t* m, const double* h, double level, FillLevel lv, double sealevel,
                     double* lfill, hipStream_t s, int8_t* reach = nullptr) {
  const int64_t N = m->N;
  double* dem = m->wd[0];
  int32_t* state = m->wi[3];
  int32_t* root = m->wi[4];
  unsigned long long* S = m->wu[0];
  const int g = nblocks(N);
  k_fill_init<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, m->outlet, m->flat, h, level, lv, dem,
                                         state, reach ? 1 : 0);
  k_fill_descent<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, dem, state, root, S);
  GOSPL_CHECK(hipGetLastError());
  int it = 0;
  if (until_stable(m, s, (int)N + 16, 3, &m->iter_hint[0], [&](int32_t* f, int pass) {
        k_jump<<<g, GOSPL_BLOCK, 0, s>>>(N, root, f, pass);
      }, &it))
    return 1;
  // relax the basin levels over the cells that touch another basin
  int32_t* blist = m->wi[5];
  int32_t* bcount = m->dflag + 44;
  {
    int8_t* bflag = reinterpret_cast<int8_t*>(m->wi[6]);
    k_basin_boundary<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, state, root, bflag);
    GOSPL_CHECK(hipGetLastError());
    size_t bytes = m->cub_tmp_bytes;
    rocprim::counting_iterator<int32_t> cit(0);
    GOSPL_CHECK(rocprim::select(m->cub_tmp, bytes, cit, bflag, blist, bcount, (size_t)N, s));
  }
  // one list entry per thread; the length stays on the device (workgroups past it return)
  const int gb = g;
  if (until_stable(m, s, (int)N + 16, 2, &m->iter_hint[1], [&](int32_t* f, int pass) {
        k_fill_relax_list<<<gb, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, dem, state, root, S, blist,
                                                      bcount, f, pass);
      }, &it))
    return 1;
  k_fill_final<<<g, GOSPL_BLOCK, 0, s>>>(N, dem, state, root, S, level, lv, sealevel, lfill, reach);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_eps_fill(gospl_mesh_t* m, const double* h, double level, double* fillz, int32_t* passes,
                   gospl_stream_t s) {
  if (!m) {
    set_error("gospl_eps_fill: null handle");
    return 1;
  }
  const int64_t N = m->N;
  int8_t* reach = reinterpret_cast<int8_t*>(m->wi[7]);
  if (fill_body(m, h, level, FillLevel{nullptr, 0, 0.0, nullptr}, -INFINITY, fillz, s, reach)) return 1;
  const int32_t* state = m->wi[3];
  const int g = nblocks(N);
  int it = 0;
  if (until_stable(m, s, (int)N + 16, 4, &m->iter_hint[3], [&](int32_t* f, int pass) {
        k_eps_relax<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, h, state, reach, fillz, f, pass);
      }, &it))
    return 1;
  if (passes) *passes = it;
  return 0;
}

int gospl_fill_depressions(gospl_mesh_t* m, const double* h, double level, double sealevel,
                           double* lfill, gospl_stream_t s) {
  return fill_body(m, h, level, FillLevel{nullptr, 0, 0.0, nullptr}, sealevel, lfill, s);
}

int gospl_fill_elevation(gospl_mesh_t* m, const double* h, int32_t add_eps, double floor_level,
                         double sealevel, double* lfill, gospl_stream_t s) {
  unsigned long long* key = m->wu[1] + 8;  // free while the fill runs
  GOSPL_CHECK(hipMemsetAsync(key, 0xff, sizeof(unsigned long long), s));
  k_min_key<<<FILL_MIN_BLOCKS, GOSPL_BLOCK, 0, s>>>(m->N, h, key);
  GOSPL_CHECK(hipGetLastError());
  return fill_body(m, h, 0.0, FillLevel{key, add_eps ? 1 : 0, floor_level, nullptr}, sealevel, lfill, s);
}

int gospl_fill_elevation_seg(gospl_mesh_t* m, const double* h, int32_t add_eps, double floor_level,
                             double sealevel, const int32_t* seg, int32_t nseg, double* lfill, gospl_stream_t s) {
  if (nseg < 1 || !seg) {
    set_error("gospl_fill_elevation_seg: need a segment array and nseg >= 1");
    return 1;
  }
  if (ensure_seg(m, nseg)) return 1;
  GOSPL_CHECK(hipMemsetAsync(m->seg_keys, 0xff, sizeof(unsigned long long) * nseg, s));
  k_min_key_seg<<<nblocks(m->N), GOSPL_BLOCK, 0, s>>>(m->N, h, seg, m->seg_keys);
  GOSPL_CHECK(hipGetLastError());
  return fill_body(m, h, 0.0, FillLevel{m->seg_keys, add_eps ? 1 : 0, floor_level, seg}, sealevel, lfill, s);
}

}  // extern "C"

namespace gospl {

// Depression information with its three levels: pits are the equal-fill
// components above `label_lvl` (label_pits' lvl), cells filled to `sealevel`
// or below are restored, and the per-pit sums clip elevations below
// `clip_lvl` to it (goSPL: sea level, sea level, sea level for water; the fill
// level, sea level and no clip for sediment, pitfilling.py:718-721, 828-832).
static int pit_info_body(gospl_mesh_t* m, const double* hl, double* lfill, double label_lvl, double sealevel,
                         double clip_lvl, int32_t* pitids, int32_t* flatdirs, int32_t* ptdir, int8_t* lspill,
                         int64_t cap, int32_t* spill, double* pitparams, double* filled_lvl,
                         double* filled_vol, int64_t* npits_out, gospl_stream_t s, int* ptdir_hint) {
  const int64_t N = m->N;
  const int g = nblocks(N);
  int32_t* parent = m->wi[3];
  int32_t* cmin = m->wi[4];
  int32_t* cnon = m->wi[5];
  int32_t* mark = m->wi[6];
  int32_t* rank = m->wi[7];
  int8_t* noflow = reinterpret_cast<int8_t*>(m->wi[0]);
  int8_t* qual = reinterpret_cast<int8_t*>(m->wi[1]);
  int32_t* best = m->wi[2];

  k_pit_init<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, lfill, label_lvl, parent, noflow, cmin,
                                        cnon);
  k_cc_hook<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, lfill, parent);
  k_cc_compress<<<g, GOSPL_BLOCK, 0, s>>>(N, parent, noflow, m->gid, m->outlet, m->flat, cmin, cnon);
  GOSPL_CHECK(hipMemsetAsync(mark, 0, sizeof(int32_t) * N, s));
  k_pit_mark<<<g, GOSPL_BLOCK, 0, s>>>(N, parent, cmin, cnon, mark);
  GOSPL_CHECK(hipGetLastError());
  size_t bytes = m->cub_tmp_bytes;
  GOSPL_CHECK(rocprim::exclusive_scan(m->cub_tmp, bytes, mark, rank, 0, (size_t)N, rocprim::plus<int32_t>(), s));
  // npits = rank[N-1] + mark[N-1]
  GOSPL_CHECK(hipMemcpyAsync(m->hflag + 16, rank + (N - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipMemcpyAsync(m->hflag + 17, mark + (N - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  int32_t* ncells = m->dflag + 40;
  int32_t* cells = mark;  // the marks are consumed by the scan above
  k_pit_ids<<<g, GOSPL_BLOCK, 0, s>>>(N, parent, cmin, cnon, rank, m->outlet, m->flat, pitids);
  GOSPL_CHECK(hipGetLastError());
  // compact list of pit cells in ascending order
  if (pit_cells(m, pitids, cells, ncells, s)) return 1;
  GOSPL_CHECK(hipMemcpyAsync(m->hflag + 18, ncells, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  const int64_t npits = (int64_t)m->hflag[16] + m->hflag[17];
  const int npc = m->hflag[18];
  *npits_out = npits;
  const int64_t np1 = npits + 1;
  if (np1 > cap) {
    set_error("gospl_pit_info: %lld pits exceed the per-pit capacity %lld", (long long)npits,
              (long long)cap);
    return 1;
  }
  // one launch clears the grouping's segment table and the spill search's
  // buffers
  if (ensure_groups(m, npits + 1)) return 1;
  FillList fl{{m->pg_seg, best, spill, lspill},
              {(int64_t)sizeof(int32_t) * 2 * (m->pg_cap + 64), (int64_t)sizeof(int32_t) * np1,
               (int64_t)sizeof(int32_t) * np1, N},
              {0, 0x7f, 0xff, 0},
              4};
  if (fill_list(fl, s)) return 1;
  // cells grouped by pit for the per-pit reductions
  if (pit_group(m, pitids, npits, cells, ncells, npc, s, true)) return 1;
  // spill points
  k_spill_best<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, lfill, pitids, m->outlet, m->flat,
                                          m->gid, best, qual);
  k_spill_set<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, lfill, pitids, m->outlet, m->flat,
                                         m->gid, best, qual, spill, lspill);
  GOSPL_CHECK(hipGetLastError());
  // flat routing distances
  k_ptdir_init<<<g, GOSPL_BLOCK, 0, s>>>(N, lspill, ptdir);
  GOSPL_CHECK(hipGetLastError());
  if (npits > 0 && npc > 0) {
    int it = 0;
    if (until_stable(m, s, npc + 32, 16, ptdir_hint, [&](int32_t* f, int pass) {
          k_ptdir_relax<<<(int)((npc + PTDIR_BLOCK - 1) / PTDIR_BLOCK), PTDIR_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, lfill, ncells,
                                                             cells, ptdir, f, pass);
        }, &it))
      return 1;
  }
  k_flatdirs<<<g, GOSPL_BLOCK, 0, s>>>(N, m->W, m->nbr, m->deg, lfill, pitids, ptdir, m->gid, lspill,
                                        flatdirs);
  // marine restore, then the pit parameters and depth-volume curves
  k_pit_marine<<<g, GOSPL_BLOCK, 0, s>>>(N, hl, lfill, flatdirs, sealevel);
  GOSPL_CHECK(hipGetLastError());
  return pit_group_params(m, hl, lfill, clip_lvl, cap, pitparams, filled_lvl, filled_vol, s);
}

}  // namespace gospl

extern "C" {

int gospl_pit_info(gospl_mesh_t* m, const double* hl, double* lfill, double sealevel,
                   int32_t* pitids, int32_t* flatdirs, int32_t* ptdir, int8_t* lspill, int64_t cap,
                   int32_t* spill, double* pitparams, double* filled_lvl, double* filled_vol,
                   int64_t* npits_out, gospl_stream_t s) {
  return pit_info_body(m, hl, lfill, sealevel, sealevel, sealevel, pitids, flatdirs, ptdir, lspill, cap, spill,
                       pitparams, filled_lvl, filled_vol, npits_out, s, &m->iter_hint[2]);
}

int gospl_pit_info_sed(gospl_mesh_t* m, const double* hl, double* lfill, double level, double sealevel,
                       int32_t* pitids, int32_t* flatdirs, int32_t* ptdir, int8_t* lspill, int64_t cap,
                       int32_t* spill, double* pitparams, double* filled_lvl, double* filled_vol,
                       int64_t* npits_out, gospl_stream_t s) {
  return pit_info_body(m, hl, lfill, level, sealevel, -INFINITY, pitids, flatdirs, ptdir, lspill, cap, spill,
                       pitparams, filled_lvl, filled_vol, npits_out, s, &m->iter_hint[4]);
}

}  // extern "C"
