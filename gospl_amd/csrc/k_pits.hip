// Per-depression reductions with a fixed summation order.
//
// goSPL sums per pit with numpy group-by reductions (pitfilling.py:521-555
// volumes and depths, getpitvol functions.F90:1971-2003 for the depth-volume
// curves, flowplex.py:596-600 for the cascade inflow).  Here the cells of
// every pit are grouped once: the cells with a pit id, in ascending device
// order (a select, i.e. a scan, so no contended counter), are stably
// radix-sorted on the pit id, which leaves each pit's cells contiguous and in
// ascending order.  One workgroup per pit then reduces its cells: strided
// per-thread partial sums followed by a fixed LDS tree.  The result does not
// depend on scheduling (no floating-point atomics), so runs are reproducible
// bit for bit.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "common.h"

namespace gospl {

namespace {

struct IsPitCell {
  const int32_t* pitids;
  __device__ __forceinline__ bool operator()(const int32_t& i) const { return pitids[i] > 0; }
};

}  // namespace

__global__ void k_pg_keys(const int32_t* __restrict__ cells, const int32_t* __restrict__ ncells,
                          const int32_t* __restrict__ pitids, unsigned long long* __restrict__ keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= *ncells) return;
  const int c = cells[t];
  keys[t] = ((unsigned long long)(unsigned int)pitids[c] << 32) | (unsigned int)c;
}

// start/end of each pit's run in the sorted keys
__global__ void k_pg_segments(int64_t n, const unsigned long long* __restrict__ keys,
                              int32_t* __restrict__ start, int32_t* __restrict__ end) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const unsigned int p = (unsigned int)(keys[t] >> 32);
  if (t == 0 || (unsigned int)(keys[t - 1] >> 32) != p) start[p] = (int32_t)t;
  if (t == n - 1 || (unsigned int)(keys[t + 1] >> 32) != p) end[p] = (int32_t)(t + 1);
}

template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*sh)[GOSPL_BLOCK]) {
#pragma unroll
  for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int o = GOSPL_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
#pragma unroll
      for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + o];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = sh[k][0];
}

__device__ __forceinline__ void block_max2(double& a, double& b, double (*sh)[GOSPL_BLOCK]) {
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = GOSPL_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
    }
    __syncthreads();
  }
  a = sh[0][0];
  b = sh[1][0];
}

// Pit volume, maximum filled elevation and maximum depth (pitfilling.py:521-555,
// h clipped at sea level as in fillElevation, :872-876).  Block b -> pit b
// (row 0, no pit, gets the values of an empty pit).
__global__ __launch_bounds__(GOSPL_BLOCK) void k_pit_reduce_params(
    const unsigned long long* __restrict__ keys, const int32_t* __restrict__ start,
    const int32_t* __restrict__ end, const double* __restrict__ hl, const double* __restrict__ fill,
    const double* __restrict__ area, double sl, int64_t cap, double* __restrict__ params) {
  __shared__ double sh[2][GOSPL_BLOCK];
  const int p = blockIdx.x;
  const int b = start[p], e = end[p];
  double v[1] = {0.0};
  double hm = -INFINITY, dh = -INFINITY;
  // eight cells at a time, loads first, then the plain loop's additions in
  // its order
  constexpr int U = 8;
  int i = b + (int)threadIdx.x;
  for (; i + (U - 1) * GOSPL_BLOCK < e; i += U * GOSPL_BLOCK) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = (int)(keys[i + u * GOSPL_BLOCK] & 0xffffffffull);
    double f[U], h[U], a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      f[u] = fill[c[u]];
      h[u] = hl[c[u]];
      a[u] = area[c[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double hc = h[u] < sl ? sl : h[u];
      v[0] = v[0] + (f[u] - hc) * a[u];
      hm = fmax(hm, f[u]);
      dh = fmax(dh, f[u] - hc);
    }
  }
  for (; i < e; i += GOSPL_BLOCK) {
    const int c = (int)(keys[i] & 0xffffffffull);
    const double f = fill[c];
    const double h = hl[c];
    const double hc = h < sl ? sl : h;
    v[0] = v[0] + (f - hc) * area[c];
    hm = fmax(hm, f);
    dh = fmax(dh, f - hc);
  }
  block_sum<1>(v, sh);
  __syncthreads();
  block_max2(hm, dh, sh);
  if (threadIdx.x == 0) {
    params[p] = v[0];
    params[cap + p] = (e > b) ? hm : -1.0e8;
    params[2 * cap + p] = (e > b) ? dh : 0.0;
  }
}

// filled_lvl rows (pitfilling.py:877-881).
__global__ void k_pit_levels(int64_t npit, int64_t cap, const double* __restrict__ params,
                             double* __restrict__ lvl) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npit) return;
  const double hmax = params[cap + p];
  const double dh = params[2 * cap + p];
  double c[6] = {0, 0, 0, 0, 0, 0};
  if (params[p] > 0.0) {
    c[0] = hmax - dh;
    const double d5 = dh / 5.0;
    for (int k = 1; k < 6; ++k) c[k] = c[k - 1] + d5;
  }
  for (int k = 0; k < 5; ++k) lvl[k * cap + p] = c[k + 1];
}

// getpitvol (functions.F90:1971-2003) on h clipped at sea level: volumes
// below the first four levels; the fifth is the pit volume.
__global__ __launch_bounds__(GOSPL_BLOCK) void k_pit_reduce_vol(
    const unsigned long long* __restrict__ keys, const int32_t* __restrict__ start,
    const int32_t* __restrict__ end, const double* __restrict__ hl, const double* __restrict__ area,
    double sl, int64_t cap, const double* __restrict__ params, const double* __restrict__ lvl,
    double* __restrict__ fvol) {
  __shared__ double sh[4][GOSPL_BLOCK];
  const int p = blockIdx.x;
  const int b = start[p], e = end[p];
  double l[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) l[k] = lvl[k * cap + p];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  constexpr int U = 8;
  int i = b + (int)threadIdx.x;
  for (; i + (U - 1) * GOSPL_BLOCK < e; i += U * GOSPL_BLOCK) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = (int)(keys[i + u * GOSPL_BLOCK] & 0xffffffffull);
    double hh[U], aa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      hh[u] = hl[c[u]];
      aa[u] = area[c[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double h = hh[u] < sl ? sl : hh[u];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (l[k] > h) v[k] = v[k] + (l[k] - h) * aa[u];
    }
  }
  for (; i < e; i += GOSPL_BLOCK) {
    const int c = (int)(keys[i] & 0xffffffffull);
    double h = hl[c];
    if (h < sl) h = sl;
    const double a = area[c];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (l[k] > h) v[k] = v[k] + (l[k] - h) * a;
  }
  block_sum<4>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) fvol[k * cap + p] = v[k];
    fvol[4 * cap + p] = params[p];
  }
}

// Cascade inflow per pit: sum of fa over its lsink cells (flowplex.py:596-600).
__global__ __launch_bounds__(GOSPL_BLOCK) void k_pit_reduce_inflow(
    const unsigned long long* __restrict__ keys, const int32_t* __restrict__ start,
    const int32_t* __restrict__ end, const double* __restrict__ fa,
    const int8_t* __restrict__ lsink, double* __restrict__ inv) {
  __shared__ double sh[1][GOSPL_BLOCK];
  const int p = blockIdx.x;
  const int b = start[p], e = end[p];
  double v[1] = {0.0};
  // each thread's cells in increasing position, eight at a time with all
  // their loads issued before the additions (the order of the sum is that of
  // the plain strided loop; a depression of a million cells is otherwise a
  // chain of dependent gathers on one CU)
  constexpr int U = 8;
  int i = b + (int)threadIdx.x;
  for (; i + (U - 1) * GOSPL_BLOCK < e; i += U * GOSPL_BLOCK) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = (int)(keys[i + u * GOSPL_BLOCK] & 0xffffffffull);
    int8_t f[U];
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      f[u] = lsink[c[u]];
      x[u] = fa[c[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (f[u]) v[0] = v[0] + x[u];
  }
  for (; i < e; i += GOSPL_BLOCK) {
    const int c = (int)(keys[i] & 0xffffffffull);
    if (lsink[c]) v[0] = v[0] + fa[c];
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) inv[p] = v[0];
}

int scan_tmp_bytes(int64_t n, size_t* bytes) {
  *bytes = 0;
  GOSPL_CHECK(rocprim::exclusive_scan(nullptr, *bytes, (int32_t*)nullptr, (int32_t*)nullptr, 0, (size_t)n,
                                      rocprim::plus<int32_t>()));
  return 0;
}

int pit_tmp_bytes(int64_t N, size_t* bytes) {
  size_t b1 = 0, b2 = 0;
  rocprim::counting_iterator<int32_t> it(0);
  // (the select's workspace does not depend on its predicate)
  GOSPL_CHECK(rocprim::select(nullptr, b1, it, (int32_t*)nullptr, (int32_t*)nullptr, (size_t)N,
                              IsPitCell{nullptr}));
  GOSPL_CHECK(rocprim::radix_sort_keys(nullptr, b2, (unsigned long long*)nullptr,
                                       (unsigned long long*)nullptr, (size_t)N, 32, 64));
  size_t b3 = 0;
  GOSPL_CHECK(rocprim::select(nullptr, b3, it, (const int8_t*)nullptr, (int32_t*)nullptr,
                              (int32_t*)nullptr, (size_t)N));
  *bytes = b1 > b2 ? b1 : b2;
  if (b3 > *bytes) *bytes = b3;
  return 0;
}

int pit_cells(gospl_mesh* m, const int32_t* pitids, int32_t* cells, int32_t* ncells, hipStream_t s) {
  size_t bytes = m->cub_tmp_bytes;
  rocprim::counting_iterator<int32_t> it(0);
  GOSPL_CHECK(rocprim::select(m->cub_tmp, bytes, it, cells, ncells, (size_t)m->N, IsPitCell{pitids}, s));
  return 0;
}

__global__ void k_fill_list(FillList f) {
  const int e = blockIdx.y;
  unsigned char* p = static_cast<unsigned char*>(f.p[e]);
  const int64_t nb = f.nbytes[e];
  const uint32_t w = 0x01010101u * f.byte[e];
  const int64_t nw = nb >> 2;
  uint32_t* pw = reinterpret_cast<uint32_t*>(p);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x)
    pw[i] = w;
  if (blockIdx.x == 0 && (int64_t)threadIdx.x < (nb & 3)) p[(nw << 2) + threadIdx.x] = f.byte[e];
}

int fill_list(const FillList& f, hipStream_t s) {
  if (f.n <= 0) return 0;
  if (f.n > 8) {
    set_error("fill_list: %d buffers (at most 8)", f.n);
    return 1;
  }
  int64_t most = 0;
  for (int k = 0; k < f.n; ++k) {
    if (reinterpret_cast<uintptr_t>(f.p[k]) & 3) {
      set_error("fill_list: buffer %d is not 4-byte aligned", k);
      return 1;
    }
    if (f.nbytes[k] > most) most = f.nbytes[k];
  }
  int g = nblocks((most + 3) / 4);
  if (g > 1024) g = 1024;  // grid-stride beyond
  k_fill_list<<<dim3(g, f.n), GOSPL_BLOCK, 0, s>>>(f);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int ensure_groups(gospl_mesh* m, int64_t rows) {
  if (!m->pg_keys[0]) {
    GOSPL_CHECK(hipMalloc(&m->pg_keys[0], sizeof(unsigned long long) * (m->N + 64)));
    GOSPL_CHECK(hipMalloc(&m->pg_keys[1], sizeof(unsigned long long) * (m->N + 64)));
  }
  if (m->pg_cap < rows) {
    if (m->pg_seg) GOSPL_CHECK(hipFree(m->pg_seg));
    m->pg_seg = nullptr;
    GOSPL_CHECK(hipMalloc(&m->pg_seg, sizeof(int32_t) * 2 * (rows + 64)));
    m->pg_cap = rows;
  }
  return 0;
}

int pit_group(gospl_mesh* m, const int32_t* pitids, int64_t npits, const int32_t* cells,
              const int32_t* ncells_dev, int64_t ncells, hipStream_t s, bool cleared) {
  if (ensure_groups(m, npits + 1)) return 1;
  int32_t* start = m->pg_seg;
  int32_t* end = m->pg_seg + (m->pg_cap + 64);
  if (!cleared) GOSPL_CHECK(hipMemsetAsync(m->pg_seg, 0, sizeof(int32_t) * 2 * (m->pg_cap + 64), s));
  m->pg_ids = pitids;
  m->pg_npits = npits;
  m->pg_n = ncells;
  if (ncells == 0) return 0;
  const int g = nblocks(ncells);
  k_pg_keys<<<g, GOSPL_BLOCK, 0, s>>>(cells, ncells_dev, pitids, m->pg_keys[1]);
  GOSPL_CHECK(hipGetLastError());
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) <= npits) ++bits;
  size_t bytes = m->cub_tmp_bytes;
  GOSPL_CHECK(rocprim::radix_sort_keys(m->cub_tmp, bytes, m->pg_keys[1], m->pg_keys[0], (size_t)ncells, 32,
                                       32 + bits, s));
  k_pg_segments<<<g, GOSPL_BLOCK, 0, s>>>(ncells, m->pg_keys[0], start, end);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int pit_group_params(gospl_mesh* m, const double* hl, const double* fill, double sl, int64_t cap,
                     double* pitparams, double* filled_lvl, double* filled_vol, hipStream_t s) {
  const int64_t npits = m->pg_npits;
  const int32_t* start = m->pg_seg;
  const int32_t* end = m->pg_seg + (m->pg_cap + 64);
  // the three kernels write rows 0..npits of every table (one block or thread
  // per pit, empty pits included); rows past npits are left as they were
  k_pit_reduce_params<<<(int)(npits + 1), GOSPL_BLOCK, 0, s>>>(m->pg_keys[0], start, end, hl, fill,
                                                         m->area, sl, cap, pitparams);
  k_pit_levels<<<nblocks(npits + 1), GOSPL_BLOCK, 0, s>>>(npits + 1, cap, pitparams, filled_lvl);
  k_pit_reduce_vol<<<(int)(npits + 1), GOSPL_BLOCK, 0, s>>>(m->pg_keys[0], start, end, hl, m->area, sl,
                                                      cap, pitparams, filled_lvl, filled_vol);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

// Rows 0..np1-1 of the per-pit tables, one block of np1 after the other:
// spill (as f64, exact), pitparams rows 0-2, filled_lvl rows 0-4, filled_vol
// rows 0-4.
__global__ void k_pit_pack(int64_t np1, int64_t cap, const int32_t* __restrict__ spill,
                           const double* __restrict__ params, const double* __restrict__ lvl,
                           const double* __restrict__ vol, double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np1) return;
  out[p] = (double)spill[p];
  for (int r = 0; r < 3; ++r) out[(1 + r) * np1 + p] = params[r * cap + p];
  for (int r = 0; r < 5; ++r) out[(4 + r) * np1 + p] = lvl[r * cap + p];
  for (int r = 0; r < 5; ++r) out[(9 + r) * np1 + p] = vol[r * cap + p];
}

}  // namespace gospl

using namespace gospl;

extern "C" int gospl_pit_tables(gospl_mesh_t* m, int64_t npits, int64_t cap, const int32_t* spill,
                                const double* pitparams, const double* filled_lvl, const double* filled_vol,
                                double* out, gospl_stream_t stream) {
  hipStream_t s = stream;
  const int64_t np1 = npits + 1;
  if (npits < 0 || np1 > cap || !out) {
    set_error("gospl_pit_tables: bad arguments (npits=%lld, cap=%lld)", (long long)npits, (long long)cap);
    return 1;
  }
  const int64_t n = 14 * np1;
  if (m->pack_cap < n) {
    if (m->pack_buf) GOSPL_CHECK(hipFree(m->pack_buf));
    m->pack_buf = nullptr;
    m->pack_cap = 0;
    const int64_t want = n < 4096 ? 4096 : 2 * n;
    GOSPL_CHECK(hipMalloc(&m->pack_buf, sizeof(double) * want));
    m->pack_cap = want;
  }
  k_pit_pack<<<nblocks(np1), GOSPL_BLOCK, 0, s>>>(np1, cap, spill, pitparams, filled_lvl, filled_vol, m->pack_buf);
  GOSPL_CHECK(hipGetLastError());
  GOSPL_CHECK(hipMemcpyAsync(out, m->pack_buf, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  GOSPL_CHECK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int gospl_pit_inflow(gospl_mesh_t* m, const double* fa, const int8_t* lsink,
                                const int32_t* pitids, int64_t npit, double* inv,
                                gospl_stream_t stream) {
  hipStream_t s = stream;
  GOSPL_CHECK(hipMemsetAsync(inv, 0, sizeof(double) * (npit > 0 ? npit : 1), s));
  if (npit <= 1) return 0;
  if (m->pg_ids != pitids || m->pg_npits != npit - 1) {
    // no grouping from gospl_pit_info for these ids: build it
    int32_t* cells = m->wi[6];
    int32_t* ncells = m->dflag + 40;
    if (pit_cells(m, pitids, cells, ncells, s)) return 1;
    GOSPL_CHECK(hipMemcpyAsync(m->hflag + 18, ncells, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GOSPL_CHECK(hipStreamSynchronize(s));
    if (pit_group(m, pitids, npit - 1, cells, ncells, m->hflag[18], s)) return 1;
  }
  k_pit_reduce_inflow<<<(int)npit, GOSPL_BLOCK, 0, s>>>(
      m->pg_keys[0], m->pg_seg, m->pg_seg + (m->pg_cap + 64), fa, lsink, inv);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}
