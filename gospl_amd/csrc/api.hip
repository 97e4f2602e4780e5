// Mesh handle, error reporting and the small per-step kernels of the
// flow-accumulation driver (runoff source, pit-overflow cascade pieces,
// discharge finalisation).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>


#include "common.h"

namespace gospl {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// Runoff source b = rain * area, zero at sea (flowplex.py:794-795).
__global__ void k_rain_source(int64_t N, const double* __restrict__ rain,
                              const double* __restrict__ area, const double* __restrict__ lfill,
                              double sl, double* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  b[i] = (lfill[i] <= sl) ? 0.0 : rain[i] * area[i];
}

// Water level of pits (flowplex.py:624-653).
__global__ void k_pit_water(int64_t N, const int32_t* __restrict__ pit,
                            const int8_t* __restrict__ full, const double* __restrict__ level,
                            const double* __restrict__ lfill, double* __restrict__ water) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int32_t p = pit[i];
  if (p < 0) return;
  if (full[p]) {
    water[i] = lfill[i];
  } else {
    double l = level[p];
    if (!isnan(l) && water[i] <= l) water[i] = l;
  }
}

// One cascade pass's device update from a packed staging buffer
// (flowplex.py:602-666): the pit water levels as k_pit_water, with
// full[p] = stage[p] != 0 and level[p] = stage[npit + p], and the cleared
// right-hand side.  The spill values are written by k_cascade_spill after it.
__global__ void k_cascade_seed(int64_t N, const int32_t* __restrict__ pit, const double* __restrict__ stage,
                               int64_t npit, const double* __restrict__ lfill, double* __restrict__ water,
                               double* __restrict__ rhs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  rhs[i] = 0.0;
  int32_t p = pit[i];
  if (p < 0 || p >= npit) return;
  if (stage[p] != 0.0) {
    water[i] = lfill[i];
  } else {
    double l = stage[npit + p];
    if (!isnan(l) && water[i] <= l) water[i] = l;
  }
}

// rhs[v_j] = vol_j / dt at the spill vertices v_j (distinct), the volumes a
// full pit passes on (flowplex.py:612-622, 667).
__global__ void k_cascade_spill(int64_t N, const double* __restrict__ vid, const double* __restrict__ vol,
                                int64_t nspill, double dt, double* __restrict__ rhs) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nspill) return;
  const double v = vid[j];
  if (!(v >= 0.0 && v < (double)N)) return;  // the host checks the ids; never write out of range
  rhs[(int64_t)v] = vol[j] / dt;
}

// nFA = q*dt ; FAL += (nFA masked where hl < water)/dt (flowplex.py:674-680).
__global__ void k_cascade_add(int64_t N, const double* __restrict__ q, const double* __restrict__ hl,
                              const double* __restrict__ water, double dt, double* __restrict__ nfa,
                              double* __restrict__ fal) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double v = q[i] * dt;
  nfa[i] = v;
  double f = (hl[i] < water[i]) ? 0.0 : v;
  fal[i] = fal[i] + f / dt;
}

// Discharge at filled pits and its maximum (flowplex.py:906-918).  A fixed
// grid strides over the mesh, so the maximum takes one atomic per workgroup
// of that grid (same-address atomics serialise).
#define FINALIZE_BLOCKS 512
__global__ void k_finalize_a(int64_t N, const double* __restrict__ hl, const double* __restrict__ water,
                             const int32_t* __restrict__ pit, int has_pits, double* __restrict__ fal,
                             unsigned long long* __restrict__ maxkey) {
  double v = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    double f = fal[i];
    if (has_pits && hl[i] <= water[i] && pit[i] > -1) f = 0.0;
    fal[i] = f;
    v = v > f ? v : f;
  }
  // wave-level max, block-level max, then one atomic per block
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = v > w ? v : w;
  }
  __shared__ double s_max[GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s_max[0];
    for (int k = 1; k < GOSPL_BLOCK / 64; ++k) b = b > s_max[k] ? b : s_max[k];
    if (b > -INFINITY) atomicMax(maxkey, dkey(b));
  }
}

__global__ void k_finalize_b(int64_t N, const double* __restrict__ hl, const double* __restrict__ water,
                             const int32_t* __restrict__ pit, int has_pits,
                             const unsigned long long* __restrict__ maxkey, double* __restrict__ fal,
                             double* __restrict__ fillfal) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double f = fal[i];
  double g = f;
  if (has_pits && hl[i] <= water[i] && pit[i] > -1) g = dkey_inv(*maxkey) * 0.1;
  fal[i] = f < 0.0 ? 0.0 : f;
  fillfal[i] = g < 0.0 ? 0.0 : g;
}

// gospl_flow_finalize per segment (ensembles): each segment's own pit flag
// and discharge maximum; padding vertices (segment < 0) are only clamped.
__global__ __launch_bounds__(GOSPL_BLOCK) void k_finalize_seg_a(int64_t N, const double* __restrict__ hl,
                                                               const double* __restrict__ water,
                                                               const int32_t* __restrict__ pit,
                                                               const int32_t* __restrict__ seg,
                                                               const int8_t* __restrict__ has_pits,
                                                               double* __restrict__ fal,
                                                               unsigned long long* __restrict__ keys) {
  __shared__ double s_max[GOSPL_BLOCK / 64];
  __shared__ int s_lo, s_hi;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = i < N ? seg[i] : -1;
  double v = -INFINITY;
  if (sg >= 0) {
    double f = fal[i];
    if (has_pits[sg] && hl[i] <= water[i] && pit[i] > -1) f = 0.0;
    fal[i] = f;
    v = f;
  }
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
  if (lo != hi) {
    if (sg >= 0 && v > -INFINITY) atomicMax(keys + sg, dkey(v));
    return;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_down(v, o, 64);
    v = v > w ? v : w;
  }
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s_max[0];
    for (int k = 1; k < GOSPL_BLOCK / 64; ++k) b = b > s_max[k] ? b : s_max[k];
    if (b > -INFINITY) atomicMax(keys + lo, dkey(b));
  }
}

__global__ void k_finalize_seg_b(int64_t N, const double* __restrict__ hl, const double* __restrict__ water,
                                 const int32_t* __restrict__ pit, const int32_t* __restrict__ seg,
                                 const int8_t* __restrict__ has_pits, const unsigned long long* __restrict__ keys,
                                 double* __restrict__ fal, double* __restrict__ fillfal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int sg = seg[i];
  const double f = fal[i];
  double g = f;
  if (sg >= 0 && has_pits[sg] && hl[i] <= water[i] && pit[i] > -1) g = dkey_inv(keys[sg]) * 0.1;
  fal[i] = f < 0.0 ? 0.0 : f;
  fillfal[i] = g < 0.0 ? 0.0 : g;
}

// gospl_flow_finalize without a handle or any allocation: the discharge
// maximum goes through `scratch`, one partial maximum (order-preserving key)
// per workgroup of k_fin_partial, each workgroup striding over the mesh; every
// call writes all the slots it reads, so nothing needs clearing.
#define NOALLOC_PARTIALS 1024
static __device__ __align__(256) unsigned long long scratch[NOALLOC_PARTIALS];

__global__ void k_fin_partial(int64_t N, const double* __restrict__ hl, const double* __restrict__ water,
                              const int32_t* __restrict__ pit, int has_pits, double* __restrict__ fal,
                              unsigned long long* __restrict__ part) {
  double v = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    double f = fal[i];
    if (has_pits && hl[i] <= water[i] && pit[i] > -1) f = 0.0;
    fal[i] = f;
    v = v > f ? v : f;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = v > w ? v : w;
  }
  __shared__ double s_max[GOSPL_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s_max[0];
    for (int k = 1; k < GOSPL_BLOCK / 64; ++k) b = b > s_max[k] ? b : s_max[k];
    part[blockIdx.x] = dkey(b);
  }
}

__global__ void k_fin_apply(int64_t N, const double* __restrict__ hl, const double* __restrict__ water,
                            const int32_t* __restrict__ pit, int has_pits,
                            const unsigned long long* __restrict__ part, int nparts,
                            double* __restrict__ fal, double* __restrict__ fillfal) {
  __shared__ unsigned long long s_key[GOSPL_BLOCK / 64];
  unsigned long long k = 0ull;
  for (int t = threadIdx.x; t < nparts; t += blockDim.x) k = k > part[t] ? k : part[t];
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(k, o, 64);
    k = k > w ? k : w;
  }
  if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = k;
  __syncthreads();
  for (int w = 0; w < GOSPL_BLOCK / 64; ++w) k = k > s_key[w] ? k : s_key[w];
  const double fmaxv = dkey_inv(k);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double f = fal[i];
  double g = f;
  if (has_pits && hl[i] <= water[i] && pit[i] > -1) g = fmaxv * 0.1;
  fal[i] = f < 0.0 ? 0.0 : f;
  fillfal[i] = g < 0.0 ? 0.0 : g;
}

// Address of `scratch` on each device, looked up once per device with that
// device current (the host-side name is not a device address).
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static int scratch_on_device(void** out) {
  int dev = 0;
  GOSPL_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("gospl_flow_finalize_noalloc: device %d outside 0..63", dev);
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GOSPL_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Counts in *asym the neighbour relations that are not mutual.
__global__ void k_nbr_symmetry(int64_t N, int W, const int32_t* __restrict__ nbr,
                               const int32_t* __restrict__ deg, int32_t* __restrict__ asym) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int dg = deg[i];
  for (int q = 0; q < W && q < dg; ++q) {
    const int j = nbr[(int64_t)q * N + i];
    if (j < 0) continue;
    bool found = false;
    const int dj = deg[j];
    for (int p = 0; p < W && p < dj; ++p) found = found || nbr[(int64_t)p * N + j] == (int)i;
    if (!found) atomicAdd(asym, 1);
  }
}

}  // namespace gospl

using namespace gospl;

extern "C" {

const char* gospl_last_error(void) { return g_err; }

}  // extern "C"

namespace gospl {

// Counts the neighbour relations of the (W, N) table that are not mutual; the
// receiver-DAG sweeps find a donor's row entry through the receiver's
// neighbour list, so every relation must be mutual (as in a triangulation
// whose vertices have at most GOSPL_MAXW neighbours).
static int check_symmetry(int64_t N, int W, const int32_t* nbr, const int32_t* deg, const char* who) {
  int32_t* asym = nullptr;
  GOSPL_CHECK(hipMalloc(&asym, sizeof(int32_t)));
  int32_t h = 0;
  hipError_t e = hipMemset(asym, 0, sizeof(int32_t));
  if (e == hipSuccess) {
    k_nbr_symmetry<<<nblocks(N), GOSPL_BLOCK>>>(N, W, nbr, deg, asym);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&h, asym, sizeof(int32_t), hipMemcpyDeviceToHost);
  (void)hipFree(asym);
  if (e != hipSuccess) {
    set_error("%s: neighbour check failed: %s", who, hipGetErrorString(e));
    return 1;
  }
  if (h) {
    set_error("%s: %d neighbour relations are not mutual (a vertex with more than %d neighbours?)", who,
              (int)h, GOSPL_MAXW);
    return 1;
  }
  return 0;
}

// Device workspace of a new handle (every pointer starts null, so a failure
// part way leaves a handle gospl_mesh_destroy can release).
static int mesh_alloc(gospl_mesh* m) {
  const int64_t N = m->N;
  GOSPL_CHECK(hipGetDevice(&m->device));
  hipDeviceProp_t prop;
  GOSPL_CHECK(hipGetDeviceProperties(&prop, m->device));
  m->num_cu = prop.multiProcessorCount;
  for (int k = 0; k < 8; ++k) GOSPL_CHECK(hipMalloc(&m->wi[k], sizeof(int32_t) * (N + 64)));
  for (int k = 0; k < 5; ++k) GOSPL_CHECK(hipMalloc(&m->wdag[k], sizeof(int32_t) * (N + 64)));
  for (int k = 0; k < 2; ++k) GOSPL_CHECK(hipMalloc(&m->wu[k], sizeof(unsigned long long) * (N + 64)));
  for (int k = 0; k < 4; ++k) GOSPL_CHECK(hipMalloc(&m->wd[k], sizeof(double) * (N + 64)));
  GOSPL_CHECK(hipMalloc(&m->counts, sizeof(int32_t) * (N + 64)));
  GOSPL_CHECK(hipMalloc(&m->dflag, sizeof(int32_t) * 64));
  GOSPL_CHECK(hipMalloc(&m->red_part, sizeof(unsigned long long) * 3 * (nblocks(N) + 64)));
  GOSPL_CHECK(hipHostMalloc((void**)&m->hflag, sizeof(int32_t) * 64, hipHostMallocDefault));
  // rocPRIM workspace: scans of N + 1 ints, selects over N ids, sorts of N 64-bit keys
  size_t bytes = 0, b2 = 0;
  if (scan_tmp_bytes(N + 1, &bytes)) return 1;
  if (pit_tmp_bytes(N, &b2)) return 1;
  if (b2 > bytes) bytes = b2;
  m->cub_tmp_bytes = bytes + 256;
  GOSPL_CHECK(hipMalloc(&m->cub_tmp, m->cub_tmp_bytes));
  GOSPL_CHECK(hipMalloc(&m->dag_tmp, m->cub_tmp_bytes));
  if (gospl_dag_init_occupancy(m)) return 1;
  if (gospl_ensure_coef(m, 3)) return 1;
  return 0;
}

}  // namespace gospl

extern "C" {

int gospl_mesh_create(int64_t N, int32_t W, int32_t flat, const int32_t* nbr, const int32_t* deg,
                      const double* elgt, const double* vdist, const double* area,
                      const int32_t* gid, const int8_t* outlet, gospl_mesh_t** out) {
  if (!out || N <= 0 || W <= 0 || W > GOSPL_MAXW || N >= (int64_t)INT32_MAX - 2) {
    set_error("gospl_mesh_create: bad arguments (N=%lld, W=%d)", (long long)N, (int)W);
    return 1;
  }
  if (check_symmetry(N, W, nbr, deg, "gospl_mesh_create")) return 1;
  gospl_mesh* m = new gospl_mesh();
  std::memset(m, 0, sizeof(*m));
  m->N = N;
  m->W = W;
  m->flat = flat;
  m->nbr = nbr;
  m->deg = deg;
  m->elgt = elgt;
  m->vdist = vdist;
  m->area = area;
  m->gid = gid;
  m->tgid = gid;
  m->outlet = outlet;
  if (mesh_alloc(m)) {
    gospl_mesh_destroy(m);  // releases what was allocated before the failure
    return 1;
  }
  *out = m;
  return 0;
}

int gospl_ensure_coef(gospl_mesh* m, int rows) {
  if (m->wcoef_rows >= rows) return 0;
  if (m->wcoef) GOSPL_CHECK(hipFree(m->wcoef));
  m->wcoef = nullptr;
  GOSPL_CHECK(hipMalloc(&m->wcoef, sizeof(double) * rows * (m->N + 64)));
  m->wcoef_rows = rows;
  return 0;
}

int gospl_mesh_touch(gospl_mesh_t* m) {
  if (!m) {
    set_error("gospl_mesh_touch: null handle");
    return 1;
  }
  // the lent neighbour table may have changed: check it again, and drop the
  // reverse-neighbour positions built from it (rebuilt by the next sweep)
  if (check_symmetry(m->N, m->W, m->nbr, m->deg, "gospl_mesh_touch")) return 1;
  if (m->rev) {
    GOSPL_CHECK(hipFree(m->rev));
    m->rev = nullptr;
  }
  m->hvalid = 0;
  return 0;
}

int gospl_mesh_destroy(gospl_mesh_t* m) {
  if (!m) return 0;
  // best effort: release everything even if one call fails
  auto dfree = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (int k = 0; k < 8; ++k) dfree(m->wi[k]);
  for (int k = 0; k < 5; ++k) dfree(m->wdag[k]);
  dfree(m->dag_tmp);
  if (m->side) {
    (void)hipStreamDestroy(m->side);
    for (int k = 0; k < 3; ++k) (void)hipEventDestroy(m->side_ev[k]);
  }
  for (int k = 0; k < 2; ++k) dfree(m->wu[k]);
  for (int k = 0; k < 4; ++k) dfree(m->wd[k]);
  dfree(m->counts);
  dfree(m->dflag);
  dfree(m->chain_ctl);
  dfree(m->chain_gq);
  dfree(m->trace_v);
  if (m->hflag) (void)hipHostFree(m->hflag);
  dfree(m->cub_tmp);
  dfree(m->wcoef);
  if (m->tev[0]) (void)hipEventDestroy(m->tev[0]);
  if (m->tev[1]) (void)hipEventDestroy(m->tev[1]);
  dfree(m->wdon);
  dfree(m->wacc);
  dfree(m->wslot);
  dfree(m->wrec);
  dfree(m->rev);
  dfree(m->trace_buf);
  dfree(m->pg_keys[0]);
  dfree(m->pg_keys[1]);
  dfree(m->pg_seg);
  dfree(m->pack_buf);
  dfree(m->hcoef);
  dfree(m->rx_ctl);
  if (m->rx_hctl) (void)hipHostFree(m->rx_hctl);
  dfree(m->hdiag);
  dfree(m->hcoef32);
  dfree(m->adj_hcoef);
  dfree(m->adj_hdiag);
  dfree(m->nld_buf);
  dfree(m->sed_mask);
  dfree(m->sed_buf);
  dfree(m->kry_buf);
  dfree(m->kry_d);
  if (m->kry_h) (void)hipHostFree(m->kry_h);
  dfree(m->red_part);
  dfree(m->sea_d);
  dfree(m->sea_don);
  dfree(m->sea_nd);
  dfree(m->sea_ctl);
  dfree(m->sea_pts);
  dfree(m->sea_i);
  dfree(m->sea_cell);
  dfree(m->sea_bvh);
  if (m->sea_h) (void)hipHostFree(m->sea_h);
  dfree(m->seg_keys);
  dfree(m->seg_ctl);
  dfree(m->seg_aux);
  free(m->seg_haux);
  if (m->flow_htab) (void)hipHostFree(m->flow_htab);
  if (m->flow_hstage) (void)hipHostFree(m->flow_hstage);
  dfree(m->flow_dstage);
  dfree(m->hdiag32);
  dfree(m->hsea);
  delete m;
  return 0;
}

int gospl_rain_source(gospl_mesh_t* m, const double* rain, const double* lfill, double sl,
                      double* b, gospl_stream_t stream) {
  k_rain_source<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, rain, m->area, lfill, sl, b);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_pit_water(gospl_mesh_t* m, const int32_t* pitids, const int8_t* full, const double* level,
                    const double* lfill, double* water, gospl_stream_t stream) {
  k_pit_water<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, pitids, full, level, lfill, water);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_cascade_seed(gospl_mesh_t* m, const double* stage, int64_t npit, int64_t nspill,
                       const int32_t* pitids, const double* lfill, double* water, double dt, double* rhs,
                       gospl_stream_t stream) {
  if (npit < 0 || nspill < 0 || nspill > npit) {
    set_error("gospl_cascade_seed: bad sizes npit=%lld nspill=%lld", (long long)npit, (long long)nspill);
    return 1;
  }
  k_cascade_seed<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, pitids, stage, npit, lfill, water, rhs);
  if (nspill > 0)
    k_cascade_spill<<<nblocks(nspill), GOSPL_BLOCK, 0, stream>>>(m->N, stage + 2 * npit,
                                                                 stage + 2 * npit + nspill, nspill, dt, rhs);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_cascade_add(gospl_mesh_t* m, const double* q, const double* hl, const double* water,
                      double dt, double* nfa, double* fal, gospl_stream_t stream) {
  k_cascade_add<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, q, hl, water, dt, nfa, fal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_flow_finalize_noalloc(int64_t N, const double* hl, const double* water, const int32_t* pitids,
                                int32_t has_pits, double* fal, double* fillfal, gospl_stream_t stream) {
  if (N <= 0 || N > (int64_t)INT32_MAX) {
    set_error("gospl_flow_finalize_noalloc: N=%lld outside 1..%d", (long long)N, INT32_MAX);
    return 1;
  }
  void* part = nullptr;
  if (scratch_on_device(&part)) return 1;
  const int g = nblocks(N) < NOALLOC_PARTIALS ? nblocks(N) : NOALLOC_PARTIALS;
  k_fin_partial<<<g, GOSPL_BLOCK, 0, stream>>>(N, hl, water, pitids, has_pits, fal,
                                               static_cast<unsigned long long*>(part));
  k_fin_apply<<<nblocks(N), GOSPL_BLOCK, 0, stream>>>(N, hl, water, pitids, has_pits,
                                                      static_cast<const unsigned long long*>(part), g, fal,
                                                      fillfal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_flow_finalize(gospl_mesh_t* m, const double* hl, const double* water,
                        const int32_t* pitids, int32_t has_pits, double* fal, double* fillfal,
                        gospl_stream_t stream) {
  unsigned long long* key = m->wu[1];
  GOSPL_CHECK(hipMemsetAsync(key, 0, sizeof(unsigned long long), stream));
  k_finalize_a<<<FINALIZE_BLOCKS, GOSPL_BLOCK, 0, stream>>>(m->N, hl, water, pitids, has_pits, fal, key);
  k_finalize_b<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, hl, water, pitids, has_pits, key, fal,
                                                           fillfal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

int gospl_mesh_set_tiebreak(gospl_mesh_t* m, const int32_t* tgid) {
  if (!m) {
    set_error("gospl_mesh_set_tiebreak: null handle");
    return 1;
  }
  m->tgid = tgid ? tgid : m->gid;
  return 0;
}

int gospl_flow_finalize_seg(gospl_mesh_t* m, const double* hl, const double* water, const int32_t* pitids,
                            const int32_t* seg, int32_t nseg, const int8_t* has_pits, double* fal, double* fillfal,
                            gospl_stream_t stream) {
  if (nseg < 1 || !seg || !has_pits) {
    set_error("gospl_flow_finalize_seg: need segment and per-segment pit flag arrays, nseg >= 1");
    return 1;
  }
  if (ensure_seg(m, nseg)) return 1;
  GOSPL_CHECK(hipMemsetAsync(m->seg_keys, 0, sizeof(unsigned long long) * nseg, stream));
  k_finalize_seg_a<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, hl, water, pitids, seg, has_pits, fal,
                                                              m->seg_keys);
  k_finalize_seg_b<<<nblocks(m->N), GOSPL_BLOCK, 0, stream>>>(m->N, hl, water, pitids, seg, has_pits, m->seg_keys,
                                                              fal, fillfal);
  GOSPL_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
