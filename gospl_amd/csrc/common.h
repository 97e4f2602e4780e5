// Shared declarations of the gospl HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gospl_hip.h"

#define GOSPL_MAXW 12        // neighbour cap of the FV stencil
#define GOSPL_BLOCK 256

// Mesh handle: device geometry owned by the caller + library workspace.
struct gospl_mesh {
  int64_t N;
  int W;                      // ELL width (max degree)
  int flat;
  const int32_t* nbr;         // (W, N) device indices, -1 padded
  const int32_t* deg;         // (N)
  const double* elgt;         // (W, N)
  const double* vdist;        // (W, N)
  const double* area;         // (N)
  const int32_t* gid;         // (N) input-mesh ids
  const int32_t* tgid;        // (N) ids of the receivers' slope tie-break (gid unless set)
  const int8_t* outlet;       // (N) draining border flags (flat meshes)
  int device;
  int num_cu;
  int coop_blocks;            // co-resident blocks for the persistent DAG kernels
  int dag_threads;            // threads per block of the persistent DAG kernel
  int dag_interleave;         // frontier slices interleaved over blocks
  int dag_tail;               // frontier size below which one workgroup finishes the sweep
  int dag_coop;               // launch the persistent DAG kernel with the cooperative API
  int dag_csr;                // use the CSR rows even when packed records apply
  int dag_chain;              // flow accumulation as a dataflow sweep (records only)
  int chain_blocks[2];        // its grid, by sweep (flow, stream power)
  int chain_resident[2];      // workgroups of it that fit on the device at once
  int chain_pool;             // its pool capacity per wavefront
  int chain_share;            // pool size from which batches go to the global queue
  int chain_pollers;          // idle wavefronts that stay to poll it
  int chain_spin;             // bound of its waits
  int32_t* chain_ctl;         // its control words
  int32_t* chain_gq;          // its global queue (-1 between sweeps)
  long long* trace_v;         // its per-vertex solve times (GOSPL_DAG_TRACE)
  // workspace
  int32_t* wi[8];             // 8 x N int32
  int32_t* wdag[5];           // 5 x N int32 of the DAG sweeps alone (they may overlap pit information)
  void* dag_tmp;              // their rocPRIM workspace
  hipStream_t side;           // second stream of the native flow driver (default priority)
  int dag_pending_chain;      // what the issued sweep was (read by its finish)
  int dag_pending_trace;
  int dag_pending_ltrace;
  hipEvent_t side_ev[3];
  unsigned long long* wu[2];  // 2 x N uint64
  double* wd[4];              // 4 x N double
  int32_t* counts;            // N + 2 level counters
  int32_t* dflag;             // small device flag block (64 ints)
  int32_t* hflag;             // pinned host mirror (64 ints)
  double* wcoef;              // (wcoef_rows, N) coefficient rows
  int wcoef_rows;
  int32_t* wdon;              // donor rows of the receiver DAG
  double* wacc;               // weighted donor values, one per donor entry
  int32_t* wslot;             // (nrcv, N) slot of each (vertex, receiver) in its receiver's row
  double* wrec;               // packed vertex records of the DAG sweep (16 words per vertex)
  int8_t* rev;                // (W, N) position of i in the neighbour list of its neighbour q
  int64_t wdon_cap;
  long long* trace_buf;       // per-level trace (GOSPL_DAG_TRACE)
  int iter_hint[6];           // passes the relaxation loops needed last time (one slot per loop and mode)
  // cells grouped by depression (sorted (pit id << 32 | cell) keys + runs)
  unsigned long long* pg_keys[2];
  int32_t* pg_seg;            // start (pg_cap + 64) then end (pg_cap + 64)
  int64_t pg_cap;
  const int32_t* pg_ids;      // pit-id buffer the grouping was built from
  int64_t pg_npits;
  int64_t pg_n;               // grouped cells
  double* hcoef;              // hillslope operator rows (W, N), kept between calls
  double* hdiag;              // its diagonal
  float* hcoef32;             // FP32 copies of both (refined solve's corrections)
  float* hdiag32;
  double* adj_hcoef;          // hillslope rows from given sea flags (adjoint, frozen replay)
  double* adj_hdiag;
  // segmented (ensemble) calls: per-segment keys and control words
  unsigned long long* seg_keys;
  double* seg_ctl;
  int32_t* seg_aux;
  int seg_cap;
  int32_t* seg_haux;          // host copy of the last seg_aux upload
  int seg_haux_n;
  // native flow driver: pinned per-pit tables and staging, device staging
  double* flow_htab;
  double* flow_hstage;
  double* flow_dstage;
  int64_t flow_cap;           // doubles in each
  int64_t flow_n1;            // rows of the last call's tables
  int8_t* hsea;               // the sea flags the rows were built from
  double hpar[3];             // cda, cdm, dt they were built with
  int hvalid;
  double* pack_buf;           // per-pit tables packed for one read-back
  int64_t pack_cap;           // its capacity (doubles)
  int timing;                 // record HIP events around the DAG sweep kernels
  hipEvent_t tev[2];
  double t_ms[4];             // [0] flow-accumulation sweep, [1] SPL sweep, [2], [3] their refined solves
  int64_t t_calls[4];
  void* cub_tmp;              // rocPRIM workspace (scans, selects, radix sorts)
  double* rx_ctl;             // refined solves: control block (device, 64 doubles)
  double* rx_hctl;            // its pinned host mirror
  int rx_part_cap;            // partial-sum slots available in wd[0]
  size_t cub_tmp_bytes;
  double* nld_buf;            // non-linear diffusion: rows (W, N), diagonal and four N vectors
  int8_t* sed_mask;           // sediment routing: (N) int8 cell flags
  double* sed_buf;            // sediment routing: two N vectors
  double* kry_buf;            // coupled h / q solves: Krylov basis, its preconditioned copies, work vectors
  double* kry_h;              // pinned host block of their reductions
  double* kry_d;              // device block of their reductions
  double* sea_d;              // marine routing: six N vectors
  double* sea_ctl;            // its partial sums and control values
  double* sea_pts;            // coastline points and their grid-sorted copy (3 doubles each)
  int32_t* sea_i;             // coastline counts / offsets
  int32_t* sea_cell;          // grid cell of each coastline point
  int64_t sea_pcap;           // points the two hold
  double* sea_bvh;            // bounding boxes of the sorted points' buckets (6 doubles per tree node)
  int64_t sea_bvh_cap;        // nodes it holds
  unsigned long long* sea_h;  // pinned host block of the marine read-backs
  int32_t* sea_don;           // marine routing: donors of each vertex (W, N), by input id
  int8_t* sea_nd;             // their count
  unsigned long long* red_part;  // per-workgroup partial maxima of the convergence tests
};

namespace gospl {

void set_error(const char* fmt, ...);

// depression grouping (k_pits.hip)
int pit_tmp_bytes(int64_t N, size_t* bytes);  // rocPRIM workspace of the grouping
int scan_tmp_bytes(int64_t n, size_t* bytes);  // rocPRIM workspace of an int32 exclusive scan
int pit_cells(gospl_mesh* m, const int32_t* pitids, int32_t* cells, int32_t* ncells, hipStream_t s);
int ensure_groups(gospl_mesh* m, int64_t rows);  // grouping buffers for `rows` pits
// `cleared`: the caller has already zeroed the segment table (2 * (pg_cap + 64) ints)
int pit_group(gospl_mesh* m, const int32_t* pitids, int64_t npits, const int32_t* cells,
              const int32_t* ncells_dev, int64_t ncells, hipStream_t s, bool cleared = false);
int pit_group_params(gospl_mesh* m, const double* hl, const double* fill, double sl, int64_t cap,
                     double* pitparams, double* filled_lvl, double* filled_vol, hipStream_t s);

// receiver-DAG workspace (k_dag.hip): packed records + reverse neighbour
// positions; donor rows / overflow entries of capacity `cap`
int ensure_records(gospl_mesh* m, hipStream_t s);
int ensure_donors(gospl_mesh* m, int64_t cap);
int ensure_seg(gospl_mesh* m, int nseg);  // per-segment keys / control words (k_fill.hip)
// One exact sweep of the receiver DAG (k_dag.hip), synchronising `s`: mode 0
// the flow system (I - W^T) x = rhs (coef = W, self receivers dropped, a
// draining cycle iterated to its fixed point), mode 1 the erosion system
// diag x + off x_rcv = rhs (coef = off).  *unsolved (HOST, may be NULL): the
// vertices left unsolved (an undrained cycle).
int dag_solve_ext(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef, const double* diag,
                  const double* rhs, double* x, int64_t* unsolved, hipStream_t s);
// The stream-power rows of SPL._eroMats (off (nrcv, N), diag (N)) and the
// state update of erodepSPL from the new surface x (k_dag.hip).
int spl_coef_launch(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* wgt, const double* dist,
                    const double* fal, const double* kfac, double K, double mexp, double dt, double sl,
                    const double* lfill, const double* h, double* off, double* diag, hipStream_t s);
int spl_update_launch(gospl_mesh* m, const double* x, const int8_t* lsink, double dt, double* h, double* cumed,
                      double* eb, double* eblocal, hipStream_t s);
// Kbr of the non-linear stream power (K kfac dt FAL^m elimiter, 0 at sea; k_dag.hip).
int nl_kbr_launch(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* h, const double* fal, const double* kfac,
                  double K, double mexp, double dt, double sl, const double* lfill, double* kbr, hipStream_t s);
// gospl_flowacc in two halves (k_dag.hip): enqueue, then wait and read the status
int flowacc_issue(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* wgt, const double* b, double* q,
                  hipStream_t s);
int flowacc_finish(gospl_mesh* m, int nrcv, const int32_t* rcv, const double* wgt, int32_t* nlevels,
                   int64_t* unsolved, hipStream_t s);

// Refined solve on the receiver DAG (k_refine.hip): x (in: initial guess,
// out: solution) of the flow system (mode 0: (I - W^T) x = rhs, coef = W) or
// the stream-power system (mode 1: diag x + off x_rcv = rhs, coef = off) by
// correction sweeps in 64-bit fixed point and FP64 residuals, until
// ||rhs - A x||_2 <= rtol ||rhs||_2 or maxpass corrections.
struct RefineStatus {
  int passes;        // correction sweeps run
  double relres;     // final ||r||_2 / ||rhs||_2
  int fallback;      // the sweep left vertices unsolved (receiver cycle) or a value overflowed
  int aborted;       // a bounded wait of the dataflow kernel ran out
};
int refine_solve(gospl_mesh* m, int mode, int nrcv, const int32_t* rcv, const double* coef,
                 const double* diag, const double* rhs, double* x, double rtol, int maxpass,
                 RefineStatus* st, hipStream_t s);
}  // namespace gospl

extern "C" {
int gospl_ensure_coef(gospl_mesh* m, int rows);        // internal (not exported in the header)
int gospl_dag_init_occupancy(gospl_mesh* m);           // internal
}

namespace gospl {

inline int nblocks(int64_t n, int b = GOSPL_BLOCK) {
  int64_t g = (n + b - 1) / b;
  return (int)(g < 1 ? 1 : g);
}

// Several buffers set to a byte value in one launch (each hipMemsetAsync is a
// launch of its own, and back-to-back small ones leave the GPU waiting on the
// host).  Buffers are 4-byte aligned; blockIdx.y selects the buffer.
struct FillList {
  void* p[8];
  int64_t nbytes[8];
  uint8_t byte[8];
  int n;
};

__global__ void k_fill_list(FillList f);

// Enqueues the fills on `s` (no-op for an empty list).
int fill_list(const FillList& f, hipStream_t s);

// Total order on doubles as unsigned 64-bit keys (no NaNs expected).
__device__ __forceinline__ unsigned long long dkey(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}
inline double dkey_inv_host(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  __builtin_memcpy(&d, &b, sizeof(d));
  return d;
}

// bound of the hillslope solve's iterate-to-convergence mode
#define GOSPL_HILL_MAX_SWEEPS 1000000

// Append `value` for every lane with `pred` set, one atomicAdd per wavefront.
// Must be reached by all lanes of the wavefront.
__device__ __forceinline__ void wave_append(bool pred, int value, int32_t* counter, int32_t* out) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(mask));
  base = __shfl(base, leader);
  if (pred) out[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

}  // namespace gospl

#define GOSPL_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      gospl::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                     \
      return 1;                                                                 \
    }                                                                           \
  } while (0)
