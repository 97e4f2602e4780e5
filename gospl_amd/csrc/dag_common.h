// Device helpers shared by the receiver-DAG sweeps (k_dag.hip: the exact
// FP64 sweeps; k_refine.hip: the fixed-point correction sweeps of the
// refined solves).
#pragma once

#include "common.h"

namespace gospl {

// Small int32 arrays cleared by the first workgroup of a launch.
struct ZeroList {
  int32_t* p[4];
  int len[4];
  int n;
};

// Appends into a frontier list in global memory; `blk` (LDS, may be null)
// accumulates the workgroup's appended count.
struct GlobalAppend {
  int32_t* cnt;
  int32_t* out;
  int32_t* blk;
  __device__ __forceinline__ int reserve(int n) const {
    if (blk) atomicAdd(blk, n);
    return atomicAdd(cnt, n);
  }
  __device__ __forceinline__ void write(int idx, int value) const { out[idx] = value; }
};

#define CHAIN_THREADS 256
#define CHAIN_POOL 512
#define CHAIN_SPIN_MAX (1 << 21)
#define CHAIN_POLLERS 32  // idle wavefronts that stay to poll the queue (256: a few % slower)
// control words, one 128-byte line each
#define CQ_QUEUE 0    // 64-bit: batches reserved << 32 | batches claimed
#define CQ_ACTIVE 32  // wavefronts holding work
#define CQ_IDLERS 64  // idle wavefronts polling the queue
#define CQ_ABORT 96
#define CQ_WORDS 128
#define CQ_RETRIES 112  // traced sweeps: passes that found a row incomplete

__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent_i(const int32_t* p) {
  return __hip_atomic_load(const_cast<int32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_i(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS written by some lanes of a wavefront and read by others: order the
// accesses (one wavefront's LDS operations complete in issue order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace gospl
