// Block barrier for LDS exchanges that leaves global loads in flight.
//
// __syncthreads() is a workgroup release/acquire around s_barrier, and on
// this target the release waits for every outstanding vector memory
// operation (s_waitcnt vmcnt(0)): a load issued for the next step (a
// prefetch) is drained at every barrier, so its latency is never hidden.
// lds_barrier() waits only for the wave's own LDS operations (lgkmcnt), then
// synchronises; empty asm statements with a memory clobber keep the compiler
// from moving memory accesses across it.  Data written to LDS by direct
// global-to-LDS loads is not covered: wait for those with an explicit vmcnt
// before the barrier.  Global memory written before the barrier is not made
// visible to other waves by it.
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt unconstrained
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace mf
