// coherent.h -- data passed between the workgroups of one launch.  Such data is written and
// read only with device-scope relaxed atomics, which are coherent across the XCDs' L2 caches
// without a cache write-back or invalidation (what a device-scope fence costs).  A writer
// waits for its stores to complete (wait_memory) before it signals another workgroup through
// an atomic; a reader issues its loads after it has seen that signal.
#pragma once

#include <hip/hip_runtime.h>

// this wave's outstanding memory operations complete (stores acknowledged at their coherence
// point) before any later one is issued
__device__ inline void wait_memory()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <typename T>
__device__ inline T ld_shared_wg(const T *p)
{
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ inline void st_shared_wg(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
