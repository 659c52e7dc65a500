// Atomic-throughput probe: the ceiling the ray cast's table updates are measured against.
// Every thread issues a fixed number of non-returning int32 atomic adds of 1 into a table of
// V words, at addresses from a per-thread linear congruential sequence:
//   pattern 0  scattered over the whole table (device scope)
//   pattern 1  64 hot words (device scope): same-address contention
//   pattern 2  scattered, but block b only inside region b mod 8 of 8 equal regions (device
//              scope): the address stream of each XCD (workgroups are dispatched to the 8
//              XCDs round-robin) stays in one part of the table
//   pattern 3  as 0 with workgroup-scope atomics (a timing point only: such adds are not
//              atomic across workgroups of different XCDs)
//   pattern 4  each wavefront instruction: its 64 lanes on 64 consecutive words at a random
//              place (two 128-byte lines)
//   pattern 5  each wavefront instruction: 8 random 128-byte lines, 8 lanes on consecutive
//              words of each
#include <hip/hip_runtime.h>

#include "gvom_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_probe_atomics(int32_t *__restrict__ table, uint32_t V, int per_thread,
                                                       int pattern) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = tid * 2654435761u + 12345u;
    const uint32_t region = V / 8u;
    const uint32_t rbase = (blockIdx.x & 7u) * region;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t hw = (blockIdx.x * 4u + (threadIdx.x >> 6)) * 2654435761u + 777u;   // per wavefront
    for (int k = 0; k < per_thread; ++k) {
        h = h * 1664525u + 1013904223u;
        hw = hw * 1664525u + 1013904223u;
        const uint32_t r = h >> 3;
        uint32_t idx;
        if (pattern == 1) idx = r & 63u;
        else if (pattern == 2) idx = rbase + r % region;
        else if (pattern == 4) idx = ((hw >> 3) % (V / 64u)) * 64u + lane;
        else if (pattern == 5) {
            const uint32_t g8 = lane >> 3, hl = (hw ^ (g8 * 0x9E3779B9u)) * 2246822519u;
            idx = ((hl >> 3) % (V / 32u)) * 32u + (lane & 7u);
        } else idx = r % V;
        if (pattern == 3)
            __hip_atomic_fetch_add(table + idx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            __hip_atomic_fetch_add(table + idx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

hipError_t launch_probe_atomics(int32_t *table, uint32_t V, int blocks, int per_thread, int pattern, hipStream_t s) {
    k_probe_atomics<<<blocks, 256, 0, s>>>(table, V, per_thread, pattern);
    return hipGetLastError();
}
