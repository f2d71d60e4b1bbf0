// Atomic-throughput kernels for mi355nufft_measure_atomics: the rates the
// scatter's roofline is quoted against.  Every index is masked into its
// array and every loop has a fixed trip count.
#pragma once

#include <hip/hip_runtime.h>

namespace mi355 {

// LDS atomic adds: each wave issues `iters` instructions at scattered
// addresses (a multiplicative hash per lane), 32 KiB of LDS per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void lds_atomic_probe_kernel(T* __restrict__ out, int iters) {
    constexpr int N = 32768 / (int)sizeof(T);
    __shared__ T t[N];
    for (int i = threadIdx.x; i < N; i += blockDim.x) t[i] = (T)0;
    __syncthreads();
    unsigned h = threadIdx.x * 2654435761u + blockIdx.x * 40503u;
    for (int it = 0; it < iters; ++it) {
        h = h * 1664525u + 1013904223u;
        atomicAdd(&t[(h >> 8) & (N - 1)], (T)1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += blockDim.x) out[(size_t)blockIdx.x * N + i] = t[i];
}

// Global float atomic adds without return, coalesced over n elements.
__global__ __launch_bounds__(256) static void global_atomic_probe_kernel(float* __restrict__ p, long long n,
                                                                         int reps) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (int r = 0; r < reps; ++r)
        for (long long i = tid; i < n; i += stride) unsafeAtomicAdd(p + i, 1.0f);
}

}  // namespace mi355
