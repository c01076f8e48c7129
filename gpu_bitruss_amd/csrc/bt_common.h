// Device helpers shared by the kernels (wave/block scans, searches).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bt {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t len, uint32_t key) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        uint32_t const mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Largest j in [0, len) with s[j] <= t, given s[0] <= t and s non-decreasing.
__device__ __forceinline__ int upper_index(const uint32_t* s, int len, uint32_t t) {
    int lo = 0, hi = len - 1;
    while (lo < hi) {
        int const mid = (lo + hi + 1) >> 1;
        if (s[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int const lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t const y = __shfl_up(v, d, kWave);
        if (lane >= d) v += y;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        uint64_t const o = __shfl_xor(v, d, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Exclusive scan over a block of BLOCK threads (BLOCK a multiple of 64).
// s_part needs BLOCK/64 + 1 slots. All threads must call it.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_part, uint32_t& total) {
    uint32_t const incl = wave_incl_scan(x);
    if constexpr (BLOCK == kWave) {
        total = __shfl(incl, kWave - 1, kWave);
        return incl - x;
    } else {
        int const lane = threadIdx.x & (kWave - 1);
        int const wid = threadIdx.x / kWave;
        if (lane == kWave - 1) s_part[wid] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = 0;
            for (int i = 0; i < BLOCK / kWave; i++) {
                uint32_t const t = s_part[i];
                s_part[i] = acc;
                acc += t;
            }
            s_part[BLOCK / kWave] = acc;
        }
        __syncthreads();
        total = s_part[BLOCK / kWave];
        uint32_t const r = incl - x + s_part[wid];
        __syncthreads();
        return r;
    }
}

// Sum of a u64 over the block; the result is valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* s_part) {
    v = wave_sum_u64(v);
    if constexpr (BLOCK == kWave) {
        return v;
    } else {
        int const lane = threadIdx.x & (kWave - 1);
        int const wid = threadIdx.x / kWave;
        if (lane == 0) s_part[wid] = v;
        __syncthreads();
        uint64_t r = 0;
        if (threadIdx.x == 0)
            for (int i = 0; i < BLOCK / kWave; i++) r += s_part[i];
        __syncthreads();
        return r;
    }
}

__device__ __forceinline__ uint64_t atomic_add_u64(uint64_t* p, uint64_t v) {
    return atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

__device__ __forceinline__ uint64_t atomic_sub_u64(uint64_t* p, uint64_t v) {
    return atomicAdd(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(0ull - v));
}

__device__ __forceinline__ void atomic_max_u64(uint64_t* p, uint64_t v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

__device__ __forceinline__ void atomic_min_u64(uint64_t* p, uint64_t v) {
    atomicMin(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

inline unsigned grid_for(uint64_t work, unsigned block, unsigned cap) {
    uint64_t g = (work + block - 1) / block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

}  // namespace bt
