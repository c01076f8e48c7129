// Bloom index for the peeling algorithms.
//
// Every butterfly is counted exactly once by the priority counting of
// bt_count.hip, at its highest-priority vertex r with far end w (the
// vertex opposite r). For a fixed (r, w) the middles v (v < r, adjacent to
// both) with edges r-v and v-w form a "bloom" (a (2,k)-biclique); its
// butterflies are the pairs of middles. The counting pass emits one record
// (key (r, w), value (id(r-v), id(v-w))) per wedge whose far end has at
// least two middles; sorting the records by key makes each bloom a
// contiguous run. Each edge then gets the list of blooms it belongs to.
//
// Arrays (all u32 unless noted):
//   vals1[R] (u64)   sorted middles: id(r-v) << 32 | id(v-w)
//   bstart[NB+1]     first record of each bloom (sentinel R)
//   bk[R]            at a bloom's first record s: number of middles |
//                    (r is an upper vertex) << 31; a bloom is named by s
//   bclaim[R + ...]  at s: claim stamp of the peel (0 here); then the
//                    peel's dummy stamps (kClaimDummiesPerBlock per block)
//   bgraph[R]        batches only, at s: the graph of the bloom
//   inc_off[m+1] (u64) per-edge offsets into inc
//   inc[2R]          first records of the blooms holding the edge (or, for
//                    the h-index rounds, the records holding it)
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

namespace {

__global__ void k_heads(uint64_t R, const uint64_t* __restrict__ keys, uint32_t* __restrict__ flag) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < R; i += uint64_t(gridDim.x) * blockDim.x)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_bloom_start(uint64_t R, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bid,
                              uint32_t* __restrict__ bstart) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < R; i += uint64_t(gridDim.x) * blockDim.x)
        if (i == 0 || keys[i] != keys[i - 1]) bstart[bid[i] - 1] = uint32_t(i);
}

__global__ void k_bloom_size(uint32_t nbloom, uint32_t R, int nb, const uint64_t* __restrict__ keys,
                             const uint8_t* __restrict__ r_up, uint32_t* __restrict__ bstart,
                             uint32_t* __restrict__ bk, uint32_t* __restrict__ bclaim,
                             const uint64_t* __restrict__ rec, const uint32_t* __restrict__ egraph,
                             uint32_t* __restrict__ bgraph, uint32_t* __restrict__ kmax) {
    uint32_t const b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) bstart[nbloom] = R;
    uint32_t size = 0;
    if (b < nbloom) {
        uint32_t const s = bstart[b];
        uint32_t const e = (b + 1 < nbloom) ? bstart[b + 1] : R;
        uint32_t const r = uint32_t(keys[s] >> nb);
        size = e - s;
        bk[s] = size | (uint32_t(r_up[r]) << 31);
        bclaim[s] = 0;
        if (egraph) bgraph[s] = egraph[uint32_t(rec[s] >> 32)];  // batches: the bloom's graph
    }
    // largest bloom: one atomic per wave
    uint64_t const mx = wave_max_u64(size);
    if ((threadIdx.x & (kWave - 1)) == 0 && mx) atomicMax(kmax, uint32_t(mx));
}

__global__ void k_inc_count(uint64_t R, const uint64_t* __restrict__ rec, uint32_t* __restrict__ cnt) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < R; i += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t const v = rec[i];
        atomicAdd(&cnt[uint32_t(v >> 32)], 1u);
        atomicAdd(&cnt[uint32_t(v)], 1u);
    }
}

__global__ void k_inc_fill(uint64_t R, const uint64_t* __restrict__ rec, const uint32_t* __restrict__ bid,
                           const uint32_t* __restrict__ bstart, const uint64_t* __restrict__ inc_off,
                           uint32_t* __restrict__ cur, uint32_t* __restrict__ inc, bool records) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < R; i += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t const v = rec[i];
        uint32_t const b = records ? uint32_t(i) : bstart[bid[i] - 1];
        uint32_t const e1 = uint32_t(v >> 32), e2 = uint32_t(v);
        inc[inc_off[e1] + atomicAdd(&cur[e1], 1u)] = b;
        inc[inc_off[e2] + atomicAdd(&cur[e2], 1u)] = b;
    }
}

int bits_of(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) b++;
    return b;
}

}  // namespace

int bt_index_build(bt_dgraph* dg, bool inc_records) {
    bt_ctx* c = dg->ctx;
    bt_index& ix = c->ix;
    hipStream_t s = c->stream;
    uint32_t const m = dg->m;
    uint64_t const R = ix.R;
    if (R > kMaxIndexRecords) return BT_ERR_OVERFLOW;
    constexpr unsigned T = 256;
    unsigned const grid = grid_for(R ? R : 1, T, unsigned(c->cu_count) * 16);
    BT_TRY(bt_buf_reserve(ix.inc_off, sizeof(uint64_t) * (size_t(m) + 1)));
    BT_TRY(bt_buf_reserve(ix.inc_cur, sizeof(uint32_t) * (size_t(m) + 1)));
    uint64_t* inc_off = static_cast<uint64_t*>(ix.inc_off.ptr);
    uint32_t* cur = static_cast<uint32_t*>(ix.inc_cur.ptr);
    if (R == 0) {
        ix.NB = 0;
        BT_HIP_TRY(hipMemsetAsync(dg->ctl + kCtlBloomMax, 0, sizeof(uint64_t), s));
        BT_HIP_TRY(hipMemsetAsync(inc_off, 0, sizeof(uint64_t) * (size_t(m) + 1), s));
        BT_TRY(bt_buf_reserve(ix.bstart, 64));
        BT_TRY(bt_buf_reserve(ix.bk, 64));
        BT_TRY(bt_buf_reserve(ix.bclaim, sizeof(uint32_t) * size_t(c->cu_count) * kClaimDummiesPerBlock));
        BT_TRY(bt_buf_reserve(ix.touched, 64));
        BT_TRY(bt_buf_reserve(ix.inc, 64));
        BT_TRY(bt_buf_reserve(ix.vals1, 64));
        return BT_OK;
    }
    BT_TRY(bt_buf_reserve(ix.keys1, sizeof(uint64_t) * R));
    BT_TRY(bt_buf_reserve(ix.vals1, sizeof(uint64_t) * R));
    BT_TRY(bt_buf_reserve(ix.bid, sizeof(uint32_t) * R));
    auto* k0 = static_cast<uint64_t*>(ix.keys0.ptr);
    auto* k1 = static_cast<uint64_t*>(ix.keys1.ptr);
    auto* v0 = static_cast<uint64_t*>(ix.vals0.ptr);
    auto* v1 = static_cast<uint64_t*>(ix.vals1.ptr);
    auto* bid = static_cast<uint32_t*>(ix.bid.ptr);
    auto* flag = reinterpret_cast<uint32_t*>(k0);  // keys0 is free after the sort
    int const nb = bits_of(dg->n);
    size_t t1 = 0, t2 = 0, t3 = 0;
    BT_HIP_TRY(rocprim::radix_sort_pairs(nullptr, t1, k0, k1, v0, v1, size_t(R), 0, 2 * nb, s));
    BT_HIP_TRY(rocprim::inclusive_scan(nullptr, t2, flag, bid, size_t(R), rocprim::plus<uint32_t>(), s));
    BT_HIP_TRY(rocprim::exclusive_scan(nullptr, t3, cur, inc_off, uint64_t(0), size_t(m) + 1, rocprim::plus<uint64_t>(),
                                       s));
    BT_TRY(bt_buf_reserve(c->sort_tmp, std::max(t1, std::max(t2, t3))));
    BT_HIP_TRY(rocprim::radix_sort_pairs(c->sort_tmp.ptr, t1, k0, k1, v0, v1, size_t(R), 0, 2 * nb, s));
    k_heads<<<grid, T, 0, s>>>(R, k1, flag);
    BT_HIP_TRY(rocprim::inclusive_scan(c->sort_tmp.ptr, t2, flag, bid, size_t(R), rocprim::plus<uint32_t>(), s));
    uint32_t nbloom = 0;
    BT_HIP_TRY(hipMemcpyAsync(&nbloom, bid + (R - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    BT_HIP_TRY(hipStreamSynchronize(s));
    ix.NB = nbloom;
    BT_TRY(bt_buf_reserve(ix.bstart, sizeof(uint32_t) * (size_t(nbloom) + 1)));
    BT_TRY(bt_buf_reserve(ix.bk, sizeof(uint32_t) * R));
    BT_TRY(bt_buf_reserve(ix.bclaim, sizeof(uint32_t) * (R + size_t(c->cu_count) * kClaimDummiesPerBlock)));
    BT_TRY(bt_buf_reserve(ix.touched, sizeof(uint32_t) * 3 * std::max<size_t>(nbloom, 1)));
    BT_TRY(bt_buf_reserve(ix.inc, sizeof(uint32_t) * 2 * R));
    auto* bstart = static_cast<uint32_t*>(ix.bstart.ptr);
    auto* bk = static_cast<uint32_t*>(ix.bk.ptr);
    auto* bclaim = static_cast<uint32_t*>(ix.bclaim.ptr);
    auto* inc = static_cast<uint32_t*>(ix.inc.ptr);
    k_bloom_start<<<grid, T, 0, s>>>(R, k1, bid, bstart);
    uint32_t* bgraph = nullptr;
    if (dg->egraph) {
        BT_TRY(bt_buf_reserve(ix.bgraph, sizeof(uint32_t) * R));
        bgraph = static_cast<uint32_t*>(ix.bgraph.ptr);
    }
    uint32_t* kmax = reinterpret_cast<uint32_t*>(dg->ctl + kCtlBloomMax);
    BT_HIP_TRY(hipMemsetAsync(kmax, 0, sizeof(uint32_t), s));
    k_bloom_size<<<grid_for(nbloom, T, 1u << 30), T, 0, s>>>(nbloom, uint32_t(R), nb, k1, dg->r_up, bstart, bk, bclaim,
                                                              v1, dg->egraph, bgraph, kmax);
    // edge -> blooms incidence lists
    BT_HIP_TRY(hipMemsetAsync(cur, 0, sizeof(uint32_t) * (size_t(m) + 1), s));
    k_inc_count<<<grid, T, 0, s>>>(R, v1, cur);
    BT_HIP_TRY(rocprim::exclusive_scan(c->sort_tmp.ptr, t3, cur, inc_off, uint64_t(0), size_t(m) + 1,
                                       rocprim::plus<uint64_t>(), s));
    BT_HIP_TRY(hipMemsetAsync(cur, 0, sizeof(uint32_t) * (size_t(m) + 1), s));
    k_inc_fill<<<grid, T, 0, s>>>(R, v1, bid, bstart, inc_off, cur, inc, inc_records);
    BT_HIP_TRY(hipGetLastError());
    return BT_OK;
}
