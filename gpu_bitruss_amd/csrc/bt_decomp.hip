// Bitruss decomposition on gfx950: GBiD multi-stage peeling (msp) and the
// h-index iteration (hidx).
//
// Reference semantics:
//   msp  CPU c_bitruss_msp (cpu_bitruss.cpp:536-597) == GPU bitruss_msp
//        (bitruss_msp.cu:440-583): supports from ebfc; support-0 edges are
//        final; for each level the frontier is {live e : support == level};
//        a sub-round peels the frontier at once with the per-pivot
//        attribution rule (cpu_bitruss.cpp:301-351) and the clamp of
//        update_edge_support (:275-297); edges that fall to the level form
//        the next frontier; the result is the final support.
//   hidx CPU c_bitruss_hindex (cpu_bitruss.cpp:125-239) == GPU
//        g_bitruss_hindex (bitruss_hidx.cu:51-256): Jacobi rounds of the
//        h-index of {min of the other three supports} over the butterflies
//        of each edge whose three other edges are non-zero, until nothing
//        changes.
//
// Decrement rule on the device: an edge support only decreases during a
// sub-round (atomic subtract, no add-back); the one decrement that takes it
// from level+1 down to level appends it to `next`; the advance step then
// sets every `next` edge to exactly `level`. The net effect is the
// reference's s' = max(level, s - d), with the same `next` set.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

namespace {

constexpr int kPeelBlock = 256;
constexpr uint32_t kUCap = 4096;  // N(u) staged in LDS up to this degree

enum : uint8_t { kLive = 0, kCurr = 1, kDone = 2 };

// control words in dg->ctl
enum {
    C_ALIVE = 0,
    C_MIN = 1,
    C_CURR = 2,
    C_NEXT = 3,
    C_CHANGED = 4,
};

struct Csr {
    const uint32_t* __restrict__ off;
    const uint32_t* __restrict__ nbr;
    const uint32_t* __restrict__ eids;
    const uint32_t* __restrict__ edges;
};

struct PeelShared {
    uint32_t un[kUCap];
    uint32_t ue[kUCap];
    uint32_t start[kPeelBlock];
    uint32_t wpos[kPeelBlock];
    uint32_t wv[kPeelBlock];
    uint32_t part[kPeelBlock / kWave + 1];
};

// Enumerates, for pivot (u, v), every butterfly u-v-w-x (w in N(v)\{u},
// x in N(u) & N(w) \ {v}) whose wedge passes wedge_ok(w, wv), calling
// visit(wv, wx, ux). N(u) is staged in LDS when it fits; the wedges of
// N(v) are expanded over the block in chunks, each thread scanning entries
// of N(w) and probing N(u) by binary search.
template <class WedgeOk, class Visit>
__device__ __forceinline__ void for_each_butterfly(const Csr& g, uint32_t u, uint32_t v, PeelShared& sh,
                                                   WedgeOk wedge_ok, Visit visit) {
    int const tid = threadIdx.x;
    uint32_t const ub = g.off[u], ulen = g.off[u + 1] - ub;
    uint32_t const vb = g.off[v], vlen = g.off[v + 1] - vb;
    bool const staged = ulen <= kUCap;
    if (staged)
        for (uint32_t i = tid; i < ulen; i += kPeelBlock) {
            sh.un[i] = g.nbr[ub + i];
            sh.ue[i] = g.eids[ub + i];
        }
    const uint32_t* un = staged ? sh.un : g.nbr + ub;
    __syncthreads();
    for (uint32_t base = 0; base < vlen; base += kPeelBlock) {
        uint32_t const idx = base + tid;
        uint32_t w = 0, wv = 0, wlen = 0, wb = 0;
        if (idx < vlen) {
            w = g.nbr[vb + idx];
            wv = g.eids[vb + idx];
            if (w != u && wedge_ok(w, wv)) {
                wb = g.off[w];
                wlen = g.off[w + 1] - wb;
            }
        }
        uint32_t chunk = 0;
        uint32_t const excl = block_excl_scan<kPeelBlock>(wlen, sh.part, chunk);
        sh.start[tid] = excl;
        sh.wpos[tid] = wb;
        sh.wv[tid] = wv;
        __syncthreads();
        for (uint32_t t = tid; t < chunk; t += kPeelBlock) {
            int const j = upper_index(sh.start, kPeelBlock, t);
            uint32_t const pos = sh.wpos[j] + (t - sh.start[j]);
            uint32_t const x = g.nbr[pos];
            if (x == v) continue;
            uint32_t const k = lower_bound_u32(un, ulen, x);
            if (k == ulen || un[k] != x) continue;
            uint32_t const ux = staged ? sh.ue[k] : g.eids[ub + k];
            visit(sh.wv[j], g.eids[pos], ux);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------- msp ---

__global__ void k_msp_init(uint32_t m, const uint64_t* __restrict__ sup, uint8_t* __restrict__ state,
                           uint64_t* __restrict__ ctl) {
    uint64_t live = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        bool const z = sup[e] == 0;  // process_support_zero (cpu_bitruss.cpp:241-256)
        state[e] = z ? kDone : kLive;
        live += z ? 0 : 1;
    }
    live = wave_sum_u64(live);
    if ((threadIdx.x & (kWave - 1)) == 0 && live) atomic_add_u64(&ctl[C_ALIVE], live);
}

__global__ void k_msp_min(uint32_t m, const uint64_t* __restrict__ sup, const uint8_t* __restrict__ state,
                          uint64_t* __restrict__ ctl) {
    uint64_t mn = ~0ull;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
        if (state[e] == kLive && sup[e] < mn) mn = sup[e];
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        uint64_t const o = __shfl_xor(mn, d, kWave);
        mn = o < mn ? o : mn;
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && mn != ~0ull) atomic_min_u64(&ctl[C_MIN], mn);
}

// scan_edge (cpu_bitruss.cpp:258-272) at `level`
__global__ void k_msp_collect(uint32_t m, uint64_t level, const uint64_t* __restrict__ sup,
                              uint8_t* __restrict__ state, uint32_t* __restrict__ curr,
                              uint64_t* __restrict__ ctl) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        if (state[e] == kLive && sup[e] == level) {
            state[e] = kCurr;
            curr[atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[C_CURR]), 1ull)] = e;
        }
    }
}

struct Decrementer {
    uint64_t* sup;
    uint32_t* next;
    uint64_t* ctl;
    uint64_t level;
    __device__ __forceinline__ void operator()(uint32_t e) const {
        uint64_t const old = atomic_sub_u64(&sup[e], 1);
        if (old == level + 1)
            next[atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[C_NEXT]), 1ull)] = e;
    }
};

// One sub-round of peel_edges (cpu_bitruss.cpp:353-424 /
// process_sub_level_kernel, bitruss_msp.cu:187-274); one workgroup per pivot.
__global__ __launch_bounds__(kPeelBlock) void k_msp_peel(const uint32_t* __restrict__ curr, uint32_t curr_len,
                                                         uint64_t level, Csr g, const uint8_t* __restrict__ state,
                                                         uint64_t* __restrict__ sup, uint32_t* __restrict__ next,
                                                         uint64_t* __restrict__ ctl) {
    __shared__ PeelShared sh;
    Decrementer dec{sup, next, ctl, level};
    for (uint32_t ci = blockIdx.x; ci < curr_len; ci += gridDim.x) {
        uint32_t const uv = curr[ci];
        uint32_t const u = g.edges[2ull * uv], v = g.edges[2ull * uv + 1];
        auto wedge_ok = [&](uint32_t, uint32_t wv) {
            uint8_t const s = state[wv];
            // processed wedges are skipped; a wedge whose v-w edge is in the
            // frontier with a smaller id is left to that pivot (its rule
            // never fires from this pivot, bitruss_msp.cu:229)
            return s != kDone && !(s == kCurr && uv > wv);
        };
        auto visit = [&](uint32_t wv, uint32_t wx, uint32_t ux) {
            uint8_t const swx = state[wx], sux = state[ux];
            if (swx == kDone || sux == kDone) return;
            int const pwv = state[wv] != kCurr, pwx = swx != kCurr, pux = sux != kCurr;
            int const cnt = pwv + pwx + pux;
            // ppeel_one_edge / ppeel_two_edge / ppeel_three_edge
            // (cpu_bitruss.cpp:301-351)
            if (cnt == 1) {
                if (pwx && uv < ux && ux < wv) dec(wx);
                if (pwv && uv < ux && ux < wx) dec(wv);
                if (pux && uv < wv && wv < wx) dec(ux);
            } else if (cnt == 2) {
                if (pwv && pwx && uv < ux) { dec(wv); dec(wx); }
                if (pwv && pux && uv < wx) { dec(wv); dec(ux); }
                if (pwx && pux && uv < wv) { dec(wx); dec(ux); }
            } else if (cnt == 3) {
                dec(wv);
                dec(wx);
                dec(ux);
            }
        };
        for_each_butterfly(g, u, v, sh, wedge_ok, visit);
    }
}

// Commit the frontier (update_processed_edges, bitruss_msp.cu:277-286) and
// promote `next` to the new frontier at exactly `level`.
__global__ void k_msp_advance(const uint32_t* __restrict__ curr, uint32_t curr_len,
                              const uint32_t* __restrict__ next, uint32_t next_len, uint64_t level,
                              uint8_t* __restrict__ state, uint64_t* __restrict__ sup) {
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < curr_len; i += stride) state[curr[i]] = kDone;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < next_len; i += stride) {
        uint32_t const e = next[i];
        state[e] = kCurr;
        sup[e] = level;
    }
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------ hidx ---

constexpr int kHBins = 1024;

struct HidxShared {
    PeelShared peel;
    uint32_t hist[kHBins];
    uint64_t lo, hi, width, above;
    int done;
    uint64_t h;
};

// One Jacobi round: curr[e] = h-index of {min(prev[wv], prev[wx], prev[ux])}
// over e's butterflies whose three other edges are non-zero
// (compute_bitruss_hindex, cpu_bitruss.cpp:125-194). The h-index never
// exceeds prev[e] (the iteration is non-increasing), so values are clipped
// to prev[e] and binned into kHBins bins; when a bin is wider than one value
// the enumeration is repeated inside the bin that holds the answer.
__global__ __launch_bounds__(kPeelBlock) void k_hidx_round(uint32_t m, Csr g, const uint64_t* __restrict__ prev,
                                                           uint64_t* __restrict__ curr,
                                                           uint64_t* __restrict__ ctl,
                                                           const uint32_t* __restrict__ egraph,
                                                           uint32_t* __restrict__ glast, uint32_t round) {
    __shared__ HidxShared sh;
    uint32_t last_g = ~0u;  // thread 0: graph whose change this block last recorded
    for (uint32_t e = blockIdx.x; e < m; e += gridDim.x) {
        uint64_t const pe = prev[e];
        if (pe == 0) {  // cpu_bitruss.cpp:128
            if (threadIdx.x == 0) curr[e] = 0;
            continue;
        }
        uint32_t const u = g.edges[2ull * e], v = g.edges[2ull * e + 1];
        if (threadIdx.x == 0) {
            sh.lo = 0;
            sh.hi = pe;
            sh.above = 0;
            sh.done = 0;
            sh.width = (pe + kHBins) / kHBins;  // ceil((pe + 1) / kHBins)
        }
        __syncthreads();
        while (true) {
            uint64_t const lo = sh.lo, hi = sh.hi, width = sh.width;
            for (int i = threadIdx.x; i < kHBins; i += kPeelBlock) sh.hist[i] = 0;
            __syncthreads();
            auto wedge_ok = [&](uint32_t, uint32_t wv) { return prev[wv] != 0; };
            auto visit = [&](uint32_t wv, uint32_t wx, uint32_t ux) {
                uint64_t const a = prev[wx], b = prev[ux];
                if (a == 0 || b == 0) return;
                uint64_t val = prev[wv];
                val = a < val ? a : val;
                val = b < val ? b : val;
                // clipped to prev[e] (a bound on the answer) in every pass,
                // so a value above prev[e] counts at prev[e] in the top bin
                if (val > pe) val = pe;
                if (val > hi) return;  // already counted in `above`
                if (val < lo) return;
                atomicAdd(&sh.hist[(val - lo) / width], 1u);
            };
            for_each_butterfly(g, u, v, sh.peel, wedge_ok, visit);
            __syncthreads();
            if (threadIdx.x == 0) {
                // largest bin b whose lower end a_b has #{vals >= a_b} >= a_b
                uint64_t const nb = (hi - lo) / width + 1;
                uint64_t suffix = sh.above;
                uint64_t b = 0;
                for (uint64_t i = nb; i-- > 0;) {
                    uint64_t const above_i = suffix;
                    suffix += sh.hist[i];
                    if (suffix >= lo + i * width) {
                        b = i;
                        sh.above = above_i;
                        break;
                    }
                }
                uint64_t const a = lo + b * width;
                if (width == 1) {
                    sh.h = a;
                    sh.done = 1;
                } else {
                    sh.lo = a;
                    uint64_t const top = a + width - 1;
                    sh.hi = top < hi ? top : hi;
                    uint64_t const span = sh.hi - sh.lo + 1;
                    sh.width = (span + kHBins - 1) / kHBins;
                }
            }
            __syncthreads();
            if (sh.done) break;
        }
        if (threadIdx.x == 0) {
            curr[e] = sh.h;
            if (sh.h != pe) {
                ctl[C_CHANGED] = 1;
                if (egraph && egraph[e] != last_g) {  // batches: per-graph round count
                    last_g = egraph[e];
                    __hip_atomic_store(&glast[last_g], round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

// msp entry: supports (and, for the index path, bloom records) first, then
// the peel. The index path needs cooperative launch and room for the index
// (about 44 bytes per priority wedge); otherwise the merge path runs.
// rule 1 (mspp) and 2 (the textbook peel) run on the index path only.
int bt_msp_impl(bt_dgraph* dg, uint32_t* levels_out, int rule) {
    bt_ctx* c = dg->ctx;
    bool const want_index = c->coop && (c->msp_path != BT_MSP_PATH_MERGE || rule != 0);
    BT_TRY(bt_count_impl(dg, true, nullptr, nullptr, want_index));
    if (want_index && c->ix.R != kNoIndex) {
        int st = bt_msp_index_peel(dg, levels_out, rule);
        if (st != BT_ERR_BUSY) return st;
        // the peel's share of the CUs was held by other work (its grid
        // barrier timed out): once more, from the supports, as one
        // cooperative launch on the whole device
        int const share = c->peel_blocks;
        c->peel_blocks = 0;
        st = bt_count_impl(dg, true, nullptr, nullptr, true);
        if (st == BT_OK) st = c->ix.R != kNoIndex ? bt_msp_index_peel(dg, levels_out, rule) : BT_ERR_BUDGET;
        c->peel_blocks = share;
        return st == BT_ERR_INTERNAL ? BT_ERR_BUSY : st;
    }
    if (c->msp_path == BT_MSP_PATH_INDEX || rule != 0) return BT_ERR_BUDGET;
    return bt_msp_merge_peel(dg, levels_out);
}

// Merge-path peel: one workgroup per frontier edge per sub-round, the host
// steps the levels and sub-rounds (process_sub_level, bitruss_msp.cu:290-309).
int bt_msp_merge_peel(bt_dgraph* dg, uint32_t* levels_out) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const m = dg->m;
    bt_timings const count_t = c->last;
    bt_timer tm(c);
    BT_TRY(tm.start(5));
    dg->has_support = false;  // dg->sup becomes the working support
    Csr g{dg->off, dg->nbr, dg->eids, dg->edges};
    constexpr unsigned T = 256;
    unsigned const scan_grid = grid_for(m, T, unsigned(c->cu_count) * 8);
    unsigned const peel_cap = unsigned(c->cu_count) * 8;
    BT_HIP_TRY(hipMemsetAsync(dg->ctl, 0, sizeof(uint64_t) * 8, s));
    uint32_t level_max = 0;
    uint64_t sub_rounds = 0, levels_scanned = 0;
    if (m) {
        k_msp_init<<<scan_grid, T, 0, s>>>(m, dg->sup, dg->state, dg->ctl);
        BT_HIP_TRY(hipGetLastError());
        uint64_t alive = 0;
        BT_TRY(bt_read_ctl(dg, C_ALIVE, 1, &alive));
        uint32_t* curr = dg->list0;
        uint32_t* next = dg->list1;
        while (alive > 0) {
            // next level: the smallest live support (the reference steps the
            // level by one and scans; empty levels change nothing)
            k_set_u64<<<1, 64, 0, s>>>(dg->ctl + C_MIN, ~0ull);
            k_msp_min<<<scan_grid, T, 0, s>>>(m, dg->sup, dg->state, dg->ctl);
            uint64_t level = 0;
            BT_TRY(bt_read_ctl(dg, C_MIN, 1, &level));
            if (level == ~0ull || level == 0) return BT_ERR_INTERNAL;
            levels_scanned++;
            k_set_u64<<<1, 64, 0, s>>>(dg->ctl + C_CURR, 0);
            k_msp_collect<<<scan_grid, T, 0, s>>>(m, level, dg->sup, dg->state, curr, dg->ctl);
            uint64_t curr_len = 0;
            BT_TRY(bt_read_ctl(dg, C_CURR, 1, &curr_len));
            if (curr_len == 0) return BT_ERR_INTERNAL;  // the minimum must be collected
            if (level > 0xFFFFFFFFull) return BT_ERR_OVERFLOW;
            level_max = uint32_t(level);
            while (curr_len > 0) {
                if (curr_len == alive) {  // bitruss_msp.cu:517-520
                    alive = 0;
                    break;
                }
                k_set_u64<<<1, 64, 0, s>>>(dg->ctl + C_NEXT, 0);
                k_msp_peel<<<grid_for(curr_len, 1, peel_cap), kPeelBlock, 0, s>>>(
                    curr, uint32_t(curr_len), level, g, dg->state, dg->sup, next, dg->ctl);
                BT_HIP_TRY(hipGetLastError());
                uint64_t next_len = 0;
                BT_TRY(bt_read_ctl(dg, C_NEXT, 1, &next_len));
                k_msp_advance<<<grid_for(curr_len + next_len, T, unsigned(c->cu_count) * 4), T, 0, s>>>(
                    curr, uint32_t(curr_len), next, uint32_t(next_len), level, dg->state, dg->sup);
                BT_HIP_TRY(hipGetLastError());
                sub_rounds++;
                alive -= curr_len;
                uint32_t* t = curr;
                curr = next;
                next = t;
                curr_len = next_len;
            }
        }
        BT_HIP_TRY(hipMemcpyAsync(dg->res, dg->sup, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, s));
    }
    BT_TRY(tm.start(6));
    BT_HIP_TRY(hipStreamSynchronize(s));
    c->last = count_t;
    c->last.peel_ms = tm.ms(5, 6);
    c->last.total_ms = count_t.total_ms + c->last.peel_ms;
    c->last.sub_rounds = sub_rounds;
    c->last.levels_scanned = levels_scanned;
    if (levels_out) *levels_out = level_max;
    return BT_OK;
}

int bt_hidx_impl(bt_dgraph* dg, uint32_t* rounds_out) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const m = dg->m;
    // the bloom index when it fits (the path option applies to hidx too)
    bool const want_index = c->msp_path != BT_MSP_PATH_MERGE;
    BT_TRY(bt_count_impl(dg, true, nullptr, nullptr, want_index));
    if (want_index && c->ix.R != kNoIndex) return bt_hidx_index(dg, rounds_out);
    if (c->msp_path == BT_MSP_PATH_INDEX) return BT_ERR_BUDGET;
    bt_timings const count_t = c->last;
    bt_timer tm(c);
    BT_TRY(tm.start(5));
    Csr g{dg->off, dg->nbr, dg->eids, dg->edges};
    // the reference runs one round even on an edgeless graph
    // (bitruss_hidx.cu:218-231)
    uint32_t rounds = m ? 0 : 1;
    if (m) {
        BT_HIP_TRY(hipMemcpyAsync(dg->res, dg->sup, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, s));
        uint64_t* prev = dg->res;
        uint64_t* curr = dg->aux;
        unsigned const grid = grid_for(m, 1, unsigned(c->cu_count) * 8);
        while (true) {
            rounds++;
            k_set_u64<<<1, 64, 0, s>>>(dg->ctl + C_CHANGED, 0);
            k_hidx_round<<<grid, kPeelBlock, 0, s>>>(m, g, prev, curr, dg->ctl, dg->egraph, dg->graph_last, rounds);
            BT_HIP_TRY(hipGetLastError());
            uint64_t changed = 0;
            BT_TRY(bt_read_ctl(dg, C_CHANGED, 1, &changed));
            uint64_t* t = prev;
            prev = curr;
            curr = t;
            if (!changed) break;
        }
        if (prev != dg->res)
            BT_HIP_TRY(hipMemcpyAsync(dg->res, prev, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, s));
    }
    BT_TRY(tm.start(6));
    BT_HIP_TRY(hipStreamSynchronize(s));
    c->last = count_t;
    c->last.peel_ms = tm.ms(5, 6);
    c->last.total_ms = count_t.total_ms + c->last.peel_ms;
    c->last.sub_rounds = rounds;
    if (rounds_out) *rounds_out = rounds;
    return BT_OK;
}
