// Butterfly counting (bfc) and per-edge support (ebfc) for gfx950.
//
// Semantics: the reference's vertex-priority counting (CPU cpu_bfc.cpp:24-203,
// GPU bfc_vpp.cu:33-221): every butterfly is counted once at its vertex of
// highest priority u (priority = (degree, id)), as the pairs of wedges
// u-v-w with v, w below u that share the far end w; with cnt[w] such wedges
// a root contributes C(cnt[w],2) butterflies, and each of its wedges adds
// cnt[w]-1 to the supports of u-v and v-w.
//
// Layout (MI355X design, not the reference's): the CSR is relabelled by
// priority rank once per call, each row sorted by rank, so for root r the
// lower-priority neighbours are a row prefix (r_lo[r] entries) and for each
// such v the valid far ends are the prefix of row v below r (r_p entries).
// Roots are bucketed by their wedge count W_r (bounds and workgroup sizes
// per graph-size class from the context's tuning table, bt_tune.hip):
//   A  W_r <= wave_wedges  (256..2048): one 64-lane wavefront per root,
//                          packed LDS hash of 2 * wave_wedges slots
//   B  W_r <= block_wedges (4096..16384): one workgroup of block_threads
//                          per root, LDS hash of 2 * block_wedges slots
//   C  the rest: hub roots; chunks of hub_threads row entries are spread
//                over the whole grid and count into dense per-root tables
//                in HBM
// An LDS hash slot packs (w+1) in the high bits and the count in the low
// `cb` bits; roots with r_lo >= 2^cb go to bucket C. A hash never holds
// more keys than half its slots (distinct far ends <= wedges of the root).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

namespace {


// ------------------------------------------------------------------ prep ---

__global__ void k_rank_keys(uint32_t n, const uint32_t* __restrict__ deg, uint64_t* __restrict__ keys) {
    uint32_t const x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) keys[x] = (uint64_t(deg[x]) << 32) | x;
}

__global__ void k_rank_scatter(uint32_t n, uint32_t u_num, const uint64_t* __restrict__ sorted,
                               uint32_t* __restrict__ rank, uint32_t* __restrict__ r_deg,
                               uint8_t* __restrict__ r_up) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint64_t const k = sorted[i];
        uint32_t const x = uint32_t(k);
        rank[x] = i;
        r_deg[i] = uint32_t(k >> 32);
        r_up[i] = x < u_num ? 1 : 0;
    }
}

__global__ void k_set_last(uint32_t* __restrict__ a, uint32_t idx, uint32_t value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) a[idx] = value;
}

__global__ void k_pair_keys(uint32_t m, const uint32_t* __restrict__ edges, const uint32_t* __restrict__ rank,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    uint64_t const ru = rank[edges[2ull * e]];
    uint64_t const rv = rank[edges[2ull * e + 1]];
    keys[2ull * e] = (ru << 32) | rv;
    keys[2ull * e + 1] = (rv << 32) | ru;
    vals[2ull * e] = e;
    vals[2ull * e + 1] = e;
}

__global__ void k_pair_unpack(uint32_t cnt, const uint64_t* __restrict__ keys, uint32_t* __restrict__ r_nbr) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) r_nbr[i] = uint32_t(keys[i]);
}

__global__ void k_lo_deg(uint32_t n, const uint32_t* __restrict__ r_off, const uint32_t* __restrict__ r_nbr,
                         uint32_t* __restrict__ r_lo) {
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t const b = r_off[r];
    r_lo[r] = lower_bound_u32(r_nbr + b, r_off[r + 1] - b, r);
}

// r_p[i] for entry i = (r, v) of row r: #entries of row v below r, when v < r.
__global__ void k_prefix_p(uint32_t cnt, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ r_off,
                           const uint32_t* __restrict__ r_nbr, uint32_t* __restrict__ r_p) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    uint32_t const r = uint32_t(keys[i] >> 32);
    uint32_t const v = r_nbr[i];
    uint32_t p = 0;
    if (v < r) {
        uint32_t const b = r_off[v];
        p = lower_bound_u32(r_nbr + b, r_off[v + 1] - b, r);
    }
    r_p[i] = p;
}

// One wavefront per root: W_r = sum of r_p over the row prefix.
__global__ void k_root_w(uint32_t n, const uint32_t* __restrict__ r_off, const uint32_t* __restrict__ r_lo,
                         const uint32_t* __restrict__ r_p, uint64_t* __restrict__ r_w, uint64_t* __restrict__ ctl) {
    uint32_t const r = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    int const lane = threadIdx.x & (kWave - 1);
    if (r >= n) return;
    uint32_t const b = r_off[r], lo = r_lo[r];
    uint64_t s = 0;
    for (uint32_t i = lane; i < lo; i += kWave) s += r_p[b + i];
    s = wave_sum_u64(s);
    if (lane == 0) {
        r_w[r] = s;
        if (s) atomic_add_u64(&ctl[3], s);
    }
}

__global__ void k_classify(uint32_t n, const uint64_t* __restrict__ r_w, const uint32_t* __restrict__ r_lo,
                           uint32_t lo_limit, uint32_t wa, uint32_t wb, uint32_t* __restrict__ la,
                           uint32_t* __restrict__ lb, uint32_t* __restrict__ lc, uint64_t* __restrict__ ctl) {
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t const w = r_w[r];
    if (w == 0) return;
    uint32_t const lo = r_lo[r];
    if (lo < lo_limit && w <= wa) {
        la[atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[0]), 1ull)] = r;
    } else if (lo < lo_limit && w <= wb) {
        lb[atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[1]), 1ull)] = r;
    } else {
        lc[atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[2]), 1ull)] = r;
    }
}

// --------------------------------------------------------------- counting ---

// Packed LDS hash: slot = ((w+1) << cb) | count, 0 = empty.
struct LdsHash {
    uint32_t* tab;
    uint32_t mask;   // capacity - 1
    int shift;       // 32 - log2(capacity)
    int cb;
    __device__ __forceinline__ uint32_t slot0(uint32_t w) const { return (w * 2654435761u) >> shift; }
    // increments the count of w, returns the count before the increment
    // The table holds at most half its capacity of keys (distinct far ends
    // <= wedges of the root <= capacity / 2), so probing ends; the probe cap
    // only guards against a broken invariant.
    __device__ __forceinline__ uint32_t inc(uint32_t w) const {
        uint32_t const key = w + 1;
        uint32_t s = slot0(w);
        for (uint32_t probe = 0; probe <= mask; probe++) {
            uint32_t cur = tab[s];
            if (cur == 0) {
                uint32_t const prev = atomicCAS(&tab[s], 0u, (key << cb) | 1u);
                if (prev == 0) return 0;
                cur = prev;
            }
            if ((cur >> cb) == key) {
                uint32_t const old = atomicAdd(&tab[s], 1u);
                return old & ((1u << cb) - 1u);
            }
            s = (s + 1) & mask;
        }
        return 0;
    }
    __device__ __forceinline__ uint32_t get(uint32_t w) const {
        uint32_t const key = w + 1;
        uint32_t s = slot0(w);
        for (uint32_t probe = 0; probe <= mask; probe++) {
            uint32_t const cur = tab[s];
            if ((cur >> cb) == key) return cur & ((1u << cb) - 1u);
            if (cur == 0) return 0;
            s = (s + 1) & mask;
        }
        return 0;
    }
};

struct DenseTable {
    uint32_t* tab;
    __device__ __forceinline__ uint32_t inc(uint32_t w) const { return atomicAdd(&tab[w], 1u); }
    __device__ __forceinline__ uint32_t get(uint32_t w) const { return tab[w]; }
};

struct RowView {
    const uint32_t* __restrict__ r_off;
    const uint32_t* __restrict__ r_nbr;
    const uint32_t* __restrict__ r_eid;
    const uint32_t* __restrict__ r_p;
};

// Optional output of the wedges that belong to blooms (far end w with
// cnt >= 2): one record per wedge, key = (r << nb) | w, value =
// (id(r-v) << 32) | id(v-w). Sorting the records by key groups the blooms.
struct Emit {
    uint64_t* keys = nullptr;
    uint64_t* vals = nullptr;
    unsigned long long* count = nullptr;
    int nb = 0;
};

// Per-block scratch of expand_entries.
template <int BLOCK>
struct ExpandShared {
    uint32_t start[BLOCK];
    uint32_t pos[BLOCK];
    uint64_t acc[BLOCK];
    uint32_t part[BLOCK / kWave + 1];
    uint64_t red[BLOCK / kWave + 1];
    unsigned long long base;
};

// Expands the wedges of entries [i0, i1) of root row r in chunks of BLOCK
// entries. PASS 1 counts far ends into the table and returns the sum of the
// old counts (the butterflies); PASS 2 adds cnt-1 to the supports of v-w
// (one atomic per wedge) and u-v (aggregated per entry), and emits the
// bloom wedge records when em.keys is set.
template <int BLOCK, int PASS, class Table>
__device__ __forceinline__ uint64_t expand_entries(const RowView& g, uint32_t r, uint32_t i0, uint32_t i1,
                                                   const Table& tab, uint64_t* __restrict__ sup, const Emit& em,
                                                   ExpandShared<BLOCK>& sh) {
    uint64_t tot = 0;
    int const tid = threadIdx.x;
    for (uint32_t base = i0; base < i1; base += BLOCK) {
        uint32_t const idx = base + tid;
        bool const valid = idx < i1;
        uint32_t const p = valid ? g.r_p[idx] : 0;
        uint32_t const v = valid ? g.r_nbr[idx] : 0;
        uint32_t chunk = 0;
        uint32_t const excl = block_excl_scan<BLOCK>(p, sh.part, chunk);
        sh.start[tid] = excl;
        sh.pos[tid] = valid ? g.r_off[v] : 0;
        if (PASS == 2) sh.acc[tid] = 0;
        __syncthreads();
        uint32_t my_emit = 0;
        for (uint32_t t = tid; t < chunk; t += BLOCK) {
            int const j = upper_index(sh.start, BLOCK, t);
            uint32_t const pos = sh.pos[j] + (t - sh.start[j]);
            uint32_t const w = g.r_nbr[pos];
            if (PASS == 1) {
                tot += tab.inc(w);
            } else {
                uint32_t const c = tab.get(w);
                if (c > 1) {
                    atomic_add_u64(&sup[g.r_eid[pos]], uint64_t(c - 1));
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sh.acc[j]), (unsigned long long)(c - 1));
                    my_emit++;
                }
            }
        }
        __syncthreads();
        if (PASS == 2) {
            if (valid && sh.acc[tid]) atomic_add_u64(&sup[g.r_eid[idx]], sh.acc[tid]);
            if (em.keys) {
                uint32_t total_emit = 0;
                uint32_t const ex = block_excl_scan<BLOCK>(my_emit, sh.part, total_emit);
                if (tid == 0) sh.base = total_emit ? atomicAdd(em.count, (unsigned long long)total_emit) : 0ull;
                __syncthreads();
                uint64_t at = sh.base + ex;
                for (uint32_t t = tid; t < chunk; t += BLOCK) {
                    int const j = upper_index(sh.start, BLOCK, t);
                    uint32_t const pos = sh.pos[j] + (t - sh.start[j]);
                    uint32_t const w = g.r_nbr[pos];
                    if (tab.get(w) > 1) {
                        em.keys[at] = (uint64_t(r) << em.nb) | w;
                        em.vals[at] = (uint64_t(g.r_eid[base + j]) << 32) | g.r_eid[pos];
                        at++;
                    }
                }
            }
            __syncthreads();
        }
    }
    return tot;
}

// Device-wide work queue: the workgroup's next run of `grab` units of
// [0, count) from the counter q (every workgroup takes runs until the
// counter passes count), so that uneven units spread over the workgroups as
// they finish. Returns the run's first unit (>= count: none left).
template <int BLOCK>
__device__ __forceinline__ uint32_t take_units(unsigned* q, uint32_t grab, uint32_t* s_take) {
    if (threadIdx.x == 0) *s_take = atomicAdd(q, grab);
    __syncthreads();
    uint32_t const first = *s_take;
    __syncthreads();
    return first;
}

// Buckets A and B: one root per workgroup iteration, table in LDS; roots
// come from a device-wide queue in runs of GRAB.
template <int BLOCK, uint32_t CAP>
__global__ __launch_bounds__(BLOCK) void k_count_lds(const uint32_t* __restrict__ roots, uint32_t nroots,
                                                     RowView g, const uint32_t* __restrict__ r_lo,
                                                     uint64_t* __restrict__ sup, uint64_t* __restrict__ total,
                                                     int cb, int per_edge, Emit em, unsigned* queue) {
    __shared__ uint32_t tab[CAP];
    __shared__ ExpandShared<BLOCK> sh;
    __shared__ uint32_t s_take;
    constexpr int kLog = __builtin_ctz(CAP);
    constexpr uint32_t GRAB = BLOCK == kWave ? 16u : 2u;  // wave roots are light
    LdsHash h{tab, CAP - 1, 32 - kLog, cb};
    uint64_t tot = 0;
    uint32_t ri = 0, rend = 0;
    for (;;) {
        if (ri == rend) {
            ri = take_units<BLOCK>(queue, GRAB, &s_take);
            if (ri >= nroots) break;
            rend = ri + GRAB < nroots ? ri + GRAB : nroots;
        }
        uint32_t const r = roots[ri++];
        uint32_t const b = g.r_off[r], lo = r_lo[r];
        for (uint32_t i = threadIdx.x; i < CAP; i += BLOCK) tab[i] = 0;
        __syncthreads();
        tot += expand_entries<BLOCK, 1>(g, r, b, b + lo, h, sup, em, sh);
        if (per_edge) expand_entries<BLOCK, 2>(g, r, b, b + lo, h, sup, em, sh);
        __syncthreads();
    }
    uint64_t const s = block_sum_u64<BLOCK>(tot, sh.red);
    if (threadIdx.x == 0 && s) atomic_add_u64(total, s);
}

// Bucket C: item = (root r, entry range [i0, i1), table offset); the dense
// table of r has r slots (far ends have rank < r).
struct HubItem {
    uint32_t r, i0, i1, pad;
    uint64_t toff;
};

template <int BLOCK, int PASS>
__global__ __launch_bounds__(BLOCK) void k_count_dense(const HubItem* __restrict__ items, uint32_t nitems,
                                                       RowView g, uint32_t* __restrict__ pool,
                                                       uint64_t* __restrict__ sup, uint64_t* __restrict__ total,
                                                       Emit em, unsigned* queue) {
    __shared__ ExpandShared<BLOCK> sh;
    __shared__ uint32_t s_take;
    uint64_t tot = 0;
    for (;;) {
        uint32_t const it = take_units<BLOCK>(queue, 1u, &s_take);
        if (it >= nitems) break;
        HubItem const h = items[it];
        DenseTable t{pool + h.toff};
        tot += expand_entries<BLOCK, PASS>(g, h.r, h.i0, h.i1, t, sup, em, sh);
    }
    if (PASS == 1) {
        uint64_t const s = block_sum_u64<BLOCK>(tot, sh.red);
        if (threadIdx.x == 0 && s) atomic_add_u64(total, s);
    }
}

__global__ void k_max_u64(uint32_t m, const uint64_t* __restrict__ a, uint64_t* __restrict__ out) {
    uint64_t mx = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        mx = a[i] > mx ? a[i] : mx;
    mx = wave_max_u64(mx);
    if ((threadIdx.x & (kWave - 1)) == 0 && mx) atomic_max_u64(out, mx);
}

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) b++;
    return b;
}

// Launch of buckets A and B for the tuned settings: the LDS table has twice
// the bucket's wedge bound in slots.
struct LdsLaunch {
    const uint32_t* roots;
    uint32_t nroots;
    RowView g;
    const uint32_t* r_lo;
    uint64_t* sup;
    uint64_t* total;
    int cb, per_edge;
    Emit em;
    unsigned grid_cap;
    hipStream_t s;
    unsigned* queue;  // zeroed device word: the roots taken so far
    template <int BLOCK, uint32_t CAP>
    void run() const {
        k_count_lds<BLOCK, CAP><<<grid_for(nroots, 1, grid_cap), BLOCK, 0, s>>>(roots, nroots, g, r_lo, sup, total,
                                                                               cb, per_edge, em, queue);
    }
    int bucket_a(uint32_t wave_wedges) const {
        switch (wave_wedges) {
            case 256: run<kWave, 512>(); break;
            case 512: run<kWave, 1024>(); break;
            case 1024: run<kWave, 2048>(); break;
            case 2048: run<kWave, 4096>(); break;
            default: return BT_ERR_INVALID;
        }
        return BT_OK;
    }
    int bucket_b(uint32_t block_wedges, uint32_t threads) const {
        if (threads == 256) {
            switch (block_wedges) {
                case 4096: run<256, 8192>(); break;
                case 8192: run<256, 16384>(); break;
                case 16384: run<256, 32768>(); break;
                default: return BT_ERR_INVALID;
            }
        } else if (threads == 512) {
            switch (block_wedges) {
                case 4096: run<512, 8192>(); break;
                case 8192: run<512, 16384>(); break;
                case 16384: run<512, 32768>(); break;
                default: return BT_ERR_INVALID;
            }
        } else {
            return BT_ERR_INVALID;
        }
        return BT_OK;
    }
};

// queues: two zeroed device words, one per pass
template <int BLOCK>
void launch_dense(unsigned grid, hipStream_t s, const HubItem* items, uint32_t nitems, const RowView& g,
                  uint32_t* pool, uint64_t* sup, uint64_t* total, const Emit& em, bool per_edge, unsigned* queues) {
    k_count_dense<BLOCK, 1><<<grid, BLOCK, 0, s>>>(items, nitems, g, pool, sup, total, em, queues);
    if (per_edge) k_count_dense<BLOCK, 2><<<grid, BLOCK, 0, s>>>(items, nitems, g, pool, sup, total, em, queues + 1);
}

}  // namespace

int bt_prep_priority(bt_dgraph* dg, const bt_fixed_scratch* fixed) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const n = dg->n, m = dg->m;
    if (n == 0) return BT_OK;
    constexpr unsigned T = 256;
    int const nb = bits_for(n);
    // temporary storage for the exact calls below
    size_t t1 = 0, t2 = 0, t3 = 0;
    BT_HIP_TRY(rocprim::radix_sort_keys(nullptr, t1, dg->r_key, dg->r_key2, n, 0, 64, s));
    if (m)
        BT_HIP_TRY(rocprim::radix_sort_pairs(nullptr, t2, dg->r_key, dg->r_key2, dg->r_val2, dg->r_eid,
                                             2 * size_t(m), 0, 32 + nb, s));
    BT_HIP_TRY(rocprim::exclusive_scan(nullptr, t3, dg->list0, dg->r_off, 0u, n, rocprim::plus<uint32_t>(), s));
    size_t const t_need = std::max(t1, std::max(t2, t3));
    void* tmp = nullptr;
    if (fixed) {
        if (t_need > fixed->bytes) return BT_ERR_SCRATCH;
        tmp = fixed->ptr;
    } else {
        BT_TRY(bt_buf_reserve(c->sort_tmp, t_need));
        tmp = c->sort_tmp.ptr;
    }
    // rank = position in (degree, id) order
    k_rank_keys<<<grid_for(n, T, 1u << 30), T, 0, s>>>(n, dg->deg, dg->r_key);
    BT_HIP_TRY(rocprim::radix_sort_keys(tmp, t1, dg->r_key, dg->r_key2, n, 0, 64, s));
    k_rank_scatter<<<grid_for(n, T, 1u << 30), T, 0, s>>>(n, dg->u_num, dg->r_key2, dg->rank, dg->list0, dg->r_up);
    BT_HIP_TRY(rocprim::exclusive_scan(tmp, t3, dg->list0, dg->r_off, 0u, n, rocprim::plus<uint32_t>(), s));
    k_set_last<<<1, 64, 0, s>>>(dg->r_off, n, 2 * m);
    if (m) {
        // rows of the relabelled CSR: sort directed pairs by (rank, rank)
        k_pair_keys<<<grid_for(m, T, 1u << 30), T, 0, s>>>(m, dg->edges, dg->rank, dg->r_key, dg->r_val2);
        BT_HIP_TRY(rocprim::radix_sort_pairs(tmp, t2, dg->r_key, dg->r_key2, dg->r_val2, dg->r_eid,
                                             2 * size_t(m), 0, 32 + nb, s));
        k_pair_unpack<<<grid_for(2ull * m, T, 1u << 30), T, 0, s>>>(2 * m, dg->r_key2, dg->r_nbr);
    }
    k_lo_deg<<<grid_for(n, T, 1u << 30), T, 0, s>>>(n, dg->r_off, dg->r_nbr, dg->r_lo);
    if (m) k_prefix_p<<<grid_for(2ull * m, T, 1u << 30), T, 0, s>>>(2 * m, dg->r_key2, dg->r_off, dg->r_nbr, dg->r_p);
    BT_HIP_TRY(hipGetLastError());
    return BT_OK;
}

int bt_count_impl(bt_dgraph* dg, bool per_edge, uint64_t* total_out, uint64_t* smax_out, bool emit,
                  const bt_fixed_scratch* fixed) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const n = dg->n, m = dg->m;
    c->last = bt_timings{};
    bt_timer tm(c);
    BT_TRY(tm.start(0));
    dg->has_support = false;
    if (m) BT_HIP_TRY(hipMemsetAsync(dg->sup, 0, sizeof(uint64_t) * m, s));
    BT_HIP_TRY(hipMemsetAsync(dg->ctl, 0, sizeof(uint64_t) * 64, s));
    if (m == 0 || n == 0) {
        BT_TRY(tm.start(1));
        BT_HIP_TRY(hipStreamSynchronize(s));
        if (total_out) *total_out = 0;
        if (smax_out) *smax_out = 0;
        dg->has_support = per_edge;
        c->last.total_ms = tm.ms(0, 1);
        return BT_OK;
    }
    if (fixed) emit = false;  // the bloom index is never built into fixed scratch
    BT_TRY(bt_prep_priority(dg, fixed));
    BT_TRY(tm.start(1));
    constexpr unsigned T = 256;
    k_root_w<<<grid_for(uint64_t(n) * kWave, T, 1u << 30), T, 0, s>>>(n, dg->r_off, dg->r_lo, dg->r_p, dg->r_w,
                                                                        dg->ctl);
    int const kb = bits_for(uint64_t(n));  // bits of w+1 <= n
    int const cb = 32 - kb;
    uint32_t const lo_limit = cb >= 31 ? 0x7FFFFFFFu : (1u << cb);
    uint32_t* la = dg->list0;
    uint32_t* lb = dg->list1;
    uint32_t* lc = dg->r_val2;
    bt_count_config const cfg = bt_count_settings(dg);
    if (!bt_count_config_valid(cfg)) return BT_ERR_INVALID;
    k_classify<<<grid_for(n, T, 1u << 30), T, 0, s>>>(n, dg->r_w, dg->r_lo, cb > 0 ? lo_limit : 0, cfg.wave_wedges,
                                                      cfg.block_wedges, la, lb, lc, dg->ctl);
    uint64_t cnts[4];
    BT_TRY(bt_read_ctl(dg, 0, 4, cnts));
    uint64_t const na = cnts[0], nb = cnts[1], nc = cnts[2];
    c->last.wedges_count = cnts[3];
    RowView g{dg->r_off, dg->r_nbr, dg->r_eid, dg->r_p};
    Emit em;
    c->ix.R = kNoIndex;
    if (emit && per_edge) {
        // at most one record per priority wedge, 16 bytes each in emission
        // order: these buffers may take index_mem_frac of the free device
        // memory; the rest of the index is sized once the records are counted
        bt_index& ix = c->ix;
        size_t const cap = std::max<uint64_t>(cnts[3], 1);
        size_t free_b = 0, total_b = 0;
        BT_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        // the share is of the memory the index can have: what is free plus
        // every index buffer the context already holds from an earlier call
        // (they are reused), so that a repeated call on a large graph finds
        // the same room as the first
        size_t const have = ix.keys0.bytes + ix.vals0.bytes + ix.keys1.bytes + ix.vals1.bytes + ix.bid.bytes +
                            ix.bstart.bytes + ix.bk.bytes + ix.bclaim.bytes + ix.inc_off.bytes + ix.inc_cur.bytes +
                            ix.inc.bytes + ix.touched.bytes + ix.bgraph.bytes + ix.live.bytes + ix.c2s.bytes +
                            ix.hdirty.bytes;
        if (double(cap) * 16.0 > (double(free_b) + double(have)) * c->index_mem_frac) emit = false;
    }
    if (emit && per_edge) {
        bt_index& ix = c->ix;
        size_t const cap = std::max<uint64_t>(cnts[3], 1);
        BT_TRY(bt_buf_reserve(ix.keys0, cap * sizeof(uint64_t)));
        BT_TRY(bt_buf_reserve(ix.vals0, cap * sizeof(uint64_t)));
        em.keys = static_cast<uint64_t*>(ix.keys0.ptr);
        em.vals = static_cast<uint64_t*>(ix.vals0.ptr);
        em.count = reinterpret_cast<unsigned long long*>(dg->ctl + 10);
        em.nb = bits_for(n);
    }
    uint64_t* total = dg->ctl + 8;
    unsigned const max_grid = unsigned(c->cu_count) * 16;
    BT_TRY(tm.start(2));
    // work-queue counters in control words 4..7 (zeroed above): bucket A,
    // bucket B, the hub items' two passes (reset per batch of hub roots)
    unsigned* queues = reinterpret_cast<unsigned*>(dg->ctl + 4);
    if (na)
        BT_TRY((LdsLaunch{la, uint32_t(na), g, dg->r_lo, dg->sup, total, cb, per_edge, em, max_grid * 4, s, queues}
                    .bucket_a(cfg.wave_wedges)));
    if (nb)
        BT_TRY((LdsLaunch{lb, uint32_t(nb), g, dg->r_lo, dg->sup, total, cb, per_edge, em, max_grid, s, queues + 2}
                    .bucket_b(cfg.block_wedges, cfg.block_threads)));
    BT_HIP_TRY(hipGetLastError());
    if (nc) {
        // hub roots: build (root, entry chunk) items on the host, batch them
        // so that their dense tables fit the pool
        std::vector<uint32_t> hroots(nc), hlo(nc), hoff(nc);
        BT_HIP_TRY(hipMemcpyAsync(hroots.data(), lc, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, s));
        BT_HIP_TRY(hipStreamSynchronize(s));
        std::sort(hroots.begin(), hroots.end());
        std::vector<uint32_t> all_lo(n), all_off(n + 1);
        BT_HIP_TRY(hipMemcpyAsync(all_lo.data(), dg->r_lo, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        BT_HIP_TRY(hipMemcpyAsync(all_off.data(), dg->r_off, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, s));
        BT_HIP_TRY(hipStreamSynchronize(s));
        size_t const pool_entries_max = size_t(256) << 20;  // 1 GiB of u32 counters
        size_t max_table = 0;
        for (uint32_t r : hroots) max_table = std::max<size_t>(max_table, r);
        size_t pool_entries = std::max(max_table, std::min<size_t>(pool_entries_max, size_t(n) * nc));
        uint32_t* pool = nullptr;
        if (fixed) {
            // the sort storage is done with (stream order): the tables take
            // the whole region, at least the largest root's table
            pool_entries = fixed->bytes / sizeof(uint32_t);
            if (max_table > pool_entries) return BT_ERR_SCRATCH;
            pool = static_cast<uint32_t*>(fixed->ptr);
        } else {
            BT_TRY(bt_buf_reserve(c->pool, pool_entries * sizeof(uint32_t)));
            pool = static_cast<uint32_t*>(c->pool.ptr);
        }
        HubItem* ditems = reinterpret_cast<HubItem*>(dg->r_key);  // scratch: 2m+n u64 >= items
        size_t const items_cap = (size_t(2) * m + n) * sizeof(uint64_t) / sizeof(HubItem);
        size_t k = 0;
        while (k < nc) {
            std::vector<HubItem> items;
            size_t used = 0;
            size_t k2 = k;
            while (k2 < nc && used + hroots[k2] <= pool_entries) {
                uint32_t const r = hroots[k2];
                uint32_t const b = all_off[r], lo = all_lo[r];
                for (uint32_t i = 0; i < lo; i += cfg.hub_threads)
                    items.push_back(HubItem{r, b + i, b + std::min<uint32_t>(lo, i + cfg.hub_threads), 0, used});
                used += r;
                k2++;
            }
            if (k2 == k || items.size() > items_cap) return BT_ERR_INTERNAL;
            BT_HIP_TRY(hipMemsetAsync(pool, 0, used * sizeof(uint32_t), s));
            BT_HIP_TRY(hipMemsetAsync(queues + 4, 0, 2 * sizeof(unsigned), s));
            BT_HIP_TRY(hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(HubItem), hipMemcpyHostToDevice, s));
            unsigned const gC = grid_for(items.size(), 1, max_grid);
            if (cfg.hub_threads == 512)
                launch_dense<512>(gC, s, ditems, uint32_t(items.size()), g, pool, dg->sup, total, em, per_edge,
                                  queues + 4);
            else
                launch_dense<256>(gC, s, ditems, uint32_t(items.size()), g, pool, dg->sup, total, em, per_edge,
                                  queues + 4);
            BT_HIP_TRY(hipGetLastError());
            // the host vector `items` is reused only after this batch's copy
            BT_HIP_TRY(hipStreamSynchronize(s));
            k = k2;
        }
    }
    BT_TRY(tm.start(3));
    if (per_edge) {
        uint64_t* smax = dg->ctl + 9;
        k_max_u64<<<grid_for(m, T, unsigned(c->cu_count) * 8), T, 0, s>>>(m, dg->sup, smax);
        BT_HIP_TRY(hipGetLastError());
    }
    BT_TRY(tm.start(4));
    uint64_t res[3];
    BT_TRY(bt_read_ctl(dg, 8, 3, res));
    if (total_out) *total_out = res[0];
    if (smax_out) *smax_out = res[1];
    c->ix.R = em.keys ? res[2] : kNoIndex;
    if (em.keys) {
        // the rest of the index, per record: sorted keys and values (16 B),
        // bloom ids, sizes and claim stamps (12 B) and incidences (8 B); it
        // must fit the free device memory with a tenth to spare, and records
        // are named by u32 indices
        bt_index& ix = c->ix;
        size_t free_b = 0, total_b = 0;
        BT_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        size_t const have = ix.keys1.bytes + ix.vals1.bytes + ix.bid.bytes + ix.bk.bytes + ix.bclaim.bytes +
                            ix.inc.bytes;
        double const need = double(res[2]) * 36.0;
        if (res[2] > kMaxIndexRecords || need > (double(free_b) + double(have)) * 0.9) c->ix.R = kNoIndex;
    }
    dg->has_support = per_edge;
    c->last.prep_ms = tm.ms(0, 1);
    c->last.count_kernel_ms = tm.ms(2, 3);
    c->last.count_ms = tm.ms(1, 4);
    c->last.total_ms = tm.ms(0, 4);
    return BT_OK;
}

// ---- allocation-free ebfc ---------------------------------------------------
//
// For callers that cannot allocate device memory: bt_ebfc_static takes the
// whole scratch of the call (sort storage, then the hub roots' tables) from
// this array. Each device holds its own copy for the life of the process;
// host code uses the address hipGetSymbolAddress returns on that device,
// never the array's name (a host-side stand-in, not a device address).
static __device__ __align__(256) uint8_t scratch[BT_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxStaticDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxStaticDevices] = {};

// The array's address on the context's device (looked up once per device,
// with that device current).
int static_scratch(bt_ctx* c, void** out) {
    if (c->device < 0 || c->device >= kMaxStaticDevices) return BT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        BT_HIP_TRY(hipSetDevice(c->device));
        void* p = nullptr;
        BT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return BT_ERR_DEVICE;
        g_scratch_dev[c->device] = p;
    }
    *out = g_scratch_dev[c->device];
    return BT_OK;
}
}  // namespace

extern "C" int bt_ebfc_static(bt_ctx* ctx, bt_dgraph* dg, uint64_t* support_out, uint64_t* total,
                              uint64_t* support_max) {
    if (!ctx || !dg || dg->ctx != ctx) return BT_ERR_INVALID;
    void* p = nullptr;
    BT_TRY(static_scratch(ctx, &p));
    BT_HIP_TRY(hipSetDevice(ctx->device));
    bt_fixed_scratch const fx{p, size_t(BT_STATIC_SCRATCH_BYTES)};
    uint64_t t = 0, sm = 0;
    BT_TRY(bt_count_impl(dg, true, &t, &sm, false, &fx));
    if (support_out && dg->m) {
        BT_HIP_TRY(hipMemcpyAsync(support_out, dg->sup, sizeof(uint64_t) * dg->m, hipMemcpyDeviceToHost,
                                  ctx->stream));
        BT_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    if (total) *total = t;
    if (support_max) *support_max = sm;
    return BT_OK;
}
