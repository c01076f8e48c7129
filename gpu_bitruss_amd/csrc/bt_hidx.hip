// h-index bitruss numbers (Jacobi rounds) over the bloom index.
//
// Reference semantics: g_bitruss_hindex / bitruss_decomposition_h_index
// (bitruss_hidx.cu:51-256) and c_bitruss_hindex / compute_bitruss_hindex
// (cpu_bitruss.cpp:125-239): start from the butterfly supports; in each round
// every edge e with prev[e] > 0 takes the h-index of the values
// min(prev of the three other edges) over its butterflies whose three other
// edges are all non-zero; edges with prev[e] == 0 stay 0; rounds repeat until
// one changes nothing (that round is counted, so an edgeless graph takes one
// round, bitruss_hidx.cu:218-231).
//
// With the bloom index a butterfly is a pair of middles (i, j) of one bloom.
// For edge e in middle i (partner edge e' of the same middle) the values are
// min(prev[e'], mv[j]) for every other middle j of the bloom, where
// mv[j] = min(prev of j's two edges) is computed once per round. A zero
// value is a butterfly the reference skips; it never counts towards an
// h-index >= 1, so it is simply not binned.
//
// h-index of one edge's values: it never exceeds prev[e] (the rounds are
// non-increasing), so values are clipped to prev[e] and binned into kHB
// bins of an LDS histogram; when the answer's bin is wider than one value
// the values are streamed again inside that bin (two passes up to
// prev[e] ~ 10^6, a third above). One wave per edge; the wave walks the
// concatenation of the edge's blooms 64 middles at a time.
//
// Sorted chunks: that pass costs the sizes of the edge's blooms, which on a
// skewed graph run to tens of thousands of middles per bloom. Each round
// therefore also keeps every bloom's mv values (clamped to 32 bits, exact
// for thresholds below 2^32) in descending order: whole for blooms up to
// kSmallBloom middles (a thread sorts one), in chunks of kChunk above (a
// workgroup sorts one), re-sorting only the blooms with a change in the
// previous round. The number of an edge's values >= t is then, per bloom
// with prev[partner] >= t, a binary search per chunk (the chunks of one
// bloom spread over the wave) minus the edge's own middle, and the h-index
// is found by a search over t after a first count at t = prev[e] (most
// edges keep their value in the late rounds), each count narrowing the
// range from both sides. An edge is evaluated again only when a value in
// one of its blooms dropped from at least its h-index (k_hidx_mark). Per
// edge the wave estimates both costs from its blooms' sizes and takes the
// cheaper way; both give the same h-index. Edges in more than kHeavyInc
// blooms are evaluated by a whole workgroup each (hidx_heavy).
//
// Round loop: the reference reads a change flag back after every round
// (bitruss_hidx.cu:219-231). Here the rounds are captured once into a HIP
// graph of K rounds (BT_OPT_HIDX_GRAPH_ROUNDS, default 8) whose kernels keep
// the loop state on the device: a round counter, the change flag and a done
// flag that the first round without a change sets, after which the graph's
// remaining rounds return at once. The host launches the graph and reads
// the state back once per K rounds. The buffers alternate by the position
// of a round in the graph (K is even), and the round without a change
// leaves both buffers equal, so the result is in either.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>

#include <algorithm>

#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

namespace {

constexpr int kHT = 512;              // block size
constexpr int kHW = kHT / kWave;      // waves per block
constexpr uint32_t kHB = 1024;        // histogram bins per wave
constexpr uint32_t kSmallBloom = 32;  // blooms up to this size: sorted whole, a thread each
constexpr uint32_t kChunk = 2048;     // larger blooms: mv values sorted in chunks of this length
constexpr int kST = 256;              // block size of the chunk sort
constexpr uint32_t kCacheInc = 256;   // incidences whose bloom data a wave keeps in LDS (in its histogram)
constexpr uint32_t kSearchCost = 12;  // cost of a binary search in one chunk, in values streamed
constexpr uint64_t kHeavyInc = 256;   // edges with more incidences: a workgroup each (sorted chunks only)
constexpr uint32_t kHeavyCache = kHW * kHB / 3;  // incidences a heavy edge's workgroup caches in its histograms
constexpr uint32_t kHeavyGlobal = 61440;         // and the next ones in its own HBM rows (hcache)

// Device-side loop state (control words of the graph): round counter,
// change flag of the current round, done flag.
struct LoopState {
    unsigned long long* round;
    unsigned long long* changed;
    unsigned long long* done;
    unsigned long long* work;  // next heavy edge of the round
};

__global__ void k_hidx_loop_init(LoopState st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *st.round = 0;
        *st.changed = 0;
        *st.done = 0;
    }
}

// Start of a round: count it and clear its change flag (unless done).
__global__ void k_hidx_round_begin(LoopState st) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !*st.done) {
        *st.round += 1;
        *st.changed = 0;
        *st.work = 0;
    }
}

// End of a round: a round without a change ends the loop.
__global__ void k_hidx_round_end(LoopState st) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !*st.done && !*st.changed) *st.done = 1;
}

// Clears a round's bloom marks and largest changed values (a kernel, not a
// memset node: the graph replays it).
__global__ void k_hidx_clear(uint8_t* __restrict__ p, uint32_t* __restrict__ q, size_t n,
                             const unsigned long long* done) {
    if (done && *done) return;
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        p[i] = 0;
        q[i] = 0;
    }
}

// mv of every record, and its 32-bit copy smv (null: none) for the sorted
// chunks. A bloom without a change in the previous round (dirty_in, null:
// every bloom) keeps the values it has; for a bloom with a change, dmax
// receives the largest old mv among its records whose mv changed (clamped
// to 32 bits, 0xFFFFFFFF standing for any larger value). In graph rounds
// round 1 ignores dirty_in.
__global__ void k_hidx_mv(uint32_t R, const uint64_t* __restrict__ rec, const uint64_t* __restrict__ prev,
                          uint64_t* __restrict__ mv, uint32_t* __restrict__ smv, const uint32_t* __restrict__ bid,
                          const uint8_t* __restrict__ dirty_in, uint32_t* __restrict__ dmax,
                          const unsigned long long* dev_round, const unsigned long long* done) {
    if (done && *done) return;
    if (dev_round && *dev_round == 1) dirty_in = nullptr;
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < R; p += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t const b0 = dirty_in ? bid[p] - 1 : 0u;
        if (dirty_in && !dirty_in[b0]) continue;
        uint64_t const r = rec[p];
        uint64_t const a = prev[uint32_t(r >> 32)], b = prev[uint32_t(r)];
        uint64_t const v = a < b ? a : b;
        if (dirty_in) {
            uint64_t const old = mv[p];
            if (old != v) atomicMax(&dmax[b0], old < 0xFFFFFFFFull ? uint32_t(old) : 0xFFFFFFFFu);
        }
        mv[p] = v;
        if (smv) smv[p] = v < 0xFFFFFFFFull ? uint32_t(v) : 0xFFFFFFFFu;
    }
}

// Marks the edges to evaluate again: both edges x of every record of a
// bloom whose largest changed old value d is >= prev[x]. An edge's count at
// t = prev[x] changes only when one of its butterflies' values drops from
// >= t to < t: either a middle's mv did (its old mv >= t) or the partner's
// value did, which changes the edge's own middle from an old mv >= t. Its
// count at prev[x] stays >= prev[x] otherwise, and so does its h-index.
// (An edge lies in a bloom exactly when it is an edge of one of its
// records.) Round 1 evaluates every edge.
__global__ void k_hidx_mark(uint32_t R, const uint64_t* __restrict__ rec, const uint64_t* __restrict__ prev,
                            const uint32_t* __restrict__ bid, const uint8_t* __restrict__ dirty_in,
                            const uint32_t* __restrict__ dmax, uint8_t* __restrict__ emark,
                            const unsigned long long* dev_round, const unsigned long long* done) {
    if (done && *done) return;
    if (!dirty_in || (dev_round && *dev_round == 1)) return;
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < R; p += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t const b0 = bid[p] - 1;
        if (!dirty_in[b0]) continue;
        uint32_t const d = dmax[b0];
        if (d == 0) continue;
        uint64_t const r = rec[p];
        uint32_t const x = uint32_t(r >> 32), y = uint32_t(r);
        if (d == 0xFFFFFFFFu || d >= prev[x]) emark[x] = 1;
        if (d == 0xFFFFFFFFu || d >= prev[y]) emark[y] = 1;
    }
}

// The chunks of the blooms above kSmallBloom middles, as bloom << 32 |
// chunk: counted into *cnt (list null), then listed at *cnt's positions
// (the order is free: each chunk is sorted on its own).
__global__ void k_hidx_chunks(uint32_t NB, const uint32_t* __restrict__ bstart, unsigned long long* cnt,
                              uint64_t* __restrict__ list) {
    for (size_t b = blockIdx.x * size_t(blockDim.x) + threadIdx.x; b < NB; b += size_t(gridDim.x) * blockDim.x) {
        uint32_t const k = bstart[b + 1] - bstart[b];
        if (k <= kSmallBloom) continue;
        uint32_t const nch = (k + kChunk - 1) / kChunk;
        unsigned long long const at = atomicAdd(cnt, (unsigned long long)nch);
        if (list)
            for (uint32_t c = 0; c < nch; c++) list[at + c] = (uint64_t(b) << 32) | c;
    }
}

// The heavy edges (above kHeavyInc incidences, support below 2^32): counted
// into *cnt (list null), then listed.
__global__ void k_hidx_heavy_list(uint32_t m, const uint64_t* __restrict__ inc_off, const uint64_t* __restrict__ sup,
                                  unsigned long long* cnt, uint32_t* __restrict__ list) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        if (inc_off[e + 1] - inc_off[e] <= kHeavyInc || sup[e] > 0xFFFFFFFFull) continue;
        unsigned long long const at = atomicAdd(cnt, 1ull);
        if (list) list[at] = e;
    }
}

// Sorts the blooms of 2..kSmallBloom middles of smv in descending order, a
// thread per bloom (insertion sort in its LDS row), skipping the blooms
// without a change in the previous round.
__global__ __launch_bounds__(256) void k_hidx_sort_small(uint32_t NB, const uint32_t* __restrict__ bstart,
                                                         uint32_t* __restrict__ smv,
                                                         const uint8_t* __restrict__ dirty_in,
                                                         const unsigned long long* dev_round,
                                                         const unsigned long long* done) {
    if (done && *done) return;
    if (dev_round && *dev_round == 1) dirty_in = nullptr;
    __shared__ uint32_t s_rows[256 * (kSmallBloom + 1)];
    uint32_t* v = s_rows + threadIdx.x * (kSmallBloom + 1);
    for (size_t b = blockIdx.x * size_t(blockDim.x) + threadIdx.x; b < NB; b += size_t(gridDim.x) * blockDim.x) {
        uint32_t const b0 = bstart[b], k = bstart[b + 1] - b0;
        if (k < 2 || k > kSmallBloom || (dirty_in && !dirty_in[b])) continue;
        for (uint32_t j = 0; j < k; j++) {
            uint32_t const x = smv[b0 + j];
            uint32_t i = j;
            while (i > 0 && v[i - 1] < x) {
                v[i] = v[i - 1];
                i--;
            }
            v[i] = x;
        }
        for (uint32_t j = 0; j < k; j++) smv[b0 + j] = v[j];
    }
}

// Sorts each listed chunk of smv in descending order (a bitonic sort in
// LDS, padded with zeros to a power of two), skipping the blooms without
// a change in the previous round. One workgroup per chunk.
__global__ __launch_bounds__(kST) void k_hidx_sort(const uint64_t* __restrict__ list, uint64_t n,
                                                   const uint32_t* __restrict__ bstart, uint32_t* __restrict__ smv,
                                                   const uint8_t* __restrict__ dirty_in,
                                                   const unsigned long long* dev_round,
                                                   const unsigned long long* done) {
    if (done && *done) return;
    if (dev_round && *dev_round == 1) dirty_in = nullptr;
    __shared__ uint32_t s[kChunk];
    for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
        uint64_t const it = list[i];
        uint32_t const b = uint32_t(it >> 32), ch = uint32_t(it);
        if (dirty_in && !dirty_in[b]) continue;  // the same for the whole workgroup
        uint32_t const b0 = bstart[b], k = bstart[b + 1] - b0;
        uint32_t const start = b0 + ch * kChunk;
        uint32_t const len = std::min(kChunk, k - ch * kChunk);
        uint32_t P = 64;
        while (P < len) P <<= 1;
        __syncthreads();  // the previous chunk's values are written out
        for (uint32_t j = threadIdx.x; j < P; j += kST) s[j] = j < len ? smv[start + j] : 0u;
        for (uint32_t size = 2; size <= P; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (uint32_t j = threadIdx.x; j < P / 2; j += kST) {
                    uint32_t const lo = 2 * j - (j & (stride - 1)), hi = lo + stride;
                    bool const desc = (lo & size) == 0;  // the last merge (size P) is descending throughout
                    uint32_t const x = s[lo], y = s[hi];
                    if ((x < y) == desc) {
                        s[lo] = y;
                        s[hi] = x;
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < len; j += kST) smv[start + j] = s[j];
    }
}

struct HidxArgs {
    uint32_t m;
    const uint64_t* prev;
    uint64_t* curr;
    const uint64_t* inc_off;
    const uint32_t* inc;     // per incidence: record (middle) holding the edge
    const uint64_t* rec;     // middles: id(r-v) << 32 | id(v-w)
    const uint32_t* bid;     // per record: bloom index + 1
    const uint32_t* bstart;  // per bloom: first record, then R
    const uint64_t* mv;      // per record: min(prev) of its two edges
    // per record: mv clamped to 32 bits, in descending chunks of kChunk for
    // the blooms above kSmallBloom middles (null: every edge streams)
    const uint32_t* smv;
    unsigned long long* changed;
    const uint32_t* egraph;  // batches: graph of each edge (else null)
    uint32_t* glast;         // batches: per graph, last round with a change
    uint32_t round;
    // per bloom: a value of one of its edges changed in the previous round
    // (null: every bloom, the first round), and the marks of this round
    const uint8_t* dirty_in;
    uint8_t* dirty_out;
    // per edge: one of its blooms changed in the previous round (set by
    // k_hidx_mv, cleared here when the edge is evaluated)
    uint8_t* emark;
    // edges above kHeavyInc incidences with a support below 2^32 (sorted
    // chunks only): taken from the work counter by the first heavy_blocks
    // workgroups, a workgroup per edge; the others skip them
    const uint64_t* sup;
    const uint32_t* heavy;
    uint32_t n_heavy, heavy_blocks;
    unsigned long long* heavy_next;
    uint32_t* hcache;  // per heavy workgroup: 3 rows of kHeavyGlobal (pp, base, k)
    // graph rounds: the device loop state (null: round and dirty_in above
    // hold); the round number is read from it, and round 1 ignores dirty_in
    const unsigned long long* dev_round;
    const unsigned long long* done;
};

// Streams the values of edge e's butterflies (one wave), clips them to cap
// (= prev[e], which bounds the answer) and bins those in [lo, hi] into
// hist[(v - lo) / width]; values above hi are the ones already counted
// above the range.
__device__ __forceinline__ void bin_values(const HidxArgs& a, uint32_t e, uint64_t q0, uint64_t q1, uint64_t lo,
                                           uint64_t hi, uint64_t width, uint64_t cap, uint32_t* hist, uint32_t* s_incl,
                                           uint32_t* s_base, uint32_t* s_skip, uint64_t* s_pp) {
    int const lane = threadIdx.x & (kWave - 1);
    for (uint64_t b0 = q0; b0 < q1; b0 += kWave) {
        uint64_t const q = b0 + lane;
        uint32_t len = 0, base = 0, skip = 0;
        uint64_t pp = 0;
        if (q < q1) {
            uint32_t const p = a.inc[q];
            uint64_t const r = a.rec[p];
            uint32_t const partner = uint32_t(r >> 32) == e ? uint32_t(r) : uint32_t(r >> 32);
            pp = a.prev[partner];
            uint32_t const b = a.bid[p] - 1;
            base = a.bstart[b];
            skip = p;
            // a zero partner makes every value of this bloom zero
            len = pp ? a.bstart[b + 1] - base : 0u;
        }
        uint32_t const incl = wave_incl_scan(len);
        uint32_t const total = __shfl(incl, kWave - 1, kWave);
        __builtin_amdgcn_wave_barrier();
        s_incl[lane] = incl;
        s_base[lane] = base - (incl - len);  // record index minus flattened position
        s_skip[lane] = skip;
        s_pp[lane] = pp;
        __builtin_amdgcn_wave_barrier();
        uint32_t const nb = b0 + kWave <= q1 ? uint32_t(kWave) : uint32_t(q1 - b0);
        for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
            uint32_t const t = t0 + lane;
            if (t < total) {
                // the bloom of flattened position t: first prefix above t
                uint32_t j = 0;
#pragma unroll
                for (uint32_t step = kWave / 2; step > 0; step >>= 1) {
                    uint32_t const k = j + step;
                    j = (k <= nb && s_incl[k - 1] <= t) ? k : j;
                }
                uint32_t const idx = s_base[j] + t;
                if (idx != s_skip[j]) {
                    uint64_t v = a.mv[idx];
                    uint64_t const pj = s_pp[j];
                    v = pj < v ? pj : v;
                    v = v < cap ? v : cap;
                    if (v != 0 && v >= lo && v <= hi) atomicAdd(&hist[(v - lo) / width], 1u);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Bloom data of edge e's incidence q: its partner's value clamped to 32 bits
// (0: the bloom gives e no non-zero value), the bloom's first record and
// its size (not read for a zero partner).
struct IncInfo {
    uint32_t pp, base, k;
};

__device__ __forceinline__ IncInfo inc_info(const HidxArgs& a, uint32_t e, uint64_t q) {
    uint32_t const p = a.inc[q];
    uint64_t const r = a.rec[p];
    uint32_t const partner = uint32_t(r >> 32) == e ? uint32_t(r) : uint32_t(r >> 32);
    uint64_t const pp = a.prev[partner];
    IncInfo x{pp < 0xFFFFFFFFull ? uint32_t(pp) : 0xFFFFFFFFu, 0u, 0u};
    if (x.pp) {
        uint32_t const b = a.bid[p] - 1;
        x.base = a.bstart[b];
        x.k = a.bstart[b + 1] - x.base;
    }
    return x;
}

// mv values >= t in one item of a bloom: the whole bloom when it is small,
// else its chunk ch (a binary search in descending order either way).
__device__ __forceinline__ uint32_t item_count(const uint32_t* __restrict__ smv, uint32_t base, uint32_t k,
                                               uint32_t ch, uint32_t t) {
    const uint32_t* s = smv + base + ch * kChunk;
    uint32_t lo = 0, hi = std::min(kChunk, k - ch * kChunk);
    while (lo < hi) {  // the first position below t
        uint32_t const mid = (lo + hi) >> 1;
        if (s[mid] >= t)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Number of edge e's butterfly values >= t (1 <= t <= prev[e]), summed over
// the wave. Each bloom with prev[partner] >= t gives its mv values >= t less
// the edge's own middle (mv = min(prev[e], partner) >= t). The wave takes
// 64 incidences at a time (the bloom data of the first kCacheInc from its
// LDS cache, if any: pp, base and k in three rows of cap, then those of
// the next kHeavyGlobal from gcache, if any) and
// spreads their items (a small bloom, or one chunk of a larger one) over
// its lanes, so that the chunks of one large bloom are searched side by
// side. s_incl, s_base and s_k are the wave's 64-entry LDS rows. A wave of
// a workgroup that shares the edge takes the 64 incidences from q0 + first
// on, every stride, and returns its own share.
__device__ uint64_t count_at_least(const HidxArgs& a, uint32_t e, uint64_t q0, uint64_t q1, uint32_t t,
                                   const uint32_t* cache, uint32_t* s_incl, uint32_t* s_base, uint32_t* s_k,
                                   uint64_t first = 0, uint64_t stride = kWave, uint32_t cap = kCacheInc,
                                   const uint32_t* gcache = nullptr) {
    int const lane = threadIdx.x & (kWave - 1);
    uint64_t c = 0, own = 0;
    for (uint64_t b0 = q0 + first; b0 < q1; b0 += stride) {
        uint64_t const q = b0 + lane;
        IncInfo x{0u, 0u, 0u};
        if (q < q1) {
            uint64_t const i = q - q0;
            if (cache && i < cap)
                x = IncInfo{cache[i], cache[cap + i], cache[2 * cap + i]};
            else if (gcache && i >= cap && i - cap < kHeavyGlobal)
                x = IncInfo{gcache[i - cap], gcache[kHeavyGlobal + i - cap], gcache[2 * kHeavyGlobal + i - cap]};
            else
                x = inc_info(a, e, q);
        }
        uint32_t const items = x.pp >= t ? (x.k <= kSmallBloom ? 1u : (x.k + kChunk - 1) / kChunk) : 0u;
        own += __builtin_popcountll(__ballot(items != 0));
        if (!__ballot(items > 1)) {  // one item per bloom: each lane its own
            if (items) c += item_count(a.smv, x.base, x.k, 0, t);
            continue;
        }
        uint32_t const incl = wave_incl_scan(items);
        uint32_t const total = __shfl(incl, kWave - 1, kWave);
        __builtin_amdgcn_wave_barrier();
        s_incl[lane] = incl;
        s_base[lane] = x.base;
        s_k[lane] = x.k;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
            uint32_t const pos = t0 + lane;
            if (pos < total) {
                // the incidence of item pos: first prefix above pos
                uint32_t j = 0;
#pragma unroll
                for (uint32_t step = kWave / 2; step > 0; step >>= 1) {
                    uint32_t const k = j + step;
                    j = s_incl[k - 1] <= pos ? k : j;
                }
                uint32_t const first = j ? s_incl[j - 1] : 0u;
                c += item_count(a.smv, s_base[j], s_k[j], pos - first, t);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    return wave_sum_u64(c) - own;
}

// The sorted-chunk way for edge e (whole wave): fills the cache, estimates
// the cost of a count against the values one stream pass reads, counts at
// t = prev[e] and, while it stays the cheaper way, searches the h-index.
// Returns false (and leaves *h) when the stream passes should do it.
__device__ bool hidx_sorted(const HidxArgs& a, uint32_t e, uint64_t q0, uint64_t q1, uint32_t pe, uint32_t* cache,
                            uint32_t* s_incl, uint32_t* s_base, uint32_t* s_k, uint64_t* h) {
    int const lane = threadIdx.x & (kWave - 1);
    uint64_t cost = 0, vals = 0;
    for (uint64_t b0 = q0; b0 < q1; b0 += kWave) {
        uint64_t const q = b0 + lane;
        if (q < q1) {
            IncInfo const x = inc_info(a, e, q);
            uint64_t const i = q - q0;
            if (i < kCacheInc) {
                cache[i] = x.pp;
                cache[kCacheInc + i] = x.base;
                cache[2 * kCacheInc + i] = x.k;
            }
            if (x.pp) {
                vals += x.k - 1;
                // a binary search per item: about log2 of its length, plus one
                cost += x.k <= kSmallBloom ? uint64_t(33 - __builtin_clz(x.k))
                                           : uint64_t((x.k + kChunk - 1) / kChunk) * kSearchCost;
            }
        }
    }
    cost = wave_sum_u64(cost);
    vals = wave_sum_u64(vals);
    __builtin_amdgcn_wave_barrier();
    // the stream passes: one up to kHB - 1, two up to kHB^2 - 1, else three
    uint64_t const passes = pe < kHB ? 1 : (uint64_t(pe) < uint64_t(kHB) * kHB ? 2 : 3);
    if (cost >= vals * passes) return false;
    uint64_t const c0 = count_at_least(a, e, q0, q1, pe, cache, s_incl, s_base, s_k);
    if (c0 >= pe) {
        *h = pe;
        return true;
    }
    // The count C(t) does not grow with t, so every count bounds the answer
    // from both sides: C(x) < x rules out t >= x and admits t <= C(x)
    // (C(t) >= C(x) >= t there); C(x) >= x admits x and rules out
    // t > C(x). The search starts from [C(pe), pe - 1].
    uint32_t lo = uint32_t(c0), hi = pe - 1;
    uint64_t const steps = 32 - __builtin_clz(hi - lo + 1);
    if (steps * cost >= vals * passes) return false;
    while (lo < hi) {
        uint32_t const mid = lo + (hi - lo + 1) / 2;
        uint64_t const cm = count_at_least(a, e, q0, q1, mid, cache, s_incl, s_base, s_k);
        if (cm >= mid) {
            lo = mid;
            hi = cm < hi ? uint32_t(cm) : hi;
        } else {
            hi = mid - 1;
            lo = cm > lo ? uint32_t(cm) : lo;
        }
    }
    *h = lo;
    return true;
}

// The heavy edges of a round (one workgroup): each edge taken from the work
// counter is evaluated by the whole workgroup, its waves sharing the
// incidences of every count, with the bloom data of its first kHeavyCache
// incidences cached in the workgroup's histograms (cache); the search is
// that of hidx_sorted. s_part holds kHW partial counts.
__device__ void hidx_heavy(const HidxArgs& a, uint32_t* cache, uint32_t* s_incl, uint32_t* s_base, uint32_t* s_k,
                           uint64_t* s_part) {
    __shared__ uint32_t s_item;
    int const lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    bool changed = false;
    for (;;) {
        if (threadIdx.x == 0) s_item = uint32_t(atomicAdd(a.heavy_next, 1ull));
        __syncthreads();
        uint32_t const i = s_item;
        __syncthreads();
        if (i >= a.n_heavy) break;
        uint32_t const e = a.heavy[i];
        uint64_t const pe = a.prev[e];
        if (pe == 0) {  // cpu_bitruss.cpp:128
            if (threadIdx.x == 0) a.curr[e] = 0;
            continue;
        }
        if (a.dirty_in) {
            bool const mk = a.emark[e] != 0;
            __syncthreads();
            if (!mk) {
                if (threadIdx.x == 0) a.curr[e] = pe;
                continue;
            }
            if (threadIdx.x == 0) a.emark[e] = 0;
        }
        uint64_t const q0 = a.inc_off[e], q1 = a.inc_off[e + 1];
        uint64_t const qc = std::min(q1, q0 + kHeavyCache + kHeavyGlobal);
        uint32_t* gcache = a.hcache + size_t(blockIdx.x) * 3 * kHeavyGlobal;
        for (uint64_t q = q0 + threadIdx.x; q < qc; q += kHT) {
            IncInfo const x = inc_info(a, e, q);
            uint64_t const i = q - q0;
            if (i < kHeavyCache) {
                cache[i] = x.pp;
                cache[kHeavyCache + i] = x.base;
                cache[2 * kHeavyCache + i] = x.k;
            } else {
                gcache[i - kHeavyCache] = x.pp;
                gcache[kHeavyGlobal + i - kHeavyCache] = x.base;
                gcache[2 * kHeavyGlobal + i - kHeavyCache] = x.k;
            }
        }
        __syncthreads();
        auto count = [&](uint32_t t) {
            uint64_t const c = count_at_least(a, e, q0, q1, t, cache, s_incl, s_base, s_k, uint64_t(wv) * kWave, kHT,
                                              kHeavyCache, gcache);
            if (lane == 0) s_part[wv] = c;
            __syncthreads();
            uint64_t tot = 0;
            for (int w = 0; w < kHW; w++) tot += s_part[w];
            __syncthreads();
            return tot;
        };
        uint64_t const c0 = count(uint32_t(pe));
        uint64_t h = pe;
        if (c0 < pe) {  // the two-sided search of hidx_sorted
            uint32_t lo = uint32_t(c0), hi = uint32_t(pe) - 1;
            while (lo < hi) {
                uint32_t const mid = lo + (hi - lo + 1) / 2;
                uint64_t const cm = count(mid);
                if (cm >= mid) {
                    lo = mid;
                    hi = cm < hi ? uint32_t(cm) : hi;
                } else {
                    hi = mid - 1;
                    lo = cm > lo ? uint32_t(cm) : lo;
                }
            }
            h = lo;
        }
        if (threadIdx.x == 0) a.curr[e] = h;
        if (h != pe) {
            changed = true;
            for (uint64_t q = q0 + threadIdx.x; q < q1; q += kHT) a.dirty_out[a.bid[a.inc[q]] - 1] = 1;
            if (a.egraph && threadIdx.x == 0)
                __hip_atomic_store(&a.glast[a.egraph[e]], a.round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (changed && threadIdx.x == 0) __hip_atomic_store(a.changed, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kHT) void k_hidx_round(HidxArgs a) {
    if (a.done) {
        if (*a.done) return;
        a.round = uint32_t(*a.dev_round);
        if (a.round == 1) a.dirty_in = nullptr;
    }
    __shared__ uint32_t s_hist[kHW][kHB];
    __shared__ uint32_t s_incl[kHW][kWave], s_base[kHW][kWave], s_skip[kHW][kWave];
    __shared__ uint64_t s_pp[kHW][kWave];
    int const lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    uint32_t* hist = s_hist[wv];
    bool changed = false;
    if (blockIdx.x < a.heavy_blocks) {
        hidx_heavy(a, &s_hist[0][0], s_incl[wv], s_base[wv], s_skip[wv], reinterpret_cast<uint64_t*>(s_pp));
        return;
    }
    uint32_t const nw = (gridDim.x - a.heavy_blocks) * kHW;
    uint32_t last_g = ~0u;  // graph whose change this wave last recorded
    for (uint32_t e = (blockIdx.x - a.heavy_blocks) * kHW + wv; e < a.m; e += nw) {
        uint64_t const q0 = a.inc_off[e], q1 = a.inc_off[e + 1];
        if (a.n_heavy && q1 - q0 > kHeavyInc && a.sup[e] <= 0xFFFFFFFFull) continue;  // a heavy edge
        uint64_t const pe = a.prev[e];
        if (pe == 0) {  // cpu_bitruss.cpp:128
            if (lane == 0) a.curr[e] = 0;
            continue;
        }
        // no bloom of e saw a change in the previous round: the values of
        // e's butterflies, and so its h-index, are those of that round
        if (a.dirty_in) {
            if (!a.emark[e]) {
                if (lane == 0) a.curr[e] = pe;
                continue;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) a.emark[e] = 0;
        }
        uint64_t h = 0;
        bool const sorted = a.smv && pe <= 0xFFFFFFFFull &&
                            hidx_sorted(a, e, q0, q1, uint32_t(pe), hist, s_incl[wv], s_base[wv], s_skip[wv], &h);
        // candidates [lo, hi]; `above` values lie above hi
        uint64_t lo = 0, hi = pe, above = 0;
        uint64_t width = (pe + kHB) / kHB;  // ceil((pe + 1) / kHB)
        for (; !sorted;) {
            for (uint32_t i = lane; i < kHB; i += kWave) hist[i] = 0;
            __builtin_amdgcn_wave_barrier();
            bin_values(a, e, q0, q1, lo, hi, width, pe, hist, s_incl[wv], s_base[wv], s_skip[wv], s_pp[wv]);
            __builtin_amdgcn_wave_barrier();
            // largest bin b whose lower end x_b = lo + b * width has
            // #{values >= x_b} >= x_b; lane L holds bins [L * per, (L+1) * per)
            uint32_t const nbins = uint32_t((hi - lo) / width + 1);
            uint32_t const per = (nbins + kWave - 1) / kWave;
            uint32_t const b_lo = lane * per, b_hi = std::min(nbins, b_lo + per);
            uint64_t mine = 0;
            for (uint32_t i = b_lo; i < b_hi; i++) mine += hist[i];
            // values in the bins of higher lanes
            uint64_t suf = mine;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                uint64_t const o = __shfl_down(suf, d, kWave);
                if (lane + d < kWave) suf += o;
            }
            uint64_t acc = above + (suf - mine);
            int found = -1;
            uint64_t found_above = 0;
            for (uint32_t i = b_hi; i-- > b_lo;) {
                uint64_t const above_i = acc;
                acc += hist[i];
                if (acc >= lo + uint64_t(i) * width) {
                    found = int(i);
                    found_above = above_i;
                    break;
                }
            }
            // the highest lane with a hit holds the answer's bin (bin 0 always
            // qualifies: its lower end lo has #{values >= lo} >= lo)
            uint64_t const hits = __ballot(found >= 0);
            int const src = hits ? 63 - __builtin_clzll(hits) : 0;
            uint32_t const b = uint32_t(__shfl(found, src, kWave));
            uint64_t const ab = __shfl(found_above, src, kWave);
            uint64_t const x = lo + uint64_t(b) * width;
            __builtin_amdgcn_wave_barrier();
            if (width == 1) {
                h = x;
                break;
            }
            above = ab;
            lo = x;
            hi = std::min(x + width - 1, hi);
            width = (hi - lo + kHB) / kHB;  // ceil((hi - lo + 1) / kHB)
        }
        if (lane == 0) a.curr[e] = h;
        if (h != pe) {
            changed = true;
            for (uint64_t q = q0 + lane; q < q1; q += kWave) a.dirty_out[a.bid[a.inc[q]] - 1] = 1;
            if (a.egraph && a.egraph[e] != last_g) {  // batches: per-graph round count
                last_g = a.egraph[e];
                if (lane == 0) __hip_atomic_store(&a.glast[last_g], a.round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    if (changed && lane == 0) __hip_atomic_store(a.changed, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

// The sorted chunks of a call: smv (null: every edge streams its values)
// and the list of chunks to sort each round.
struct SortedChunks {
    uint32_t* smv = nullptr;
    const uint64_t* list = nullptr;
    uint64_t n = 0;
    unsigned grid = 1;
    unsigned grid_small = 1;          // the small blooms' sort
    const uint32_t* heavy = nullptr;  // the heavy edges and their workgroups
    uint32_t n_heavy = 0, heavy_blocks = 0;
    uint32_t* hcache = nullptr;  // the workgroups' HBM rows of bloom data
};

// smv lives in the peel's claim stamps (bclaim, at least R u32) and the
// chunk list in its claimed-bloom lists (touched): the index build sizes
// both for every call and h-index rounds use neither otherwise.
static int hidx_sorted_setup(bt_dgraph* dg, SortedChunks* sc) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    bt_index& ix = c->ix;
    *sc = SortedChunks{};
    if (!c->hidx_sorted || ix.R == 0 || ix.NB == 0 || ix.bclaim.bytes < sizeof(uint32_t) * ix.R) return BT_OK;
    uint32_t const NB = uint32_t(ix.NB);
    auto* cnt = reinterpret_cast<unsigned long long*>(dg->ctl + 19);
    auto const* bstart = static_cast<const uint32_t*>(ix.bstart.ptr);
    unsigned const grid = grid_for(NB, 256, unsigned(c->cu_count) * 16);
    BT_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint64_t), s));
    k_hidx_chunks<<<grid, 256, 0, s>>>(NB, bstart, cnt, nullptr);
    BT_HIP_TRY(hipGetLastError());
    uint64_t* h = static_cast<uint64_t*>(c->h_pinned);
    BT_HIP_TRY(hipMemcpyAsync(h, cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    BT_HIP_TRY(hipStreamSynchronize(s));
    uint64_t const n = h[0];
    sc->smv = static_cast<uint32_t*>(ix.bclaim.ptr);
    sc->grid_small = grid_for(NB, 256, unsigned(c->cu_count) * 8);
    if (n) {
        BT_TRY(bt_buf_reserve(ix.touched, sizeof(uint64_t) * n));
        BT_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint64_t), s));
        k_hidx_chunks<<<grid, 256, 0, s>>>(NB, bstart, cnt, static_cast<uint64_t*>(ix.touched.ptr));
        BT_HIP_TRY(hipGetLastError());
        sc->list = static_cast<const uint64_t*>(ix.touched.ptr);
        sc->n = n;
        sc->grid = unsigned(std::min<uint64_t>(n, uint64_t(c->cu_count) * 16));
    }
    // the heavy edges, listed in the peel's per-block live-edge segments
    // (live, u32 per edge, unused by h-index rounds)
    uint32_t const m = dg->m;
    auto const* inc_off = static_cast<const uint64_t*>(ix.inc_off.ptr);
    unsigned const grid_m = grid_for(std::max<uint32_t>(m, 1), 256, unsigned(c->cu_count) * 16);
    BT_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint64_t), s));
    k_hidx_heavy_list<<<grid_m, 256, 0, s>>>(m, inc_off, dg->sup, cnt, nullptr);
    BT_HIP_TRY(hipGetLastError());
    BT_HIP_TRY(hipMemcpyAsync(h, cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    BT_HIP_TRY(hipStreamSynchronize(s));
    uint32_t const nh = uint32_t(h[0]);
    if (nh) {
        BT_TRY(bt_buf_reserve(ix.live, sizeof(uint32_t) * nh));
        BT_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint64_t), s));
        k_hidx_heavy_list<<<grid_m, 256, 0, s>>>(m, inc_off, dg->sup, cnt, static_cast<uint32_t*>(ix.live.ptr));
        BT_HIP_TRY(hipGetLastError());
        sc->heavy = static_cast<const uint32_t*>(ix.live.ptr);
        sc->n_heavy = nh;
        sc->heavy_blocks = std::min<uint32_t>(nh, uint32_t(c->cu_count) * 2);
        // the rows in the peel's sorted both-frontier lists (c2s, unused by
        // h-index rounds)
        BT_TRY(bt_buf_reserve(ix.c2s, sizeof(uint32_t) * 3 * size_t(kHeavyGlobal) * sc->heavy_blocks));
        sc->hcache = static_cast<uint32_t*>(ix.c2s.ptr);
    }
    return BT_OK;
}

// The rounds as a captured graph of K rounds (K even), launched until a
// round without a change; *rounds receives the rounds counted on the device.
static int hidx_graph_loop(bt_dgraph* dg, uint64_t* mv, const SortedChunks& sc, uint8_t* emark,
                           uint32_t* const dmax[2], unsigned long long* changed, uint8_t* const dirty[2], size_t nbm,
                           unsigned grid_mv, unsigned grid_e, uint32_t* rounds) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    bt_index& ix = c->ix;
    uint32_t const R = uint32_t(ix.R), m = dg->m;
    int const K = c->hidx_graph_rounds + (c->hidx_graph_rounds & 1);
    LoopState const st{reinterpret_cast<unsigned long long*>(dg->ctl + 17), changed,
                       reinterpret_cast<unsigned long long*>(dg->ctl + 18),
                       reinterpret_cast<unsigned long long*>(dg->ctl + 21)};
    k_hidx_loop_init<<<1, 64, 0, s>>>(st);
    BT_HIP_TRY(hipGetLastError());
    unsigned const grid_clear = grid_for(nbm, 256, unsigned(c->cu_count) * 8);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BT_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; k++) {
        // position k: round parity k & 1 (rounds start at 1 on position 0)
        uint64_t* prev = (k & 1) ? dg->aux : dg->res;
        uint64_t* curr = (k & 1) ? dg->res : dg->aux;
        uint8_t* out = dirty[(k + 1) & 1];  // round r writes dirty[r & 1]
        uint8_t* in = dirty[k & 1];         // and reads dirty[(r + 1) & 1]
        k_hidx_round_begin<<<1, 64, 0, s>>>(st);
        uint32_t* dm = dmax[(k + 1) & 1];  // round r's largest changed values: dmax[r & 1]
        k_hidx_clear<<<grid_clear, 256, 0, s>>>(out, dm, nbm, st.done);
        if (R)
            k_hidx_mv<<<grid_mv, 256, 0, s>>>(R, static_cast<const uint64_t*>(ix.vals1.ptr), prev, mv, sc.smv,
                                              static_cast<const uint32_t*>(ix.bid.ptr), in, dm, st.round, st.done);
        if (R)
            k_hidx_mark<<<grid_mv, 256, 0, s>>>(R, static_cast<const uint64_t*>(ix.vals1.ptr), prev,
                                                static_cast<const uint32_t*>(ix.bid.ptr), in, dm, emark, st.round,
                                                st.done);
        if (sc.n)
            k_hidx_sort<<<sc.grid, kST, 0, s>>>(sc.list, sc.n, static_cast<const uint32_t*>(ix.bstart.ptr), sc.smv,
                                                 in, st.round, st.done);
        if (sc.smv)
            k_hidx_sort_small<<<sc.grid_small, 256, 0, s>>>(uint32_t(ix.NB), static_cast<const uint32_t*>(ix.bstart.ptr),
                                                            sc.smv, in, st.round, st.done);
        HidxArgs a{m,
                   prev,
                   curr,
                   static_cast<const uint64_t*>(ix.inc_off.ptr),
                   static_cast<const uint32_t*>(ix.inc.ptr),
                   static_cast<const uint64_t*>(ix.vals1.ptr),
                   static_cast<const uint32_t*>(ix.bid.ptr),
                   static_cast<const uint32_t*>(ix.bstart.ptr),
                   mv,
                   sc.smv,
                   changed,
                   dg->egraph,
                   dg->graph_last,
                   0,
                   in,
                   out,
                   emark,
                   dg->sup,
                   sc.heavy,
                   sc.n_heavy,
                   sc.heavy_blocks,
                   st.work,
                   sc.hcache,
                   st.round,
                   st.done};
        k_hidx_round<<<grid_e + sc.heavy_blocks, kHT, 0, s>>>(a);
        k_hidx_round_end<<<1, 64, 0, s>>>(st);
    }
    hipError_t const ce = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        BT_HIP_TRY(ce);
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    BT_HIP_TRY(ie);
    uint64_t* h = static_cast<uint64_t*>(c->h_pinned);
    bool const progress = getenv("BT_PEEL_PROGRESS") != nullptr;
    auto const t_first = std::chrono::steady_clock::now();
    int st_out = BT_OK;
    for (;;) {
        hipError_t e = hipGraphLaunch(exec, s);
        if (e == hipSuccess) e = hipMemcpyAsync(h, dg->ctl + 17, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            bt_set_hip_error(e, "hidx graph rounds", __FILE__, __LINE__);
            st_out = BT_ERR_DEVICE;
            break;
        }
        if (progress) {
            double const sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_first).count();
            std::fprintf(stderr, "[bt_hidx] %llu rounds%s, %.1f s\n", (unsigned long long)h[0],
                         h[1] ? " (done)" : "", sec);
            std::fflush(stderr);
        }
        if (h[1]) break;
    }
    (void)hipGraphExecDestroy(exec);
    if (st_out == BT_OK) *rounds = uint32_t(h[0]);
    return st_out;
}

int bt_hidx_index(bt_dgraph* dg, uint32_t* rounds_out) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const m = dg->m;
    bt_timings const count_t = c->last;
    bt_timer tm(c);
    BT_TRY(tm.start(5));
    BT_TRY(bt_index_build(dg, true));
    bt_index& ix = c->ix;
    uint32_t const R = uint32_t(ix.R);
    // the reference runs one round even on an edgeless graph
    uint32_t rounds = m ? 0 : 1;
    if (m) {
        // mv reuses the unsorted record keys (R u64, free after the sort)
        BT_TRY(bt_buf_reserve(ix.keys0, sizeof(uint64_t) * std::max<uint32_t>(R, 1)));
        uint64_t* mv = static_cast<uint64_t*>(ix.keys0.ptr);
        BT_HIP_TRY(hipMemcpyAsync(dg->res, dg->sup, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, s));
        uint64_t* prev = dg->res;
        uint64_t* curr = dg->aux;
        unsigned long long* changed = reinterpret_cast<unsigned long long*>(dg->ctl + 16);
        unsigned long long* work = reinterpret_cast<unsigned long long*>(dg->ctl + 21);  // next heavy edge
        unsigned const grid_mv = grid_for(std::max<uint32_t>(R, 1), 256, unsigned(c->cu_count) * 16);
        unsigned const grid_e = std::max(1u, std::min(unsigned(c->cu_count) * 2, (m + kHW - 1) / kHW));
        uint64_t* h_changed = static_cast<uint64_t*>(c->h_pinned);
        // per-bloom change marks of the previous and the current round
        size_t const nbm = std::max<size_t>(ix.NB, 1);
        BT_TRY(bt_buf_reserve(ix.hdirty, 2 * nbm));
        uint8_t* dirty[2] = {static_cast<uint8_t*>(ix.hdirty.ptr), static_cast<uint8_t*>(ix.hdirty.ptr) + nbm};
        SortedChunks sc;
        BT_TRY(hidx_sorted_setup(dg, &sc));
        // per-edge marks in the index build's fill cursors (u32 per edge + 1,
        // free after the build)
        uint8_t* emark = static_cast<uint8_t*>(ix.inc_cur.ptr);
        BT_HIP_TRY(hipMemsetAsync(emark, 0, m, s));
        // per bloom and round parity: the largest old mv among the records
        // whose mv changed, in the peel's live-middle counts (bk, u32 per
        // record; a bloom has at least two records)
        BT_TRY(bt_buf_reserve(ix.bk, sizeof(uint32_t) * 2 * nbm));
        uint32_t* dmax[2] = {static_cast<uint32_t*>(ix.bk.ptr), static_cast<uint32_t*>(ix.bk.ptr) + nbm};
        bool const progress = getenv("BT_PEEL_PROGRESS") != nullptr;
        auto const t_first = std::chrono::steady_clock::now();
        if (c->hidx_graph_rounds > 0) {
            BT_TRY(hidx_graph_loop(dg, mv, sc, emark, dmax, changed, dirty, nbm, grid_mv, grid_e, &rounds));
            prev = dg->res;  // the last round left both buffers equal
        }
        while (c->hidx_graph_rounds == 0) {
            rounds++;
            BT_HIP_TRY(hipMemsetAsync(changed, 0, sizeof(uint64_t), s));
            BT_HIP_TRY(hipMemsetAsync(work, 0, sizeof(uint64_t), s));
            BT_HIP_TRY(hipMemsetAsync(dirty[rounds & 1], 0, nbm, s));
            BT_HIP_TRY(hipMemsetAsync(dmax[rounds & 1], 0, sizeof(uint32_t) * nbm, s));
            uint8_t const* in = rounds > 1 ? dirty[(rounds + 1) & 1] : nullptr;
            if (R)
                k_hidx_mv<<<grid_mv, 256, 0, s>>>(R, static_cast<const uint64_t*>(ix.vals1.ptr), prev, mv, sc.smv,
                                                  static_cast<const uint32_t*>(ix.bid.ptr), in, dmax[rounds & 1],
                                                  nullptr, nullptr);
            if (R)
                k_hidx_mark<<<grid_mv, 256, 0, s>>>(R, static_cast<const uint64_t*>(ix.vals1.ptr), prev,
                                                    static_cast<const uint32_t*>(ix.bid.ptr), in, dmax[rounds & 1],
                                                    emark, nullptr, nullptr);
            if (sc.n)
                k_hidx_sort<<<sc.grid, kST, 0, s>>>(sc.list, sc.n, static_cast<const uint32_t*>(ix.bstart.ptr),
                                                     sc.smv, in, nullptr, nullptr);
            if (sc.smv)
                k_hidx_sort_small<<<sc.grid_small, 256, 0, s>>>(
                    uint32_t(ix.NB), static_cast<const uint32_t*>(ix.bstart.ptr), sc.smv, in, nullptr, nullptr);
            HidxArgs a{m,
                       prev,
                       curr,
                       static_cast<const uint64_t*>(ix.inc_off.ptr),
                       static_cast<const uint32_t*>(ix.inc.ptr),
                       static_cast<const uint64_t*>(ix.vals1.ptr),
                       static_cast<const uint32_t*>(ix.bid.ptr),
                       static_cast<const uint32_t*>(ix.bstart.ptr),
                       mv,
                       sc.smv,
                       changed,
                       dg->egraph,
                       dg->graph_last,
                       rounds,
                       in,
                       dirty[rounds & 1],
                       emark,
                       dg->sup,
                       sc.heavy,
                       sc.n_heavy,
                       sc.heavy_blocks,
                       work,
                       sc.hcache,
                       nullptr,
                       nullptr};
            k_hidx_round<<<grid_e + sc.heavy_blocks, kHT, 0, s>>>(a);
            BT_HIP_TRY(hipGetLastError());
            BT_HIP_TRY(hipMemcpyAsync(h_changed, changed, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
            BT_HIP_TRY(hipStreamSynchronize(s));
            std::swap(prev, curr);
            if (progress) {
                double const sec =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t_first).count();
                std::fprintf(stderr, "[bt_hidx] round %u: %s, %.1f s\n", rounds,
                             *h_changed ? "values changed" : "no change", sec);
                std::fflush(stderr);
            }
            if (!*h_changed) break;
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
