// GBiD (msp) peeling over the bloom index in one persistent launch.
//
// Reference semantics: c_bitruss_msp / bitruss_msp (cpu_bitruss.cpp:536-597,
// bitruss_msp.cu:440-583) with the per-pivot attribution rule of
// cpu_bitruss.cpp:301-351. Per butterfly that rule amounts to (derived by
// placing a surviving edge in each of the three pivot roles):
//   1 frontier edge  -> the 3 survivors lose 1
//   2 frontier edges -> the 2 survivors lose 1
//   3 frontier edges -> the survivor f loses 1 iff (o<q<p) or (p<o<q) or
//                       (q<o<p), o the frontier edge opposite f, p the one
//                       sharing f's upper vertex, q the one sharing f's
//                       lower vertex
// Each butterfly lies in exactly one bloom, so a sub-round visits every
// bloom that holds a frontier edge once and applies the rule to whole
// groups of middles at a time (middles with 0/1/2 frontier edges: n0/n1/n2):
//   a middle with no frontier edge: both edges lose n1 + n2
//   a middle with one frontier edge: its survivor loses n0 + (n1 - 1) plus
//     the 3-frontier-edge rule over the n2 middles with two frontier edges
// after which only the n0 middles stay alive in the bloom (compaction).
//
// Device state (no clamping, no add-back): supports only decrease; the one
// decrement that takes an edge from above the level to at most the level
// puts it in the next sub-round (stamp = sub-round + 1) and fixes its result
// to the level; that is the reference's max(level, s - d) and `next` set.
// Edge stamps: 0 live, 1 support zero, X = frontier of sub-round X; an
// edge with 0 < stamp < X is done during sub-round X. A sub-round whose
// frontier claims no bloom changes nothing and is skipped (the stamp
// counter then advances by two so those edges count as done).
//
// Control: one cooperative launch, one block per CU; phases are separated
// by a grid barrier, and after each barrier every block reads the same
// per-block words and takes the same decision. Hot words are not written
// by many blocks (the two shared live-edge lists take one atomic per
// 2,048-edge chunk):
//   - the barrier: each block stores its arrival epoch in its own slot,
//     block 0 watches all slots (one thread each) and releases with one
//     store (about 1.6 us for 256 blocks; a shared counter serialises the
//     256 arrivals at about 12 ns each);
//   - claimed blooms: each block appends to its own list and publishes its
//     length; readers locate entry t by a search over the block prefix
//     (only a block that overflows its list reserves in a shared one);
//   - the next level: per-block minima (kept supports, decremented
//     supports) in private slots, reduced by every block at a level's end.
// The blooms a frontier edge belongs to are claimed by the whole block at
// the end of a phase (its incidence lists concatenated and walked by 512
// lanes), so a hub edge's long list does not serialise one wave; likewise a
// bloom with more than 1,024 middles is walked by the whole block.
// Time slices: the launch returns at the first level boundary after its
// slice and is launched again to resume from the place it left in ctl.
// Coherence across the 8 XCDs without cache-wide fences: every word that
// one block writes and another reads later is accessed with relaxed
// device-scope atomics, and a barrier first waits for the block's own
// memory operations (s_waitcnt 0). Read-only index data uses plain loads.
#include <hip/hip_runtime.h>

#include <chrono>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

namespace {

constexpr int kMB = 512;                // block size (one block per CU)
constexpr int kWPB = kMB / kWave;       // waves per block
constexpr int kC2Cap = 128;             // both-frontier middles kept in LDS per wave
constexpr int kClaimUnroll = 4;         // incidences per lane per claim step
static_assert(kMB * kClaimUnroll == int(kClaimDummiesPerBlock), "dummy claim stamps per block");
constexpr int kXCap = 2048;             // frontier edges buffered per wave
constexpr int kBL = 1024;               // frontier edges per block claim batch
constexpr uint32_t kBlkCap = 16384;     // claimed blooms a block keeps in its own list
constexpr uint32_t kBigCap = 64;        // large blooms a block walks together per sub-round
constexpr uint32_t kWindowEdges = 4u << 20;  // live edges from which the collect uses a window
constexpr int kSp = 32;                 // u32 spacing of the barrier slots (128 B)
constexpr uint64_t kInf = ~0ull;
// Attribution of a butterfly with three frontier edges to its survivor f
// (o the frontier edge opposite f, p the one sharing f's upper vertex, q the
// one sharing f's lower vertex):
//   kRuleMsp    bitruss_msp (cpu_bitruss.cpp:301-351): f loses 1 iff
//               (o<q<p) or (p<o<q) or (q<o<p)
//   kRuleMspp   bitruss_mspp (bitruss_msp.cu:600-722): the same, except that
//               process_edge (:366-378) turned the pivot o around (its upper
//               degree above 10x its lower one), which exchanges p and q in
//               o's own condition: (o<p<q) or (p<o<q) or (q<o<p)
//   kRuleExact  the textbook peel (bitruss_peeling_cpu, cpu_bitruss.cpp:
//               34-122): every destroyed butterfly costs each survivor 1, so
//               f loses 1; the levels then give the true bitruss numbers
constexpr uint32_t kRuleMsp = 0, kRuleMspp = 1, kRuleExact = 2;
constexpr uint64_t kBarrierLimit = 30ull * 100000000ull;  // 30 s of the 100 MHz clock
constexpr unsigned kRelErr = 0x80000000u;                   // release word: error bit
constexpr uint32_t kSpan = 1024;        // graphs of a block's segment with an LDS minimum

struct MspCtl {
    unsigned rel;                    // barrier release epoch | kRelErr
    unsigned err;
    unsigned ovf_n[4];               // shared-list claims by sub-round % 3
    unsigned pad[2];
    unsigned long long subrounds;
    unsigned long long levels;
    unsigned long long level_max;
    // diagnostics (block 0's clock, 100 MHz ticks)
    unsigned long long t_collect, t_peel, t_work_collect, t_work_peel;
    unsigned long long pmax[2], pmax_sum;
    unsigned long long hist[24];
    // time slices: block 0 sets `yield` before each peel barrier once the
    // launch has run its slice; at the next level boundary every block sees
    // it, and block 0 leaves the place to resume from (sub-round, level
    // step, level, barrier epoch)
    unsigned long long yield, yielded;
    unsigned long long r_X, r_k, r_L, r_epoch;
    unsigned long long t_launch;     // block 0's clock at the launch
    unsigned long long sorted;       // bloom visits that sorted both-frontier lists
    unsigned long long live_n[2];    // plain graphs: lengths of the two collect lists
    unsigned long long high_n[2];    // plain graphs: lengths of the two lists above the window
    unsigned long long hp, W;        // current list above the window, the window's top
    unsigned long long partial;      // stopped after max_level with live edges left
    // diagnostics: per peel phase the slowest block's claim step and whole
    // work (max over blocks, summed by block 0), middles walked by waves
    // and by whole blocks, large-bloom visits
    unsigned long long cmax[2], cmax_sum, wmax[2], wmax_sum, mid_wave, mid_block, n_block;
};
static_assert(sizeof(MspCtl) <= 128 * sizeof(uint64_t), "control block overlays dg->ctl");

struct MspArgs {
    uint32_t m;
    uint32_t R;            // records (bloom middles)
    uint32_t nbloom;       // capacity of one shared claim list
    uint32_t diag;         // record per-phase diagnostics
    uint64_t* sup;
    uint64_t* res;
    uint32_t* stamp;
    const uint64_t* inc_off;
    const uint32_t* inc;   // per incidence: first record of the bloom
    uint64_t* rec;         // middles, compacted in place
    uint32_t* hdr;         // at a bloom's first record: live middles | upper << 31
    uint32_t* claim;       // at a bloom's first record: last claiming sub-round;
                           // then G * kMB * kClaimUnroll dummy stamps
    uint32_t* area;        // [3][G][kBlkCap] per-block claim lists by sub-round % 3
    uint32_t* ovf;         // [3][nbloom] shared claim lists (block list overflow)
    uint32_t* cnt;         // [3][G] per-block claim-list lengths
    uint64_t* slots;       // [2 level parities][2][G]: min kept, min decremented support
    uint32_t* arrive;      // [G * kSp] barrier arrival epochs
    uint32_t* live;
    uint32_t* live_cnt;
    MspCtl* ctl;
    // batches (the graph is a disjoint union; each graph peels its own levels)
    uint32_t ngraph;
    const uint32_t* egraph;  // m: graph of each edge
    const uint32_t* bgraph;  // at a bloom's first record: its graph
    uint64_t* glev;          // [2][ngraph] per-graph level of the current / next step
    uint32_t* any;           // per block: live edges left
    // time slices (plain graphs): a launch returns at the first level
    // boundary after slice_ticks of the 100 MHz clock (0: never), and a
    // launch with resume set continues from the place left in ctl
    uint64_t slice_ticks;
    uint32_t resume;
    // per wave: scratch of 3 * (largest bloom) u32 for the sorted lists of
    // both-frontier middles, and the bits of the largest edge id
    uint32_t* c2s;
    uint32_t c2s_stride;
    int id_bits;
    uint32_t c2cap;        // both-frontier middles summed from the LDS list (<= kC2Cap)
    uint32_t window_min;   // live edges from which a split opens a window
    uint64_t max_level;    // plain graphs: stop after this level (0: never)
    uint32_t big_min;      // blooms with more middles are walked by the whole block
    // attribution rule for a butterfly with three frontier edges (kRule*);
    // mspp reads the orientation flag of each edge from flipbits
    uint32_t rule;
    const uint32_t* flipbits;
};

// relaxed device-scope accesses (coherent across XCDs)
template <class T>
__device__ __forceinline__ T ald(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ void ast(T* p, U v) {
    __hip_atomic_store(p, T(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T afetch_add(T* p, U v) {
    return __hip_atomic_fetch_add(p, T(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T afetch_max(T* p, U v) {
    return __hip_atomic_fetch_max(p, T(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until this wave's loads, stores and atomics have completed.
__device__ __forceinline__ void drain() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << (threadIdx.x & (kWave - 1))) - 1ull;
}

// Per-block min reduction; every thread gets the result.
__device__ __forceinline__ uint64_t block_min_all(uint64_t v, uint64_t* s_red) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        uint64_t const o = __shfl_xor(v, d, kWave);
        v = o < v ? o : v;
    }
    int const lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) s_red[wid] = v;
    __syncthreads();
    uint64_t r = kInf;
    for (int i = 0; i < kWPB; i++) r = s_red[i] < r ? s_red[i] : r;
    __syncthreads();
    return r;
}

// Grid barrier for a cooperative launch (every block calls it with the same
// epoch sequence 1, 2, ...). Waits are bounded by kBarrierLimit: a timeout
// sets the error word and the error bit of the release, and every block
// returns false, so a broken invariant ends the kernel instead of hanging.
__device__ __forceinline__ bool grid_sync(const MspArgs& a, unsigned G, unsigned epoch, int* s_ok) {
    MspCtl* ctl = a.ctl;
    drain();
    __syncthreads();
    if (blockIdx.x == 0) {
        int ok = 1;
        uint32_t const t = threadIdx.x;
        if (t != 0 && t < G) {
            uint64_t const t0 = wall_clock64();
            while (ald(&a.arrive[t * kSp]) != epoch) {
                if (wall_clock64() - t0 > kBarrierLimit) {
                    ok = 0;
                    break;
                }
            }
        }
        ok = __syncthreads_and(ok);
        if (t == 0) {
            if (!ok) ast(&ctl->err, 1u);
            ast(&ctl->rel, epoch | (ok ? 0u : kRelErr));
        }
        return ok != 0;
    }
    if (threadIdx.x == 0) {
        ast(&a.arrive[blockIdx.x * kSp], epoch);
        int ok = 1;
        uint64_t const t0 = wall_clock64();
        unsigned r;
        while (((r = ald(&ctl->rel)) & ~kRelErr) != epoch) {
            __builtin_amdgcn_s_sleep(1);
            if (wall_clock64() - t0 > kBarrierLimit) {
                ast(&ctl->err, 1u);
                ok = 0;
                break;
            }
        }
        *s_ok = ok && !(r & kRelErr);
    }
    __syncthreads();
    return *s_ok != 0;
}

// Largest j in [0, n) with a[j] <= q (a non-decreasing, a[0] <= q),
// branch-free over LDS; N is a power of two >= n.
template <int N>
__device__ __forceinline__ uint32_t last_le(const uint32_t* a, uint32_t n, uint32_t q) {
    uint32_t j = 0;
#pragma unroll
    for (uint32_t step = N / 2; step > 0; step >>= 1) {
        uint32_t const k = j + step;
        uint32_t const v = a[k < n ? k : n - 1];
        j = (k < n && v <= q) ? k : j;
    }
    return j;
}

// Claim stamp index for a lane with no incidence in a claim step: one per
// (block, thread, unroll slot) past the R real stamps.
__device__ __forceinline__ uint32_t dummy_claim(const MspArgs& a, int u) {
    return a.R + (blockIdx.x * kMB + threadIdx.x) * kClaimUnroll + uint32_t(u);
}

// Claim list of one sub-round Y being filled by this block: its own list
// first, the shared list beyond kBlkCap.
struct Out {
    uint32_t* own;        // area + (Y % 3, B)
    uint32_t* shared;     // ovf + Y % 3
    unsigned* shared_n;   // ctl->ovf_n + Y % 3
    unsigned* cnt;        // LDS counter
    // appends, for one wave, the entries v[u] of the lanes with app[u]
    template <int U>
    __device__ __forceinline__ void append(const bool* app, const uint32_t* v) const {
        int const lane = threadIdx.x & (kWave - 1);
        uint64_t bal[U];
        unsigned tot = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            bal[u] = __ballot(app[u]);
            tot += unsigned(__popcll(bal[u]));
        }
        if (!tot) return;
        unsigned base = 0, gbase = 0;
        if (lane == 0) {
            base = atomicAdd(cnt, tot);
            unsigned const over_start = base > kBlkCap ? base : kBlkCap;
            if (base + tot > kBlkCap) gbase = afetch_add(shared_n, base + tot - over_start);
        }
        base = __shfl(base, 0, kWave);
        gbase = __shfl(gbase, 0, kWave);
        unsigned const over_start = base > kBlkCap ? base : kBlkCap;
        unsigned before = base;
#pragma unroll
        for (int u = 0; u < U; u++) {
            unsigned const pos = before + unsigned(__popcll(bal[u] & lanemask_lt()));
            if (app[u]) {
                if (pos < kBlkCap) ast(&own[pos], v[u]);
                else ast(&shared[gbase + (pos - over_start)], v[u]);
            }
            before += unsigned(__popcll(bal[u]));
        }
    }
};

// Per-wave LDS scratch for flattening the incidence lists of a batch of
// up to 64 edges.
struct WaveScan {
    uint32_t* incl;  // inclusive prefix of list lengths
    uint64_t* off;   // incidence index minus flattened position
};

// Claims, for sub-round Y, every bloom that holds one of the n edges in xl
// (LDS), with one wave: used when a wave's frontier buffer fills up before
// the block claims together. Branch-free, so that all loads and then all
// atomics of a step are in flight together: lanes past the end claim a
// private dummy stamp instead (the same address for all of them would
// serialise their atomics) and append nothing. Of all the claims of one
// bloom for sub-round Y exactly one finds the stamp below Y and appends it.
__device__ __forceinline__ void claim_edges(const uint32_t* xl, uint32_t n, uint32_t Y, const MspArgs& a,
                                            const Out& out, const WaveScan& ws) {
    int const lane = threadIdx.x & (kWave - 1);
    for (uint32_t b0 = 0; b0 < n; b0 += kWave) {
        uint32_t const cnt = n - b0 < uint32_t(kWave) ? n - b0 : uint32_t(kWave);
        uint64_t p0 = 0;
        uint32_t len = 0;
        if (uint32_t(lane) < cnt) {
            uint32_t const e = xl[b0 + lane];
            p0 = a.inc_off[e];
            len = uint32_t(a.inc_off[e + 1] - p0);
        }
        uint32_t const incl = wave_incl_scan(len);
        uint32_t const total = __shfl(incl, kWave - 1, kWave);
        __builtin_amdgcn_wave_barrier();
        ws.incl[lane] = incl;
        ws.off[lane] = p0 - (incl - len);
        __builtin_amdgcn_wave_barrier();
        for (uint32_t q0 = 0; q0 < total; q0 += kWave * kClaimUnroll) {
            uint32_t sb[kClaimUnroll], sc[kClaimUnroll];
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) {
                uint32_t const qr = q0 + u * kWave + lane;
                uint32_t const q = qr < total ? qr : total - 1;
                // the first edge whose prefix passes q: one past the last <= q
                uint32_t const j = ws.incl[0] > q ? 0u : last_le<kWave>(ws.incl, cnt, q) + 1;
                sb[u] = a.inc[q + ws.off[j]];
                sc[u] = qr < total ? sb[u] : dummy_claim(a, u);
            }
            uint32_t old[kClaimUnroll];
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) old[u] = afetch_max(&a.claim[sc[u]], Y);
            bool app[kClaimUnroll];
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) app[u] = q0 + u * kWave + lane < total && old[u] < Y;
            out.append<kClaimUnroll>(app, sb);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Per-wave buffer of frontier edges of sub-round Y whose blooms the block
// claims at the end of the phase. Room is made before a batch of puts (the
// wave claims a buffer that could overflow itself), so that the puts carry
// no claim code: every inlined copy of it would grow the kernel, whose
// instruction fetches are a cost of their own.
struct CrossBuf {
    uint32_t* xl;
    uint32_t n;
    WaveScan ws;
    // after this, `room` more edges fit (room <= kXCap)
    __device__ __forceinline__ void ensure(uint32_t room, uint32_t Y, const MspArgs& a, const Out& out) {
        if (n + room > uint32_t(kXCap)) {
            __builtin_amdgcn_wave_barrier();
            claim_edges(xl, n, Y, a, out, ws);
            n = 0;
        }
    }
    __device__ __forceinline__ void put(bool has, uint32_t e) {
        uint64_t const b = __ballot(has);
        if (has) xl[n + __popcll(b & lanemask_lt())] = e;
        n += __popcll(b);
    }
};

// The whole block claims, for sub-round Y, the blooms of the frontier edges
// left in the waves' buffers (s_xl[w][0, s_xn[w])): their incidence lists
// are concatenated (prefix in s_incl, index offset in s_off) and walked by
// all lanes. All threads call, after a __syncthreads that follows the
// waves' last pushes.
__device__ __forceinline__ void claim_block(const MspArgs& a, uint32_t Y, const Out& out,
                                            uint32_t (*s_xl)[kXCap], const uint32_t* s_xn, uint32_t* s_incl,
                                            uint64_t* s_off, uint32_t* s_part) {
    uint32_t const tid = threadIdx.x;
    uint32_t pre[kWPB + 1];
    pre[0] = 0;
#pragma unroll
    for (int w = 0; w < kWPB; w++) pre[w + 1] = pre[w] + s_xn[w];
    uint32_t const n_all = pre[kWPB];
    // batches of kBL edges; thread t handles entries 2t and 2t + 1 of a batch
    for (uint32_t b0 = 0; b0 < n_all; b0 += kBL) {
        uint32_t const n = n_all - b0 < uint32_t(kBL) ? n_all - b0 : uint32_t(kBL);
        uint64_t p0[2] = {0, 0};
        uint32_t len[2] = {0, 0};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t const j = 2 * tid + h;
            if (j < n) {
                uint32_t const jj = b0 + j;
                int w = 0;
#pragma unroll
                for (int i = 1; i < kWPB; i++) w += jj >= pre[i] ? 1 : 0;
                uint32_t const e = s_xl[w][jj - pre[w]];
                p0[h] = a.inc_off[e];
                len[h] = uint32_t(a.inc_off[e + 1] - p0[h]);
            }
        }
        uint32_t total = 0;
        uint32_t const ex = block_excl_scan<kMB>(len[0] + len[1], s_part, total);
        if (2 * tid < n) {
            s_incl[2 * tid] = ex + len[0];
            s_off[2 * tid] = p0[0] - ex;
        }
        if (2 * tid + 1 < n) {
            s_incl[2 * tid + 1] = ex + len[0] + len[1];
            s_off[2 * tid + 1] = p0[1] - (ex + len[0]);
        }
        __syncthreads();
        for (uint32_t q0 = 0; q0 < total; q0 += kMB * kClaimUnroll) {
            uint32_t sb[kClaimUnroll], sc[kClaimUnroll];
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) {
                uint32_t const qr = q0 + u * kMB + tid;
                uint32_t const q = qr < total ? qr : total - 1;
                uint32_t const j = s_incl[0] > q ? 0u : last_le<kBL>(s_incl, n, q) + 1;
                sb[u] = a.inc[q + s_off[j]];
                sc[u] = qr < total ? sb[u] : dummy_claim(a, u);
            }
            uint32_t old[kClaimUnroll];
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) old[u] = afetch_max(&a.claim[sc[u]], Y);
            bool app[kClaimUnroll];
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) app[u] = q0 + u * kMB + tid < total && old[u] < Y;
            out.append<kClaimUnroll>(app, sb);
        }
        __syncthreads();
    }
}

// the next level's collect list and its length (plain graphs), for the
// rare appends from Dec::settle; in LDS so that no register holds them
__shared__ uint32_t* s_low_list;
__shared__ unsigned long long* s_low_len;

struct Dec {
    uint64_t* sup;
    uint32_t* stamp;
    uint64_t* res;
    uint64_t L;
    uint32_t X;
    uint64_t tmin;  // smallest decremented support still above the level
    // plain graphs: an edge whose support enters the collect window
    // (<= W) joins the next level's collect list
    uint64_t W;
    // subtracts d from the support of x (atomic, returns the old support)
    __device__ __forceinline__ uint64_t issue(uint32_t x, uint64_t d) const {
        return afetch_add(&sup[x], uint64_t(0) - d);
    }
    // after issue: true when x fell to the level with this decrement
    __device__ __forceinline__ bool settle(uint32_t x, uint64_t d, uint64_t old) {
        uint64_t const nw = old - d;
        if (old > L && nw <= L) {
            ast(&stamp[x], X + 1);
            res[x] = L;
            return true;
        }
        if (nw > L && nw < tmin) tmin = nw;
        if (old > W && nw <= W) {
            uint32_t const pos = uint32_t(afetch_add(s_low_len, 1ull));
            ast(&s_low_list[pos], x);
        }
        return false;
    }
    __device__ __forceinline__ bool operator()(uint32_t x, uint64_t d) { return settle(x, d, issue(x, d)); }
};

__device__ __forceinline__ int classify(uint32_t s1, uint32_t s2, uint32_t X) {
    bool const dead = (s1 != 0 && s1 < X) || (s2 != 0 && s2 < X);
    if (dead) return -1;
    return int(s1 == X) + int(s2 == X);
}

__device__ __forceinline__ bool rule3(uint32_t o, uint32_t p, uint32_t q) {
    return (o < q && q < p) || (p < o && o < q) || (q < o && o < p);
}

__device__ __forceinline__ bool flipped(const MspArgs& a, uint32_t e) {
    return (a.flipbits[e >> 5] >> (e & 31u)) & 1u;
}

// Decrement owed by the survivor of a one-frontier-edge middle (g is the
// frontier edge of the same middle) for the both-frontier middles. Under
// kRuleMspp a turned-around o exchanges the roles of p and q, which is the
// rule of the other side.
__device__ __forceinline__ uint64_t rule_sum(const uint64_t* c2, uint32_t n2, bool froot, uint32_t up,
                                             uint32_t g, const MspArgs& a) {
    if (a.rule == kRuleExact) return n2;
    uint64_t d = 0;
    for (uint32_t t = 0; t < n2; t++) {
        uint64_t const rr = c2[t];
        uint32_t const o = froot ? uint32_t(rr) : uint32_t(rr >> 32);
        uint32_t const sr = froot ? uint32_t(rr >> 32) : uint32_t(rr);
        bool const u2 = (a.rule == kRuleMspp && flipped(a, o)) ? !up : bool(up);
        d += u2 ? rule3(o, sr, g) : rule3(o, g, sr);
    }
    return d;
}

// Orders the wave's global-memory accesses to its scratch before the next
// phase reads them from other lanes (the wave's CU holds the L1 they use).
__device__ __forceinline__ void scratch_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Stable LSD radix sort of n u32 keys below 2^bits in global memory, by one
// wave, 8 bits per pass with a 256-bin LDS histogram; tmp holds n keys.
__device__ __forceinline__ void wave_radix_sort(uint32_t* keys, uint32_t n, uint32_t* tmp, uint32_t* hist, int bits) {
    int const lane = threadIdx.x & (kWave - 1);
    uint32_t* src = keys;
    uint32_t* dst = tmp;
    for (int sh = 0; sh < bits; sh += 8) {
        for (int i = lane; i < 256; i += kWave) hist[i] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < n; i += kWave) atomicAdd(&hist[(src[i] >> sh) & 255u], 1u);
        __builtin_amdgcn_wave_barrier();
        // bins 4 * lane .. 4 * lane + 3 become their first output positions
        uint32_t h[4], sum = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            h[j] = hist[4 * lane + j];
            sum += h[j];
        }
        uint32_t pos = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            hist[4 * lane + j] = pos;
            pos += h[j];
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t b = 0; b < n; b += kWave) {
            uint32_t const i = b + lane;
            bool const v = i < n;
            uint32_t const key = v ? src[i] : 0u;
            uint32_t const d = (key >> sh) & 255u;
            // lanes with the same digit keep their order
            uint64_t peers = __ballot(v);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                uint64_t const m = __ballot((d >> t) & 1u);
                peers &= ((d >> t) & 1u) ? m : ~m;
            }
            uint32_t const rank = __popcll(peers & lanemask_lt());
            uint32_t const at = v ? hist[d] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (v) dst[at + rank] = key;
            if (v && rank == 0) hist[d] = at + uint32_t(__popcll(peers));
            __builtin_amdgcn_wave_barrier();
        }
        scratch_fence();
        uint32_t* const t = src;
        src = dst;
        dst = t;
    }
    if (src != keys) {
        for (uint32_t i = lane; i < n; i += kWave) keys[i] = src[i];
        scratch_fence();
    }
}

// Entries of the sorted list l[0, n) below g.
__device__ __forceinline__ uint32_t count_below(const uint32_t* l, uint32_t n, uint32_t g) {
    uint32_t pos = 0;
    for (uint32_t step = n ? (1u << (31 - __builtin_clz(n))) : 0u; step; step >>= 1)
        if (pos + step <= n && l[pos + step - 1] < g) pos += step;
    return pos;
}

// The both-frontier middles of a bloom as four sorted id lists in the
// wave's scratch: e1 and e2 of the middles with e1 < e2 (set P), then of
// those with e2 < e1 (set Q). rule_sum over them is then a few counts of
// entries below g (the middles of a bloom have distinct edges, so no entry
// equals g):
//   up,  froot:  |Q| - #(e1Q < g) + #(e2P < g)
//   up,  !froot: |P| - #(e2P < g) + #(e1Q < g)
//   !up, froot:  |Q| + #(e1Q < g) - #(e2Q < g) + #(e2P < g)
//   !up, !froot: |P| + #(e2P < g) - #(e1P < g) + #(e1Q < g)
// so a bloom with more both-frontier middles than the LDS list holds costs
// a sort and a few binary searches per one-frontier middle, not a pass
// over the bloom for each.
struct C2Lists {
    uint32_t *e1p, *e2p, *e1q, *e2q;
    uint32_t np, nq;
    uint32_t* tmp;  // sort buffer
    uint32_t cp, cq;
    __device__ __forceinline__ C2Lists(uint32_t* scratch, uint32_t np_, uint32_t nq_)
        : e1p(scratch), e2p(scratch + np_), e1q(scratch + 2 * np_), e2q(scratch + 2 * np_ + nq_), np(np_), nq(nq_),
          tmp(scratch + 2 * (np_ + nq_)), cp(0), cq(0) {}
    // appends middle r of the lanes where c2 holds (whole wave)
    __device__ __forceinline__ void put(bool c2, uint64_t r) {
        uint32_t const e1 = uint32_t(r >> 32), e2 = uint32_t(r);
        bool const p = c2 && e1 < e2, q = c2 && e2 < e1;
        uint64_t const bp = __ballot(p), bq = __ballot(q);
        if (p) {
            uint32_t const k = cp + __popcll(bp & lanemask_lt());
            e1p[k] = e1;
            e2p[k] = e2;
        }
        if (q) {
            uint32_t const k = cq + __popcll(bq & lanemask_lt());
            e1q[k] = e1;
            e2q[k] = e2;
        }
        cp += __popcll(bp);
        cq += __popcll(bq);
    }
    // after the puts (whole wave)
    __device__ __forceinline__ void sort(uint32_t* hist, int bits) {
        scratch_fence();
#pragma unroll 1
        for (int i = 0; i < 4; i++)
            wave_radix_sort(i == 0 ? e1p : i == 1 ? e2p : i == 2 ? e1q : e2q, i < 2 ? np : nq, tmp, hist, bits);
    }
    // kRuleMspp: the rule pairwise over the unsorted lists (the counts
    // above assume one orientation for every o)
    __device__ __forceinline__ uint64_t rule_pairs(bool froot, uint32_t up, uint32_t g, const MspArgs& a) const {
        uint64_t d = 0;
        for (int h = 0; h < 2; h++) {
            const uint32_t* e1 = h ? e1q : e1p;
            const uint32_t* e2 = h ? e2q : e2p;
            uint32_t const n = h ? cq : cp;
            for (uint32_t t = 0; t < n; t++) {
                uint32_t const o = froot ? e2[t] : e1[t];
                uint32_t const sr = froot ? e1[t] : e2[t];
                bool const u2 = flipped(a, o) ? !up : bool(up);
                d += u2 ? rule3(o, sr, g) : rule3(o, g, sr);
            }
        }
        return d;
    }
    __device__ __forceinline__ uint64_t rule(bool froot, uint32_t up, uint32_t g) const {
        if (up)
            return froot ? uint64_t(nq - count_below(e1q, nq, g)) + count_below(e2p, np, g)
                         : uint64_t(np - count_below(e2p, np, g)) + count_below(e1q, nq, g);
        return froot ? uint64_t(nq + count_below(e1q, nq, g) - count_below(e2q, nq, g)) + count_below(e2p, np, g)
                     : uint64_t(np + count_below(e2p, np, g) - count_below(e1p, np, g)) + count_below(e1q, nq, g);
    }
};

// scratch of the calling wave for its C2Lists (3 entries per bloom middle)
__device__ __forceinline__ uint32_t* wave_scratch(const MspArgs& a) {
    uint32_t const w = __builtin_amdgcn_readfirstlane(blockIdx.x * kWPB + threadIdx.x / kWave);
    return a.c2s + size_t(w) * a.c2s_stride;
}

// Processes a claimed bloom with kb <= 64 * RPT middles held in registers:
// all middles and their stamps are loaded up front, so the bloom costs a
// few dependent memory steps whatever its size. r0 holds middle `lane`.
template <int RPT>
__device__ __forceinline__ bool process_bloom(const MspArgs& a, uint32_t s, uint32_t kb, uint32_t up, uint64_t r0,
                                              uint32_t X, Dec& dec, uint64_t* c2, CrossBuf& xb,
                                              const Out& st) {
    int const lane = threadIdx.x & (kWave - 1);
    uint64_t* rec = a.rec + s;
    uint64_t r[RPT];
    uint32_t s1[RPT], s2[RPT];
    int c[RPT];
    r[0] = r0;
#pragma unroll
    for (int i = 1; i < RPT; i++) {
        uint32_t const j = i * kWave + lane;
        r[i] = j < kb ? ald(&rec[j]) : 0;
    }
    // every stamp load is in flight before the first is used
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        uint32_t const j = i * kWave + lane;
        s1[i] = j < kb ? ald(&a.stamp[uint32_t(r[i] >> 32)]) : 0u;
        s2[i] = j < kb ? ald(&a.stamp[uint32_t(r[i])]) : 0u;
    }
#pragma unroll
    for (int i = 0; i < RPT; i++) c[i] = (i * kWave + lane) < int(kb) ? classify(s1[i], s2[i], X) : -1;
    uint32_t n0 = 0, n1 = 0, n2 = 0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        uint64_t const b0 = __ballot(c[i] == 0), b1 = __ballot(c[i] == 1), b2 = __ballot(c[i] == 2);
        if (c[i] == 2) {
            uint32_t const pos = n2 + __popcll(b2 & lanemask_lt());
            if (pos < uint32_t(kC2Cap)) c2[pos] = r[i];
        }
        n0 += __popcll(b0);
        n1 += __popcll(b1);
        n2 += __popcll(b2);
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t const dec0 = n1 + n2;
    // more both-frontier middles than the LDS list holds, and middles that
    // need the rule: the streaming path (nothing is written yet)
    if (a.rule != kRuleExact && n2 > a.c2cap && n1) return false;
    // decrements: the first target is r-v (c = 0) or the survivor (c = 1),
    // the second is v-w (c = 0 only, by dec0)
    uint32_t t1[RPT], d1[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        uint32_t const e1 = uint32_t(r[i] >> 32), e2 = uint32_t(r[i]);
        t1[i] = e1;
        d1[i] = 0;
        if (c[i] == 0) {
            d1[i] = dec0;
        } else if (c[i] == 1) {
            bool const froot = s1[i] != X;  // the survivor is r-v
            uint32_t const g = froot ? e2 : e1;
            t1[i] = froot ? e1 : e2;
            d1[i] = n0 + (n1 - 1) +
                    uint32_t(rule_sum(c2, n2, froot, up, g, a));
        }
    }
    // the atomics of several rows (8, or 4 in the largest tier) are in
    // flight before the first result is used; more would not fit the
    // register budget
    constexpr int HB = RPT < 8 ? RPT : (RPT > 8 ? 4 : 8);
#pragma unroll
    for (int h = 0; h < RPT; h += HB) {
        uint64_t o1[HB], o2[HB];
#pragma unroll
        for (int i = 0; i < HB; i++) {
            o1[i] = d1[h + i] ? dec.issue(t1[h + i], d1[h + i]) : 0;
            o2[i] = (c[h + i] == 0 && dec0) ? dec.issue(uint32_t(r[h + i]), dec0) : 0;
        }
#pragma unroll
        for (int i = 0; i < HB; i++) {
            bool const x1 = d1[h + i] && dec.settle(t1[h + i], d1[h + i], o1[i]);
            bool const x2 = c[h + i] == 0 && dec0 && dec.settle(uint32_t(r[h + i]), dec0, o2[i]);
            xb.put(x1, t1[h + i]);
            xb.put(x2, uint32_t(r[h + i]));
        }
    }
    // compaction: only the middles without frontier edges stay
    if (n0 != kb) {
        uint32_t wpos = 0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            uint64_t const keepm = __ballot(c[i] == 0);
            if (c[i] == 0) ast(&rec[wpos + __popcll(keepm & lanemask_lt())], r[i]);
            wpos += __popcll(keepm);
        }
        if (lane == 0) {
            ast(&a.hdr[s], n0 | (up << 31));
            // fewer than two middles left: no butterfly remains, and the
            // top claim stamp keeps the bloom out of every later claim list
            if (n0 < 2) afetch_max(&a.claim[s], ~0u);
        }
    }
    return true;
}

// Blooms above 1024 middles: a counting pass, then a decrement pass, then
// (when the both-frontier list overflowed) a separate compaction pass.
__device__ __forceinline__ void process_large(const MspArgs& a, uint32_t s, uint32_t kb, uint32_t up, uint32_t X, Dec& dec,
                              uint64_t* c2, CrossBuf& xb, const Out& st) {
    int const lane = threadIdx.x & (kWave - 1);
    uint64_t* rec = a.rec + s;
    uint32_t n0 = 0, n1 = 0, n2 = 0, np = 0;
    for (uint32_t j0 = 0; j0 < kb; j0 += kWave) {
        uint32_t const j = j0 + lane;
        int c = -1;
        uint64_t r = 0;
        if (j < kb) {
            r = ald(&rec[j]);
            c = classify(ald(&a.stamp[uint32_t(r >> 32)]), ald(&a.stamp[uint32_t(r)]), X);
        }
        uint64_t const b0 = __ballot(c == 0), b1 = __ballot(c == 1), b2 = __ballot(c == 2);
        if (c == 2) {
            uint32_t const pos = n2 + __popcll(b2 & lanemask_lt());
            if (pos < uint32_t(kC2Cap)) c2[pos] = r;
        }
        n0 += __popcll(b0);
        n1 += __popcll(b1);
        n2 += __popcll(b2);
        np += __popcll(__ballot(c == 2 && uint32_t(r >> 32) < uint32_t(r)));
    }
    __builtin_amdgcn_wave_barrier();
    uint64_t const dec0 = uint64_t(n1) + n2;
    bool const c2ok = n2 <= a.c2cap || a.rule == kRuleExact;
    C2Lists lists(wave_scratch(a), np, n2 - np);
    if (!c2ok && n1) {
        for (uint32_t j0 = 0; j0 < kb; j0 += kWave) {
            uint32_t const j = j0 + lane;
            int c = -1;
            uint64_t r = 0;
            if (j < kb) {
                r = ald(&rec[j]);
                c = classify(ald(&a.stamp[uint32_t(r >> 32)]), ald(&a.stamp[uint32_t(r)]), X);
            }
            lists.put(c == 2, r);
        }
        if (a.rule == kRuleMsp) lists.sort(reinterpret_cast<uint32_t*>(c2), a.id_bits);
        else scratch_fence();
        if (lane == 0) afetch_add(&a.ctl->sorted, 1ull);
    }
    uint32_t wpos = 0;
    for (uint32_t j0 = 0; j0 < kb; j0 += kWave) {
        xb.ensure(2 * kWave, X + 1, a, st);
        uint32_t const j = j0 + lane;
        int c = -1;
        uint64_t r = 0;
        uint32_t e1 = 0, e2 = 0, s1 = 0;
        if (j < kb) {
            r = ald(&rec[j]);
            e1 = uint32_t(r >> 32);
            e2 = uint32_t(r);
            s1 = ald(&a.stamp[e1]);
            c = classify(s1, ald(&a.stamp[e2]), X);
        }
        bool x1 = false, x2 = false;
        uint32_t f = e1;
        if (c == 0 && dec0) {
            x1 = dec(e1, dec0);
            x2 = dec(e2, dec0);
        } else if (c == 1) {
            bool const froot = s1 != X;
            f = froot ? e1 : e2;
            uint32_t const g = froot ? e2 : e1;
            uint64_t const d =
                uint64_t(n0) + (n1 - 1) +
                (c2ok ? rule_sum(c2, n2, froot, up, g, a)
                      : a.rule == kRuleMsp ? lists.rule(froot, up, g) : lists.rule_pairs(froot, up, g, a));
            if (d) x1 = dec(f, d);
        }
        // compaction in place: the writes stay below the records read so far
        uint64_t const keepm = __ballot(c == 0);
        if (c == 0) ast(&rec[wpos + __popcll(keepm & lanemask_lt())], r);
        wpos += __popcll(keepm);
        xb.put(x1, f);
        xb.put(x2, e2);
    }
    if (lane == 0) {
        ast(&a.hdr[s], wpos | (up << 31));
        if (wpos < 2) afetch_max(&a.claim[s], ~0u);
    }
}

// kWindow (plain graphs with many edges): the collect scans a support
// window instead of every live edge; a separate instance keeps the code of
// smaller graphs free of it (register pressure).
template <bool kBatch, bool kWindow = false>
__global__ __launch_bounds__(kMB) void k_msp_persist(MspArgs a) {
    __shared__ uint64_t s_gmin[kBatch ? kSpan : 1];
    __shared__ uint32_t s_part[kWPB + 1];
    __shared__ uint64_t s_red[kWPB + 1];
    __shared__ uint64_t s_c2[kWPB][kC2Cap];
    __shared__ uint32_t s_xl[kWPB][kXCap];
    __shared__ uint32_t s_xn[kWPB];
    __shared__ uint32_t s_wsi[kWPB][kWave];
    __shared__ uint64_t s_wso[kWPB][kWave];
    __shared__ uint32_t s_incl[kBL];
    __shared__ uint64_t s_off[kBL];
    __shared__ uint32_t s_pfx[kMB + 1];
    __shared__ unsigned s_on;
    __shared__ int s_ok;
    __shared__ uint32_t s_gbase;
    __shared__ unsigned s_pick;
    __shared__ uint32_t s_big[kBigCap], s_big_bg[kBigCap];
    __shared__ unsigned s_nbig, s_bcnt[4];
    __shared__ uint64_t s_bc2[kC2Cap];
    __shared__ unsigned long long s_dmid[3];
    uint32_t const G = gridDim.x, B = blockIdx.x, tid = threadIdx.x;
    int const wv = tid / kWave;
    uint32_t const gw = B * kWPB + wv, nw = G * kWPB;
    uint32_t const seg0 = uint32_t(uint64_t(a.m) * B / G), seg1 = uint32_t(uint64_t(a.m) * (B + 1) / G);
    MspCtl* ctl = a.ctl;
    if (tid == 0) s_on = 0;
    WaveScan const ws{s_wsi[wv], s_wso[wv]};
    unsigned epoch = a.resume ? unsigned(ctl->r_epoch) : 0u;
    if (a.slice_ticks && B == 0 && tid == 0) ctl->t_launch = wall_clock64();
    // claim list of sub-round Y filled by this block
    auto out_for = [&](uint32_t Y) {
        return Out{a.area + (size_t(Y % 3) * G + B) * kBlkCap, a.ovf + size_t(Y % 3) * a.nbloom, &ctl->ovf_n[Y % 3],
                   &s_on};
    };
    // publishes this block's claim-list length for sub-round Y (whole block)
    auto close_out = [&](uint32_t Y) {
        __syncthreads();
        if (tid == 0) {
            ast(&a.cnt[size_t(Y % 3) * G + B], s_on < kBlkCap ? s_on : kBlkCap);
            s_on = 0;
        }
    };
    // after a barrier: prefix of the blocks' list lengths for sub-round Y in
    // s_pfx[0, G]; returns the number of claimed blooms (shared list included)
    auto open_list = [&](uint32_t Y) -> uint32_t {
        uint32_t const v = tid < G ? ald(&a.cnt[size_t(Y % 3) * G + tid]) : 0u;
        uint32_t const sh = ald(&ctl->ovf_n[Y % 3]);
        uint32_t total = 0;
        uint32_t const ex = block_excl_scan<kMB>(v, s_part, total);
        if (tid <= G) s_pfx[tid] = tid < G ? ex : total;
        __syncthreads();
        return total + sh;
    };
    // entry t of the claim list of sub-round Y (after open_list(Y))
    auto list_entry = [&](uint32_t Y, uint32_t t) -> uint32_t {
        uint32_t const own = s_pfx[G];
        if (t >= own) return ald(&a.ovf[size_t(Y % 3) * a.nbloom + (t - own)]);
        uint32_t const b = last_le<kMB>(s_pfx, G, t);
        return ald(&a.area[(size_t(Y % 3) * G + b) * kBlkCap + (t - s_pfx[b])]);
    };
    // minimum over the blocks' slots of level parity p (whole block)
    auto level_min = [&](uint32_t p) -> uint64_t {
        const uint64_t* sl = a.slots + size_t(p) * 2 * G;
        uint64_t v = kInf;
        if (tid < G) {
            uint64_t const x = ald(&sl[tid]), y = ald(&sl[G + tid]);
            v = x < y ? x : y;
        }
        return block_min_all(v, s_red);
    };
    auto set_slots = [&](uint32_t p, uint64_t kept, uint64_t dec) {
        if (tid == 0) {
            ast(&a.slots[size_t(p) * 2 * G + B], kept);
            ast(&a.slots[size_t(p) * 2 * G + G + B], dec);
        }
    };

    // ---- init: support-0 edges are final (process_support_zero); the rest
    // are live: in this block's segment of the live array (batches), or
    // appended to the list above the window (plain graphs: the first level
    // splits it; every level's collect spreads its list evenly over the
    // blocks); the first level is their min support (slots of parity 1, so
    // that level 0 can write parity 0)
    if (!a.resume) {
        uint32_t w = seg0;
        uint64_t mn = kInf;
        for (uint32_t base = seg0; base < seg1; base += kMB) {
            uint32_t const i = base + tid;
            bool keep = false;
            if (i < seg1) {
                uint64_t const v = ald(&a.sup[i]);
                if (v == 0) {
                    ast(&a.stamp[i], 1u);
                    a.res[i] = 0;
                } else {
                    ast(&a.stamp[i], 0u);
                    keep = true;
                    mn = v < mn ? v : mn;
                }
            }
            uint32_t tot = 0;
            uint32_t const ex = block_excl_scan<kMB>(keep ? 1u : 0u, s_part, tot);
            if constexpr (kBatch) {
                if (keep) a.live[w + ex] = i;
            } else {
                // with a window all go above it (the first level splits them),
                // else to collect list 0
                unsigned long long* const len = kWindow ? &ctl->high_n[0] : &ctl->live_n[0];
                uint32_t* const list = kWindow ? a.live + 2 * size_t(a.m) : a.live;
                if (tid == 0) s_gbase = tot ? uint32_t(afetch_add(len, (unsigned long long)tot)) : 0u;
                __syncthreads();
                if (keep) ast(&list[s_gbase + ex], i);  // read by other blocks
                __syncthreads();
            }
            w += tot;
        }
        if (tid == 0) a.live_cnt[B] = w - seg0;
        set_slots(1, block_min_all(mn, s_red), kInf);
        if (!grid_sync(a, G, ++epoch, &s_ok)) return;
    }
    uint32_t X = a.resume ? uint32_t(ctl->r_X) : 2u, k = a.resume ? uint32_t(ctl->r_k) : 0u;
    uint64_t W = a.resume ? ctl->W : 0ull;  // top of the collect window (plain graphs)
    uint32_t nt = 0;            // claimed blooms of sub-round X
    uint64_t t_col = 0, t_pe = 0, w_col = 0, w_pe = 0, pmax_sum = 0;
    // diagnostics: middles walked by waves and by the whole block, and the
    // block walks (in LDS, so that no register holds them)
    if (tid < 3) s_dmid[tid] = 0;
    // adds this block's middle counts to the control block (diagnostics)
    auto flush_diag = [&]() {
        if (!a.diag) return;
        __syncthreads();
        if (tid == 0) {
            afetch_add(&ctl->mid_wave, s_dmid[0]);
            afetch_add(&ctl->mid_block, s_dmid[1]);
            afetch_add(&ctl->n_block, s_dmid[2]);
            s_dmid[0] = s_dmid[1] = s_dmid[2] = 0;
        }
    };

    // ---- one peel sub-round X over the claimed blooms; level_of(bg) is the
    // level of the bloom's graph (bg = its graph in a batch, else 0); the
    // number of blooms claimed for X + 1 is returned to every thread
    auto peel_sub_round = [&](auto level_of, uint64_t& tmin, uint64_t& dec_written, uint32_t p) -> uint32_t {
        uint64_t const tp0 = wall_clock64();
        if (B == 0 && tid == 0) ast(&ctl->ovf_n[(X + 2) % 3], 0u);
        Out const out = out_for(X + 1);
        Dec dec{a.sup, a.stamp, a.res, 0, X, tmin, kWindow ? W : kInf};
        if (kWindow && tid == 0) {
            s_low_list = a.live + size_t(p ^ 1u) * a.m;
            s_low_len = &ctl->live_n[p ^ 1u];
        }
        uint64_t* c2 = s_c2[wv];
        CrossBuf xb{s_xl[wv], 0, ws};
        int const lane = tid & (kWave - 1);
        uint64_t const tw0 = wall_clock64();
        // the block's blooms are list entries B, B + G, B + 2G, ...: each
        // wave starts with one and takes the next slot from an LDS counter,
        // so a wave held up by a large bloom leaves the rest to the others.
        // Software pipeline: the list entry two blooms ahead and the header
        // and first 64 middles of the next one load while the current bloom
        // is processed.
        uint32_t const nb_blk = nt > B ? (nt - B + G - 1) / G : 0u;
        if (tid == 0) {
            s_pick = kWPB;
            s_nbig = 0;
        }
        __syncthreads();
        auto grab = [&]() -> uint32_t {
            uint32_t v = 0;
            if (lane == 0) v = atomicAdd(&s_pick, 1u);
            return __builtin_amdgcn_readfirstlane(v);
        };
        uint32_t s_next = 0, s_next2 = 0, hdr_next = 0, bg_next = 0;
        uint64_t r0_next = 0;
        uint32_t j_nxt = grab();
        bool more = uint32_t(wv) < nb_blk;
        if (more) {
            s_next = list_entry(X, B + G * uint32_t(wv));
            hdr_next = ald(&a.hdr[s_next]);
            r0_next = (s_next + lane < a.R) ? ald(&a.rec[s_next + lane]) : 0;
            if constexpr (kBatch) bg_next = a.bgraph[s_next];
        }
        if (j_nxt < nb_blk) s_next2 = list_entry(X, B + G * j_nxt);
        while (more) {
            uint32_t const s = s_next, hdr = hdr_next, bg = bg_next;
            uint64_t const r0 = r0_next;
            more = j_nxt < nb_blk;
            if (more) {
                s_next = s_next2;
                hdr_next = ald(&a.hdr[s_next]);
                r0_next = (s_next + lane < a.R) ? ald(&a.rec[s_next + lane]) : 0;
                if constexpr (kBatch) bg_next = a.bgraph[s_next];
                j_nxt = grab();
                if (j_nxt < nb_blk) s_next2 = list_entry(X, B + G * j_nxt);
            }
            uint32_t const kb = hdr & 0x7FFFFFFFu, up = hdr >> 31;
            if (kb < 2) continue;
            // a large bloom is walked by the whole block after this loop,
            // unless the block's list of them is full
            if (kb > a.big_min) {
                uint32_t q = 0;
                if (lane == 0) q = atomicAdd(&s_nbig, 1u);
                q = __builtin_amdgcn_readfirstlane(q);
                if (q < kBigCap) {
                    if (lane == 0) {
                        s_big[q] = s;
                        s_big_bg[q] = bg;
                    }
                    continue;
                }
            }
            dec.L = level_of(bg);
            if (a.diag && lane == 0) atomicAdd(&s_dmid[0], (unsigned long long)kb);
            // a bloom frees at most two edges per middle
            xb.ensure(2 * (kb < uint32_t(16 * kWave) ? kb : uint32_t(16 * kWave)), X + 1, a, out);
            bool done = false;
            if (kb <= uint32_t(kWave)) done = process_bloom<1>(a, s, kb, up, r0, X, dec, c2, xb, out);
            else if (kb <= uint32_t(4 * kWave)) done = process_bloom<4>(a, s, kb, up, r0, X, dec, c2, xb, out);
            else if (kb <= uint32_t(16 * kWave)) done = process_bloom<16>(a, s, kb, up, r0, X, dec, c2, xb, out);
            if (!done) process_large(a, s, kb, up, X, dec, c2, xb, out);
        }
        // ---- the large blooms, each by the whole block: a counting pass
        // over the middles (UB per thread in flight, so a block step covers
        // UB * 512 middles with one round of dependent loads), then a
        // decrement pass in the same layout; the both-frontier middles for
        // the rule are gathered in LDS. A bloom whose both-frontier middles
        // overflow that list goes to one wave (process_large: sorted lists).
        // Once at most half of it is left the decrement pass also compacts
        // it (each block step reads its middles before any is written, and
        // writes land below them), so the walks stay proportional to its
        // live middles; middles keep no order.
        __syncthreads();
        uint32_t const nbig = s_nbig < kBigCap ? s_nbig : kBigCap;
        constexpr int UB = 4;
        for (uint32_t bi = 0; bi < nbig; bi++) {
            uint32_t const s = s_big[bi];
            uint32_t const hdr = ald(&a.hdr[s]);
            uint32_t const kb = hdr & 0x7FFFFFFFu, up = hdr >> 31;
            uint64_t* rec = a.rec + s;
            dec.L = level_of(s_big_bg[bi]);
            if (a.diag && tid == 0) {
                s_dmid[1] += kb;
                s_dmid[2]++;
            }
            if (tid < 4) s_bcnt[tid] = 0;
            __syncthreads();
            uint32_t w0 = 0, w1 = 0, wp = 0;
            for (uint32_t j0 = uint32_t(wv) * kWave; j0 < kb; j0 += UB * kMB) {
                uint64_t r[UB];
                uint32_t s1[UB], s2[UB];
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    uint32_t const j = j0 + u * kMB + lane;
                    r[u] = j < kb ? ald(&rec[j]) : 0;
                }
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    uint32_t const j = j0 + u * kMB + lane;
                    s1[u] = j < kb ? ald(&a.stamp[uint32_t(r[u] >> 32)]) : 0u;
                    s2[u] = j < kb ? ald(&a.stamp[uint32_t(r[u])]) : 0u;
                }
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    uint32_t const j = j0 + u * kMB + lane;
                    int const c = j < kb ? classify(s1[u], s2[u], X) : -1;
                    uint64_t const b2 = __ballot(c == 2);
                    if (b2) {
                        uint32_t base = 0;
                        if (lane == 0) base = atomicAdd(&s_bcnt[2], uint32_t(__popcll(b2)));
                        base = __builtin_amdgcn_readfirstlane(base);
                        uint32_t const pos = base + __popcll(b2 & lanemask_lt());
                        if (c == 2 && pos < uint32_t(kC2Cap)) s_bc2[pos] = r[u];
                    }
                    w0 += __popcll(__ballot(c == 0));
                    w1 += __popcll(__ballot(c == 1));
                    wp += __popcll(__ballot(c == 2 && uint32_t(r[u] >> 32) < uint32_t(r[u])));
                }
            }
            if (lane == 0) {
                atomicAdd(&s_bcnt[0], w0);
                atomicAdd(&s_bcnt[1], w1);
                atomicAdd(&s_bcnt[3], wp);
            }
            __syncthreads();
            uint32_t const n0 = s_bcnt[0], n1 = s_bcnt[1], n2 = s_bcnt[2];
            if (a.rule != kRuleExact && n2 > a.c2cap && n1) {
                if (wv == 0) process_large(a, s, kb, up, X, dec, c2, xb, out);
                __syncthreads();
                continue;
            }
            uint64_t const dec0 = uint64_t(n1) + n2;
            // compact in this pass: a bloom at most half alive that keeps a
            // butterfly (below two middles the top claim stamp retires it)
            bool const compact = n0 >= 2 && 2 * n0 <= kb;
            uint32_t wpos = 0;  // kept middles written so far (whole block)
            for (uint32_t it0 = 0; it0 < kb; it0 += UB * kMB) {
                xb.ensure(2 * UB * kWave, X + 1, a, out);
                uint32_t const j0 = it0 + uint32_t(wv) * kWave;
                uint64_t r[UB];
                uint32_t s1[UB], s2[UB];
                int c[UB];
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    uint32_t const j = j0 + u * kMB + lane;
                    r[u] = j < kb ? ald(&rec[j]) : 0;
                }
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    uint32_t const j = j0 + u * kMB + lane;
                    s1[u] = j < kb ? ald(&a.stamp[uint32_t(r[u] >> 32)]) : 0u;
                    s2[u] = j < kb ? ald(&a.stamp[uint32_t(r[u])]) : 0u;
                }
                // targets: the first is r-v (c = 0) or the survivor (c = 1),
                // the second v-w (c = 0 only)
                uint32_t t1[UB];
                uint64_t d1[UB];
                uint32_t nk = 0;
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    uint32_t const j = j0 + u * kMB + lane;
                    c[u] = j < kb ? classify(s1[u], s2[u], X) : -1;
                    uint32_t const e1 = uint32_t(r[u] >> 32), e2 = uint32_t(r[u]);
                    t1[u] = e1;
                    d1[u] = 0;
                    if (c[u] == 0) {
                        d1[u] = dec0;
                        nk++;
                    } else if (c[u] == 1) {
                        bool const froot = s1[u] != X;  // the survivor is r-v
                        uint32_t const g = froot ? e2 : e1;
                        t1[u] = froot ? e1 : e2;
                        d1[u] = uint64_t(n0) + (n1 - 1) + rule_sum(s_bc2, n2, froot, up, g, a);
                    }
                }
                uint64_t o1[UB], o2[UB];
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    o1[u] = d1[u] ? dec.issue(t1[u], d1[u]) : 0;
                    o2[u] = (c[u] == 0 && dec0) ? dec.issue(uint32_t(r[u]), dec0) : 0;
                }
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    bool const x1 = d1[u] && dec.settle(t1[u], d1[u], o1[u]);
                    bool const x2 = c[u] == 0 && dec0 && dec.settle(uint32_t(r[u]), dec0, o2[u]);
                    xb.put(x1, t1[u]);
                    xb.put(x2, uint32_t(r[u]));
                }
                if (compact) {
                    // every middle of this block step has been read (the
                    // scan's barriers) before the kept ones are written
                    uint32_t tot = 0;
                    uint32_t pos = wpos + block_excl_scan<kMB>(nk, s_part, tot);
#pragma unroll
                    for (int u = 0; u < UB; u++)
                        if (c[u] == 0) ast(&rec[pos++], r[u]);
                    wpos += tot;
                }
            }
            __syncthreads();
            if (tid == 0) {
                if (n0 < 2) {
                    // no butterfly is left: the top claim stamp keeps the
                    // bloom out of every later claim list
                    afetch_max(&a.claim[s], ~0u);
                } else if (compact) {
                    ast(&a.hdr[s], wpos | (up << 31));
                }
            }
            __syncthreads();
        }
        tmin = dec.tmin;
        uint64_t const tw = wall_clock64() - tw0;
        if (lane == 0) s_xn[wv] = xb.n;
        if constexpr (!kBatch) {
            // publish this block's smallest decremented support of the level
            uint64_t const bdec = block_min_all(tmin, s_red);
            if (tid == 0 && bdec < dec_written) ast(&a.slots[size_t(p) * 2 * G + G + B], bdec);
            dec_written = bdec < dec_written ? bdec : dec_written;
        } else {
            __syncthreads();
        }
        uint64_t const tc0 = wall_clock64();
        claim_block(a, X + 1, out, s_xl, s_xn, s_incl, s_off, s_part);
        close_out(X + 1);
        uint64_t const tp1 = wall_clock64();
        if (a.diag) {
            uint64_t const mx = ~block_min_all(~tw, s_red);  // slowest wave of the block
            if (tid == 0) {
                afetch_max(&ctl->pmax[X & 1], (unsigned long long)mx);
                afetch_max(&ctl->cmax[X & 1], (unsigned long long)(tp1 - tc0));
                afetch_max(&ctl->wmax[X & 1], (unsigned long long)(tp1 - tp0));
            }
        }
        if (a.slice_ticks && B == 0 && tid == 0) ast(&ctl->yield, tp1 - ctl->t_launch >= a.slice_ticks ? 1u : 0u);
        if (!grid_sync(a, G, ++epoch, &s_ok)) return ~0u;
        uint32_t const nn = open_list(X + 1);
        uint64_t const tp2 = wall_clock64();
        t_pe += tp2 - tp0;
        w_pe += tp1 - tp0;
        if (a.diag && B == 0 && tid == 0) {
            unsigned long long const pm = ald(&ctl->pmax[X & 1]);
            ast(&ctl->pmax[X & 1], 0ull);
            pmax_sum += pm;
            ctl->cmax_sum += ald(&ctl->cmax[X & 1]);
            ast(&ctl->cmax[X & 1], 0ull);
            ctl->wmax_sum += ald(&ctl->wmax[X & 1]);
            ast(&ctl->wmax[X & 1], 0ull);
            int lb = 0;
            while (lb < 23 && (pm >> (lb + 1))) lb++;
            ctl->hist[lb]++;
        }
        return nn;
    };

    if constexpr (!kBatch) {
        uint64_t L = a.resume ? ctl->r_L : level_min(1);
        for (;;) {
            // the slice is over: `yield` was set before the last barrier and
            // is rewritten only after the next one, so every block agrees
            if (a.slice_ticks && ald(&ctl->yield)) {
                if (B == 0 && tid == 0) {
                    ctl->r_X = X;
                    ctl->r_k = k;
                    ctl->r_L = L;
                    ctl->r_epoch = epoch;
                    ctl->t_collect += t_col;
                    ctl->t_peel += t_pe;
                    ctl->t_work_collect += w_col;
                    ctl->t_work_peel += w_pe;
                    ctl->pmax_sum += pmax_sum;
                    ast(&ctl->yielded, 1u);
                }
                flush_diag();
                return;
            }
            uint32_t const p = k & 1;
            // ---- window: the collect scans only the live edges with support
            // <= W (list p); the others wait in the list above the window,
            // and a decrement that takes an edge into the window appends it
            // to the next level's list. When list p is empty, every live
            // edge is above the window (a live edge that entered it is in
            // list p until it is peeled), so the level is not known from the
            // slots: the minimum of the list above is taken (phase 1), and
            // the list is split at the new window top (phase 2), dropping
            // the edges already peeled. The same when the slots hold no
            // level (list p's edges are all peeled). A level taken from the
            // slots while list p has no live edge can exceed W: its collect
            // then finds nothing, changes nothing, and the next level splits.
            if (!kWindow && L == kInf) break;
            if (kWindow && (ald(&ctl->live_n[p]) == 0 || L == kInf)) {
                uint32_t const hp = uint32_t(ald(&ctl->hp));
                const uint32_t* hcur = a.live + size_t(2 + hp) * a.m;
                uint32_t* hnxt = a.live + size_t(3 - hp) * a.m;
                uint32_t* lcur = a.live + size_t(p) * a.m;
                uint32_t const hcnt = uint32_t(ald(&ctl->high_n[hp]));
                constexpr int U = 4;
                uint64_t mn = kInf;
                for (uint32_t base = B * (kMB * U); base < hcnt; base += G * (kMB * U)) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        if (i < hcnt) {
                            uint32_t const e = ald(&hcur[i]);
                            if (ald(&a.stamp[e]) == 0) {
                                uint64_t const v = ald(&a.sup[e]);
                                mn = v < mn ? v : mn;
                            }
                        }
                    }
                }
                set_slots(p, block_min_all(mn, s_red), kInf);
                if (!grid_sync(a, G, ++epoch, &s_ok)) return;
                L = level_min(p);
                if (L == kInf) break;
                // a window pays on many live edges; below that, one list
                // holds them all (no edge is above the window)
                W = hcnt > a.window_min ? L + (L / 4 > 16 ? L / 4 : 16) : kInf;
                for (uint32_t base = B * (kMB * U); base < hcnt; base += G * (kMB * U)) {
                    uint32_t e[U];
                    int where[U];  // 0 dropped, 1 window, 2 above
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        e[u] = i < hcnt ? ald(&hcur[i]) : 0u;
                        where[u] = 0;
                        if (i < hcnt && ald(&a.stamp[e[u]]) == 0) where[u] = ald(&a.sup[e[u]]) <= W ? 1 : 2;
                    }
                    uint32_t n1 = 0, n2 = 0;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        n1 += where[u] == 1 ? 1u : 0u;
                        n2 += where[u] == 2 ? 1u : 0u;
                    }
                    uint32_t t1 = 0, t2 = 0;
                    uint32_t const x1 = block_excl_scan<kMB>(n1, s_part, t1);
                    uint32_t const x2 = block_excl_scan<kMB>(n2, s_part, t2);
                    if (tid == 0) {
                        s_bcnt[0] = t1 ? uint32_t(afetch_add(&ctl->live_n[p], (unsigned long long)t1)) : 0u;
                        s_bcnt[1] = t2 ? uint32_t(afetch_add(&ctl->high_n[hp ^ 1u], (unsigned long long)t2)) : 0u;
                    }
                    __syncthreads();
                    uint32_t q1 = s_bcnt[0] + x1, q2 = s_bcnt[1] + x2;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (where[u] == 1) ast(&lcur[q1++], e[u]);
                        if (where[u] == 2) ast(&hnxt[q2++], e[u]);
                    }
                    __syncthreads();
                }
                if (!grid_sync(a, G, ++epoch, &s_ok)) return;
                // every block has read the old list's length and the window
                // is the same in every block; block 0 records them for the
                // next split and for a resumed launch
                if (B == 0 && tid == 0) {
                    ast(&ctl->high_n[hp], 0ull);
                    ast(&ctl->hp, hp ^ 1u);
                    ast(&ctl->W, W);
                }
            }
            // a prefix of the decomposition only: levels up to max_level
            if (a.max_level && L > a.max_level) {
                if (B == 0 && tid == 0) ctl->partial = 1;
                break;
            }
            // ---- collect: frontier of level L from the shared live list
            {
                uint64_t const tp0 = wall_clock64();
                if (B == 0 && tid == 0) ast(&ctl->ovf_n[(X + 1) % 3], 0u);
                Out const out = out_for(X);
                // live list p is read in chunks spread over all blocks, the
                // edges that stay live are appended to list 1 - p (device-
                // scope accesses: other blocks, on other XCDs, read them)
                const uint32_t* cur = a.live + size_t(p) * a.m;
                uint32_t* nxt = a.live + size_t(p ^ 1u) * a.m;
                uint32_t const cnt = uint32_t(ald(&ctl->live_n[p]));
                uint64_t mn = kInf;
                CrossBuf xb{s_xl[wv], 0, ws};
                constexpr int U = 4;
                for (uint32_t base = B * (kMB * U); base < cnt; base += G * (kMB * U)) {
                    xb.ensure(U * kWave, X, a, out);
                    // an edge finished at an earlier level has support below L
                    // (it was peeled at a level < L with support <= that level);
                    // a live edge has support >= L
                    uint32_t e[U];
                    uint64_t v[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        e[u] = i < cnt ? ald(&cur[i]) : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        v[u] = i < cnt ? ald(&a.sup[e[u]]) : 0ull;
                    }
                    uint32_t nk = 0;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        bool const in = i < cnt;
                        nk += (in && v[u] > L) ? 1u : 0u;
                        if (in && v[u] > L) mn = v[u] < mn ? v[u] : mn;
                    }
                    uint32_t tot = 0;
                    uint32_t const ex = block_excl_scan<kMB>(nk, s_part, tot);
                    if (tid == 0) s_gbase = tot ? uint32_t(afetch_add(&ctl->live_n[p ^ 1u], (unsigned long long)tot)) : 0u;
                    __syncthreads();
                    uint32_t pos = s_gbase + ex;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        bool const in = i < cnt;
                        if (in && v[u] > L) ast(&nxt[pos++], e[u]);
                        bool const take = in && v[u] == L;
                        if (take) {
                            ast(&a.stamp[e[u]], X);
                            a.res[e[u]] = L;
                        }
                        xb.put(take, e[u]);
                    }
                    __syncthreads();
                }
                set_slots(p, block_min_all(mn, s_red), kInf);
                if ((tid & (kWave - 1)) == 0) s_xn[wv] = xb.n;
                __syncthreads();
                claim_block(a, X, out, s_xl, s_xn, s_incl, s_off, s_part);
                close_out(X);
                uint64_t const tp1 = wall_clock64();
                if (!grid_sync(a, G, ++epoch, &s_ok)) return;
                // every block has read list p's length: it is the next
                // level's append target (the peel barriers come between)
                if (B == 0 && tid == 0) ast(&ctl->live_n[p], 0ull);
                nt = open_list(X);
                uint64_t const tp2 = wall_clock64();
                t_col += tp2 - tp0;
                w_col += tp1 - tp0;
            }
            // ---- peel sub-rounds X, X + 1, ... until one claims nothing
            uint64_t tmin = kInf;  // this thread's smallest decremented support of the level
            uint64_t dec_written = kInf;
            for (;;) {
                uint32_t const nn = peel_sub_round([&](uint32_t) { return L; }, tmin, dec_written, p);
                if (nn == ~0u) return;
                if (!nn) break;
                nt = nn;
                X++;
            }
            // no bloom claimed for X + 1: the level is over (X + 1 is skipped)
            L = level_min(p);
            X += 2;
            k++;
        }
    } else {
        // Batch: level steps. In step t every graph with live edges peels
        // its own next level L_g = min support of its live edges, so each
        // graph goes through exactly the levels and sub-rounds of its plain
        // run while the phases are shared. Liveness is the stamp (0 = live).
        uint32_t const gA = seg1 > seg0 ? a.egraph[seg0] : 0u;
        uint32_t const gB = seg1 > seg0 ? a.egraph[seg1 - 1] : 0u;
        uint32_t const span = seg1 > seg0 ? gB - gA + 1 : 0u;
        bool const lds_min = span <= kSpan;
        for (;;) {
            uint64_t* gl = a.glev + size_t(k & 1) * a.ngraph;
            uint64_t* gl_next = a.glev + size_t((k + 1) & 1) * a.ngraph;
            // ---- level: drop finished edges, per-graph min of live supports
            {
                uint64_t const tp0 = wall_clock64();
                for (uint32_t i = tid; i < span; i += kMB) ast(&gl_next[gA + i], kInf);
                if (lds_min)
                    for (uint32_t i = tid; i < span; i += kMB) s_gmin[i] = kInf;
                __syncthreads();
                uint32_t const cnt = a.live_cnt[B];
                uint32_t w = seg0;
                constexpr int U = 4;
                for (uint32_t base = seg0; base < seg0 + cnt; base += kMB * U) {
                    uint32_t e[U], st[U], g[U];
                    uint64_t v[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        e[u] = i < seg0 + cnt ? a.live[i] : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        bool const in = i < seg0 + cnt;
                        st[u] = in ? ald(&a.stamp[e[u]]) : 1u;
                        v[u] = in ? ald(&a.sup[e[u]]) : 0ull;
                        g[u] = in ? a.egraph[e[u]] : 0u;
                    }
                    uint32_t nk = 0;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (st[u] != 0) continue;
                        nk++;
                        if (lds_min) atomicMin(reinterpret_cast<unsigned long long*>(&s_gmin[g[u] - gA]),
                                               (unsigned long long)v[u]);
                        else __hip_atomic_fetch_min(&gl[g[u]], v[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    uint32_t tot = 0;
                    uint32_t pos = w + block_excl_scan<kMB>(nk, s_part, tot);
                    w += tot;
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (st[u] == 0) a.live[pos++] = e[u];
                }
                if (tid == 0) {
                    a.live_cnt[B] = w - seg0;
                    ast(&a.any[B], w > seg0 ? 1u : 0u);
                }
                __syncthreads();
                if (lds_min)
                    for (uint32_t i = tid; i < span; i += kMB)
                        if (s_gmin[i] != kInf)
                            __hip_atomic_fetch_min(&gl[gA + i], s_gmin[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (!grid_sync(a, G, ++epoch, &s_ok)) return;
                // any live edge left anywhere?
                uint32_t const any = tid < G ? ald(&a.any[tid]) : 0u;
                if (!__syncthreads_or(int(any))) break;
                uint64_t const tp2 = wall_clock64();
                t_col += tp2 - tp0;
            }
            // ---- collect: each graph's frontier of its level
            {
                uint64_t const tp0 = wall_clock64();
                if (B == 0 && tid == 0) ast(&ctl->ovf_n[(X + 1) % 3], 0u);
                Out const out = out_for(X);
                uint32_t const cnt = a.live_cnt[B];
                CrossBuf xb{s_xl[wv], 0, ws};
                constexpr int U = 4;
                for (uint32_t base = seg0; base < seg0 + cnt; base += kMB * U) {
                    xb.ensure(U * kWave, X, a, out);
                    uint32_t e[U];
                    uint64_t v[U], L[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        e[u] = i < seg0 + cnt ? a.live[i] : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        uint32_t const i = base + u * kMB + tid;
                        bool const in = i < seg0 + cnt;
                        v[u] = in ? ald(&a.sup[e[u]]) : 0ull;
                        L[u] = in ? ald(&gl[a.egraph[e[u]]]) : kInf;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        bool const take = v[u] == L[u];
                        if (take) {
                            ast(&a.stamp[e[u]], X);
                            a.res[e[u]] = L[u];
                        }
                        xb.put(take, e[u]);
                    }
                }
                if ((tid & (kWave - 1)) == 0) s_xn[wv] = xb.n;
                __syncthreads();
                claim_block(a, X, out, s_xl, s_xn, s_incl, s_off, s_part);
                close_out(X);
                uint64_t const tp1 = wall_clock64();
                if (!grid_sync(a, G, ++epoch, &s_ok)) return;
                nt = open_list(X);
                uint64_t const tp2 = wall_clock64();
                t_col += tp2 - tp0;
                w_col += tp1 - tp0;
            }
            // ---- peel sub-rounds until no graph claims a bloom
            uint64_t tmin = kInf, dec_written = kInf;
            for (;;) {
                uint32_t const nn = peel_sub_round([&](uint32_t bg) { return ald(&gl[bg]); }, tmin, dec_written, 0);
                if (nn == ~0u) return;
                if (!nn) break;
                nt = nn;
                X++;
            }
            X += 2;
            k++;
        }
    }
    flush_diag();
    if (B == 0 && tid == 0) {
        ctl->subrounds = X - 1;
        ctl->levels = k;
        ctl->t_collect += t_col;
        ctl->t_peel += t_pe;
        ctl->t_work_collect += w_col;
        ctl->t_work_peel += w_pe;
        ctl->pmax_sum += pmax_sum;
    }
}

__global__ void k_msp_init(MspCtl* ctl, uint32_t* arrive, uint32_t n_arrive, uint64_t* glev, uint32_t n_glev) {
    uint32_t const first = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (uint32_t i = first; i < n_arrive; i += stride) arrive[i] = 0;
    for (uint32_t i = first; i < n_glev; i += stride) glev[i] = kInf;
    if (first == 0) {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(ctl);
        for (size_t j = 0; j < sizeof(MspCtl) / sizeof(unsigned long long); j++) w[j] = 0;
    }
}

// after a prefix of the decomposition: the edges not peeled yet report their
// current support
__global__ void k_prefix_state(uint32_t m, const uint32_t* __restrict__ stamp, const uint64_t* __restrict__ sup,
                               uint64_t* __restrict__ res) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        if (stamp[i] == 0) res[i] = sup[i];
}

// One bit per edge: the pivot orientation of bitruss_mspp is turned around
// (deg(upper) > 10 deg(lower), process_edge in bitruss_msp.cu:366-378).
__global__ void k_orient_bits(uint32_t m, const uint32_t* __restrict__ edges, const uint32_t* __restrict__ deg,
                              uint32_t* __restrict__ bits) {
    uint32_t const words = m / 32 + 1;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gridDim.x * blockDim.x) {
        uint32_t b = 0;
        for (uint32_t i = 0; i < 32; i++) {
            uint32_t const e = w * 32 + i;
            if (e < m && uint64_t(deg[edges[2ull * e]]) > 10ull * deg[edges[2ull * e + 1]]) b |= 1u << i;
        }
        bits[w] = b;
    }
}

__global__ void k_max_res(uint32_t m, const uint64_t* __restrict__ a, uint64_t* __restrict__ out) {
    uint64_t mx = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        mx = a[i] > mx ? a[i] : mx;
    mx = wave_max_u64(mx);
    if ((threadIdx.x & (kWave - 1)) == 0 && mx) atomic_max_u64(out, mx);
}

}  // namespace

// Peels with the bloom index; the supports and index records come from the
// counting call that precedes it (bt_count_impl with emit).
int bt_msp_index_peel(bt_dgraph* dg, uint32_t* levels_out, int rule) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const m = dg->m;
    bt_timings const count_t = c->last;
    bt_timer tm(c);
    BT_TRY(tm.start(5));
    dg->has_support = false;
    BT_TRY(bt_index_build(dg));
    bt_index& ix = c->ix;
    // one block per CU keeps the grid barrier short; block 0 watches every
    // other block with one thread
    unsigned const G = unsigned(c->peel_blocks > 0 && c->peel_blocks < c->cu_count ? c->peel_blocks : c->cu_count);
    if (G == 0 || G >= unsigned(kMB)) return BT_ERR_DEVICE;
    int per_cu = 0;
    bool const batch = dg->egraph != nullptr && dg->ngraph > 0 && c->batch_levels == BT_BATCH_LEVELS_PER_GRAPH;
    // the window for plain graphs with many edges (the live edges at the
    // start are at most m); BT_OPT_PEEL_WINDOW_EDGES moves the threshold
    // (0 exercises the window on small graphs)
    uint64_t const window_edges = c->peel_window_edges;
    bool const window = !batch && m > window_edges;
    const void* kern = batch    ? reinterpret_cast<const void*>(k_msp_persist<true>)
                       : window ? reinterpret_cast<const void*>(k_msp_persist<false, true>)
                                : reinterpret_cast<const void*>(k_msp_persist<false>);
    BT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kMB, 0));
    if (per_cu < 1) return BT_ERR_DEVICE;
    uint32_t const nbloom = uint32_t(std::max<uint64_t>(ix.NB, 1));
    // largest bloom: the waves' scratch for sorted both-frontier lists
    uint32_t kmax = 0;
    BT_HIP_TRY(hipMemcpyAsync(&kmax, dg->ctl + kCtlBloomMax, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    BT_HIP_TRY(hipStreamSynchronize(s));
    // BT_OPT_PEEL_C2CAP (0..128) lowers the LDS list's share (0 exercises
    // the sorted lists on small graphs)
    uint32_t const c2cap = uint32_t(std::min(std::max(c->peel_c2cap, 0), kC2Cap));
    uint32_t const c2s_stride = kmax > c2cap ? (3 * kmax + 63) & ~63u : 0u;
    if (c2s_stride) BT_TRY(bt_buf_reserve(ix.c2s, sizeof(uint32_t) * size_t(G) * kWPB * c2s_stride));
    // claim stamps: R real ones, then the dummies (both sized by the index build)
    if (ix.bclaim.bytes < sizeof(uint32_t) * (ix.R + size_t(G) * kClaimDummiesPerBlock)) return BT_ERR_INTERNAL;
    // plain graphs: two collect lists and two lists above the window
    BT_TRY(bt_buf_reserve(ix.live, sizeof(uint32_t) * 4 * std::max<size_t>(m, 1)));
    BT_TRY(bt_buf_reserve(ix.area, sizeof(uint32_t) * 3 * size_t(G) * kBlkCap));
    BT_TRY(bt_buf_reserve(ix.touched, sizeof(uint32_t) * 3 * size_t(nbloom)));
    // per-block words: live counts, list lengths, level slots, barrier slots
    auto up256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t const o_cnt = up256(sizeof(uint32_t) * G);
    size_t const o_slots = o_cnt + up256(sizeof(uint32_t) * 3 * G);
    size_t const o_arrive = o_slots + up256(sizeof(uint64_t) * 4 * G);
    size_t const n_arrive = size_t(G) * kSp;
    size_t const o_any = o_arrive + up256(sizeof(uint32_t) * n_arrive);
    BT_TRY(bt_buf_reserve(ix.seg, o_any + sizeof(uint32_t) * G));
    uint32_t const n_glev = batch ? 2 * dg->ngraph : 0u;
    if (batch) BT_TRY(bt_buf_reserve(ix.glev, sizeof(uint64_t) * n_glev));
    char* seg = static_cast<char*>(ix.seg.ptr);
    MspCtl* ctl = reinterpret_cast<MspCtl*>(dg->ctl);
    MspArgs a{};
    a.m = m;
    a.R = uint32_t(ix.R);
    a.nbloom = nbloom;
    a.sup = dg->sup;
    a.res = dg->res;
    a.stamp = dg->list0;
    a.inc_off = static_cast<const uint64_t*>(ix.inc_off.ptr);
    a.inc = static_cast<const uint32_t*>(ix.inc.ptr);
    a.rec = static_cast<uint64_t*>(ix.vals1.ptr);
    a.hdr = static_cast<uint32_t*>(ix.bk.ptr);
    a.claim = static_cast<uint32_t*>(ix.bclaim.ptr);
    a.area = static_cast<uint32_t*>(ix.area.ptr);
    a.ovf = static_cast<uint32_t*>(ix.touched.ptr);
    a.live = static_cast<uint32_t*>(ix.live.ptr);
    a.live_cnt = reinterpret_cast<uint32_t*>(seg);
    a.cnt = reinterpret_cast<uint32_t*>(seg + o_cnt);
    a.slots = reinterpret_cast<uint64_t*>(seg + o_slots);
    a.arrive = reinterpret_cast<uint32_t*>(seg + o_arrive);
    a.ctl = ctl;
    a.c2s = c2s_stride ? static_cast<uint32_t*>(ix.c2s.ptr) : nullptr;
    a.c2s_stride = c2s_stride;
    a.c2cap = c2cap;
    a.window_min = uint32_t(std::min<uint64_t>(window_edges / 4, 0xFFFFFFFFull));
    // blooms beyond the register tiers (default 1,024 middles) are walked
    // by the whole block
    a.big_min = uint32_t(std::max(c->peel_big_middles, 1));
    // BT_OPT_PEEL_MAX_LEVEL: stop after that level and report the state of
    // the prefix (peeled edges their result, the others their current
    // support) with BT_PARTIAL; plain graphs only (not batch unions)
    a.max_level = dg->ngraph == 0 ? c->peel_max_level : 0;
    a.rule = uint32_t(rule);
    if (rule == kRuleMspp) {
        // process_edge (bitruss_msp.cu:366-378): an edge whose upper
        // endpoint's degree is above ten times its lower one's is peeled
        // with the endpoints exchanged
        BT_TRY(bt_buf_reserve(ix.flip, sizeof(uint32_t) * (size_t(m) / 32 + 1)));
        a.flipbits = static_cast<const uint32_t*>(ix.flip.ptr);
        k_orient_bits<<<grid_for(size_t(m) / 32 + 1, 256, unsigned(c->cu_count) * 8), 256, 0, s>>>(
            m, dg->edges, dg->deg, static_cast<uint32_t*>(ix.flip.ptr));
        BT_HIP_TRY(hipGetLastError());
    }
    a.id_bits = 1;
    while (a.id_bits < 32 && ((m ? m - 1 : 0) >> a.id_bits)) a.id_bits++;
    if (batch) {
        a.ngraph = dg->ngraph;
        a.egraph = dg->egraph;
        a.bgraph = static_cast<const uint32_t*>(ix.bgraph.ptr);
        a.glev = static_cast<uint64_t*>(ix.glev.ptr);
        a.any = reinterpret_cast<uint32_t*>(seg + o_any);
    }
    // per-phase timing diagnostics on stderr (results are the same)
    a.diag = getenv("BT_PEEL_TRACE") ? 1u : 0u;
    k_msp_init<<<grid_for(std::max<size_t>(n_arrive, n_glev), 256, 1u << 16), 256, 0, s>>>(
        ctl, a.arrive, uint32_t(n_arrive), a.glev, n_glev);
    BT_HIP_TRY(hipGetLastError());
    // plain graphs run in time slices: the kernel returns at a level
    // boundary once its slice is over and is launched again to resume
    a.slice_ticks = (!batch && c->peel_slice_ms > 0) ? uint64_t(c->peel_slice_ms) * 100000ull : 0ull;
    bool const progress = getenv("BT_PEEL_PROGRESS") != nullptr;
    auto const t_first = std::chrono::steady_clock::now();
    unsigned slices = 1;
    // events 6..7 bracket the persistent kernel's launches
    BT_TRY(tm.start(6));
    void* params[] = {&a};
    bool plain_full = false;  // a whole-device peel launched without the cooperative guarantee
    {
        bt_cu_lease lease(c->device, c->cu_count, int(G));
        for (;; slices++) {
            if (G < unsigned(c->cu_count)) {
                // Part of the device: the runtime runs cooperative launches
                // one at a time, so this one is a plain launch. Its blocks
                // are all resident together: each needs a CU of its own
                // (LDS), the lease keeps the persistent peels' blocks within
                // the CU count, and other kernels on the device finish by
                // themselves.
                BT_HIP_TRY(hipLaunchKernel(kern, dim3(G), dim3(kMB), params, 0, s));
            } else if (hipLaunchCooperativeKernel(kern, dim3(G), dim3(kMB), params, 0, s) != hipSuccess) {
                // the runtime refuses cooperative launches while another
                // process holds the device ("cannot be performed in the
                // present state"): a plain launch of one block per CU, whose
                // blocks are resident together unless that other process
                // keeps CUs busy, in which case the bounded grid barrier
                // ends the peel with BT_ERR_BUSY
                (void)hipGetLastError();
                BT_HIP_TRY(hipLaunchKernel(kern, dim3(G), dim3(kMB), params, 0, s));
                plain_full = true;
            }
            if (!a.slice_ticks) break;
            uint64_t st[2 + 4];  // yield, yielded, r_X, r_k, r_L, r_epoch
            unsigned err = 0;
            BT_HIP_TRY(hipMemcpyAsync(st, &ctl->yield, sizeof(st), hipMemcpyDeviceToHost, s));
            BT_HIP_TRY(hipMemcpyAsync(&err, &ctl->err, sizeof(err), hipMemcpyDeviceToHost, s));
            BT_HIP_TRY(hipStreamSynchronize(s));
            if (err || !st[1]) break;
            if (progress) {
                double const sec =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t_first).count();
                std::fprintf(stderr, "[bt_peel] slice %u: level %llu, level steps %llu, sub-rounds %llu, %.1f s",
                             slices, (unsigned long long)st[4], (unsigned long long)st[3],
                             (unsigned long long)st[2], sec);
                if (a.diag) {
                    // block 0's clock: phase time, its own work, the slowest waves
                    MspCtl h{};
                    BT_HIP_TRY(hipMemcpy(&h, ctl, sizeof(MspCtl), hipMemcpyDeviceToHost));
                    std::fprintf(stderr,
                                 "; collect %.1f ms (block 0 work %.1f), peel %.1f ms (block 0 work %.1f, "
                                 "slowest waves %.1f), sorted %llu",
                                 h.t_collect * 1e-5, h.t_work_collect * 1e-5, h.t_peel * 1e-5, h.t_work_peel * 1e-5,
                                 h.pmax_sum * 1e-5, h.sorted);
                }
                std::fprintf(stderr, "\n");
                std::fflush(stderr);
            }
            a.resume = 1;
            BT_HIP_TRY(hipMemsetAsync(&ctl->yield, 0, 2 * sizeof(uint64_t), s));
        }
        BT_TRY(tm.start(7));
        BT_HIP_TRY(hipStreamSynchronize(s));
    }
    if (a.max_level && m)
        k_prefix_state<<<grid_for(m, 256, unsigned(c->cu_count) * 8), 256, 0, s>>>(m, a.stamp, a.sup, dg->res);
    // the level count the reference logs is the largest result
    BT_HIP_TRY(hipMemsetAsync(&ctl->level_max, 0, sizeof(uint64_t), s));
    if (m)
        k_max_res<<<grid_for(m, 256, unsigned(c->cu_count) * 8), 256, 0, s>>>(
            m, dg->res, reinterpret_cast<uint64_t*>(&ctl->level_max));
    BT_HIP_TRY(hipGetLastError());
    MspCtl h{};
    BT_HIP_TRY(hipMemcpyAsync(&h, ctl, sizeof(MspCtl), hipMemcpyDeviceToHost, s));
    BT_HIP_TRY(hipStreamSynchronize(s));
    // a barrier that timed out: in a plain launch, other work held CUs the
    // peel's workgroups needed (they were not resident together)
    if (h.err) return (G < unsigned(c->cu_count) || plain_full) ? BT_ERR_BUSY : BT_ERR_INTERNAL;
    if (a.diag) {
        std::fprintf(stderr,
                     "[bt_peel] index %.3f ms; phases: collect %.3f ms (block0 work %.3f), peel %.3f ms (block0 work "
                     "%.3f), slowest-wave sum %.3f ms, sub-rounds %llu, levels %llu, blocks %u, R %llu, blooms "
                     "%llu\n",
                     tm.ms(5, 6), h.t_collect * 1e-5, h.t_work_collect * 1e-5, h.t_peel * 1e-5, h.t_work_peel * 1e-5,
                     h.pmax_sum * 1e-5, h.subrounds, h.levels, G, (unsigned long long)ix.R,
                     (unsigned long long)ix.NB);
        std::fprintf(stderr,
                     "[bt_peel] per peel phase, slowest block: claim step %.3f ms, whole work %.3f ms (sums); "
                     "middles walked by waves %llu, by whole blocks %llu in %llu visits\n",
                     h.cmax_sum * 1e-5, h.wmax_sum * 1e-5, h.mid_wave, h.mid_block, h.n_block);
        std::fprintf(stderr, "[bt_peel] slowest-wave histogram (log2 10ns ticks: phases):");
        for (int i = 0; i < 24; i++)
            if (h.hist[i]) std::fprintf(stderr, " %d:%llu", i, h.hist[i]);
        std::fprintf(stderr, "\n");
    }
    c->last = count_t;
    c->last.peel_ms = tm.ms(5, 7);
    c->last.peel_kernel_ms = tm.ms(6, 7);
    c->last.peel_kernel_launches = slices;
    c->last.sorted_blooms = h.sorted;
    c->last.total_ms = count_t.total_ms + c->last.peel_ms;
    c->last.sub_rounds = h.subrounds;
    c->last.levels_scanned = h.levels;
    if (h.level_max > 0xFFFFFFFFull) return BT_ERR_OVERFLOW;
    if (levels_out) *levels_out = uint32_t(h.level_max);
    return h.partial ? BT_PARTIAL : BT_OK;
}
