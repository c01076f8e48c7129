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
// by a grid barrier and every block takes the same decision after each
// barrier from the same control words. Per sub-round one barrier; per level
// one more (collect). No word is written by every block in a phase except
// the barrier counter and one claim-list reservation per block:
//   - claimed blooms are staged per block in LDS and appended with one
//     reservation at the end of the phase;
//   - per-block minima (kept supports, decremented supports) go to private
//     slots; the last block to arrive at a level's final barrier reduces
//     them into the next level.
// Coherence across the 8 XCDs without cache-wide fences: every word that
// one block writes and another reads later is accessed with relaxed
// device-scope atomics, and a barrier first waits for the block's own
// memory operations (s_waitcnt 0). Read-only index data uses plain loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

namespace {

constexpr int kMB = 512;           // block size (one block per CU)
constexpr int kWPB = kMB / kWave;  // waves per block
constexpr int kC2Cap = 128;        // both-frontier middles kept in LDS per wave
constexpr int kClaimUnroll = 4;    // incidences per lane per claim step
constexpr int kXCap = 128;         // crossing edges buffered per wave
constexpr int kStage = 4096;       // claimed blooms staged per block
constexpr uint64_t kInf = ~0ull;
constexpr uint64_t kBarrierLimit = 30ull * 100000000ull;  // 30 s of the 100 MHz clock

struct MspCtl {
    unsigned bar_count;
    unsigned bar_gen;
    unsigned err;
    unsigned pad0;
    unsigned touched_n[4];           // claim-list lengths by sub-round % 3
    unsigned long long next_level;   // set by the last block at a level's end
    unsigned long long subrounds;
    unsigned long long levels;
    unsigned long long level_max;
    // diagnostics (block 0's clock, 100 MHz ticks)
    unsigned long long t_collect, t_peel, t_work_collect, t_work_peel;
    unsigned long long pmax[2], pmax_sum;
    unsigned long long hist[24];
};
static_assert(sizeof(MspCtl) <= 64 * sizeof(uint64_t), "control block overlays dg->ctl");

struct MspArgs {
    uint32_t m;
    uint32_t nbloom;       // capacity of one claim list
    uint32_t R;            // records (bloom middles)
    uint32_t diag;         // record per-phase diagnostics
    uint64_t* sup;
    uint64_t* res;
    uint32_t* stamp;
    const uint32_t* inc_off;
    const uint32_t* inc;   // per incidence: first record of the bloom
    uint64_t* rec;         // middles, compacted in place
    uint32_t* hdr;         // at a bloom's first record: live middles | upper << 31
    uint32_t* claim;       // at a bloom's first record: last claiming sub-round
    uint32_t* touched;     // 3 claim lists of bloom first records
    uint32_t* live;
    uint32_t* live_cnt;
    uint64_t* slots;       // per block: [0, G) min kept support, [G, 2G) min decremented support
    MspCtl* ctl;
    uint64_t* trace;       // optional: per sub-round (claimed, slowest-wave ticks)
    uint32_t trace_cap;
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

// Control words a block needs after a barrier; thread 0 reads them once
// (every thread of the grid loading one word would serialise on it).
struct Snap {
    unsigned touched;               // claim-list length of the next peel
    int ok;
    unsigned long long next_level;
};

enum SyncMode { kSyncPlain = 0, kSyncLevelIfIdle = 1, kSyncLevel = 2 };

// Grid barrier for a cooperative launch (every block arrives once per
// phase). The last block to arrive reduces the per-block minima into
// next_level when the level ends (mode kSyncLevel, or kSyncLevelIfIdle when
// claim list `idx` is empty). Returns false (and sets err) when a wait
// exceeds kBarrierLimit, so a broken invariant ends the kernel instead of
// hanging it. Afterwards `snap` holds touched_n[idx] and next_level.
__device__ __forceinline__ bool grid_sync(const MspArgs& a, unsigned nblocks, int mode, unsigned idx, Snap* snap,
                          uint64_t* s_red) {
    __shared__ int s_last;
    __shared__ unsigned s_gen;
    MspCtl* ctl = a.ctl;
    drain();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned const gen = ald(&ctl->bar_gen);
        drain();
        unsigned const arrived = afetch_add(&ctl->bar_count, 1u);
        s_gen = gen;
        int last = arrived == nblocks - 1;
        if (last && mode == kSyncLevelIfIdle) last = ald(&ctl->touched_n[idx]) == 0 ? 2 : 1;
        else if (last) last = mode == kSyncLevel ? 2 : 1;
        s_last = last;
    }
    __syncthreads();
    int const last = s_last;
    if (last == 2) {
        // level ends: next level = min over blocks of kept / decremented supports
        uint64_t v = kInf;
        for (uint32_t i = threadIdx.x; i < 2 * nblocks; i += blockDim.x) {
            uint64_t const x = ald(&a.slots[i]);
            v = x < v ? x : v;
        }
        v = block_min_all(v, s_red);
        if (threadIdx.x == 0) ast(&ctl->next_level, (unsigned long long)v);
    }
    if (threadIdx.x == 0) {
        int ok = 1;
        if (last) {
            ast(&ctl->bar_count, 0u);
            drain();
            afetch_add(&ctl->bar_gen, 1u);
        } else {
            uint64_t const t0 = wall_clock64();
            while (ald(&ctl->bar_gen) == s_gen) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > kBarrierLimit) {
                    ast(&ctl->err, 1u);
                    ok = 0;
                    break;
                }
            }
        }
        snap->ok = ok && ald(&ctl->err) == 0;
        snap->touched = ald(&ctl->touched_n[idx]);
        snap->next_level = ald(&ctl->next_level);
    }
    __syncthreads();
    return snap->ok != 0;
}

// Block-local staging of claimed blooms; one reservation per phase in the
// global claim list (overflow beyond the LDS buffer is written directly).
struct Stage {
    uint32_t* buf;        // LDS, kStage entries
    unsigned* cnt;        // LDS counter
    // stores the entry of each lane with `app`: `before` entries of the
    // batch precede this ballot, the batch has `tot` entries and owns LDS
    // slots [base, base + tot); slots beyond kStage go to the global list
    __device__ __forceinline__ void append_wave(bool app, uint32_t v, uint64_t bal, unsigned before, unsigned tot,
                                                unsigned base, uint32_t* tl, unsigned* tn, unsigned& gbase) const {
        unsigned const pos = base + before + __popcll(bal & lanemask_lt());
        unsigned const over_start = base > unsigned(kStage) ? base : unsigned(kStage);
        if (app) {
            if (pos < unsigned(kStage)) buf[pos] = v;
            else ast(&tl[gbase + (pos - over_start)], v);
        }
        (void)tot;
        (void)tn;
    }
    // at the end of a phase, with the whole block
    __device__ __forceinline__ void flush(uint32_t* tl, unsigned* tn, unsigned* s_tmp) const {
        __syncthreads();
        unsigned const n = *cnt < unsigned(kStage) ? *cnt : unsigned(kStage);
        if (threadIdx.x == 0) *s_tmp = n ? afetch_add(tn, n) : 0u;
        __syncthreads();
        unsigned const gb = *s_tmp;
        for (unsigned i = threadIdx.x; i < n; i += blockDim.x) ast(&tl[gb + i], buf[i]);
        __syncthreads();
        if (threadIdx.x == 0) *cnt = 0;
        __syncthreads();
    }
};

// Claims, for sub-round X, every bloom that holds edge ee (the whole wave
// works on the edge's incidence list); a bloom is claimed once per
// sub-round (max on its claim stamp). All lanes of the wave must call.
__device__ __forceinline__ void claim_edge(uint32_t ee, uint32_t X, const MspArgs& a, const Stage& st) {
    int const lane = threadIdx.x & (kWave - 1);
    uint32_t* tl = a.touched + size_t(X % 3) * a.nbloom;
    unsigned* tn = &a.ctl->touched_n[X % 3];
    uint32_t const p0 = a.inc_off[ee], p1 = a.inc_off[ee + 1];
    for (uint32_t q = p0; q < p1; q += kWave * kClaimUnroll) {
        uint32_t sb[kClaimUnroll];
        bool app[kClaimUnroll];
#pragma unroll
        for (int u = 0; u < kClaimUnroll; u++) {
            uint32_t const p = q + u * kWave + lane;
            sb[u] = p < p1 ? a.inc[p] : 0;
        }
#pragma unroll
        for (int u = 0; u < kClaimUnroll; u++) {
            uint32_t const p = q + u * kWave + lane;
            app[u] = p < p1 && afetch_max(&a.claim[sb[u]], X) < X;
        }
        uint64_t bal[kClaimUnroll];
        unsigned tot = 0;
#pragma unroll
        for (int u = 0; u < kClaimUnroll; u++) {
            bal[u] = __ballot(app[u]);
            tot += unsigned(__popcll(bal[u]));
        }
        if (tot) {
            // LDS slots for the batch; the part beyond kStage is reserved in
            // the global list at once
            unsigned base = 0, gbase = 0;
            if (lane == 0) {
                base = atomicAdd(st.cnt, tot);
                unsigned const over_start = base > unsigned(kStage) ? base : unsigned(kStage);
                unsigned const n_over = base + tot > unsigned(kStage) ? base + tot - over_start : 0u;
                if (n_over) gbase = afetch_add(tn, n_over);
            }
            base = __shfl(base, 0, kWave);
            gbase = __shfl(gbase, 0, kWave);
            unsigned before = 0;
#pragma unroll
            for (int u = 0; u < kClaimUnroll; u++) {
                st.append_wave(app[u], sb[u], bal[u], before, tot, base, tl, tn, gbase);
                before += unsigned(__popcll(bal[u]));
            }
        }
    }
}

// Per-wave buffer of edges that fell to the level; their blooms are claimed
// for the next sub-round after each bloom (or when the buffer fills).
struct CrossBuf {
    uint32_t* xl;
    uint32_t n;
    __device__ __forceinline__ void flush(uint32_t X, const MspArgs& a, const Stage& st) {
        if (!n) return;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t t = 0; t < n; t++) claim_edge(xl[t], X, a, st);
        n = 0;
    }
    __device__ __forceinline__ void push(bool has, uint32_t e, uint32_t X, const MspArgs& a, const Stage& st) {
        uint64_t const b = __ballot(has);
        uint32_t const cnt = __popcll(b);
        if (!cnt) return;
        if (n + cnt > uint32_t(kXCap)) flush(X, a, st);
        if (has) xl[n + __popcll(b & lanemask_lt())] = e;
        n += cnt;
    }
};

struct Dec {
    uint64_t* sup;
    uint32_t* stamp;
    uint64_t* res;
    uint64_t L;
    uint32_t X;
    uint64_t tmin;  // smallest decremented support still above the level
    // subtracts d from the support of x; true when x falls to the level now
    __device__ __forceinline__ bool operator()(uint32_t x, uint64_t d) {
        uint64_t const old = afetch_add(&sup[x], uint64_t(0) - d);
        uint64_t const nw = old - d;
        if (old > L && nw <= L) {
            ast(&stamp[x], X + 1);
            res[x] = L;
            return true;
        }
        if (nw > L && nw < tmin) tmin = nw;
        return false;
    }
};

__device__ __forceinline__ int classify(uint32_t s1, uint32_t s2, uint32_t X) {
    bool const dead = (s1 != 0 && s1 < X) || (s2 != 0 && s2 < X);
    if (dead) return -1;
    return int(s1 == X) + int(s2 == X);
}

__device__ __forceinline__ bool rule3(uint32_t o, uint32_t p, uint32_t q) {
    return (o < q && q < p) || (p < o && o < q) || (q < o && o < p);
}

// Decrement owed by the survivor of a one-frontier-edge middle (g is the
// frontier edge of the same middle) for the both-frontier middles.
__device__ __forceinline__ uint64_t rule_sum(const uint64_t* c2, uint32_t n2, bool froot, uint32_t up,
                                             uint32_t g) {
    uint64_t d = 0;
    for (uint32_t t = 0; t < n2; t++) {
        uint64_t const rr = c2[t];
        uint32_t const o = froot ? uint32_t(rr) : uint32_t(rr >> 32);
        uint32_t const sr = froot ? uint32_t(rr >> 32) : uint32_t(rr);
        d += up ? rule3(o, sr, g) : rule3(o, g, sr);
    }
    return d;
}

// Same sum read from the bloom in global memory (more both-frontier
// middles than the LDS list holds).
__device__ __forceinline__ uint64_t rule_scan(const MspArgs& a, const uint64_t* rec, uint32_t kb, uint32_t X, bool froot,
                              uint32_t up, uint32_t g) {
    uint64_t d = 0;
    for (uint32_t t = 0; t < kb; t++) {
        uint64_t const rr = ald(&rec[t]);
        if (classify(ald(&a.stamp[uint32_t(rr >> 32)]), ald(&a.stamp[uint32_t(rr)]), X) != 2) continue;
        uint32_t const o = froot ? uint32_t(rr) : uint32_t(rr >> 32);
        uint32_t const sr = froot ? uint32_t(rr >> 32) : uint32_t(rr);
        d += up ? rule3(o, sr, g) : rule3(o, g, sr);
    }
    return d;
}

// Processes a claimed bloom with kb <= 64 * RPT middles held in registers:
// all middles and their stamps are loaded up front, so the bloom costs a
// few dependent memory steps whatever its size. r0 holds middle `lane`.
template <int RPT>
__device__ __forceinline__ void process_bloom(const MspArgs& a, uint32_t s, uint32_t kb, uint32_t up, uint64_t r0,
                                              uint32_t X, Dec& dec, uint64_t* c2, CrossBuf& xb,
                                              const Stage& st) {
    int const lane = threadIdx.x & (kWave - 1);
    uint64_t* rec = a.rec + s;
    uint64_t r[RPT];
    uint32_t st1[RPT];
    int c[RPT];
    r[0] = r0;
#pragma unroll
    for (int i = 1; i < RPT; i++) {
        uint32_t const j = i * kWave + lane;
        r[i] = j < kb ? ald(&rec[j]) : 0;
    }
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        uint32_t const j = i * kWave + lane;
        c[i] = -1;
        st1[i] = 0;
        if (j < kb) {
            st1[i] = ald(&a.stamp[uint32_t(r[i] >> 32)]);
            c[i] = classify(st1[i], ald(&a.stamp[uint32_t(r[i])]), X);
        }
    }
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
    uint64_t const dec0 = uint64_t(n1) + n2;
    bool const c2ok = n2 <= uint32_t(kC2Cap);
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        uint32_t const e1 = uint32_t(r[i] >> 32), e2 = uint32_t(r[i]);
        bool x1 = false, x2 = false;
        uint32_t f = e1;
        if (c[i] == 0 && dec0) {
            x1 = dec(e1, dec0);
            x2 = dec(e2, dec0);
        } else if (c[i] == 1) {
            bool const froot = st1[i] != X;  // the survivor is r-v
            f = froot ? e1 : e2;
            uint32_t const g = froot ? e2 : e1;
            uint64_t const d = uint64_t(n0) + (n1 - 1) +
                               (c2ok ? rule_sum(c2, n2, froot, up, g) : rule_scan(a, rec, kb, X, froot, up, g));
            if (d) x1 = dec(f, d);
        }
        xb.push(x1, f, X + 1, a, st);
        xb.push(x2, e2, X + 1, a, st);
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
        if (lane == 0) ast(&a.hdr[s], n0 | (up << 31));
    }
}

// Blooms above 1024 middles: a counting pass, then a decrement pass, then
// (when the both-frontier list overflowed) a separate compaction pass.
__device__ __forceinline__ void process_large(const MspArgs& a, uint32_t s, uint32_t kb, uint32_t up, uint32_t X, Dec& dec,
                              uint64_t* c2, CrossBuf& xb, const Stage& st) {
    int const lane = threadIdx.x & (kWave - 1);
    uint64_t* rec = a.rec + s;
    uint32_t n0 = 0, n1 = 0, n2 = 0;
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
    }
    __builtin_amdgcn_wave_barrier();
    uint64_t const dec0 = uint64_t(n1) + n2;
    bool const inline_compact = n2 <= uint32_t(kC2Cap);
    uint32_t wpos = 0;
    for (uint32_t j0 = 0; j0 < kb; j0 += kWave) {
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
            uint64_t const d = uint64_t(n0) + (n1 - 1) +
                               (inline_compact ? rule_sum(c2, n2, froot, up, g) : rule_scan(a, rec, kb, X, froot, up, g));
            if (d) x1 = dec(f, d);
        }
        if (inline_compact) {
            uint64_t const keepm = __ballot(c == 0);
            if (c == 0) ast(&rec[wpos + __popcll(keepm & lanemask_lt())], r);
            wpos += __popcll(keepm);
        }
        xb.push(x1, f, X + 1, a, st);
        xb.push(x2, e2, X + 1, a, st);
    }
    if (!inline_compact) {
        for (uint32_t j0 = 0; j0 < kb; j0 += kWave) {
            uint32_t const j = j0 + lane;
            int c = -1;
            uint64_t r = 0;
            if (j < kb) {
                r = ald(&rec[j]);
                c = classify(ald(&a.stamp[uint32_t(r >> 32)]), ald(&a.stamp[uint32_t(r)]), X);
            }
            uint64_t const keepm = __ballot(c == 0);
            if (c == 0) ast(&rec[wpos + __popcll(keepm & lanemask_lt())], r);
            wpos += __popcll(keepm);
        }
    }
    if (lane == 0) ast(&a.hdr[s], wpos | (up << 31));
}

__global__ __launch_bounds__(kMB) void k_msp_persist(MspArgs a) {
    __shared__ uint32_t s_part[kWPB + 1];
    __shared__ uint64_t s_red[kWPB + 1];
    __shared__ uint64_t s_c2[kWPB][kC2Cap];
    __shared__ uint32_t s_xl[kWPB][kXCap];
    __shared__ uint32_t s_stage[kStage];
    __shared__ unsigned s_stage_n, s_tmp;
    __shared__ Snap snap;
    uint32_t const G = gridDim.x, B = blockIdx.x, tid = threadIdx.x;
    int const lane = tid & (kWave - 1), wv = tid / kWave;
    uint32_t const gw = B * kWPB + wv, nw = G * kWPB;
    uint32_t const seg0 = uint32_t(uint64_t(a.m) * B / G), seg1 = uint32_t(uint64_t(a.m) * (B + 1) / G);
    MspCtl* ctl = a.ctl;
    uint64_t* slot_m1 = a.slots + B;
    uint64_t* slot_dec = a.slots + G + B;
    if (tid == 0) s_stage_n = 0;
    Stage const st{s_stage, &s_stage_n};

    // ---- init: support-0 edges are final (process_support_zero); the rest
    // form this block's live segment; the first level is their min support
    {
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
            if (keep) a.live[w + ex] = i;
            w += tot;
        }
        if (tid == 0) a.live_cnt[B] = w - seg0;
        uint64_t const bm = block_min_all(mn, s_red);
        if (tid == 0) {
            ast(slot_m1, (unsigned long long)bm);
            ast(slot_dec, (unsigned long long)kInf);
        }
    }
    if (!grid_sync(a, G, kSyncLevel, 2, &snap, s_red)) return;
    uint64_t L = snap.next_level;
    uint32_t X = 2, k = 0;
    bool collect = true;
    uint64_t tmin = kInf;       // this thread's smallest decremented support of the level
    uint64_t dec_written = kInf;
    uint32_t nt = 0;            // claimed blooms of the sub-round peeled next
    uint64_t t_col = 0, t_pe = 0, w_col = 0, w_pe = 0, pmax_sum = 0;
    while (L != kInf) {
        uint64_t const tp0 = wall_clock64();
        if (collect) {
            // ---- collect: frontier of level L from this block's live segment
            if (B == 0 && tid == 0) ast(&ctl->touched_n[(X + 1) % 3], 0u);
            uint32_t const cnt = a.live_cnt[B];
            __syncthreads();
            uint32_t w = seg0;
            uint64_t mn = kInf;
            for (uint32_t base = seg0; base < seg0 + cnt; base += kMB) {
                uint32_t const i = base + tid;
                uint32_t e = 0;
                bool keep = false, take = false;
                if (i < seg0 + cnt) {
                    // an edge finished at an earlier level has support below L
                    // (it was peeled at a level < L with support <= that level);
                    // a live edge has support >= L
                    e = a.live[i];
                    uint64_t const v = ald(&a.sup[e]);
                    if (v == L) {
                        take = true;
                    } else if (v > L) {
                        keep = true;
                        mn = v < mn ? v : mn;
                    }
                }
                uint32_t tot = 0;
                uint32_t const ex = block_excl_scan<kMB>(keep ? 1u : 0u, s_part, tot);
                if (keep) a.live[w + ex] = e;
                w += tot;
                if (take) {
                    ast(&a.stamp[e], X);
                    a.res[e] = L;
                }
                uint64_t pend = __ballot(take);
                while (pend) {
                    int const src = __ffsll((long long)pend) - 1;
                    pend &= pend - 1;
                    claim_edge(__shfl(e, src, kWave), X, a, st);
                }
            }
            if (tid == 0) a.live_cnt[B] = w - seg0;
            uint64_t const bm = block_min_all(mn, s_red);
            if (tid == 0) {
                ast(slot_m1, (unsigned long long)bm);
                ast(slot_dec, (unsigned long long)kInf);
            }
            tmin = kInf;
            dec_written = kInf;
            st.flush(a.touched + size_t(X % 3) * a.nbloom, &ctl->touched_n[X % 3], &s_tmp);
        } else {
            // ---- peel sub-round X: every claimed bloom once
            if (B == 0 && tid == 0) ast(&ctl->touched_n[(X + 2) % 3], 0u);
            const uint32_t* tl = a.touched + size_t(X % 3) * a.nbloom;
            Dec dec{a.sup, a.stamp, a.res, L, X, tmin};
            uint64_t* c2 = s_c2[wv];
            CrossBuf xb{s_xl[wv], 0};
            uint64_t const tw0 = wall_clock64();
            for (uint32_t ti = gw; ti < nt; ti += nw) {
                uint32_t const s = ald(&tl[ti]);
                // the header and the first 64 middles are loaded together
                uint32_t const hdr = ald(&a.hdr[s]);
                uint64_t const r0 = (s + lane < a.R) ? ald(&a.rec[s + lane]) : 0;
                uint32_t const kb = hdr & 0x7FFFFFFFu, up = hdr >> 31;
                if (kb < 2) continue;
                if (kb <= uint32_t(kWave)) process_bloom<1>(a, s, kb, up, r0, X, dec, c2, xb, st);
                else if (kb <= uint32_t(4 * kWave)) process_bloom<4>(a, s, kb, up, r0, X, dec, c2, xb, st);
                else if (kb <= uint32_t(16 * kWave)) process_bloom<16>(a, s, kb, up, r0, X, dec, c2, xb, st);
                else process_large(a, s, kb, up, X, dec, c2, xb, st);
                xb.flush(X + 1, a, st);
            }
            tmin = dec.tmin;
            uint64_t const tw = wall_clock64() - tw0;
            st.flush(a.touched + size_t((X + 1) % 3) * a.nbloom, &ctl->touched_n[(X + 1) % 3], &s_tmp);
            // publish this block's smallest decremented support of the level
            uint64_t const bdec = block_min_all(tmin, s_red);
            if (tid == 0 && bdec < dec_written) ast(slot_dec, (unsigned long long)bdec);
            dec_written = bdec < dec_written ? bdec : dec_written;
            if (a.diag) {
                uint64_t const mx = ~block_min_all(~tw, s_red);  // slowest wave of the block
                if (tid == 0) afetch_max(&ctl->pmax[X & 1], (unsigned long long)mx);
            }
        }
        uint64_t const tp1 = wall_clock64();
        // a collect fills the claim list of X; a peel fills that of X + 1
        unsigned const idx = collect ? X % 3 : (X + 1) % 3;
        if (!grid_sync(a, G, collect ? kSyncPlain : kSyncLevelIfIdle, idx, &snap, s_red)) return;
        uint64_t const tp2 = wall_clock64();
        nt = snap.touched;
        if (collect) {
            t_col += tp2 - tp0;
            w_col += tp1 - tp0;
            collect = false;
            continue;
        }
        t_pe += tp2 - tp0;
        w_pe += tp1 - tp0;
        if (a.diag && B == 0 && tid == 0) {
            unsigned long long const pm = ald(&ctl->pmax[X & 1]);
            ast(&ctl->pmax[X & 1], 0ull);
            pmax_sum += pm;
            int lb = 0;
            while (lb < 23 && (pm >> (lb + 1))) lb++;
            ctl->hist[lb]++;
            if (a.trace && X < a.trace_cap) a.trace[X] = (uint64_t(nt) << 32) | (pm & 0xFFFFFFFFull);
        }
        if (nt) {
            X++;
            continue;
        }
        // no bloom claimed for X + 1: the level is over (X + 1 is skipped)
        L = snap.next_level;
        X += 2;
        k++;
        collect = true;
    }
    if (B == 0 && tid == 0) {
        ctl->subrounds = X - 1;
        ctl->levels = k;
        ctl->t_collect = t_col;
        ctl->t_peel = t_pe;
        ctl->t_work_collect = w_col;
        ctl->t_work_peel = w_pe;
        ctl->pmax_sum = pmax_sum;
    }
}

__global__ void k_msp_ctl_init(MspCtl* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(ctl);
        for (size_t i = 0; i < sizeof(MspCtl) / sizeof(unsigned long long); i++) w[i] = 0;
        ctl->next_level = kInf;
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
int bt_msp_index_peel(bt_dgraph* dg, uint32_t* levels_out) {
    bt_ctx* c = dg->ctx;
    hipStream_t s = c->stream;
    uint32_t const m = dg->m;
    bt_timings const count_t = c->last;
    bt_timer tm(c);
    BT_TRY(tm.start(5));
    dg->has_support = false;
    BT_TRY(bt_index_build(dg));
    BT_TRY(tm.start(6));
    bt_index& ix = c->ix;
    unsigned const G = unsigned(c->cu_count);  // one block per CU keeps the grid barrier short
    int per_cu = 0;
    BT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_msp_persist),
                                                            kMB, 0));
    if (per_cu < 1) return BT_ERR_DEVICE;
    BT_TRY(bt_buf_reserve(ix.live, sizeof(uint32_t) * std::max<size_t>(m, 1)));
    size_t const cnt_bytes = (sizeof(uint32_t) * (G + 1) + 7) & ~size_t(7);
    BT_TRY(bt_buf_reserve(ix.seg, cnt_bytes + sizeof(uint64_t) * 2 * G));
    MspCtl* ctl = reinterpret_cast<MspCtl*>(dg->ctl);
    MspArgs a{};
    a.m = m;
    a.nbloom = uint32_t(std::max<uint64_t>(ix.NB, 1));
    a.R = uint32_t(ix.R);
    a.sup = dg->sup;
    a.res = dg->res;
    a.stamp = dg->list0;
    a.inc_off = static_cast<const uint32_t*>(ix.inc_off.ptr);
    a.inc = static_cast<const uint32_t*>(ix.inc.ptr);
    a.rec = static_cast<uint64_t*>(ix.vals1.ptr);
    a.hdr = static_cast<uint32_t*>(ix.bk.ptr);
    a.claim = static_cast<uint32_t*>(ix.bclaim.ptr);
    a.touched = static_cast<uint32_t*>(ix.touched.ptr);
    a.live = static_cast<uint32_t*>(ix.live.ptr);
    a.live_cnt = static_cast<uint32_t*>(ix.seg.ptr);
    a.slots = reinterpret_cast<uint64_t*>(static_cast<char*>(ix.seg.ptr) + cnt_bytes);
    a.ctl = ctl;
    static bt_buf trace_buf;
    bool const tracing = getenv("BT_PEEL_TRACE_FILE") != nullptr;
    a.diag = (getenv("BT_PEEL_TRACE") || tracing) ? 1u : 0u;
    if (tracing) {
        a.trace_cap = m + 8;
        BT_TRY(bt_buf_reserve(trace_buf, sizeof(uint64_t) * a.trace_cap));
        a.trace = static_cast<uint64_t*>(trace_buf.ptr);
        BT_HIP_TRY(hipMemsetAsync(a.trace, 0, sizeof(uint64_t) * a.trace_cap, s));
    }
    k_msp_ctl_init<<<1, 64, 0, s>>>(ctl);
    BT_HIP_TRY(hipGetLastError());
    void* params[] = {&a};
    BT_HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_msp_persist), dim3(G), dim3(kMB), params,
                                          0, s));
    BT_TRY(tm.start(7));
    // the level count the reference logs is the largest result
    BT_HIP_TRY(hipMemsetAsync(&ctl->level_max, 0, sizeof(uint64_t), s));
    if (m)
        k_max_res<<<grid_for(m, 256, unsigned(c->cu_count) * 8), 256, 0, s>>>(
            m, dg->res, reinterpret_cast<uint64_t*>(&ctl->level_max));
    BT_HIP_TRY(hipGetLastError());
    MspCtl h{};
    BT_HIP_TRY(hipMemcpyAsync(&h, ctl, sizeof(MspCtl), hipMemcpyDeviceToHost, s));
    BT_HIP_TRY(hipStreamSynchronize(s));
    if (h.err) return BT_ERR_INTERNAL;
    if (a.diag) {
        std::fprintf(stderr,
                     "[bt_peel] index %.3f ms; phases: collect %.3f ms (block0 work %.3f), peel %.3f ms (block0 work "
                     "%.3f), slowest-wave sum %.3f ms, sub-rounds %llu, levels %llu, blocks %u, R %llu, blooms "
                     "%llu\n",
                     tm.ms(5, 6), h.t_collect * 1e-5, h.t_work_collect * 1e-5, h.t_peel * 1e-5, h.t_work_peel * 1e-5,
                     h.pmax_sum * 1e-5, h.subrounds, h.levels, G, (unsigned long long)ix.R,
                     (unsigned long long)ix.NB);
        std::fprintf(stderr, "[bt_peel] slowest-wave histogram (log2 10ns ticks: phases):");
        for (int i = 0; i < 24; i++)
            if (h.hist[i]) std::fprintf(stderr, " %d:%llu", i, h.hist[i]);
        std::fprintf(stderr, "\n");
    }
    if (tracing) {
        std::vector<uint64_t> tr(std::min<uint64_t>(h.subrounds + 2, a.trace_cap));
        BT_HIP_TRY(hipMemcpy(tr.data(), a.trace, sizeof(uint64_t) * tr.size(), hipMemcpyDeviceToHost));
        FILE* f = std::fopen(getenv("BT_PEEL_TRACE_FILE"), "w");
        if (f) {
            for (size_t x = 0; x < tr.size(); x++)
                std::fprintf(f, "%zu %llu %llu\n", x, (unsigned long long)(tr[x] >> 32),
                             (unsigned long long)(tr[x] & 0xFFFFFFFFull));
            std::fclose(f);
        }
    }
    c->last = count_t;
    c->last.peel_ms = tm.ms(5, 7);
    c->last.peel_kernel_ms = tm.ms(6, 7);
    c->last.peel_kernel_launches = 1;
    c->last.total_ms = count_t.total_ms + c->last.peel_ms;
    c->last.sub_rounds = h.subrounds;
    c->last.levels_scanned = h.levels;
    if (h.level_max > 0xFFFFFFFFull) return BT_ERR_OVERFLOW;
    if (levels_out) *levels_out = uint32_t(h.level_max);
    return BT_OK;
}
