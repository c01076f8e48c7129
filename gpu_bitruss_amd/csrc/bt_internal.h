// Internal declarations shared by the HIP translation units of
// libbitruss_hip.so. Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <mutex>
#include <utility>
#include <vector>
#include <map>
#include <set>

#include <cstddef>
#include <cstdint>
#include <string>

#include "bitruss_hip.h"

#define BT_HIP_TRY(expr)                                   \
    do {                                                   \
        hipError_t bt_e_ = (expr);                         \
        if (bt_e_ != hipSuccess) {                         \
            bt_set_hip_error(bt_e_, #expr, __FILE__, __LINE__); \
            return BT_ERR_DEVICE;                          \
        }                                                  \
    } while (0)

#define BT_TRY(expr)                     \
    do {                                 \
        int bt_s_ = (expr);              \
        if (bt_s_ != BT_OK) return bt_s_; \
    } while (0)

void bt_set_hip_error(hipError_t e, const char* what, const char* file, int line);

// A device buffer that grows on demand and is reused between calls.
struct bt_ctx;
// A growable device buffer. A buffer owned by a context grows through the
// context's block cache (the old block goes back to the cache, reused in
// stream order), so growing it never waits for other contexts' work.
struct bt_buf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bt_ctx* owner = nullptr;
};
int bt_buf_reserve(bt_buf& b, size_t bytes);
void bt_buf_free(bt_buf& b);

// XCDs of an MI355X (8 x 32 CUs). Workgroups of a launch are dispatched
// round-robin over them, so a grid that is a multiple of kXcds takes the
// same number of CUs on each.
constexpr int kXcds = 8;

// Claim stamps past the R real ones that the peel's claim steps use for
// lanes with nothing to claim (one per thread and unroll slot of a block).
constexpr uint32_t kClaimDummiesPerBlock = 2048;

// Records and blooms are named by u32 record indices, and the claim stamps
// of the peel's dummy lanes follow the records: the index holds at most
// 2^32 - 1 - (dummies of the largest grid) records.
constexpr uint64_t kMaxIndexRecords = 0xFFFFFFFFull - 1024ull * kClaimDummiesPerBlock;
// dg->ctl word (u32) where the index build leaves its largest bloom size
constexpr int kCtlBloomMax = 120;

// Bloom index of a graph (wedges grouped by (root, far end) with >= 2
// middles), rebuilt by each msp call; buffers are reused across calls.
struct bt_index {
    bt_buf keys0, keys1;  // record keys (emission order, sorted)
    bt_buf vals0, vals1;  // record values; vals1 = sorted middles (id(r-v) << 32 | id(v-w))
    bt_buf bid;           // u32 per record: inclusive head count (bloom id + 1)
    bt_buf bstart;        // u32 per bloom: first record, plus a sentinel R
    bt_buf bk;            // u32 per record, used at bloom starts: live middles | root-is-upper << 31
    bt_buf bclaim;        // u32 per record, used at bloom starts: last claiming sub-round
    bt_buf inc_off;       // u64 per edge + 1: incidence offsets (2R may pass 2^32)
    bt_buf inc_cur;       // u32 per edge: fill cursors
    bt_buf inc;           // u32 per incidence: first record of the bloom
    bt_buf touched;       // u32 3 x blooms: shared claimed-bloom lists
    bt_buf area;          // u32 3 x blocks x cap: per-block claimed-bloom lists
    bt_buf bgraph;        // u32 per record, used at bloom starts: graph of the bloom (batches)
    bt_buf glev;          // u64 2 x graphs: per-graph levels of the peel (batches)
    bt_buf live;          // u32 per edge: per-block live-edge segments
    bt_buf seg;           // u32 per block + 1: segment offsets, then counts
    bt_buf c2s;           // u32 per peel wave: sorted both-frontier lists
    bt_buf hdirty;        // u8 2 x blooms: h-index rounds' change marks
    bt_buf flip;          // u32 per 32 edges: mspp's turned-around pivots
    uint64_t R = 0, NB = 0;
};

struct bt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 0;
    bt_timings last{};
    hipEvent_t ev[8]{};
    // scratch shared by all graphs of the context
    bt_buf sort_tmp;   // rocprim temporary storage
    bt_buf pool;       // dense per-root count tables for hub roots
    bt_buf pinned_dummy;
    uint64_t* h_pinned = nullptr;  // small pinned host buffer for read-backs
    bt_index ix;
    int coop = 0;                  // cooperative launch supported
    int msp_path = 0;              // BT_MSP_PATH_*
    double index_mem_frac = 0.5;   // share of free device memory the index may use
    uint64_t batch_wedges = 0;     // wedge bound per batch union (0: the index limits)
    int batch_levels = 1;          // BT_BATCH_LEVELS_*: how a union's msp walks levels
    int batch_lanes = 2;           // batches: lanes (contexts on CU shares) run side by side
    int peel_blocks = 0;           // CUs the persistent peel uses (0: all)
    int peel_slice_ms = 2000;      // time slice of a persistent peel launch (0: one launch)
    uint64_t peel_max_level = 0;   // msp stops after this level (0: never); BT_PARTIAL
    int peel_c2cap = 128;          // both-frontier middles summed from the on-chip list
    uint64_t peel_window_edges = uint64_t(4) << 20;  // live edges from which the collect uses a window
    int peel_big_middles = 1024;   // blooms beyond this are walked by a whole workgroup
    int hidx_graph_rounds = 8;     // h-index rounds per captured graph (0: one by one)
    int hidx_sorted = 1;           // h-index counts over sorted chunks of the larger blooms (0: streams only)
    // counting launch settings per graph-size class (bt_tune_*), and the
    // settings an autotuning run imposes on the calls it times (null: the
    // table decides)
    bt_count_config tune[BT_TUNE_CLASSES]{};
    const bt_count_config* count_force = nullptr;
    // device blocks given back by freed graphs and batches, reused by the
    // next ones (hipFree would wait for every stream on the device)
    std::mutex cache_mu;
    std::multimap<size_t, void*> cache;
    size_t cache_bytes = 0;
    // graphs and batches made on this context and not freed yet: destroying
    // the context releases their device memory and detaches them, so that
    // freeing them afterwards only deletes the handle
    std::mutex live_mu;
    std::set<bt_dgraph*> live_graphs;
    std::set<bt_batch*> live_batches;
};

// Releases a batch's device memory and lane contexts and detaches it from
// its context (bt_ctx_destroy of that context; bt_batch_free then deletes it).
void bt_batch_detach(bt_batch* b);

// Device memory for graphs and batches from the context's cache of freed
// blocks (best fit within twice the request), else hipMalloc; *got is the
// block's size, to be passed back to bt_ctx_release.
int bt_ctx_alloc(bt_ctx* c, size_t bytes, void** p, size_t* got);
// Give a block back to the context's cache once the work queued on its
// stream no longer needs it (the caller synchronises the stream first).
void bt_ctx_release(bt_ctx* c, void* p, size_t bytes);
// Free every cached block of the context (after its stream has drained).
int bt_ctx_trim_cache(bt_ctx* c);

// Device-resident graph plus the per-graph scratch of the algorithms.
struct bt_dgraph {
    bt_ctx* ctx = nullptr;
    uint32_t u_num = 0, l_num = 0, n = 0, m = 0;
    // the uploaded reference CSR (original ids)
    uint32_t* nbr = nullptr;    // 2m
    uint32_t* off = nullptr;    // n+1
    uint32_t* deg = nullptr;    // n
    uint32_t* edges = nullptr;  // 2m
    uint32_t* eids = nullptr;   // 2m
    // priority-relabelled CSR (rank = position in (degree, id) order)
    uint32_t* rank = nullptr;   // n   : original id -> rank
    uint32_t* r_off = nullptr;  // n+1
    uint32_t* r_nbr = nullptr;  // 2m  : neighbour ranks, ascending per row
    uint32_t* r_eid = nullptr;  // 2m
    uint32_t* r_lo = nullptr;   // n   : neighbours of lower rank (a row prefix)
    uint8_t* r_up = nullptr;    // n   : rank belongs to an upper vertex
    uint32_t* r_p = nullptr;    // 2m  : per lower neighbour v of r, #N(v) below r
    uint64_t* r_w = nullptr;    // n   : priority wedges rooted at r
    uint64_t* sup = nullptr;    // m   : ebfc supports (then msp working supports)
    uint64_t* res = nullptr;    // m   : hidx / msp results
    uint64_t* aux = nullptr;    // m   : second hidx buffer
    uint8_t* state = nullptr;   // m   : msp edge state
    uint32_t* list0 = nullptr;  // m   : frontier buffers / root lists
    uint32_t* list1 = nullptr;  // m
    uint64_t* ctl = nullptr;    // 128 control words
    uint64_t* r_key = nullptr;  // 2m  : sort keys (prep only)
    uint64_t* r_key2 = nullptr; // 2m
    uint32_t* r_val2 = nullptr; // 2m
    bool has_support = false;   // sup holds the ebfc result of this graph
    // batches (the graph is a disjoint union): graph of each edge, and per
    // graph the last hidx round that changed one of its edges
    const uint32_t* egraph = nullptr;  // m
    uint32_t* graph_last = nullptr;    // graphs of the batch
    uint32_t ngraph = 0;               // graphs of the batch (0: not a batch)
    std::vector<std::pair<void*, size_t>> blocks;  // device blocks of the arrays above
};

// per-call helpers
struct bt_timer {
    bt_ctx* ctx;
    explicit bt_timer(bt_ctx* c) : ctx(c) {}
    int start(int slot) { return hipEventRecord(ctx->ev[slot], ctx->stream) == hipSuccess ? BT_OK : BT_ERR_DEVICE; }
    double ms(int a, int b);
};

// A fixed device region that a call takes its scratch from instead of the
// context's growable buffers (the allocation-free entry point).
struct bt_fixed_scratch {
    void* ptr;
    size_t bytes;
};

// Algorithm entry points implemented in the .hip files (inputs checked by
// the C-ABI layer). With `fixed`, the counting's scratch (sort storage, hub
// tables) comes from that region, nothing is allocated, and a graph that
// needs more returns BT_ERR_SCRATCH.
int bt_prep_priority(bt_dgraph* dg, const bt_fixed_scratch* fixed = nullptr);
int bt_count_impl(bt_dgraph* dg, bool per_edge, uint64_t* total, uint64_t* support_max, bool emit = false,
                  const bt_fixed_scratch* fixed = nullptr);
// inc_records: incidences name the record holding the edge instead of the
// bloom's first record (the h-index rounds need the edge's middle)
int bt_index_build(bt_dgraph* dg, bool inc_records = false);
// rule: 0 msp, 1 mspp, 2 the textbook (exact) peel
int bt_msp_index_peel(bt_dgraph* dg, uint32_t* levels, int rule = 0);
int bt_msp_merge_peel(bt_dgraph* dg, uint32_t* levels);
// ix.R value meaning "records were not emitted" (index path unavailable)
constexpr uint64_t kNoIndex = ~0ull;
int bt_msp_impl(bt_dgraph* dg, uint32_t* levels, int rule = 0);
int bt_hidx_impl(bt_dgraph* dg, uint32_t* rounds);
int bt_hidx_index(bt_dgraph* dg, uint32_t* rounds);

// Cooperative (grid-barrier) kernels share a device's CUs: each of their
// blocks needs a CU of its own (LDS), so kernels whose block counts add up
// to at most the CU count are resident together, while more could each
// hold CUs that the other's blocks wait for. A kernel takes its CUs before
// the launch and gives them back when it has finished.
class bt_cu_lease {
public:
    bt_cu_lease(int device, int total_cus, int want);
    ~bt_cu_lease();
    bt_cu_lease(const bt_cu_lease&) = delete;
    bt_cu_lease& operator=(const bt_cu_lease&) = delete;

private:
    int device_, want_;
};

// Read back a few control words (synchronises the stream).
int bt_read_ctl(bt_dgraph* dg, int first, int count, uint64_t* out);

// The counting settings a call on dg uses: the forced ones of an
// autotuning run, else the context's entry for dg's size class.
bt_count_config bt_count_settings(const bt_dgraph* dg);
// Whether cfg is one of the settings the counting kernels are built for.
bool bt_count_config_valid(const bt_count_config& cfg);
