/*
 * bitruss_hip.h — C-ABI of the MI355X (gfx950) bitruss library.
 *
 * The reference has no plugin or operator API: its boundary is a set of
 * C++ free functions on a mutable Graph (src/bfc/bfc.cuh:17-23,
 * src/bitruss/bitruss.cuh:17-24) called only from main.cu:96-129. Each
 * entry point below replaces one of them; the reference's CUDA functions
 * allocate, copy and free all device memory per call and exit() on error,
 * while here the device graph stays resident between calls and every call
 * returns a status (0 = BT_OK, text via bt_status_str in bt_graph.h).
 *
 * Results are bit-exact with the reference CPU semantics, widened to
 * uint64 (the reference keeps supports in int32).
 *
 * Threading: one bt_ctx per host thread; calls on one context are
 * serialised on its HIP stream. Contexts on the same device may be used
 * concurrently from different threads.
 */
#ifndef BITRUSS_HIP_H
#define BITRUSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "bt_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bt_ctx bt_ctx;       /* device, stream, scratch pool  */
typedef struct bt_dgraph bt_dgraph; /* a graph resident on a device  */

/* Per-call measurements of the last algorithm call on a context
 * (HIP events on the context's stream). */
typedef struct bt_timings {
    double total_ms;               /* whole call, inputs resident            */
    double prep_ms;                /* priority relabel of the CSR            */
    double count_ms;               /* butterfly / support counting           */
    double peel_ms;                /* msp peeling or hidx rounds             */
    double peel_kernel_ms;         /* sum of the peel kernel durations       */
    uint64_t peel_kernel_launches; /* launches of the peel kernel (msp: its
                                      time slices)                           */
    double count_kernel_ms;        /* sum of the counting kernel durations   */
    uint64_t count_kernel_launches;
    uint64_t wedges_count;         /* priority wedges counted                */
    uint64_t wedges_peel;          /* reserved                               */
    uint64_t sub_rounds;           /* msp sub-rounds / hidx rounds           */
    uint64_t levels_scanned;       /* msp levels visited                     */
    uint64_t sorted_blooms;        /* msp: bloom visits that sorted their
                                      both-frontier middles                  */
} bt_timings;

/* Number of HIP devices (0 when none). Replaces cudaGetDeviceCount use
 * in main.cu:54-55. */
int bt_device_count(int* count);

/* Device properties for `--device_info` (utility.h:127-146 prints the same
 * table from cudaGetDeviceProperties). */
typedef struct bt_device_props {
    char name[256];
    char arch[64];                 /* gcnArchName, e.g. gfx950:sramecc+:xnack- */
    int memory_bus_width_bits;
    int memory_clock_khz;
    uint64_t total_global_mem;
    int multiprocessor_count;
    uint64_t shared_mem_per_block;
    int max_threads_per_multiprocessor;
    int warp_size;
    int l2_cache_bytes;
} bt_device_props;
int bt_device_info(int device, bt_device_props* out);

/* Create a context on `device` (main.cu:58-65 device selection): one
 * HIP stream and a reusable scratch pool. */
int bt_ctx_create(int device, bt_ctx** out);
/* Destroying a context gives back the device memory of the graphs and
 * batches still uploaded on it and detaches them: their handles stay valid
 * only for bt_graph_free / bt_batch_free, which then just delete them. */
int bt_ctx_destroy(bt_ctx* ctx);
int bt_ctx_last_timings(const bt_ctx* ctx, bt_timings* out);
/* Give back the device memory the context keeps between calls: the blocks
 * of freed graphs and batches it caches for reuse, and its index and
 * scratch buffers (they grow again on the next call). No counterpart in the
 * reference, which frees everything per call. */
int bt_ctx_trim(bt_ctx* ctx);

/* Context options (no counterpart in the reference, whose launch shapes are
 * compile-time macros in src/util/config.h:6-13).
 *   BT_OPT_MSP_PATH        BT_MSP_PATH_AUTO (default): msp and hidx run over the
 *                          bloom index when it fits (msp: one persistent
 *                          launch), else over the merge enumeration of the
 *                          reference's wedges; _MERGE / _INDEX force one path
 *                          (_INDEX fails with BT_ERR_BUDGET when it cannot run)
 *   BT_OPT_INDEX_MEM_PCT   share of free device memory the emitted index
 *                          records may use (1..95, default 50; the sorted
 *                          index must then fit the free memory)
 *   BT_OPT_BATCH_WEDGES    batches uploaded afterwards put consecutive graphs
 *                          in one union up to this bound of priority wedges
 *                          (0, the default: as many as one bloom index holds)
 *   BT_OPT_BATCH_LEVELS    how msp walks the levels of a batch union:
 *                          BT_BATCH_LEVELS_UNION (default): the union's levels
 *                          in order (a graph sits out the levels it does not
 *                          have); BT_BATCH_LEVELS_PER_GRAPH: in each step
 *                          every graph peels its own next level (more phases
 *                          of less work; slower on the measured mixes).
 *                          Results are the same; only the speed differs.
 *   BT_OPT_PEEL_BLOCKS     CUs (one workgroup each) the persistent msp peel
 *                          occupies (0, the default: all). Peels of several
 *                          contexts run side by side while their CU counts
 *                          fit the device; otherwise they wait for CUs.
 *   BT_OPT_PEEL_SLICE_MS   the persistent msp peel of a graph returns at the
 *                          first level boundary after this many milliseconds
 *                          and is launched again to resume (default 2000;
 *                          0: one launch). Results are the same.
 *   BT_OPT_BATCH_LANES     batches uploaded afterwards spread their graphs
 *                          over this many lanes (1..64, default 2), each a
 *                          context with its own stream and an equal share of
 *                          the CUs for its persistent peels; the lanes run
 *                          side by side. Results are the same. Streams beyond
 *                          the process's hardware queues (GPU_MAX_HW_QUEUES,
 *                          4 by default) share queues, which runs their
 *                          kernels one after the other.
 *   BT_OPT_PEEL_MAX_LEVEL  msp of a plain graph stops after this level (0, the
 *                          default: never) and bt_msp returns BT_PARTIAL with
 *                          the state of that prefix: the edges peeled so far
 *                          hold their result, the others their current
 *                          support. For checking a prefix of a decomposition
 *                          too long to compare whole.
 *   BT_OPT_PEEL_C2CAP      both-frontier middles of a bloom summed from the
 *                          on-chip list before the sorted lists take over
 *                          (0..128, default 128; 0 forces the sorted lists).
 *                          Results are the same.
 *   BT_OPT_PEEL_WINDOW_EDGES  live edges from which the level collect scans a
 *                          support window (default 4,194,304; 0 uses it on
 *                          every graph). Results are the same.
 *   BT_OPT_HIDX_GRAPH_ROUNDS  h-index rounds per captured HIP graph (default 8;
 *                          odd values round up): the round loop keeps its
 *                          state on the device and the host reads it back
 *                          once per graph launch; 0 runs the rounds one by
 *                          one with a read-back after each, as the reference
 *                          (bitruss_hidx.cu:219-231)
 *   BT_OPT_PEEL_BIG_MIDDLES  blooms with more middles are walked by a whole
 *                          workgroup (default 1,024). Results are the same.
 *   BT_OPT_HIDX_SORTED     1 (default): each h-index round keeps the values
 *                          of the blooms above 32 middles in sorted chunks,
 *                          so that an edge counts its values >= x by binary
 *                          searches where that is cheaper than a pass; 0:
 *                          every edge streams its values. Results and round
 *                          counts are the same. */
enum { BT_OPT_MSP_PATH = 1, BT_OPT_INDEX_MEM_PCT = 2, BT_OPT_BATCH_WEDGES = 3, BT_OPT_BATCH_LEVELS = 4,
       BT_OPT_PEEL_BLOCKS = 5, BT_OPT_PEEL_SLICE_MS = 6, BT_OPT_BATCH_LANES = 7, BT_OPT_PEEL_MAX_LEVEL = 8,
       BT_OPT_PEEL_C2CAP = 9, BT_OPT_PEEL_WINDOW_EDGES = 10, BT_OPT_PEEL_BIG_MIDDLES = 11,
       BT_OPT_HIDX_GRAPH_ROUNDS = 12, BT_OPT_HIDX_SORTED = 13 };
enum { BT_MSP_PATH_AUTO = 0, BT_MSP_PATH_MERGE = 1, BT_MSP_PATH_INDEX = 2 };
enum { BT_BATCH_LEVELS_PER_GRAPH = 0, BT_BATCH_LEVELS_UNION = 1 };
int bt_ctx_set_option(bt_ctx* ctx, int option, int64_t value);

/* Copy a host CSR (the reference Graph arrays, graph.h:43-51) to the
 * device and keep it resident. Replaces the per-call cudaMalloc/cudaMemcpy
 * of bfc_vpp.cu:233-262, bitruss_hidx.cu:196-211, bitruss_msp.cu:447-465. */
int bt_graph_upload(bt_ctx* ctx, const bt_host_graph* g, bt_dgraph** out);
int bt_graph_free(bt_dgraph* dg);

/* Total butterfly count. Replaces bfc_vpp (bfc_vpp.cu:321-385), whose
 * total is only logged (bfc_vpp.cu:366). */
int bt_bfc(bt_ctx* ctx, bt_dgraph* dg, uint64_t* total);

/* Per-edge butterfly support. Replaces bfc_evpp (bfc_vpp.cu:233-319),
 * which writes g->edge_support and g->support_max. support_out (m
 * entries) may be NULL to keep the result on the device; total and
 * support_max may be NULL. */
int bt_ebfc(bt_ctx* ctx, bt_dgraph* dg, uint64_t* support_out, uint64_t* total,
            uint64_t* support_max);

/* Allocation-free per-edge support. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same results as
 * bt_ebfc on an uploaded graph, but the call's scratch (sort storage and
 * hub-root count tables) is a fixed device array of BT_STATIC_SCRATCH_BYTES
 * per device, held from the first call until the process ends, and nothing
 * is allocated. A graph that needs more (roughly m above 5 M edges: the sort
 * storage takes about 24 B per edge) returns BT_ERR_SCRATCH. Every call on
 * one device shares that array: calls that overlap on one device (from
 * several threads or contexts) race and must be serialised by the caller. */
#define BT_STATIC_SCRATCH_BYTES (128u << 20)
int bt_ebfc_static(bt_ctx* ctx, bt_dgraph* dg, uint64_t* support_out, uint64_t* total,
                   uint64_t* support_max);

/* ---- counting launch settings per graph-size class -----------------------
 *
 * Replace the compile-time launch constants of the reference's counting
 * kernels (BLK_NUMS / BLK_DIM, src/util/config.h:9-10). Roots are split by
 * their priority-wedge count: up to wave_wedges, one 64-lane wavefront per
 * root with an LDS hash table of 2 * wave_wedges slots; up to block_wedges,
 * one workgroup of block_threads per root with an LDS table of
 * 2 * block_wedges slots; above (the hubs), chunks of hub_threads row
 * entries spread over the device, counting into dense tables in HBM. The
 * settings change only the speed, never a result. A graph's size class is
 * floor(log4(edges per graph)) (a batch union: its edges over its graphs),
 * clamped to BT_TUNE_CLASSES - 1. A context starts from the table tuned on
 * MI355X (bt_tune_default); bt_autotune measures every candidate on a
 * graph and keeps the fastest for that graph's class. */
#define BT_TUNE_CLASSES 16
typedef struct bt_count_config {
    uint32_t wave_wedges;   /* 256, 512, 1024 or 2048 */
    uint32_t block_wedges;  /* 4096, 8192 or 16384 */
    uint32_t block_threads; /* 256 or 512 */
    uint32_t hub_threads;   /* 256 or 512 */
} bt_count_config;
int bt_tune_class(uint64_t edges, uint32_t graphs);
int bt_tune_default(int size_class, bt_count_config* out);
/* the candidates bt_autotune measures: *n receives their number; out (may
 * be NULL) receives them, at most cap entries */
int bt_tune_candidates(bt_count_config* out, int cap, int* n);
int bt_tune_get(const bt_ctx* ctx, int size_class, bt_count_config* out);
int bt_tune_set(bt_ctx* ctx, int size_class, const bt_count_config* cfg);
/* Times the per-edge support count of dg under every candidate (best of
 * reps runs each), checks that every candidate's supports equal the first
 * one's (BT_ERR_INTERNAL otherwise), and sets the fastest for dg's class.
 * best and best_ms may be NULL; ms (may be NULL) receives the time of each
 * candidate in bt_tune_candidates order. */
int bt_autotune(bt_ctx* ctx, bt_dgraph* dg, int reps, bt_count_config* best, double* best_ms, double* ms);

/* h-index bitruss decomposition. Replaces g_bitruss_hindex
 * (bitruss_hidx.cu:179-256), which writes g->bitruss. bitruss_out (m
 * entries) may be NULL; rounds (may be NULL) receives the number of
 * rounds run. */
int bt_hidx(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* rounds);

/* GBiD multi-stage peeling with the reference's attribution rule.
 * Replaces bitruss_msp (bitruss_msp.cu:440-583), which writes
 * g->bitruss. bitruss_out (m entries) may be NULL; levels (may be NULL)
 * receives the value the reference logs as "levels" (the largest
 * bitruss number, 0 for a butterfly-free graph). With BT_OPT_PEEL_MAX_LEVEL
 * set and levels beyond it, the result is a prefix state and the status is
 * BT_PARTIAL (the outputs are written). BT_ERR_BUSY: a peel given part of
 * the device (BT_OPT_PEEL_BLOCKS) found its workgroups not resident
 * together twice (CUs held by other work, e.g. another process), once on
 * its share and once on the whole device. */
int bt_msp(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels);

/* The reference's baseline msp. Replaces bitruss_mspp
 * (bitruss_msp.cu:600-722): bt_msp's peel, except that an edge whose upper
 * endpoint has more than ten times the degree of its lower one is taken as
 * a pivot the other way round (process_edge, bitruss_msp.cu:366-378), which
 * changes the attribution of butterflies with three frontier edges. levels
 * as bt_msp (the reference logs "(baseline) bitruss msp with N levels").
 * Runs on the bloom index only (BT_ERR_BUDGET when it does not fit). */
int bt_mspp(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels);

/* True bitruss numbers by peeling. Replaces bitruss_peeling_cpu
 * (cpu_bitruss.cpp:34-122), the reference's sequential textbook peel: the
 * same level-synchronous peel as bt_msp in which every destroyed butterfly
 * costs each of its surviving edges exactly 1, which gives the same numbers
 * as peeling one edge of minimum support at a time (and as bt_hidx).
 * levels receives the largest bitruss number. Bloom index only. */
int bt_peel(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels);

/* ---- batches: many independent graphs decomposed together --------------
 *
 * No counterpart in the reference, which runs one graph per process. A batch
 * keeps k graphs on the device as disjoint unions (the upper vertices of
 * every graph, then the lower vertices; edge ids offset per graph) and runs
 * each algorithm once over each union. Butterflies never cross graphs, and
 * vertex priorities and edge ids keep their order inside each graph, so the
 * slice of a batch result that belongs to graph i equals the plain result
 * of graph i: msp sees the levels of other graphs as empty levels, and the
 * hidx round count of graph i is the round after its last change. */
typedef struct bt_batch bt_batch;

enum { BT_ALGO_BFC = 1, BT_ALGO_EBFC = 2, BT_ALGO_HIDX = 3, BT_ALGO_MSP = 4 };

/* Per-graph scalars of a batch run: the values the plain calls return. */
typedef struct bt_batch_result {
    uint64_t total;        /* bfc / ebfc: butterflies of the graph        */
    uint64_t support_max;  /* ebfc                                        */
    uint32_t levels;       /* msp: as bt_msp's `levels`                   */
    uint32_t rounds;       /* hidx: as bt_hidx's `rounds`                 */
} bt_batch_result;

/* Copy k host graphs (an array of k structs) to the device as one batch.
 * The graphs are spread over BT_OPT_BATCH_LANES lanes by their priority
 * wedges (a bound of their bloom records), largest first onto the lightest
 * lane; a lane's graphs share one union while their wedges fit a 32-bit
 * index and the lane's share of 85% of the free device memory (the lanes'
 * indexes are built side by side, at most 52 bytes per wedge), else the
 * lane holds several unions, run one after the other. The lanes run side
 * by side. */
int bt_batch_upload(bt_ctx* ctx, const bt_host_graph* graphs, size_t k, bt_batch** out);
int bt_batch_free(bt_batch* b);
/* Number of unions the batch was split into (over all lanes). */
int bt_batch_chunks(const bt_batch* b, size_t* chunks);

/* Run `algo` (BT_ALGO_*) on every graph of the batch. `outs` (k pointers, or
 * NULL) receives the per-edge results of graph i in outs[i] (m_i entries:
 * ebfc supports, hidx or msp bitruss numbers; a NULL entry skips that
 * graph). `results` (k entries, or NULL) receives the per-graph scalars. */
int bt_batch_run(bt_ctx* ctx, bt_batch* b, int algo, uint64_t* const* outs, bt_batch_result* results);

/* ---- serving: a request loop with continuous batching ------------------
 *
 * No counterpart in the reference. A server owns three lanes on one device,
 * each a context and a worker thread. Small requests (fewer than
 * large_edges edges) are decomposed together: a lane's worker takes every
 * pending request of the oldest request's algorithm (up to max_batch; on
 * the heavy lane up to batch_wedges priority wedges) and runs them as one
 * batch, while requests that arrive meanwhile wait for the next batch. Small
 * requests with fewer than light_wedges priority wedges go to the light
 * lane, the others to the heavy lane; large requests are batched on a lane
 * of their own (the light lane has two workers: requests that arrive while
 * one worker's batch runs are taken by the other at once). The lanes'
 * persistent peels run on disjoint CU shares
 * (light_cus, heavy_cus, large_cus, each rounded down to a multiple of the
 * 8 XCDs, at least 8), so no lane waits for another's. A request is
 * admitted when the estimated device bytes of the running requests plus its
 * own fit budget_bytes (the graph, ~140 B per edge, and its bloom index,
 * ~48 B per priority wedge); a request that alone exceeds the budget is
 * refused with BT_ERR_BUDGET. Results are those of the plain calls. */
typedef struct bt_server bt_server;

typedef struct bt_serve_config {
    int device;
    uint64_t budget_bytes;  /* admission budget (0: 64 GiB)                  */
    uint32_t large_edges;   /* edges from which a request runs alone
                               (0: 1,000,000)                               */
    uint32_t max_batch;     /* requests per batch at most (0: 1024)         */
    uint64_t light_wedges;  /* priority wedges below which a small request
                               takes the light lane (0: 2,000,000)          */
    uint64_t batch_wedges;  /* priority wedges a heavy-lane batch takes at
                               most, so that batches stay short (its first
                               request always fits; 0: 100,000,000)         */
    uint32_t light_cus;     /* CUs of the light lane's peels (0: 1/4)       */
    uint32_t heavy_cus;     /* CUs of the heavy lane's peels (0: 3/8)       */
    uint32_t large_cus;     /* CUs of the large lane's peels (0: the rest)  */
    uint32_t light_workers; /* workers (contexts) of the light lane, 1 or 2
                               (0: 2): while one worker's batch runs, the
                               other takes the requests that arrive, and
                               their copy, support count and bloom index
                               overlap the running peel                     */
} bt_serve_config;

typedef struct bt_serve_stats {
    uint64_t submitted, completed, failed;
    uint64_t batches, batched_requests;  /* light and heavy lanes        */
    uint64_t large_runs;                 /* batches of the large lane      */
    uint64_t peak_bytes;    /* largest admitted estimate in flight          */
    uint64_t in_use_bytes;  /* admitted estimate in flight now              */
} bt_serve_stats;

int bt_serve_start(const bt_serve_config* cfg, bt_server** out);
/* Queue a request; the graph is copied, so the caller's arrays may be
 * released on return. *ticket names the request. */
int bt_serve_submit(bt_server* s, const bt_host_graph* g, int algo, uint64_t* ticket);
/* *done = 1 once the request has finished (successfully or not). */
int bt_serve_poll(bt_server* s, uint64_t ticket, int* done);
/* Block until the request finishes, then return its status and copy its
 * per-edge result (m entries; not for BT_ALGO_BFC) to out (may be NULL),
 * its scalars to result and the time from submit to completion to
 * latency_ms (both may be NULL). The ticket is released. */
int bt_serve_wait(bt_server* s, uint64_t ticket, uint64_t* out, bt_batch_result* result, double* latency_ms);
int bt_serve_stats_get(bt_server* s, bt_serve_stats* stats);
/* Finish every queued request, then stop the workers and free the server
 * (tickets not waited for are dropped). */
int bt_serve_stop(bt_server* s);

#ifdef __cplusplus
}
#endif

#endif /* BITRUSS_HIP_H */
