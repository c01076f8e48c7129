/*
 * bt_oracle.h — CPU restatement of the reference's bitruss algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker: only tests/, the smoke
 * check in __graft_entry__.py and the CPU-baseline leg of bench.py load it.
 * The product library (libbitruss_hip.so) never links or calls it.
 *
 * Semantics follow the reference CPU code, with uint64 supports/counts and
 * byte flags (the reference uses int32 and std::vector<bool>):
 *   or_ebfc      src/bfc/cpu_bfc.cpp:24-122
 *   or_bfc       src/bfc/cpu_bfc.cpp:137-203
 *   or_hidx      src/bitruss/cpu_bitruss.cpp:125-239
 *   or_msp       src/bitruss/cpu_bitruss.cpp:241-597
 *   or_mspp      src/bitruss/bitruss_msp.cu:366-378, 600-722 (baseline msp)
 *   or_peel_seq  src/bitruss/cpu_bitruss.cpp:34-122 (textbook peel)
 */
#ifndef BT_ORACLE_H
#define BT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_graph {
    uint32_t u_num, l_num, n, m;
    const uint32_t* neighbors; /* 2m  */
    const uint32_t* offsets;   /* n+1 */
    const uint32_t* degrees;   /* n   */
    const uint32_t* edges;     /* 2m  */
    const uint32_t* edge_ids;  /* 2m  */
} or_graph;

/* Work counters (units of work of the benchmark). */
typedef struct or_stats {
    uint64_t w_cnt;      /* priority wedges: trip count of cpu_bfc.cpp:72     */
    uint64_t w_peel;     /* msp wedges: pivots' w with wv unprocessed         */
    uint64_t i_peel;     /* msp intersection reads: sum deg_cur(w) per wedge  */
    uint64_t b_peel;     /* msp butterflies reaching the attribution rule     */
    uint64_t d_peel;     /* msp update_edge_support calls                     */
    uint64_t sum_curr;   /* msp total frontier size over all sub-rounds       */
    uint64_t sub_rounds; /* msp sub-rounds                                    */
    uint64_t w_h;        /* hidx wedges, summed over rounds                   */
    uint64_t i_h;        /* hidx intersection reads, summed over rounds       */
    uint64_t b_h;        /* hidx butterflies, summed over rounds              */
} or_stats;

int or_ebfc(const or_graph* g, int threads, uint64_t* support, uint64_t* total,
            uint64_t* support_max, or_stats* stats);
int or_bfc(const or_graph* g, int threads, uint64_t* total, or_stats* stats);
int or_hidx(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* rounds,
            or_stats* stats);
int or_msp(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* levels,
           or_stats* stats);
/* bitruss_mspp (bitruss_msp.cu:600-722): or_msp with each pivot taken the
 * other way round when deg(upper) > 10 deg(lower) (process_edge, :366-378)
 * and without the compaction */
int or_mspp(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* levels,
            or_stats* stats);
/* The msp computation cut short for timing: support counting, then peel
 * levels until the first level that ends past max_seconds. `work` (m)
 * holds the supports as they stand; *complete is 1 when every edge got
 * its result. The stats count the work done. */
int or_msp_sample(const or_graph* g, int threads, double max_seconds, uint64_t* work,
                  uint32_t* levels_done, int* complete, or_stats* stats);
/* msp through level max_level only: state[e] is the result of a peeled
 * edge and the current support of the others (cpu_bitruss.cpp:536-597 with
 * the level loop cut) */
int or_msp_levels(const or_graph* g, int threads, uint64_t max_level, uint64_t* state, uint32_t* levels_done,
                  int* complete, or_stats* stats);
int or_peel_seq(const or_graph* g, uint64_t* bitruss);
/* msp over butterfly blooms (bt_oracle_bloom.c): the results, level count
 * and work counters of or_msp, for graphs on which the reference's
 * per-pivot merges take hours. max_level > 0 stops after that level with
 * the prefix state of or_msp_levels. flags: OR_BLOOM_SORTED sums the
 * 3-frontier-edge rule from sorted id lists wherever it applies (a test
 * switch); OR_BLOOM_EXACT decrements the single survivor always, which is
 * the textbook peel (the results of or_peel_seq: true bitruss numbers);
 * OR_BLOOM_PLANT_SWAP puts a known slip into the rule (p and q exchanged)
 * so that tests can show the comparison with or_msp catches it. */
#define OR_BLOOM_SORTED 1
#define OR_BLOOM_EXACT 2
#define OR_BLOOM_PLANT_SWAP 4
int or_msp_bloom(const or_graph* g, int threads, uint64_t max_level, int flags, uint64_t* state,
                 uint32_t* levels_out, int* complete, or_stats* stats);
/* The same with checkpoints: for each ck_levels[i] (ascending) ck is called
 * once with the state after every level up to it (the prefix state of
 * or_msp_levels with max_level = ck_levels[i]), so one long run records
 * prefixes at several depths. */
typedef void (*or_checkpoint_fn)(uint64_t level, const uint64_t* state, uint32_t m, void* user);
int or_msp_bloom_ck(const or_graph* g, int threads, uint64_t max_level, int flags, uint64_t* state,
                    uint32_t* levels_out, int* complete, or_stats* stats, const uint64_t* ck_levels, uint32_t n_ck,
                    or_checkpoint_fn ck, void* ck_user);

#ifdef __cplusplus
}
#endif

#endif /* BT_ORACLE_H */
