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

#ifdef __cplusplus
}
#endif

#endif /* BT_ORACLE_H */
