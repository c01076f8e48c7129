/*
 * bt_graph.h — host-side bipartite graph C-ABI: the reference's Graph CSR,
 * its text->.bin converter, its .bin reader/writer, its consistency check,
 * and the seeded synthetic generators used by the tests and the benchmark.
 *
 * Every function returns an int status (0 = BT_OK); bt_status_str() gives
 * the text. Nothing here exits the process (the reference exit()s on every
 * error; its CLI behaviour is reproduced by the `coh` host on top of this).
 *
 * Layout of bt_host_graph is the reference's `class Graph` public CSR
 * (src/graph/graph.h:43-51): vertex ids are u32, upper vertices are
 * 0..u_num-1, lower vertices are u_num..n-1, every adjacency list is sorted
 * ascending, edge e joins edges[2e] (upper) and edges[2e+1] (lower), and
 * edge_ids[offsets[x]+i] is the id of the edge x--neighbors[offsets[x]+i]
 * (the same id from both endpoints).
 */
#ifndef BT_GRAPH_H
#define BT_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared by every entry point of libbitruss_hip.so. */
enum bt_status {
    BT_OK = 0,
    BT_ERR_INVALID = 1,     /* bad argument or malformed graph            */
    BT_ERR_IO = 2,          /* file cannot be opened / read / written     */
    BT_ERR_NOMEM = 3,       /* host allocation failed                     */
    BT_ERR_DEVICE = 4,      /* a HIP runtime call failed                  */
    BT_ERR_NO_DEVICE = 5,   /* no HIP device, or the device id is invalid */
    BT_ERR_BUDGET = 6,      /* request exceeds the device-memory budget   */
    BT_ERR_OVERFLOW = 7,    /* a count does not fit its type              */
    BT_ERR_INTERNAL = 8,    /* an internal invariant failed               */
    BT_ERR_STATE = 9,       /* call not valid in the current state        */
    BT_PARTIAL = 10,        /* msp stopped at the level bound of
                               BT_OPT_PEEL_MAX_LEVEL: the results are the
                               state of that prefix, not the decomposition */
    BT_ERR_BUSY = 11,       /* the blocks of a persistent peel on part of
                               the device were not resident together (its
                               grid barrier timed out): other work held
                               the CUs it was given                         */
    BT_ERR_SCRATCH = 12     /* the graph needs more than the fixed static
                               scratch of an allocation-free entry point    */
};

/* Text for a status code (static storage, never NULL). */
const char* bt_status_str(int status);

/* Largest raw vertex id the converter accepts, exclusive
 * (MAX_IDS in src/util/config.h:6). */
#define BT_MAX_IDS 99999999u

typedef struct bt_host_graph {
    uint32_t u_num, l_num, u_max, m, n, support_max;
    uint32_t* neighbors; /* 2m  */
    uint32_t* offsets;   /* n+1 */
    uint32_t* degrees;   /* n   */
    uint32_t* edges;     /* 2m  : (upper, lower) per edge id */
    uint32_t* edge_ids;  /* 2m  : parallel to neighbors     */
} bt_host_graph;

/* Raw edge list: pairs (upper[i], lower[i]) of raw (file) vertex ids in
 * input order. */
typedef struct bt_edge_list {
    uint64_t count;
    uint32_t* upper;
    uint32_t* lower;
} bt_edge_list;

/* Build the CSR from raw pairs with the reference's process_graph semantics
 * (src/graph/graph.cpp:74-242): first-appearance relabel of upper and of
 * lower raw ids, lower ids shifted by u_num, sorted lists, edge ids in
 * (upper id, neighbour) order, lower-side ids by the same binary search.
 * Raw ids must be < BT_MAX_IDS. Duplicate pairs are kept, as in the
 * reference. Arrays of *out are malloc'ed; release with bt_graph_release. */
int bt_graph_from_edges(const uint32_t* raw_upper, const uint32_t* raw_lower,
                        uint64_t count, bt_host_graph* out);

/* Parse a KONECT-style text edge list (lines starting with '%' skipped, the
 * first two integers of each line are "upper lower") and build the CSR as
 * bt_graph_from_edges does (src/graph/graph.cpp:74-114). Blank lines are
 * skipped; a non-blank line without two non-negative integers is
 * BT_ERR_INVALID. */
int bt_graph_read_text(const char* path, bt_host_graph* out);

/* Read / write the reference .bin layout byte for byte
 * (src/graph/graph.cpp:293-318 write, 335-375 read). The edge_support and
 * bitruss arrays of the file are written as zeros and skipped on read.
 * The reader validates sizes and CSR invariants and fails with
 * BT_ERR_INVALID instead of reading out of bounds. */
int bt_graph_read_bin(const char* path, bt_host_graph* out);
int bt_graph_write_bin(const char* path, const bt_host_graph* g);

/* Structural validation: offsets monotone and consistent with degrees,
 * ids in range, sorted lists, and the edge-id consistency that the
 * reference's check_graph tests (src/graph/graph.cpp:13-52). */
int bt_graph_check(const bt_host_graph* g);

/* Priority wedges of a graph: #{(u,v,w): v in N(u), w in N(v), v < u,
 * w < u} for the vertex priority (degree, id) of cpu_bfc.cpp:72-90, the
 * number of wedges butterfly counting enumerates (an upper bound of the
 * bloom records of the peel). O(m log deg) on the host. */
int bt_graph_wedges(const bt_host_graph* g, uint64_t* out);

/* Disjoint union of k graphs (an array of k structs) in the batch layout:
 * the upper vertices of graph 0, 1, ..., then the lower vertices of graph
 * 0, 1, ...; each graph's ids keep their order, its edge ids are offset by
 * e_off[i] (k + 1 entries, filled here). No counterpart in the reference.
 * Fails with BT_ERR_OVERFLOW when 2m or n does not fit 32 bits. */
int bt_graph_union(const bt_host_graph* graphs, uint64_t k, bt_host_graph* out, uint64_t* e_off);

/* Free the arrays of a graph filled by this library and zero it. */
void bt_graph_release(bt_host_graph* g);
void bt_edge_list_release(bt_edge_list* e);

/* ---- seeded generators (integer PRNG: splitmix64-seeded xoshiro256**) --- */

/* m distinct pairs drawn uniformly from [0,nu) x [0,nv) without
 * replacement, in draw order. m must be <= nu*nv. */
int bt_gen_uniform(uint32_t nu, uint32_t nv, uint64_t m, uint64_t seed,
                   bt_edge_list* out);

/* Bipartite Chung-Lu: expected-degree weights w_i = min(cap, c*(i+1)^(-1/(g-1)))
 * on each side, each side scaled so its weights sum to m_target, and edge
 * (u,v) present independently with p = min(1, w_u*w_v/m_target), sampled
 * in O(nu + m) by geometric skipping (Miller & Hagberg 2011). Edges are
 * distinct, emitted upper-major. */
int bt_gen_chung_lu(uint32_t nu, uint32_t nv, double gamma_u, double gamma_v,
                    double m_target, double cap_u, double cap_v, uint64_t seed,
                    bt_edge_list* out);

/* Named workloads:
 *   BT_WL_SMALL_UNIFORM   1,000 x 1,500 uniform, m = 10,000, seed 0
 *   BT_WL_POWERLAW_2M     Chung-Lu 200,000 x 100,000, gamma 2.3/2.3,
 *                         m ~ 2,000,000, expected degree cap 5,000, seed 1
 *   BT_WL_SKEWED_60M      Chung-Lu 4,000,000 (gamma 2.5) x 1,000,000
 *                         (gamma 2.0), m ~ 60,000,000, cap 200,000, seed 2
 *   BT_WL_MIX_GRAPH       one graph of the 4,096-graph mix; `index` in
 *                         [0,4096) uses seed 1000+index: |U|,|V| log-uniform
 *                         in [50,5000], density log-uniform in [0.5%,5%],
 *                         even seeds uniform, odd seeds Chung-Lu gamma 2.2
 *   BT_WL_LARGE_REQUEST   one large serving request: Chung-Lu gamma 2.2,
 *                         |U| in [100k,500k], |V| in [50k,200k],
 *                         m in [1M,5M], drawn from seed `index`          */
enum bt_workload {
    BT_WL_SMALL_UNIFORM = 1,
    BT_WL_POWERLAW_2M = 2,
    BT_WL_SKEWED_60M = 3,
    BT_WL_MIX_GRAPH = 4,
    BT_WL_LARGE_REQUEST = 5
};
int bt_gen_workload(int workload, uint64_t index, bt_edge_list* out);

#ifdef __cplusplus
}
#endif

#endif /* BT_GRAPH_H */
