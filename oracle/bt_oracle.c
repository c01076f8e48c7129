/*
 * bt_oracle.c — CPU restatement of the reference's bitruss algorithms.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that tests/, the smoke check and the
 * CPU-baseline leg of bench.py run. The product library never links it.
 *
 * Each function restates the reference CPU code (ShunyangLi/GPU_Bitruss)
 * with uint64 supports and byte flags; the reference uses int32 supports
 * and std::vector<bool> flags that race when threads > 1. Results do not
 * depend on the thread count: every concurrent update is an atomic
 * increment/decrement whose net effect is order independent.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   or_bfc / or_ebfc  brute force over upper-vertex pairs (total =
 *                     sum C(|N(a)&N(b)|,2), support by the same pairs),
 *                     K_{a,b} closed forms, hand-derived aw.graph answers,
 *                     zero-butterfly graphs.
 *   or_hidx           equals the textbook sequential peel (or_peel_seq) and
 *                     an independent Python peel on random small graphs;
 *                     K_{a,b}; aw.graph; the 7-edge quirk graph (all ones).
 *   or_msp            an independent butterfly-centric Python restatement of
 *                     the attribution rule (derived, not retyped: it
 *                     enumerates butterflies and applies the closed-form
 *                     survivor condition) on random small graphs; K_{a,b};
 *                     aw.graph; the quirk graph ([2,1,1,1,1,1,1], where msp
 *                     differs from the true bitruss numbers).
 *   or_peel_seq       independent Python textbook peel; K_{a,b}.
 * The compiled reference cannot be built here (its sources include
 * argparse and CUDA headers that the image lacks), so no function is
 * checked against reference binaries.
 */
#include "bt_oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_EINVAL 1
#define OR_ENOMEM 3

/* x has lower priority than u: cpu_bfc.cpp:59-60 (deg, then id). */
static inline int below(const uint32_t* deg, uint32_t x, uint32_t u) {
    return deg[x] < deg[u] || (deg[x] == deg[u] && x < u);
}

static int check_graph(const or_graph* g) {
    if (!g) return OR_EINVAL;
    if (g->n != g->u_num + g->l_num) return OR_EINVAL;
    if (g->m && (!g->neighbors || !g->edges || !g->edge_ids)) return OR_EINVAL;
    if (g->n && (!g->offsets || !g->degrees)) return OR_EINVAL;
    return OR_OK;
}

static int set_threads(int threads) { return threads < 1 ? 1 : threads; }

/* Shared body of ebfc/bfc: cpu_bfc.cpp:36-115 (pass 2 only when support). */
static int count_body(const or_graph* g, int threads, uint64_t* support, uint64_t* total_out,
                      uint64_t* wcnt_out) {
    const uint32_t* deg = g->degrees;
    const uint32_t* nbr = g->neighbors;
    const uint32_t* off = g->offsets;
    const uint32_t* eid = g->edge_ids;
    const uint32_t n = g->n;
    uint64_t total = 0, wcnt = 0;
    int fail = 0;
    threads = set_threads(threads);
#pragma omp parallel num_threads(threads) reduction(+ : total, wcnt) reduction(| : fail)
    {
        uint32_t* last_cnt = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
        uint32_t* last_use = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
        uint32_t idx = 0;
        if (!last_cnt || !last_use) fail = 1;
#pragma omp for schedule(dynamic, 64)
        for (int64_t uu = 0; uu < (int64_t)n; uu++) {
            if (fail) continue;
            const uint32_t u = (uint32_t)uu;
            for (uint32_t j = 0; j < idx; j++) last_cnt[last_use[j]] = 0;
            idx = 0;
            /* pass 1: count wedges u-v-w with v, w below u (cpu_bfc.cpp:56-78) */
            for (uint32_t i = off[u]; i < off[u + 1]; i++) {
                const uint32_t v = nbr[i];
                if (!below(deg, v, u)) continue;
                for (uint32_t j = off[v]; j < off[v + 1]; j++) {
                    const uint32_t w = nbr[j];
                    if (!below(deg, w, u)) continue;
                    wcnt++;
                    total += last_cnt[w];
                    if (++last_cnt[w] == 1) last_use[idx++] = w;
                }
            }
            if (!support) continue;
            /* pass 2: add cnt[w]-1 to uv and vw (cpu_bfc.cpp:80-107) */
            for (uint32_t i = off[u]; i < off[u + 1]; i++) {
                const uint32_t v = nbr[i];
                if (!below(deg, v, u)) continue;
                const uint32_t uv = eid[i];
                for (uint32_t j = off[v]; j < off[v + 1]; j++) {
                    const uint32_t w = nbr[j];
                    if (!below(deg, w, u)) continue;
                    if (last_cnt[w] == 0) continue;
                    const uint64_t dlt = (uint64_t)last_cnt[w] - 1;
                    if (!dlt) continue;
                    const uint32_t vw = eid[j];
                    __atomic_fetch_add(&support[uv], dlt, __ATOMIC_RELAXED);
                    __atomic_fetch_add(&support[vw], dlt, __ATOMIC_RELAXED);
                }
            }
        }
        free(last_cnt);
        free(last_use);
    }
    if (fail) return OR_ENOMEM;
    if (total_out) *total_out = total;
    if (wcnt_out) *wcnt_out = wcnt;
    return OR_OK;
}

int or_ebfc(const or_graph* g, int threads, uint64_t* support, uint64_t* total,
            uint64_t* support_max, or_stats* stats) {
    if (check_graph(g) || (g->m && !support)) return OR_EINVAL;
    memset(support, 0, sizeof(uint64_t) * (g->m ? g->m : 1));
    uint64_t t = 0, w = 0;
    int st = count_body(g, threads, support, &t, &w);
    if (st) return st;
    uint64_t mx = 0;
    for (uint32_t e = 0; e < g->m; e++)
        if (support[e] > mx) mx = support[e];
    if (total) *total = t;
    if (support_max) *support_max = mx;
    if (stats) stats->w_cnt += w;
    return OR_OK;
}

int or_bfc(const or_graph* g, int threads, uint64_t* total, or_stats* stats) {
    if (check_graph(g)) return OR_EINVAL;
    uint64_t t = 0, w = 0;
    int st = count_body(g, threads, NULL, &t, &w);
    if (st) return st;
    if (total) *total = t;
    if (stats) stats->w_cnt += w;
    return OR_OK;
}

/* ------------------------------------------------------------------ hidx */

/* h-index of the multiset vals[0..len): the largest h with at least h
 * values >= h (cpu_bitruss.cpp:176-185 sorts descending and scans; the
 * counting form below gives the same number). hist has len+1 slots. */
static uint64_t h_index(const uint64_t* vals, uint64_t len, uint64_t* hist) {
    if (len == 0) return 0;
    memset(hist, 0, sizeof(uint64_t) * (len + 1));
    for (uint64_t i = 0; i < len; i++) hist[vals[i] < len ? vals[i] : len]++;
    uint64_t acc = 0;
    for (uint64_t h = len; h >= 1; h--) {
        acc += hist[h];
        if (acc >= h) return h;
    }
    return 0;
}

typedef struct {
    uint64_t* vals;
    uint64_t* hist;
    uint64_t cap;
} hbuf;

static int hbuf_push(hbuf* b, uint64_t len, uint64_t x) {
    if (len == b->cap) {
        uint64_t nc = b->cap ? 2 * b->cap : 1024;
        uint64_t* nv = (uint64_t*)realloc(b->vals, nc * sizeof(uint64_t));
        if (!nv) return OR_ENOMEM;
        b->vals = nv;
        b->cap = nc;
    }
    b->vals[len] = x;
    return OR_OK;
}

int or_hidx(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* rounds,
            or_stats* stats) {
    if (check_graph(g) || (g->m && !bitruss)) return OR_EINVAL;
    const uint32_t m = g->m;
    uint64_t* prev = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
    uint64_t* curr = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
    if (!prev || !curr) {
        free(prev);
        free(curr);
        return OR_ENOMEM;
    }
    /* cpu_bitruss.cpp:202: supports from ebfc */
    int st = or_ebfc(g, threads, prev, NULL, NULL, stats);
    if (st) {
        free(prev);
        free(curr);
        return st;
    }
    const uint32_t* nbr = g->neighbors;
    const uint32_t* off = g->offsets;
    const uint32_t* eid = g->edge_ids;
    threads = set_threads(threads);
    uint32_t round = 0;
    uint64_t wh = 0, ih = 0, bh = 0;
    int changed = 1, fail = 0;
    while (changed && !fail) {
        /* one Jacobi round over all edges (cpu_bitruss.cpp:213-231) */
        round++;
        changed = 0;
#pragma omp parallel num_threads(threads) reduction(| : changed, fail) reduction(+ : wh, ih, bh)
        {
            hbuf b = {NULL, NULL, 0};
            uint64_t hcap = 0;
#pragma omp for schedule(dynamic, 64)
            for (int64_t ee = 0; ee < (int64_t)m; ee++) {
                const uint32_t e = (uint32_t)ee;
                curr[e] = 0;
                if (fail || prev[e] == 0) continue; /* :128 */
                const uint32_t u = g->edges[2ull * e], v = g->edges[2ull * e + 1];
                const uint32_t* un = nbr + off[u];
                const uint32_t ulen = off[u + 1] - off[u];
                uint64_t len = 0;
                for (uint32_t i = off[v]; i < off[v + 1]; i++) {
                    const uint32_t w = nbr[i];
                    if (w == u) continue;
                    const uint32_t wv = eid[i];
                    const uint32_t* wn = nbr + off[w];
                    const uint32_t wlen = off[w + 1] - off[w];
                    wh++;
                    ih += wlen;
                    /* merge N(w) and N(u), skipping the shared v (:153-173) */
                    uint32_t j = 0, k = 0;
                    while (j < wlen && k < ulen) {
                        if (wn[j] < un[k]) j++;
                        else if (wn[j] > un[k]) k++;
                        else if (wn[j] == v) { j++; k++; }
                        else {
                            const uint32_t wx = eid[off[w] + j];
                            const uint32_t ux = eid[off[u] + k];
                            j++;
                            k++;
                            bh++;
                            if (prev[ux] == 0 || prev[wx] == 0 || prev[wv] == 0) continue;
                            uint64_t mn = prev[ux];
                            if (prev[wx] < mn) mn = prev[wx];
                            if (prev[wv] < mn) mn = prev[wv];
                            if (hbuf_push(&b, len, mn)) { fail = 1; break; }
                            len++;
                        }
                    }
                    if (fail) break;
                }
                if (fail) continue;
                if (len + 1 > hcap) {
                    free(b.hist);
                    hcap = len + 1;
                    b.hist = (uint64_t*)malloc(hcap * sizeof(uint64_t));
                    if (!b.hist) { fail = 1; hcap = 0; continue; }
                }
                const uint64_t h = h_index(b.vals, len, b.hist);
                curr[e] = h;
                if (h != prev[e]) changed = 1;
            }
            free(b.vals);
            free(b.hist);
        }
        if (fail) break;
        uint64_t* t = prev;
        prev = curr;
        curr = t;
    }
    if (!fail) memcpy(bitruss, prev, sizeof(uint64_t) * m);
    free(prev);
    free(curr);
    if (fail) return OR_ENOMEM;
    if (rounds) *rounds = round;
    if (stats) {
        stats->w_h += wh;
        stats->i_h += ih;
        stats->b_h += bh;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------- msp */

typedef struct {
    uint64_t* sup;
    uint32_t* next;
    uint32_t* next_len; /* atomic append index */
    uint8_t* in_next;
    uint64_t level;
    uint64_t dcalls;
} peel_ctx;

/* update_edge_support (cpu_bitruss.cpp:275-297): decrement, join `next`
 * when the old value was level+1, undo when the old value was <= level. */
static inline void dec_edge(peel_ctx* c, uint32_t e) {
    const uint64_t cur = __atomic_fetch_sub(&c->sup[e], 1, __ATOMIC_RELAXED);
    c->dcalls++;
    if (cur == c->level + 1) {
        const uint32_t at = __atomic_fetch_add(c->next_len, 1u, __ATOMIC_RELAXED);
        c->next[at] = e;
        c->in_next[e] = 1;
    }
    if (cur <= c->level) __atomic_fetch_add(&c->sup[e], 1, __ATOMIC_RELAXED);
}

/* ppeel_one_edge / ppeel_two_edge / ppeel_three_edge
 * (cpu_bitruss.cpp:301-351), the literal per-pivot attribution rule. */
static inline void attribute(peel_ctx* c, uint32_t uv, uint32_t wv, uint32_t wx, uint32_t ux,
                             int pwv, int pwx, int pux) {
    const int cnt = pwv + pwx + pux;
    if (cnt == 1) {
        if (pwx && uv < ux && ux < wv) dec_edge(c, wx);
        if (pwv && uv < ux && ux < wx) dec_edge(c, wv);
        if (pux && uv < wv && wv < wx) dec_edge(c, ux);
    } else if (cnt == 2) {
        if (pwv && pwx && uv < ux) { dec_edge(c, wv); dec_edge(c, wx); }
        if (pwv && pux && uv < wx) { dec_edge(c, wv); dec_edge(c, ux); }
        if (pwx && pux && uv < wv) { dec_edge(c, wx); dec_edge(c, ux); }
    } else if (cnt == 3) {
        dec_edge(c, wv);
        dec_edge(c, wx);
        dec_edge(c, ux);
    }
}

/* max_seconds > 0 stops after the first level that ends past that time
 * (a bounded sample of the decomposition, for timing); max_level > 0 stops
 * after that level (the state of a prefix of the decomposition: peeled
 * edges hold their result, the others their current support); *complete
 * says whether every edge was peeled. */
static int msp_body(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* levels_out,
                    or_stats* stats, double max_seconds, uint64_t max_level, int* complete, int baseline) {
    if (check_graph(g) || (g->m && !bitruss)) return OR_EINVAL;
    const double t_start = omp_get_wtime();
    const uint32_t m = g->m, n = g->n;
    uint64_t* sup = bitruss; /* the result is the final support (:536-597) */
    int st = or_ebfc(g, threads, sup, NULL, NULL, stats);
    if (st) return st;
    threads = set_threads(threads);

    uint8_t* processed = (uint8_t*)calloc(m ? m : 1, 1);
    uint8_t* in_curr = (uint8_t*)calloc(m ? m : 1, 1);
    uint8_t* in_next = (uint8_t*)calloc(m ? m : 1, 1);
    uint32_t* curr = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    /* working CSR: neighbors / edge_ids / degrees compacted in place by
     * the compaction step; offsets are not changed (:500-530, :579-589) */
    uint32_t* nbr = (uint32_t*)malloc(sizeof(uint32_t) * (2ull * m + 1));
    uint32_t* eid = (uint32_t*)malloc(sizeof(uint32_t) * (2ull * m + 1));
    uint32_t* deg = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1ull));
    if (!processed || !in_curr || !in_next || !curr || !next || !nbr || !eid || !deg) {
        free(processed); free(in_curr); free(in_next); free(curr); free(next);
        free(nbr); free(eid); free(deg);
        return OR_ENOMEM;
    }
    if (m) {
        memcpy(nbr, g->neighbors, sizeof(uint32_t) * 2ull * m);
        memcpy(eid, g->edge_ids, sizeof(uint32_t) * 2ull * m);
    }
    if (n) memcpy(deg, g->degrees, sizeof(uint32_t) * n);
    const uint32_t* off = g->offsets;

    /* process_support_zero (:241-256) */
    int64_t edge_num = m;
    for (uint32_t e = 0; e < m; e++)
        if (sup[e] == 0) { processed[e] = 1; edge_num--; }

    uint64_t level = 1;
    uint64_t mcur = m;
    uint64_t wp = 0, ip = 0, bp = 0, dp = 0, sc = 0, rounds = 0;
    while (edge_num > 0) {
        /* scan_edge (:258-272) */
        uint32_t curr_len = 0;
        for (uint32_t e = 0; e < m; e++) {
            if (processed[e]) continue;
            if (sup[e] == level) { in_curr[e] = 1; curr[curr_len++] = e; }
        }
        while (curr_len > 0) {
            edge_num -= curr_len;
            sc += curr_len;
            rounds++;
            uint32_t next_len = 0;
            /* peel_edges (:353-424) */
#pragma omp parallel num_threads(threads) reduction(+ : wp, ip, bp, dp)
            {
                peel_ctx c = {sup, next, &next_len, in_next, level, 0};
#pragma omp for schedule(dynamic, 1)
                for (int64_t ci = 0; ci < (int64_t)curr_len; ci++) {
                    const uint32_t uv = curr[ci];
                    if (processed[uv]) continue;
                    uint32_t u = g->edges[2ull * uv], v = g->edges[2ull * uv + 1];
                    /* bitruss_mspp's process_edge (bitruss_msp.cu:366-378):
                     * the pivot is turned around when its upper endpoint's
                     * degree is above ten times its lower one's */
                    if (baseline && g->degrees[u] > 10ull * g->degrees[v]) {
                        const uint32_t t = u;
                        u = v;
                        v = t;
                    }
                    const uint32_t* un = nbr + off[u];
                    const uint32_t ulen = deg[u];
                    for (uint32_t x = 0; x < deg[v]; x++) {
                        const uint32_t w = nbr[off[v] + x];
                        if (w == u) continue;
                        const uint32_t wv = eid[off[v] + x];
                        if (processed[wv]) continue;
                        const uint32_t* wn = nbr + off[w];
                        const uint32_t wlen = deg[w];
                        wp++;
                        ip += wlen;
                        uint32_t j = 0, k = 0;
                        while (j < wlen && k < ulen) {
                            if (wn[j] < un[k]) j++;
                            else if (wn[j] > un[k]) k++;
                            else if (wn[j] == v) { j++; k++; }
                            else {
                                const uint32_t wx = eid[off[w] + j];
                                const uint32_t ux = eid[off[u] + k];
                                j++;
                                k++;
                                if (processed[ux] || processed[wx]) continue;
                                bp++;
                                attribute(&c, uv, wv, wx, ux, !in_curr[wv], !in_curr[wx],
                                          !in_curr[ux]);
                            }
                        }
                    }
                }
                dp += c.dcalls;
            }
            /* mark the frontier processed (:419-423), then swap (:572-576) */
            for (uint32_t i = 0; i < curr_len; i++) {
                processed[curr[i]] = 1;
                in_curr[curr[i]] = 0;
            }
            uint32_t* t = curr;
            curr = next;
            next = t;
            uint8_t* tf = in_curr;
            in_curr = in_next;
            in_next = tf;
            curr_len = next_len;
        }
        /* compaction trigger (:579); affects speed only: it drops edges
         * whose support is <= level, which are all processed by now
         * (bitruss_mspp runs without it) */
        if (!baseline && (double)edge_num <= (double)mcur * 0.5 && (double)mcur >= (double)m * 0.1) {
            uint64_t kept = 0;
            for (uint32_t x = 0; x < n; x++) {
                uint32_t k = 0;
                for (uint32_t i = 0; i < deg[x]; i++) {
                    const uint32_t e = eid[off[x] + i];
                    if (sup[e] > level) {
                        nbr[off[x] + k] = nbr[off[x] + i];
                        eid[off[x] + k] = e;
                        k++;
                    }
                }
                deg[x] = k;
                kept += k;
            }
            mcur = kept / 2;
        }
        level++;
        if (max_seconds > 0 && omp_get_wtime() - t_start > max_seconds) break;
        if (max_level > 0 && level > max_level) break;
    }
    if (complete) *complete = edge_num <= 0;
    free(processed); free(in_curr); free(in_next); free(curr); free(next);
    free(nbr); free(eid); free(deg);
    if (levels_out) *levels_out = (uint32_t)(level - 1);
    if (stats) {
        stats->w_peel += wp;
        stats->i_peel += ip;
        stats->b_peel += bp;
        stats->d_peel += dp;
        stats->sum_curr += sc;
        stats->sub_rounds += rounds;
    }
    return OR_OK;
}

int or_msp(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* levels_out,
           or_stats* stats) {
    return msp_body(g, threads, bitruss, levels_out, stats, 0.0, 0, NULL, 0);
}

int or_mspp(const or_graph* g, int threads, uint64_t* bitruss, uint32_t* levels_out, or_stats* stats) {
    return msp_body(g, threads, bitruss, levels_out, stats, 0.0, 0, NULL, 1);
}

int or_msp_sample(const or_graph* g, int threads, double max_seconds, uint64_t* work,
                  uint32_t* levels_done, int* complete, or_stats* stats) {
    return msp_body(g, threads, work, levels_done, stats, max_seconds, 0, complete, 0);
}

int or_msp_levels(const or_graph* g, int threads, uint64_t max_level, uint64_t* state, uint32_t* levels_done,
                  int* complete, or_stats* stats) {
    if (max_level == 0) return OR_EINVAL;
    return msp_body(g, threads, state, levels_done, stats, 0.0, max_level, complete, 0);
}

/* ------------------------------------------------------- textbook peel */

/* bitruss_peeling_cpu (cpu_bitruss.cpp:34-122): repeatedly take an
 * unassigned edge of minimum support, fix its bitruss number, and
 * decrement the other edges of its live butterflies while they exceed it. */
int or_peel_seq(const or_graph* g, uint64_t* bitruss) {
    if (check_graph(g) || (g->m && !bitruss)) return OR_EINVAL;
    const uint32_t m = g->m;
    uint64_t* sup = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
    uint8_t* done = (uint8_t*)calloc(m ? m : 1, 1);
    if (!sup || !done) { free(sup); free(done); return OR_ENOMEM; }
    int st = or_ebfc(g, 1, sup, NULL, NULL, NULL);
    if (st) { free(sup); free(done); return st; }
    const uint32_t* nbr = g->neighbors;
    const uint32_t* off = g->offsets;
    const uint32_t* eid = g->edge_ids;
    for (uint32_t e = 0; e < m; e++) {
        bitruss[e] = 0;
        if (sup[e] == 0) done[e] = 1;
    }
    for (;;) {
        uint32_t best = m;
        for (uint32_t e = 0; e < m; e++)
            if (!done[e] && (best == m || sup[e] < sup[best])) best = e;
        if (best == m) break;
        const uint32_t e = best;
        bitruss[e] = sup[e];
        const uint32_t u = g->edges[2ull * e], v = g->edges[2ull * e + 1];
        for (uint32_t i = off[v]; i < off[v + 1]; i++) {
            const uint32_t w = nbr[i];
            if (w == u) continue;
            const uint32_t wv = eid[i];
            if (done[wv]) continue;
            for (uint32_t j = off[w]; j < off[w + 1]; j++) {
                const uint32_t x = nbr[j];
                if (x == v) continue;
                const uint32_t wx = eid[j];
                if (done[wx]) continue;
                /* x in N(u)? binary search (:84-92) */
                uint32_t lo = off[u], hi = off[u + 1];
                uint32_t ux = m;
                while (lo < hi) {
                    uint32_t mid = lo + (hi - lo) / 2;
                    if (nbr[mid] < x) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < off[u + 1] && nbr[lo] == x) ux = eid[lo];
                if (ux == m || done[ux]) continue;
                if (sup[ux] > sup[e]) sup[ux]--;
                if (sup[wx] > sup[e]) sup[wx]--;
                if (sup[wv] > sup[e]) sup[wv]--;
            }
        }
        done[e] = 1;
    }
    free(sup);
    free(done);
    return OR_OK;
}
