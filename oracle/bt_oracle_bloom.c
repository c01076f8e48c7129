/*
 * bt_oracle_bloom.c — the msp (GBiD) decomposition restated over butterfly
 * blooms, on the host.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for graphs on which or_msp (the
 * reference's per-pivot merges, cpu_bitruss.cpp:353-424) takes hours. Only
 * tests/, the golden-vector script and the CPU-baseline leg of bench.py
 * load it; the product library never links it.
 *
 * Semantics: c_bitruss_msp (cpu_bitruss.cpp:536-597) — supports from ebfc,
 * support-zero edges final (process_support_zero, :241-256), then level by
 * level the frontier = unprocessed edges with support == level (scan_edge,
 * :258-272), peeled in sub-rounds; next = the edges that fell to the level
 * (update_edge_support, :275-297: s' = max(level, s - d)); the result is the
 * final support. Empty levels change nothing and are skipped: the next
 * level is the smallest live support.
 *
 * Every butterfly is counted once at its vertex of highest priority r
 * (priority = (degree, id), cpu_bfc.cpp:59-60); with w the vertex of r's
 * side, the butterflies at (r, w) are the pairs of middles of the bloom
 * (r, w) = { v in N(r) & N(w) : v below r }. In a sub-round, a butterfly
 * whose four edges are unprocessed and c of which are in the frontier is
 * seen by each of its c frontier edges as the pivot of ppeel_*_edge
 * (cpu_bitruss.cpp:301-351). Placing the survivors in the pivot roles gives
 * per butterfly:
 *   c = 1: the three survivors lose 1 (3 decrement calls);
 *   c = 2: the two survivors lose 1, issued once by the pivot of smaller id
 *          (2 calls);
 *   c = 3: the survivor f loses 1 iff (o<q<p) or (p<o<q) or (q<o<p), o the
 *          frontier edge opposite f, p the one sharing f's upper vertex, q
 *          the one sharing f's lower vertex (the single pivot whose literal
 *          condition holds makes the call);
 *   c = 4: nothing.
 * A bloom visit groups its live middles by frontier edges (n0, n1, n2):
 * both edges of a 0-middle lose n1 + n2; the survivor of a 1-middle loses
 * n0 + n1 - 1 plus the c = 3 rule over the 2-middles. Afterwards only the
 * 0-middles stay live in the bloom.
 *
 * The work counters equal or_msp's (the reference loops' trip counts):
 *   W_peel  per pivot uv (u upper, v lower): live edges at v, minus uv;
 *   I_peel  per such wedge u-v-w: deg(w) in the working CSR, which the
 *           compaction of cpu_bitruss.cpp:579-589 shrinks to the live edges;
 *   B_peel  per bloom visit: (n1 + 2 n2)(k - 1), k = n0 + n1 + n2;
 *   D_peel  per bloom visit: 3 n0 n1 + 2 n0 n2 + n1 (n1 - 1) + the c = 3
 *           decrements;
 *   sum_curr, sub_rounds as counted by the reference's loop.
 *
 * Pinned (tests/test_oracle.py): results, level count and every counter
 * equal or_msp's on random graphs, the quirk-firing seeds, K_{a,b}, the
 * hand-derived aw/quirk7 answers, the mix graphs and the 2M-edge graph's
 * golden; the level-bounded prefix equals or_msp_levels; the sorted-list
 * form of the c = 3 rule equals the pairwise form.
 */
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bt_oracle.h"

#define OR_OK 0
#define OR_EINVAL 1
#define OR_ENOMEM 3

#define DONE 0xFFu         /* edge state: processed (1, 2: frontier)     */
#define KWIN 8192u         /* levels per bucket window                    */
#define SORT_MIN 48u       /* 2-middles from which the rule uses sorted lists */

typedef struct {
    uint32_t* a;
    size_t n, cap;
} vec32;

static int vpush(vec32* v, uint32_t x) {
    if (v->n == v->cap) {
        size_t const c = v->cap ? 2 * v->cap : 256;
        uint32_t* p = (uint32_t*)realloc(v->a, c * sizeof(uint32_t));
        if (!p) return 1;
        v->a = p;
        v->cap = c;
    }
    v->a[v->n++] = x;
    return 0;
}

/* ------------------------------------------------------------- the index */

typedef struct {
    uint32_t n, m;
    uint32_t* order;   /* rank -> vertex                                  */
    uint32_t* rank;    /* vertex -> rank (position in (degree, id) order) */
    uint64_t* aoff;    /* by rank: adjacency offsets                      */
    uint32_t* anbr;    /* neighbour ranks, ascending per row              */
    uint32_t* aeid;    /* edge ids, parallel to anbr                      */
    uint64_t* boff;    /* by root rank: its blooms [boff[r], boff[r+1])   */
    uint32_t* bw;      /* far-end rank | root-is-upper << 31, ascending   */
    uint64_t* bstart;  /* first record                                    */
    uint32_t* blive;   /* live middles                                    */
    uint32_t* bclaim;  /* last claiming sub-round                         */
    uint64_t* rec;     /* middles: eid(r-v) << 32 | eid(v-w)              */
    uint64_t nb, nr;
} bix;

static void bix_free(bix* b) {
    free(b->order); free(b->rank); free(b->aoff); free(b->anbr); free(b->aeid);
    free(b->boff); free(b->bw); free(b->bstart); free(b->blive); free(b->bclaim); free(b->rec);
    memset(b, 0, sizeof(*b));
}

static int cmp_u64(const void* x, const void* y) {
    uint64_t const a = *(const uint64_t*)x, b = *(const uint64_t*)y;
    return a < b ? -1 : a > b;
}
static int cmp_u32(const void* x, const void* y) {
    uint32_t const a = *(const uint32_t*)x, b = *(const uint32_t*)y;
    return a < b ? -1 : a > b;
}

/* Rank order and the rank-relabelled adjacency (rows ascending). */
static int build_adjacency(const or_graph* g, int threads, bix* b) {
    uint32_t const n = g->n, m = g->m;
    b->n = n;
    b->m = m;
    b->order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    b->rank = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    b->aoff = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1ull));
    b->anbr = (uint32_t*)malloc(sizeof(uint32_t) * (2ull * m + 1));
    b->aeid = (uint32_t*)malloc(sizeof(uint32_t) * (2ull * m + 1));
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!b->order || !b->rank || !b->aoff || !b->anbr || !b->aeid || !key) {
        free(key);
        return OR_ENOMEM;
    }
    for (uint32_t x = 0; x < n; x++) key[x] = ((uint64_t)g->degrees[x] << 32) | x;
    qsort(key, n, sizeof(uint64_t), cmp_u64);
    for (uint32_t r = 0; r < n; r++) {
        b->order[r] = (uint32_t)key[r];
        b->rank[(uint32_t)key[r]] = r;
    }
    free(key);
    b->aoff[0] = 0;
    for (uint32_t r = 0; r < n; r++) b->aoff[r + 1] = b->aoff[r] + g->degrees[b->order[r]];
    int bad = 0;
#pragma omp parallel num_threads(threads) reduction(| : bad)
    {
        uint64_t* row = NULL;
        size_t cap = 0;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t r = 0; r < (int64_t)n; r++) {
            uint32_t const x = b->order[r];
            uint32_t const d = g->degrees[x];
            if (d > cap) {
                free(row);
                cap = d;
                row = (uint64_t*)malloc(sizeof(uint64_t) * cap);
                if (!row) { bad = 1; cap = 0; continue; }
            }
            for (uint32_t i = 0; i < d; i++)
                row[i] = ((uint64_t)b->rank[g->neighbors[g->offsets[x] + i]] << 32) | g->edge_ids[g->offsets[x] + i];
            qsort(row, d, sizeof(uint64_t), cmp_u64);
            for (uint32_t i = 0; i < d; i++) {
                b->anbr[b->aoff[r] + i] = (uint32_t)(row[i] >> 32);
                b->aeid[b->aoff[r] + i] = (uint32_t)row[i];
            }
        }
        free(row);
    }
    return bad ? OR_ENOMEM : OR_OK;
}

/* Blooms of every root in two passes over its priority wedges r-v-w
 * (v, w below r): count, then fill. Per root, blooms are ascending by far
 * end and a bloom's middles ascending by rank. */
static int build_blooms(const or_graph* g, int threads, bix* b) {
    uint32_t const n = b->n;
    uint32_t* nbr_cnt = (uint32_t*)calloc(n + 1ull, sizeof(uint32_t));
    uint64_t* nrec = (uint64_t*)calloc(n + 1ull, sizeof(uint64_t));
    b->boff = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1ull));
    if (!nbr_cnt || !nrec || !b->boff) { free(nbr_cnt); free(nrec); return OR_ENOMEM; }
    int bad = 0;
    for (int pass = 0; pass < 2 && !bad; pass++) {
        if (pass == 1) {
            uint64_t nb = 0, nr = 0;
            for (uint32_t r = 0; r < n; r++) {
                b->boff[r] = nb;
                nb += nbr_cnt[r];
                uint64_t const t = nrec[r];
                nrec[r] = nr;  /* now the first record of root r */
                nr += t;
            }
            b->boff[n] = nb;
            b->nb = nb;
            b->nr = nr;
            b->bw = (uint32_t*)malloc(sizeof(uint32_t) * (nb ? nb : 1));
            b->bstart = (uint64_t*)malloc(sizeof(uint64_t) * (nb + 1));
            b->blive = (uint32_t*)malloc(sizeof(uint32_t) * (nb ? nb : 1));
            b->bclaim = (uint32_t*)calloc(nb ? nb : 1, sizeof(uint32_t));
            b->rec = (uint64_t*)malloc(sizeof(uint64_t) * (nr ? nr : 1));
            if (!b->bw || !b->bstart || !b->blive || !b->bclaim || !b->rec) { bad = 1; break; }
            b->bstart[nb] = nr;
        }
#pragma omp parallel num_threads(threads) reduction(| : bad)
        {
            uint32_t* cnt = (uint32_t*)calloc(n + 1ull, sizeof(uint32_t));
            uint32_t* touched = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1ull));
            uint64_t* pos = pass ? (uint64_t*)malloc(sizeof(uint64_t) * (n + 1ull)) : NULL;
            if (!cnt || !touched || (pass && !pos)) bad = 1;
#pragma omp for schedule(dynamic, 256)
            for (int64_t rr = 0; rr < (int64_t)n; rr++) {
                if (bad) continue;
                uint32_t const r = (uint32_t)rr;
                uint32_t nt = 0;
                for (uint64_t i = b->aoff[r]; i < b->aoff[r + 1] && b->anbr[i] < r; i++) {
                    uint32_t const v = b->anbr[i];
                    for (uint64_t j = b->aoff[v]; j < b->aoff[v + 1] && b->anbr[j] < r; j++) {
                        uint32_t const w = b->anbr[j];
                        if (cnt[w]++ == 0) touched[nt++] = w;
                    }
                }
                if (pass == 0) {
                    uint32_t nb = 0;
                    uint64_t nr = 0;
                    for (uint32_t t = 0; t < nt; t++) {
                        uint32_t const w = touched[t];
                        if (cnt[w] >= 2) { nb++; nr += cnt[w]; }
                        cnt[w] = 0;
                    }
                    nbr_cnt[r] = nb;
                    nrec[r] = nr;
                    continue;
                }
                /* blooms of r ascending by far end */
                uint32_t k = 0;
                for (uint32_t t = 0; t < nt; t++)
                    if (cnt[touched[t]] >= 2) touched[k++] = touched[t];
                    else cnt[touched[t]] = 0;
                qsort(touched, k, sizeof(uint32_t), cmp_u32);
                uint32_t const up = b->order[r] < g->u_num;
                uint64_t cur = nrec[r];
                for (uint32_t t = 0; t < k; t++) {
                    uint32_t const w = touched[t];
                    uint64_t const bi = b->boff[r] + t;
                    b->bw[bi] = w | (up << 31);
                    b->bstart[bi] = cur;
                    b->blive[bi] = cnt[w];
                    pos[w] = cur;
                    cur += cnt[w];
                }
                for (uint64_t i = b->aoff[r]; i < b->aoff[r + 1] && b->anbr[i] < r; i++) {
                    uint32_t const v = b->anbr[i];
                    for (uint64_t j = b->aoff[v]; j < b->aoff[v + 1] && b->anbr[j] < r; j++) {
                        uint32_t const w = b->anbr[j];
                        if (cnt[w] >= 2) b->rec[pos[w]++] = ((uint64_t)b->aeid[i] << 32) | b->aeid[j];
                    }
                }
                for (uint32_t t = 0; t < k; t++) cnt[touched[t]] = 0;
            }
            free(cnt); free(touched); free(pos);
        }
    }
    free(nbr_cnt);
    free(nrec);
    return bad ? OR_ENOMEM : OR_OK;
}

/* Bloom (r, w) or -1: a search over r's blooms, ascending by far end. */
static inline int64_t find_bloom(const bix* b, uint32_t r, uint32_t w) {
    uint64_t lo = b->boff[r], hi = b->boff[r + 1];
    while (lo < hi) {
        uint64_t const mid = (lo + hi) >> 1;
        uint32_t const x = b->bw[mid] & 0x7FFFFFFFu;
        if (x < w) lo = mid + 1;
        else if (x > w) hi = mid;
        else return (int64_t)mid;
    }
    return -1;
}

/* Bloom (r, w) or -1, searching r's blooms from *pos on (w ascending over
 * successive calls): a galloping search that leaves *pos at the first
 * bloom with far end >= w. */
static inline int64_t find_bloom_from(const bix* b, uint32_t r, uint32_t w, uint64_t* pos) {
    uint64_t lo = *pos;
    uint64_t const end = b->boff[r + 1];
    uint64_t step = 1, hi = lo;
    while (hi < end && (b->bw[hi] & 0x7FFFFFFFu) < w) {
        lo = hi + 1;
        hi += step;
        step <<= 1;
    }
    if (hi > end) hi = end;
    while (lo < hi) {
        uint64_t const mid = (lo + hi) >> 1;
        if ((b->bw[mid] & 0x7FFFFFFFu) < w) lo = mid + 1;
        else hi = mid;
    }
    *pos = lo;
    return lo < end && (b->bw[lo] & 0x7FFFFFFFu) == w ? (int64_t)lo : -1;
}

/* First position in [lo, hi) of anbr with a value > x. */
static inline uint64_t upper_pos(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t x) {
    while (lo < hi) {
        uint64_t const mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* -------------------------------------------------------- the c = 3 rule */

static inline int rule3(uint32_t o, uint32_t p, uint32_t q) {
    return (o < q && q < p) || (p < o && o < q) || (q < o && o < p);
}

/* Decrement owed to the survivor of a 1-middle by the n2 2-middles
 * (a = their r-v ids, b = their v-w ids), pairwise. froot: the survivor is
 * the middle's r-v edge; g: its frontier edge; up: r is an upper vertex. */
static uint64_t rule_pairs(const uint32_t* a, const uint32_t* bb, uint32_t n2, int froot, int up, uint32_t g) {
    uint64_t d = 0;
    for (uint32_t j = 0; j < n2; j++) {
        uint32_t const o = froot ? bb[j] : a[j];
        uint32_t const s = froot ? a[j] : bb[j]; /* shares an end with the survivor */
        d += up ? rule3(o, s, g) : rule3(o, g, s);
    }
    return d;
}

/* Entries of the ascending list l[0, n) below g. */
static inline uint32_t below_cnt(const uint32_t* l, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t const mid = (lo + hi) >> 1;
        if (l[mid] < g) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* The same sum by counts: with P the 2-middles with a < b and Q those with
 * b < a, the three orderings of rule3 over {a, b, g} reduce to
 *   froot, up:    #{P: b < g} + #{Q: a > g}
 *   froot, !up:   #{P: b < g} + #{Q: a < g} + #{Q: b > g}
 *   !froot, up:   #{P: b > g} + #{Q: a < g}
 *   !froot, !up:  #{P: b < g} + #{P: a > g} + #{Q: a < g}
 * over four ascending id lists (the ids of a bloom's middles are distinct,
 * and g is not one of them). */
typedef struct {
    uint32_t *aP, *bP, *aQ, *bQ;
    uint32_t nP, nQ;
} sorted2;

static uint64_t rule_sorted(const sorted2* s, int froot, int up, uint32_t g) {
    uint32_t const bPl = below_cnt(s->bP, s->nP, g), aQl = below_cnt(s->aQ, s->nQ, g);
    if (froot && up) return (uint64_t)bPl + (s->nQ - aQl);
    if (froot) return (uint64_t)bPl + aQl + (s->nQ - below_cnt(s->bQ, s->nQ, g));
    if (up) return (uint64_t)(s->nP - bPl) + aQl;
    return (uint64_t)bPl + (s->nP - below_cnt(s->aP, s->nP, g)) + aQl;
}

/* ---------------------------------------------------------------- the peel */

typedef struct {
    uint64_t w, i, bt, d, sc, sub;
    uint64_t mid, visits; /* middles walked and bloom visits (progress lines) */
} counters;

typedef struct {
    vec32 claimed, touched, next;
    vec32* bucket;  /* KWIN per thread */
    uint32_t *c2a, *c2b, *tmp;
    uint8_t* cls;
    size_t c2cap;
    counters k;
    int oom;
} tstate;

static int grow_c2(tstate* t, size_t need) {
    if (need <= t->c2cap) return 0;
    size_t c = t->c2cap ? t->c2cap : 1024;
    while (c < need) c *= 2;
    free(t->c2a); free(t->c2b); free(t->tmp); free(t->cls);
    t->c2a = (uint32_t*)malloc(sizeof(uint32_t) * c);
    t->c2b = (uint32_t*)malloc(sizeof(uint32_t) * c);
    t->tmp = (uint32_t*)malloc(sizeof(uint32_t) * 4 * c);
    t->cls = (uint8_t*)malloc(c);
    t->c2cap = (t->c2a && t->c2b && t->tmp && t->cls) ? c : 0;
    return t->c2cap ? 0 : 1;
}

typedef struct {
    const or_graph* g;
    bix* b;
    uint64_t* sup;
    uint8_t* st;       /* 0 live, DONE processed, else the frontier stamp of
                        * its sub-round (1 and 2 alternate) */
    uint32_t* dec;     /* m: an edge's decrements in the current sub-round */
    uint32_t* livedeg; /* unprocessed edges per vertex */
    uint32_t* degcur;  /* degree in the working CSR (compaction) */
    uint64_t* S;       /* per vertex v: sum of degcur(w) over live edges wv */
    int flags;         /* OR_BLOOM_* test switches */
} peel;

/* Sums in one shared array (relaxed atomic adds; d > 0): the thread whose
 * add finds the sum at zero lists the edge for the apply step. A sub-round
 * takes at most the edge's support from it, below 2^32 (checked at the
 * start). */
static inline void add_dec(const peel* P, tstate* t, uint32_t e, uint64_t d) {
    if (__atomic_fetch_add(&P->dec[e], (uint32_t)d, __ATOMIC_RELAXED) == 0) t->oom |= vpush(&t->touched, e);
}

/* Claims, for sub-round X, every live bloom that holds frontier edge e,
 * and adds the edge's W_peel and I_peel as the reference's pivot. */
static void claim_edge(const peel* P, tstate* t, uint32_t e, uint32_t X) {
    const bix* b = P->b;
    uint32_t const u = P->g->edges[2ull * e], v = P->g->edges[2ull * e + 1];
    t->k.w += P->livedeg[v] - 1;
    t->k.i += P->S[v] - P->degcur[u];
    uint32_t const ru = b->rank[u], rv = b->rank[v];
    uint32_t const hi = ru > rv ? ru : rv, lo = ru > rv ? rv : ru;
#define CLAIM(bi)                                                                                 \
    do {                                                                                          \
        int64_t const bx_ = (bi);                                                                 \
        if (bx_ >= 0 && b->blive[bx_] >= 2 && __atomic_load_n(&b->bclaim[bx_], __ATOMIC_RELAXED) < X && \
            __atomic_exchange_n(&b->bclaim[bx_], X, __ATOMIC_RELAXED) < X)                         \
            t->oom |= vpush(&t->claimed, (uint32_t)bx_);                                          \
    } while (0)
    /* e = r-v with r = hi: blooms (hi, w), w in N(lo) below hi */
    uint64_t pos = b->boff[hi];
    for (uint64_t j = b->aoff[lo]; j < b->aoff[lo + 1] && b->anbr[j] < hi; j++)
        CLAIM(find_bloom_from(b, hi, b->anbr[j], &pos));
    /* e = v-w with v = lo, w = hi: roots r in N(lo) above hi */
    for (uint64_t j = upper_pos(b->anbr, b->aoff[lo], b->aoff[lo + 1], hi); j < b->aoff[lo + 1]; j++)
        CLAIM(find_bloom(b, b->anbr[j], hi));
    /* e = v-w with v = hi, w = lo: roots r in N(hi) above hi */
    for (uint64_t j = upper_pos(b->anbr, b->aoff[hi], b->aoff[hi + 1], hi); j < b->aoff[hi + 1]; j++)
        CLAIM(find_bloom(b, b->anbr[j], lo));
#undef CLAIM
}

/* One claimed bloom in sub-round X at level L. One pass classifies the
 * middles into t->cls (0, 1 survivor r-v, 2 survivor v-w, 3 both frontier,
 * 4 dead), prefetching the stamps of middles PF ahead; the decrement and
 * compaction passes read the classes instead of the stamps. */
#define PF 16
static void visit_bloom(const peel* P, tstate* t, uint32_t bi, uint8_t X) {
    bix* b = P->b;
    uint64_t* rec = b->rec + b->bstart[bi];
    uint32_t const k = b->blive[bi];
    int const up = (int)(b->bw[bi] >> 31);
    uint32_t n0 = 0, n1 = 0, n2 = 0, nP = 0;
    if (t->oom |= grow_c2(t, k)) return;
    t->k.mid += k;
    t->k.visits++;
    uint8_t* cls = t->cls;
    const uint8_t* st = P->st;
    for (uint32_t i = 0; i < k; i++) {
        if (i + PF < k) {
            __builtin_prefetch(&st[(uint32_t)(rec[i + PF] >> 32)], 0, 0);
            __builtin_prefetch(&st[(uint32_t)rec[i + PF]], 0, 0);
        }
        uint32_t const e1 = (uint32_t)(rec[i] >> 32), e2 = (uint32_t)rec[i];
        uint8_t const s1 = st[e1], s2 = st[e2];
        uint8_t c;
        if (s1 == DONE || s2 == DONE) c = 4;
        else if (s1 == X && s2 == X) {
            c = 3;
            t->c2a[n2] = e1;
            t->c2b[n2] = e2;
            n2++;
            nP += e1 < e2;
        } else if (s1 == X) { c = 2; n1++; }       /* survivor v-w */
        else if (s2 == X) { c = 1; n1++; }        /* survivor r-v */
        else { c = 0; n0++; }
        cls[i] = c;
    }
    uint64_t const kl = (uint64_t)n0 + n1 + n2;
    uint64_t rsum = 0;
    if (n1 + n2) {
        t->k.bt += ((uint64_t)n1 + 2ull * n2) * (kl - 1);
        int const use_sorted = n1 && n2 && ((P->flags & OR_BLOOM_SORTED) || n2 >= SORT_MIN);
        sorted2 s2l = {0};
        if (use_sorted) {
            s2l.aP = t->tmp;
            s2l.bP = t->tmp + nP;
            s2l.aQ = t->tmp + 2 * nP;
            s2l.bQ = t->tmp + 2 * nP + (n2 - nP);
            s2l.nP = nP;
            s2l.nQ = n2 - nP;
            uint32_t ip = 0, iq = 0;
            for (uint32_t j = 0; j < n2; j++) {
                if (t->c2a[j] < t->c2b[j]) { s2l.aP[ip] = t->c2a[j]; s2l.bP[ip] = t->c2b[j]; ip++; }
                else { s2l.aQ[iq] = t->c2a[j]; s2l.bQ[iq] = t->c2b[j]; iq++; }
            }
            qsort(s2l.aP, s2l.nP, sizeof(uint32_t), cmp_u32);
            qsort(s2l.bP, s2l.nP, sizeof(uint32_t), cmp_u32);
            qsort(s2l.aQ, s2l.nQ, sizeof(uint32_t), cmp_u32);
            qsort(s2l.bQ, s2l.nQ, sizeof(uint32_t), cmp_u32);
        }
        uint64_t const d0 = (uint64_t)n1 + n2;
        for (uint32_t i = 0; i < k; i++) {
            if (i + PF < k && cls[i + PF] <= 2) {
                __builtin_prefetch(&P->dec[(uint32_t)(rec[i + PF] >> 32)], 1, 0);
                __builtin_prefetch(&P->dec[(uint32_t)rec[i + PF]], 1, 0);
            }
            uint8_t const c = cls[i];
            if (c > 2) continue;
            uint32_t const e1 = (uint32_t)(rec[i] >> 32), e2 = (uint32_t)rec[i];
            if (c == 0) {
                add_dec(P, t, e1, d0);
                add_dec(P, t, e2, d0);
                continue;
            }
            int const froot = c == 1; /* the survivor is r-v */
            uint32_t const surv = froot ? e1 : e2, g = froot ? e2 : e1;
            uint64_t r = n2 ? (use_sorted ? rule_sorted(&s2l, froot, up, g)
                                          : rule_pairs(t->c2a, t->c2b, n2, froot, up, g))
                            : 0;
            /* OR_BLOOM_EXACT: every destroyed butterfly costs its survivor
             * 1, the textbook peel (cpu_bitruss.cpp:34-122), whose levels are
             * the true bitruss numbers; the PLANT flag is a known slip for
             * the tests: p and q exchanged */
            if (P->flags & OR_BLOOM_EXACT) r = n2;
            if (P->flags & OR_BLOOM_PLANT_SWAP) r = rule_pairs(t->c2a, t->c2b, n2, froot, !up, g);
            rsum += r;
            uint64_t const d = (uint64_t)n0 + (n1 - 1) + r;
            if (d) add_dec(P, t, surv, d);
        }
        t->k.d += 3ull * n0 * n1 + 2ull * n0 * n2 + (uint64_t)n1 * (n1 - 1) + rsum;
    }
    /* only the 0-middles stay live */
    if (n0 != k) {
        uint32_t w = 0;
        for (uint32_t i = 0; i < k; i++)
            if (cls[i] == 0) rec[w++] = rec[i];
        b->blive[bi] = w;
    }
}

/* The working-CSR degrees become the live degrees (the compaction of
 * cpu_bitruss.cpp:579-589 keeps the edges with support > level), and S is
 * recomputed. */
static void compact_counts(const peel* P, int threads) {
    const or_graph* g = P->g;
    uint32_t const n = g->n;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t x = 0; x < (int64_t)n; x++) P->degcur[x] = P->livedeg[x];
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024)
    for (int64_t x = 0; x < (int64_t)n; x++) {
        uint64_t s = 0;
        for (uint32_t i = g->offsets[x]; i < g->offsets[x + 1]; i++)
            if (P->st[g->edge_ids[i]] != DONE) s += P->degcur[g->neighbors[i]];
        P->S[x] = s;
    }
}

int or_msp_bloom(const or_graph* g, int threads, uint64_t max_level, int flags, uint64_t* state,
                 uint32_t* levels_out, int* complete, or_stats* stats) {
    return or_msp_bloom_ck(g, threads, max_level, flags, state, levels_out, complete, stats, NULL, 0, NULL, NULL);
}

int or_msp_bloom_ck(const or_graph* g, int threads, uint64_t max_level, int flags, uint64_t* state,
                    uint32_t* levels_out, int* complete, or_stats* stats, const uint64_t* ck_levels, uint32_t n_ck,
                    or_checkpoint_fn ck, void* ck_user) {
    if (!g || g->n != g->u_num + g->l_num || (g->m && (!state || !g->neighbors || !g->edges || !g->edge_ids)))
        return OR_EINVAL;
    if (threads < 1) threads = 1;
    uint32_t const m = g->m, n = g->n;
    int rc = or_ebfc(g, threads, state, NULL, NULL, stats);
    if (rc) return rc;
    bix b;
    memset(&b, 0, sizeof(b));
    peel P = {g, &b, state, NULL, NULL, NULL, NULL, NULL, flags};
    P.st = (uint8_t*)calloc(m ? m : 1, 1);
    P.dec = (uint32_t*)calloc(m ? m : 1, sizeof(uint32_t));
    P.livedeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    P.degcur = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    P.S = (uint64_t*)calloc(n ? n : 1, sizeof(uint64_t));
    uint32_t* fr = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));   /* frontier */
    uint32_t* cl = NULL;                                                /* claimed blooms */
    tstate* ts = (tstate*)calloc((size_t)threads, sizeof(tstate));
    if (!P.st || !P.dec || !P.livedeg || !P.degcur || !P.S || !fr || !ts) { rc = OR_ENOMEM; goto out; }
    for (int t = 0; t < threads; t++) {
        ts[t].bucket = (vec32*)calloc(KWIN, sizeof(vec32));
        if (!ts[t].bucket) { rc = OR_ENOMEM; goto out; }
    }
    for (uint32_t e = 0; e < m; e++)
        if (state[e] >> 32) { rc = OR_EINVAL; goto out; } /* supports beyond the u32 sums */
    /* OR_PROGRESS in the environment: progress lines on stderr */
    int const progress = getenv("OR_PROGRESS") != NULL;
    double const interval = progress && atof(getenv("OR_PROGRESS")) > 0 ? atof(getenv("OR_PROGRESS")) : 30.0;
    double const t0 = omp_get_wtime();
    double t_last = t0;
    if ((rc = build_adjacency(g, threads, &b)) || (rc = build_blooms(g, threads, &b))) goto out;
    if (progress)
        fprintf(stderr, "[or_msp_bloom] index: %llu blooms, %llu middles, %.1f s\n", (unsigned long long)b.nb,
                (unsigned long long)b.nr, omp_get_wtime() - t0);
    cl = (uint32_t*)malloc(sizeof(uint32_t) * (b.nb ? b.nb : 1));
    if (!cl) { rc = OR_ENOMEM; goto out; }

    /* process_support_zero (cpu_bitruss.cpp:241-256) */
    uint64_t live = 0;
    for (uint32_t e = 0; e < m; e++) {
        if (state[e] == 0) P.st[e] = DONE;
        else {
            live++;
            P.livedeg[g->edges[2ull * e]]++;
            P.livedeg[g->edges[2ull * e + 1]]++;
        }
    }
    for (uint32_t x = 0; x < n; x++) P.degcur[x] = g->degrees[x];
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024)
    for (int64_t x = 0; x < (int64_t)n; x++) {
        uint64_t s = 0;
        for (uint32_t i = g->offsets[x]; i < g->offsets[x + 1]; i++)
            if (P.st[g->edge_ids[i]] != DONE) s += P.degcur[g->neighbors[i]];
        P.S[x] = s;
    }

    uint64_t mcur = m;          /* edges in the working CSR */
    uint64_t last_level = 0;    /* last level peeled (0: none) */
    uint64_t L = 0, base = 0;   /* current level; bucket window [base, base + KWIN) */
    int have_window = 0;
    uint32_t X = 1;             /* sub-round stamp */
    int stopped = 0;
    uint32_t ck_next = 0;       /* next checkpoint */
    double t_claim = 0, t_visit = 0, t_apply = 0;
    while (live > 0) {
        /* ---- next level: the smallest live support, from the buckets */
        uint32_t nfr = 0;
        for (;;) {
            if (have_window) {
                for (uint64_t lv = L + 1; lv < base + KWIN && nfr == 0; lv++) {
                    for (int t = 0; t < threads; t++) {
                        vec32* bk = &ts[t].bucket[lv - base];
                        for (size_t i = 0; i < bk->n; i++) {
                            uint32_t const e = bk->a[i];
                            if (P.st[e] == 0 && state[e] == lv) fr[nfr++] = e;
                        }
                        bk->n = 0;
                    }
                    if (nfr) L = lv;
                }
                if (nfr) break;
            }
            /* refill: the window starts at the smallest live support */
            uint64_t mn = ~0ull;
#pragma omp parallel for num_threads(threads) reduction(min : mn)
            for (int64_t e = 0; e < (int64_t)m; e++)
                if (P.st[e] == 0 && state[e] < mn) mn = state[e];
            if (mn == ~0ull) break;
            base = mn;
            L = mn - 1;
            have_window = 1;
#pragma omp parallel num_threads(threads)
            {
                tstate* t = &ts[omp_get_thread_num()];
#pragma omp for schedule(static)
                for (int64_t e = 0; e < (int64_t)m; e++)
                    if (P.st[e] == 0 && state[e] < base + KWIN) t->oom |= vpush(&t->bucket[state[e] - base], (uint32_t)e);
            }
            for (int t = 0; t < threads; t++)
                if (ts[t].oom) { rc = OR_ENOMEM; goto out; }
        }
        if (nfr == 0) break; /* no live edge left (cannot happen while live > 0) */
        /* checkpoints: the state after every level up to ck_levels[i] */
        while (ck && ck_next < n_ck && L > ck_levels[ck_next]) ck(ck_levels[ck_next++], state, m, ck_user);
        if (max_level && L > max_level) { stopped = 1; break; }
        /* a run of empty levels ends with the compaction check of
         * cpu_bitruss.cpp:579 (nothing changes in between, so once) */
        if (L > last_level + 1 && (double)live <= (double)mcur * 0.5 && (double)mcur >= (double)m * 0.1) {
            compact_counts(&P, threads);
            mcur = live;
        }
        for (uint32_t i = 0; i < nfr; i++) P.st[fr[i]] = (uint8_t)(1 + (X & 1));
        /* ---- sub-rounds of level L */
        while (nfr > 0) {
            live -= nfr;
            uint32_t ncl = 0;
#pragma omp parallel num_threads(threads)
            {
                tstate* t = &ts[omp_get_thread_num()];
                t->claimed.n = t->touched.n = t->next.n = 0;
                double const tq0 = progress ? omp_get_wtime() : 0.0;
#pragma omp for schedule(dynamic, 16)
                for (int64_t i = 0; i < (int64_t)nfr; i++) claim_edge(&P, t, fr[i], X);
#pragma omp single
                {
                    if (progress) t_claim += omp_get_wtime() - tq0;
                    for (int q = 0; q < threads; q++) {
                        memcpy(cl + ncl, ts[q].claimed.a, sizeof(uint32_t) * ts[q].claimed.n);
                        ncl += (uint32_t)ts[q].claimed.n;
                    }
                }
                double const tv0 = progress ? omp_get_wtime() : 0.0;
#pragma omp for schedule(dynamic, 1)
                for (int64_t i = 0; i < (int64_t)ncl; i++) visit_bloom(&P, t, cl[i], (uint8_t)(1 + (X & 1)));
                if (progress && omp_get_thread_num() == 0) t_visit += omp_get_wtime() - tv0;
                double const ta0 = progress ? omp_get_wtime() : 0.0;
                /* s' = max(L, s - d) for the edges this thread listed (the
                 * visits ended at the loop's barrier); the edges that reach
                 * L are the next sub-round's frontier */
                for (size_t i = 0; i < t->touched.n; i++) {
                    uint32_t const e = t->touched.a[i];
                    uint64_t const d = P.dec[e], s = state[e];
                    P.dec[e] = 0;
                    if (s - L <= d) {
                        state[e] = L;
                        P.st[e] = (uint8_t)(1 + ((X + 1) & 1));
                        t->oom |= vpush(&t->next, e);
                    } else {
                        state[e] = s - d;
                        if (s - d < base + KWIN) t->oom |= vpush(&t->bucket[s - d - base], e);
                    }
                }
                /* the frontier is processed: live degrees and S follow */
#pragma omp for schedule(static)
                for (int64_t i = 0; i < (int64_t)nfr; i++) {
                    uint32_t const e = fr[i];
                    uint32_t const a = g->edges[2ull * e], c = g->edges[2ull * e + 1];
                    P.st[e] = DONE;
                    __atomic_fetch_sub(&P.livedeg[a], 1u, __ATOMIC_RELAXED);
                    __atomic_fetch_sub(&P.livedeg[c], 1u, __ATOMIC_RELAXED);
                    __atomic_fetch_sub(&P.S[a], (uint64_t)P.degcur[c], __ATOMIC_RELAXED);
                    __atomic_fetch_sub(&P.S[c], (uint64_t)P.degcur[a], __ATOMIC_RELAXED);
                }
                if (progress && omp_get_thread_num() == 0) t_apply += omp_get_wtime() - ta0;
            }
            for (int t = 0; t < threads; t++)
                if (ts[t].oom) { rc = OR_ENOMEM; goto out; }
            ts[0].k.sc += nfr;
            ts[0].k.sub++;
            nfr = 0;
            for (int t = 0; t < threads; t++) {
                memcpy(fr + nfr, ts[t].next.a, sizeof(uint32_t) * ts[t].next.n);
                nfr += (uint32_t)ts[t].next.n;
            }
            X++;
        }
        last_level = L;
        if (progress && omp_get_wtime() - t_last > interval) {
            t_last = omp_get_wtime();
            uint64_t mid = 0, vis = 0;
            for (int q = 0; q < threads; q++) {
                mid += ts[q].k.mid;
                vis += ts[q].k.visits;
            }
            fprintf(stderr, "[or_msp_bloom] %llu bloom visits, %llu middles walked\n", (unsigned long long)vis,
                    (unsigned long long)mid);
            fprintf(stderr,
                    "[or_msp_bloom] level %llu, sub-rounds %llu, live edges %llu, %.1f s (claims %.1f, blooms %.1f, "
                    "apply %.1f)\n",
                    (unsigned long long)L, (unsigned long long)ts[0].k.sub, (unsigned long long)live, t_last - t0,
                    t_claim, t_visit, t_apply);
        }
        /* the compaction check at the end of the level (cpu_bitruss.cpp:579) */
        if ((double)live <= (double)mcur * 0.5 && (double)mcur >= (double)m * 0.1) {
            compact_counts(&P, threads);
            mcur = live;
        }
    }
    /* the checkpoints at or past the last level see the final state */
    while (!stopped && ck && ck_next < n_ck) ck(ck_levels[ck_next++], state, m, ck_user);
    if (complete) *complete = !stopped && live == 0;
    if (levels_out) *levels_out = (uint32_t)(stopped ? max_level : last_level);
    if (stats) {
        for (int t = 0; t < threads; t++) {
            stats->w_peel += ts[t].k.w;
            stats->i_peel += ts[t].k.i;
            stats->b_peel += ts[t].k.bt;
            stats->d_peel += ts[t].k.d;
            stats->sum_curr += ts[t].k.sc;
            stats->sub_rounds += ts[t].k.sub;
        }
    }
out:
    if (ts) {
        for (int t = 0; t < threads; t++) {
            free(ts[t].claimed.a); free(ts[t].touched.a); free(ts[t].next.a);
            if (ts[t].bucket)
                for (uint32_t i = 0; i < KWIN; i++) free(ts[t].bucket[i].a);
            free(ts[t].bucket);
            free(ts[t].c2a); free(ts[t].c2b); free(ts[t].tmp); free(ts[t].cls);
        }
        free(ts);
    }
    free(P.st); free(P.dec); free(P.livedeg); free(P.degcur); free(P.S); free(fr); free(cl);
    bix_free(&b);
    return rc;
}
